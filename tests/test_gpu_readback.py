"""svo_read_image_async: the image copies that overlap the next frame hold
exactly the images blocking reads give, for graph and direct warp frames
and for full frames, and the double-buffered image leaves the rest of the
frame state untouched."""
import numpy as np
import pytest

import svo_amd as S
from helpers import CONFIG1, assert_equal_u32, params, path, scene_octree


@pytest.fixture(scope="module")
def scene1():
    return scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])


@pytest.mark.gpu
@pytest.mark.parametrize("graph,pinned", [(True, True), (True, False), (False, True)])
def test_async_image_reads_match_blocking_reads(scene1, graph, pinned):
    words, root = scene1
    rx, ry = CONFIG1["res"]
    n = 10
    pr = path(CONFIG1, n)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, CONFIG1["depth"])
            c.resize(rx, ry)
            c.set_graph(graph)
        pins = [S.PinnedImage(rx, ry) for _ in range(n + 3)] if pinned else []
        got = [b.array for b in pins] if pinned else [np.zeros((ry, rx), np.uint32) for _ in range(n + 3)]
        want = []
        for f in range(n):
            print(f"graph={graph} warp frame {f}", flush=True)
            p = params(CONFIG1, pr, f)
            a.frame_warp(p)
            a.read_image_async(got[f])
            b.frame_warp(p)
            want.append(b.image())
            if f % 3 == 2:  # a blocking read before the next frame sees the same frame
                assert_equal_u32(a.image(), want[f], f"blocking read after async, frame {f}")
        for k in range(3):  # full frames through the same two buffers
            p = params(CONFIG1, pr, n - 1 - k)
            a.frame_full(p)
            a.read_image_async(got[n + k])
            b.frame_full(p)
            want.append(b.image())
        a.sync()
        for f in range(n + 3):
            assert_equal_u32(got[f], want[f], f"async image {f}")
        assert_equal_u32(a.screen(), b.screen(), "screen planes")
        got = None
        for p in pins:
            p.close()
    finally:
        a.close()
        b.close()


@pytest.mark.gpu
def test_async_read_then_resize(scene1):
    """A resize after an async read waits for the copy and starts over with
    one image buffer."""
    words, root = scene1
    c = S.Context(0)
    try:
        c.upload_octree(words, root, CONFIG1["depth"])
        c.resize(*CONFIG1["res"])
        pr = path(CONFIG1, 4)
        out = np.zeros((CONFIG1["res"][1], CONFIG1["res"][0]), np.uint32)
        c.frame_warp(params(CONFIG1, pr, 0))
        c.read_image_async(out)
        c.resize(200, 136)
        c.frame_warp(params(CONFIG1, pr, 0))
        small = np.zeros((136, 200), np.uint32)
        c.read_image_async(small)
        c.sync()
        assert_equal_u32(small, c.image(), "after resize")
        with pytest.raises(ValueError):
            c.read_image_async(out)  # wrong size for the new resolution
    finally:
        c.close()


@pytest.mark.gpu
def test_async_read_then_per_stage_calls_keep_the_image(scene1):
    """Per-stage calls that write part of the image or none (projection,
    compaction, rays, copy, fillhole) leave the image that svo_read returns
    alone after an async read; svo_colorize, which rewrites all of it, then
    renders into the other buffer and equals a context that never read."""
    words, root = scene1
    rx, ry = CONFIG1["res"]
    pr = path(CONFIG1, 4)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, CONFIG1["depth"])
            c.resize(rx, ry)
        for f in range(3):
            a.frame_warp(params(CONFIG1, pr, f))
            b.frame_warp(params(CONFIG1, pr, f))
        out = np.zeros((ry, rx), np.uint32)
        a.read_image_async(out)
        a.sync()
        img = a.image()
        assert_equal_u32(out, img, "async copy")
        p = params(CONFIG1, pr, 3)
        for c in (a, b):
            c.memset(0, S.HOLE, 4 * c.N)
            c.proj(2, p)
            c.counthole()
            c.sumids()
            c.writeids()
            c.raycast_holes(p)
            c.raycast_rect(p, 0, 0, 64, 64)
            c.copy_cache(2)
            c.fillhole()
        assert_equal_u32(a.image(), img, "image after per-stage calls")
        for c in (a, b):
            c.colorize()
        assert_equal_u32(a.image(), b.image(), "colorize after an async read")
        assert_equal_u32(a.screen(), b.screen(), "screen planes")
    finally:
        a.close()
        b.close()


@pytest.mark.gpu
def test_async_reads_of_learned_and_batch_frames(scene1):
    """Learned frames and batch frames also render into the other image
    buffer after an async read: every copy equals the image of a context
    that never reads asynchronously."""
    from svo_amd import learned as L
    words, root = scene1
    rx, ry = CONFIG1["res"]
    pr = path(CONFIG1, 8)
    cnn = S.Cnn(0)
    L.load(cnn, L.WEIGHTS)
    ctxs = [S.Context(0) for _ in range(4)]
    batch = None
    try:
        for c in ctxs:
            c.upload_octree(words, root, CONFIG1["depth"])
            c.resize(rx, ry)
        a, b = ctxs[0], ctxs[1]
        got = [np.zeros((ry, rx), np.uint32) for _ in range(6)]
        want = []
        for f in range(6):
            p = params(CONFIG1, pr, f)
            a.frame_warp_learned(cnn, p, 64)
            a.read_image_async(got[f])
            b.frame_warp_learned(cnn, p, 64)
            want.append(b.image())
        a.sync()
        for f in range(6):
            assert_equal_u32(got[f], want[f], f"learned frame {f}")
        # a batch of one view (ctxs[2]) against ctxs[3] alone
        batch = S.Batch([ctxs[2]])
        got = [np.zeros((ry, rx), np.uint32) for _ in range(6)]
        want = []
        for f in range(6):
            p = params(CONFIG1, pr, f)
            batch.frame_warp([p])
            batch.sync()
            ctxs[2].read_image_async(got[f])
            ctxs[3].frame_warp(p)
            want.append(ctxs[3].image())
        ctxs[2].sync()
        for f in range(6):
            assert_equal_u32(got[f], want[f], f"batch frame {f}")
    finally:
        if batch is not None:
            batch.close()
        for c in ctxs:
            c.close()
        cnn.close()

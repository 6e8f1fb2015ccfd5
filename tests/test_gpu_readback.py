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

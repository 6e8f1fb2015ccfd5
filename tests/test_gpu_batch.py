"""Batched views: svo_batch_frame_warp over several contexts of one scene
must give, per view, the words svo_frame_warp gives on that context alone
(and so the oracle's), for mixed resolutions and frame counters."""
import numpy as np
import pytest

import oracle as O
import svo_amd as S
from helpers import CONFIG1, assert_equal_f32_bits, assert_equal_u32, path, scene_octree

pytestmark = pytest.mark.gpu

# mixed sizes, including ones that are not multiples of 16 or 8
RES = [(512, 256), (320, 240), (200, 136), (128, 64), (640, 480), (264, 88)]


def view_params(pr, view, f):
    # each view follows the same path at its own offset and frame counter
    row = pr[(3 * view + f) % len(pr)]
    yaw, pitch, roll, x, y, z = [float(v) for v in row]
    p, _ = S.camera(yaw, pitch, roll, (x, y, z), CONFIG1["depth"], f + view)
    return p


def make_views(words, root, res):
    out = []
    for rx, ry in res:
        c = S.Context(0)
        c.upload_octree(words, root, CONFIG1["depth"])
        c.resize(rx, ry)
        out.append(c)
    return out


def compare(a, b, what):
    assert_equal_u32(a.screen(), b.screen(), f"{what} screen")
    assert_equal_f32_bits(a.coords(), b.coords(), f"{what} coords")
    assert_equal_u32(a.ids(), b.ids(), f"{what} ids")
    assert_equal_u32(a.image(), b.image(), f"{what} image")


def test_batch_matches_single_views():
    words, root = scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])
    pr = path(CONFIG1, 32)
    batched = make_views(words, root, RES)
    single = make_views(words, root, RES)
    batch = S.Batch(batched)
    try:
        for f in range(8):
            ps = [view_params(pr, v, f) for v in range(len(RES))]
            print(f"batch frame {f}", flush=True)
            batch.frame_warp(ps)
            for c, p in zip(single, ps):
                c.frame_warp(p)
            for v, (a, b) in enumerate(zip(batched, single)):
                a.sync()
                b.sync()
                compare(a, b, f"frame {f} view {v} {RES[v]}")
    finally:
        batch.close()
        for c in batched + single:
            c.close()


def test_batch_interleaved_with_single_frames_vs_oracle():
    """Single-view frames queued between batch frames stay ordered; one view
    is also checked against the oracle directly."""
    words, root = scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])
    pr = path(CONFIG1, 32)
    res = [(200, 136), (128, 64)]
    views = make_views(words, root, res)
    oracles = [O.Frame(rx, ry, words, root, CONFIG1["depth"], nthreads=8) for rx, ry in res]
    batch = S.Batch(views)
    try:
        for f in range(6):
            ps = [view_params(pr, v, f) for v in range(len(res))]
            print(f"interleaved frame {f}", flush=True)
            if f % 3 == 2:
                views[0].frame_warp(ps[0])  # alone on its own stream
                views[1].frame_warp(ps[1])
            else:
                batch.frame_warp(ps)
            for v in range(len(res)):
                oracles[v].frame_warp(O.params_from(ps[v]))
                views[v].sync()
                assert_equal_u32(views[v].screen(), oracles[v].screen, f"frame {f} view {v} screen")
                assert_equal_u32(views[v].image(), oracles[v].image, f"frame {f} view {v} image")
    finally:
        batch.close()
        for c in views:
            c.close()
        for o in oracles:
            o.close()


def test_batch_rejects_mismatched_octrees():
    a_words, a_root = scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])
    b_words, b_root = scene_octree(1, 7, 9)
    a = S.Context(0)
    b = S.Context(0)
    try:
        a.upload_octree(a_words, a_root, CONFIG1["depth"])
        a.resize(64, 64)
        b.upload_octree(b_words, b_root, 9)
        b.resize(64, 64)
        with pytest.raises(S.SvoError):
            S.Batch([a, b])
        with pytest.raises(S.SvoError):
            S.Batch([a, a])
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("max_depth", [0.0, 60.0])
def test_batch_learned_matches_single_views(max_depth):
    """Batched learned frames (gate of all views in one launch, gate pixels in
    the shared ray launch, the CNN tiles of all views in one launch) equal
    per-view learned frames, with and without the gate's depth rule."""
    from svo_amd import learned as L
    words, root = scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])
    pr = path(CONFIG1, 32)
    res = RES[:4]
    batched = make_views(words, root, res)
    single = make_views(words, root, res)
    for c in batched + single:
        c.set_gate_depth(max_depth)
    cnn = S.Cnn(0)
    L.load(cnn, L.WEIGHTS)
    batch = S.Batch(batched)
    try:
        for f in range(6):
            ps = [view_params(pr, v, f) for v in range(len(res))]
            print(f"batch learned frame {f}", flush=True)
            batch.frame_warp_learned(cnn, ps, 64)
            for c, p in zip(single, ps):
                c.frame_warp_learned(cnn, p, 64)
            for v, (a, b) in enumerate(zip(batched, single)):
                a.sync()
                b.sync()
                compare(a, b, f"learned frame {f} view {v} {res[v]}")
                assert a.gate_counts() == b.gate_counts()
    finally:
        batch.close()
        for c in batched + single:
            c.close()
        cnn.close()


def test_batch_refuses_a_resized_view():
    words, root = scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])
    views = make_views(words, root, [(128, 64), (64, 64)])
    batch = S.Batch(views)
    try:
        pr = path(CONFIG1, 4)
        batch.frame_warp([view_params(pr, v, 0) for v in range(2)])
        batch.sync()
        views[1].resize(96, 64)
        with pytest.raises(S.SvoError):
            batch.frame_warp([view_params(pr, v, 1) for v in range(2)])
    finally:
        batch.close()
        for c in views:
            c.close()

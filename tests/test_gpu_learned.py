"""Learned fill on the GPU: the gate path against the oracle (bit-exact when
every tile with holes is sent back to raycasting), and training of the BF16
MFMA CNN on the pipeline's own (warp, full raycast) pairs: the loss falls and
the trained network beats the method without it on held-out frames."""
import numpy as np
import pytest

import oracle as O
import svo_amd as S
from svo_amd import learned as L
from svo_amd import workloads as W
from helpers import assert_equal_f32_bits, assert_equal_u32

pytestmark = pytest.mark.gpu
NAME = "flythrough_256"


@pytest.mark.parametrize("max_holes,max_depth", [(0, 0.0), (0, 60.0), (64, 60.0)])
def test_gate_all_raycast_bitexact(max_holes, max_depth):
    """The gate's ray lists against the oracle's gated frame: every residual
    hole raycast (max_holes 0), with and without the depth rule; and the
    depth rule with CNN tiles left unfilled (no network: fillhole2 colours)."""
    words, root, depth, _ = W.octree(NAME)
    rx, ry = W.WORKLOADS[NAME]["res"]
    rows = W.path(NAME, 10)
    o = O.Frame(rx, ry, words, root, depth, nthreads=8)
    with S.Context(0) as c:
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
        c.set_gate_depth(max_depth)
        for f in range(10):
            p = W.frame_params(NAME, rows, f)
            c.frame_warp_learned(None, p, max_holes)
            o.frame_warp_gated(O.params_from(p), max_holes, max_depth)
            c.sync()
            assert_equal_u32(c.screen(), o.screen, f"frame {f} screen")
            assert_equal_f32_bits(c.coords(), o.coords, f"frame {f} coords")
            assert_equal_u32(c.image(), o.image, f"frame {f} image")
            if max_holes == 0:
                assert not (c.screen(2) == 0xFFFFFF00).any()
    o.close()


def test_training_improves_fill():
    cnn = S.Cnn(0, seed=3)
    before_all, before_holes = L.evaluate(cnn, NAME, n_frames=24)
    losses = L.train(cnn, [(NAME, 0, 64), (NAME, 1, 64), (NAME, 2, 64), (NAME, 3, 64)], log_every=16,
                     log=print)
    got = [l for l in losses if l is not None]
    assert np.mean(got[-3:]) < 0.5 * np.mean(got[:2]), got
    after_all, after_holes = L.evaluate(cnn, NAME, n_frames=24)
    print("before", before_all, before_holes, "after", after_all, after_holes)
    # the trained CNN beats fillhole2 on the pixels it fills, and overall
    assert after_holes["learned"] > after_holes["gate"] + 0.5
    assert after_all["learned"] > after_all["gate"]
    assert after_all["gate"] >= after_all["plain"]
    cnn.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_fused_tile_fill_matches_per_layer_fill(mode):
    """The frame fill with the fused tile kernel and with the per-layer
    kernels: identical screens, images equal outside CNN pixels and within
    bf16 rounding (a few 8-bit levels) on them."""
    cnn = S.Cnn(0)
    L.load(cnn, L.WEIGHTS)
    words, root, depth, _ = W.octree(NAME)
    rx, ry = W.WORKLOADS[NAME]["res"]
    rows = W.path(NAME, 12)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, depth)
            c.resize(rx, ry)
        n_cnn = 0
        for f in range(12):
            p = W.frame_params(NAME, rows, f)
            cnn.set_fused(mode=mode)
            a.frame_warp_learned(cnn, p, 64)
            a.sync()
            cnn.set_fused(False)
            b.frame_warp_learned(cnn, p, 64)
            b.sync()
            assert_equal_u32(a.screen(), b.screen(), f"frame {f} screen")
            holes = (a.screen(2) == 0xFFFFFF00).reshape(-1)
            ia, ib = a.image().reshape(-1), b.image().reshape(-1)
            assert np.array_equal(ia[~holes], ib[~holes]), f"frame {f}: non-CNN pixels differ"
            ca = np.stack([(ia[holes] >> s) & 255 for s in (0, 8, 16)], 1).astype(int)
            cb = np.stack([(ib[holes] >> s) & 255 for s in (0, 8, 16)], 1).astype(int)
            if holes.any():
                d = np.abs(ca - cb)
                assert d.max() <= 4 and d.mean() < 0.5, (f, d.max(), d.mean())
            n_cnn += int(holes.sum())
        assert n_cnn > 0
    finally:
        a.close()
        b.close()
        cnn.close()


@pytest.mark.parametrize("overlap_wgs", [None, 0, 1, 7])
def test_cnn_alongside_rays_equals_sequential(overlap_wgs):
    """The CNN on the second stream alongside the ray launch (with the
    default, the full or a small workgroup count) and the CNN run after the
    fill on the frame's stream (per-stage timing) give identical screens and
    images."""
    cnn = S.Cnn(0)
    L.load(cnn, L.WEIGHTS)
    words, root, depth, _ = W.octree(NAME)
    rx, ry = W.WORKLOADS[NAME]["res"]
    rows = W.path(NAME, 12)
    a, b = S.Context(0), S.Context(0)
    try:
        for c in (a, b):
            c.upload_octree(words, root, depth)
            c.resize(rx, ry)
        if overlap_wgs is not None:
            a.set_cnn_overlap(overlap_wgs)
        b.enable_timing(True, mode=1)
        n_cnn = 0
        for f in range(12):
            p = W.frame_params(NAME, rows, f)
            a.frame_warp_learned(cnn, p, 64)
            b.frame_warp_learned(cnn, p, 64)
            a.sync()
            b.sync()
            assert_equal_u32(a.screen(), b.screen(), f"frame {f} screen")
            assert_equal_u32(a.image(), b.image(), f"frame {f} image")
            n_cnn += a.gate_counts()[0]
        assert n_cnn > 0
    finally:
        a.close()
        b.close()
        cnn.close()


def test_train_step_needs_a_learned_frame():
    cnn = S.Cnn(0)
    words, root, depth, _ = W.octree(NAME)
    rx, ry = W.WORKLOADS[NAME]["res"]
    p = W.frame_params(NAME, W.path(NAME, 2), 0)
    warp, full = S.Context(0), S.Context(0)
    try:
        for c in (warp, full):
            c.upload_octree(words, root, depth)
            c.resize(rx, ry)
        warp.frame_warp(p)
        full.frame_full(p)
        with pytest.raises(S.SvoError):
            cnn.train_step(warp, full, p, 4, 32)
        warp.frame_warp_learned(None, p, 64)
        loss = cnn.train_step(warp, full, p, 4, 32, want_loss=True)
        assert np.isfinite(loss)
    finally:
        warp.close()
        full.close()
        cnn.close()

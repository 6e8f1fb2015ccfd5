"""Every benchmark configuration through the GPU pipeline against the oracle.

  config 2  flythrough_2048 (1920x1024, depth 11): the whole 300-frame path,
            plane 0, the id list and the image every frame, all four planes
            and coordinates every 8th frame (the bench times frames 5..24)
  config 3  cave_4096 (1920x1024, depth 12, fast camera): the first 10 frames
            (2 clears + 8 warp frames), every buffer every frame
  config 4  learned fill at 1920x1024 on the depth-11 flythrough: the gate
            with every residual hole raycast is the oracle's gated frame word
            for word, with and without the CNN; the committed network beats
            fillhole2 on the hole pixels of a held-out path
  config 5  24 batched views of mixed resolution on the depth-11 scene, plain
            and learned (gate all-raycast): each view against its own oracle
            frame, 2 clears + 2 warp frames per view

The octree of the large scenes comes from the product builder (depth 11) or
the GPU builder (depth 12); both are checked word for word elsewhere.  The
scene-builder arithmetic of the reference (octree.h:74-76) wraps its 24-bit
node handles above 2^24 pointer nodes, which both scenes exceed, so the
oracle renders the same word array the GPU renders.
"""
import numpy as np
import pytest

import oracle as O
import svo_amd as S
from svo_amd import learned as L
from svo_amd import workloads as W
from helpers import assert_equal_f32_bits, assert_equal_u32

pytestmark = pytest.mark.gpu
FLY = "flythrough_2048"
CAVE = "cave_4096"


def _ids_prefix(ids, S_):
    """counts, offsets and the hole list (ids[0] = its length after sumids)"""
    return ids[:2 * S_ + int(ids[0])]


def _compare_frame(c, o, what, full):
    if full:
        assert_equal_u32(c.screen(), o.screen, f"{what} screen")
        assert_equal_f32_bits(c.coords(), o.coords, f"{what} coords")
    else:
        assert_equal_u32(c.screen(0), o.screen[0], f"{what} plane 0")
    S_ = (o.res_x // 16) * (o.res_y // 16)
    assert_equal_u32(_ids_prefix(c.ids(), S_), _ids_prefix(o.ids, S_), f"{what} ids")
    assert_equal_u32(c.image(), o.image, f"{what} image")


def test_config2_whole_path_vs_oracle():
    words, root, depth, _ = W.octree(FLY)
    rx, ry = W.WORKLOADS[FLY]["res"]
    n = W.WORKLOADS[FLY]["frames"]
    rows = W.path(FLY, n)
    o = O.Frame(rx, ry, words, root, depth, nthreads=16)
    with S.Context(0) as c:
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
        for f in range(n):
            p = W.frame_params(FLY, rows, f)
            if f % 50 == 0:
                print(f"config 2 frame {f}", flush=True)
            c.frame_warp(p)
            o.frame_warp(O.params_from(p))
            c.sync()
            _compare_frame(c, o, f"config 2 frame {f}", full=(f % 8 == 0 or f == n - 1))
    o.close()


def test_config3_cave_vs_oracle():
    words, root, depth, _ = W.octree(CAVE)
    rx, ry = W.WORKLOADS[CAVE]["res"]
    rows = W.path(CAVE, 10)
    o = O.Frame(rx, ry, words, root, depth, nthreads=16)
    holes = []
    with S.Context(0) as c:
        c.upload_octree(words, root, depth)
        c.resize(rx, ry)
        for f in range(10):
            p = W.frame_params(CAVE, rows, f)
            print(f"config 3 frame {f}", flush=True)
            c.frame_warp(p)
            o.frame_warp(O.params_from(p))
            c.sync()
            _compare_frame(c, o, f"config 3 frame {f}", full=True)
            if f >= 2:
                holes.append(int(o.ids[0]) / (rx * ry))
    o.close()
    print("config 3 hole-list fractions", [round(h, 3) for h in holes])
    assert max(holes) > 0.01  # the fast camera opens holes from the first warp frames on


@pytest.mark.parametrize("with_cnn,max_depth", [(False, 0.0), (True, 0.0), (False, 1500.0)])
def test_config4_gate_all_raycast_vs_oracle(with_cnn, max_depth):
    """max_cnn_holes = 0: every tile with a residual hole is raycast in the
    frame's ray launch, so the frame equals the oracle's gated frame; with the
    CNN passed in, the gate gives it no tile and nothing changes; with the
    depth rule, tiles holding a reprojected pixel deeper than max_depth world
    units are raycast whole, as in the oracle."""
    words, root, depth, _ = W.octree(FLY)
    rx, ry = W.WORKLOADS[FLY]["res"]
    rows = W.path(FLY, 12)
    cnn = None
    if with_cnn:
        cnn = S.Cnn(0)
        L.load(cnn, L.WEIGHTS)
    o = O.Frame(rx, ry, words, root, depth, nthreads=16)
    try:
        with S.Context(0) as c:
            c.upload_octree(words, root, depth)
            c.resize(rx, ry)
            c.set_gate_depth(max_depth)
            for f in range(12):
                p = W.frame_params(FLY, rows, f)
                print(f"config 4 frame {f} (cnn {with_cnn}, depth rule {max_depth})", flush=True)
                c.frame_warp_learned(cnn, p, 0)
                o.frame_warp_gated(O.params_from(p), 0, max_depth)
                c.sync()
                assert_equal_u32(c.screen(), o.screen, f"config 4 frame {f} screen")
                assert_equal_f32_bits(c.coords(), o.coords, f"config 4 frame {f} coords")
                assert_equal_u32(c.image(), o.image, f"config 4 frame {f} image")
                assert c.gate_counts()[0] == 0
                assert not (c.screen(2) == S.HOLE).any()
    finally:
        o.close()
        if cnn is not None:
            cnn.close()


def test_config4_trained_net_beats_fillhole2_on_holes():
    """The committed network against the method without it at 1920x1024 on
    the depth-11 flythrough (held-out path, gate threshold 64): better on
    the pixels still holes after the rays, and better overall."""
    cnn = S.Cnn(0)
    L.load(cnn, L.WEIGHTS)
    allp, holes = L.evaluate(cnn, FLY, seed_off=1000, n_frames=24, skip=8, max_cnn_holes=64)
    cnn.close()
    print("config 4 PSNR all pixels", allp, "hole pixels", holes)
    assert holes["learned"] > holes["gate"] + 1.0
    assert holes["learned"] > holes["plain"] + 1.0
    assert allp["learned"] > allp["gate"] >= allp["plain"] - 0.05


@pytest.fixture(scope="module")
def views24():
    views = W.batch_views()
    words, root, depth, _ = W.octree(views[0]["scene"])
    paths = [W.view_path(v) for v in views]
    return views, words, root, depth, paths


@pytest.mark.parametrize("learned", [False, True])
def test_config5_batch_vs_oracle(views24, learned):
    """The 24-view batch (resolutions drawn from 320x240 .. 1920x1024) in one
    launch per stage and one shared ray launch; each view's buffers equal its
    own oracle frame (plain warp, or the gate with every residual hole
    raycast)."""
    views, words, root, depth, paths = views24
    ctxs, orcs = [], []
    batch = None
    try:
        for v in views:
            c = S.Context(0)
            c.upload_octree(words, root, depth)
            c.resize(*v["res"])
            ctxs.append(c)
            orcs.append(O.Frame(*v["res"], words, root, depth, nthreads=16))
        batch = S.Batch(ctxs)
        for f in range(4):
            ps = [W.frame_params(v["scene"], paths[i], f) for i, v in enumerate(views)]
            print(f"config 5 batch frame {f} (learned {learned})", flush=True)
            if learned:
                batch.frame_warp_learned(None, ps, 0)
            else:
                batch.frame_warp(ps)
            batch.sync()
            for i, (c, o, p) in enumerate(zip(ctxs, orcs, ps)):
                if learned:
                    o.frame_warp_gated(O.params_from(p), 0)
                else:
                    o.frame_warp(O.params_from(p))
                what = f"config 5 frame {f} view {i} {views[i]['res']}"
                assert_equal_u32(c.screen(), o.screen, f"{what} screen")
                assert_equal_f32_bits(c.coords(), o.coords, f"{what} coords")
                S_ = (o.res_x // 16) * (o.res_y // 16)
                assert_equal_u32(_ids_prefix(c.ids(), S_), _ids_prefix(o.ids, S_), f"{what} ids")
                assert_equal_u32(c.image(), o.image, f"{what} image")
    finally:
        if batch is not None:
            batch.close()
        for c in ctxs:
            c.close()
        for o in orcs:
            o.close()

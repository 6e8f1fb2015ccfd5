"""GPU parity: the HIP path through the C ABI vs the CPU oracle, bit-exact
on integer words and on the FP32 coordinate bits (identical op order)."""
import numpy as np
import pytest

import oracle as O
import svo_amd as S
from helpers import CONFIG1, assert_equal_f32_bits, assert_equal_u32, params, path, scene_octree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene1():
    return scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def test_full_frame_config1_bitexact(ctx, scene1):
    words, root = scene1
    rx, ry = CONFIG1["res"]
    ctx.upload_octree(words, root, 8)
    ctx.resize(rx, ry)
    pr = path(CONFIG1, 4)
    for f in (0, 3):
        p = params(CONFIG1, pr, f)
        print(f"full frame {f}", flush=True)
        ctx.frame_full(p)
        ctx.sync()
        o = O.Frame(rx, ry, words, root, 8, nthreads=8)
        o.frame_full(O.params_from(p))
        assert_equal_u32(ctx.screen(0), o.screen[0], "screen plane 0")
        assert_equal_f32_bits(ctx.coords(0)[..., :3], o.coords[0][..., :3], "coords xyz")
        assert_equal_u32(ctx.image(), o.image, "image")
        o.close()


def test_warp_sequence_config1_bitexact(ctx, scene1):
    words, root = scene1
    rx, ry = CONFIG1["res"]
    ctx.upload_octree(words, root, 8)
    ctx.resize(rx, ry)
    o = O.Frame(rx, ry, words, root, 8, nthreads=8)
    pr = path(CONFIG1, 12)
    for f in range(12):
        p = params(CONFIG1, pr, f)
        print(f"warp frame {f}", flush=True)
        ctx.frame_warp(p)
        ctx.sync()
        o.frame_warp(O.params_from(p))
        assert_equal_u32(ctx.screen(), o.screen, f"frame {f} screen planes")
        assert_equal_f32_bits(ctx.coords(), o.coords, f"frame {f} coords planes")
        assert_equal_u32(ctx.ids(), o.ids, f"frame {f} id buffer")
        assert_equal_u32(ctx.image(), o.image, f"frame {f} image")
    o.close()


def test_warp_graph_vs_direct_launches(scene1):
    """Frames >= 2 replay one captured graph; the same frames as direct
    launches must give identical buffers, also across a re-capture (tuning
    change), a resize and a timing mode switch."""
    words, root = scene1
    rx, ry = CONFIG1["res"]
    g = S.Context(0)
    d = S.Context(0)
    try:
        for c in (g, d):
            c.upload_octree(words, root, 8)
            c.resize(rx, ry)
        d.set_graph(False)
        pr = path(CONFIG1, 20)
        for f in range(20):
            if f == 7:
                g.set_tuning(refill=8, iters=16)  # forces a new capture
                d.set_tuning(refill=8, iters=16)
            if f == 11:
                g.enable_timing(mode=2)
            if f == 14:
                g.enable_timing(False)
            p = params(CONFIG1, pr, f)
            print(f"graph frame {f}", flush=True)
            g.frame_warp(p)
            d.frame_warp(p)
            g.sync()
            d.sync()
            if f == 12:
                assert g.frame_time() > 0.0
            assert_equal_u32(g.screen(), d.screen(), f"frame {f} screen")
            assert_equal_f32_bits(g.coords(), d.coords(), f"frame {f} coords")
            assert_equal_u32(g.ids(), d.ids(), f"frame {f} ids")
            assert_equal_u32(g.image(), d.image(), f"frame {f} image")
        # a new resolution re-captures; restart the sequence on both
        for c in (g, d):
            c.resize(200, 136)
        for f in range(6):
            p = params(CONFIG1, pr, f)
            g.frame_warp(p)
            d.frame_warp(p)
            g.sync()
            d.sync()
            assert_equal_u32(g.screen(), d.screen(), f"resized frame {f} screen")
            assert_equal_u32(g.image(), d.image(), f"resized frame {f} image")
    finally:
        g.close()
        d.close()


@pytest.mark.parametrize("spread,refill,wg,iters,cache,lanes,block", [(64, 16, 4, 64, 0, 64, 16),
                                                                      (1, 48, 8, 8, 736, 32, 64),
                                                                      (4, 1, 1, 1, 4096, 16, 32),
                                                                      (16, 64, 32, 4096, 96, 8, 8),
                                                                      (1, 8, 2, 16, 0, 1, 64)])
def test_launch_knobs_do_not_change_results(scene1, spread, refill, wg, iters, cache, lanes, block):
    """Ray hand-out order, refill threshold, workgroups per CU, trips per
    refill check, the LDS node cache, lanes per wave and the ray block shape
    change only the schedule."""
    words, root = scene1
    rx, ry = CONFIG1["res"]
    pr = path(CONFIG1, 6)
    c = S.Context(0)
    o = O.Frame(rx, ry, words, root, 8, nthreads=8)
    try:
        c.upload_octree(words, root, 8)
        c.resize(rx, ry)
        c.set_tuning(refill, wg, iters, cache)
        c.set_ray_spread(spread, only_if_fits=(spread != 16))
        c.set_ray_lanes(lanes)
        c.set_ray_block(block)
        for f in range(6):
            p = params(CONFIG1, pr, f)
            print(f"knobs {spread, refill, wg, iters, cache, lanes, block} frame {f}", flush=True)
            c.frame_warp(p)
            o.frame_warp(O.params_from(p))
            c.sync()
            assert_equal_u32(c.screen(), o.screen, f"frame {f} screen")
            assert_equal_f32_bits(c.coords(), o.coords, f"frame {f} coords")
        p = params(CONFIG1, pr, 5)
        c.frame_full(p)
        o.frame_full(O.params_from(p))
        c.sync()
        assert_equal_u32(c.screen(0), o.screen[0], "full frame")
    finally:
        c.close()
        o.close()


@pytest.mark.parametrize("long_trips,block", [(1, 8), (16, 8), (512, 8), (16, 32), (1, 64)])
def test_full_frame_long_first_bitexact(scene1, long_trips, block):
    """Full frames that raycast the previous frame's long-ray blocks first
    (1: every block, so the front list is the whole screen): identical to the
    oracle, with the setting switched off for one frame, across a warp frame
    in between and a resize to a size that is not a multiple of 8."""
    words, root = scene1
    pr = path(CONFIG1, 6)
    c = S.Context(0)
    try:
        c.upload_octree(words, root, 8)
        c.set_ray_block(block)
        for rx, ry in ((CONFIG1["res"][0], CONFIG1["res"][1]), (203, 133)):
            c.resize(rx, ry)
            o = O.Frame(rx, ry, words, root, 8, nthreads=8)
            for f in range(6):
                p = params(CONFIG1, pr, f)
                c.set_long_first(0 if f == 4 else long_trips)
                print(f"long first {long_trips} block {block} {rx}x{ry} frame {f}", flush=True)
                if f == 2 and rx % 8 == 0:
                    c.frame_warp(p)  # warp state in between (sizes the warp tests cover)
                c.frame_full(p)
                c.sync()
                o.frame_full(O.params_from(p))
                assert_equal_u32(c.screen(0), o.screen[0], f"{rx}x{ry} frame {f} screen")
                assert_equal_f32_bits(c.coords(0)[..., :3], o.coords[0][..., :3], f"{rx}x{ry} frame {f} coords")
                assert_equal_u32(c.image(), o.image, f"{rx}x{ry} frame {f} image")
            o.close()
    finally:
        c.close()

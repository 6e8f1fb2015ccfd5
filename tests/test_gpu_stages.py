"""GPU parity of the per-stage entry points (reference kernel names) and of
edge cases, against the CPU oracle on the same seeded inputs."""
import numpy as np
import pytest

import oracle as O
import svo_amd as S
from helpers import CONFIG1, assert_equal_f32_bits, assert_equal_u32, params, path, scene_octree

pytestmark = pytest.mark.gpu
HOLE = 0xFFFFFF00


@pytest.fixture(scope="module")
def scene1():
    return scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def _compare_all(c, o, what):
    assert_equal_u32(c.screen(), o.screen, f"{what}: screen")
    assert_equal_f32_bits(c.coords(), o.coords, f"{what}: coords")
    assert_equal_u32(c.ids(), o.ids, f"{what}: ids")


def test_stagewise_frames(ctx, scene1):
    """raycast.h:147-438 issued stage by stage through the C ABI."""
    words, root = scene1
    rx, ry = 256, 128
    ctx.upload_octree(words, root, 8)
    ctx.resize(rx, ry)
    o = O.Frame(rx, ry, words, root, 8, nthreads=8)
    N = rx * ry
    pr = path(CONFIG1, 6)
    for f in range(6):
        p = params(CONFIG1, pr, f)
        op = O.params_from(p)
        print(f"stagewise frame {f}", flush=True)
        if f < 2:
            ctx.memset(0, HOLE, 16 * N)
            o.memset(0, HOLE, 4 * N)
        ctx.memset(0, HOLE, 4 * N)
        o.memset(0, HOLE, N)
        for plane in (1, 2):
            ctx.proj(plane, p)
            o.proj(plane, op)
            ctx.sync()
            _compare_all(ctx, o, f"frame {f} proj {plane}")
        ctx.counthole(); o.counthole()
        ctx.sumids(); o.sumids()
        ctx.sync()
        assert ctx.hole_count() == o.ids[0]
        ctx.writeids(); o.writeids()
        ctx.sync()
        _compare_all(ctx, o, f"frame {f} compaction")
        ctx.raycast_holes(p); o.raycast_holes(op)
        ax, ay, ex, ey = S.tile_schedule(rx, ry, f)
        assert (ax, ay, ex, ey) == O.tile_schedule(rx, ry, f)
        ctx.raycast_rect(p, ax, ay, ex, ey); o.raycast_rect(op, ax, ay, ex, ey)
        ctx.sync()
        _compare_all(ctx, o, f"frame {f} rays")
        ctx.copy_cache(2); o.copy_plane(2)
        ctx.fillhole(); o.fillhole_snapshot()
        ctx.colorize(); o.colorize()
        ctx.sync()
        _compare_all(ctx, o, f"frame {f} filter")
        assert_equal_u32(ctx.image(), o.image, f"frame {f} image")
    o.close()


def test_proj_against_populated_plane0(ctx, scene1):
    """proj without the per-frame clear: sources must beat the existing
    plane-0 depth key (kernel.cl:571) to win."""
    words, root = scene1
    rx, ry = 128, 96
    ctx.upload_octree(words, root, 8)
    ctx.resize(rx, ry)
    o = O.Frame(rx, ry, words, root, 8)
    pr = path(CONFIG1, 4)
    p0 = params(CONFIG1, pr, 0)
    ctx.frame_full(p0)
    o.frame_full(O.params_from(p0))
    ctx.copy_cache(2); o.copy_plane(2)
    rng = np.random.default_rng(5)
    plane0 = rng.integers(0, 2**32, size=(ry, rx), dtype=np.uint64).astype(np.uint32)
    plane0[rng.random((ry, rx)) < 0.3] = HOLE
    ctx.write(S.BUF_SCREEN, 0, plane0)
    o.screen[0][:] = plane0
    p2 = params(CONFIG1, pr, 3)
    ctx.proj(2, p2)
    o.proj(2, O.params_from(p2))
    ctx.sync()
    _compare_all(ctx, o, "populated proj")
    o.close()


@pytest.mark.parametrize("res", [(200, 120), (48, 40), (333, 77)])
def test_ragged_resolution_warp(ctx, scene1, res):
    """Resolutions that are not multiples of 16 (partial blocks are never
    compacted, kernel.cl:243-244)."""
    words, root = scene1
    rx, ry = res
    ctx.upload_octree(words, root, 8)
    ctx.resize(rx, ry)
    o = O.Frame(rx, ry, words, root, 8, nthreads=8)
    pr = path(CONFIG1, 5)
    for f in range(5):
        p = params(CONFIG1, pr, f)
        ctx.frame_warp(p)
        o.frame_warp(O.params_from(p))
        ctx.sync()
        _compare_all(ctx, o, f"{res} frame {f}")
        assert_equal_u32(ctx.image(), o.image, f"{res} frame {f} image")
    o.close()


def test_empty_scene(ctx):
    """No voxels: every ray misses and returns the root colour word."""
    words, root = S.build_octree(np.zeros((0, 3), np.uint16), np.zeros(0, np.uint32), 8)
    ctx.upload_octree(words, root, 8)
    ctx.resize(64, 32)
    p, _ = S.camera(0.3, 0.2, 0.0, (3.0, 10.0, 4.0), 8, 0)
    ctx.frame_full(p)
    ctx.sync()
    o = O.Frame(64, 32, words, root, 8)
    o.frame_full(O.params_from(p))
    assert_equal_u32(ctx.screen(0), o.screen[0], "empty scene screen")
    assert np.all(ctx.screen(0) == 0xFF000000)
    o.close()


def test_depth_variants(ctx):
    """Depths 9 and 12 (block boundaries move with D) on small random scenes."""
    rng = np.random.default_rng(11)
    for D in (9, 12):
        n = 40000
        size = 1 << D
        xyz = np.stack([rng.integers(0, size, n), rng.integers(0, size // 3, n),
                        rng.integers(0, size, n)], 1).astype(np.uint16)
        rgba = (rng.integers(1, 255, n) | (rng.integers(0, 2**24, n) << 8)).astype(np.uint32)
        words, root = S.build_octree(xyz, rgba, D)
        ctx.upload_octree(words, root, D)
        ctx.resize(96, 64)
        o = O.Frame(96, 64, words, root, D, nthreads=8)
        for f in range(3):
            p, _ = S.camera(0.4 + 0.01 * f, 0.3, 0.0, (size / 16 * 0.3, size / 16 * 0.5, size / 16 * 0.2), D, f)
            ctx.frame_warp(p)
            o.frame_warp(O.params_from(p))
            ctx.sync()
            _compare_all(ctx, o, f"D={D} frame {f}")
        o.close()


def test_error_paths(ctx):
    with pytest.raises(S.SvoError):
        ctx.resize(8, 8)
    with pytest.raises(S.SvoError):
        ctx.upload_octree(np.zeros(5, np.uint32), 0, 8)
    bad = np.zeros(2097152 + 64, np.uint32)
    bad[0] = (5 << 9) | 256 | 1  # a block ref directly under the root (blocks start at depth D-6)
    with pytest.raises(S.SvoError):
        ctx.upload_octree(bad, 0x01, 8)
    with pytest.raises(S.SvoError):
        ctx.proj(0, S.FrameParams())
    with pytest.raises(S.SvoError):
        ctx.set_tuning(refill=65)


@pytest.mark.parametrize("res,add,frame", [((96, 64), (0, 0), 5), ((98, 66), (49, 33), 12), ((200, 136), (100, 0), 8)])
def test_quad_raycast_vs_oracle(ctx, scene1, res, add, frame):
    """svo_raycast_quads (raycast_fine, kernel.cl:696-843) on a random
    hole/non-hole screen: all four planes and coordinates bit-exact against
    the oracle, odd offsets included (picks past the row or frame)."""
    words, root = scene1
    rx, ry = res
    ctx.upload_octree(words, root, 8)
    ctx.resize(rx, ry)
    p, _ = S.camera(0.7, 0.3, 0.0, (4.0, 30.0, 9.0), 8, frame)
    rng = np.random.default_rng(rx + frame)
    before = np.where(rng.random(4 * rx * ry) < 0.4, 0xFFFFFF00, 0x12345600).astype(np.uint32)
    ctx.write(S.BUF_SCREEN, 0, before)
    o = O.Frame(rx, ry, words, root, 8, nthreads=8)
    o.screen.reshape(-1)[:] = before
    print(f"quads {res} add {add}", flush=True)
    ctx.raycast_quads(p, *add)
    ctx.sync()
    o.raycast_quads(O.params_from(p), *add)
    assert_equal_u32(ctx.screen(), o.screen, "screen planes")
    assert_equal_f32_bits(ctx.coords()[..., :3], o.coords[..., :3], "coords")
    o.close()


@pytest.mark.parametrize("res", [(64, 48), (203, 133)])
def test_fillhole_punch_vs_oracle(ctx, scene1, res):
    """raycast_fillhole (kernel.cl:342-401) on a warped plane 0 with random
    motion buffers (about half the pixels move): bit-exact with the oracle,
    and zero motion changes nothing."""
    words, root = scene1
    rx, ry = res
    ctx.upload_octree(words, root, 8)
    ctx.resize(rx, ry)
    o = O.Frame(rx, ry, words, root, 8, nthreads=8)
    pr = path(CONFIG1, 4)
    for f in range(3):
        p = params(CONFIG1, pr, f)
        ctx.frame_warp(p)
        o.frame_warp(O.params_from(p))
    ctx.sync()
    _compare_all(ctx, o, f"{res} before punch")
    before = ctx.screen(0).copy()
    ctx.fillhole_punch(None, None)
    ctx.sync()
    assert_equal_u32(ctx.screen(0), before, "zero motion")
    rng = np.random.default_rng(rx)
    mx = (rng.integers(0, 2, (ry, rx)) * rng.integers(-3, 4, (ry, rx))).astype(np.int32)
    my = (rng.integers(0, 2, (ry, rx)) * rng.integers(-3, 4, (ry, rx))).astype(np.int32)
    ctx.fillhole_punch(mx, my)
    o.fillhole_punch(mx, my)
    ctx.sync()
    assert_equal_u32(ctx.screen(), o.screen, f"{res} punched screen")
    punched = int(((ctx.screen(0) == 0xFFFFFF00) & (before != 0xFFFFFF00)).sum())
    print(f"{res}: {punched} pixels punched", flush=True)
    assert punched > 0
    o.close()


def _np_compaction_vec(scr, rx, ry):
    """counthole / sumids / writeids by array operations (kernel.cl:234-340):
    4 ids per all-hole 2x2 quad, blocks of 16x16 in raster order, quads in
    row-major order inside a block."""
    bx, by = rx // 16, ry // 16
    s = scr[:by * 16, :bx * 16] == 0xFFFFFF00
    q = s[0::2, 0::2] & s[0::2, 1::2] & s[1::2, 0::2] & s[1::2, 1::2]
    qb = q.reshape(by, 8, bx, 8).transpose(0, 2, 1, 3).reshape(by * bx, 8, 8)
    counts = (4 * qb.reshape(by * bx, 64).sum(1)).astype(np.uint32)
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))[:-1]]).astype(np.uint32)
    b, j, i = np.nonzero(qb)
    v = ((b % bx) * 16 + 2 * i) | (((b // bx) * 16 + 2 * j) << 16)
    ids = np.stack([v, v + 1, v + 1 + (1 << 16), v + (1 << 16)], 1).reshape(-1).astype(np.uint32)
    return counts, offs, int(counts.sum(dtype=np.int64)), ids


@pytest.mark.parametrize("res", [(203, 133), (2048, 1088), (8192, 8192)])
def test_compaction_stages_vs_array_formulation(ctx, scene1, res):
    """counthole, sumids and writeids on a random hole mask: block counts,
    exclusive offsets (several scan rounds of 8192 blocks at the larger sizes),
    total and the id list equal an independent array formulation."""
    words, root = scene1
    rx, ry = res
    ctx.upload_octree(words, root, 8)
    ctx.resize(rx, ry)
    rng = np.random.default_rng(rx + ry)
    scr = np.where(rng.random((ry, rx)) < 0.9, 0xFFFFFF00, rng.integers(0, 2**31, (ry, rx))).astype(np.uint32)
    ctx.write(S.BUF_SCREEN, 0, scr)
    print(f"compaction {rx}x{ry}", flush=True)
    ctx.counthole()
    ctx.sync()
    Sn = (rx // 16) * (ry // 16)
    counts, offs, H, ids = _np_compaction_vec(scr, rx, ry)
    got = ctx.ids()
    assert np.array_equal(got[:Sn], counts)
    ctx.sumids()
    ctx.writeids()
    ctx.sync()
    got = ctx.ids()
    assert int(got[0]) == H
    assert np.array_equal(got[Sn + 1:2 * Sn], offs[1:])  # ids[S] = 0 as well; ids[0] holds the total
    assert int(got[Sn]) == 0
    assert np.array_equal(got[2 * Sn:2 * Sn + H], ids)

"""Pins of the CPU oracle's per-frame kernels by closed forms, exhaustive
enumeration, brute force and invariants (none of these re-types the oracle's
own code)."""
import numpy as np
import pytest

import oracle as O
import svo_amd as S

HOLE = 0xFFFFFF00


def test_bitcount_exhaustive():
    """kernel.cl:23-29: SWAR popcount of the low byte."""
    for b in range(1 << 12):
        assert O.bitcount8(b) == bin(b & 0xFF).count("1")


def test_colorize_exhaustive():
    """kernel.cl:959-971: tint table x float(a & 0xf8), truncated, clamped."""
    tab = np.array([[1.0, 1.0, 1.0], [1.0, 0.7, 0.3], [1.5, 0.8, 0.1], [0.2, 0.8, 0.2]], np.float32)
    rng = np.random.default_rng(0)
    for a in list(range(256)) + list(rng.integers(0, 2**32, 500)):
        a = int(a)
        i = np.float32(a & 0xF8)
        r, g, b = [min(255, int(i * tab[a & 3][k])) for k in range(3)]
        assert O.colorize_word(a) == b + (g << 8) + (r << 16)


@pytest.mark.parametrize("res", [(512, 256), (1920, 1024), (320, 240), (1280, 720), (17, 33)])
def test_tile_schedule_closed_form(res):
    """raycast.h:363-366 + ocl.h:188-198: the 8x4 tiles cover the screen in 32
    frames; extents are res/8, res/4 rounded up to 16."""
    rx, ry = res
    seen = np.zeros((ry, rx), bool)
    for f in range(32):
        ax, ay, ex, ey = O.tile_schedule(rx, ry, f)
        assert (ax, ay, ex, ey) == S.tile_schedule(rx, ry, f)
        assert ax == (rx // 8) * (f % 8) and ay == (ry // 4) * ((f // 8) % 4)
        assert ex % 16 == 0 and ex - 16 < rx // 8 <= ex
        assert ey % 16 == 0 and ey - 16 < ry // 4 <= ey
        seen[ay:ay + ey, ax:ax + ex] = True
    assert seen[:ry // 4 * 4, :rx // 8 * 8].all()
    ax, ay, _, _ = O.tile_schedule(rx, ry, 32)
    assert (ax, ay) == (0, 0)


def _np_compaction(screen, rx, ry):
    """Independent formulation: quad hole mask by array slicing."""
    bx, by = rx // 16, ry // 16
    s = screen[:by * 16, :bx * 16] == HOLE
    q = s[0::2, 0::2] & s[0::2, 1::2] & s[1::2, 0::2] & s[1::2, 1::2]   # [by*8, bx*8]
    qb = q.reshape(by, 8, bx, 8).transpose(0, 2, 1, 3)                  # block, qy, qx
    counts = 4 * qb.reshape(by * bx, 64).sum(1)
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
    ids = []
    for b in range(by * bx):
        y0, x0 = (b // bx) * 16, (b % bx) * 16
        for j in range(8):
            for i in range(8):
                if qb[b // bx, b % bx, j, i]:
                    v = (x0 + 2 * i) | ((y0 + 2 * j) << 16)
                    ids += [v, v + 1, v + 1 + (1 << 16), v + (1 << 16)]
    return counts.astype(np.uint32), offs.astype(np.uint32), int(counts.sum()), np.array(ids, np.uint32)


@pytest.mark.parametrize("res,p", [((64, 48), 0.9), ((96, 64), 0.5), ((200, 120), 0.97), ((40, 40), 1.0)])
def test_compaction_brute_force(res, p):
    rx, ry = res
    rng = np.random.default_rng(rx * ry)
    words = np.zeros(2097152 + 16, np.uint32)
    f = O.Frame(rx, ry, words, 0, 8)
    scr = np.where(rng.random((ry, rx)) < p, HOLE, rng.integers(0, 2**31, (ry, rx))).astype(np.uint32)
    f.screen[0][:] = scr
    f.counthole()
    S_ = f.S
    counts_ref, offs_ref, H, ids_ref = _np_compaction(scr, rx, ry)
    assert np.array_equal(f.ids[:S_], counts_ref)
    f.sumids()
    assert f.ids[0] == H
    assert np.array_equal(f.ids[S_:2 * S_], offs_ref)
    f.writeids()
    assert np.array_equal(f.ids[2 * S_:2 * S_ + H], ids_ref)
    f.close()


def _fill_case(center_nbrs, rx=16, ry=16):
    words = np.zeros(2097152 + 16, np.uint32)
    f = O.Frame(rx, ry, words, 0, 8)
    scr = np.full((ry, rx), HOLE, np.uint32)
    for (dx, dy), v in center_nbrs.items():
        scr[8 + dy, 8 + dx] = v
    f.screen[0][:] = scr
    f.fillhole_snapshot()
    out = int(f.screen[0][8, 8])
    f.close()
    return out


def test_fillhole_closed_form_rules():
    """kernel.cl:428-469 case by case on a single hole at (8,8)."""
    a, b, c, d = 0x12345679, 0xAB, 0x3C, 0xFFFF01F2
    # all four neighbours: (c1&3) + ((sum of &0xfc)>>2 & 0xfc)
    v = _fill_case({(1, 0): a, (-1, 0): b, (0, 1): c, (0, -1): d})
    assert v == (a & 3) + ((((a & 0xFC) + (b & 0xFC) + (c & 0xFC) + (d & 0xFC)) >> 2) & 0xFC)
    # horizontal pair
    v = _fill_case({(1, 0): a, (-1, 0): b, (0, 1): c})
    assert v == (a & 3) + ((((a & 0xFC) + (b & 0xFC)) >> 1) & 0xFC)
    # vertical pair
    v = _fill_case({(0, 1): c, (0, -1): d, (1, 0): a})
    assert v == (c & 3) + ((((c & 0xFC) + (d & 0xFC)) >> 1) & 0xFC)
    # 2x2 order (1,0) before (0,1) before (1,1); the copied word is unmodified
    assert _fill_case({(0, 1): c, (1, 1): d}) == c
    assert _fill_case({(1, 1): d}) == d
    # 5x5, x offset outer, y offset inner: (-2,+2) precedes (-1,-2)
    assert _fill_case({(-1, -2): a, (-2, 2): b}) == b
    assert _fill_case({(2, -2): a}) == a
    # nothing within reach: stays a hole
    assert _fill_case({(3, 0): a}) == HOLE


def test_fillhole_snapshot_vs_inplace_order():
    """Holes (7,8) and (8,8): in the serial in-place order (8,8) averages its
    four neighbours including the freshly filled (7,8); on the snapshot it
    only has the vertical pair."""
    words = np.zeros(2097152 + 16, np.uint32)
    f = O.Frame(16, 16, words, 0, 8)
    scr = np.full((16, 16), 0x44, np.uint32)
    scr[8, 7] = HOLE
    scr[8, 8] = HOLE
    scr[8, 6], scr[9, 7], scr[7, 7] = 0x10, 0x20, 0x30      # left of, below, above (7,8)
    scr[8, 9], scr[9, 8], scr[7, 8] = 0x80, 0x41, 0x61      # right of, below, above (8,8)
    f.screen[0][:] = scr
    f.fillhole_inplace()
    inplace = f.screen[0].copy()
    f.screen[0][:] = scr
    f.fillhole_snapshot()
    snap = f.screen[0].copy()
    f.close()
    F = (0x20 & 3) + (((0x20 + 0x30) >> 1) & 0xFC)
    assert inplace[8, 7] == snap[8, 7] == F
    assert inplace[8, 8] == (0x80 & 3) + (((0x80 + (F & 0xFC) + 0x40 + 0x60) >> 2) & 0xFC)
    assert snap[8, 8] == (0x41 & 3) + (((0x40 + 0x60) >> 1) & 0xFC)


def test_camera_matrix_identities():
    """raycast.h:159-162, :322-326: the proj vectors are the rows of m, the
    ray vectors its columns (transpose), m is a rotation."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        yaw, pitch, roll = rng.uniform(-3, 3, 3)
        pos = rng.uniform(-100, 100, 3)
        p, wr = O.camera(yaw, pitch, roll, pos, 11, 0)
        q, wr2 = S.camera(yaw, pitch, roll, pos, 11, 0)
        assert bytes(p) == bytes(q) and wr == wr2
        rows = np.array([[p.row0.x, p.row0.y, p.row0.z], [p.row1.x, p.row1.y, p.row1.z],
                         [p.row2.x, p.row2.y, p.row2.z]])
        cols = np.array([[p.col0.x, p.col0.y, p.col0.z], [p.col1.x, p.col1.y, p.col1.z],
                         [p.col2.x, p.col2.y, p.col2.z]])
        assert np.array_equal(rows.T, cols)
        assert np.allclose(rows @ rows.T, np.eye(3), atol=1e-6)
        assert all(0 <= v < 2 * 2048 / 16 for v in wr)
        assert np.allclose(np.mod(np.asarray(pos), 256.0), wr, atol=1e-4)


def test_zero_motion_reprojection():
    """Warping a fully raycast frame with the same camera maps every cached
    pixel (except the last row/column, kernel.cl:559-562) onto itself."""
    xyz, rgba = S.scene(1, 1001, 8)
    words, root = S.build_octree(xyz, rgba, 8)
    rx, ry = 160, 96
    p, _ = O.camera(0.7, 0.3, 0.0, (4.0, 30.0, 9.0), 8, 5)
    f = O.Frame(rx, ry, words, root, 8, nthreads=4)
    f.frame_full(p)
    f.copy_plane(2)
    src = f.screen[2].copy()
    xyzs = f.coords[2].copy()
    f.memset(0, HOLE, rx * ry)
    f.proj(2, p)
    dst = f.screen[0]
    q = xyzs[..., :3] - np.array([p.m0.x, p.m0.y, p.m0.z], np.float32)
    z = (q[..., 0] * np.float32(p.row2.x) + q[..., 1] * np.float32(p.row2.y)) + q[..., 2] * np.float32(p.row2.z)
    key = ((z * np.float32(1000)).astype(np.int64).astype(np.uint32) << 8) + (src & 0xFF)
    inner = np.zeros_like(dst, bool)
    inner[:-1, :-1] = True
    ok = (dst == key) | ~inner | (z < np.float32(0.05))
    assert ok.mean() > 0.995, ok.mean()
    same = inner & (dst == key)
    assert np.array_equal(f.coords[0][same][:, :3], xyzs[same][:, :3])
    f.close()


def _key_min_proj(screen, coords, planes, p, rx, ry):
    """Independent formulation of the warp: keep, per target, the minimum of
    (depth key, source order) -- the sequential first-smallest-wins rule."""
    N = rx * ry
    best = {}
    for pl in planes:
        for i in range(N):
            col = int(screen[pl].reshape(-1)[i])
            if col == HOLE:
                continue
            c = coords[pl].reshape(-1, 4)[i]
            q = (c[:3] - np.array([p.m0.x, p.m0.y, p.m0.z], np.float32)).astype(np.float32)
            r = [np.float32((q[0] * np.float32(v[0]) + q[1] * np.float32(v[1])) + q[2] * np.float32(v[2]))
                 for v in ((p.row0.x, p.row0.y, p.row0.z), (p.row1.x, p.row1.y, p.row1.z),
                           (p.row2.x, p.row2.y, p.row2.z))]
            if r[2] < np.float32(0.05):
                continue
            sx = int(np.float32(np.float32(r[0] * np.float32(ry)) / r[2]) + np.float32(rx) / np.float32(2))
            sy = int(np.float32(np.float32(r[1] * np.float32(ry)) / r[2]) + np.float32(ry) / np.float32(2))
            if not (0 <= sx < rx - 1 and 0 <= sy < ry - 1):
                continue
            d = (int(np.float32(r[2] * np.float32(1000))) & 0xFFFFFFFF) << 8 & 0xFFFFFFFF
            k = (d, pl * N + i)
            t = sy * rx + sx
            if d < 0xFFFFFF00 and (t not in best or k < best[t][0]):
                best[t] = (k, (d + (col & 0xFF)) & 0xFFFFFFFF, c[:3].copy(), r[2])
    return best


def test_proj_first_smallest_wins_random():
    """Random caches with many collisions: orc_proj (the reference's
    pre-check / atom_min / re-check run sequentially) equals the key-min
    formulation, ties going to the earlier source."""
    rng = np.random.default_rng(9)
    rx, ry = 24, 16
    words = np.zeros(2097152 + 16, np.uint32)
    p, _ = O.camera(0.2, 0.1, 0.0, (2.0, 3.0, 1.0), 8, 0)
    for trial in range(3):
        f = O.Frame(rx, ry, words, 0, 8)
        f.memset(0, HOLE, 4 * rx * ry)
        for pl in (1, 2):
            scr = rng.integers(0, 2**32, (ry, rx), dtype=np.uint64).astype(np.uint32)
            scr[rng.random((ry, rx)) < 0.2] = HOLE
            f.screen[pl][:] = scr
            # points in front of the camera, quantized so that depth keys collide
            f.coords[pl][..., 0] = np.round(rng.uniform(-3, 7, (ry, rx)), 1)
            f.coords[pl][..., 1] = np.round(rng.uniform(0, 6, (ry, rx)), 1)
            f.coords[pl][..., 2] = np.round(rng.uniform(3, 6, (ry, rx)), 0)
        scr0 = f.screen.copy()
        crd0 = f.coords.copy()
        f.proj(1, p)
        f.proj(2, p)
        best = _key_min_proj(scr0, crd0, (1, 2), p, rx, ry)
        out = f.screen[0].reshape(-1)
        for t in range(rx * ry):
            if t in best:
                assert out[t] == best[t][1]
                assert np.array_equal(f.coords[0].reshape(-1, 4)[t][:3], best[t][2])
            else:
                assert out[t] == HOLE
        f.close()


@pytest.mark.parametrize("res,add,frame", [((96, 64), (0, 0), 5), ((98, 66), (49, 33), 12), ((96, 64), (48, 32), 8)])
def test_quad_raycast_picks_and_flat_writes(res, add, frame):
    """raycast_fine (kernel.cl:696-843): the traced pixel of each quad is its
    first hole in the order (0,0), (1,0), (0,1), (1,1), else the frame's
    fallback corner; the ray of pixel (x, y) lands at flat offset y*res_x + x,
    also when x == res_x or y == res_y (odd offsets reach the next row and
    plane 1).  Checked by brute force against single-ray traces."""
    xyz, rgba = S.scene(1, 1001, 8)
    words, root = S.build_octree(xyz, rgba, 8)
    rx, ry = res
    ax, ay = add
    p, _ = O.camera(0.7, 0.3, 0.0, (4.0, 30.0, 9.0), 8, frame)
    f = O.Frame(rx, ry, words, root, 8, nthreads=4)
    rng = np.random.default_rng(rx + frame)
    before = np.where(rng.random(4 * rx * ry) < 0.4, HOLE, 0x12345600).astype(np.uint32)
    f.screen.reshape(-1)[:] = before
    f.raycast_quads(p, ax, ay)
    after = f.screen.reshape(-1).copy()
    gx_n, gy_n = ((rx // 4 + 15) // 16) * 16, ((ry // 4 + 15) // 16) * 16
    picks = {}
    for gy in range(gy_n):
        for gx in range(gx_n):
            x, y = 2 * gx + ax, 2 * gy + ay
            if x >= rx or y >= ry:
                continue
            for dx, dy in ((0, 0), (1, 0), (0, 1), (1, 1)):
                if before[(y + dy) * rx + x + dx] == HOLE:
                    x, y = x + dx, y + dy
                    break
            else:
                x, y = x + ((frame >> 2) & 1), y + ((frame >> 3) & 1)
            picks[y * rx + x] = (x, y)
    changed = set(np.nonzero(after != before)[0].tolist())
    assert changed == set(picks)
    assert any(x == rx or y == ry for x, y in picks.values()) == (ax % 2 == 1)
    crd = f.coords.reshape(-1, 4)
    for ofs, (x, y) in list(picks.items())[::7]:
        w, pos, _ = O.trace_pixel(words, root, 8, rx, ry, p, x, y)
        assert int(after[ofs]) == w
        assert np.array_equal(crd[ofs, :3].view(np.uint32), np.array(pos, np.float32).view(np.uint32))
    f.close()


# ---- raycast_fillhole (kernel.cl:342-401), the disabled hole-punching pass:
# each case is derived from the kernel's rule by hand
def _punch_case(z0, nbr, motion_p, nbr_motion, hole=False, px=(8, 8), res=(16, 16)):
    """Pixel px with depth z0; nbr: {(i, j): depth} for its (i, j) in {-1,0}^2
    neighbours (others at z0); motions as (dx, dy).  Returns whether the pixel
    became a hole."""
    rx, ry = res
    words, root = S.build_octree(np.array([[1, 1, 1]], np.uint16), np.array([0x11223344], np.uint32), 8)
    o = O.Frame(rx, ry, words, root, 8, nthreads=1)
    o.screen[0][:] = 0xFF000000 + 7
    o.coords[0][..., 3] = z0
    mx = np.zeros((ry, rx), np.int32)
    my = np.zeros((ry, rx), np.int32)
    x, y = px
    for (i, j), z in nbr.items():
        o.coords[0][y + j, x + i, 3] = z
        mx[y + j, x + i], my[y + j, x + i] = nbr_motion.get((i, j), (0, 0))
    mx[y, x], my[y, x] = motion_p
    if hole:
        o.screen[0][y, x] = 0xFFFFFF00
    o.fillhole_punch(mx, my)
    after = o.screen[0].copy()
    o.close()
    return after[y, x] == 0xFFFFFF00


def test_punch_depth_step_with_different_motion():
    assert _punch_case(10.0, {(0, -1): 5.0}, (0, 0), {(0, -1): (1, 0)})
    assert _punch_case(10.0, {(-1, 0): 5.0}, (0, 0), {(-1, 0): (0, -2)})


def test_punch_needs_different_motion():
    assert not _punch_case(10.0, {(0, -1): 5.0}, (3, 4), {(0, -1): (3, 4)})


def test_punch_needs_a_depth_step():
    # |z1 - z0| must exceed min(z0, z1) * 0.00625 * 4 = 0.2475 for z1 = 9.9
    assert not _punch_case(10.0, {(-1, -1): 9.9}, (0, 0), {(-1, -1): (5, 5)})
    assert _punch_case(10.0, {(-1, -1): 9.7}, (0, 0), {(-1, -1): (5, 5)})
    # a farther neighbour is never the reference
    assert not _punch_case(10.0, {(-1, -1): 20.0}, (0, 0), {(-1, -1): (5, 5)})


def test_punch_picks_the_first_nearest_neighbour():
    # i outer, j inner, strictly smaller: (-1,-1) comes before (0,-1) at equal depth
    assert _punch_case(10.0, {(-1, -1): 5.0, (0, -1): 5.0}, (0, 0), {(-1, -1): (1, 0), (0, -1): (0, 0)})
    assert not _punch_case(10.0, {(-1, -1): 5.0, (0, -1): 5.0}, (0, 0), {(-1, -1): (0, 0), (0, -1): (1, 0)})
    # the nearest wins over an earlier, less near one
    assert not _punch_case(10.0, {(-1, -1): 6.0, (-1, 0): 5.0}, (0, 0), {(-1, -1): (1, 0), (-1, 0): (0, 0)})


def test_punch_skips_holes_and_borders():
    assert _punch_case(10.0, {(0, -1): 5.0}, (0, 0), {(0, -1): (1, 0)}, hole=True)  # stays a hole
    # 3 <= x < res_x - 4: x = 2 and x = 12 (res 16) are never touched
    assert not _punch_case(10.0, {(0, -1): 5.0}, (0, 0), {(0, -1): (1, 0)}, px=(2, 8))
    assert not _punch_case(10.0, {(0, -1): 5.0}, (0, 0), {(0, -1): (1, 0)}, px=(12, 8))
    assert _punch_case(10.0, {(0, -1): 5.0}, (0, 0), {(0, -1): (1, 0)}, px=(11, 11))


def test_punch_without_motion_is_a_no_op():
    rx, ry = 40, 24
    words, root = S.build_octree(np.array([[1, 1, 1]], np.uint16), np.array([0x11223344], np.uint32), 8)
    o = O.Frame(rx, ry, words, root, 8, nthreads=1)
    rng = np.random.default_rng(3)
    o.screen[0][:] = rng.integers(0, 2**32, (ry, rx), dtype=np.uint64).astype(np.uint32)
    o.coords[0][..., 3] = rng.uniform(1, 100, (ry, rx)).astype(np.float32)
    before = o.screen[0].copy()
    o.fillhole_punch(None, None)
    assert np.array_equal(o.screen[0], before)
    o.close()


def test_gated_frame_depth_rule():
    """The gate's depth rule on the oracle (no reference counterpart): a limit
    no reprojected pixel reaches changes nothing; a limit every reprojected
    pixel exceeds sends every reprojected pixel to raycasting (inside the
    refresh rectangle it is traced anyway), so after the frame's rays plane 2
    (the unfiltered frame) holds only raycast words, each equal to the full
    raycast's, and holes."""
    from helpers import CONFIG1, params, path, scene_octree
    words, root = scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])
    rx, ry = 256, 128
    pr = path(CONFIG1, 6)
    frames = {k: O.Frame(rx, ry, words, root, CONFIG1["depth"], nthreads=8) for k in ("plain", "far", "near", "full")}
    try:
        for f in range(6):
            p = O.params_from(params(CONFIG1, pr, f))
            frames["plain"].frame_warp_gated(p, 64)
            frames["far"].frame_warp_gated(p, 64, 1e9)
            frames["near"].frame_warp_gated(p, 64, 1e-3)
            frames["full"].frame_full(p)
            assert np.array_equal(frames["far"].screen, frames["plain"].screen)
            near2, full0 = frames["near"].screen[2], frames["full"].screen[0]
            traced = near2 != O.HOLE
            assert np.array_equal(near2[traced], full0[traced]), f"frame {f}"
            if f >= 2:  # the plain gated frame keeps warped words that the depth rule replaces
                plain2 = frames["plain"].screen[2]
                assert (plain2 != full0).sum() > 100 and traced.mean() > 0.9
    finally:
        for fr in frames.values():
            fr.close()

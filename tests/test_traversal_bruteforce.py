"""The oracle's CAST_RAY restatement (kernel.cl:114-214, fetchChildOffset
:32-62, fetchColor :64-112) against an independent FP64 voxel-grid DDA
(Amanatides-Woo) on a small dense scene: for rays that hit within the LOD-1
range, the hit point lies on the first voxel the FP64 march enters and the
returned colour byte is that voxel's leaf byte.  Rays that stop at a coarser
level (LOD) return the colour the octree defines for the cube they stop in:
its creator voxel's colour (the first voxel of the stream inside it,
octree.h:46-48, :75-81) from level 3 up, the first occupied child's leaf byte
at level 2 (kernel.cl:100-109)."""
import math

import numpy as np
import pytest

import oracle as O
import svo_amd as S

D = 8
SIZE = 1 << D


@pytest.fixture(scope="module")
def scene():
    rng = np.random.default_rng(77)
    occ = np.zeros((SIZE, SIZE, SIZE), bool)          # [x, y, z]
    col = np.zeros((SIZE, SIZE, SIZE), np.uint32)
    occ[:, 0:4, :] = True
    col[:, 0:4, :] = rng.integers(1, 2**32, (SIZE, 4, SIZE), dtype=np.uint64).astype(np.uint32) | 1
    for _ in range(40):
        x0, z0 = rng.integers(0, SIZE - 40, 2)
        y0 = rng.integers(4, 120)
        w, h, d = rng.integers(3, 40, 3)
        occ[x0:x0 + w, y0:y0 + h, z0:z0 + d] = True
        col[x0:x0 + w, y0:y0 + h, z0:z0 + d] = int(rng.integers(1, 2**32)) | 1
    xs, ys, zs = np.nonzero(occ)
    order = np.lexsort((ys, xs, zs))                   # stream order z -> x -> y
    xyz = np.stack([xs, ys, zs], 1)[order].astype(np.uint16)
    rgba = col[xs, ys, zs][order]
    words, root = S.build_octree(xyz, rgba, D)
    return occ, col, words, root, xyz, rgba


def ray_dir32(p, x, y, rx, ry):
    f32 = np.float32
    d1x = (f32(x - rx // 2) + f32(0.5)) * f32(p.fovx) / f32(ry)
    d1y = (f32(y - ry // 2) + f32(0.5)) * f32(p.fovy) / f32(ry)
    cols = [(p.col0.x, p.col0.y, p.col0.z), (p.col1.x, p.col1.y, p.col1.z), (p.col2.x, p.col2.y, p.col2.z)]
    return [float((d1x * f32(c[0]) + d1y * f32(c[1])) + f32(1) * f32(c[2])) for c in cols]


def march(occ, o, d, max_dist):
    """First occupied voxel (2 ray units per voxel, x and z wrapping) along
    o + t*d in FP64; None if none within max_dist or the ray leaves in y."""
    n = math.sqrt(sum(v * v for v in d))
    cell = [math.floor(o[i] / 2) for i in range(3)]
    step = [1 if d[i] > 0 else -1 for i in range(3)]
    tmax, tdel = [], []
    for i in range(3):
        if d[i] == 0:
            tmax.append(math.inf); tdel.append(math.inf); continue
        nb = (cell[i] + (1 if d[i] > 0 else 0)) * 2
        tmax.append((nb - o[i]) / d[i]); tdel.append(2 / abs(d[i]))
    t = 0.0
    while t * n <= max_dist:
        if not (0 <= cell[1] < SIZE):
            return None
        if occ[cell[0] % SIZE, cell[1], cell[2] % SIZE]:
            return cell, t
        a = int(np.argmin(tmax))
        t = tmax[a]
        cell[a] += step[a]
        tmax[a] += tdel[a]
    return None


def test_hits_match_fp64_dda(scene):
    occ, col, words, root, _, _ = scene
    rng = np.random.default_rng(5)
    rx, ry = 2048, 1024                    # lodswitch = 4096 ray units
    checked = agree = 0
    errs = []
    for cam in range(4):
        yaw = float(rng.uniform(-3, 3))
        pitch = float(rng.uniform(0.25, 0.7))
        pos = (float(rng.uniform(0, 32)), float(rng.uniform(17, 24)), float(rng.uniform(0, 32)))
        p, wr = O.camera(yaw, pitch, 0.0, pos, D, 0)
        o = [float(np.float32(v) * np.float32(16)) for v in wr]
        for _ in range(120):
            x, y = int(rng.integers(0, rx)), int(rng.integers(0, ry))
            d = ray_dir32(p, x, y, rx, ry)
            if min(abs(v) for v in d) < 2**-20:
                continue
            hit = march(occ, o, d, 3000.0)
            if hit is None:
                continue
            cell, t = hit
            word, xyz, st = O.trace_pixel(words, root, D, rx, ry, p, x, y)
            checked += 1
            pt = [v * 16.0 for v in xyz]
            on_box = all(cell[i] * 2 - 1e-3 <= pt[i] + (SIZE * 2 * round((cell[i] * 2 - pt[i]) / (SIZE * 2)) if i != 1 else 0)
                         <= cell[i] * 2 + 2 + 1e-3 for i in range(3))
            colour_ok = ((word - 0xFF000000) & 0xFF) == (int(col[cell[0] % SIZE, cell[1], cell[2] % SIZE]) & 0xFF)
            agree += bool(on_box and colour_ok)
            if on_box and colour_ok:
                # FP32 hit point vs the FP64 entry point o + t d of the same voxel
                # (x and z compared modulo the 2*SIZE wrap)
                p64 = [o[i] + t * d[i] for i in range(3)]
                err = []
                for i in range(3):
                    e = pt[i] - p64[i]
                    if i != 1:
                        e -= SIZE * 2 * round(e / (SIZE * 2))
                    err.append(e)
                abs_err = math.sqrt(sum(e * e for e in err))
                rng_len = t * math.sqrt(sum(v * v for v in d))
                errs.append((abs_err, abs_err / max(abs(v) for v in p64), abs_err / rng_len))
    errs = np.array(errs)
    print(f'checked {checked} agree {agree}; hit point vs FP64: max {errs[:, 0].max():.3g} ray units, '
          f'max {errs[:, 1].max():.3g} relative to the coordinate, max {errs[:, 2].max():.3g} relative to the '
          f'ray length; median {np.median(errs[:, 1]):.3g}')
    assert checked > 300
    assert agree >= 0.99 * checked, (agree, checked)
    # FP32 world coordinates of the hits stay within 1e-5 of the coordinate
    # magnitude of the FP64 march's entry points
    assert errs[:, 1].max() < 1e-5, errs[:, 1].max()


def test_lod_hits_return_node_colours(scene):
    occ, col, words, root, xyz, rgba = scene
    # creator voxel of every occupied cube of edge 2^lv voxels: the first voxel
    # of the stream inside it
    creator = {}
    for lv in range(2, D + 1):
        e = 1 << lv
        keys = (xyz[:, 0].astype(np.int64) // e) * (SIZE * SIZE) + (xyz[:, 1].astype(np.int64) // e) * SIZE + \
            xyz[:, 2] // e
        u, first = np.unique(keys, return_index=True)
        creator[lv] = dict(zip(u.tolist(), first.tolist()))
    rng = np.random.default_rng(9)
    rx, ry = 64, 32                        # lodswitch = 128 ray units: most hits stop at a coarse level
    by_level = {}
    for cam in range(4):
        yaw = float(rng.uniform(-3, 3))
        pitch = float(rng.uniform(0.05, 0.5))
        pos = (float(rng.uniform(0, 32)), float(rng.uniform(17, 24)), float(rng.uniform(0, 32)))
        p, _ = O.camera(yaw, pitch, 0.0, pos, D, 0)
        for y in range(ry):
            for x in range(rx):
                d = ray_dir32(p, x, y, rx, ry)
                if min(abs(v) for v in d) < 2**-20:
                    continue
                word, pt, st = O.trace_pixel(words, root, D, rx, ry, p, x, y)
                if not st.last_hit or st.last_rek < 2:
                    continue
                rek = int(st.last_rek)
                lv = rek - 1                   # the cube the ray stopped in: edge 2^lv voxels
                ncell = SIZE >> lv
                # its index from the ray's integer position in the folded frame
                # (folding p -> 2*SIZE - p maps cell k to ncell-1-k)
                c = []
                for i, ip in enumerate((st.last_ix, st.last_iy, st.last_iz)):
                    k = int(ip) >> rek
                    c.append(ncell - 1 - k if (int(st.last_sign) >> i) & 1 else k)
                # the hit point lies on that cube (x and z modulo the 2*SIZE-unit wrap)
                for i in range(3):
                    lo, hi = c[i] * (1 << rek), (c[i] + 1) * (1 << rek)
                    v = pt[i] * 16.0
                    if i != 1:
                        v -= 2 * SIZE * math.floor((v - 0.5 * (lo + hi)) / (2 * SIZE) + 0.5)
                    assert lo - 0.01 <= v <= hi + 0.01, (x, y, i, v, lo, hi)
                cx, cy, cz = c[0] % ncell, c[1], c[2] % ncell
                got = (word - 0xFF000000) & 0xFFFFFFFF
                if lv == 1:
                    exp = next(int(col[2 * cx + (j & 1), 2 * cy + ((j >> 1) & 1), 2 * cz + (j >> 2)]) & 0xFF
                               for j in range(8) if occ[2 * cx + (j & 1), 2 * cy + ((j >> 1) & 1), 2 * cz + (j >> 2)])
                    ok = (got & 0xFF) == exp
                else:
                    i = creator[lv][cx * SIZE * SIZE + cy * SIZE + cz]
                    # level 3 returns a byte of a packed byte-node word; above, the colour word
                    ok = (got & 0xFF) == (int(rgba[i]) & 0xFF) if lv == 2 else (got & 0xFFFFFF) == (int(rgba[i]) & 0xFFFFFF)
                t = by_level.setdefault(rek, [0, 0])
                t[0] += 1
                t[1] += bool(ok)
    print("level: (rays, agreeing)", by_level)
    assert sum(v[0] for v in by_level.values()) > 3000
    assert all(k in by_level for k in (2, 3, 4, 5)), by_level
    assert all(v[0] == v[1] for v in by_level.values()), by_level



def test_pops_go_above_the_highest_changed_bit(scene):
    """Every DDA step of the restatement, logged: a step pops when the XOR of
    the three axes' position changes (kernel.cl:197-201) has its highest bit
    at or above the step's level, and resumes one level above that bit.  The
    geometric rule takes the OR of the changes instead: the lowest level
    whose node is certain to hold the new cell.  The two differ only when
    two axes flip the same highest bit in one step (an edge or corner
    crossing).  Every step where they agree pins the log2, the +1 and the
    2^D mask of kernel.cl:197-208 against the geometric rule (most steps
    change two or three axes, so this covers more than single-axis steps);
    the steps where they differ are counted, including the ones where the
    XOR cancels a bit the OR keeps so that the reference skips a pop the
    geometry calls for."""
    occ, col, words, root, xyz, rgba = scene
    dand = SIZE - 1
    rng = np.random.default_rng(21)
    rx, ry = 256, 128
    agree = quirk = skipped = multi = pops = 0
    for cam in range(3):
        yaw = float(rng.uniform(-3, 3))
        pitch = float(rng.uniform(0.05, 0.5))
        pos = (float(rng.uniform(0, 32)), float(rng.uniform(17, 24)), float(rng.uniform(0, 32)))
        p, _ = O.camera(yaw, pitch, 0.0, pos, D, 0)
        for y in range(0, ry, 8):
            for x in range(0, rx, 8):
                d = ray_dir32(p, x, y, rx, ry)
                if min(abs(v) for v in d) < 2**-20:
                    continue
                _, _, steps, _ = O.trace_steps(words, root, D, rx, ry, p, x, y)
                for bx, by, bz, ix, iy, iz, lv, xor1, far in steps.tolist():
                    c = [(bx ^ ix) & dand, (by ^ iy) & dand, (bz ^ iz) & dand]
                    or_lv = (c[0] | c[1] | c[2]).bit_length()
                    xor_lv = (c[0] ^ c[1] ^ c[2]).bit_length()
                    assert xor1 == xor_lv  # the reference's rule, literally
                    if far:
                        continue
                    popped = xor_lv - 1 >= lv
                    pops += popped
                    multi += popped and sum(v != 0 for v in c) > 1
                    if xor_lv == or_lv:
                        agree += popped
                        if popped:
                            assert (c[0] | c[1] | c[2]) >> lv, "a pop without leaving the node"
                    else:
                        quirk += 1
                        skipped += (or_lv - 1 >= lv) and not popped
    assert agree > 10000 and multi > agree // 2
    assert quirk <= agree // 100
    print(f"pops: {pops}, {agree} at the geometric level ({multi} with several axes changed); "
          f"{quirk} steps where the XOR level differs from the OR level, {skipped} of them skip a pop")

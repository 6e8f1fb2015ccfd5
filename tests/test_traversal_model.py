"""The oracle's CAST_RAY (kernel.cl:114-214) against an independent FP64 model
of what the traversal computes, written from the geometry rather than from
the macro:

  * the ray is o + t d in FP64 (octant-folded so d > 0, as the reference
    folds it, kernel.cl:123-125); occupancy comes from a pyramid of the
    scene's voxels (no octree words);
  * at a point the model descends through the cells of edge 2^l ray units
    that contain it, from the level it resumed at; an occupied cell at a
    level l <= lod is the hit; an empty cell is left through its exit face
    (largest-empty-cell skipping);
  * leaving a cell is a step event at Euclidean distance t|d|; it resumes at
    one level above the highest position bit (below 2^D) the step changed,
    when that level is above the step's (a pop), else at the same level;
  * the LOD switch (lodswitch = 2 res_x, doubled once per switch) is applied
    at pops only (kernel.cl:209, :663);
  * the ray ends by distance (> 400000 after a step, before any pop,
    :203), or when its y position crosses 2^(D+1) (after the pop, :211);
    a ray that ends without a hit returns the colour of the node it ends in
    (:667-672): the node whose children it was testing, i.e. the cube of
    edge 2^(level+1) ray units around its last position (around the position
    before the step for a distance exit).

For each ray the model's step sequence (level, resume level) is compared
with the oracle's step log; where they agree (the FP32 and FP64 positions
cross the same faces) the exit kind, the exit level, the exit cube and the
returned colour must agree too.  Planted mistakes in the oracle (LOD
switched on every trip or on descends, a non-hit exit returning its
parent's colour) fail the comparison.
"""
import math

import numpy as np
import pytest

import oracle as O
import svo_amd as S

D = 8
SIZE = 1 << D
SCALE = 2 * SIZE          # ray units across the world (2 per voxel)
DAND = SIZE - 1
FAR = 400000.0


@pytest.fixture(scope="module")
def scene():
    return make_scene()


def make_scene():
    """A terrain layer plus boxes whose tops stay below 96 voxels, so the
    upper root children and the band of y 96..127 are empty and
    near-horizontal rays there travel to the distance limit (at the root, or
    at level 6 inside the lower root child); voxel colours are random 32-bit
    words."""
    rng = np.random.default_rng(404)
    occ = np.zeros((SIZE, SIZE, SIZE), bool)          # [x, y, z]
    col = np.zeros((SIZE, SIZE, SIZE), np.uint32)
    occ[:, 0:3, :] = True
    col[:, 0:3, :] = rng.integers(1, 2**32, (SIZE, 3, SIZE), dtype=np.uint64).astype(np.uint32)
    for _ in range(60):
        x0, z0 = rng.integers(0, SIZE - 48, 2)
        y0 = rng.integers(3, 60)
        w, d = rng.integers(2, 48, 2)
        h = rng.integers(2, 37)
        occ[x0:x0 + w, y0:y0 + h, z0:z0 + d] = True
        col[x0:x0 + w, y0:y0 + h, z0:z0 + d] = rng.integers(1, 2**32, (w, h, d), dtype=np.uint64).astype(np.uint32)
    xs, ys, zs = np.nonzero(occ)
    order = np.lexsort((ys, xs, zs))                   # stream order z -> x -> y
    xyz = np.stack([xs, ys, zs], 1)[order].astype(np.uint16)
    rgba = col[xs, ys, zs][order]
    words, root = S.build_octree(xyz, rgba, D)
    # occupancy pyramid: level lv = cubes of edge 2^lv voxels
    pyr = [occ]
    for lv in range(1, D + 1):
        a = pyr[-1]
        n = a.shape[0] // 2
        pyr.append(a.reshape(n, 2, n, 2, n, 2).any(axis=(1, 3, 5)))
    # creator voxel (first in the stream) of every occupied cube, octree.h:46-48, :75-81
    creator = {}
    for lv in range(1, D + 1):
        e = 1 << lv
        n = SIZE >> lv
        keys = (xyz[:, 0].astype(np.int64) // e) * n * n + (xyz[:, 1].astype(np.int64) // e) * n + xyz[:, 2] // e
        u, first = np.unique(keys, return_index=True)
        creator[lv] = dict(zip(u.tolist(), first.tolist()))
    return dict(occ=occ, col=col, words=words, root=root, pyr=pyr, creator=creator, rgba=rgba)


def ray_dir32(p, x, y, rx, ry):
    f32 = np.float32
    d1x = (f32(x - rx // 2) + f32(0.5)) * f32(p.fovx) / f32(ry)
    d1y = (f32(y - ry // 2) + f32(0.5)) * f32(p.fovy) / f32(ry)
    cols = [(p.col0.x, p.col0.y, p.col0.z), (p.col1.x, p.col1.y, p.col1.z), (p.col2.x, p.col2.y, p.col2.z)]
    return [float((d1x * f32(c[0]) + d1y * f32(c[1])) + f32(1) * f32(c[2])) for c in cols]


def unfold_cube(k, lvl_units, signs):
    """Folded cell indices of edge lvl_units ray units -> unfolded cube
    indices (x and z wrapped to the world)."""
    n = SCALE // lvl_units
    out = []
    for i in range(3):
        v = k[i]
        if i != 1:
            v %= n
        if (signs >> i) & 1:
            v = n - 1 - v
        out.append(v)
    return out


def model_trace(sc, o, d, signs, lodsw0, max_steps=200000):
    """FP64 model (module docstring).  o, d: folded origin and direction
    (d > 0).  Returns (kind, level, folded cube, steps, flags) with kind in
    hit / far / y, the cube of edge 2^level (hit) or 2^(level+1) (far, y)
    ray units, steps = [(level, resume level)], and flags naming where FP32
    rounding may legitimately decide otherwise: "geom" (a step crosses an
    edge or corner, another axis ends within FP32 rounding of an integer, or
    two axes flip the same highest bit), "lod" (a pop's distance within FP32
    rounding of lodswitch), "far" (the last distance within FP32 rounding of
    the limit)."""
    pyr = sc["pyr"]
    dn = math.sqrt(sum(v * v for v in d))
    lvl, lod, lodsw = D, 1, float(lodsw0)
    pos = list(o)
    steps, flags = [], set()
    quirks = sc.setdefault("quirks", [0, 0])  # XOR-level steps, skipped pops (of the last ray)
    quirks[0] = quirks[1] = 0

    def occupied(l, k):
        n = SCALE >> l
        if not 0 <= k[1] < n:
            return False
        u = unfold_cube(k, 1 << l, signs)
        return bool(pyr[l - 1][u[0], u[1], u[2]])

    for _ in range(max_steps):
        k = [math.floor(pos[i] / (1 << lvl)) for i in range(3)]
        if occupied(lvl, k):
            if lvl <= lod:
                return "hit", lvl, k, steps, flags
            lvl -= 1
            continue
        # leave the empty cell through its exit face
        hi = [(k[i] + 1) * (1 << lvl) for i in range(3)]
        tt = [(hi[i] - o[i]) / d[i] for i in range(3)]
        a = min(range(3), key=lambda i: tt[i])
        t_new = tt[a]
        srt = sorted(tt)
        if srt[1] - srt[0] <= 2.0**-20 * max(1.0, srt[0]):
            flags.add("geom")  # an edge or corner: the FP32 order of the faces is not fixed
        new = [o[i] + t_new * d[i] for i in range(3)]
        new[a] = float(hi[a])
        ib = [math.floor(v) for v in pos]
        ia = [math.floor(v) for v in new]
        for i in range(3):
            if i != a and abs(new[i] - round(new[i])) <= 2.0**-19 * max(1.0, abs(new[i])):
                flags.add("geom")  # another axis ends on an integer: FP32 truncation may differ
        ch = [(ib[i] ^ ia[i]) & DAND for i in range(3)]
        xo = ch[0] ^ ch[1] ^ ch[2]
        orl = (ch[0] | ch[1] | ch[2]).bit_length()
        if orl != xo.bit_length():
            flags.add("geom")  # two axes flip the same highest bit: the XOR cancels it
            quirks[0] += 1
            quirks[1] += orl - 1 >= lvl > xo.bit_length() - 1  # a pop the geometry calls for, skipped
        resume = xo.bit_length()                  # highest changed bit + 1
        dist = t_new * dn
        steps.append((lvl, resume))
        if abs(dist - FAR) < 1e-3 * FAR:
            flags.add("far")
        if dist > FAR:
            return "far", lvl, [math.floor(pos[i] / (2 << lvl)) for i in range(3)], steps, flags
        pos = new
        if resume - 1 >= lvl:                      # pop
            lvl = resume
            if abs(dist - lodsw) < 1e-3 * lodsw:
                flags.add("lod")
            if dist > lodsw:                       # LOD switch, on pops only
                lodsw *= 2.0
                lod += 1
        if (ib[1] ^ ia[1]) & (2 * SIZE):           # y crossed 2^(D+1)
            return "y", lvl, [math.floor(pos[i] / (2 << lvl)) for i in range(3)], steps, flags
    return "cap", lvl, None, steps, {"geom"}


def oracle_trace(sc, p, x, y, rx, ry):
    word, xyz, log, st = O.trace_steps(sc["words"], sc["root"], D, rx, ry, p, x, y)
    steps = []
    for r in log.tolist():
        # the oracle pops when the XOR level is at or above the step's level
        steps.append((r[6], r[7]))
    if st.last_hit:
        kind = "hit"
        k = [int(st.last_ix) >> int(st.last_rek), int(st.last_iy) >> int(st.last_rek), int(st.last_iz) >> int(st.last_rek)]
    else:
        last = log[-1].tolist()
        if last[8]:
            kind, src = "far", last[0:3]
        else:
            kind, src = "y", last[3:6]
            assert ((last[1] ^ last[4]) >> (D + 1)) & 1
        k = [int(v) >> (int(st.last_rek) + 1) for v in src]
    return word, kind, int(st.last_rek), k, steps, st


def expected_colour(sc, kind, lvl, cube, signs):
    """The colour word the octree defines for the cube a ray stopped in, or
    None where the low levels' byte packing makes it a byte (checked on its
    low byte)."""
    units = (1 << lvl) if kind == "hit" else (2 << lvl)
    lv = int(math.log2(units)) - 1                 # cube edge 2^lv voxels
    if lv == D:  # the root: created empty before any voxel (raycast.h:14-16)
        return 0, 0xFFFFFFFF
    u = unfold_cube(cube, units, signs)
    n = SIZE >> lv
    if lv == 0:
        return int(sc["col"][u[0], u[1], u[2]]) & 0xFF, 0xFF
    i = sc["creator"][lv][u[0] * n * n + u[1] * n + u[2]]
    if lv == 1:   # a depth-(D-1) node: its first occupied leaf's byte (kernel.cl:100-109)
        for j in range(8):
            a, b, c = 2 * u[0] + (j & 1), 2 * u[1] + ((j >> 1) & 1), 2 * u[2] + (j >> 2)
            if sc["occ"][a, b, c]:
                return int(sc["col"][a, b, c]) & 0xFF, 0xFF
    if lv == 2:   # a byte node: its colour byte (octree.h:117)
        return int(sc["rgba"][i]) & 0xFF, 0xFF
    return int(sc["rgba"][i]), 0xFFFFFFFF


CAMERAS = [
    # (yaw, pitch, position in voxels, res_x, res_y, pixel stride): LOD hits
    # at many levels (small screens: lodswitch 2*res_x), distance and y exits
    (0.7, 0.35, (30.0, 110.0, 40.0), 64, 32, 1),
    (-2.1, 0.20, (200.0, 95.0, 17.0), 96, 48, 1),
    (2.6, 0.05, (90.0, 60.0, 230.0), 128, 64, 2),
    (1.3, -0.25, (10.0, 30.0, 10.0), 128, 64, 2),
    # horizon rows near the screen's centre column (pitch 0, yaw a multiple
    # of 90 degrees: rays close to an axis, whose FP32 steps cross the same
    # faces as the FP64 march): distance exits at the root (camera above half
    # height) and at level 6 inside the lower root child (the empty band
    # 96..127), and y exits
    (0.0, 0.0, (77.0, 200.0, 150.0), 2048, 1024, 0),
    (1.5707964, 0.0, (140.0, 190.0, 60.0), 4096, 2048, 0),
    (0.0, 0.0, (20.0, 112.0, 90.0), 8192, 4096, 0),
    (3.1415927, 0.0, (220.0, 108.0, 12.0), 8192, 4096, 0),
    (-1.5707964, 0.0, (30.0, 118.0, 200.0), 8192, 4096, 0),
]


def _rays(sc, cams):
    for yaw, pitch, pos, rx, ry, stride in cams:
        p, wr = O.camera(yaw, pitch, 0.0, pos, D, 0)
        if stride:
            pix = [(x, y) for y in range(0, ry, stride) for x in range(0, rx, stride)]
        else:  # the rows around the horizon, 32 columns around the centre
            pix = [(x, y) for y in range(ry // 2 - 3, ry // 2 + 3) for x in range(rx // 2 - 96, rx // 2 + 96, 6)]
        for x, y in pix:
            d = ray_dir32(p, x, y, rx, ry)
            if min(abs(v) for v in d) < 2**-20:
                continue
            yield p, wr, rx, ry, x, y, d


def _compare(sc, cams=CAMERAS):
    """Rays whose step sequences agree must agree on the exit and its colour
    (unless a pop's or the last distance lies within FP32 rounding of
    lodswitch or the limit); rays whose sequences differ must show an FP32
    rounding-sensitive step."""
    stats = {"rays": 0, "same_steps": 0, "diverged_geom": 0, "diverged": 0, "lod_or_far_tie": 0,
             "kinds": {}, "levels": {}, "colour_checked": 0}
    bad = []
    for p, wr, rx, ry, x, y, d in _rays(sc, cams):
        o32 = [np.float32(v) * np.float32(16) for v in wr]
        signs = 0
        o, dd = [], []
        for i in range(3):
            if d[i] < 0:
                signs |= 1 << i
                o.append(float(np.float32(SCALE) - o32[i]))
                dd.append(-d[i])
            else:
                o.append(float(o32[i]))
                dd.append(d[i])
        kind, lvl, cube, msteps, flags = model_trace(sc, o, dd, signs, 2 * rx)
        word, okind, olvl, ocube, osteps, st = oracle_trace(sc, p, x, y, rx, ry)
        stats["rays"] += 1
        n = min(len(msteps), len(osteps))
        if msteps[:n] != osteps[:n]:
            if "geom" in flags:
                stats["diverged_geom"] += 1
            else:
                stats["diverged"] += 1
                bad.append(("steps", x, y, rx, len(msteps), len(osteps)))
            continue
        if (okind, olvl, ocube) != (kind, lvl, cube) or len(msteps) != len(osteps):
            if flags & {"lod", "far"}:
                stats["lod_or_far_tie"] += 1
            else:
                bad.append(("exit", x, y, rx, kind, lvl, cube, okind, olvl, ocube))
            continue
        stats["same_steps"] += 1
        stats["xor_steps"] = stats.get("xor_steps", 0) + sc["quirks"][0]
        stats["xor_skipped_pops"] = stats.get("xor_skipped_pops", 0) + sc["quirks"][1]
        stats["kinds"][kind] = stats["kinds"].get(kind, 0) + 1
        stats["levels"][(kind, lvl)] = stats["levels"].get((kind, lvl), 0) + 1
        if kind != "hit" and lvl < 3:
            continue  # exits at the byte levels read past their node (kernel.cl:77-89)
        exp, mask = expected_colour(sc, kind, lvl, cube, signs)
        stats["colour_checked"] += 1
        if ((word - 0xFF000000) & mask) != exp:
            bad.append(("colour", x, y, rx, kind, lvl, hex(word), hex(exp)))
    return stats, bad


def test_traversal_matches_fp64_model(scene):
    stats, bad = _compare(scene)
    print(stats)
    assert not bad, bad[:10]
    assert stats["same_steps"] >= 0.9 * stats["rays"], stats
    k = stats["kinds"]
    assert k.get("hit", 0) > 2000 and k.get("far", 0) > 100 and k.get("y", 0) > 50, k
    levels = stats["levels"]
    assert all(("hit", l) in levels for l in range(1, 7)), levels
    assert ("far", D) in levels and ("far", 6) in levels and ("y", D) in levels, levels
    # steps where two axes flip the same highest bit, on rays whose whole
    # sequence agrees: there the oracle resumes at the XOR level (kernel.cl:
    # 197-202), as the model does
    assert stats["xor_steps"] > 100, stats


@pytest.mark.parametrize("mutation", [1, 2, 3])
def test_planted_traversal_mistakes_fail(scene, mutation):
    """LOD switched on every trip (1) or on descends (2) instead of on pops,
    or a ray ending without a hit returning its parent node's colour (3):
    each breaks the agreement with the model."""
    O.set_mutation(mutation)
    try:  # LOD hits at many levels, and distance exits at levels 6 and 7
        stats, bad = _compare(scene, [CAMERAS[1], CAMERAS[6]])
    finally:
        O.set_mutation(0)
    print(mutation, stats, len(bad), bad[:3])
    assert len(bad) > 20

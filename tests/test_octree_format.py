"""Octree builder parity: the oracle restatement and the product's C++
builder against the reference octree.h (golden fixtures generated from
oracle/_ref, plus the compiled reference itself when present), and a
hand-derived word array for a one-voxel scene."""
import glob
import os

import numpy as np
import pytest

import oracle as O
import svo_amd as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _load(path):
    z = np.load(path)
    words = np.zeros(int(z["n_words"]), np.uint32)
    words[z["nz_index"]] = z["nz_value"]
    return z["xyz"], z["rgba"], int(z["depth"]), words, int(z["root"])


FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "octree_*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[7:-4] for p in FIXTURES])
def test_builders_match_reference_fixture(path):
    xyz, rgba, depth, words, root = _load(path)
    w1, r1 = O.build_octree(xyz, rgba, depth)
    w2, r2 = S.build_octree(xyz, rgba, depth)
    assert r1 == root and r2 == root
    assert np.array_equal(w1, words), "oracle restatement differs from the reference words"
    assert np.array_equal(w2, words), "product builder differs from the reference words"


def test_one_voxel_hand_derived():
    """octree.h:232-293 worked by hand for one voxel at the origin, D = 8:
    normal region = depth-1 record at 0 and the root at 10; the block of the
    depth-2 node starts at word 2^21 (raycast.h:38) and holds the copied block
    root record, the byte node (colour, mask, leaf byte) and the compact
    records of depths 5, 4, 3 in post order."""
    C = 0x11223301
    words, root = S.build_octree(np.array([[0, 0, 0]], np.uint16), np.array([C], np.uint32), 8)
    LR = 2097152
    exp = {0: (LR >> 6 << 9) | 256 | 1, 8: C, 9: C, 10: 1,
           LR + 0: C, LR + 1: (7 << 9) + 257, LR + 2: 0x00010101,
           LR + 3: C, LR + 4: (2 << 9) + 257, LR + 5: C, LR + 6: (3 << 9) + 257,
           LR + 7: C, LR + 8: (5 << 9) + 257}
    assert root == (10 << 9) | 1
    assert len(words) == LR + 9
    nz = {int(i): int(words[i]) for i in np.nonzero(words)[0]}
    assert nz == exp


@pytest.mark.skipif(not O.ref_available(), reason="reference octree builder not built")
@pytest.mark.parametrize("depth,n,seed", [(8, 20000, 1), (9, 30000, 2), (11, 50000, 3), (12, 20000, 4)])
def test_builders_match_compiled_reference(depth, n, seed):
    rng = np.random.default_rng(seed)
    size = 1 << depth
    xyz = rng.integers(0, size, (n, 3)).astype(np.uint16)
    xyz[n // 2:n // 2 + n // 10] = xyz[:n // 10]      # duplicates: last colour, first creator
    rgba = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    rgba[::97] = 0                                    # zero colour words (octree.h:122,133 skip them)
    wr, rr = O.ref_build_octree(xyz, rgba, depth)
    w1, r1 = O.build_octree(xyz, rgba, depth)
    w2, r2 = S.build_octree(xyz, rgba, depth)
    assert rr == r1 == r2
    assert np.array_equal(wr, w1)
    assert np.array_equal(wr, w2)


@pytest.mark.skipif(not O.ref_available(), reason="reference octree builder not built")
def test_scene_flythrough_256_matches_reference():
    xyz, rgba = S.scene(1, 1001, 8)
    wr, rr = O.ref_build_octree(xyz, rgba, 8)
    w2, r2 = S.build_octree(xyz, rgba, 8)
    assert rr == r2 and np.array_equal(wr, w2)


def test_empty_scene_and_errors():
    w, r = S.build_octree(np.zeros((0, 3), np.uint16), np.zeros(0, np.uint32), 8)
    assert r == 0 and len(w) == 2097152 and not w.any()
    with pytest.raises(S.SvoError):
        S.build_octree(np.array([[256, 0, 0]], np.uint16), np.array([1], np.uint32), 8)
    with pytest.raises(S.SvoError):
        S.build_octree(np.array([[0, 0, 0]], np.uint16), np.array([1], np.uint32), 7)


def test_save_load_roundtrip(tmp_path):
    xyz, rgba = S.scene(1, 1001, 8)
    w, r = S.build_octree(xyz, rgba, 8)
    p = str(tmp_path / "scene.svo")
    S.octree_save(p, w, r, 8)
    w2, r2, d2 = S.octree_load(p)
    assert r2 == r and d2 == 8 and np.array_equal(w, w2)
    bad = tmp_path / "bad.svo"
    bad.write_bytes(b"not an octree")
    with pytest.raises(S.SvoError):
        S.octree_load(str(bad))

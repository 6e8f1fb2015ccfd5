"""GPU octree builder: word-for-word equal to the reference's octree.h
builder through its committed outputs (tests/golden, written from the
reference compiled on the host), to the oracle's restatement of the builder,
and to the host builder."""
import glob
import os

import numpy as np
import pytest

import oracle as O
import svo_amd as S
from helpers import CONFIG1, CONFIG2

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "octree_*.npz")))

pytestmark = pytest.mark.gpu


def _random_stream(rng, D, n, zero_frac=0.1):
    size = 1 << D
    c = rng.integers(0, size, (max(1, n // 4), 3))
    xyz = (np.concatenate([c + rng.integers(0, 3, c.shape) for _ in range(4)]) % size).astype(np.uint16)
    rgba = rng.integers(0, 2**32, len(xyz), dtype=np.uint64).astype(np.uint32)
    rgba[rng.random(len(xyz)) < zero_frac] = 0
    return xyz, rgba


@pytest.mark.parametrize("D,n,seed", [(8, 500, 1), (8, 6000, 2), (9, 4000, 3), (10, 3000, 4), (11, 2000, 5),
                                      (12, 2000, 6), (13, 1000, 7), (8, 1, 8)])
def test_random_streams_match_host_builder(D, n, seed):
    rng = np.random.default_rng(seed)
    xyz, rgba = _random_stream(rng, D, n)
    print(f"build D={D} n={len(xyz)}", flush=True)
    w0, r0 = S.build_octree(xyz, rgba, D)
    w1, r1 = S.build_octree_gpu(xyz, rgba, D)
    assert r0 == r1
    assert len(w0) == len(w1)
    bad = np.nonzero(w0 != w1)[0]
    assert bad.size == 0, f"{bad.size} words differ, first at {bad[0]}"


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_fixtures(path):
    """The reference builder's words (octree.h:32-293 via convert_tree_blocks,
    raycast.h:38-39), stored sparsely: every nonzero word and the length."""
    z = np.load(path, allow_pickle=False)
    words, root = S.build_octree_gpu(z["xyz"], z["rgba"], int(z["depth"]))
    assert root == int(z["root"]) and len(words) == int(z["n_words"])
    nz = np.nonzero(words)[0]
    assert np.array_equal(nz, z["nz_index"]) and np.array_equal(words[nz], z["nz_value"])


@pytest.mark.parametrize("D,n,seed", [(8, 3000, 11), (10, 4000, 12), (12, 1500, 13)])
def test_random_streams_match_oracle_builder(D, n, seed):
    """Against the oracle's restatement of set_voxel + convert_tree_blocks, on
    streams with duplicate voxels and zero colour words (octree.h:122, :133
    skip those), below the reference's 2^24-node limit."""
    rng = np.random.default_rng(seed)
    xyz, rgba = _random_stream(rng, D, n)
    wo, ro = O.build_octree(xyz, rgba, D)
    w1, r1 = S.build_octree_gpu(xyz, rgba, D)
    assert r1 == ro and len(w1) == len(wo)
    assert np.array_equal(w1, wo)


def test_dense_block_and_duplicates():
    # a full 64^3 block (all byte/compact records full) written twice with
    # different colours: last colour wins at leaves, first voxel is the creator
    g = np.stack(np.meshgrid(np.arange(64), np.arange(64), np.arange(64), indexing="ij"), -1).reshape(-1, 3)
    xyz = np.concatenate([g, g]).astype(np.uint16) + 64
    rgba = np.concatenate([np.full(len(g), 0x00112233, np.uint32), np.arange(len(g), dtype=np.uint32) * 7 + 1])
    w0, r0 = S.build_octree(xyz, rgba, 8)
    w1, r1 = S.build_octree_gpu(xyz, rgba, 8)
    assert r0 == r1 and np.array_equal(w0, w1)


def test_scene_config1_and_config2():
    for cfg in (CONFIG1, CONFIG2):
        xyz, rgba = S.scene(cfg["kind"], cfg["seed"], cfg["depth"], cfg.get("spheres", 0))
        print(f"scene depth {cfg['depth']}: {len(xyz)} voxels", flush=True)
        w0, r0 = S.build_octree(xyz, rgba, cfg["depth"])
        w1, r1 = S.build_octree_gpu(xyz, rgba, cfg["depth"])
        assert r0 == r1 and len(w0) == len(w1)
        assert np.array_equal(w0, w1)


def test_empty_and_bad_input():
    w, r = S.build_octree_gpu(np.zeros((0, 3), np.uint16), np.zeros(0, np.uint32), 9)
    w0, r0 = S.build_octree(np.zeros((0, 3), np.uint16), np.zeros(0, np.uint32), 9)
    assert r == r0 and np.array_equal(w, w0)
    with pytest.raises(S.SvoError):
        S.build_octree_gpu(np.array([[256, 0, 0]], np.uint16), np.array([1], np.uint32), 8)

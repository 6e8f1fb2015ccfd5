"""Writes the golden fixtures under tests/golden/ from the REFERENCE octree
builder (src/octree/octree.h compiled on the host into oracle/_ref by
oracle/Makefile).  Fixtures are data only: the input voxel streams and the
reference's output word arrays, stored sparsely (index, value) because the
reference pre-sizes the normal region to 2^21 words.

    make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def scenes():
    rng = np.random.default_rng(20261202)
    out = {}
    # one voxel at the origin, colour 0x11223301
    out["one_voxel_d8"] = (np.array([[0, 0, 0]], np.uint16), np.array([0x11223301], np.uint32), 8)
    # the 8 voxels of one 2x2x2 cell, distinct colours, inserted in reverse order
    xyz = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)][::-1], np.uint16) + 100
    out["cell8_d8"] = (xyz, (np.arange(8, dtype=np.uint32) * 0x01010101 + 0x05000005), 8)
    # a full 4x4x4 brick (one byte node's worth) at a block boundary
    g = np.array([[x, y, z] for z in range(62, 66) for y in range(30, 34) for x in range(126, 130)], np.uint16)
    out["brick64_d8"] = (g, (rng.integers(1, 2**32, len(g), dtype=np.uint64)).astype(np.uint32), 8)
    # random sparse scene with duplicate positions (last write wins, first creates)
    n = 3000
    xyz = rng.integers(0, 256, (n, 3)).astype(np.uint16)
    xyz[1500:1600] = xyz[0:100]
    out["random3000_d8"] = (xyz, rng.integers(1, 2**32, n, dtype=np.uint64).astype(np.uint32), 8)
    n = 2000
    out["random2000_d10"] = (rng.integers(0, 1024, (n, 3)).astype(np.uint16),
                             rng.integers(1, 2**32, n, dtype=np.uint64).astype(np.uint32), 10)
    return out


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref_octree.so missing: run `make -C oracle` with the reference present")
    for name, (xyz, rgba, depth) in scenes().items():
        words, root = O.ref_build_octree(xyz, rgba, depth)
        nz = np.nonzero(words)[0].astype(np.uint32)
        np.savez_compressed(os.path.join(HERE, f"octree_{name}.npz"), xyz=xyz, rgba=rgba,
                            depth=np.int32(depth), n_words=np.uint64(len(words)), root=np.uint32(root),
                            nz_index=nz, nz_value=words[nz])
        print(name, len(words), hex(root), len(nz))


if __name__ == "__main__":
    main()

"""The cave scenes (kind 3: a slab of 1/8 of the height, kind 4: the whole
height) generated on the GPU are the host generator's voxel streams, bit for
bit (stream order, coordinates and colour words)."""
import time

import numpy as np
import pytest

import svo_amd as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,seed", [(3, 1003), (3, 77), (4, 1003)])
def test_cave_gpu_matches_host_depth8(kind, seed):
    t = time.time()
    xyz_h, rgba_h = S.scene(kind, seed, 8)
    th = time.time() - t
    t = time.time()
    xyz_g, rgba_g = S.scene_gpu(kind, seed, 8)
    print(f"cave kind {kind} depth 8 seed {seed}: {len(xyz_h)} voxels, host {th:.1f}s, gpu {time.time() - t:.2f}s",
          flush=True)
    assert len(xyz_g) == len(xyz_h)
    assert np.array_equal(xyz_g, xyz_h)
    assert np.array_equal(rgba_g, rgba_h)


def test_cave_depth12_stream_properties():
    """At depth 12 (host generation would take hours): every voxel inside the
    slab, stream order (z, x, y) strictly increasing, and the octree built on
    the GPU uploads and validates."""
    t = time.time()
    xyz, rgba = S.scene_gpu(3, 1003, 12)
    print(f"cave depth 12: {len(xyz)} voxels in {time.time() - t:.1f}s", flush=True)
    assert len(xyz) > 1000
    x, y, z = (xyz[:, i].astype(np.int64) for i in range(3))
    assert x.max() < 4096 and z.max() < 4096 and y.max() < 512
    key = (z << 24) | (x << 12) | y
    assert np.all(np.diff(key) > 0)
    assert np.all((rgba & 3) == 2)
    t = time.time()
    words, root = S.build_octree_gpu(xyz, rgba, 12)
    print(f"  built {len(words)} words in {time.time() - t:.1f}s", flush=True)
    assert len(words) < (1 << 29)
    with S.Context(0) as c:
        c.upload_octree(words, root, 12)


def test_cave_depth9_warp_frames_vs_oracle():
    """The divergence-stress scene kind (cave, fast camera) through the warp
    pipeline: bit-exact against the oracle at depth 9."""
    import oracle as O
    from helpers import assert_equal_f32_bits, assert_equal_u32
    xyz, rgba = S.scene_gpu(3, 1003, 9)
    words, root = S.build_octree_gpu(xyz, rgba, 9)
    rows = S.camera_path(2003, 9, 8, 4.0, 5.0, 32.0, 0.05)
    rx, ry = 320, 160
    o = O.Frame(rx, ry, words, root, 9, nthreads=8)
    with S.Context(0) as c:
        c.upload_octree(words, root, 9)
        c.resize(rx, ry)
        for f in range(8):
            yaw, pitch, roll, x, y, z = [float(v) for v in rows[f]]
            p, _ = S.camera(yaw, pitch, roll, (x, y, z), 9, f)
            c.frame_warp(p)
            o.frame_warp(O.params_from(p))
            c.sync()
            assert_equal_u32(c.screen(), o.screen, f"cave frame {f} screen")
            assert_equal_f32_bits(c.coords(), o.coords, f"cave frame {f} coords")
            assert_equal_u32(c.image(), o.image, f"cave frame {f} image")
    o.close()

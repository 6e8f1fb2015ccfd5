"""Shared test helpers: scenes, cameras and comparisons."""
import functools

import numpy as np

import oracle as O
import svo_amd as S

# workload configurations (scene seed = 1000 + config, path seed = 2000 + config)
CONFIG1 = dict(kind=1, seed=1001, depth=8, res=(512, 256), frames=64, vox=2.0, yaw=0.5,
               height=240.0, pitch=0.35, path_seed=2001)
CONFIG2 = dict(kind=2, seed=1002, depth=11, res=(1920, 1024), frames=300, vox=8.0, yaw=1.0,
               height=1400.0, pitch=0.30, path_seed=2002, spheres=2000)


@functools.lru_cache(maxsize=4)
def scene_octree(kind, seed, depth, spheres=0):
    xyz, rgba = S.scene(kind, seed, depth, spheres)
    words, root = S.build_octree(xyz, rgba, depth)
    return words, root


def path(cfg, n=None):
    return S.camera_path(cfg["path_seed"], cfg["depth"], n or cfg["frames"], cfg["vox"], cfg["yaw"],
                         cfg["height"], cfg["pitch"])


def params(cfg, path_rows, f):
    yaw, pitch, roll, x, y, z = [float(v) for v in path_rows[f]]
    p, _ = S.camera(yaw, pitch, roll, (x, y, z), cfg["depth"], f)
    return p


def assert_equal_u32(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    bad = np.nonzero(a.reshape(-1) != b.reshape(-1))[0]
    assert bad.size == 0, (f"{what}: {bad.size} mismatches, first at flat index {bad[0]}: "
                           f"{a.reshape(-1)[bad[0]]:#x} vs {b.reshape(-1)[bad[0]]:#x}")


def assert_equal_f32_bits(a, b, what):
    assert_equal_u32(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32), what)


def read_png(raw):
    """Decode an 8-bit RGB, filter-0 PNG with the standard library: every
    chunk CRC checked with binascii.crc32, the IDAT stream inflated by zlib
    (which checks its Adler-32)."""
    import binascii
    import struct
    import zlib
    assert raw[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(raw):
        n, = struct.unpack(">I", raw[pos:pos + 4])
        typ, data = raw[pos + 4:pos + 8], raw[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", raw[pos + 8 + n:pos + 12 + n])
        assert crc == binascii.crc32(typ + data) & 0xffffffff, typ
        chunks.append((typ, data))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    rows = zlib.decompress(chunks[1][1])
    assert len(rows) == h * (3 * w + 1)
    px = np.frombuffer(rows, np.uint8).reshape(h, 3 * w + 1)
    assert (px[:, 0] == 0).all()
    return px[:, 1:].reshape(h, w, 3)


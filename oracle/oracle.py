"""ctypes wrapper of the CPU oracle (liboracle.so) and of the reference
octree builder compiled on the host (_ref/libref_octree.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Build with `make -C oracle`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOLE = 0xFFFFFF00

_u32p = C.POINTER(C.c_uint32)


class F4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Params(C.Structure):
    """orc_params (same field order and size as svo_frame_params)."""
    _fields_ = [("m0", F4), ("row0", F4), ("row1", F4), ("row2", F4),
                ("col0", F4), ("col1", F4), ("col2", F4),
                ("fovx", C.c_float), ("fovy", C.c_float), ("frame", C.c_int32), ("pad", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("iters", C.c_uint64), ("descends", C.c_uint64),
                ("steps", C.c_uint64), ("pops", C.c_uint64), ("fetch_words", C.c_uint64),
                ("desc_by_rek", C.c_uint64 * 16), ("last_rek", C.c_int64), ("last_lod", C.c_int64),
                ("last_hit", C.c_int64), ("last_ix", C.c_int64), ("last_iy", C.c_int64), ("last_iz", C.c_int64),
                ("last_sign", C.c_int64), ("max_iters", C.c_uint64)]


class State(C.Structure):
    _fields_ = [("res_x", C.c_int), ("res_y", C.c_int), ("depth", C.c_int),
                ("oct", _u32p), ("n_words", C.c_size_t), ("root", C.c_uint32),
                ("screen", _u32p), ("coords", C.POINTER(C.c_float)), ("ids", _u32p),
                ("image", _u32p), ("nthreads", C.c_int), ("st", Stats)]


def _load(path):
    if not os.path.exists(path):
        raise ImportError(f"{path} missing; run `make -C oracle`")
    return C.CDLL(path)


lib = _load(os.path.join(HERE, "liboracle.so"))


def _sig(name, restype, *args):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(args)
    return f


_P = C.POINTER(Params)
_S = C.POINTER(State)
_camera = _sig("orc_camera", None, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float), C.c_int,
               C.c_int, _P, C.POINTER(C.c_float))
_build = _sig("orc_build_octree", C.c_int, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
              C.POINTER(_u32p), C.POINTER(C.c_size_t), C.POINTER(C.c_uint32))
_free = _sig("orc_free", None, C.c_void_p)
_trace = _sig("orc_trace_pixel", C.c_uint32, _u32p, C.c_size_t, C.c_uint32, C.c_int, C.c_int, C.c_int,
              _P, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(Stats))
_trace_pops = _sig("orc_trace_pixel_pops", C.c_uint32, _u32p, C.c_size_t, C.c_uint32, C.c_int, C.c_int, C.c_int,
                   _P, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int,
                   C.POINTER(C.c_int), C.POINTER(Stats))
set_mutation = _sig("orc_set_mutation", None, C.c_int)
_init = _sig("orc_state_init", C.c_int, _S, C.c_int, C.c_int, _u32p, C.c_size_t, C.c_uint32, C.c_int)
_sfree = _sig("orc_state_free", None, _S)
_memset = _sig("orc_memset_screen", None, _S, C.c_size_t, C.c_uint32, C.c_size_t)
_proj = _sig("orc_proj", None, _S, C.c_int, _P)
_count = _sig("orc_counthole", None, _S)
_sum = _sig("orc_sumids", None, _S)
_write = _sig("orc_writeids", None, _S)
_holes = _sig("orc_raycast_holes", None, _S, _P)
_rect = _sig("orc_raycast_rect", None, _S, _P, C.c_int, C.c_int, C.c_int, C.c_int)
_quads = _sig("orc_raycast_quads", None, _S, _P, C.c_int, C.c_int)
_punch = _sig("orc_fillhole_punch", None, _S, C.POINTER(C.c_int32), C.POINTER(C.c_int32))
_copy = _sig("orc_copy_plane", None, _S, C.c_int)
_fill_snap = _sig("orc_fillhole2_snapshot", None, _S, _u32p)
_fill_inpl = _sig("orc_fillhole2_inplace", None, _S)
_colorize = _sig("orc_colorize", None, _S)
_tiles = _sig("orc_tile_schedule", None, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
              C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int))
_warp = _sig("orc_frame_warp", None, _S, _P, C.c_int)
_full = _sig("orc_frame_full", None, _S, _P)
_gated = _sig("orc_frame_warp_gated", None, _S, _P, C.c_int, C.c_float)
bitcount8 = _sig("orc_bitcount8", C.c_uint32, C.c_uint32)
colorize_word = _sig("orc_colorize_word", C.c_uint32, C.c_uint32)


def camera(yaw, pitch, roll, pos, depth, frame):
    p = Params()
    pc = (C.c_float * 3)(*[float(v) for v in pos])
    wr = (C.c_float * 3)()
    _camera(float(yaw), float(pitch), float(roll), pc, int(depth), int(frame), C.byref(p), wr)
    return p, tuple(wr)


def params_from(fp):
    """Copy an svo_amd.FrameParams (same layout) into oracle Params."""
    p = Params()
    C.memmove(C.byref(p), C.byref(fp), C.sizeof(Params))
    return p


def _build_with(fn, free, xyz, rgba, depth):
    xyz = np.ascontiguousarray(xyz, dtype=np.uint16).reshape(-1, 3)
    rgba = np.ascontiguousarray(rgba, dtype=np.uint32).reshape(-1)
    out = _u32p()
    n = C.c_size_t()
    root = C.c_uint32()
    rc = fn(xyz.ctypes.data, rgba.ctypes.data, len(rgba), int(depth), C.byref(out), C.byref(n),
            C.byref(root))
    if rc != 0:
        raise RuntimeError(f"octree build failed ({rc})")
    a = np.ctypeslib.as_array(out, (n.value,)).copy()
    free(C.cast(out, C.c_void_p))
    return a, root.value


def build_octree(xyz, rgba, depth):
    """Oracle restatement of set_voxel + convert_tree_blocks."""
    return _build_with(_build, _free, xyz, rgba, depth)


_ref_lib = None


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libref_octree.so"))


def ref_build_octree(xyz, rgba, depth):
    """The reference's own octree.h, compiled on the host (oracle/_ref)."""
    global _ref_lib
    if _ref_lib is None:
        _ref_lib = _load(os.path.join(HERE, "_ref", "libref_octree.so"))
        _ref_lib.ref_build_octree.restype = C.c_int
        _ref_lib.ref_build_octree.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                              C.POINTER(_u32p), C.POINTER(C.c_size_t),
                                              C.POINTER(C.c_uint32)]
        _ref_lib.ref_free.argtypes = [C.c_void_p]
    return _build_with(_ref_lib.ref_build_octree, _ref_lib.ref_free, xyz, rgba, depth)


def tile_schedule(res_x, res_y, frame):
    v = [C.c_int() for _ in range(4)]
    _tiles(res_x, res_y, frame, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def trace_pixel(words, root, depth, res_x, res_y, p, x, y):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    xyz = (C.c_float * 3)()
    st = Stats()
    w = _trace(words.ctypes.data_as(_u32p), len(words), root, depth, res_x, res_y, C.byref(p), x, y,
               xyz, C.byref(st))
    return w, tuple(xyz), st


def trace_steps(words, root, depth, res_x, res_y, p, x, y, max_steps=1 << 17):
    """The same ray as trace_pixel, with every DDA step: int32 [n, 9] rows of
    (x, y, z before the step, x, y, z after, level, XOR level + 1, ends by
    distance), positions in the octant-folded frame; the step pops to the
    XOR level + 1 when that is >= its level and it does not end the ray.
    Returns (word, xyz, steps, stats)."""
    words = np.ascontiguousarray(words, dtype=np.uint32)
    xyz = (C.c_float * 3)()
    log = np.zeros((max_steps, 9), np.int32)
    n = C.c_int(0)
    st = Stats()
    w = _trace_pops(words.ctypes.data_as(_u32p), len(words), root, depth, res_x, res_y, C.byref(p), x, y, xyz,
                    log.ctypes.data_as(C.POINTER(C.c_int32)), max_steps, C.byref(n), C.byref(st))
    return w, tuple(xyz), log[:n.value].copy(), st


def trace_pixel_pops(words, root, depth, res_x, res_y, p, x, y, max_pops=1 << 17):
    """The pops of the ray: int32 [n, 8] rows of (x, y, z before the step,
    x, y, z after, level before, level popped to)."""
    w, xyz, steps, _ = trace_steps(words, root, depth, res_x, res_y, p, x, y, max_pops)
    popped = (steps[:, 7] >= steps[:, 6]) & (steps[:, 8] == 0)
    return w, xyz, steps[popped, :8].copy()


class Frame:
    """Oracle frame state: the reference's buffers (raycast.h:79-85)."""

    def __init__(self, res_x, res_y, words, root, depth, nthreads=1):
        self.words = np.ascontiguousarray(words, dtype=np.uint32)
        self.s = State()
        rc = _init(C.byref(self.s), res_x, res_y, self.words.ctypes.data_as(_u32p), len(self.words),
                   root, depth)
        if rc != 0:
            raise MemoryError("oracle state")
        self.s.nthreads = int(nthreads)
        self.res_x, self.res_y = res_x, res_y
        self.N = res_x * res_y
        self.S = (res_x // 16) * (res_y // 16)

    def close(self):
        if self.s.screen:
            _sfree(C.byref(self.s))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # views (live)
    @property
    def screen(self):
        return np.ctypeslib.as_array(self.s.screen, (4, self.res_y, self.res_x))

    @property
    def coords(self):
        return np.ctypeslib.as_array(self.s.coords, (4, self.res_y, self.res_x, 4))

    @property
    def ids(self):
        return np.ctypeslib.as_array(self.s.ids, (2 * self.S + self.N,))

    @property
    def image(self):
        return np.ctypeslib.as_array(self.s.image, (self.res_y, self.res_x))

    @property
    def stats(self):
        st = self.s.st
        d = {k: getattr(st, k) for k, _ in Stats._fields_ if k != "desc_by_rek"}
        d["desc_by_rek"] = list(st.desc_by_rek)
        return d

    # kernels
    def memset(self, word_ofs, val, n_words):
        _memset(C.byref(self.s), word_ofs, val, n_words)

    def proj(self, plane, p):
        _proj(C.byref(self.s), plane, C.byref(p))

    def counthole(self):
        _count(C.byref(self.s))

    def sumids(self):
        _sum(C.byref(self.s))

    def writeids(self):
        _write(C.byref(self.s))

    def raycast_holes(self, p):
        _holes(C.byref(self.s), C.byref(p))

    def raycast_rect(self, p, ax, ay, ex, ey):
        _rect(C.byref(self.s), C.byref(p), ax, ay, ex, ey)

    def raycast_quads(self, p, ax, ay):
        _quads(C.byref(self.s), C.byref(p), ax, ay)

    def fillhole_punch(self, motion_x=None, motion_y=None):
        """raycast_fillhole on plane 0; motion arrays of N int32 or None (zero)."""
        args = []
        keep = []
        for m in (motion_x, motion_y):
            if m is None:
                args.append(None)
            else:
                a = np.ascontiguousarray(m, dtype=np.int32).reshape(-1)
                if a.size != self.N:
                    raise ValueError("motion buffer must hold one int per pixel")
                keep.append(a)
                args.append(a.ctypes.data_as(C.POINTER(C.c_int32)))
        _punch(C.byref(self.s), *args)

    def copy_plane(self, dst):
        _copy(C.byref(self.s), dst)

    def fillhole_snapshot(self):
        snap = np.ascontiguousarray(self.screen[0]).copy()
        self._snap = snap
        _fill_snap(C.byref(self.s), snap.ctypes.data_as(_u32p))

    def fillhole_inplace(self):
        _fill_inpl(C.byref(self.s))

    def colorize(self):
        _colorize(C.byref(self.s))

    def frame_warp(self, p, inplace_filter=False):
        _warp(C.byref(self.s), C.byref(p), 1 if inplace_filter else 0)

    def frame_full(self, p):
        _full(C.byref(self.s), C.byref(p))

    def frame_warp_gated(self, p, max_holes, max_depth=0.0):
        _gated(C.byref(self.s), C.byref(p), int(max_holes), float(max_depth))

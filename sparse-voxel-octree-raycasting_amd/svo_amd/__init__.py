"""ctypes binding of libsvo.so -- the MI355X SVO raycast + image-warp pipeline.

Host-side mirror of the reference's per-frame interface (src/raycast.h driving
kernel/kernel.cl through src/ocl.h).  Every call goes through the C ABI
declared in include/svo.h; nothing here computes pixels on the CPU.  If the
compiled library is missing, importing this module raises ImportError: there
is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SVO_LIB", os.path.join(_HERE, "libsvo.so"))
if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libsvo.so not found at {LIB_PATH}; build it with "
        "`python sparse-voxel-octree-raycasting_amd/build.py`")
_lib = C.CDLL(LIB_PATH)

HOLE = 0xFFFFFF00
BUF_SCREEN, BUF_COORDS, BUF_IDS, BUF_IMAGE, BUF_OCTREE = range(5)

_u32p = C.POINTER(C.c_uint32)
_u16p = C.POINTER(C.c_uint16)
_vp = C.c_void_p


class FrameParams(C.Structure):
    """svo_frame_params: kernel arguments of one frame (kernel.cl:479-482,
    :609-614, :858-863)."""
    _fields_ = [("m0", C.c_float * 4),
                ("row0", C.c_float * 4), ("row1", C.c_float * 4), ("row2", C.c_float * 4),
                ("col0", C.c_float * 4), ("col1", C.c_float * 4), ("col2", C.c_float * 4),
                ("fovx", C.c_float), ("fovy", C.c_float),
                ("frame", C.c_int32), ("reserved", C.c_int32)]


class SvoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"svo error {code}: {msg}")
        self.code = code


def _sig(name, restype, *argtypes):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_last_error = _sig("svo_last_error", C.c_char_p)


def _check(rc):
    if rc < 0:
        raise SvoError(rc, _last_error().decode(errors="replace"))
    return rc


_ctxp = C.c_void_p
_f = {
    "create": _sig("svo_create", C.c_int, C.POINTER(_ctxp), C.c_int),
    "destroy": _sig("svo_destroy", C.c_int, _ctxp),
    "upload_octree": _sig("svo_upload_octree", C.c_int, _ctxp, _u32p, C.c_size_t, C.c_uint32, C.c_int),
    "resize": _sig("svo_resize", C.c_int, _ctxp, C.c_int, C.c_int),
    "frame_warp": _sig("svo_frame_warp", C.c_int, _ctxp, C.POINTER(FrameParams)),
    "frame_full": _sig("svo_frame_full", C.c_int, _ctxp, C.POINTER(FrameParams)),
    "memset": _sig("svo_memset", C.c_int, _ctxp, C.c_size_t, C.c_uint32, C.c_size_t),
    "proj": _sig("svo_proj", C.c_int, _ctxp, C.c_int, C.POINTER(FrameParams)),
    "set_gate_depth": _sig("svo_set_gate_depth", C.c_int, _ctxp, C.c_float),
    "proj_noalloc": _sig("svo_proj_noalloc", C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int,
                         C.POINTER(FrameParams), _vp),
    "counthole": _sig("svo_counthole", C.c_int, _ctxp),
    "sumids": _sig("svo_sumids", C.c_int, _ctxp),
    "writeids": _sig("svo_writeids", C.c_int, _ctxp),
    "raycast_holes": _sig("svo_raycast_holes", C.c_int, _ctxp, C.POINTER(FrameParams)),
    "raycast_rect": _sig("svo_raycast_rect", C.c_int, _ctxp, C.POINTER(FrameParams),
                         C.c_int, C.c_int, C.c_int, C.c_int),
    "copy_cache": _sig("svo_copy_cache", C.c_int, _ctxp, C.c_int),
    "raycast_quads": _sig("svo_raycast_quads", C.c_int, _ctxp, C.POINTER(FrameParams), C.c_int, C.c_int),
    "fillhole": _sig("svo_fillhole", C.c_int, _ctxp),
    "colorize": _sig("svo_colorize", C.c_int, _ctxp),
    "read": _sig("svo_read", C.c_int, _ctxp, C.c_int, C.c_size_t, C.c_size_t, _vp),
    "write": _sig("svo_write", C.c_int, _ctxp, C.c_int, C.c_size_t, C.c_size_t, _vp),
    "sync": _sig("svo_sync", C.c_int, _ctxp),
    "hole_count": _sig("svo_hole_count", C.c_int, _ctxp, _u32p),
    "ray_count": _sig("svo_ray_count", C.c_int, _ctxp, C.POINTER(C.c_uint64), C.c_int),
    "ray_stats": _sig("svo_ray_stats", C.c_int, _ctxp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.c_int),
    "set_pretrace": _sig("svo_set_pretrace", C.c_int, _ctxp, C.c_void_p),
    "get_pretrace": _sig("svo_get_pretrace", C.c_int, _ctxp, C.c_void_p),
    "frame_intervals": _sig("svo_frame_intervals", C.c_int, _ctxp, C.POINTER(C.c_float), C.c_int,
                            C.POINTER(C.c_int)),
    "pretrace_stats": _sig("svo_pretrace_stats", C.c_int, _ctxp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                           C.POINTER(C.c_uint32), C.c_int),
    "enable_timing": _sig("svo_enable_timing", C.c_int, _ctxp, C.c_int),
    "stage_times": _sig("svo_stage_times", C.c_int, _ctxp, C.POINTER(C.c_float), C.c_int),
    "stage_name": _sig("svo_stage_name", C.c_char_p, C.c_int),
    "frame_time": _sig("svo_frame_time", C.c_int, _ctxp, C.POINTER(C.c_float)),
    "set_graph": _sig("svo_set_graph", C.c_int, _ctxp, C.c_int),
    "set_tuning": _sig("svo_set_tuning", C.c_int, _ctxp, C.c_int, C.c_int, C.c_int, C.c_int),
    "set_ray_spread": _sig("svo_set_ray_spread", C.c_int, _ctxp, C.c_int, C.c_int),
    "set_ray_lanes": _sig("svo_set_ray_lanes", C.c_int, _ctxp, C.c_int),
    "set_long_first": _sig("svo_set_long_first", C.c_int, _ctxp, C.c_int),
    "set_ray_block": _sig("svo_set_ray_block", C.c_int, _ctxp, C.c_int),
    "set_cnn_overlap": _sig("svo_set_cnn_overlap", C.c_int, _ctxp, C.c_int),
    "read_image_async": _sig("svo_read_image_async", C.c_int, _ctxp, _u32p),
    "host_alloc": _sig("svo_host_alloc", C.c_int, C.c_size_t, C.POINTER(_vp)),
    "host_free": _sig("svo_host_free", C.c_int, _vp),
    "fillhole_punch": _sig("svo_fillhole_punch", C.c_int, _ctxp, _vp, _vp),
    "camera": _sig("svo_camera", C.c_int, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float),
                   C.c_int, C.c_int, C.POINTER(FrameParams), C.POINTER(C.c_float)),
    "tile_schedule": _sig("svo_tile_schedule", C.c_int, C.c_int, C.c_int, C.c_int,
                          C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                          C.POINTER(C.c_int)),
    "build_octree": _sig("svo_build_octree", C.c_int, _u16p, _u32p, C.c_size_t, C.c_int,
                         C.POINTER(_u32p), C.POINTER(C.c_size_t), C.POINTER(C.c_uint32)),
    "free": _sig("svo_free", None, _vp),
    "build_octree_gpu": _sig("svo_build_octree_gpu", C.c_int, C.c_int, _u16p, _u32p, C.c_size_t, C.c_int,
                             C.POINTER(_u32p), C.POINTER(C.c_size_t), C.POINTER(C.c_uint32)),
    "octree_save": _sig("svo_octree_save", C.c_int, C.c_char_p, _u32p, C.c_size_t, C.c_uint32, C.c_int),
    "octree_load": _sig("svo_octree_load", C.c_int, C.c_char_p, C.POINTER(_u32p),
                        C.POINTER(C.c_size_t), C.POINTER(C.c_uint32), C.POINTER(C.c_int)),
    "scene_generate_gpu": _sig("svo_scene_generate_gpu", C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int,
                               C.POINTER(_u16p), C.POINTER(_u32p), C.POINTER(C.c_size_t)),
    "scene_generate": _sig("svo_scene_generate", C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int,
                           C.POINTER(_u16p), C.POINTER(_u32p), C.POINTER(C.c_size_t)),
    "cnn_create": _sig("svo_cnn_create", C.c_int, C.POINTER(_ctxp), C.c_int, C.c_uint64),
    "cnn_destroy": _sig("svo_cnn_destroy", C.c_int, _ctxp),
    "cnn_param_count": _sig("svo_cnn_param_count", C.c_size_t),
    "cnn_get_params": _sig("svo_cnn_get_params", C.c_int, _ctxp, C.POINTER(C.c_float), C.c_size_t),
    "cnn_set_params": _sig("svo_cnn_set_params", C.c_int, _ctxp, C.POINTER(C.c_float), C.c_size_t),
    "cnn_forward": _sig("svo_cnn_forward", C.c_int, _ctxp, _u16p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)),
    "cnn_gradients": _sig("svo_cnn_gradients", C.c_int, _ctxp, _u16p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                          C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_float)),
    "cnn_train_step": _sig("svo_cnn_train_step", C.c_int, _ctxp, _ctxp, _ctxp, C.POINTER(FrameParams), C.c_int,
                           C.c_int, C.c_uint64, C.c_float, C.POINTER(C.c_float)),
    "cnn_forward_tiles": _sig("svo_cnn_forward_tiles", C.c_int, _ctxp, _u16p, C.c_int, C.POINTER(C.c_float)),
    "cnn_set_fused": _sig("svo_cnn_set_fused", C.c_int, _ctxp, C.c_int),
    "gate_counts": _sig("svo_gate_counts", C.c_int, _ctxp, _u32p, _u32p),
    "frame_warp_learned": _sig("svo_frame_warp_learned", C.c_int, _ctxp, _ctxp, C.POINTER(FrameParams), C.c_int),
    "batch_create": _sig("svo_batch_create", C.c_int, C.POINTER(_ctxp), C.POINTER(_ctxp), C.c_int),
    "batch_destroy": _sig("svo_batch_destroy", C.c_int, _ctxp),
    "batch_frame_warp": _sig("svo_batch_frame_warp", C.c_int, _ctxp, C.POINTER(FrameParams)),
    "batch_sync": _sig("svo_batch_sync", C.c_int, _ctxp),
    "batch_frame_warp_learned": _sig("svo_batch_frame_warp_learned", C.c_int, _ctxp, _ctxp, C.POINTER(FrameParams),
                                     C.c_int),
    "rle4_load": _sig("svo_rle4_load", C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                      C.POINTER(_u16p), C.POINTER(_u32p), C.POINTER(C.c_size_t)),
    "rle4_save": _sig("svo_rle4_save", C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int32)),
    "image_save_ppm": _sig("svo_image_save_ppm", C.c_int, C.c_char_p, _u32p, C.c_int, C.c_int),
    "image_save_png": _sig("svo_image_save_png", C.c_int, C.c_char_p, _u32p, C.c_int, C.c_int),
    "camera_path": _sig("svo_camera_path", C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_float,
                        C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float)),
}

# every symbol declared in include/svo.h
EXPORTED = [
    "svo_last_error", "svo_create", "svo_destroy", "svo_upload_octree", "svo_resize",
    "svo_frame_warp", "svo_frame_full", "svo_memset", "svo_proj", "svo_counthole", "svo_sumids",
    "svo_writeids", "svo_raycast_holes", "svo_raycast_rect", "svo_copy_cache", "svo_fillhole",
    "svo_colorize", "svo_read", "svo_write", "svo_sync", "svo_hole_count", "svo_camera",
    "svo_tile_schedule", "svo_build_octree", "svo_free", "svo_octree_save", "svo_octree_load",
    "svo_scene_generate", "svo_camera_path", "svo_enable_timing", "svo_stage_times",
    "svo_stage_name", "svo_set_tuning", "svo_ray_count",
    "svo_cnn_create", "svo_cnn_destroy", "svo_cnn_param_count", "svo_cnn_get_params", "svo_cnn_set_params",
    "svo_cnn_forward", "svo_cnn_gradients", "svo_cnn_train_step", "svo_frame_warp_learned",
    "svo_batch_create", "svo_batch_destroy", "svo_batch_frame_warp", "svo_batch_sync",
    "svo_frame_time", "svo_set_graph", "svo_rle4_load", "svo_rle4_save", "svo_image_save_ppm",
    "svo_build_octree_gpu", "svo_cnn_forward_tiles", "svo_cnn_set_fused", "svo_set_ray_spread",
    "svo_gate_counts", "svo_batch_frame_warp_learned", "svo_scene_generate_gpu",
    "svo_set_ray_lanes", "svo_set_long_first", "svo_set_ray_block", "svo_raycast_quads", "svo_fillhole_punch",
    "svo_image_save_png", "svo_set_cnn_overlap", "svo_read_image_async", "svo_host_alloc", "svo_host_free",
    "svo_proj_noalloc", "svo_set_gate_depth", "svo_ray_stats", "svo_set_pretrace", "svo_pretrace_stats",
    "svo_get_pretrace", "svo_frame_intervals",
]


class PretraceConfig(C.Structure):
    """svo_pretrace_config"""
    _fields_ = [("sets", C.c_int), ("long_trips", C.c_int), ("wg_per_cu", C.c_int), ("split_trips", C.c_int),
                ("split_lanes", C.c_int), ("only_queued", C.c_int), ("predict", C.c_int), ("whole_trips", C.c_int),
                ("queue_skip", C.c_int)]


def _take_u32(p, n):
    a = np.ctypeslib.as_array(p, (n,)).copy()
    _f["free"](C.cast(p, _vp))
    return a


# ------------------------------------------------------------------ host helpers
def camera(yaw, pitch, roll, pos, depth, frame):
    """Camera vectors for one frame (raycast.h:113-162, :322-326).
    Returns (FrameParams, wrapped position)."""
    p = FrameParams()
    pos_c = (C.c_float * 3)(*[float(v) for v in pos])
    wrapped = (C.c_float * 3)()
    _check(_f["camera"](float(yaw), float(pitch), float(roll), pos_c, int(depth), int(frame),
                        C.byref(p), wrapped))
    return p, tuple(wrapped)


def tile_schedule(res_x, res_y, frame):
    """(add_x, add_y, ext_x, ext_y) of the 8x4 refresh tile (raycast.h:363-366)."""
    v = [C.c_int() for _ in range(4)]
    _check(_f["tile_schedule"](res_x, res_y, frame, *[C.byref(x) for x in v]))
    return tuple(x.value for x in v)


def build_octree(xyz, rgba, depth):
    """set_voxel stream -> compact word array + root ref (octree.h:32-293)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.uint16).reshape(-1, 3)
    rgba = np.ascontiguousarray(rgba, dtype=np.uint32).reshape(-1)
    if len(xyz) != len(rgba):
        raise ValueError("xyz and rgba lengths differ")
    out = _u32p()
    n = C.c_size_t()
    root = C.c_uint32()
    _check(_f["build_octree"](xyz.ctypes.data_as(_u16p), rgba.ctypes.data_as(_u32p), len(rgba),
                              int(depth), C.byref(out), C.byref(n), C.byref(root)))
    return _take_u32(out, n.value), root.value


def build_octree_gpu(xyz, rgba, depth, device=0):
    """build_octree on the GPU: identical words and root ref."""
    xyz = np.ascontiguousarray(xyz, dtype=np.uint16).reshape(-1, 3)
    rgba = np.ascontiguousarray(rgba, dtype=np.uint32).reshape(-1)
    if len(xyz) != len(rgba):
        raise ValueError("xyz and rgba lengths differ")
    out = _u32p()
    n = C.c_size_t()
    root = C.c_uint32()
    _check(_f["build_octree_gpu"](int(device), xyz.ctypes.data_as(_u16p), rgba.ctypes.data_as(_u32p), len(rgba),
                                  int(depth), C.byref(out), C.byref(n), C.byref(root)))
    return _take_u32(out, n.value), root.value


def octree_save(path, words, root, depth):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    _check(_f["octree_save"](os.fsencode(path), words.ctypes.data_as(_u32p), len(words), root, depth))


def octree_load(path):
    out = _u32p()
    n = C.c_size_t()
    root = C.c_uint32()
    depth = C.c_int()
    _check(_f["octree_load"](os.fsencode(path), C.byref(out), C.byref(n), C.byref(root), C.byref(depth)))
    return _take_u32(out, n.value), root.value, depth.value


def scene(kind, seed, depth, n_spheres=0):
    """Seeded synthetic voxel stream: returns (xyz uint16 [n,3], rgba uint32 [n])."""
    xyz = _u16p()
    rgba = _u32p()
    n = C.c_size_t()
    _check(_f["scene_generate"](int(kind), int(seed), int(depth), int(n_spheres), C.byref(xyz),
                                C.byref(rgba), C.byref(n)))
    nv = n.value
    a = np.ctypeslib.as_array(xyz, (nv * 3,)).copy().reshape(-1, 3) if nv else np.zeros((0, 3), np.uint16)
    b = np.ctypeslib.as_array(rgba, (nv,)).copy() if nv else np.zeros(0, np.uint32)
    _f["free"](C.cast(xyz, _vp))
    _f["free"](C.cast(rgba, _vp))
    return a, b


def proj_noalloc(device, screen_ptr, coords_ptr, res_x, res_y, plane, p, stream=0):
    """svo_proj_noalloc on caller-owned device buffers (integer device
    addresses: 4 planes of N uint32 screen words, 4 planes of N float4);
    asynchronous on `stream` (0 = the null stream)."""
    _check(_f["proj_noalloc"](int(device), _vp(int(screen_ptr)), _vp(int(coords_ptr)), int(res_x), int(res_y),
                              int(plane), C.byref(p), _vp(int(stream)) if stream else None))


def scene_gpu(kind, seed, depth, device=0):
    """scene() for the caves (kinds 3 and 4) generated on the GPU: the same stream."""
    xyz = _u16p()
    rgba = _u32p()
    n = C.c_size_t()
    _check(_f["scene_generate_gpu"](int(device), int(kind), int(seed), int(depth), C.byref(xyz), C.byref(rgba),
                                    C.byref(n)))
    return _take_stream(xyz, rgba, n)


def _take_stream(xyz, rgba, n):
    nv = n.value
    a = np.ctypeslib.as_array(xyz, (nv * 3,)).copy().reshape(-1, 3) if nv else np.zeros((0, 3), np.uint16)
    b = np.ctypeslib.as_array(rgba, (nv,)).copy() if nv else np.zeros(0, np.uint32)
    _f["free"](C.cast(xyz, _vp))
    _f["free"](C.cast(rgba, _vp))
    return a, b


def rle4_load(path, depth, palette=1, add=(0, 0, 0)):
    """Map 0 of an RLE4 volume file as a voxel stream (xyz uint16 [n,3],
    rgba uint32 [n]) in the reference loader's order and colour rule."""
    xyz = _u16p()
    rgba = _u32p()
    n = C.c_size_t()
    _check(_f["rle4_load"](str(path).encode(), int(palette), int(add[0]), int(add[1]), int(add[2]), int(depth),
                           C.byref(xyz), C.byref(rgba), C.byref(n)))
    return _take_stream(xyz, rgba, n)


def save_ppm(path, image):
    """Write a colorized image (uint32 [h, w], b + g<<8 + r<<16) as binary PPM."""
    img = np.ascontiguousarray(image, dtype=np.uint32)
    h, w = img.shape
    _check(_f["image_save_ppm"](str(path).encode(), img.ctypes.data_as(_u32p), int(w), int(h)))


class PinnedImage:
    """A page-locked host buffer (svo_host_alloc) viewed as a uint32 image
    [res_y, res_x], for Context.read_image_async; free with close()."""

    def __init__(self, res_x, res_y):
        self._p = _vp()
        _check(_f["host_alloc"](4 * int(res_x) * int(res_y), C.byref(self._p)))
        buf = (C.c_uint32 * (int(res_x) * int(res_y))).from_address(self._p.value)
        self.array = np.ctypeslib.as_array(buf).reshape(int(res_y), int(res_x))

    def close(self):
        if self._p:
            self.array = None
            _check(_f["host_free"](self._p))
            self._p = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def save_png(path, image):
    """Write a colorized image (uint32 [h, w], b + g<<8 + r<<16) as 8-bit RGB PNG."""
    img = np.ascontiguousarray(image, dtype=np.uint32)
    h, w = img.shape
    _check(_f["image_save_png"](str(path).encode(), img.ctypes.data_as(_u32p), int(w), int(h)))


def rle4_save(path, vol):
    """Write a one-map RLE4 file; vol is int32 [sz, sx, sy] (file rows along
    the last axis), < 0 = empty, else the 16-bit colour word."""
    vol = np.ascontiguousarray(vol, dtype=np.int32)
    sz, sx, sy = vol.shape
    _check(_f["rle4_save"](str(path).encode(), int(sx), int(sy), int(sz),
                           vol.ctypes.data_as(C.POINTER(C.c_int32))))


def camera_path(seed, depth, n_frames, max_voxels_per_frame, max_yaw_deg_per_frame,
                height_voxels, pitch):
    """Seeded Catmull-Rom flythrough: float32 [n_frames, 6] rows
    (yaw, pitch, roll, x, y, z) in world units."""
    out = np.zeros((n_frames, 6), np.float32)
    _check(_f["camera_path"](int(seed), int(depth), int(n_frames), float(max_voxels_per_frame),
                             float(max_yaw_deg_per_frame), float(height_voxels), float(pitch),
                             out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


# ------------------------------------------------------------------ context
class Context:
    """One GPU context: octree, frame buffers and a HIP stream
    (the reference's ocl_init + raycast_init state)."""

    def __init__(self, device=0):
        self._h = _ctxp()
        _check(_f["create"](C.byref(self._h), int(device)))
        self.res_x = self.res_y = 0
        self.depth = 0

    # lifecycle
    def close(self):
        if self._h:
            _f["destroy"](self._h)
            self._h = _ctxp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def N(self):
        return self.res_x * self.res_y

    @property
    def S(self):
        return (self.res_x // 16) * (self.res_y // 16)

    def upload_octree(self, words, root, depth):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        _check(_f["upload_octree"](self._h, words.ctypes.data_as(_u32p), len(words), int(root), int(depth)))
        self.depth = int(depth)

    def resize(self, res_x, res_y):
        _check(_f["resize"](self._h, int(res_x), int(res_y)))
        self.res_x, self.res_y = int(res_x), int(res_y)

    # frames
    def frame_warp(self, p):
        _check(_f["frame_warp"](self._h, C.byref(p)))

    def frame_full(self, p):
        _check(_f["frame_full"](self._h, C.byref(p)))

    def frame_warp_learned(self, cnn, p, max_cnn_holes=64):
        """Warp frame with the CNN fill and per-tile gate."""
        _check(_f["frame_warp_learned"](self._h, cnn._h if cnn is not None else None, C.byref(p),
                                        int(max_cnn_holes)))

    def gate_counts(self):
        """(CNN tiles, gate rays) of the last learned frame."""
        a, b = C.c_uint32(), C.c_uint32()
        _check(_f["gate_counts"](self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # per-stage (reference kernel names)
    def memset(self, byte_ofs, val, nbytes):
        _check(_f["memset"](self._h, int(byte_ofs), int(val) & 0xFFFFFFFF, int(nbytes)))

    def proj(self, plane, p):
        _check(_f["proj"](self._h, int(plane), C.byref(p)))

    def counthole(self):
        _check(_f["counthole"](self._h))

    def sumids(self):
        _check(_f["sumids"](self._h))

    def writeids(self):
        _check(_f["writeids"](self._h))

    def raycast_holes(self, p):
        _check(_f["raycast_holes"](self._h, C.byref(p)))

    def raycast_rect(self, p, add_x, add_y, ext_x, ext_y):
        _check(_f["raycast_rect"](self._h, C.byref(p), int(add_x), int(add_y), int(ext_x), int(ext_y)))

    def raycast_quads(self, p, add_x, add_y):
        """raycast_fine: one ray per 2x2 quad of the (res/4)-rounded grid."""
        _check(_f["raycast_quads"](self._h, C.byref(p), int(add_x), int(add_y)))

    def fillhole_punch(self, motion_x=None, motion_y=None):
        """raycast_fillhole on plane 0; motion arrays of N int32 each or None
        (zero motion: no pixel changes)."""
        ptrs, keep = [], []
        for m in (motion_x, motion_y):
            if m is None:
                ptrs.append(None)
                continue
            a = np.ascontiguousarray(m, dtype=np.int32).reshape(-1)
            if a.size != self.N:
                raise ValueError("motion buffer must hold one int per pixel")
            keep.append(a)
            ptrs.append(a.ctypes.data_as(_vp))
        _check(_f["fillhole_punch"](self._h, *ptrs))

    def copy_cache(self, dst_plane):
        _check(_f["copy_cache"](self._h, int(dst_plane)))

    def fillhole(self):
        _check(_f["fillhole"](self._h))

    def colorize(self):
        _check(_f["colorize"](self._h))

    def sync(self):
        _check(_f["sync"](self._h))

    def hole_count(self):
        v = C.c_uint32()
        _check(_f["hole_count"](self._h, C.byref(v)))
        return v.value

    def set_tuning(self, refill=-1, wg_per_cu=-1, iters=-1, lds_cache_words=-1):
        """Ray-kernel knobs; -1 keeps the current value."""
        _check(_f["set_tuning"](self._h, int(refill), int(wg_per_cu), int(iters), int(lds_cache_words)))

    def set_ray_spread(self, group=1, only_if_fits=True):
        """Ray hand-out order: groups of `group` consecutive rays spread across
        waves (64 = list order); only_if_fits limits it to launches whose rays
        all fit in the resident lanes."""
        _check(_f["set_ray_spread"](self._h, int(group), 1 if only_if_fits else 0))

    def set_ray_lanes(self, lanes=64):
        """Lanes per wave that take rays (64 = all)."""
        _check(_f["set_ray_lanes"](self._h, int(lanes)))

    def set_ray_block(self, width=8):
        """Rectangle rays are handed out in blocks of width x 64/width pixels
        (8, 16, 32 or 64).  Results are the same for every shape."""
        _check(_f["set_ray_block"](self._h, int(width)))

    def set_long_first(self, long_trips=1024):
        """Full frames raycast first the 8x8 blocks whose previous full frame
        had a ray of at least long_trips trips (0 = raster order only).
        Results are the same either way."""
        _check(_f["set_long_first"](self._h, int(long_trips)))

    def set_gate_depth(self, max_depth):
        """Learned frames: raycast whole every tile holding a reprojected
        pixel deeper than max_depth world units (0 = off)."""
        _check(_f["set_gate_depth"](self._h, float(max_depth)))

    def set_cnn_overlap(self, workgroups):
        """Workgroups of the CNN tile kernel while it runs alongside the
        ray launch in learned frames (0 = two per CU).  Results are the same
        either way."""
        _check(_f["set_cnn_overlap"](self._h, int(workgroups)))

    def ray_count(self, reset=False):
        """Rays traced by the ray kernels since the last reset."""
        v = C.c_uint64()
        _check(_f["ray_count"](self._h, C.byref(v), 1 if reset else 0))
        return v.value

    def set_pretrace(self, sets=None, long_trips=None, wg_per_cu=None, split_trips=None, split_lanes=None,
                     only_queued=None, predict=None, whole_trips=None, queue_skip=None):
        """Trace each warp frame's refresh rectangle and long-ray blocks
        (predict = 0) or predicted hole quads (predict = k: from the frame k
        frames back; blocks with a known ray of at least whole_trips trips
        whole, at once) ahead on streams of their own (sets = 0: off;
        queue_skip leaves every fourth stream of its pool out);
        arguments left None keep their current values.  Results are the
        same."""
        cfg = self.get_pretrace()
        for k, v in (("sets", sets), ("long_trips", long_trips), ("wg_per_cu", wg_per_cu),
                     ("split_trips", split_trips), ("split_lanes", split_lanes), ("only_queued", only_queued),
                     ("predict", predict), ("whole_trips", whole_trips), ("queue_skip", queue_skip)):
            if v is not None:
                setattr(cfg, k, int(v))
        _check(_f["set_pretrace"](self._h, C.byref(cfg)))

    def get_pretrace(self):
        cfg = PretraceConfig()
        _check(_f["get_pretrace"](self._h, C.byref(cfg)))
        return cfg

    def pretrace_stats(self, reset=False):
        """(rays traced ahead, frame rays that took their result, loop trips
        of the longest ray traced ahead) since the last reset."""
        r = C.c_uint64()
        m = C.c_uint64()
        t = C.c_uint32()
        _check(_f["pretrace_stats"](self._h, C.byref(r), C.byref(m), C.byref(t), 1 if reset else 0))
        return r.value, m.value, t.value

    def ray_stats(self, reset=False):
        """(rays traced, loop trips of the longest of them) since the last
        reset."""
        v = C.c_uint64()
        m = C.c_uint32()
        _check(_f["ray_stats"](self._h, C.byref(v), C.byref(m), 1 if reset else 0))
        return v.value, m.value

    def enable_timing(self, on=True, mode=None):
        """on: per-stage timing (mode 1); mode=2: one event pair per frame."""
        _check(_f["enable_timing"](self._h, int(mode) if mode is not None else (1 if on else 0)))

    def frame_intervals(self, max_n=511):
        """Timing mode 3: ms between the completions of successive frames."""
        buf = (C.c_float * max(1, max_n))()
        n = C.c_int()
        _check(_f["frame_intervals"](self._h, buf, int(max_n), C.byref(n)))
        return np.array(buf[:n.value], dtype=np.float64)

    def frame_time(self):
        """GPU ms of the last frame (timing mode 1 or 2, after sync())."""
        v = C.c_float()
        _check(_f["frame_time"](self._h, C.byref(v)))
        return v.value

    def set_graph(self, on=True):
        """Run warp frames >= 2 as one captured HIP graph (default on)."""
        _check(_f["set_graph"](self._h, 1 if on else 0))

    def stage_times(self):
        buf = (C.c_float * 16)()
        n = _check(_f["stage_times"](self._h, buf, 16))
        return {_f["stage_name"](i).decode(): buf[i] for i in range(n)}

    # transfers
    def read(self, buf, byte_ofs, nbytes, dtype=np.uint32):
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        _check(_f["read"](self._h, int(buf), int(byte_ofs), int(nbytes), out.ctypes.data_as(_vp)))
        return out

    def write(self, buf, byte_ofs, arr):
        arr = np.ascontiguousarray(arr)
        _check(_f["write"](self._h, int(buf), int(byte_ofs), arr.nbytes, arr.ctypes.data_as(_vp)))

    def screen(self, plane=None):
        """uint32 words: [4, res_y, res_x] or one plane [res_y, res_x]."""
        if plane is None:
            return self.read(BUF_SCREEN, 0, 16 * self.N).reshape(4, self.res_y, self.res_x)
        return self.read(BUF_SCREEN, 4 * self.N * plane, 4 * self.N).reshape(self.res_y, self.res_x)

    def coords(self, plane=None):
        """float32 [.., res_y, res_x, 4] world hit coordinates."""
        if plane is None:
            return self.read(BUF_COORDS, 0, 64 * self.N, np.float32).reshape(4, self.res_y, self.res_x, 4)
        return self.read(BUF_COORDS, 16 * self.N * plane, 16 * self.N, np.float32).reshape(
            self.res_y, self.res_x, 4)

    def ids(self):
        return self.read(BUF_IDS, 0, 4 * (2 * self.S + self.N))

    def image(self):
        return self.read(BUF_IMAGE, 0, 4 * self.N).reshape(self.res_y, self.res_x)

    def read_image_async(self, out):
        """Start copying the last frame's image into `out` (uint32, res_y x
        res_x, C-contiguous; pinned memory lets the copy overlap the next
        frame).  `out` holds the image after sync(); keep it alive until then."""
        if out.dtype != np.uint32 or not out.flags["C_CONTIGUOUS"] or out.size != self.N:
            raise ValueError("out must be a C-contiguous uint32 array of res_x * res_y elements")
        _check(_f["read_image_async"](self._h, out.ctypes.data_as(_u32p)))


_fp = C.POINTER(C.c_float)


class Cnn:
    """Learned hole filler: BF16 MFMA CNN with its own training kernels."""

    def __init__(self, device=0, seed=1):
        self._h = _ctxp()
        _check(_f["cnn_create"](C.byref(self._h), int(device), int(seed)))
        self.n_params = int(_f["cnn_param_count"]())

    def close(self):
        if self._h:
            _f["cnn_destroy"](self._h)
            self._h = _ctxp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        _check(_f["cnn_get_params"](self._h, out.ctypes.data_as(_fp), self.n_params))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_f["cnn_set_params"](self._h, p.ctypes.data_as(_fp), len(p)))

    def forward(self, x_bf16):
        x = np.ascontiguousarray(x_bf16, dtype=np.uint16)
        B, H, W, Cc = x.shape
        out = np.zeros((B, H, W, 16), np.float32)
        _check(_f["cnn_forward"](self._h, x.ctypes.data_as(_u16p), B, H, W, out.ctypes.data_as(_fp)))
        return out

    def forward_tiles(self, x_bf16):
        """Fused tile network: x [n, 24, 24, 16] bf16 bits -> centre RGB [n, 16, 16, 3]."""
        x = np.ascontiguousarray(x_bf16, dtype=np.uint16)
        if x.shape[1:] != (24, 24, 16):
            raise ValueError("x must be [n, 24, 24, 16]")
        out = np.zeros((x.shape[0], 16, 16, 3), np.float32)
        _check(_f["cnn_forward_tiles"](self._h, x.ctypes.data_as(_u16p), x.shape[0], out.ctypes.data_as(_fp)))
        return out

    def set_fused(self, on=True, mode=None):
        """Frame fill with the fused tile kernel (default; mode 1: 4 waves
        staging each layer's weights, mode 2: 8 waves with resident weights)
        or, with on=False, the per-layer kernels."""
        _check(_f["cnn_set_fused"](self._h, int(mode) if mode is not None else (1 if on else 0)))

    def train_step(self, warp, full, p, n_crops=16, crop=64, seed=0, lr=1e-3, want_loss=False):
        """One Adam step on crops of the frame pair held by warp/full."""
        loss = C.c_float()
        _check(_f["cnn_train_step"](self._h, warp._h, full._h, C.byref(p), int(n_crops), int(crop),
                                    int(seed), float(lr), C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def gradients(self, x_bf16, target, mask):
        x = np.ascontiguousarray(x_bf16, dtype=np.uint16)
        t = np.ascontiguousarray(target, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.float32)
        B, H, W, _ = x.shape
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        _check(_f["cnn_gradients"](self._h, x.ctypes.data_as(_u16p), t.ctypes.data_as(_fp),
                                   m.ctypes.data_as(_fp), B, H, W, g.ctypes.data_as(_fp), self.n_params,
                                   C.byref(loss)))
        return g, loss.value


def image_rgb(img):
    """BGRX words (b + g<<8 + r<<16) -> uint8 [h, w, 3] RGB."""
    img = np.asarray(img, dtype=np.uint32)
    return np.stack([(img >> 16) & 255, (img >> 8) & 255, img & 255], axis=-1).astype(np.uint8)


def psnr(a, b):
    """PSNR in dB between two uint8 RGB images (inf when identical)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    mse = np.mean((a - b) ** 2)
    return float("inf") if mse == 0 else 10.0 * np.log10(255.0 ** 2 / mse)


# ------------------------------------------------------------------ batched views
class Batch:
    """Warp frames of several contexts (same octree, one device) run as one
    launch per stage and one shared persistent ray launch."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (_ctxp * len(self.contexts))(*[c._h for c in self.contexts])
        self._h = _ctxp()
        _check(_f["batch_create"](C.byref(self._h), arr, len(self.contexts)))

    def frame_warp(self, params):
        if len(params) != len(self.contexts):
            raise ValueError("one FrameParams per view")
        arr = (FrameParams * len(params))(*params)
        _check(_f["batch_frame_warp"](self._h, arr))

    def frame_warp_learned(self, cnn, params, max_cnn_holes=64):
        if len(params) != len(self.contexts):
            raise ValueError("one FrameParams per view")
        arr = (FrameParams * len(params))(*params)
        _check(_f["batch_frame_warp_learned"](self._h, cnn._h if cnn is not None else None, arr, int(max_cnn_holes)))

    def sync(self):
        _check(_f["batch_sync"](self._h))

    def close(self):
        if self._h:
            _f["batch_destroy"](self._h)
            self._h = _ctxp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

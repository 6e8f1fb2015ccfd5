"""svo_proj_noalloc (no counterpart in the reference; for callers that
cannot allocate): the warp of caller-owned device planes with the library's
static key array equals svo_proj on a context holding the same planes, bit
for bit, on two calls in a row (the second on a smaller frame), and sizes
past its limit are refused before any launch.

The test plays the caller: it owns its device buffers, allocated with the
HIP runtime libsvo.so itself links (libamdhip64.so.7), through ctypes."""
import ctypes as C

import numpy as np
import pytest

import svo_amd as S
from helpers import CONFIG1, assert_equal_f32_bits, assert_equal_u32, params, path, scene_octree

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Caller-owned device memory (hipMalloc / hipMemcpy / hipFree)."""
    _hip = None

    def __init__(self, host):
        if DeviceBuffer._hip is None:
            DeviceBuffer._hip = C.CDLL("libamdhip64.so.7")
        self.hip = DeviceBuffer._hip
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        h = np.ascontiguousarray(host)
        assert self.hip.hipMemcpy(self.ptr, h.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) == 0  # H2D

    def read(self, like):
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) == 0  # D2H
        return out

    def close(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _warp_once(words, root, res, frame):
    rx, ry = res
    pr = path(CONFIG1, frame + 1)
    with S.Context(0) as c:
        c.upload_octree(words, root, CONFIG1["depth"])
        c.resize(rx, ry)
        for f in range(frame):
            c.frame_warp(params(CONFIG1, pr, f))
        c.sync()
        scr, crd = c.screen().copy(), c.coords().copy()
        p = params(CONFIG1, pr, frame)
        # the library's own path on the context: plane 0 cleared, plane 2 warped
        c.memset(0, S.HOLE, 4 * c.N)
        c.proj(2, p)
        want_s, want_c = c.screen(), c.coords()
    scr[0] = S.HOLE
    ds, dc = DeviceBuffer(scr), DeviceBuffer(crd)
    try:
        print(f"svo_proj_noalloc {rx}x{ry} frame {frame}", flush=True)
        S.proj_noalloc(0, ds.ptr.value, dc.ptr.value, rx, ry, 2, p)
        assert_equal_u32(ds.read(scr), want_s, f"{rx}x{ry} screen")
        assert_equal_f32_bits(dc.read(crd), want_c, f"{rx}x{ry} coords")
    finally:
        ds.close()
        dc.close()
    assert (want_s[0] != S.HOLE).mean() > 0.5  # the warp filled most of plane 0


def test_noalloc_proj_matches_context_proj():
    words, root = scene_octree(CONFIG1["kind"], CONFIG1["seed"], CONFIG1["depth"])
    _warp_once(words, root, (512, 256), 5)
    _warp_once(words, root, (320, 240), 4)  # a second call on a smaller frame


def test_noalloc_proj_refuses_bad_sizes():
    d = DeviceBuffer(np.zeros(64, np.uint64))
    p = params(CONFIG1, path(CONFIG1, 1), 0)
    try:
        with pytest.raises(S.SvoError) as e:
            S.proj_noalloc(0, d.ptr.value, d.ptr.value, 4096, 1024, 2, p)  # 2^22 pixels > 2^21
        assert e.value.code == -1
        with pytest.raises(S.SvoError):
            S.proj_noalloc(0, d.ptr.value, d.ptr.value, 512, 256, 4, p)  # plane outside 1..3
    finally:
        d.close()

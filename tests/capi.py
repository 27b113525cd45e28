"""ctypes binding of libfpsx.so's C ABI (include/fpsx.h), as tests use it."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fpsample_amd", "libfpsx.so")
HEADER = os.path.join(ROOT, "include", "fpsx.h")

_sz = ctypes.c_size_t
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)

OK, EBADDIM, EBADSTART, EINVAL, ETOOBIG, EHIP, ENOMEM, ETIMEOUT = range(8)
ROWS, COLS = 0, 1
HIGHEST, LOWEST = 0, 1


class CApi:
    def __init__(self, path: str = LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (python fpsample_amd/build.py)")
        from fpsample_amd.device import hip

        hip()  # one HIP runtime for this library and a torch imported later
        L = ctypes.CDLL(path)
        L.fpsx_last_error.restype = ctypes.c_char_p
        L.fpsx_version.restype = ctypes.c_char_p
        L.fpsx_device_count.restype = ctypes.c_int
        L.fpsx_fps_sampling.argtypes = [_f32p, _sz, _sz, ctypes.c_int, _sz, _u64p, _sz, _u64p]
        L.fpsx_fps_sampling.restype = ctypes.c_int
        L.fpsx_fps_sampling_device.argtypes = [ctypes.c_void_p, _sz, _sz, _sz, _sz,
                                               ctypes.c_void_p, _sz, ctypes.c_void_p,
                                               ctypes.c_int, ctypes.c_float, ctypes.c_void_p]
        L.fpsx_fps_sampling_device.restype = ctypes.c_int
        L.fpsx_device_status.argtypes = [ctypes.c_void_p]
        L.fpsx_device_status.restype = ctypes.c_int
        L.bucket_fps_kdline.argtypes = [_f32p, _sz, _sz, _sz, _sz, _sz, _u64p]
        L.bucket_fps_kdline.restype = ctypes.c_int
        L.fpsx_kdline_permutation.argtypes = [_f32p, _sz, _sz, _sz, _u64p, _u64p, _u64p]
        L.fpsx_kdline_permutation.restype = ctypes.c_long
        L.fpsx_plan_info.argtypes = [_sz, _sz, ctypes.c_int, _sz, _ip, _ip, _ip, _ip, _ip]
        L.fpsx_plan_info.restype = ctypes.c_int
        L.fpsx_fps_npdu_sampling.argtypes = [_f32p, _sz, _sz, _sz, _sz, _sz, _u64p]
        L.fpsx_fps_npdu_sampling.restype = ctypes.c_int
        L.fpsx_fps_npdu_sampling_device.argtypes = [ctypes.c_void_p, _sz, _sz, _sz, _sz, _sz,
                                                    _sz, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_void_p]
        L.fpsx_fps_npdu_sampling_device.restype = ctypes.c_int
        L.fpsx_npdu_plan_info.argtypes = [_sz, _sz, _ip, _ip, _ip, _ip]
        L.fpsx_npdu_plan_info.restype = ctypes.c_int
        self.L = L

    def device_count(self) -> int:
        return int(self.L.fpsx_device_count())

    def last_error(self) -> str:
        return self.L.fpsx_last_error().decode()

    def fps(self, pc, m, starts, layout=COLS, n_points=None, dim=None):
        """(rc, picks) of fpsx_fps_sampling on a host array."""
        pc = np.asarray(pc, dtype=np.float32)
        arr = np.asfortranarray(pc) if layout == COLS else np.ascontiguousarray(pc)
        n = pc.shape[0] if n_points is None else n_points
        d = pc.shape[1] if dim is None else dim
        st = np.ascontiguousarray(np.atleast_1d(np.asarray(starts, dtype=np.uint64)))
        out = np.zeros(max(int(m), 1), dtype=np.uint64)
        rc = self.L.fpsx_fps_sampling(arr.ctypes.data_as(_f32p), n, d, layout, m,
                                      st.ctypes.data_as(_u64p), st.size,
                                      out.ctypes.data_as(_u64p))
        return rc, out[: int(m)]

    def npdu(self, pc, m, w, start, dim=None):
        """(rc, picks) of fpsx_fps_npdu_sampling on a row-major host array."""
        pc = np.ascontiguousarray(pc, dtype=np.float32)
        d = pc.shape[1] if dim is None else dim
        out = np.zeros(max(int(m), 1), dtype=np.uint64)
        rc = self.L.fpsx_fps_npdu_sampling(pc.ctypes.data_as(_f32p), pc.shape[0], d, m, w, start,
                                           out.ctypes.data_as(_u64p))
        return rc, out[: int(m)]

    def npdu_plan_info(self, n, d):
        v = [ctypes.c_int() for _ in range(4)]
        rc = self.L.fpsx_npdu_plan_info(n, d, *[ctypes.byref(x) for x in v])
        if rc != 0:
            raise RuntimeError(self.last_error())
        kdim, block, ppt, grid = (x.value for x in v)
        return dict(kernel_dim=kdim, block=block, ppt=ppt, grid=grid)

    def kdline(self, pc, m, start, h, dim=None):
        pc = np.ascontiguousarray(pc, dtype=np.float32)
        d = pc.shape[1] if dim is None else dim
        out = np.zeros(max(int(m), 1), dtype=np.uint64)
        rc = self.L.bucket_fps_kdline(pc.ctypes.data_as(_f32p), pc.shape[0], d, m, start, h,
                                      out.ctypes.data_as(_u64p))
        return rc, out[: int(m)]

    def kdline_perm(self, pc, h):
        pc = np.ascontiguousarray(pc, dtype=np.float32)
        n = pc.shape[0]
        perm = np.zeros(n, np.uint64)
        cap = min(n, 1 << min(h, 40))
        lo = np.zeros(cap, np.uint64)
        hi = np.zeros(cap, np.uint64)
        nl = self.L.fpsx_kdline_permutation(pc.ctypes.data_as(_f32p), n, pc.shape[1], h,
                                            perm.ctypes.data_as(_u64p), lo.ctypes.data_as(_u64p),
                                            hi.ctypes.data_as(_u64p))
        if nl < 0:
            raise RuntimeError(f"fpsx_kdline_permutation -> {nl}: {self.last_error()}")
        return perm, lo[:nl], hi[:nl]

    def plan_info(self, n, d, tie=HIGHEST, m=0):
        v = [ctypes.c_int() for _ in range(5)]
        rc = self.L.fpsx_plan_info(n, d, tie, m, *[ctypes.byref(x) for x in v])
        if rc != 0:
            raise RuntimeError(self.last_error())
        kind, block, ppt, grid, chunk = (x.value for x in v)
        return dict(kind=kind, block=block, ppt=ppt, grid=grid, chunk=chunk)


def header_functions(path: str = HEADER):
    """Names of the functions include/fpsx.h declares."""
    import re

    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)))

"""Device-resident entry points of libfpsx.so, usable without torch.

``DeviceArray`` is a minimal owner of HIP device memory (hipMalloc/hipMemcpy
through ctypes).  ``fps_sampling_device`` runs vanilla FPS on points already
in HBM, laid out SoA (coordinate j of point i at ``ptr[j * ld + i]``), and
leaves the picks in device memory; it is the form the benchmark times.  Any
object with an integer device address works as an argument (a DeviceArray, or
``tensor.data_ptr()`` of a torch tensor on the GPU).
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import Optional, Sequence, Union

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

TIE_HIGHEST = 0
TIE_LOWEST = 1
_H2D, _D2H = 1, 2

_hip = None
_fpsx = None


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetErrorString.restype = ctypes.c_char_p
    return _hip


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def lib() -> ctypes.CDLL:
    global _fpsx
    if _fpsx is None:
        L = ctypes.CDLL(os.path.join(HERE, "libfpsx.so"))
        sz, vp = ctypes.c_size_t, ctypes.c_void_p
        L.fpsx_last_error.restype = ctypes.c_char_p
        L.fpsx_fps_sampling_device.argtypes = [vp, sz, sz, sz, sz, vp, sz, vp, ctypes.c_int,
                                               ctypes.c_float, vp]
        L.fpsx_fps_sampling_device.restype = ctypes.c_int
        L.fpsx_device_status.argtypes = [vp]
        L.fpsx_timing.argtypes = [ctypes.c_int]
        L.fpsx_timing_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
        ip = ctypes.POINTER(ctypes.c_int)
        L.fpsx_plan_info.argtypes = [sz, sz, ctypes.c_int, ip, ip, ip, ip, ip]
        _fpsx = L
    return _fpsx


def _raise(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with error code {rc}: {lib().fpsx_last_error().decode()}")


class DeviceArray:
    """A block of HIP device memory holding a numpy-typed array."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(math.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _check(hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")
        return d

    def numpy(self) -> np.ndarray:
        a = np.empty(self.shape, self.dtype)
        _check(hip().hipMemcpy(a.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return a

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _addr(x) -> int:
    if isinstance(x, DeviceArray):
        return x.ptr
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    return int(x)


def fps_sampling_device(points, n_points: int, dim: int, n_samples: int, starts, n_starts: int,
                        out, ld: Optional[int] = None, tie: int = TIE_HIGHEST,
                        init: float = float("inf"), stream: Optional[int] = None) -> None:
    """Enqueue FPS on device data (see fpsx_fps_sampling_device in include/fpsx.h)."""
    rc = lib().fpsx_fps_sampling_device(_addr(points), ld if ld is not None else n_points,
                                        n_points, dim, n_samples, _addr(starts), n_starts,
                                        _addr(out), tie, init, stream)
    _raise(rc, "fpsx_fps_sampling_device")


def device_status(stream: Optional[int] = None) -> None:
    """Wait for the stream and raise if the last FPS kernel reported a failure."""
    _raise(lib().fpsx_device_status(stream), "fpsx_device_status")


def synchronize() -> None:
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def timing(enable: bool) -> None:
    _raise(lib().fpsx_timing(int(enable)), "fpsx_timing")


def timing_read():
    """(total kernel ms, launches) recorded since timing(True) or the last read."""
    ms = ctypes.c_double()
    n = ctypes.c_int()
    _raise(lib().fpsx_timing_read(ctypes.byref(ms), ctypes.byref(n)), "fpsx_timing_read")
    return ms.value, n.value


def plan_info(n_points: int, dim: int, tie: int = TIE_HIGHEST) -> dict:
    v = [ctypes.c_int() for _ in range(5)]
    _raise(lib().fpsx_plan_info(n_points, dim, tie, *[ctypes.byref(x) for x in v]),
           "fpsx_plan_info")
    kind, block, ppt, grid, chunk = (x.value for x in v)
    return dict(kind=("resident", "stream")[kind], block=block, ppt=ppt, grid=grid, chunk=chunk)

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
    """The HIP runtime libfpsx.so runs on, loaded under its unversioned name.

    Loading it by that name before libfpsx.so matters when torch shares the
    process: torch's libraries ask for "libamdhip64.so" and reuse an object
    already loaded under that name, while libfpsx.so's "libamdhip64.so.7"
    matches the soname of whichever copy is loaded first (the system's, or
    torch's bundled one when torch came first).  Without it a torch imported
    after this library maps a second HIP runtime, which then sees no device.
    """
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
        hip()  # the runtime first, under the name torch asks for
        L = ctypes.CDLL(os.path.join(HERE, "libfpsx.so"))
        sz, vp = ctypes.c_size_t, ctypes.c_void_p
        L.fpsx_last_error.restype = ctypes.c_char_p
        L.fpsx_fps_sampling_device.argtypes = [vp, sz, sz, sz, sz, vp, sz, vp, ctypes.c_int,
                                               ctypes.c_float, vp]
        L.fpsx_fps_sampling_device.restype = ctypes.c_int
        L.fpsx_fps_sampling_device_strided.argtypes = [vp, sz, sz, sz, sz, sz, vp, sz, vp, vp,
                                                       ctypes.c_int, ctypes.c_float, vp]
        L.fpsx_fps_sampling_device_strided.restype = ctypes.c_int
        L.fpsx_device_status.argtypes = [vp]
        L.fpsx_stream.restype = vp
        L.fpsx_stream.argtypes = []
        ip_ = ctypes.POINTER(ctypes.c_int)
        L.fpsx_autotune.argtypes = [vp, sz, sz, sz, sz, ctypes.c_int, sz, vp, ip_, ip_, ip_]
        L.fpsx_autotune.restype = ctypes.c_int
        L.fpsx_timing.argtypes = [ctypes.c_int]
        L.fpsx_timing_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
        ip = ctypes.POINTER(ctypes.c_int)
        L.fpsx_plan_info.argtypes = [sz, sz, ctypes.c_int, sz, ip, ip, ip, ip, ip]
        L.fpsx_fps_npdu_sampling_device.argtypes = [vp, sz, sz, sz, sz, sz, sz, vp, vp, vp, vp]
        L.fpsx_fps_npdu_sampling_device.restype = ctypes.c_int
        L.fpsx_npdu_plan_info.argtypes = [sz, sz, ip, ip, ip, ip]
        L.fpsx_npdu_plan_info.restype = ctypes.c_int
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

    @property
    def __cuda_array_interface__(self):
        return {"shape": self.shape, "typestr": self.dtype.str, "data": (self.ptr, False),
                "strides": None, "version": 3}

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


def fps_sampling_device_strided(points, stride_point: int, stride_dim: int, n_points: int,
                                dim: int, n_samples: int, starts, n_starts: int, out,
                                out_points=None, tie: int = TIE_HIGHEST,
                                init: float = float("inf"), stream: Optional[int] = None) -> None:
    """Enqueue FPS on device data of any layout (strides in elements); optionally
    write the picks' coordinates (n_samples x dim float32) to out_points."""
    rc = lib().fpsx_fps_sampling_device_strided(
        _addr(points), stride_point, stride_dim, n_points, dim, n_samples, _addr(starts),
        n_starts, _addr(out), _addr(out_points) if out_points is not None else None, tie, init,
        stream)
    _raise(rc, "fpsx_fps_sampling_device_strided")


def fps_npdu_sampling_device(points, stride_point: int, stride_dim: int, n_points: int,
                             dim: int, n_samples: int, w: int, start, out, out_points=None,
                             stream: Optional[int] = None) -> None:
    """Enqueue windowed-update FPS (fps_npdu_sampling semantics) on device data;
    `start` is a device array holding the one start index."""
    rc = lib().fpsx_fps_npdu_sampling_device(
        _addr(points), stride_point, stride_dim, n_points, dim, n_samples, w, _addr(start),
        _addr(out), _addr(out_points) if out_points is not None else None, stream)
    _raise(rc, "fpsx_fps_npdu_sampling_device")


def npdu_plan_info(n_points: int, dim: int) -> dict:
    """The windowed-update kernel chosen for a cloud (kernel_dim 0: any dimension)."""
    v = [ctypes.c_int() for _ in range(4)]
    _raise(lib().fpsx_npdu_plan_info(n_points, dim, *[ctypes.byref(x) for x in v]),
           "fpsx_npdu_plan_info")
    kdim, block, ppt, grid = (x.value for x in v)
    return dict(kernel_dim=kdim, block=block, ppt=ppt, grid=grid)


class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    pass


_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p),
                             ("deleter", ctypes.CFUNCTYPE(None, ctypes.POINTER(_DLManagedTensor)))]

_KDL_ROCM = 10  # DLPack device type of ROCm device memory
_KDL_FLOAT = 2


class _DLPackHold:
    """Owns a consumed DLPack capsule; releases the producer's tensor on close()."""

    def __init__(self, capsule):
        api = ctypes.pythonapi
        api.PyCapsule_GetPointer.restype = ctypes.c_void_p
        api.PyCapsule_GetPointer.argtypes = [ctypes.py_object, ctypes.c_char_p]
        api.PyCapsule_SetName.argtypes = [ctypes.py_object, ctypes.c_char_p]
        ptr = api.PyCapsule_GetPointer(capsule, b"dltensor")
        if not ptr:
            raise TypeError("not a DLPack tensor capsule")
        # Renamed so the capsule's own destructor leaves the tensor to us.
        api.PyCapsule_SetName(capsule, b"used_dltensor")
        self.mt = ctypes.cast(ptr, ctypes.POINTER(_DLManagedTensor))

    def close(self) -> None:
        if self.mt is not None:
            if self.mt.contents.deleter:
                self.mt.contents.deleter(self.mt)
            self.mt = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _is_torch(t) -> bool:
    return type(t).__module__.split(".")[0] == "torch" and hasattr(t, "is_cuda")


def _tensor_view(t, stream: Optional[int] = None):
    """(address, shape, strides in elements, keep-alive) of a 2-D float32
    device array: a torch tensor on the GPU, an object exporting DLPack
    (__dlpack__ / __dlpack_device__) on a ROCm device, or anything exposing
    __cuda_array_interface__.  Host tensors are refused here, before any
    launch.  For DLPack the producer is asked to order its pending writes
    before `stream` (the stream the caller's FPS will run on)."""
    if _is_torch(t):
        if not t.is_cuda:
            raise TypeError(f"points must be a GPU tensor (got device {t.device})")
        if str(t.dtype) != "torch.float32":
            raise TypeError("points must be float32")
        if t.dim() != 2:
            raise ValueError("points must be a 2-D (n_points, dim) tensor")
        return int(t.data_ptr()), tuple(t.shape), tuple(t.stride()), None
    if hasattr(t, "__dlpack__") and hasattr(t, "__dlpack_device__"):
        dev_type, _ = t.__dlpack_device__()
        if int(dev_type) != _KDL_ROCM:
            raise TypeError(f"points must be in ROCm device memory (DLPack device type {dev_type})")
        hold = _DLPackHold(t.__dlpack__(stream=stream if stream is not None else lib_stream()))
        dl = hold.mt.contents.dl_tensor
        if dl.dtype.code != _KDL_FLOAT or dl.dtype.bits != 32 or dl.dtype.lanes != 1:
            hold.close()
            raise TypeError("points must be float32")
        if dl.ndim != 2:
            hold.close()
            raise ValueError("points must be a 2-D (n_points, dim) tensor")
        shape = (int(dl.shape[0]), int(dl.shape[1]))
        strides = ((int(dl.strides[0]), int(dl.strides[1])) if dl.strides
                   else (shape[1], 1))
        return int(dl.data or 0) + int(dl.byte_offset), shape, strides, hold
    cai = getattr(t, "__cuda_array_interface__", None)
    if cai is None:
        raise TypeError("expected a GPU tensor (torch), a DLPack-exporting device array or an "
                        "object with __cuda_array_interface__")
    if np.dtype(cai["typestr"]) != np.float32:
        raise TypeError("points must be float32")
    shape = tuple(cai["shape"])
    if len(shape) != 2:
        raise ValueError("points must be a 2-D (n_points, dim) tensor")
    strides = cai.get("strides")
    if strides is None:
        strides = (shape[1] * 4, 4)
    return int(cai["data"][0]), shape, tuple(s // 4 for s in strides), None


def lib_stream() -> int:
    """The library's own stream for the current device (a hipStream_t as int)."""
    return int(lib().fpsx_stream() or 0)


def fps_sampling_tensor(points, n_samples: int, start_idx: Union[int, Sequence[int]] = 0,
                        return_points: bool = False, stream: Optional[int] = None,
                        sync: bool = True):
    """FPS on a device-resident (n_points, dim) float32 tensor, used in place
    (the zero-copy form of fps_sampling; src/fpsample/__init__.py:62 hands the
    reference an F-order float32 copy instead).

    * torch tensor on the GPU: the work runs on the caller's stream
      (torch.cuda.current_stream(), or `stream`), after whatever torch work
      produced the tensor; returns torch tensors: indices int64 (n_samples,),
      coordinates float32 (n_samples, dim) when return_points, else None.
    * an object exporting DLPack on a ROCm device: its producer orders its
      writes before the library's stream (or `stream`); returns DeviceArrays
      (indices uint64).
    * an object with __cuda_array_interface__: returns DeviceArrays.

    sync=True (default) waits for the run and raises if the kernel reported a
    failure (an exchange time-out); with sync=False the call returns once the
    work is enqueued, and device_status(stream) reports the status later.
    """
    torch_in = _is_torch(points)
    if torch_in:
        import torch

        if points.is_cuda:
            dev = points.device
            producer = torch.cuda.current_stream(dev)
            if stream is None:
                ts = producer
            else:
                # A stream of the caller's: order it after the producer's work,
                # and allocate the outputs for it.
                ts = torch.cuda.ExternalStream(int(stream), device=dev)
                ts.wait_stream(producer)
            stream = int(ts.cuda_stream)
    addr, shape, strides, hold = _tensor_view(points, stream)
    try:
        n, d = shape
        starts = np.atleast_1d(np.asarray(start_idx, dtype=np.uint64))
        if n_samples < 1 or n_samples > n:
            raise ValueError("n_samples must be in [1, n_points]")
        if starts.size < 1 or starts.size > n_samples or np.any(starts >= n):
            raise ValueError("start indices must be < n_points, 1 <= count <= n_samples")
        if torch_in:
            with torch.cuda.device(dev), torch.cuda.stream(ts):
                d_start = torch.as_tensor(starts.astype(np.int64)).to(dev, non_blocking=False)
                out = torch.empty(n_samples, dtype=torch.int64, device=dev)
                out_pts = (torch.empty((n_samples, d), dtype=torch.float32, device=dev)
                           if return_points else None)
                fps_sampling_device_strided(addr, strides[0], strides[1], n, d, n_samples,
                                            d_start, starts.size, out, out_pts, stream=stream)
        else:
            d_start = DeviceArray.from_numpy(starts)
            out = DeviceArray((n_samples,), np.uint64)
            out_pts = DeviceArray((n_samples, d), np.float32) if return_points else None
            fps_sampling_device_strided(addr, strides[0], strides[1], n, d, n_samples, d_start,
                                        starts.size, out, out_pts, stream=stream)
            # The start array and the producer's tensor must outlive the run:
            # freeing `out` (hipFree waits for the device) releases them after it.
            out._keep = (d_start, hold)
        if sync:
            device_status(stream)
    except BaseException:
        if hold is not None:
            _synchronize_quietly()
            hold.close()
        raise
    return out, out_pts


def _synchronize_quietly() -> None:
    try:
        synchronize()
    except Exception:
        pass


def autotune(points, stride_point: int, stride_dim: int, n_points: int, dim: int,
             tie: int = TIE_HIGHEST, passes: int = 256, stream: Optional[int] = None) -> dict:
    """Time every feasible kernel shape on this cloud and keep the fastest for
    clouds of the same (n_points, dim, tie); returns the chosen shape."""
    b, p, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = lib().fpsx_autotune(_addr(points), stride_point, stride_dim, n_points, dim, tie, passes,
                             stream, ctypes.byref(b), ctypes.byref(p), ctypes.byref(g))
    _raise(rc, "fpsx_autotune")
    return dict(block=b.value, ppt=p.value, grid=g.value)


def device_status(stream: Optional[int] = None) -> None:
    """Wait for the stream and raise if the last FPS kernel reported a failure."""
    _raise(lib().fpsx_device_status(stream), "fpsx_device_status")


def exchange_timeout(seconds: float) -> None:
    """Bound on one cross-workgroup exchange wait for later launches on the
    current device (seconds <= 0: the default, 2 s or FPSX_EXCHANGE_TIMEOUT_S)."""
    L = lib()
    L.fpsx_exchange_timeout.argtypes = [ctypes.c_double]
    _raise(L.fpsx_exchange_timeout(float(seconds)), "fpsx_exchange_timeout")


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


def plan_info(n_points: int, dim: int, tie: int = TIE_HIGHEST, n_samples: int = 0) -> dict:
    """The launch the planner picks (n_samples 0: as for a short run)."""
    v = [ctypes.c_int() for _ in range(5)]
    _raise(lib().fpsx_plan_info(n_points, dim, tie, n_samples, *[ctypes.byref(x) for x in v]),
           "fpsx_plan_info")
    kind, block, ppt, grid, chunk = (x.value for x in v)
    return dict(kind=("resident", "stream", "pruned", "npdu", "spruned", "pair", "spair")[kind], block=block, ppt=ppt, grid=grid,
                chunk=chunk)

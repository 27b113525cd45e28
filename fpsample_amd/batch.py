"""Batched FPS on device data: many clouds, one persistent launch.

``DeviceBatch`` uploads a list of clouds once (concatenated SoA in HBM),
creates a batch object in libfpsx.so (job tables, optionally a captured HIP
graph of the launch sequence) and runs it as often as asked.  Each cloud's
picks equal fps_sampling on that cloud alone.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import device as D

GRAPH = 1
AUTOTUNE = 2

_bound = False


def _lib():
    global _bound
    L = D.lib()
    if not _bound:
        sz, vp = ctypes.c_size_t, ctypes.c_void_p
        u64p = ctypes.POINTER(ctypes.c_uint64)
        ip = ctypes.POINTER(ctypes.c_int)
        L.fpsx_batch_create.argtypes = [vp, sz, sz, u64p, sz, u64p, u64p, vp, ctypes.c_int,
                                        ctypes.POINTER(vp)]
        L.fpsx_batch_create.restype = ctypes.c_int
        L.fpsx_batch_run.argtypes = [vp, vp]
        L.fpsx_batch_run.restype = ctypes.c_int
        L.fpsx_batch_status.argtypes = [vp, vp]
        L.fpsx_batch_status.restype = ctypes.c_int
        L.fpsx_batch_info.argtypes = [vp, ip, ip, ip, ip, ip]
        L.fpsx_batch_info.restype = ctypes.c_int
        L.fpsx_batch_destroy.argtypes = [vp]
        L.fpsx_batch_destroy.restype = None
        _bound = True
    return L


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


class DeviceBatch:
    def __init__(self, clouds: Sequence[np.ndarray], n_samples: Sequence[int],
                 start_idx: Sequence[int], graph: bool = True, autotune: bool = False):
        clouds = [np.asarray(c, dtype=np.float32) for c in clouds]
        dim = clouds[0].shape[1]
        self.sizes = [c.shape[0] for c in clouds]
        self.n_samples = [int(v) for v in n_samples]
        self.offsets = _u64(np.concatenate([[0], np.cumsum(self.sizes)]))
        total = int(self.offsets[-1])
        soa = np.empty((dim, total), np.float32)
        for c, lo in zip(clouds, self.offsets[:-1]):
            soa[:, lo:lo + c.shape[0]] = c.T
        self.d_pts = D.DeviceArray.from_numpy(soa)
        self.d_out = D.DeviceArray((sum(self.n_samples),), np.uint64)
        ms = _u64(self.n_samples)
        st = _u64(start_idx)
        self._ms, self._st = ms, st
        h = ctypes.c_void_p()
        u64p = ctypes.POINTER(ctypes.c_uint64)
        flags = (GRAPH if graph else 0) | (AUTOTUNE if autotune else 0)
        rc = _lib().fpsx_batch_create(self.d_pts.ptr, total, dim, self.offsets.ctypes.data_as(u64p),
                                      len(clouds), ms.ctypes.data_as(u64p), st.ctypes.data_as(u64p),
                                      self.d_out.ptr, flags, ctypes.byref(h))
        D._raise(rc, "fpsx_batch_create")
        self.handle = h.value

    def run(self, stream: Optional[int] = None) -> None:
        D._raise(_lib().fpsx_batch_run(self.handle, stream), "fpsx_batch_run")

    def status(self, stream: Optional[int] = None) -> None:
        D._raise(_lib().fpsx_batch_status(self.handle, stream), "fpsx_batch_status")

    def info(self) -> dict:
        v = [ctypes.c_int() for _ in range(5)]
        D._raise(_lib().fpsx_batch_info(self.handle, *[ctypes.byref(x) for x in v]),
                 "fpsx_batch_info")
        return dict(zip(("block", "ppt", "grid", "members", "graph"), (x.value for x in v)))

    def picks_flat(self) -> np.ndarray:
        return self.d_out.numpy()

    def picks(self) -> List[np.ndarray]:
        flat = self.picks_flat()
        bounds = np.concatenate([[0], np.cumsum(self.n_samples)])
        return [flat[a:b] for a, b in zip(bounds[:-1], bounds[1:])]

    def close(self) -> None:
        if self.handle:
            _lib().fpsx_batch_destroy(self.handle)
            self.handle = None
        self.d_pts.free()
        self.d_out.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""FPS serving loop (libfpsx.so's fpsx_server_*): submit clouds from any thread,
wait for picks.  Admission by HBM budget, continuous batching on the GPU."""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

from . import device as D

_bound = False


def _lib():
    global _bound
    L = D.lib()
    if not _bound:
        sz, vp = ctypes.c_size_t, ctypes.c_void_p
        u64, u64p = ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)
        L.fpsx_server_create.argtypes = [sz, ctypes.c_uint, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(vp)]
        L.fpsx_server_submit.argtypes = [vp, ctypes.POINTER(ctypes.c_float), sz, sz, sz, u64, u64p]
        L.fpsx_server_wait.argtypes = [vp, u64, u64p, ctypes.POINTER(ctypes.c_double)]
        L.fpsx_server_stats.argtypes = [vp, u64p, u64p, u64p, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(sz)]
        L.fpsx_server_destroy.argtypes = [vp]
        L.fpsx_server_destroy.restype = None
        for f in ("fpsx_server_create", "fpsx_server_submit", "fpsx_server_wait",
                  "fpsx_server_stats"):
            getattr(L, f).restype = ctypes.c_int
        _bound = True
    return L


class BudgetError(RuntimeError):
    """The request can never fit the server's HBM budget."""


class Server:
    def __init__(self, hbm_budget_bytes: int = 8 << 30, chunk_passes: int = 256,
                 block: int = 0, ppt: int = 0):
        h = ctypes.c_void_p()
        D._raise(_lib().fpsx_server_create(hbm_budget_bytes, chunk_passes, block, ppt,
                                           ctypes.byref(h)), "fpsx_server_create")
        self.handle = h.value
        self._m = {}

    def submit(self, pc: np.ndarray, n_samples: int, start_idx: int = 0) -> int:
        pc = np.ascontiguousarray(pc, dtype=np.float32)
        rid = ctypes.c_uint64()
        rc = _lib().fpsx_server_submit(self.handle, pc.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       pc.shape[0], pc.shape[1], n_samples, start_idx,
                                       ctypes.byref(rid))
        if rc == 8:
            raise BudgetError(D.lib().fpsx_last_error().decode())
        D._raise(rc, "fpsx_server_submit")
        self._m[rid.value] = n_samples
        return rid.value

    def wait(self, request_id: int) -> Tuple[np.ndarray, float]:
        out = np.empty(self._m.pop(request_id), np.uint64)
        lat = ctypes.c_double()
        D._raise(_lib().fpsx_server_wait(self.handle, request_id,
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                         ctypes.byref(lat)), "fpsx_server_wait")
        return out, lat.value

    def stats(self) -> dict:
        u = [ctypes.c_uint64() for _ in range(3)]
        busy = ctypes.c_double()
        z = [ctypes.c_size_t() for _ in range(3)]
        rc = _lib().fpsx_server_stats(self.handle, *[ctypes.byref(x) for x in u],
                                      ctypes.byref(busy), *[ctypes.byref(x) for x in z])
        return dict(launches=u[0].value, completed=u[1].value, samples=u[2].value,
                    busy_s=busy.value, hbm_in_use=z[0].value, pending=z[1].value,
                    active=z[2].value, error=rc)

    def close(self) -> None:
        if self.handle:
            _lib().fpsx_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

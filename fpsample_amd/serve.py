"""FPS serving loop (libfpsx.so's fpsx_server_*): submit clouds from any thread,
wait for picks.  Admission by HBM budget and priority, continuous batching on
the GPU; requests can carry a deadline, be cancelled, and be waited for with a
time limit."""
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
        L.fpsx_server_submit_ex.argtypes = [vp, ctypes.POINTER(ctypes.c_float), sz, sz, sz, u64,
                                            ctypes.c_int, ctypes.c_double, u64p]
        L.fpsx_server_wait_for.argtypes = [vp, u64, ctypes.c_double, u64p,
                                           ctypes.POINTER(ctypes.c_double)]
        L.fpsx_server_cancel.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_int)]
        L.fpsx_server_counters.argtypes = [vp, u64p, u64p, u64p, u64p]
        for f in ("fpsx_server_create", "fpsx_server_submit", "fpsx_server_wait",
                  "fpsx_server_stats", "fpsx_server_submit_ex", "fpsx_server_wait_for",
                  "fpsx_server_cancel", "fpsx_server_counters"):
            getattr(L, f).restype = ctypes.c_int
        _bound = True
    return L


# Return codes of the serving entries (include/fpsx.h).
_EBUDGET, _ECANCELED, _EPENDING, _EDEADLINE = 8, 9, 10, 11


class BudgetError(RuntimeError):
    """The request can never fit the server's HBM budget."""


class CancelledError(RuntimeError):
    """The request was cancelled before it finished."""


class DeadlineError(RuntimeError):
    """The request's deadline passed before it finished."""


class Server:
    def __init__(self, hbm_budget_bytes: int = 8 << 30, chunk_passes: int = 256,
                 block: int = 0, ppt: int = 0):
        h = ctypes.c_void_p()
        D._raise(_lib().fpsx_server_create(hbm_budget_bytes, chunk_passes, block, ppt,
                                           ctypes.byref(h)), "fpsx_server_create")
        self.handle = h.value
        self._m = {}

    def submit(self, pc: np.ndarray, n_samples: int, start_idx: int = 0, priority: int = 0,
               deadline: Optional[float] = None) -> int:
        """Queue one 3-D cloud.  Higher `priority` is admitted first and may
        pause lower-priority requests between chunks, or evict them to host
        memory when HBM is short (they resume later with the same picks);
        `deadline` (seconds from now) stops the request if it has not finished
        by then."""
        pc = np.ascontiguousarray(pc, dtype=np.float32)
        rid = ctypes.c_uint64()
        rc = _lib().fpsx_server_submit_ex(self.handle,
                                          pc.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                          pc.shape[0], pc.shape[1], n_samples, start_idx,
                                          int(priority), float(deadline or 0.0), ctypes.byref(rid))
        if rc == _EBUDGET:
            raise BudgetError(D.lib().fpsx_last_error().decode())
        D._raise(rc, "fpsx_server_submit_ex")
        self._m[rid.value] = n_samples
        return rid.value

    def wait(self, request_id: int, timeout: Optional[float] = None) -> Tuple[np.ndarray, float]:
        """(picks, latency in seconds).  With `timeout` (seconds), raises
        TimeoutError if the request is still running then; it can be waited
        for again.  CancelledError / DeadlineError if it ended without picks."""
        out = np.empty(self._m[request_id], np.uint64)
        lat = ctypes.c_double()
        rc = _lib().fpsx_server_wait_for(self.handle, request_id,
                                         -1.0 if timeout is None else float(timeout),
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                         ctypes.byref(lat))
        if rc == _EPENDING:
            raise TimeoutError(f"request {request_id} still running after {timeout} s")
        self._m.pop(request_id, None)
        if rc == _ECANCELED:
            raise CancelledError(f"request {request_id} was cancelled")
        if rc == _EDEADLINE:
            raise DeadlineError(f"request {request_id} missed its deadline")
        D._raise(rc, "fpsx_server_wait_for")
        return out, lat.value

    def cancel(self, request_id: int) -> bool:
        """Stop a request; True if it will end without picks (its wait raises
        CancelledError), False if it had finished or its last chunk was queued."""
        c = ctypes.c_int()
        D._raise(_lib().fpsx_server_cancel(self.handle, request_id, ctypes.byref(c)),
                 "fpsx_server_cancel")
        return bool(c.value)

    def stats(self) -> dict:
        u = [ctypes.c_uint64() for _ in range(3)]
        busy = ctypes.c_double()
        z = [ctypes.c_size_t() for _ in range(3)]
        rc = _lib().fpsx_server_stats(self.handle, *[ctypes.byref(x) for x in u],
                                      ctypes.byref(busy), *[ctypes.byref(x) for x in z])
        c = [ctypes.c_uint64() for _ in range(4)]
        _lib().fpsx_server_counters(self.handle, *[ctypes.byref(x) for x in c])
        return dict(launches=u[0].value, completed=u[1].value, samples=u[2].value,
                    busy_s=busy.value, hbm_in_use=z[0].value, pending=z[1].value,
                    active=z[2].value, cancelled=c[0].value, expired=c[1].value,
                    paused=c[2].value, evicted=c[3].value, error=rc)

    def close(self) -> None:
        if self.handle:
            _lib().fpsx_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

"""CPU oracle for farthest point sampling -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module, and only as the checker.  The fpsample_amd
package never imports it.

Two layers:

* ``fps``, ``fps_ex``, ``kdline_perm``, ``kdline_quickfps``,
  ``kdline_exhaustive`` call the C restatement in ``fps_oracle.c`` (built into
  ``oracle/liboracle.so`` by ``oracle/Makefile``).  Each C function cites the
  reference file:line it restates.
* ``fps_f64``, ``coverage_radius`` and ``accuracy_vs_f64`` are the
  double-precision yardstick: index agreement and coverage-radius relative
  error of a float32 pick sequence against the same greedy loop in double.
* ``fps_numpy`` is an independent vectorised numpy restatement of
  src/lib.cpp:111-246 (float32 elementwise ops are unfused, round-to-nearest),
  used to cross-check the C code without sharing any of it.

``load_reference()`` returns the reference itself, compiled from its own
sources into ``oracle/_ref/`` (``_fpsample`` pybind11 module from
src/lib.cpp, ``libref_bucket.so`` from src/wrapper.hpp), or None where that
build is absent.
"""
from __future__ import annotations

import ctypes
import glob
import importlib.util
import os
from typing import Optional, Sequence, Union

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB: Optional[ctypes.CDLL] = None
FLT_MAX = float(np.finfo(np.float32).max)

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        sz = ctypes.c_size_t
        L.oracle_fps_ex.argtypes = [_f32p, sz, sz, sz, _u64p, sz, ctypes.c_int,
                                    ctypes.c_float, _u64p, _f32p]
        L.oracle_fps_ex.restype = ctypes.c_int
        L.oracle_fps.argtypes = [_f32p, sz, sz, sz, _u64p, sz, _u64p]
        L.oracle_fps.restype = ctypes.c_int
        L.oracle_kdline_perm.argtypes = [_f32p, sz, sz, sz, _u64p, _u64p, _u64p]
        L.oracle_kdline_perm.restype = ctypes.c_long
        L.oracle_fps_f64.argtypes = [_f32p, sz, sz, sz, sz, _u64p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_fps_f64.restype = ctypes.c_int
        L.oracle_coverage_f64.argtypes = [_f32p, sz, sz, _u64p, sz]
        L.oracle_coverage_f64.restype = ctypes.c_double
        L.oracle_npdu.argtypes = [_f32p, sz, sz, sz, sz, sz, _u64p]
        L.oracle_npdu.restype = ctypes.c_int
        for name in ("oracle_kdline_quickfps", "oracle_kdline_exhaustive"):
            fn = getattr(L, name)
            fn.argtypes = [_f32p, sz, sz, sz, sz, sz, _u64p]
            fn.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _rows(pc: np.ndarray) -> np.ndarray:
    pc = np.ascontiguousarray(pc, dtype=np.float32)
    if pc.ndim != 2:
        raise ValueError("points must be 2-D")
    return pc


def _starts(start_idx: Union[int, Sequence[int], np.ndarray]) -> np.ndarray:
    s = np.atleast_1d(np.asarray(start_idx, dtype=np.uint64))
    if s.size == 0:
        raise ValueError("empty start list")
    return np.ascontiguousarray(s)


def fps_ex(pc: np.ndarray, n_samples: int, start_idx=0, lowest: bool = False,
           init: float = float("inf"), with_dist: bool = False):
    """General FPS (forced start prefix, tie rule, initial distance)."""
    pts = _rows(pc)
    st = _starts(start_idx)
    out = np.zeros(n_samples, dtype=np.uint64)
    dist = np.zeros(n_samples, dtype=np.float32) if with_dist else None
    rc = lib().oracle_fps_ex(
        pts.ctypes.data_as(_f32p), pts.shape[0], pts.shape[1], n_samples,
        st.ctypes.data_as(_u64p), st.size, int(bool(lowest)), ctypes.c_float(init),
        out.ctypes.data_as(_u64p), dist.ctypes.data_as(_f32p) if with_dist else None)
    if rc != 0:
        raise RuntimeError("oracle_fps_ex: bad arguments or out of memory")
    return (out, dist) if with_dist else out


def fps(pc: np.ndarray, n_samples: int, start_idx=0) -> np.ndarray:
    """Reference fps_sampling semantics (src/lib.cpp:111-246)."""
    return fps_ex(pc, n_samples, start_idx)


def kdline_perm(pc: np.ndarray, h: int):
    """(perm, leaf_lo, leaf_hi) of the KD-line build (src/_ext/KDTreeBase.h:84-207)."""
    pts = _rows(pc)
    n = pts.shape[0]
    perm = np.zeros(n, dtype=np.uint64)
    cap = min(n, 1 << min(int(h), 40))
    lo = np.zeros(cap, dtype=np.uint64)
    hi = np.zeros(cap, dtype=np.uint64)
    nl = lib().oracle_kdline_perm(pts.ctypes.data_as(_f32p), n, pts.shape[1], int(h),
                                  perm.ctypes.data_as(_u64p), lo.ctypes.data_as(_u64p),
                                  hi.ctypes.data_as(_u64p))
    if nl < 0:
        raise RuntimeError("oracle_kdline_perm: bad arguments or out of memory")
    return perm, lo[:nl], hi[:nl]


def _kdline(fn, pc, n_samples, h, start_idx):
    pts = _rows(pc)
    out = np.zeros(n_samples, dtype=np.uint64)
    rc = fn(pts.ctypes.data_as(_f32p), pts.shape[0], pts.shape[1], n_samples,
            int(start_idx), int(h), out.ctypes.data_as(_u64p))
    if rc != 0:
        raise RuntimeError("oracle kdline: bad arguments or out of memory")
    return out


def kdline_quickfps(pc, n_samples, h, start_idx=0):
    """Bucket FPS with lazy updates (src/wrapper.hpp:45-60, src/_ext/KDNode.h:84-166)."""
    return _kdline(lib().oracle_kdline_quickfps, pc, n_samples, h, start_idx)


def kdline_exhaustive(pc, n_samples, h, start_idx=0):
    """Exhaustive FPS on the KD-permuted order, lowest-position ties, FLT_MAX init."""
    return _kdline(lib().oracle_kdline_exhaustive, pc, n_samples, h, start_idx)


def npdu(pc: np.ndarray, n_samples: int, w: int, start_idx: int = 0) -> np.ndarray:
    """FPS with windowed updates (src/lib.cpp:272-340)."""
    pts = _rows(pc)
    out = np.zeros(n_samples, dtype=np.uint64)
    rc = lib().oracle_npdu(pts.ctypes.data_as(_f32p), pts.shape[0], pts.shape[1], n_samples,
                           int(w), int(start_idx), out.ctypes.data_as(_u64p))
    if rc != 0:
        raise RuntimeError("oracle_npdu: bad arguments or out of memory")
    return out


def fps_f64(pc: np.ndarray, n_samples: int, start_idx: int = 0):
    """(picks, squared pick distances) of the greedy loop in double precision
    (accuracy yardstick for the float32 result, not a reference restatement)."""
    pts = _rows(pc)
    out = np.zeros(n_samples, dtype=np.uint64)
    dist = np.zeros(n_samples, dtype=np.float64)
    rc = lib().oracle_fps_f64(pts.ctypes.data_as(_f32p), pts.shape[0], pts.shape[1], n_samples,
                              int(start_idx), out.ctypes.data_as(_u64p),
                              dist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc != 0:
        raise RuntimeError("oracle_fps_f64: bad arguments or out of memory")
    return out, dist


def coverage_radius(pc: np.ndarray, picks) -> float:
    """max over points of the distance (double) to the nearest of `picks`."""
    pts = _rows(pc)
    p = np.ascontiguousarray(np.asarray(picks, dtype=np.uint64))
    r = lib().oracle_coverage_f64(pts.ctypes.data_as(_f32p), pts.shape[0], pts.shape[1],
                                  p.ctypes.data_as(_u64p), p.size)
    if r < 0:
        raise RuntimeError("oracle_coverage_f64: bad arguments or out of memory")
    return float(r)


def accuracy_vs_f64(pc: np.ndarray, picks, start_idx: int = 0) -> dict:
    """Index agreement of `picks` with the double-precision greedy run from the
    same start, and the relative difference of their coverage radii."""
    picks = np.asarray(picks, dtype=np.uint64)
    k = picks.size
    pts = _rows(pc)
    # One more greedy step: the (k+1)-th pick's distance is the coverage
    # radius of the first k picks.
    ref, dist = fps_f64(pts, min(k + 1, pts.shape[0]), start_idx)
    r_ref = float(np.sqrt(dist[k])) if k < pts.shape[0] else 0.0
    ref = ref[:k]
    r_got = coverage_radius(pts, picks)
    agree = int(np.argmin(picks == ref)) if not np.all(picks == ref) else k
    return {"picks": k, "index_agreement": float(np.mean(picks == ref)),
            "identical_prefix": agree, "coverage_radius_f64": r_ref,
            "coverage_radius": r_got,
            "coverage_radius_rel_err": abs(r_got - r_ref) / r_ref if r_ref > 0 else 0.0}


def fps_numpy(pc: np.ndarray, n_samples: int, start_idx=0, lowest: bool = False,
              init: float = float("inf")) -> np.ndarray:
    """Vectorised numpy restatement of src/lib.cpp:111-246 (small inputs only)."""
    pts = _rows(pc)
    n, d = pts.shape
    st = [int(s) for s in _starts(start_idx)]
    dmin = np.full(n, init, dtype=np.float32)
    out = np.zeros(n_samples, dtype=np.uint64)
    last = st[0]
    out[0] = last
    used = 1
    for s in range(1, n_samples):
        acc = np.zeros(n, dtype=np.float32)
        for j in range(d):
            diff = pts[:, j] - pts[last, j]
            acc = acc + diff * diff
        np.minimum(dmin, acc, out=dmin, where=acc < dmin)
        if used < len(st):
            last = st[used]
            used += 1
        else:
            mx = dmin.max()
            ties = np.flatnonzero(dmin == mx)
            last = int(ties[0] if lowest else ties[-1])
        out[s] = last
    return out


# ---------------------------------------------------------------------------
# The reference itself (compiled from its sources into oracle/_ref)
# ---------------------------------------------------------------------------
class Reference:
    """Thin caller of the reference build: the pybind11 module of
    src/lib.cpp and the extern "C" bucket API of src/wrapper.hpp."""

    def __init__(self, module, bucket_lib):
        self.module = module
        self.bucket = bucket_lib

    def fps_sampling(self, pc: np.ndarray, n_samples: int, start_idx) -> np.ndarray:
        # what src/fpsample/__init__.py:62-65 hands to _fps_sampling
        pc = np.asfortranarray(pc, dtype=np.float32)
        if isinstance(start_idx, (list, tuple, np.ndarray)):
            start_idx = np.array(start_idx, dtype=np.uint64)
        return self.module._fps_sampling(pc, n_samples, start_idx)

    def fps_npdu(self, pc: np.ndarray, n_samples: int, w: int, start_idx: int) -> np.ndarray:
        # what src/fpsample/__init__.py:98-103 hands to _fps_npdu_sampling
        pc = np.ascontiguousarray(pc, dtype=np.float32)
        return self.module._fps_npdu_sampling(pc, n_samples, w, start_idx)

    def bucket_fps_kdline(self, pc: np.ndarray, n_samples: int, h: int,
                          start_idx: int) -> np.ndarray:
        pts = np.ascontiguousarray(pc, dtype=np.float32)
        out = np.zeros(n_samples, dtype=np.uint64)
        rc = self.bucket.bucket_fps_kdline(
            pts.ctypes.data_as(_f32p), ctypes.c_size_t(pts.shape[0]),
            ctypes.c_size_t(pts.shape[1]), ctypes.c_size_t(n_samples),
            ctypes.c_size_t(start_idx), ctypes.c_size_t(h), out.ctypes.data_as(_u64p))
        if rc != 0:
            raise RuntimeError(f"bucket_fps_kdline failed with error code {rc}")
        return out


_REF = None


def load_reference() -> Optional[Reference]:
    """The compiled reference, or None where oracle/_ref was not built."""
    global _REF
    if _REF is not None:
        return _REF
    ref_dir = os.path.join(HERE, "_ref")
    mods = glob.glob(os.path.join(ref_dir, "_fpsample*.so"))
    blib = os.path.join(ref_dir, "libref_bucket.so")
    if not mods or not os.path.exists(blib):
        return None
    spec = importlib.util.spec_from_file_location("_fpsample", mods[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    B = ctypes.CDLL(blib)
    B.bucket_fps_kdline.restype = ctypes.c_int
    _REF = Reference(mod, B)
    return _REF

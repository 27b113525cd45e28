"""MI355X farthest point sampling -- drop-in for the reference's numpy API.

    import fpsample_amd as fpsample
    idx = fpsample.fps_sampling(pc, 1024, start_idx=0)
    idx = fpsample.bucket_fps_kdline_sampling(pc, 1024, h=7, start_idx=0)

The functions keep the reference's names, signatures, defaults, assertions,
dtype coercion and the single draw from numpy's global RNG when start_idx is
None (src/fpsample/__init__.py:19-103, 174-206).  The work runs on the GPU
through the pybind11 module ``fpsample_amd._fpsample`` and the C ABI of
``libfpsx.so``; there is no CPU fallback: importing this package fails when
the native extension is missing.
"""
from __future__ import annotations

import warnings
from typing import List, Optional, Union

import numpy as np

from .device import hip as _load_hip_runtime

# The HIP runtime is loaded under its unversioned name before the extension
# (whose libfpsx.so needs it): a torch imported later then binds to this same
# runtime instead of mapping its bundled copy beside it, and a torch imported
# earlier has its copy reused here.  One runtime per process either way.
_load_hip_runtime()

try:
    from ._fpsample import (  # noqa: F401
        __doc__ as _ext_doc,
        __version__,
        _bucket_fps_kdline_sampling,
        _fps_npdu_sampling,
        _fps_sampling,
        _fps_sampling_batch,
    )
except ImportError as exc:  # pragma: no cover - exercised only on a broken install
    raise ImportError(
        "fpsample_amd: the native extension (_fpsample + libfpsx.so) is not built; "
        "run `python fpsample_amd/build.py`"
    ) from exc


def get_start_idx(n_pts: int, start_idx: Optional[Union[int, List[int]]]) -> Union[int, np.ndarray]:
    """Resolve start_idx as src/fpsample/__init__.py:19-32 does.

    None draws one index from numpy's global RNG (np.random.randint), an int is
    used as is, a list becomes a uint64 array, anything else raises ValueError.
    """
    if start_idx is None:
        start_idx = np.random.randint(low=0, high=n_pts)
    elif isinstance(start_idx, int):
        start_idx = start_idx
    elif isinstance(start_idx, list):
        start_idx = np.array(start_idx, dtype=np.uint64)
    else:
        raise ValueError("start_idx should be None, int or list")
    return start_idx


def fps_sampling(
    pc: np.ndarray, n_samples: int, start_idx: Optional[Union[int, List[int]]] = None
) -> np.ndarray:
    """Farthest point sampling on the GPU, with the picks of src/fpsample/__init__.py:35-65.

    Each pick after the first is the point whose squared distance to its
    nearest earlier pick is largest (the highest index among equal maxima).

    pc: (n_pts, D) array, any float dtype (converted to float32).
    n_samples: how many indices to return, 1 <= n_samples <= n_pts.
    start_idx: the first pick; None draws one from numpy's global RNG; a list
        forces that prefix of picks (at most n_samples, each < n_pts), after
        which sampling continues from all of them.
    Returns a uint64 array of n_samples point indices, in pick order.
    """
    assert n_samples >= 1, "n_samples should be >= 1"
    assert pc.ndim == 2
    n_pts, _ = pc.shape
    assert n_pts >= n_samples, "n_pts should be >= n_samples"
    if isinstance(start_idx, int):
        assert (
            start_idx is None or 0 <= start_idx < n_pts
        ), "start_idx should be None or 0 <= start_idx < n_pts"
    if isinstance(start_idx, list):
        assert len(start_idx) <= n_samples, "len(start_idx) should be <= n_samples"
        for idx in start_idx:
            assert 0 <= idx < n_pts, "start_idx should be None or 0 <= start_idx < n_pts"
    # Fortran order float32 is the SoA layout the GPU kernels read.
    pc = np.asfortranarray(pc, dtype=np.float32)
    start_idx = get_start_idx(n_pts, start_idx)
    return _fps_sampling(pc, n_samples, start_idx)


def fps_npdu_sampling(
    pc: np.ndarray,
    n_samples: int,
    w: Optional[int] = None,
    start_idx: Optional[Union[int, List[int]]] = None,
) -> np.ndarray:
    """
    FPS with nearest-point-distance-updating (NPDU), src/fpsample/__init__.py:68-103.

    After the start, each pick updates only the min-distances of the ``w + 1``
    points around it in index order, so the result depends on the cloud's
    index order (dimensional locality); the next pick is still the largest
    min-distance over all points, lowest index on ties.

    pc: (n_pts, D) array (converted to float32, row-major).
    n_samples: how many indices to return, 1 <= n_samples <= n_pts.
    w: update window; None means ``int(n_pts / n_samples * 16)``, and values of
        n_pts - 1 or more are lowered to n_pts - 1 with a warning.
    start_idx: the first pick; None draws one from numpy's global RNG.
    Returns a uint64 array of n_samples point indices, in pick order.
    """
    assert n_samples >= 1, "n_samples should be >= 1"
    assert pc.ndim == 2
    n_pts, _ = pc.shape
    assert n_pts >= n_samples, "n_pts should be >= n_samples"
    assert (
        start_idx is None or 0 <= start_idx < n_pts
    ), "start_idx should be None or 0 <= start_idx < n_pts"
    if isinstance(start_idx, list):
        assert len(start_idx) <= n_samples, "len(start_idx) should be <= n_samples"
    pc = np.ascontiguousarray(pc, dtype=np.float32)
    w = w or int(n_pts / n_samples * 16)
    if w >= n_pts - 1:
        warnings.warn(f"k is too large, set to {n_pts - 1}")
        w = n_pts - 1
    start_idx = get_start_idx(n_pts, start_idx)
    return _fps_npdu_sampling(pc, n_samples, w, start_idx)


def bucket_fps_kdline_sampling(
    pc: np.ndarray, n_samples: int, h: int, start_idx: Optional[Union[int, List[int]]] = None
) -> np.ndarray:
    """
    Bucket-based FPS using a KD-line tree ("QuickFPS"), src/fpsample/__init__.py:174-206.

    Returns exactly the reference's picks: the tree of height ``h`` fixes the
    point order, the first pick is the point at position ``start_idx`` of that
    order, and ties go to the lowest position.  The sampling itself runs on the
    GPU as exhaustive FPS over the tree order, which selects the same points as
    the reference's pruned bucket search.

    pc: (n_pts, D) array with 1 <= D <= 8 (converted to float32, row-major).
    n_samples: how many indices to return, 1 <= n_samples <= n_pts.
    h: tree height, 1 <= h and 2**h <= n_pts (2**h buckets).
    start_idx: position of the first pick in the tree order; None draws one
        from numpy's global RNG.
    Returns a uint64 array of n_samples point indices, in pick order.
    """
    assert n_samples >= 1, "n_samples should be >= 1"
    assert pc.ndim == 2
    n_pts, _ = pc.shape
    assert n_pts >= n_samples, "n_pts should be >= n_samples"
    assert h >= 1, "h should be >= 1"
    assert 2**h <= n_pts, "2**h should be <= n_pts"
    assert (
        start_idx is None or 0 <= start_idx < n_pts
    ), "start_idx should be None or 0 <= start_idx < n_pts"
    if isinstance(start_idx, list):
        assert len(start_idx) <= n_samples, "len(start_idx) should be <= n_samples"
    pc = np.ascontiguousarray(pc, dtype=np.float32)
    start_idx = get_start_idx(n_pts, start_idx)
    return _bucket_fps_kdline_sampling(pc, n_samples, h, start_idx)


def fps_sampling_batch(
    pcs: List[np.ndarray],
    n_samples: Union[int, List[int]],
    start_idx: Optional[Union[int, List[int]]] = None,
) -> List[np.ndarray]:
    """
    Vanilla FPS over many clouds in one GPU launch (no counterpart in the reference).

    Each cloud's picks equal ``fps_sampling(pc, n, start_idx=s)`` for that cloud
    alone.  Clouds may differ in size but must share the number of columns.

    pcs: list of (n_i, D) arrays.
    n_samples: picks per cloud, one value for all or one per cloud.
    start_idx: first pick per cloud (None draws np.random.randint(0, n_i) for
        each cloud in order; an int applies to every cloud).
    Returns a list of uint64 pick arrays, one per cloud.
    """
    assert isinstance(pcs, (list, tuple)) and len(pcs) >= 1, "pcs should be a non-empty list"
    k = len(pcs)
    ns = [int(n_samples)] * k if isinstance(n_samples, (int, np.integer)) else [int(v) for v in n_samples]
    assert len(ns) == k, "n_samples should have one entry per cloud"
    if start_idx is None:
        starts = [int(np.random.randint(low=0, high=pc.shape[0])) for pc in pcs]
    elif isinstance(start_idx, int):
        starts = [start_idx] * k
    else:
        starts = [int(v) for v in start_idx]
    assert len(starts) == k, "start_idx should have one entry per cloud"
    for pc, n, s in zip(pcs, ns, starts):
        assert pc.ndim == 2
        assert n >= 1, "n_samples should be >= 1"
        assert pc.shape[0] >= n, "n_pts should be >= n_samples"
        assert 0 <= s < pc.shape[0], "start_idx should be None or 0 <= start_idx < n_pts"
    clouds = [np.ascontiguousarray(pc, dtype=np.float32) for pc in pcs]
    return _fps_sampling_batch(clouds, ns, starts)


__doc__ = _ext_doc

__all__ = [
    "__doc__",
    "__version__",
    "fps_sampling",
    "fps_npdu_sampling",
    "bucket_fps_kdline_sampling",
    "fps_sampling_batch",
    "get_start_idx",
]

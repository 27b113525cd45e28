"""Full-length golden fixtures (tests/golden/full/*.npz, see make_full_length.py).

Each fixture was written by the reference's fps_sampling (src/lib.cpp:188-246)
on a benchmark workload at its full size.  The clouds are regenerated from
their seeds by bench.py's generators and checked against the SHA-256 stored
with the fixture before any comparison.
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "full")


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def picks_sha256(picks) -> str:
    """SHA-256 of a pick sequence as uint64 little-endian (the reference's size_t)."""
    return sha256(np.asarray(picks).astype("<u8"))


def available(name: str) -> bool:
    return os.path.exists(os.path.join(HERE, name + ".npz"))


def load(name: str) -> dict:
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    if "out" in d:
        d["out"] = d["out"].astype(np.uint64)
    if d["cloud_sha256"].ndim == 0:
        d["cloud_sha256"] = str(d["cloud_sha256"])
    return d

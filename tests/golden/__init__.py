"""Golden fixtures written by the reference (see make_golden.py)."""
from __future__ import annotations

import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_all():
    """{name: dict(pc, m, start, kind, h, out)} for every fixture."""
    raw = {}
    for path in sorted(glob.glob(os.path.join(HERE, "*.npz"))):
        with np.load(path, allow_pickle=False) as z:
            raw[os.path.basename(path)[:-4]] = {k: z[k] for k in z.files}
    out = {}
    for name, d in raw.items():
        pc = d["pc"] if "pc" in d else raw[str(d["pc_from"])]["pc"]
        out[name] = dict(pc=pc, m=int(d["m"]), start=d["start"].astype(np.uint64),
                         kind=str(d["kind"]), h=int(d["h"]), out=d["out"].astype(np.uint64))
    return out

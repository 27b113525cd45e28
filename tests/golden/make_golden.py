"""Write the golden fixtures in tests/golden/*.npz from the reference itself.

The reference (src/lib.cpp's pybind11 module and src/wrapper.hpp's C API) is
compiled from its own sources into oracle/_ref/ by oracle/Makefile; this
script calls it on seeded inputs and stores inputs and outputs.  Run it in a
container where /root/reference exists:

    make -C oracle && python tests/golden/make_golden.py

Each fixture holds: pc (float32 rows) -- or pc_from, the name of the fixture
whose pc it shares --, m, start (uint64 array), kind ("fps" or "kdline"),
h (kdline height or npdu window, 0 for fps), out (uint64 picks).  tests/golden/__init__.py
loads them.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def c1():
    # the reference bench generator (bench/test_bench.py:19-21)
    np.random.seed(42)
    return np.random.rand(4096, 3).astype(np.float32)


def lattice(n, k, seed, d=3):
    return np.random.default_rng(seed).integers(0, k, (n, d)).astype(np.float32)


def cases():
    yield "kat_line5", np.arange(5, dtype=np.float32)[:, None], 5, [0], "fps", 0
    yield "kat_pairs", np.array([[0], [0], [1], [1]], np.float32), 4, [0], "fps", 0
    yield "kat_same6", np.zeros((6, 3), np.float32), 6, [2], "fps", 0
    yield "c1_start0", c1(), 1024, [0], "fps", 0
    yield "c1_start4095", c1(), 1024, [4095], "fps", 0
    yield "c1_multi", c1(), 1024, [0, 23, 64, 128], "fps", 0
    yield "c1_full", c1()[:600], 600, [7], "fps", 0
    yield "lattice3k", lattice(3000, 6, 0), 800, [5], "fps", 0
    yield "lattice_dups", np.repeat(lattice(500, 4, 1), 3, axis=0), 1500, [0], "fps", 0
    rng = np.random.default_rng(5)
    for d in (1, 2, 4, 5, 8, 11):
        yield f"dim{d}", rng.random((700, d)).astype(np.float32), 200, [3], "fps", 0
    # denormal squares: coordinates ~1e-21 give squared gaps ~1e-42
    yield "denormal", (rng.random((500, 3)) * 1e-21).astype(np.float32), 100, [0], "fps", 0
    # huge values: squared gaps overflow to +inf; one NaN row
    big = (rng.random((300, 3)) * 4e19).astype(np.float32)
    big[17] = np.nan
    yield "overflow_nan", big, 60, [0], "fps", 0
    wide = rng.standard_normal((2000, 3)).astype(np.float32) * np.float32(1000)
    yield "gauss_multi", wide, 500, [1999, 0, 1000], "fps", 0
    # bucket entry (start_idx is a position in the tree order)
    for h in (1, 3, 5, 7):
        for s in (0, 17, 4095):
            yield f"kd_c1_h{h}_s{s}", c1(), 1024, [s], "kdline", h
    for h in (2, 4, 6):
        yield f"kd_lattice_h{h}", lattice(3000, 6, 0), 800, [5], "kdline", h
    yield "kd_dups_h3", np.repeat(lattice(300, 3, 2), 4, axis=0), 1200, [11], "kdline", 3
    for d in (1, 2, 5, 8):
        yield f"kd_dim{d}_h4", rng.random((900, d)).astype(np.float32), 300, [4], "kdline", 4
    yield "kd_same_h2", np.ones((40, 3), np.float32), 40, [3], "kdline", 2
    # NPDU (h holds the window w)
    for w, s in ((64, 0), (16, 2000), (4095, 17), (3, 4095)):
        yield f"npdu_c1_w{w}_s{s}", c1(), 1024, [s], "npdu", w
    yield "npdu_lattice_w40", lattice(3000, 6, 0), 800, [5], "npdu", 40
    yield "npdu_sorted_w32", np.sort(rng.random((2000, 2)).astype(np.float32), axis=0), 500, [100], "npdu", 32


def main():
    R = O.load_reference()
    if R is None:
        raise SystemExit("oracle/_ref is not built (needs /root/reference): make -C oracle")
    seen = {}
    for name, pc, m, start, kind, h in cases():
        pc = np.ascontiguousarray(pc, dtype=np.float32)
        key = (pc.shape, pc.tobytes())
        if kind == "fps":
            st = start[0] if len(start) == 1 else list(start)
            out = R.fps_sampling(pc, m, st)
        elif kind == "npdu":
            out = R.fps_npdu(pc, m, h, start[0])
        else:
            out = R.bucket_fps_kdline(pc, m, h, start[0])
        data = dict(m=np.int64(m), start=np.array(start, np.uint64), kind=np.array(kind),
                    h=np.int64(h), out=np.asarray(out, np.uint64))
        if key in seen:
            data["pc_from"] = np.array(seen[key])
        else:
            seen[key] = name
            data["pc"] = pc
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
        print(f"{name:20s} n={pc.shape[0]:5d} d={pc.shape[1]} m={m:5d} {kind} -> {out[:6]}")


if __name__ == "__main__":
    main()

"""Write full-length golden fixtures (tests/golden/full/*.npz) from the reference.

The reference's fps_sampling (src/lib.cpp:188-246), compiled from its own
sources into oracle/_ref/ by oracle/Makefile, is run on the benchmark
workloads at their full sizes and every pick is recorded.  The clouds are
regenerated from their seeds at test time (bench.py's generators), so each
fixture stores the SHA-256 of the float32 cloud it was made from instead of
the cloud itself:

  gauss1m   N = 2^20 Gaussian mixture (default_rng(1)), M = 16384, start 0
  gauss1m1  the same generator at N = 2^20 + 1
  gauss2m   the same generator at N = 2^21
  lattice16m  N = 2^24 lattice {0..255}^3 with 25 % duplicates (default_rng(4)), M = 4096
  uniform4k   N = 4096 uniform in [0,1)^3 (np.random.seed(42)), M = 1024, start 0
            (bench/test_bench.py:29-33)
  uniform50k  N = 50000, same generator, M = 4096, start 0:
            the reference's own 50k bench case (bench/test_bench.py:89-93)
  uniform100k N = 100000, same generator, M = 50000, start 0 (bench/test_bench.py:143-147)
  kdline50k, kdline100k  the same two clouds through the reference's bucket_fps_kdline
            (src/wrapper.hpp:118-132) with h = 7, start position 0
            (bench/test_bench.py:131-135, 178-182); kdline4k the 4k cloud with h = 5
            (bench/test_bench.py:70-74)
  npdu4k, npdu50k, npdu100k  the 4k, 50k and 100k clouds through the reference's
            fps_npdu_sampling with the wrapper's default window, start 0
            (bench/test_bench.py:42-46, 96-100, 150-154)
  batch256  the 256 ragged clouds (default_rng(2)), M_i = N_i // 4: per cloud the
            SHA-256 of its picks (uint64 little-endian) and its first 16 picks
  serve96   the 96 serving-stream requests (default_rng(3)), M = N * U{1/8,1/4,1/2}:
            per request the SHA-256 of its picks and its first 16 picks

Single clouds store every pick (uint32).  Run in a container where
/root/reference exists (a few CPU-minutes per large cloud; runs in parallel):

    make -C oracle && python tests/golden/make_full_length.py [name ...]
"""
from __future__ import annotations

import hashlib
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "full")
sys.path.insert(0, ROOT)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def picks_sha(out) -> str:
    return sha(np.asarray(out, dtype="<u8"))


def _ref():
    from oracle import oracle as O

    R = O.load_reference()
    if R is None:
        raise SystemExit("oracle/_ref is not built (needs /root/reference): make -C oracle")
    return R


def _single(task):
    name, n, m = task
    import bench

    if name.startswith("lattice"):
        pc = bench.gen_lattice_dups(n)
    elif name.startswith("uniform") or name in KDLINE or name in NPDU:
        pc = bench.gen_uniform_n(n)
    else:
        pc = bench.gen_gauss_mixture(n)
    t0 = time.time()
    if name in KDLINE:
        out = np.asarray(_ref().bucket_fps_kdline(pc, m, KDLINE[name], 0), np.uint64)
    elif name in NPDU:
        out = np.asarray(_ref().fps_npdu(pc, m, NPDU[name], 0), np.uint64)
    else:
        out = np.asarray(_ref().fps_sampling(pc, m, 0), np.uint64)
    assert out.max() < 2**32
    np.savez_compressed(os.path.join(OUT, name + ".npz"), n=np.int64(n), m=np.int64(m),
                        start=np.uint64(0), cloud_sha256=np.array(sha(pc)),
                        out=out.astype(np.uint32))
    return f"{name}: n={n} m={m} {time.time() - t0:.0f} s"


def _one_cloud(args):
    pc, m = args
    out = np.asarray(_ref().fps_sampling(pc, m, 0), np.uint64)
    head = np.zeros(16, np.uint32)
    head[:min(16, m)] = out[:16]
    return picks_sha(out), head


def _many(name, clouds, ms, pool):
    t0 = time.time()
    order = sorted(range(len(clouds)), key=lambda i: -clouds[i].shape[0] * ms[i])
    res = dict(zip(order, pool.map(_one_cloud, [(clouds[i], ms[i]) for i in order], chunksize=1)))
    np.savez_compressed(
        os.path.join(OUT, name + ".npz"),
        n=np.array([pc.shape[0] for pc in clouds], np.int64), m=np.array(ms, np.int64),
        cloud_sha256=np.array([sha(pc) for pc in clouds]),
        picks_sha256=np.array([res[i][0] for i in range(len(clouds))]),
        head=np.stack([res[i][1] for i in range(len(clouds))]))
    return f"{name}: {len(clouds)} clouds {time.time() - t0:.0f} s"


SINGLES = {"lattice16m": (1 << 24, 4096), "gauss2m": (1 << 21, 16384),
           "gauss1m": (1 << 20, 16384), "gauss1m1": ((1 << 20) + 1, 16384),
           "uniform100k": (100_000, 50_000), "uniform50k": (50_000, 4096),
           "uniform4k": (4096, 1024),
           "kdline100k": (100_000, 50_000), "kdline50k": (50_000, 4096),
           "kdline4k": (4096, 1024),
           "npdu4k": (4096, 1024), "npdu50k": (50_000, 4096), "npdu100k": (100_000, 50_000)}
# bucket_fps_kdline (src/wrapper.hpp:118-132) with these tree heights
KDLINE = {"kdline100k": 7, "kdline50k": 7, "kdline4k": 5}
# fps_npdu_sampling (src/lib.cpp:272-340) with the Python wrapper's default window
# int(N / M * 16) (src/fpsample/__init__.py:97)
NPDU = {"npdu4k": 64, "npdu50k": 195, "npdu100k": 32}


def main(names):
    import bench

    os.makedirs(OUT, exist_ok=True)
    _ref()
    names = names or list(SINGLES) + ["serve96", "batch256"]
    workers = int(os.environ.get("JOBS", "7"))
    with Pool(workers) as pool:
        pending = [pool.apply_async(_single, ((k, *SINGLES[k]),)) for k in names if k in SINGLES]
        if "serve96" in names:
            reqs = bench.gen_serving_stream()
            print(_many("serve96", [pc for pc, _ in reqs], [m for _, m in reqs], pool), flush=True)
        if "batch256" in names:
            clouds, ms = bench.gen_ragged_batch()
            print(_many("batch256", clouds, ms, pool), flush=True)
        for p in pending:
            print(p.get(), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])

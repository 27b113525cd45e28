#!/usr/bin/env python3
"""Benchmark of the FPS hot path on one MI355X.

A step is one vanilla FPS run (fps_sampling semantics, src/lib.cpp:188-246)
over one synthetic cloud already resident in HBM: one call of
fpsx_fps_sampling_device, i.e. the exchange-area clear, for large 3-D clouds
the spatial ordering of the points (bounding box, Hilbert codes, two radix
sorts, gather), and the single persistent FPS launch that makes all M-1
passes.  The kernel shape comes from the library's planner; --autotune on
times every plan first (untimed) instead.  The npdu* workloads time
fps_npdu_sampling (src/lib.cpp:272-340) the same way, through
fpsx_fps_npdu_sampling_device on row-major device data.  The kdline*
workloads time bucket_fps_kdline_sampling (src/fpsample/__init__.py:174-206),
whose entry takes host arrays: each step there is the whole call, the host
KD-line build and the PCIe copies included.  Workloads taken from the
reference's own bench (bench/test_bench.py) also report the mean time its
README publishes for them (reference_published).

Default workload (the metric's configuration): one cloud of N = 1,048,576
points, D = 3 float32, a mixture of 64 isotropic Gaussian clusters (sigma
0.02, centres uniform in [0,1)^3, numpy default_rng(1)), M = 16,384 samples,
start_idx = 0.  The same JSON line then carries the metric's second half:
  serve     the serving workload (96 requests, N log-uniform in [2^10, 2^20],
            M = N * U{1/8,1/4,1/2}): capacity as one burst, p50/p99 latency
            under Poisson arrivals at 60 % of it, and the same requests run
            one launch per request, with outputs compared;
  batch256  the 256 ragged clouds in one persistent launch against one
            launch per cloud, outputs compared.

Prints one JSON line.  value = samples/s of the whole job (M x steps / wall
time between two device synchronisations).  roofline: what bounds the
kernel.  Multi-workgroup resident and pruned kernels are bound by the
latency of their per-pass cross-workgroup exchange, not by HBM: the line
reports exchanges per second against what the exchange alone sustains at
the same grid and candidate size (fpsx_probe_exchange, measured in the same
run), plus the measured HBM traffic and LDS bytes (committed rocprofv3
counter summaries) and, as vs_vanilla_hbm, the vanilla loop's 20 B per
point-pass over the kernel time.  cpu_baseline: the reference's
fps_sampling compiled from its own sources (oracle/_ref; the oracle's C
restatement where that build is absent), single-threaded and pinned to one
core of this host, over all M samples of the same cloud when they fit the
time budget.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_POINT_PASS = 20


# ---------------------------------------------------------------- workloads
def gen_uniform_4096(seed: int = 42) -> np.ndarray:
    """N=4096 uniform in [0,1)^3 from np.random.seed(42) (bench/test_bench.py:19-21)."""
    np.random.seed(seed)
    return np.random.rand(4096, 3).astype(np.float32)


def gen_uniform_n(n: int, seed: int = 42) -> np.ndarray:
    """N uniform in [0,1)^3 from np.random.seed(42) (bench/test_bench.py:19-21)."""
    np.random.seed(seed)
    return np.random.rand(n, 3).astype(np.float32)


def gen_gauss_mixture(n: int = 1 << 20, seed: int = 1) -> np.ndarray:
    """64 isotropic Gaussian clusters, sigma 0.02, centres uniform in [0,1)^3."""
    rng = np.random.default_rng(seed)
    centers = rng.random((64, 3))
    lab = rng.integers(0, 64, n)
    return (centers[lab] + 0.02 * rng.standard_normal((n, 3))).astype(np.float32)


def gen_lattice_dups(n: int = 1 << 24, seed: int = 4) -> np.ndarray:
    """Integer lattice {0..255}^3: 75% distinct sites, 25% exact copies, shuffled."""
    rng = np.random.default_rng(seed)
    distinct = n * 3 // 4
    sites = rng.choice(1 << 24, distinct, replace=False)
    copies = sites[rng.integers(0, distinct, n - distinct)]
    codes = np.concatenate([sites, copies])
    rng.shuffle(codes)
    return np.stack([(codes >> 16) & 255, (codes >> 8) & 255, codes & 255],
                    axis=1).astype(np.float32)


def gen_ragged_batch(seed: int = 2, k: int = 256):
    """256 clouds, N_i log-uniform in [1024, 65536], M_i = N_i // 4; even i uniform
    cube, odd i unit-sphere surface + N(0, 0.01^2) noise."""
    rng = np.random.default_rng(seed)
    clouds, ms = [], []
    for i in range(k):
        n = int(np.floor(np.exp(rng.uniform(np.log(1024), np.log(65536)))))
        if i % 2 == 0:
            pc = rng.random((n, 3))
        else:
            v = rng.standard_normal((n, 3))
            pc = v / np.linalg.norm(v, axis=1, keepdims=True) + 0.01 * rng.standard_normal((n, 3))
        clouds.append(pc.astype(np.float32))
        ms.append(n // 4)
    return clouds, ms


def gen_serving_stream(seed: int = 3, k: int = 96):
    """Requests: N log-uniform in [2^10, 2^20], uniform cube, M = N * U{1/8,1/4,1/2}."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(k):
        n = int(np.floor(np.exp(rng.uniform(np.log(1 << 10), np.log(1 << 20)))))
        m = int(n * rng.choice([0.125, 0.25, 0.5]))
        reqs.append((rng.random((n, 3)).astype(np.float32), max(1, m)))
    return reqs


WORKLOADS = {
    "gauss1m": dict(gen=gen_gauss_mixture, m=16384,
                    desc="one cloud N=1048576 D=3 f32, 64 Gaussian clusters sigma 0.02, "
                         "M=16384, start_idx=0"),
    # just beyond the 1 M-point cloud (one point per thread in LDS) and twice it
    # (the wide paired kernel, 128 workgroups): the same generator, M=16384
    "gauss1m1": dict(gen=lambda: gen_gauss_mixture((1 << 20) + 1), m=16384,
                     desc="one cloud N=1048577 D=3 f32, 64 Gaussian clusters sigma 0.02, "
                          "M=16384, start_idx=0"),
    "gauss2m": dict(gen=lambda: gen_gauss_mixture(1 << 21), m=16384,
                    desc="one cloud N=2097152 D=3 f32, 64 Gaussian clusters sigma 0.02, "
                         "M=16384, start_idx=0"),
    # the reference's own vanilla bench cases (bench/test_bench.py:29-33, 89-93, 143-147;
    # its README publishes their mean times on a Xeon Gold 6326, one thread, with a
    # random start where these use start_idx 0)
    "uniform4k": dict(gen=gen_uniform_4096, m=1024,
                      desc="one cloud N=4096 D=3 f32 uniform [0,1)^3 seed 42, M=1024, start_idx=0",
                      published=(12.5875e-3, "README.md:123")),
    "uniform50k": dict(gen=lambda: gen_uniform_n(50_000), m=4096,
                       desc="one cloud N=50000 D=3 f32 uniform [0,1)^3 seed 42, M=4096, "
                            "start_idx=0",
                       published=(593.2049e-3, "README.md:136")),
    "uniform100k": dict(gen=lambda: gen_uniform_n(100_000), m=50_000,
                        desc="one cloud N=100000 D=3 f32 uniform [0,1)^3 seed 42, M=50000, "
                             "start_idx=0",
                        published=(14561.1574e-3, "README.md:148")),
    "lattice16m": dict(gen=gen_lattice_dups, m=4096,
                       desc="one cloud N=16777216 D=3 f32 lattice {0..255}^3, 25% duplicates, "
                            "M=4096, start_idx=0"),
    # bucket_fps_kdline_sampling on the same clouds (h = 5 at 4k, 7 at 50k and 100k, the
    # reference's fastest heights there; bench/test_bench.py:70-74, 131-135, 178-182): host
    # arrays in and out, so each step includes the host KD-line build and
    # the PCIe copies; the README publishes the reference's mean times
    "kdline4k": dict(gen=gen_uniform_4096, m=1024, kind="kdline", h=5,
                     desc="bucket_fps_kdline_sampling h=5 on host arrays, one cloud N=4096 "
                          "D=3 f32 uniform [0,1)^3 seed 42, M=1024, start_idx=0",
                     published=(1.8301e-3, "README.md:117")),
    "kdline50k": dict(gen=lambda: gen_uniform_n(50_000), m=4096, kind="kdline", h=7,
                      desc="bucket_fps_kdline_sampling h=7 on host arrays, one cloud N=50000 "
                           "D=3 f32 uniform [0,1)^3 seed 42, M=4096, start_idx=0",
                      published=(25.0179e-3, "README.md:130")),
    "kdline100k": dict(gen=lambda: gen_uniform_n(100_000), m=50_000, kind="kdline", h=7,
                       desc="bucket_fps_kdline_sampling h=7 on host arrays, one cloud N=100000 "
                            "D=3 f32 uniform [0,1)^3 seed 42, M=50000, start_idx=0",
                       published=(233.3726e-3, "README.md:142")),
    # windowed-update FPS on the reference's own bench cases (bench/test_bench.py:42-46,
    # 96-100, 150-154): w is the Python wrapper's default int(N / M * 16)
    "npdu4k": dict(gen=gen_uniform_4096, m=1024, kind="npdu",
                   desc="fps_npdu_sampling, one cloud N=4096 D=3 f32 uniform [0,1)^3 seed 42, "
                        "M=1024, w=64, start_idx=0",
                   published=(3.9514e-3, "README.md:120")),
    "npdu50k": dict(gen=lambda: gen_uniform_n(50_000), m=4096, kind="npdu",
                    desc="fps_npdu_sampling, one cloud N=50000 D=3 f32 uniform [0,1)^3 seed 42, "
                         "M=4096, w=195, start_idx=0",
                    published=(180.5911e-3, "README.md:134")),
    "npdu100k": dict(gen=lambda: gen_uniform_n(100_000), m=50_000, kind="npdu",
                     desc="fps_npdu_sampling, one cloud N=100000 D=3 f32 uniform [0,1)^3 seed 42, "
                          "M=50000, w=32, start_idx=0",
                     published=(4333.3156e-3, "README.md:147")),
}


# ---------------------------------------------------------------- CPU baseline
def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


class _pinned:
    """Run the enclosed CPU timing on one core (the first this process may use)."""

    def __enter__(self):
        self.saved = None
        if hasattr(os, "sched_setaffinity"):
            self.saved = os.sched_getaffinity(0)
            self.core = min(self.saved)
            os.sched_setaffinity(0, {self.core})
        else:
            self.core = None
        return self

    def __exit__(self, *exc):
        if self.saved is not None:
            os.sched_setaffinity(0, self.saved)


def cpu_baseline(pc: np.ndarray, m_full: int, target_s: float, w: int = -1,
                 h: int = -1) -> dict:
    """w < 0: fps_sampling; else fps_npdu_sampling with window w; h >= 0:
    bucket_fps_kdline with tree height h instead.  The
    reference (oracle/_ref) single-threaded on one pinned core, over all m_full
    samples when the calibrated estimate fits target_s (else the largest
    prefix that does, stated in `sample`); the oracle's C restatement is timed
    beside it on a shorter prefix (port_value)."""
    from oracle import oracle as O

    ref = O.load_reference()

    def runner(which):
        if h >= 0:
            if which == "reference":
                return lambda m: ref.bucket_fps_kdline(pc, m, h, 0)
            return lambda m: O.kdline_quickfps(pc, m, h, 0)
        if which == "reference":
            if w < 0:
                return lambda m: ref.fps_sampling(pc, m, 0)
            return lambda m: ref.fps_npdu(pc, m, w, 0)
        if w < 0:
            return lambda m: O.fps(pc, m, 0)
        return lambda m: O.npdu(pc, m, w, 0)

    kind = "reference" if ref is not None else "port"
    run = runner(kind)
    n = pc.shape[0]
    with _pinned() as pin:
        # calibrate on a few passes, then size the sample to the time budget
        m_cal = max(2, min(m_full, int(2e8 // max(n, 1)) + 2))
        t0 = time.perf_counter()
        run(m_cal)
        per_pass = (time.perf_counter() - t0) / max(m_cal - 1, 1)
        m_s = int(min(m_full, max(m_cal, target_s / max(per_pass, 1e-12))))
        times = []
        while True:
            t0 = time.perf_counter()
            run(m_s)
            times.append(time.perf_counter() - t0)
            # a whole run under a second is repeated (for up to 2 s or 200 runs)
            # and its median taken: one short run is at the mercy of the host
            if m_s < m_full or sum(times) >= min(2.0, target_s) or len(times) >= 200:
                break
        dt = float(np.median(times))
        port = None
        if kind == "reference":
            m_p = max(2, min(m_s, m_cal * 4))
            t0 = time.perf_counter()
            runner("port")(m_p)
            port = m_p / (time.perf_counter() - t0)
    what = (f"all {m_s} samples" if m_s == m_full
            else f"the first {m_s} of the {m_full} samples ({100.0 * m_s / m_full:.1f} %)")
    runs = f"median of {len(times)} runs, " if len(times) > 1 else ""
    out = {
        "value": round(m_s / dt, 3),
        "unit": "samples/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{what} of the same cloud, start_idx=0, {runs}{dt:.4g} s single-threaded on core "
                  f"{pin.core} of a {cpu_model()} (os.cpu_count()={os.cpu_count()})",
    }
    if port is not None:
        out["port_value"] = round(port, 3)
        out["port_over_reference"] = round(port / out["value"], 3)
    return out


def accuracy_and_bucket(pc: np.ndarray, gpu_picks: np.ndarray, m_full: int, prefix: int = 1024,
                        bucket_h: int = 9) -> dict:
    """CPU-leg checks of the GPU result (the oracle only as the checker):
    index agreement and coverage-radius relative error of the first `prefix`
    GPU picks against the same greedy loop in double precision; and the
    bucket entry (bucket_fps_kdline, h=9) timed on the GPU beside the
    reference's own bucket sampler on one host core, with their picks
    compared."""
    from oracle import oracle as O

    import fpsample_amd as F

    k = min(prefix, gpu_picks.size)
    acc = O.accuracy_vs_f64(pc, gpu_picks[:k], int(gpu_picks[0]))
    acc["against"] = "double-precision greedy loop (CPU), same start"
    out = {"accuracy_vs_fp64": acc}
    ref = O.load_reference()
    F.bucket_fps_kdline_sampling(pc, m_full, bucket_h, start_idx=0)  # warm-up (threads, buffers)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        gpu = F.bucket_fps_kdline_sampling(pc, m_full, bucket_h, start_idx=0)
        ts.append(time.perf_counter() - t0)
    b = {"h": bucket_h, "n_samples": m_full,
         "gpu": {"value": round(m_full / min(ts), 3), "unit": "samples/s",
                 "ms": round(min(ts) * 1e3, 3),
                 "note": "host arrays in and out: KD-line permutation on host threads, "
                         "upload, GPU sampling, download"}}
    if ref is not None:
        t0 = time.perf_counter()
        want = ref.bucket_fps_kdline(pc, m_full, bucket_h, 0)
        dt = time.perf_counter() - t0
        b["cpu_reference"] = {"value": round(m_full / dt, 3), "unit": "samples/s",
                              "ms": round(dt * 1e3, 3), "cores": 1}
        b["identical_picks"] = bool(np.array_equal(gpu, want))
        b["gpu_over_cpu"] = round(dt / min(ts), 2)
    out["bucket_kdline"] = b
    # The drop-in Python API end to end, as the reference's callers use it: a
    # float64 host array in (the wrapper casts it to float32 in Fortran order,
    # src/fpsample/__init__.py:62), numpy uint64 picks out.  Upload, spatial
    # order, kernel and download are all inside; never the headline value.
    pc64 = pc.astype(np.float64)
    F.fps_sampling(pc64, m_full, start_idx=0)  # warm-up
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        api = F.fps_sampling(pc64, m_full, start_idx=0)
        ts.append(time.perf_counter() - t0)
    out["drop_in_api"] = {"value": round(m_full / min(ts), 3), "unit": "samples/s",
                          "ms": round(min(ts) * 1e3, 3),
                          "identical_to_device_entry": bool(np.array_equal(api, gpu_picks)),
                          "note": "fpsample_amd.fps_sampling on a float64 host array: cast, "
                                  "upload over PCIe, GPU sampling, download"}
    return out


LDS_PEAK_GBS = 256 * 128 * 2.4  # 256 CUs x 128 B per clock x 2.4 GHz


def _read_summary(workload: str, plan: dict, what: str):
    """key=value lines of profiles/fps_<kind>_<workload>_<what>.txt, when that
    rocprofv3 summary was taken with the plan this run uses (else None)."""
    path = os.path.join(ROOT, "profiles", f"fps_{plan.get('kind')}_{workload}_{what}.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.strip().split("=", 1)
            kv[k] = v
    want = f"{plan.get('kind')}:{plan.get('block')}:{plan.get('ppt')}:{plan.get('grid')}"
    return kv if kv.get("plan") == want else None


def _float(kv, key):
    try:
        return float(kv[key]) if kv is not None else None
    except (KeyError, ValueError):
        return None


def probe_exchange(grid: int, words: int, rounds: int = 20000) -> float:
    """Microseconds per round of the exchange alone (fpsx_probe_exchange)."""
    import ctypes

    from fpsample_amd import device as D

    L = D.lib()
    L.fpsx_probe_exchange.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_double)]
    us = ctypes.c_double()
    best = None
    for _ in range(3):  # the fastest of three launches (the first also warms up)
        D._raise(L.fpsx_probe_exchange(grid, words, rounds, None, ctypes.byref(us)),
                 "fpsx_probe_exchange")
        best = us.value if best is None else min(best, us.value)
    return best


def launch_exchanges() -> int:
    import ctypes

    from fpsample_amd import device as D

    L = D.lib()
    L.fpsx_launch_exchanges.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    x = ctypes.c_uint64()
    D._raise(L.fpsx_launch_exchanges(None, ctypes.byref(x)), "fpsx_launch_exchanges")
    return int(x.value)


def roofline(workload: str, plan: dict, n: int, m: int, kernel_s: float, npdu: bool,
             w: int = -1) -> dict:
    """What bounds the FPS kernel, measured in this run.

    Streaming kernels (plain or generic) read and write every point's state
    every pass: bound "hbm", the vanilla loop's 20 B per point-pass over the
    kernel time against the 8 TB/s peak.  Multi-workgroup resident and pruned
    kernels keep the state on chip: bound "latency", the exchanges the
    launch made per second against the rate the exchange alone sustains at
    the same grid and candidate size (frac = exchange floor / kernel time,
    at most 1).  The measured HBM traffic and LDS bytes come from the
    committed rocprofv3 summaries for this plan."""
    if npdu:
        alg = (20 * (min(w, n - 1) + 1) + 4 * n) * (m - 1)
    else:
        alg = BYTES_PER_POINT_PASS * n * (m - 1)
    vanilla = {"algorithmic_bytes_per_launch": alg,
               "achieved": round(alg / kernel_s / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(alg / kernel_s / 1e9 / HBM_PEAK_GBS, 4)}
    traffic = _float(_read_summary(workload, plan, "traffic"), "traffic_bytes_per_launch")
    counters = _read_summary(workload, plan, "counters")
    lds_bytes = _float(counters, "lds_bytes_per_launch")
    out = {"kernel_ms": round(kernel_s * 1e3, 4), "traffic": traffic}
    if traffic is not None:
        out["hbm"] = {"achieved": round(traffic / kernel_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(traffic / kernel_s / 1e9 / HBM_PEAK_GBS, 5)}
    if lds_bytes is not None:
        out["lds"] = {"bytes_per_launch": lds_bytes,
                      "achieved": round(lds_bytes / kernel_s / 1e9, 2), "peak": LDS_PEAK_GBS,
                      "unit": "GB/s", "frac": round(lds_bytes / kernel_s / 1e9 / LDS_PEAK_GBS, 5),
                      "bank_conflict_cycles": _float(counters, "SQ_LDS_BANK_CONFLICT")}
    kind, grid = plan.get("kind"), int(plan.get("grid", 1))
    if kind == "stream" and not npdu:
        out.update(bound="hbm", achieved=vanilla["achieved"], peak=HBM_PEAK_GBS, unit="GB/s",
                   frac=vanilla["frac"], algorithmic_bytes_per_launch=alg)
        return out
    out["vs_vanilla_hbm"] = vanilla
    if grid < 2:
        # One workgroup: no exchange; the pass is the update and the
        # workgroup's reduction chain.
        out.update(bound="latency", achieved=None, peak=None, unit="exchanges/s", frac=None,
                   note="one workgroup, no cross-workgroup exchange")
        return out
    exchanges = launch_exchanges()
    if npdu:
        # The paired windowed-update kernel (10-word candidates) settles up to
        # two picks per exchange; the one-pick kernels (5 words) make m - 1.
        words = 10 if exchanges and exchanges < m - 1 else 5
    else:
        words = 10 if kind in ("pair", "spair") else 5
    floor_us = probe_exchange(grid, words)
    if not exchanges:
        out.update(bound="latency", achieved=None, peak=round(1e6 / floor_us, 1),
                   unit="exchanges/s", frac=None, exchange_floor_us=round(floor_us, 4))
        return out
    achieved = exchanges / kernel_s
    out.update(bound="latency", achieved=round(achieved, 1), peak=round(1e6 / floor_us, 1),
               unit="exchanges/s", frac=round(achieved * floor_us / 1e6, 4),
               exchanges_per_launch=exchanges, exchange_floor_us=round(floor_us, 4),
               us_per_exchange=round(kernel_s * 1e6 / exchanges, 4),
               floor_probe={"grid": grid, "words": words, "threads": 1024})
    if kind in ("spruned", "spair"):
        out["note"] = ("the floor is the exchange alone; a streamed round also waits for the memory "
                       "round trip of its listed blocks (a paired round for two: the update's, then "
                       "the lookahead's)")
    return out


def _reference_check(name: str, picks_list) -> object:
    """True/False: every pick sequence's SHA-256 equals the one the reference
    wrote into the full-length fixture `name`; None where that fixture is absent."""
    try:
        from tests.golden import full_length as FL
    except ImportError:
        return None
    if not FL.available(name):
        return None
    d = FL.load(name)
    if len(d["picks_sha256"]) != len(picks_list):
        return False
    return all(FL.picks_sha256(p) == str(h) for p, h in zip(picks_list, d["picks_sha256"]))


def _ref_fps_worker(item):
    """Process-pool task: the reference's fps_sampling on one cloud (seconds)."""
    pc, m = item
    sys.path.insert(0, ROOT)
    from oracle import oracle as O

    ref = O.load_reference()
    run = (lambda: ref.fps_sampling(pc, m, 0)) if ref is not None else (lambda: O.fps(pc, m, 0))
    t0 = time.perf_counter()
    run()
    return time.perf_counter() - t0


def batch_cpu_baseline(clouds, ms, target_s: float) -> dict:
    """The 256 ragged clouds on the host: the reference (oracle/_ref, else the
    oracle's C restatement) single-threaded on one pinned core over whole
    clouds in order until target_s has passed, and every cloud over a process
    pool of up to 16 workers (the GPU host's share), each cloud one task."""
    from oracle import oracle as O

    ref = O.load_reference()
    kind = "reference" if ref is not None else "port"
    done_m, done_t, k = 0, 0.0, 0
    with _pinned() as pin:
        for pc, m in zip(clouds, ms):
            t0 = time.perf_counter()
            if ref is not None:
                ref.fps_sampling(pc, m, 0)
            else:
                O.fps(pc, m, 0)
            done_t += time.perf_counter() - t0
            done_m += m
            k += 1
            if done_t >= target_s:
                break
    total_m = sum(ms)
    out = {"value": round(done_m / done_t, 3), "unit": "samples/s", "cores": 1, "kind": kind,
           "sample": f"the first {k} of the {len(clouds)} clouds ({100.0 * done_m / total_m:.1f} % of the "
                     f"samples), {done_t:.2f} s single-threaded on core {pin.core} of a {cpu_model()}"}
    workers = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count())
    try:
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor

        # spawned workers: fresh interpreters that never touch the GPU
        with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
            list(ex.map(_ref_fps_worker, [(clouds[0][:64], 8)] * workers))  # start every worker
            order = sorted(range(len(clouds)), key=lambda i: -clouds[i].shape[0] * ms[i])
            t0 = time.perf_counter()
            list(ex.map(_ref_fps_worker, [(clouds[i], ms[i]) for i in order]))
            wall = time.perf_counter() - t0
        out["pool"] = {"value": round(total_m / wall, 3), "unit": "samples/s", "cores": workers,
                       "sample": f"all {len(clouds)} clouds, longest first, {wall:.2f} s over {workers} "
                                 f"worker processes"}
    except Exception as exc:  # a host without process support: single-core figure only
        out["pool"] = {"error": str(exc)}
    return out


def run_batch(args) -> dict:
    """Batched ragged clouds: one persistent launch (HIP graph) vs one launch per cloud."""
    from fpsample_amd import device as D
    from fpsample_amd.batch import DeviceBatch

    clouds, ms = gen_ragged_batch()
    starts = [0] * len(clouds)
    total_m = sum(ms)
    b = DeviceBatch(clouds, ms, starts, graph=True, autotune=True)
    info = b.info()
    for _ in range(args.warmup):
        b.run()
    b.status()
    D.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    b.status()
    t1 = time.perf_counter()
    batched = b.picks()
    wall_b = (t1 - t0) / args.steps
    # one cloud per launch, same device-resident data, planner's choice per cloud
    dev = [D.DeviceArray.from_numpy(np.ascontiguousarray(pc.T)) for pc in clouds]
    outs = [D.DeviceArray((m,), np.uint64) for m in ms]
    d_start = D.DeviceArray.from_numpy(np.zeros(1, np.uint64))

    def per_cloud():
        for pc, dp, do, m in zip(clouds, dev, outs, ms):
            D.fps_sampling_device(dp, pc.shape[0], 3, m, d_start, 1, do)

    per_cloud()
    D.synchronize()
    D.device_status()
    t0 = time.perf_counter()
    reps = max(1, args.steps // 4)
    for _ in range(reps):
        per_cloud()
    D.synchronize()
    t1 = time.perf_counter()
    wall_p = (t1 - t0) / reps
    single = [o.numpy() for o in outs]
    same = all(np.array_equal(a, o) for a, o in zip(batched, single))
    b.close()
    cpu = batch_cpu_baseline(clouds, ms, min(10.0, args.cpu_seconds)) if args.cpu_seconds > 0 else None
    return {
        "metric": "samples/s (256 ragged clouds, one persistent launch)",
        "value": round(total_m / wall_b, 3),
        "unit": "samples/s",
        "ms_per_step": round(wall_b * 1e3, 4), "steps": args.steps,
        "config": {"workload": "B=256 clouds, N_i log-uniform [1024, 65536], M_i = N_i/4, D=3 f32, "
                               "half uniform cube / half unit sphere + N(0,0.01^2), start_idx=0, "
                               "HIP graph, autotuned shape",
                   "total_points": int(sum(pc.shape[0] for pc in clouds)), "total_samples": total_m,
                   "batch_kernel": info},
        "one_cloud_per_launch": {"value": round(total_m / wall_p, 3), "unit": "samples/s",
                                 "ms_per_step": round(wall_p * 1e3, 3)},
        "speedup_vs_one_cloud_per_launch": round(wall_p / wall_b, 3),
        "target_speedup": 5.0,
        "identical_output": bool(same),
        "identical_to_reference": _reference_check("batch256", batched),
        "cpu_baseline": cpu,
    }


def run_serve(args) -> dict:
    """Serving workload: capacity (every request at once), latency under
    Poisson arrivals at 60 % of that capacity, and the same requests one
    launch per request on device-resident clouds."""
    from fpsample_amd import device as D
    from fpsample_amd.serve import Server

    reqs = gen_serving_stream()
    total_m = sum(m for _, m in reqs)
    with Server(hbm_budget_bytes=16 << 30, chunk_passes=args.serve_chunk) as srv:
        srv.wait(srv.submit(reqs[0][0], reqs[0][1], 0))  # warm-up
        t0 = time.perf_counter()
        ids = [srv.submit(pc, m, 0) for pc, m in reqs]
        burst = [srv.wait(r) for r in ids]
        t_burst = time.perf_counter() - t0
        cap = len(reqs) / t_burst
        qps = 0.6 * cap
        gaps = np.random.default_rng(5).exponential(1.0 / qps, len(reqs))
        ids = []
        t0 = time.perf_counter()
        t_next = t0
        for (pc, m), g in zip(reqs, gaps):
            t_next += g
            now = time.perf_counter()
            if t_next > now:
                time.sleep(t_next - now)
            ids.append(srv.submit(pc, m, 0))
        streamed = [srv.wait(r) for r in ids]
        t_stream = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([x[1] for x in streamed]) * 1e3
    served = [x[0] for x in burst]
    same_stream = all(np.array_equal(a[0], b) for a, b in zip(streamed, served))
    # one request per launch: clouds resident first (no upload timed), the
    # planner's plan per request, launches back to back on one stream
    dev = [D.DeviceArray.from_numpy(np.ascontiguousarray(pc.T)) for pc, _ in reqs]
    outs = [D.DeviceArray((m,), np.uint64) for _, m in reqs]
    d_start = D.DeviceArray.from_numpy(np.zeros(1, np.uint64))
    D.synchronize()
    t0 = time.perf_counter()
    for (pc, m), dp, do in zip(reqs, dev, outs):
        D.fps_sampling_device(dp, pc.shape[0], 3, m, d_start, 1, do)
    D.synchronize()
    t_single = time.perf_counter() - t0
    D.device_status()
    single = [o.numpy() for o in outs]
    same = all(np.array_equal(a, b) for a, b in zip(served, single))
    return {
        "metric": "p50/p99 latency, mixed FPS stream",
        "config": {"workload": "96 requests, N log-uniform [2^10, 2^20], uniform cube, "
                               "M = N*U{1/8,1/4,1/2}, start_idx=0, HBM budget 16 GiB, "
                               f"{args.serve_chunk}-pass chunks, batch kernel 1024 x 16",
                   "total_samples": total_m, "qps": round(qps, 3)},
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "stream_requests_per_s": round(len(reqs) / t_stream, 3),
        "stream_samples_per_s": round(total_m / t_stream, 3),
        "capacity": {"requests_per_s": round(cap, 3), "samples_per_s": round(total_m / t_burst, 3),
                     "s": round(t_burst, 3)},
        "one_request_per_launch": {"samples_per_s": round(total_m / t_single, 3),
                                   "s": round(t_single, 3),
                                   "note": "clouds already in HBM, uploads not timed"},
        "speedup_vs_one_request_per_launch": round(t_single / t_burst, 3),
        "target_speedup": 5.0,
        "identical_output": bool(same and same_stream),
        "identical_to_reference": _reference_check("serve96", served),
        "server": st,
    }


def run_pingpong(args) -> dict:
    """The exchange alone (fpsx_probe_exchange) over grid sizes and candidate
    sizes: the per-pass latency floor of the multi-workgroup kernels."""
    rows = []
    for words in (5, 10):
        for g in (2, 8, 16, 32, 64, 128, 256):
            if words == 10 and g > 64:
                continue
            rows.append({"grid": g, "words": words, "us_per_round": round(probe_exchange(g, words), 4)})
    head = next(r for r in rows if r["grid"] == 64 and r["words"] == 10)
    return {"metric": "exchange-only round time, 64 workgroups, 10-word candidates (paired kernel)",
            "value": head["us_per_round"], "unit": "us", "n_gpus": 1, "steps": 1, "warmup": 2,
            "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "u64", "data": "synthetic",
            "config": {"workload": "exchange-only probe, 1024-thread workgroups, 20000 rounds"},
            "rows": rows}


def run_sweep(args) -> dict:
    """The single-cloud planner over cloud sizes: Gaussian-mixture clouds of
    2^14 .. 2^21 points (the gauss1m generator), M = 16384 picks (N/2 below
    2^15), points resident in HBM; per size the plan, the step and kernel
    times and the kernel's microseconds per pick."""
    from fpsample_amd import device as D

    rows = []
    for n in (1 << 14, 1 << 16, 1 << 18, 1 << 20, (1 << 20) + 1, 1 << 21):
        m = min(16384, n // 2)
        pc = gen_gauss_mixture(n)
        d_pts = D.DeviceArray.from_numpy(np.ascontiguousarray(pc.T))
        d_start = D.DeviceArray.from_numpy(np.zeros(1, np.uint64))
        d_out = D.DeviceArray((m,), np.uint64)
        plan = D.plan_info(n, 3, n_samples=m)
        for _ in range(2):
            D.fps_sampling_device(d_pts, n, 3, m, d_start, 1, d_out)
        D.synchronize()
        D.timing(True)
        steps = 5
        t0 = time.perf_counter()
        for _ in range(steps):
            D.fps_sampling_device(d_pts, n, 3, m, d_start, 1, d_out)
        D.synchronize()
        wall = (time.perf_counter() - t0) / steps
        kernel_ms, launches = D.timing_read()
        D.timing(False)
        D.device_status()
        kms = kernel_ms / max(launches, 1)
        rows.append({"n_points": n, "n_samples": m, "plan": plan,
                     "ms_per_step": round(wall * 1e3, 4), "kernel_ms": round(kms, 4),
                     "us_per_pick": round(kms * 1e3 / (m - 1), 4),
                     "samples_per_s": round(m / wall, 1)})
        for d in (d_pts, d_start, d_out):
            d.free()
    head = next(r for r in rows if r["n_points"] == 1 << 20)
    return {"metric": "samples/s over cloud sizes (planner's choice)", "value": head["samples_per_s"],
            "unit": "samples/s", "n_gpus": 1, "steps": 5, "warmup": 2,
            "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "Gaussian mixtures (64 clusters, sigma 0.02) of 2^14..2^21 points, "
                                   "M = min(16384, N/2), start_idx=0"},
            "rows": rows}


def run_copy(args) -> dict:
    """Calibration: device-to-device copies of a known 256 MiB (the runtime's
    plain copy kernel), for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import ctypes

    from fpsample_amd import device as D

    nbytes = 256 << 20
    src = D.DeviceArray((nbytes // 4,), np.float32)
    dst = D.DeviceArray((nbytes // 4,), np.float32)
    h = D.hip()
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    for _ in range(args.warmup):
        D._check(h.hipMemcpy(dst.ptr, src.ptr, nbytes, 3), "hipMemcpy D2D")
    D.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        D._check(h.hipMemcpy(dst.ptr, src.ptr, nbytes, 3), "hipMemcpy D2D")
    D.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {"metric": "device copy GB/s (read + write)", "value": round(2 * nbytes / dt / 1e9, 2),
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(dt * 1e3, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": "hipMemcpy device-to-device, 268435456 bytes per copy"}}


METRIC = "samples/s at N=1M,M=16K (+% roofline); p50/p99 latency, mixed FPS stream"


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="gauss1m",
                    choices=sorted(WORKLOADS) + ["batch256", "serve", "copy256m", "pingpong", "sweep"])
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--extras", default="on", choices=["on", "off"],
                    help="gauss1m: also run the serve and batch256 legs into the same line")
    ap.add_argument("--serve-chunk", type=int, default=256,
                    help="serve: passes per request per launch of the serving loop")
    ap.add_argument("--autotune", default="off", choices=["on", "off"],
                    help="time every plan on this cloud first (off: the planner's choice)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: FPS of one cloud runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path).", file=sys.stderr)
        return 2

    if args.workload == "batch256":
        print(json.dumps(run_batch(args)), flush=True)
        return 0
    if args.workload == "serve":
        print(json.dumps(run_serve(args)), flush=True)
        return 0
    if args.workload == "copy256m":
        print(json.dumps(run_copy(args)), flush=True)
        return 0
    if args.workload == "pingpong":
        print(json.dumps(run_pingpong(args)), flush=True)
        return 0
    if args.workload == "sweep":
        print(json.dumps(run_sweep(args)), flush=True)
        return 0

    from fpsample_amd import device as D

    wl = WORKLOADS[args.workload]
    m = wl["m"]
    pc = wl["gen"]()
    n = pc.shape[0]
    npdu = wl.get("kind") == "npdu"
    kdline = wl.get("kind") == "kdline"
    w = int(n / m * 16) if npdu else -1
    d_start = D.DeviceArray.from_numpy(np.zeros(1, np.uint64))
    d_out = D.DeviceArray((m,), np.uint64)
    if npdu:
        d_pts = D.DeviceArray.from_numpy(np.ascontiguousarray(pc))  # row-major, as the reference
        plan = dict(kind="npdu", **D.npdu_plan_info(n, 3))

        def step():
            D.fps_npdu_sampling_device(d_pts, 3, 1, n, 3, m, w, d_start, d_out)
    elif kdline:
        import fpsample_amd as F

        plan = dict(D.plan_info(n, 3, tie=D.TIE_LOWEST, n_samples=m), tie="lowest",
                    h=wl["h"])
        host_out = [None]

        def step():
            host_out[0] = F.bucket_fps_kdline_sampling(pc, m, wl["h"], start_idx=0)
    else:
        soa = np.ascontiguousarray(pc.T)  # (3, N): coordinate j of point i at j*N + i
        d_pts = D.DeviceArray.from_numpy(soa)
        if args.autotune == "on":
            D.autotune(d_pts, 1, n, n, 3, passes=m)  # untimed: picks the plan for this cloud
        plan = D.plan_info(n, 3, n_samples=m)

        def step():
            D.fps_sampling_device(d_pts, n, 3, m, d_start, 1, d_out)

    for _ in range(args.warmup):
        step()
    D.synchronize()
    D.device_status()

    D.timing(True)
    D.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    D.synchronize()
    t1 = time.perf_counter()
    kernel_ms, launches = D.timing_read()
    D.timing(False)
    D.device_status()

    out = host_out[0] if kdline else d_out.numpy()
    assert (kdline or out[0] == 0) and (npdu or len(np.unique(out)) == m), \
        "picks failed the sanity check"

    wall = t1 - t0
    ms_per_step = wall / args.steps * 1e3
    value = m * args.steps / wall
    kernel_s = kernel_ms / max(launches, 1) / 1e3
    # right after the timed launches: the last launch's exchange count, then the probe
    roof = roofline(args.workload, plan, n, m, kernel_s, npdu, w)
    result = {
        "metric": METRIC if args.workload == "gauss1m" else f"samples/s ({args.workload})",
        "value": round(value, 3),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "n_points": n, "n_samples": m, "dim": 3,
                   "plan": plan},
        "roofline": roof,
        "cpu_baseline": None,
    }
    try:
        from tests.golden import full_length as FL

        if FL.available(args.workload):
            result["identical_to_reference"] = bool(
                np.array_equal(out, FL.load(args.workload)["out"]))
    except ImportError:
        pass
    if "published" in wl:
        mean_s, where = wl["published"]
        result["reference_published"] = {
            "value": round(m / mean_s, 3), "unit": "samples/s", "cores": 1,
            "hardware": "Xeon Gold 6326 @ 2.90 GHz", "source": where,
            "gpu_over_published": round(value / (m / mean_s), 2)}
    if kdline:
        result["note"] = ("host arrays in and out: each step is the whole call, the host "
                          "KD-line build, the copies over PCIe and the GPU launch")
    if args.cpu_baseline == "on":
        result["cpu_baseline"] = cpu_baseline(pc, m, args.cpu_seconds, w, wl.get("h", -1))
        if args.workload == "gauss1m":
            result.update(accuracy_and_bucket(pc, out, m))
        elif not npdu and not kdline and n <= 65536:
            from oracle import oracle as O

            result["accuracy_vs_fp64"] = O.accuracy_vs_f64(pc, out, 0)
    if args.workload == "gauss1m" and args.extras == "on":
        result["serve"] = run_serve(args)
        result["batch256"] = run_batch(args)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

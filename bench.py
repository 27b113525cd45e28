#!/usr/bin/env python3
"""Benchmark of the FPS hot path on one MI355X.

A step is one vanilla FPS run (fps_sampling semantics, src/lib.cpp:188-246)
over one synthetic cloud already resident in HBM: one call of
fpsx_fps_sampling_device, i.e. the exchange-area clear, for large 3-D clouds
the spatial ordering of the points (bounding box, Morton codes, two radix
sorts, gather), and the single persistent FPS launch that makes all M-1
passes.  The kernel shape comes from the library's planner; --autotune on
times every shape first (untimed) instead.  The npdu* workloads time
fps_npdu_sampling (src/lib.cpp:272-340) the same way, through
fpsx_fps_npdu_sampling_device on row-major device data.

Default workload (the metric's configuration): one cloud of N = 1,048,576
points, D = 3 float32, a mixture of 64 isotropic Gaussian clusters (sigma
0.02, centres uniform in [0,1)^3, numpy default_rng(1)), M = 16,384 samples,
start_idx = 0.

Prints one JSON line.  value = samples/s of the whole job (M x steps / wall
time between two device synchronisations).  roofline: algorithmic bytes of
the vanilla loop, 20 B per point-pass (x, y, z and min-distance read, the
min-distance written) x N x (M - 1) per launch, divided by the FPS kernel's
average duration measured with HIP events on the stream it runs on, against
the 8 TB/s HBM peak.  cpu_baseline: the reference's fps_sampling compiled
from its own sources (oracle/_ref) -- or the oracle's C restatement where
that build is absent -- timed single-threaded on this host over a bounded
number of samples of the same cloud (its per-sample cost does not depend on
how many samples are drawn).  The same CPU leg reports accuracy_vs_fp64
(index agreement and coverage-radius relative error of the GPU picks against
the greedy loop in double precision) and, for the 1 M-point cloud,
bucket_kdline: the bucket entry (h = 9) on the GPU beside the reference's
bucket sampler on one host core, with their picks compared.
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
    "uniform4k": dict(gen=gen_uniform_4096, m=1024,
                      desc="one cloud N=4096 D=3 f32 uniform [0,1)^3 seed 42, M=1024, start_idx=0"),
    "lattice16m": dict(gen=gen_lattice_dups, m=4096,
                       desc="one cloud N=16777216 D=3 f32 lattice {0..255}^3, 25% duplicates, "
                            "M=4096, start_idx=0"),
    # windowed-update FPS on the reference's own bench cases (bench/test_bench.py:43-46,
    # 151-154): w is the Python wrapper's default int(N / M * 16)
    "npdu4k": dict(gen=gen_uniform_4096, m=1024, kind="npdu",
                   desc="fps_npdu_sampling, one cloud N=4096 D=3 f32 uniform [0,1)^3 seed 42, "
                        "M=1024, w=64, start_idx=0"),
    "npdu100k": dict(gen=lambda: gen_uniform_n(100_000), m=50_000, kind="npdu",
                     desc="fps_npdu_sampling, one cloud N=100000 D=3 f32 uniform [0,1)^3 seed 42, "
                          "M=50000, w=32, start_idx=0"),
}


# ---------------------------------------------------------------- CPU baseline
def cpu_baseline(pc: np.ndarray, m_full: int, target_s: float, w: int = -1) -> dict:
    """w < 0: fps_sampling; else fps_npdu_sampling with window w."""
    from oracle import oracle as O

    ref = O.load_reference()
    if ref is not None:
        kind = "reference"
        if w < 0:
            run = lambda m: ref.fps_sampling(pc, m, 0)  # noqa: E731
        else:
            run = lambda m: ref.fps_npdu(pc, m, w, 0)  # noqa: E731
    else:
        kind = "port"
        if w < 0:
            run = lambda m: O.fps(pc, m, 0)  # noqa: E731
        else:
            run = lambda m: O.npdu(pc, m, w, 0)  # noqa: E731
    n = pc.shape[0]
    # calibrate on a few passes, then size the sample to about target_s seconds
    m_cal = max(2, min(m_full, int(2e8 // max(n, 1)) + 2))
    t0 = time.perf_counter()
    run(m_cal)
    dt = time.perf_counter() - t0
    per_pass = dt / max(m_cal - 1, 1)
    m_s = int(min(m_full, max(m_cal, target_s / max(per_pass, 1e-12))))
    t0 = time.perf_counter()
    run(m_s)
    dt = time.perf_counter() - t0
    return {
        "value": round(m_s / dt, 3),
        "unit": "samples/s",
        "cores": 1,
        "kind": kind,
        "sample": f"first {m_s} of the {m_full} samples of the same cloud, start_idx=0, "
                  f"{dt:.2f} s single-threaded on {platform.processor() or platform.machine()} "
                  f"(os.cpu_count()={os.cpu_count()})",
    }


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
    return out


def _read_traffic(workload: str, plan: dict):
    """HBM bytes per launch of the FPS kernel from the committed rocprofv3
    counter summary (FETCH_SIZE and WRITE_SIZE, scaled by a copy of known
    size), when that summary was taken with the plan this run uses."""
    path = os.path.join(ROOT, "profiles", f"fps_{plan.get('kind')}_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.strip().split("=", 1)
            kv[k] = v
    want = f"{plan.get('kind')}:{plan.get('block')}:{plan.get('ppt')}:{plan.get('grid')}"
    if kv.get("plan") != want:
        return None
    try:
        return float(kv["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


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
    same = all(np.array_equal(a, o.numpy()) for a, o in zip(batched, outs))
    alg = sum(BYTES_PER_POINT_PASS * pc.shape[0] * (m - 1) for pc, m in zip(clouds, ms))
    achieved = alg / wall_b / 1e9
    return {
        "metric": "samples/s (256 ragged clouds, one persistent launch)",
        "value": round(total_m / wall_b, 3),
        "unit": "samples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall_b * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "B=256 clouds, N_i log-uniform [1024, 65536], M_i = N_i/4, D=3 f32, "
                               "half uniform cube / half unit sphere + N(0,0.01^2), start_idx=0",
                   "total_points": int(sum(pc.shape[0] for pc in clouds)), "total_samples": total_m,
                   "batch_kernel": info},
        "one_cloud_per_launch": {"value": round(total_m / wall_p, 3), "unit": "samples/s",
                                 "ms_per_step": round(wall_p * 1e3, 3)},
        "speedup_vs_one_cloud_per_launch": round(wall_p / wall_b, 3),
        "identical_output": bool(same),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
    }


def run_serve(args) -> dict:
    """Serving stream: Poisson arrivals at ~60% of the measured capacity."""
    from fpsample_amd.serve import Server

    reqs = gen_serving_stream()
    with Server(hbm_budget_bytes=16 << 30, chunk_passes=256) as srv:
        # capacity: the first third of the requests submitted as one burst
        cal = reqs[: len(reqs) // 3]
        t0 = time.perf_counter()
        ids = [srv.submit(pc, m, 0) for pc, m in cal]
        for r in ids:
            srv.wait(r)
        cap = len(cal) / (time.perf_counter() - t0)
        qps = 0.6 * cap
        rng = np.random.default_rng(5)
        stream = reqs[len(reqs) // 3:]
        gaps = rng.exponential(1.0 / qps, len(stream))
        ids = []
        t0 = time.perf_counter()
        t_next = t0
        for (pc, m), g in zip(stream, gaps):
            t_next += g
            now = time.perf_counter()
            if t_next > now:
                time.sleep(t_next - now)
            ids.append(srv.submit(pc, m, 0))
        lats = [srv.wait(r)[1] for r in ids]
        t1 = time.perf_counter()
        st = srv.stats()
    lat = np.array(lats) * 1e3
    samples = sum(m for _, m in stream)
    return {
        "metric": "p50/p99 latency, mixed FPS stream",
        "value": round(samples / (t1 - t0), 3),
        "unit": "samples/s",
        "n_gpus": 1, "steps": len(stream), "warmup": len(cal),
        "ms_per_step": round((t1 - t0) / len(stream) * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "Poisson arrivals, N log-uniform [2^10, 2^20], uniform cube, "
                               "M = N*U{1/8,1/4,1/2}, start_idx=0, HBM budget 16 GiB, 256-pass chunks",
                   "qps": round(qps, 3), "measured_capacity_rps": round(cap, 3)},
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "requests_per_s": round(len(stream) / (t1 - t0), 3),
        "server": st,
    }


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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="gauss1m", choices=sorted(WORKLOADS) + ["batch256", "serve", "copy256m"])
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--autotune", default="off", choices=["on", "off"],
                    help="time every kernel shape on this cloud first (off: the planner's "
                         "per-pass time model, which picks the same shape for gauss1m)")
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

    from fpsample_amd import device as D

    wl = WORKLOADS[args.workload]
    m = wl["m"]
    pc = wl["gen"]()
    n = pc.shape[0]
    npdu = wl.get("kind") == "npdu"
    w = int(n / m * 16) if npdu else -1
    d_start = D.DeviceArray.from_numpy(np.zeros(1, np.uint64))
    d_out = D.DeviceArray((m,), np.uint64)
    if npdu:
        d_pts = D.DeviceArray.from_numpy(np.ascontiguousarray(pc))  # row-major, as the reference
        plan = dict(kind="npdu", **D.npdu_plan_info(n, 3))

        def step():
            D.fps_npdu_sampling_device(d_pts, 3, 1, n, 3, m, w, d_start, d_out)
    else:
        soa = np.ascontiguousarray(pc.T)  # (3, N): coordinate j of point i at j*N + i
        d_pts = D.DeviceArray.from_numpy(soa)
        if args.autotune == "on":
            D.autotune(d_pts, 1, n, n, 3)  # untimed: picks the kernel shape for this cloud
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

    out = d_out.numpy()
    assert out[0] == 0 and (npdu or len(np.unique(out)) == m), "picks failed the sanity check"

    wall = t1 - t0
    ms_per_step = wall / args.steps * 1e3
    value = m * args.steps / wall
    kernel_s = kernel_ms / max(launches, 1) / 1e3
    if npdu:
        # per pass: the window's w+1 points updated (20 B each) and the argmax
        # over every min-distance (4 B per point), src/lib.cpp:302-315
        alg_bytes = (20 * (min(w, n - 1) + 1) + 4 * n) * (m - 1)
    else:
        alg_bytes = BYTES_PER_POINT_PASS * n * (m - 1)
    achieved = alg_bytes / kernel_s / 1e9
    result = {
        "metric": "samples/s at N=1M,M=16K (+% roofline)" if args.workload == "gauss1m"
        else f"samples/s ({args.workload})",
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
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _read_traffic(args.workload, plan),
            "kernel_ms": round(kernel_s * 1e3, 4),
            "algorithmic_bytes_per_launch": alg_bytes,
        },
        "cpu_baseline": None,
    }
    if args.cpu_baseline == "on":
        result["cpu_baseline"] = cpu_baseline(pc, m, args.cpu_seconds, w)
        if args.workload == "gauss1m":
            result.update(accuracy_and_bucket(pc, out, m))
        elif not npdu and n <= 65536:
            from oracle import oracle as O

            result["accuracy_vs_fp64"] = O.accuracy_vs_f64(pc, out, 0)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

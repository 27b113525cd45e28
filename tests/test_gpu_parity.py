"""GPU parity: the HIP kernels against the oracle and the reference's golden vectors.

Every comparison is bit-exact on indices (the picks are integers).  Sizes the
oracle finishes in seconds are compared whole; the full benchmark sizes are
checked through an oracle-compared prefix plus size-independent properties
(forced-prefix resume reproduces the run, picks distinct while distances are
positive).  FPSX_PLAN forces each kernel shape so every instantiation family
the planner can pick is exercised.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.capi import COLS, ROWS

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def plan_env(monkeypatch):
    def set_plan(p):
        if p is None:
            monkeypatch.delenv("FPSX_PLAN", raising=False)
        else:
            monkeypatch.setenv("FPSX_PLAN", p)
    yield set_plan
    monkeypatch.delenv("FPSX_PLAN", raising=False)


def run_fps(gpu, pc, m, starts, layout=COLS):
    rc, out = gpu.fps(pc, m, starts, layout=layout)
    assert rc == 0, f"rc={rc}: {gpu.last_error()}"
    return out


def test_golden_through_python_api(gpu, golden):
    import fpsample_amd as F

    for name, g in golden.items():
        log(f"golden {name}")
        st = g["start"]
        if g["kind"] == "fps":
            arg = int(st[0]) if st.size == 1 else [int(x) for x in st]
            out = F.fps_sampling(g["pc"], g["m"], start_idx=arg)
        elif g["kind"] == "npdu":
            out = F.fps_npdu_sampling(g["pc"], g["m"], w=g["h"], start_idx=int(st[0]))
        else:
            out = F.bucket_fps_kdline_sampling(g["pc"], g["m"], g["h"], start_idx=int(st[0]))
        assert out.dtype == np.uint64
        np.testing.assert_array_equal(out, g["out"], err_msg=name)


def test_golden_through_capi_rows(gpu, golden):
    for name, g in golden.items():
        if g["kind"] != "fps":
            continue
        log(f"golden rows {name}")
        np.testing.assert_array_equal(run_fps(gpu, g["pc"], g["m"], g["start"], layout=ROWS),
                                      g["out"], err_msg=name)


def _clouds():
    rng = np.random.default_rng(2024)
    yield "uniform", rng.random((5000, 3)).astype(np.float32)
    yield "lattice_dups", np.repeat(rng.integers(0, 5, (1500, 3)), 4, axis=0).astype(np.float32)
    yield "gauss", (rng.standard_normal((3333, 3)) * 100).astype(np.float32)


RESIDENT_SHAPES = ["resident:256:1", "resident:256:4", "resident:256:16", "resident:256:32",
                   "resident:512:8", "resident:1024:1", "resident:1024:4", "resident:1024:16"]


@pytest.mark.parametrize("shape", RESIDENT_SHAPES)
def test_resident_shapes(gpu, plan_env, shape):
    # Small per-workgroup shares force multi-workgroup grids (cross-WG exchange).
    plan_env(shape)
    for name, pc in _clouds():
        for starts in (0, [len(pc) - 1, 5, 9]):
            log(f"{shape} {name} starts={starts}")
            want = O.fps(pc, 700, starts)
            np.testing.assert_array_equal(run_fps(gpu, pc, 700, starts), want,
                                          err_msg=f"{shape} {name}")


@pytest.mark.parametrize("shape", ["stream:1024", "stream:256", "stream:1024:3", "stream:1024:40",
                                   "generic:256", "generic:1024:7"])
def test_stream_shapes(gpu, plan_env, shape):
    plan_env(shape)
    for name, pc in _clouds():
        log(f"{shape} {name}")
        want = O.fps(pc, 500, [7, 1])
        np.testing.assert_array_equal(run_fps(gpu, pc, 500, [7, 1]), want,
                                      err_msg=f"{shape} {name}")


@pytest.mark.parametrize("d", [1, 2, 3, 4, 5, 6, 7, 8, 9, 13, 16, 40])
def test_dimensions_auto_plan(gpu, d):
    rng = np.random.default_rng(d)
    for n in (37, 1000, 20000):
        pc = rng.standard_normal((n, d)).astype(np.float32)
        m = min(n, 300)
        log(f"dim {d} n {n} plan {gpu.plan_info(n, d)}")
        np.testing.assert_array_equal(run_fps(gpu, pc, m, 3 % n), O.fps(pc, m, 3 % n),
                                      err_msg=f"d={d} n={n}")


@pytest.mark.parametrize("d", [1, 2, 3, 5, 8])
def test_dimensions_forced_stream(gpu, plan_env, d):
    plan_env("stream:1024:5")
    rng = np.random.default_rng(50 + d)
    pc = rng.integers(0, 4, (3000, d)).astype(np.float32)
    np.testing.assert_array_equal(run_fps(gpu, pc, 600, 0), O.fps(pc, 600, 0))


def test_edge_sizes(gpu):
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 63, 64, 65, 255, 256, 257, 1023, 1024, 1025, 4097, 16384, 16385, 40000):
        pc = rng.random((n, 3)).astype(np.float32)
        for m in sorted({k for k in (1, 2, min(n, 5), min(n, 300)) if k <= n}):
            log(f"edge n={n} m={m}")
            np.testing.assert_array_equal(run_fps(gpu, pc, m, n - 1), O.fps(pc, m, n - 1),
                                          err_msg=f"n={n} m={m}")


def test_all_points_then_duplicates(gpu):
    # More samples than distinct points: the reference repeats the highest index.
    pc = np.repeat(np.array([[0, 0, 0], [1, 1, 1], [2, 0, 1]], np.float32), 5, axis=0)
    np.testing.assert_array_equal(run_fps(gpu, pc, 15, 0), O.fps(pc, 15, 0))


def test_special_values(gpu):
    rng = np.random.default_rng(12)
    pc = (rng.random((2000, 3)) * 4e19).astype(np.float32)   # squares overflow to +inf
    pc[[5, 700]] = np.nan
    pc[9] = np.inf
    np.testing.assert_array_equal(run_fps(gpu, pc, 300, 0), O.fps(pc, 300, 0))
    tiny = (rng.random((2000, 3)) * 1e-21).astype(np.float32)  # denormal squares
    np.testing.assert_array_equal(run_fps(gpu, tiny, 300, 0), O.fps(tiny, 300, 0))


def test_kdline_capi_against_oracle(gpu):
    rng = np.random.default_rng(13)
    for pc, h in ((rng.random((20000, 3)).astype(np.float32), 9),
                  (rng.integers(0, 6, (5000, 3)).astype(np.float32), 5),
                  (rng.random((3000, 7)).astype(np.float32), 3)):
        for s in (0, len(pc) // 2):
            rc, out = gpu.kdline(pc, 400, s, h)
            assert rc == 0, gpu.last_error()
            np.testing.assert_array_equal(out, O.kdline_quickfps(pc, 400, h, s))


def test_kdline_capi_error_codes(gpu):
    pc = np.zeros((10, 9), np.float32)
    assert gpu.kdline(pc, 3, 0, 2)[0] == 1
    assert gpu.kdline(pc[:, :3], 3, 10, 2)[0] == 2


def c2_cloud(n=1 << 20, seed=1):
    rng = np.random.default_rng(seed)
    centers = rng.random((64, 3))
    lab = rng.integers(0, 64, n)
    return (centers[lab] + 0.02 * rng.standard_normal((n, 3))).astype(np.float32)


def c5_cloud(n=1 << 24, seed=4):
    rng = np.random.default_rng(seed)
    distinct = n * 3 // 4
    sites = rng.choice(1 << 24, distinct, replace=False)
    copies = sites[rng.integers(0, distinct, n - distinct)]
    codes = np.concatenate([sites, copies])
    rng.shuffle(codes)
    return np.stack([(codes >> 16) & 255, (codes >> 8) & 255, codes & 255], axis=1).astype(np.float32)


def test_large_single_cloud_properties(gpu):
    pc = c2_cloud()
    m = 16384
    log(f"C2 plan {gpu.plan_info(len(pc), 3)}")
    out = run_fps(gpu, pc, m, 0)
    np.testing.assert_array_equal(out[:128], O.fps(pc, 128, 0))
    assert len(np.unique(out)) == m
    k = 9000
    resumed = run_fps(gpu, pc, m, out[:k])
    np.testing.assert_array_equal(resumed, out)


def test_lattice_ties_streaming_size(gpu):
    pc = c5_cloud()
    m = 4096
    log(f"C5 plan {gpu.plan_info(len(pc), 3)}")
    out = run_fps(gpu, pc, m, 0)
    np.testing.assert_array_equal(out[:24], O.fps(pc, 24, 0))
    assert len(np.unique(out)) == m
    resumed = run_fps(gpu, pc, m, out[:3000])
    np.testing.assert_array_equal(resumed, out)


def test_device_strided_rows_and_fused_gather(gpu, plan_env):
    from fpsample_amd import device as D

    rng = np.random.default_rng(21)
    for d, plan in ((3, None), (3, "stream:1024:4"), (5, None), (11, None)):
        plan_env(plan)
        pc = rng.standard_normal((6000, d)).astype(np.float32)
        rows = D.DeviceArray.from_numpy(pc)               # row-major: strides (d, 1)
        want = O.fps(pc, 400, 17)
        idx, pts = D.fps_sampling_tensor(rows, 400, 17, return_points=True)
        got = idx.numpy()
        np.testing.assert_array_equal(got, want, err_msg=f"d={d} plan={plan}")
        np.testing.assert_array_equal(pts.numpy(), pc[want.astype(np.int64)])
        soa = D.DeviceArray.from_numpy(np.ascontiguousarray(pc.T))
        ds = D.DeviceArray.from_numpy(np.array([17, 3], np.uint64))
        out = D.DeviceArray((400,), np.uint64)
        opts = D.DeviceArray((400, d), np.float32)
        D.fps_sampling_device_strided(soa, 1, 6000, 6000, d, 400, ds, 2, out, opts)
        D.synchronize()
        D.device_status()
        want2 = O.fps(pc, 400, [17, 3])
        np.testing.assert_array_equal(out.numpy(), want2)
        np.testing.assert_array_equal(opts.numpy(), pc[want2.astype(np.int64)])


PRUNED_SHAPES = ["pruned:1024:16", "pruned:1024:8", "pruned:1024:4", "pruned:256:16", "pruned:256:32"]


@pytest.mark.parametrize("shape", PRUNED_SHAPES)
def test_pruned_shapes(gpu, plan_env, shape):
    # Spatial order + per-wave box pruning must not change a single pick.
    plan_env(shape)
    rng = np.random.default_rng(31)
    clouds = [("gauss", c2_cloud(40000, seed=3)),
              ("lattice_dups", np.repeat(rng.integers(0, 6, (6000, 3)), 3, axis=0).astype(np.float32)),
              ("uniform", rng.random((30000, 3)).astype(np.float32))]
    for name, pc in clouds:
        for starts in (0, [len(pc) - 1, 7, 11]):
            log(f"{shape} {name} starts={starts}")
            want = O.fps(pc, 2000, starts)
            np.testing.assert_array_equal(run_fps(gpu, pc, 2000, starts), want,
                                          err_msg=f"{shape} {name}")


def test_pruned_special_values(gpu, plan_env):
    plan_env("pruned:1024:4")
    rng = np.random.default_rng(32)
    pc = (rng.random((9000, 3)) * 4e19).astype(np.float32)
    pc[[5, 700, 4000]] = np.nan
    pc[9] = np.inf
    pc[10, 1] = -np.inf
    np.testing.assert_array_equal(run_fps(gpu, pc, 600, 0), O.fps(pc, 600, 0))
    tiny = (rng.random((9000, 3)) * 1e-21).astype(np.float32)
    np.testing.assert_array_equal(run_fps(gpu, tiny, 600, 0), O.fps(tiny, 600, 0))


def test_pruned_auto_plan_and_bucket_order(gpu):
    rng = np.random.default_rng(33)
    pc = c2_cloud(120000, seed=5)
    info = gpu.plan_info(len(pc), 3, m=2048)
    log(f"auto plan {info}")
    assert info["kind"] == 5 and info["ppt"] == 4  # paired pruned kernel, 30 workgroups
    np.testing.assert_array_equal(run_fps(gpu, pc, 2048, 77), O.fps(pc, 2048, 77))
    # bucket entry: lowest-position ties over the KD-line order, pruned plan
    lat = rng.integers(0, 40, (100000, 3)).astype(np.float32)
    rc, out = gpu.kdline(lat, 1500, 123, 7)
    assert rc == 0, gpu.last_error()
    np.testing.assert_array_equal(out, O.kdline_quickfps(lat, 1500, 7, 123))


def test_small_clouds_auto_plans(gpu):
    # The planner's choices below 64 K points: one workgroup paired at the
    # wave level (kind 5, grid 1, the smallest share holding the cloud) from
    # 4 K points and 1024 samples or 8 K points and 512 samples; the paired
    # kernel at 4 points per thread above 16 K points; the resident kernel
    # otherwise.  Every choice against the oracle, both tie rules (the bucket
    # entry ties lowest).
    rng = np.random.default_rng(35)
    cases = [(4096, 1024, 5, 4), (5000, 600, 5, 8), (8192, 2048, 5, 8), (12000, 3000, 5, 16),
             (16384, 4096, 5, 16), (16385, 1024, 5, 4), (40000, 2500, 5, 4),
             (4096, 256, 0, None), (3000, 1500, 0, None), (30000, 500, 0, None)]
    for n, m, kind, ppt in cases:
        info = gpu.plan_info(n, 3, m=m)
        assert info["kind"] == kind and (ppt is None or info["ppt"] == ppt), (n, m, info)
        for name, pc in (("uniform", rng.random((n, 3)).astype(np.float32)),
                         ("gauss", c2_cloud(n, seed=n)),
                         ("dups", rng.integers(0, 5, (n, 3)).astype(np.float32))):
            s0 = int(rng.integers(n))
            log(f"small auto n={n} m={m} {name} start={s0} plan={info}")
            np.testing.assert_array_equal(run_fps(gpu, pc, m, s0), O.fps(pc, m, s0),
                                          err_msg=f"n={n} m={m} {name}")
    pc = (rng.random((12000, 3)) * 4e19).astype(np.float32)
    pc[[5, 700, 4000]] = np.nan
    pc[9] = np.inf
    np.testing.assert_array_equal(run_fps(gpu, pc, 3000, 0), O.fps(pc, 3000, 0))
    lat = rng.integers(0, 12, (12000, 3)).astype(np.float32)
    rc, out = gpu.kdline(lat, 3000, 5, 7)
    assert rc == 0, gpu.last_error()
    np.testing.assert_array_equal(out, O.kdline_quickfps(lat, 3000, 7, 5))


def test_bucket_entry_full_size_matches_oracle(gpu):
    # 1 M points, h = 9: threaded KD-line permutation on the host, pruned GPU
    # plan with lowest-position ties; the oracle's QuickFPS restatement decides.
    import fpsample_amd as F

    pc = c2_cloud()
    m = 2048
    t = O.kdline_quickfps(pc, m, 9, 5)
    got = F.bucket_fps_kdline_sampling(pc, m, 9, start_idx=5)
    np.testing.assert_array_equal(got, t)


def test_accuracy_against_fp64_yardstick(gpu):
    # Reported error of the float32 result: index agreement and coverage-radius
    # relative error against the same greedy loop in double precision.
    # Tolerance: coverage radius within 1e-5 relative (a late divergence of
    # near-equal candidates may swap picks but barely moves the radius).
    import bench

    for name, pc, m in (("uniform4k", bench.gen_uniform_4096(), 1024),
                        ("gauss1m_prefix", c2_cloud(), 256)):
        out = run_fps(gpu, pc, m, 0)
        acc = O.accuracy_vs_f64(pc, out, 0)
        log(f"{name}: {acc}")
        assert acc["coverage_radius_rel_err"] <= 1e-5, acc
        assert acc["index_agreement"] >= 0.9, acc


SPRUNED_PLANS = ["spruned", "spruned:1", "spruned:3", "spruned:16", "spair", "spair:2", "spair:3", "spair:16"]


@pytest.mark.parametrize("plan", SPRUNED_PLANS)
def test_stream_pruned_plans(gpu, plan_env, plan):
    # Streaming kernel with per-block pruning: points in HBM in spatial order,
    # only blocks a pick can change are read; "spair" settles up to two picks
    # per exchange.  Both tie rules, forced prefix.
    from fpsample_amd import device as D

    plan_env(plan)
    rng = np.random.default_rng(41)
    clouds = [("gauss", c2_cloud(60000, seed=7)),
              ("lattice_dups", np.repeat(rng.integers(0, 12, (9000, 3)), 3, axis=0).astype(np.float32)),
              ("ragged_tail", rng.random((17 * 1024 + 5, 3)).astype(np.float32))]
    for name, pc in clouds:
        n = len(pc)
        log(f"spruned {plan} {name} n={n} plan={gpu.plan_info(n, 3, m=700)}")
        np.testing.assert_array_equal(run_fps(gpu, pc, 700, 3), O.fps(pc, 700, 3), err_msg=name)
        np.testing.assert_array_equal(run_fps(gpu, pc, 700, [5, 9, n - 1]), O.fps(pc, 700, [5, 9, n - 1]))
        soa = D.DeviceArray.from_numpy(np.ascontiguousarray(pc.T))
        ds = D.DeviceArray.from_numpy(np.array([11], np.uint64))
        out = D.DeviceArray((500,), np.uint64)
        D.fps_sampling_device(soa, n, 3, 500, ds, 1, out, tie=D.TIE_LOWEST,
                              init=float(np.finfo(np.float32).max))
        D.synchronize()
        D.device_status()
        want = O.fps_ex(pc, 500, 11, lowest=True, init=float(np.finfo(np.float32).max))
        np.testing.assert_array_equal(out.numpy(), want, err_msg=f"{name} lowest")


@pytest.mark.parametrize("plan", ["spruned:4", "spair:4"])
def test_stream_pruned_special_values(gpu, plan_env, plan):
    plan_env(plan)
    rng = np.random.default_rng(42)
    pc = (rng.random((40000, 3)) * 4e19).astype(np.float32)
    pc[[5, 7000, 39999]] = np.nan
    pc[9] = np.inf
    pc[10, 1] = -np.inf
    np.testing.assert_array_equal(run_fps(gpu, pc, 600, 0), O.fps(pc, 600, 0))
    tiny = (rng.random((30000, 3)) * 1e-21).astype(np.float32)
    np.testing.assert_array_equal(run_fps(gpu, tiny, 600, 0), O.fps(tiny, 600, 0))


def test_lattice_streaming_kernels_agree(gpu, plan_env):
    # 16.7 M lattice points with 25 % duplicates, 4096 picks: the paired
    # pruned streaming kernel (auto plan), the one-pick pruned streaming
    # kernel and the plain streaming kernel must give the same sequence; the
    # oracle checks the prefix.
    pc = c5_cloud()
    plan_env(None)
    info = gpu.plan_info(len(pc), 3, m=4096)
    log(f"C5 auto plan {info}")
    assert info["kind"] == 6
    a = run_fps(gpu, pc, 4096, 0)
    plan_env("spruned")
    b = run_fps(gpu, pc, 4096, 0)
    np.testing.assert_array_equal(a, b)
    plan_env("stream:1024")
    b = run_fps(gpu, pc, 4096, 0)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a[:24], O.fps(pc, 24, 0))


@pytest.mark.parametrize("plan", ["spair:3", "spair:8"])
def test_stream_pair_large_share(gpu, plan_env, plan):
    # Over 960 blocks per workgroup at 3 workgroups (the box tests reach the
    # last wave, which also speculates; several steps of every block scan)
    # and 512 at 8; clustered and lattice clouds of 2^21 points.
    plan_env(plan)
    rng = np.random.default_rng(45)
    n = 1 << 21
    for name, pc in (("gauss", c2_cloud(n, seed=21)),
                     ("lattice", rng.integers(0, 128, (n, 3)).astype(np.float32))):
        log(f"spair large share {plan} {name} plan={gpu.plan_info(n, 3, m=1200)}")
        np.testing.assert_array_equal(run_fps(gpu, pc, 1200, 5), O.fps(pc, 1200, 5), err_msg=name)


@pytest.mark.parametrize("plan", ["spair:2", "spair:7"])
def test_stream_pair_exhausts_cloud(gpu, plan_env, plan):
    # Every point picked from a cloud of duplicates: the later rounds pick
    # among zero min-distances, where pairs must fall back exactly as the
    # reference's argmax does.  Both tie rules, a forced prefix.
    from fpsample_amd import device as D

    plan_env(plan)
    rng = np.random.default_rng(43)
    pc = np.repeat(rng.integers(0, 6, (1500, 3)), 3, axis=0).astype(np.float32)
    n = len(pc)
    log(f"spair exhaust {plan} n={n}")
    np.testing.assert_array_equal(run_fps(gpu, pc, n, 0), O.fps(pc, n, 0))
    np.testing.assert_array_equal(run_fps(gpu, pc, n, [4, 4, 17]), O.fps(pc, n, [4, 4, 17]))
    soa = D.DeviceArray.from_numpy(np.ascontiguousarray(pc.T))
    ds = D.DeviceArray.from_numpy(np.array([2], np.uint64))
    out = D.DeviceArray((n,), np.uint64)
    D.fps_sampling_device(soa, n, 3, n, ds, 1, out, tie=D.TIE_LOWEST)
    D.synchronize()
    D.device_status()
    np.testing.assert_array_equal(out.numpy(), O.fps_ex(pc, n, 2, lowest=True), err_msg="lowest")


@pytest.mark.parametrize("plan", ["pruned:1024:4", "pair:1024:4", "spruned:3", "spair:3"])
def test_fused_gather_pruned_kinds(gpu, plan_env, plan):
    # The pruned, paired (both picks of a round) and streaming pruned kernels
    # write the picks' coordinates from their own pick records; two starts,
    # an odd sample count, SoA device data.
    from fpsample_amd import device as D

    plan_env(plan)
    n, m = 20000, 2001
    pc = c2_cloud(n, seed=9)
    want = O.fps(pc, m, [17, 3])
    soa = D.DeviceArray.from_numpy(np.ascontiguousarray(pc.T))
    ds = D.DeviceArray.from_numpy(np.array([17, 3], np.uint64))
    out = D.DeviceArray((m,), np.uint64)
    opts = D.DeviceArray((m, 3), np.float32)
    log(f"fused gather {plan}")
    D.fps_sampling_device_strided(soa, 1, n, n, 3, m, ds, 2, out, opts)
    D.synchronize()
    D.device_status()
    np.testing.assert_array_equal(out.numpy(), want, err_msg=plan)
    np.testing.assert_array_equal(opts.numpy(), pc[want.astype(np.int64)], err_msg=plan)


SOLO_PLANS = ["solo:1024:4", "solo:1024:8", "solo:1024:16", "NO_PAIR"]


@pytest.mark.parametrize("plan", SOLO_PLANS)
def test_one_workgroup_paired_plans(gpu, plan_env, monkeypatch, plan):
    # One workgroup paired at the wave level (and the one-pick workgroup
    # under FPSX_NO_PAIR=1): uniform, clustered, heavy duplicates (every
    # point picked, so the last rounds pick among zero min-distances), special
    # values; forced prefix, both tie rules, a large initial distance.
    from fpsample_amd import device as D

    if plan == "NO_PAIR":
        plan_env(None)
        monkeypatch.setenv("FPSX_NO_PAIR", "1")
    else:
        plan_env(plan)
    rng = np.random.default_rng(44)
    n = 4096
    dups = np.repeat(rng.integers(0, 5, (n // 4, 3)), 4, axis=0).astype(np.float32)
    special = (rng.random((n, 3)) * 4e19).astype(np.float32)
    special[[5, 700, 4000]] = np.nan
    special[9] = np.inf
    clouds = [("uniform", rng.random((n, 3)).astype(np.float32), 1024),
              ("gauss", c2_cloud(n, seed=12), 2048),
              ("dups_all", dups, n),
              ("special", special, 1500)]
    for name, pc, m in clouds:
        info = gpu.plan_info(len(pc), 3, m=m)
        log(f"solo {plan} {name} m={m} plan={info}")
        assert info["grid"] == 1, info
        np.testing.assert_array_equal(run_fps(gpu, pc, m, 7), O.fps(pc, m, 7), err_msg=name)
        np.testing.assert_array_equal(run_fps(gpu, pc, m, [3, 3, 100]), O.fps(pc, m, [3, 3, 100]),
                                      err_msg=name + " starts")
        soa = D.DeviceArray.from_numpy(np.ascontiguousarray(pc.T))
        ds = D.DeviceArray.from_numpy(np.array([11], np.uint64))
        out = D.DeviceArray((m,), np.uint64)
        D.fps_sampling_device(soa, len(pc), 3, m, ds, 1, out, tie=D.TIE_LOWEST,
                              init=float(np.finfo(np.float32).max))
        D.synchronize()
        D.device_status()
        want = O.fps_ex(pc, m, 11, lowest=True, init=float(np.finfo(np.float32).max))
        np.testing.assert_array_equal(out.numpy(), want, err_msg=name + " lowest")
        # Both picks of a round write their coordinates (fused gather).
        opts = D.DeviceArray((m, 3), np.float32)
        ds2 = D.DeviceArray.from_numpy(np.array([17, 3], np.uint64))
        D.fps_sampling_device_strided(soa, 1, len(pc), len(pc), 3, m, ds2, 2, out, opts)
        D.synchronize()
        D.device_status()
        want = O.fps(pc, m, [17, 3])
        np.testing.assert_array_equal(out.numpy(), want, err_msg=name + " gather")
        np.testing.assert_array_equal(opts.numpy(), pc[want.astype(np.int64)], err_msg=name + " gather")
    if plan != "NO_PAIR":
        # Clouds far below a workgroup's share (empty waves), one and two picks.
        for n_small, m_small in ((37, 37), (1000, 700), (5, 1), (64, 2)):
            pc = rng.random((n_small, 3)).astype(np.float32)
            log(f"solo {plan} n={n_small} m={m_small}")
            np.testing.assert_array_equal(run_fps(gpu, pc, m_small, 0), O.fps(pc, m_small, 0),
                                          err_msg=f"n={n_small} m={m_small}")

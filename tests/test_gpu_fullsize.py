"""GPU parity at the benchmark workloads' full sizes: every pick against the
reference's own output.

The fixtures (tests/golden/full/, written by make_full_length.py from the
reference's fps_sampling, src/lib.cpp:188-246, compiled from its sources)
hold every pick of
  - the 1 M-point Gaussian mixture, 16,384 picks, and the same generator at
    2^20 + 1 and 2^21 points;
  - the 16.7 M-point lattice with 25 % duplicates, 4,096 picks (heavy ties);
  - the reference's own 50k and 100k bench cases (uniform, np.random.seed(42);
    4,096 of 50,000 and 50,000 of 100,000 picks), through fps_sampling and
    through bucket_fps_kdline_sampling with h = 7, and the 4k case with h = 5;
and a SHA-256 per cloud of every pick of
  - the 256 ragged clouds of the batch workload (M_i = N_i / 4);
  - the 96 requests of the serving workload (N up to 2^20, M = N/8..N/2).
Each cloud is regenerated from its seed and its SHA-256 checked first.  Each
plan the library can take for these clouds is forced in turn.
"""
import numpy as np
import pytest

import bench
from tests.capi import COLS
from tests.golden import full_length as FL

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def first_diff(got, want):
    bad = np.nonzero(np.asarray(got) != np.asarray(want))[0]
    return int(bad[0]) if bad.size else -1


def check_single(gpu, monkeypatch, name, pc, env):
    d = FL.load(name)
    assert FL.sha256(pc) == d["cloud_sha256"], f"{name}: regenerated cloud differs from the fixture's"
    for k in ("FPSX_PLAN", "FPSX_NO_PAIR", "FPSX_NO_PRUNE"):
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    m = int(d["m"])
    log(f"{name}: n={pc.shape[0]} m={m} env={env} plan={gpu.plan_info(pc.shape[0], 3, m=m)}")
    rc, out = gpu.fps(np.asfortranarray(pc), m, 0, layout=COLS)
    assert rc == 0, gpu.last_error()
    i = first_diff(out, d["out"])
    assert i < 0, f"{name} {env}: first differing pick {i} (got {out[i]}, reference {d['out'][i]})"


GAUSS_PLANS = [{}, {"FPSX_NO_PAIR": "1"}, {"FPSX_NO_PRUNE": "1"}]


@pytest.fixture(scope="module")
def gauss1m():
    return bench.gen_gauss_mixture(1 << 20)


@pytest.mark.parametrize("env", GAUSS_PLANS, ids=["auto", "no_pair", "no_prune"])
def test_gauss1m_every_pick(gpu, monkeypatch, gauss1m, env):
    check_single(gpu, monkeypatch, "gauss1m", gauss1m, env)


@pytest.mark.parametrize("n,name,env", [((1 << 20) + 1, "gauss1m1", {}),
                                        ((1 << 20) + 1, "gauss1m1", {"FPSX_PLAN": "pairw2:1024:16"}),
                                        (1 << 21, "gauss2m", {}),
                                        (1 << 21, "gauss2m", {"FPSX_PLAN": "pairw4:1024:16"}),
                                        (1 << 21, "gauss2m", {"FPSX_NO_PAIR": "1"})],
                         ids=["gauss1m1-auto", "gauss1m1-pairw2", "gauss2m-auto", "gauss2m-pairw4",
                              "gauss2m-no_pair"])
def test_gauss_beyond_1m_every_pick(gpu, monkeypatch, n, name, env):
    # The planner's choices (2^20 + 1 points: 61 workgroups of 1024 x 17, one
    # point per thread in LDS; 2^21: two polling waves over 128 workgroups),
    # the two-wave form forced at 65 workgroups, four waves forced, and the
    # one-pick pruned kernel.
    check_single(gpu, monkeypatch, name, bench.gen_gauss_mixture(n), env)


@pytest.fixture(scope="module")
def lattice16m():
    return bench.gen_lattice_dups(1 << 24)


@pytest.mark.parametrize("env", [{}, {"FPSX_PLAN": "spruned"}, {"FPSX_PLAN": "stream:1024"},
                                 {"FPSX_PLAN": "spruned:128"}],
                         ids=["auto", "spruned", "stream1024", "spruned128"])
def test_lattice16m_every_pick(gpu, monkeypatch, lattice16m, env):
    check_single(gpu, monkeypatch, "lattice16m", lattice16m, env)


@pytest.mark.parametrize("n,name,env", [(50_000, "uniform50k", {}),
                                        (50_000, "uniform50k", {"FPSX_NO_PAIR": "1"}),
                                        (100_000, "uniform100k", {}),
                                        (100_000, "uniform100k", {"FPSX_NO_PAIR": "1"}),
                                        (100_000, "uniform100k", {"FPSX_NO_PRUNE": "1"})],
                         ids=["50k-auto", "50k-no_pair", "100k-auto", "100k-no_pair",
                              "100k-no_prune"])
def test_reference_bench_cases_every_pick(gpu, monkeypatch, n, name, env):
    check_single(gpu, monkeypatch, name, bench.gen_uniform_n(n), env)


@pytest.mark.parametrize("n,name,h", [(4096, "kdline4k", 5), (50_000, "kdline50k", 7),
                                      (100_000, "kdline100k", 7)])
def test_reference_kdline_bench_cases_every_pick(gpu, monkeypatch, n, name, h):
    # bucket_fps_kdline_sampling with the reference's bench heights: the KD-line
    # order built on the host, lowest-position ties and FLT_MAX initial
    # distances on the GPU, against the reference's bucket search.
    import fpsample_amd as F

    d = FL.load(name)
    pc = bench.gen_uniform_n(n)
    assert FL.sha256(pc) == d["cloud_sha256"]
    for k in ("FPSX_PLAN", "FPSX_NO_PAIR", "FPSX_NO_PRUNE"):
        monkeypatch.delenv(k, raising=False)
    m = int(d["m"])
    log(f"{name}: n={n} m={m} h={h} plan={gpu.plan_info(n, 3, m=m, tie=1)}")
    out = F.bucket_fps_kdline_sampling(pc, m, h, start_idx=0)
    i = first_diff(out, d["out"])
    assert i < 0, f"{name}: first differing pick {i} (got {out[i]}, reference {d['out'][i]})"


def check_many(name, clouds, ms, got):
    d = FL.load(name)
    assert len(clouds) == len(d["n"])
    bad = []
    for i, (pc, m, g) in enumerate(zip(clouds, ms, got)):
        assert FL.sha256(pc) == str(d["cloud_sha256"][i]), f"{name} cloud {i}: regenerated cloud differs"
        assert int(d["m"][i]) == m and len(g) == m
        if FL.picks_sha256(g) != str(d["picks_sha256"][i]):
            head = d["head"][i][:min(16, m)]
            bad.append((i, pc.shape[0], m, first_diff(g[:head.size], head)))
    assert not bad, f"{name}: {len(bad)} clouds differ from the reference (cloud, n, m, first diff in head): {bad[:8]}"


def test_batch256_every_pick_autotuned_graph(gpu):
    from fpsample_amd.batch import DeviceBatch

    clouds, ms = bench.gen_ragged_batch()
    b = DeviceBatch(clouds, ms, [0] * len(clouds), graph=True, autotune=True)
    log(f"batch256: {b.info()}")
    for rep in range(2):  # the captured graph, replayed twice
        b.run()
        b.status()
        check_many("batch256", clouds, ms, b.picks())
    b.close()


def test_serve96_every_pick_default_shape(gpu):
    # The serving loop's default shape (1024 x 16 members, teams of up to 64,
    # 256-pass chunks with min-distances saved and restored between them),
    # all 96 requests of the serving workload in flight together.
    from fpsample_amd.serve import Server

    reqs = bench.gen_serving_stream()
    with Server(hbm_budget_bytes=16 << 30, chunk_passes=256) as srv:
        ids = [srv.submit(pc, m, 0) for pc, m in reqs]
        got = [srv.wait(r)[0] for r in ids]
        st = srv.stats()
    log(f"serve96: {st}")
    assert st["completed"] == len(reqs) and st["error"] == 0
    check_many("serve96", [pc for pc, _ in reqs], [m for _, m in reqs], got)

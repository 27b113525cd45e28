"""GPU parity of the paired pruned kernel (up to two picks per exchange).

The paired kernel must give exactly the one-pick loop's picks: every shape is
forced through FPSX_PLAN and compared with the oracle on clustered, uniform,
duplicate-heavy (exact ties, repeated picks once every distance is zero) and
special-value clouds, with single and multi-start prefixes, both tie rules
(the bucket entry uses lowest-position ties), odd and even sample counts, and
at the benchmark size against the one-pick pruned kernel.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.capi import COLS

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


def gauss_cloud(n, seed):
    rng = np.random.default_rng(seed)
    centers = rng.random((64, 3))
    lab = rng.integers(0, 64, n)
    return (centers[lab] + 0.02 * rng.standard_normal((n, 3))).astype(np.float32)


# 1024 x 17: 16 points per thread in registers and one in LDS.
PAIR_SHAPES = ["pair:1024:16", "pair:1024:8", "pair:1024:4", "pair:1024:17"]


@pytest.mark.parametrize("shape", PAIR_SHAPES)
def test_pair_shapes(gpu, plan_env, shape):
    plan_env(shape)
    rng = np.random.default_rng(41)
    clouds = [("gauss", gauss_cloud(40000, 3)),
              ("lattice_dups", np.repeat(rng.integers(0, 6, (6000, 3)), 3, axis=0).astype(np.float32)),
              ("uniform", rng.random((30000, 3)).astype(np.float32))]
    for name, pc in clouds:
        info = gpu.plan_info(len(pc), 3, m=2001)
        assert info["kind"] == 5 and 2 <= info["grid"] <= 64, info
        for m, starts in ((2001, 0), (2000, [len(pc) - 1, 7, 11])):
            log(f"{shape} {name} m={m} starts={starts} plan={info}")
            want = O.fps(pc, m, starts)
            np.testing.assert_array_equal(run_fps(gpu, pc, m, starts), want,
                                          err_msg=f"{shape} {name} m={m}")


def test_pair_long_forced_prefix(gpu, plan_env):
    # A long forced prefix (multi-start / resume) leaves every wave's
    # candidate cached before the first exchange; workgroups the last forced
    # pick does not touch must still publish their best.
    plan_env("pair:1024:4")
    pc = gauss_cloud(40000, 6)
    want = O.fps(pc, 1500, 0)
    for k in (2, 100, 900):
        log(f"forced prefix of {k}")
        np.testing.assert_array_equal(run_fps(gpu, pc, 1500, want[:k]), want)


def test_pair_all_distances_zero(gpu, plan_env):
    # 20 distinct sites, 40000 points: after 20 picks every min-distance is
    # zero and the loop repeats the highest index; pairs must do the same.
    plan_env("pair:1024:4")
    rng = np.random.default_rng(42)
    sites = rng.random((20, 3)).astype(np.float32)
    pc = sites[rng.integers(0, 20, 40000)]
    for m in (64, 65):
        log(f"zero distances m={m}")
        np.testing.assert_array_equal(run_fps(gpu, pc, m, 5), O.fps(pc, m, 5))


def test_pair_special_values(gpu, plan_env):
    plan_env("pair:1024:4")
    rng = np.random.default_rng(43)
    pc = (rng.random((9000, 3)) * 4e19).astype(np.float32)
    pc[[5, 700, 4000]] = np.nan
    pc[9] = np.inf
    pc[10, 1] = -np.inf
    log("special values: nan/inf/huge")
    np.testing.assert_array_equal(run_fps(gpu, pc, 600, 0), O.fps(pc, 600, 0))
    tiny = (rng.random((9000, 3)) * 1e-21).astype(np.float32)
    log("special values: denormal distances")
    np.testing.assert_array_equal(run_fps(gpu, tiny, 600, 0), O.fps(tiny, 600, 0))


def test_pair_lowest_ties_bucket_entry(gpu, plan_env):
    # Bucket entry: lowest-position ties over the KD-line order.
    plan_env("pair:1024:4")
    rng = np.random.default_rng(44)
    lat = rng.integers(0, 40, (100000, 3)).astype(np.float32)
    log("bucket entry, lattice, pair:1024:4")
    rc, out = gpu.kdline(lat, 1500, 123, 7)
    assert rc == 0, gpu.last_error()
    np.testing.assert_array_equal(out, O.kdline_quickfps(lat, 1500, 7, 123))


def test_pair_lds_row_lowest_ties_and_specials(gpu, plan_env):
    # The LDS row of 1024 x 17: lowest-position ties through the bucket entry
    # on a lattice, and NaN / inf / huge coordinates.
    plan_env("pair:1024:17")
    rng = np.random.default_rng(46)
    lat = rng.integers(0, 40, (100000, 3)).astype(np.float32)
    log("bucket entry, lattice, pair:1024:17")
    rc, out = gpu.kdline(lat, 1500, 123, 7)
    assert rc == 0, gpu.last_error()
    np.testing.assert_array_equal(out, O.kdline_quickfps(lat, 1500, 7, 123))
    pc = (rng.random((60000, 3)) * 4e19).astype(np.float32)
    pc[[5, 17500, 40000]] = np.nan
    pc[9] = np.inf
    log("special values, pair:1024:17")
    np.testing.assert_array_equal(run_fps(gpu, pc, 700, [0, 3]), O.fps(pc, 700, [0, 3]))


@pytest.mark.parametrize("n", [(1 << 20) + 1, 64 * 17408])
def test_pair_lds_row_planner_sizes(gpu, plan_env, n):
    # Just beyond 2^20 points and the largest cloud 64 workgroups of 1024 x 17
    # hold: the planner keeps the one-wave paired kernel, oracle-compared on a
    # prefix (clusters and 10 % exact duplicates, a forced two-pick prefix).
    plan_env(None)
    pc = gauss_cloud(n, 12)
    pc[::10] = pc[1::10][: len(pc[::10])]
    info = gpu.plan_info(n, 3, m=16384)
    log(f"n={n} planner(16384)={info}")
    assert info["kind"] == 5 and info["ppt"] == 17 and info["grid"] <= 64, info
    np.testing.assert_array_equal(run_fps(gpu, pc, 16384, [3, 0])[:600], O.fps(pc, 600, [3, 0]))


def test_pair_full_size_matches_one_pick_kernel(gpu, plan_env):
    # The benchmark cloud (2^20 Gaussian-cluster points, 16384 picks): the
    # paired kernel and the one-pick pruned kernel agree on every pick, and
    # the first 512 match the oracle.
    pc = gauss_cloud(1 << 20, 1)
    plan_env("pruned:1024:16")
    one = run_fps(gpu, pc, 16384, 0)
    plan_env("pair:1024:16")
    log(f"full size plan={gpu.plan_info(len(pc), 3, m=16384)}")
    two = run_fps(gpu, pc, 16384, 0)
    np.testing.assert_array_equal(two, one)
    np.testing.assert_array_equal(two[:512], O.fps(pc, 512, 0))


def test_pair_rejects_grids_outside_range(gpu, plan_env):
    plan_env("pair:1024:16")
    pc = np.random.default_rng(45).random((9000, 3)).astype(np.float32)
    rc, _ = gpu.fps(pc, 100, 0)
    assert rc != 0 and "2..256" in gpu.last_error()  # one workgroup: nothing to pair across


def test_pair_random_configurations(gpu, plan_env):
    # Randomised: the paired kernel against the one-pick pruned kernel (itself
    # checked against the oracle above) on clouds of 70 K - 1 M points,
    # clustered, uniform, anisotropic and duplicate-heavy, random sample
    # counts, starts and forced prefixes; the lowest-tie rule through the
    # bucket entry on a few of them.
    rng = np.random.default_rng(2024)
    for case in range(16):
        n = int(rng.integers(70_000, 1 << 20))
        kind = case % 4
        if kind == 0:
            pc = gauss_cloud(n, int(rng.integers(1 << 30)))
        elif kind == 1:
            pc = rng.random((n, 3)).astype(np.float32)
        elif kind == 2:
            pc = (rng.standard_normal((n, 3)) * np.array([10.0, 1.0, 0.01])).astype(np.float32)
        else:
            sites = rng.integers(0, 64, (n // 4, 3))
            pc = sites[rng.integers(0, n // 4, n)].astype(np.float32)
        m = int(rng.integers(1024, 12_000))
        starts = int(rng.integers(n)) if case % 3 else [int(x) for x in rng.integers(0, n, 5)]
        plan_env("pruned:1024:16")
        one = run_fps(gpu, pc, m, starts)
        plan_env(None)
        info = gpu.plan_info(n, 3, m=m)
        log(f"random case {case}: n={n} kind={kind} m={m} starts={starts} plan={info}")
        assert info["kind"] == 5, info
        np.testing.assert_array_equal(run_fps(gpu, pc, m, starts), one, err_msg=f"case {case}")
        if case % 4 == 1:
            h = int(rng.integers(3, 10))
            plan_env("pruned:1024:16")
            rc, want = gpu.kdline(pc, 2000, h, 11)
            assert rc == 0, gpu.last_error()
            plan_env(None)
            rc, got = gpu.kdline(pc, 2000, h, 11)
            assert rc == 0, gpu.last_error()
            np.testing.assert_array_equal(got, want, err_msg=f"bucket case {case}")


@pytest.mark.parametrize("pw", [2, 4])
def test_wide_pair_small_grids(gpu, plan_env, pw):
    # The several-polling-wave form forced on grids of 2..40 workgroups (the
    # other polling waves see no workgroup): clustered, uniform, duplicate-
    # heavy and special-value clouds, multi-start prefixes, both tie rules.
    plan_env(f"pairw{pw}:1024:16")
    rng = np.random.default_rng(60 + pw)
    big = (rng.random((70000, 3)) * 4e19).astype(np.float32)
    big[rng.integers(0, 70000, 3)] = np.nan
    clouds = [("gauss", gauss_cloud(600_000, 5), [0]),
              ("lattice_dups", np.repeat(rng.integers(0, 7, (12000, 3)), 4, axis=0).astype(np.float32), [5, 7]),
              ("uniform", rng.random((33000, 3)).astype(np.float32), [32999]),
              ("special", big, [1])]
    for name, pc, st in clouds:
        for m in (700, 1201):
            log(f"pairw{pw} {name} n={len(pc)} m={m}")
            np.testing.assert_array_equal(run_fps(gpu, pc, m, st), O.fps(pc, m, st),
                                          err_msg=f"pairw{pw} {name} m={m}")
    # lowest-position ties (the bucket entry's rule) through the KD-line order
    pc = np.repeat(rng.integers(0, 5, (9000, 3)), 3, axis=0).astype(np.float32)
    rc, got = gpu.kdline(pc, 900, 4, 4)
    assert rc == 0, gpu.last_error()
    np.testing.assert_array_equal(got, O.kdline_quickfps(pc, 900, 4, 4))


@pytest.mark.parametrize("n,plan", [(1_200_000, None), (1_200_000, "pairw4:1024:16"),
                                    (3_000_000, None)])
def test_wide_pair_beyond_1m(gpu, plan_env, n, plan):
    # 74 and 184 workgroups: the planner's 2- and 4-wave forms (and 4 waves
    # forced at 74), oracle-compared on a prefix; the clouds have clusters
    # and 10 % exact duplicates.
    plan_env(plan)
    pc = gauss_cloud(n, 9)
    pc[::10] = pc[1::10][: len(pc[::10])]
    m = 500
    info = gpu.plan_info(n, 3, m=16384)
    log(f"n={n} plan={plan} planner(16384)={info}")
    if plan is None:
        assert info["kind"] == 5 and info["grid"] > 64
    np.testing.assert_array_equal(run_fps(gpu, pc, m, [0]), O.fps(pc, m, 0), err_msg=f"n={n} {plan}")

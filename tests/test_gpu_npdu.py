"""GPU parity of windowed-update FPS (fps_npdu_sampling, src/lib.cpp:272-340).

Bit-exact against the reference's golden vectors and against the oracle
(oracle.npdu, pinned by those vectors and by hand-derived known answers).
FPSX_NPDU_PLAN forces every kernel shape: the 3-D kernel with coordinates in
registers and the any-dimension kernel that reads only the window, each with
one and with several workgroups.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def npdu_plan(monkeypatch):
    def set_plan(p):
        if p is None:
            monkeypatch.delenv("FPSX_NPDU_PLAN", raising=False)
        else:
            monkeypatch.setenv("FPSX_NPDU_PLAN", p)
    yield set_plan
    monkeypatch.delenv("FPSX_NPDU_PLAN", raising=False)


def run(gpu, pc, m, w, s):
    rc, out = gpu.npdu(pc, m, w, s)
    assert rc == 0, f"rc={rc}: {gpu.last_error()}"
    return out


def test_golden_npdu_capi(gpu, golden):
    n = 0
    for name, g in golden.items():
        if g["kind"] != "npdu":
            continue
        log(f"golden {name}")
        np.testing.assert_array_equal(run(gpu, g["pc"], g["m"], g["h"], int(g["start"][0])),
                                      g["out"], err_msg=name)
        n += 1
    assert n >= 6


SHAPES = ["3:256:1", "3:256:4", "3:256:16", "3:1024:1", "3:1024:4", "3:1024:8", "3:1024:16",
          "0:256:1", "0:256:4", "0:1024:4", "0:1024:16", "0:1024:32"]


@pytest.mark.parametrize("shape", SHAPES)
def test_npdu_shapes(gpu, npdu_plan, shape):
    npdu_plan(shape)
    rng = np.random.default_rng(41)
    clouds = [("uniform", rng.random((5000, 3)).astype(np.float32)),
              ("sorted_gauss", np.sort(rng.standard_normal((4000, 3)).astype(np.float32), axis=0)),
              ("lattice_dups", np.repeat(rng.integers(0, 5, (1200, 3)), 3, axis=0).astype(np.float32))]
    for name, pc in clouds:
        n = len(pc)
        log(f"npdu {shape} {name} plan={gpu.npdu_plan_info(n, 3)}")
        for w, s in ((0, 0), (1, n - 1), (7, 123), (64, n // 2), (1000, 5), (10 ** 12, 9)):
            np.testing.assert_array_equal(run(gpu, pc, 600, w, s), O.npdu(pc, 600, w, s),
                                          err_msg=f"{shape} {name} w={w} s={s}")


@pytest.mark.parametrize("d", [1, 2, 4, 5, 8, 13])
def test_npdu_dimensions(gpu, d):
    rng = np.random.default_rng(42 + d)
    pc = rng.standard_normal((3000, d)).astype(np.float32)
    log(f"npdu d={d} plan={gpu.npdu_plan_info(3000, d)}")
    for w in (3, 48, 5000):
        np.testing.assert_array_equal(run(gpu, pc, 700, w, 11), O.npdu(pc, 700, w, 11),
                                      err_msg=f"d={d} w={w}")


def test_npdu_edge_sizes(gpu):
    rng = np.random.default_rng(43)
    for n in (1, 2, 3, 63, 64, 65, 256, 257, 1024, 1025, 16384, 16385, 20000):
        pc = rng.random((n, 3)).astype(np.float32)
        for m in sorted({k for k in (1, 2, min(n, 5), min(n, 300)) if k <= n}):
            log(f"npdu edge n={n} m={m}")
            w = max(1, int(n / m * 16))
            np.testing.assert_array_equal(run(gpu, pc, m, w, n - 1), O.npdu(pc, m, w, n - 1),
                                          err_msg=f"n={n} m={m}")


def test_npdu_special_values(gpu, npdu_plan):
    rng = np.random.default_rng(44)
    pc = (rng.random((3000, 3)) * 4e19).astype(np.float32)   # squares overflow to +inf
    pc[[5, 700]] = np.nan
    pc[9] = np.inf
    for plan in (None, "0:1024:4"):
        npdu_plan(plan)
        np.testing.assert_array_equal(run(gpu, pc, 400, 32, 0), O.npdu(pc, 400, 32, 0))
    same = np.ones((2000, 3), np.float32)                      # every distance 0: index order
    np.testing.assert_array_equal(run(gpu, same, 50, 16, 3), O.npdu(same, 50, 16, 3))


def test_npdu_large_multi_workgroup(gpu, npdu_plan):
    # Several workgroups exchange candidates every pass; the window is the
    # Python wrapper's default, n / m * 16.
    rng = np.random.default_rng(45)
    n, m = 200_000, 2048
    pc = np.sort(rng.random((n, 3)).astype(np.float32), axis=0)
    w = int(n / m * 16)
    want = O.npdu(pc, m, w, 1234)
    for plan in (None, "0:1024:32", "3:1024:4"):
        npdu_plan(plan)
        log(f"npdu large plan={plan} -> {gpu.npdu_plan_info(n, 3)}")
        np.testing.assert_array_equal(run(gpu, pc, m, w, 1234), want, err_msg=str(plan))


PAIR_SHAPES = ["pair:1024:2", "pair:1024:4", "pair:1024:8", "pair:1024:16"]


@pytest.mark.parametrize("shape", PAIR_SHAPES)
def test_npdu_pair_shapes(gpu, npdu_plan, shape):
    # Paired form: up to two picks per exchange, same picks as the reference.
    npdu_plan(shape)
    rng = np.random.default_rng(46)
    clouds = [("uniform", rng.random((70000, 3)).astype(np.float32)),
              ("sorted_gauss", np.sort(rng.standard_normal((60000, 3)).astype(np.float32), axis=0)),
              ("lattice_dups", np.repeat(rng.integers(0, 6, (12000, 3)), 3, axis=0).astype(np.float32))]
    for name, pc in clouds:
        n = len(pc)
        info = gpu.npdu_plan_info(n, 3)
        log(f"npdu {shape} {name} plan={info}")
        assert info["grid"] >= 2
        for m, w, s in ((1500, 0, 0), (1501, 1, n - 1), (1500, 7, 123), (1200, 64, n // 2),
                        (1001, 5000, 5), (800, 10 ** 12, 9), (1500, int(n / 1500 * 16), 77)):
            np.testing.assert_array_equal(run(gpu, pc, m, w, s), O.npdu(pc, m, w, s),
                                          err_msg=f"{shape} {name} m={m} w={w} s={s}")


def test_npdu_pair_special_values(gpu, npdu_plan):
    npdu_plan("pair:1024:4")
    rng = np.random.default_rng(47)
    pc = (rng.random((30000, 3)) * 4e19).astype(np.float32)   # squares overflow to +inf
    pc[[5, 700, 20000]] = np.nan
    pc[9] = np.inf
    log("npdu pair special values")
    np.testing.assert_array_equal(run(gpu, pc, 600, 32, 0), O.npdu(pc, 600, 32, 0))
    same = np.ones((20000, 3), np.float32)                      # every distance 0: index order
    np.testing.assert_array_equal(run(gpu, same, 300, 16, 3), O.npdu(same, 300, 16, 3))


def test_npdu_pair_reference_benchmark_size(gpu, npdu_plan):
    # The reference's own NPDU benchmark size (bench/test_bench.py:151-154):
    # 100000 points, 50000 picks, w = 32; the paired and the one-pick
    # kernels agree on every pick and the first 2000 match the oracle.
    import bench

    pc = bench.gen_uniform_n(100000)
    npdu_plan("3:1024:8")
    one = run(gpu, pc, 50000, 32, 0)
    npdu_plan(None)
    log(f"npdu 100k auto plan={gpu.npdu_plan_info(100000, 3)}")
    two = run(gpu, pc, 50000, 32, 0)
    np.testing.assert_array_equal(two, one)
    np.testing.assert_array_equal(two[:2000], O.npdu(pc, 2000, 32, 0))


@pytest.mark.parametrize("n,m,w,name", [(4096, 1024, 64, "npdu4k"), (50_000, 4096, 195, "npdu50k"),
                                        (100_000, 50_000, 32, "npdu100k")],
                         ids=["4k", "50k", "100k"])
@pytest.mark.parametrize("plan", [None, "3:1024:8", "stream"], ids=["auto", "one_pick", "stream"])
def test_npdu_reference_bench_cases_every_pick(gpu, npdu_plan, n, m, w, name, plan):
    # The reference's own NPDU bench cases (bench/test_bench.py:42-46, 96-100,
    # 150-154; the wrapper's default window int(N / M * 16)), every pick against
    # the reference's output (full-length fixtures), under the planner's choice,
    # the one-pick register kernel and the streaming form.
    import bench
    from tests.golden import full_length as FL

    d = FL.load(name)
    pc = bench.gen_uniform_n(n)
    assert FL.sha256(pc) == d["cloud_sha256"]
    npdu_plan(plan)
    log(f"{name}: n={n} m={m} w={w} plan={plan} info={gpu.npdu_plan_info(n, 3)}")
    out = run(gpu, pc, m, w, 0)
    bad = np.nonzero(out != d["out"])[0]
    assert bad.size == 0, f"{name} {plan}: first differing pick {bad[0]}"


def test_npdu_exchange_count(gpu, npdu_plan):
    # fpsx_launch_exchanges after a windowed-update launch: the paired kernel
    # counts its rounds (between (m - 1) / 2 and m - 1 exchanges), the one-pick
    # kernel makes one per pick after the start.
    import ctypes

    import bench

    L = gpu.L
    L.fpsx_launch_exchanges.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    x = ctypes.c_uint64()
    pc = bench.gen_uniform_n(50_000)
    for plan, pair in ((None, True), ("3:1024:8", False)):
        npdu_plan(plan)
        run(gpu, pc, 4096, 195, 0)
        assert L.fpsx_launch_exchanges(None, ctypes.byref(x)) == 0, gpu.last_error()
        log(f"npdu 50k plan={plan}: {x.value} exchanges")
        if pair:
            assert (4096 - 1) // 2 <= x.value < 4096 - 1
        else:
            assert x.value == 4096 - 1


def test_npdu_python_api_and_device_entry(gpu):
    import fpsample_amd as F
    from fpsample_amd import device as D

    rng = np.random.default_rng(46)
    pc = rng.random((4096, 3)).astype(np.float64)
    np.random.seed(3)
    got = F.fps_npdu_sampling(pc, 1024)
    np.random.seed(3)
    s = np.random.randint(low=0, high=4096)
    want = O.npdu(pc.astype(np.float32), 1024, int(4096 / 1024 * 16), s)
    np.testing.assert_array_equal(got, want)
    assert got.dtype == np.uint64
    for d in (3, 6):
        q = rng.standard_normal((5000, d)).astype(np.float32)
        rows = D.DeviceArray.from_numpy(q)
        ds = D.DeviceArray.from_numpy(np.array([77], np.uint64))
        out = D.DeviceArray((500,), np.uint64)
        opts = D.DeviceArray((500, d), np.float32)
        D.fps_npdu_sampling_device(rows, d, 1, 5000, d, 500, 40, ds, out, opts)
        D.synchronize()
        D.device_status()
        want = O.npdu(q, 500, 40, 77)
        np.testing.assert_array_equal(out.numpy(), want, err_msg=f"d={d}")
        np.testing.assert_array_equal(opts.numpy(), q[want.astype(np.int64)])


def test_npdu_error_codes(gpu):
    pc = np.zeros((10, 3), np.float32)
    assert gpu.npdu(pc, 3, 4, 10)[0] == 2      # start out of range
    assert gpu.npdu(pc, 11, 4, 0)[0] == 3      # n_samples > n_points
    assert gpu.npdu(pc, 3, 4, 0, dim=0)[0] == 1


@pytest.mark.parametrize("plan", ["stream", "stream:1", "stream:4", "stream:13", "stream0",
                                  "stream0:4", "streambig", "streambig:3"])
def test_npdu_streaming_form(gpu, npdu_plan, plan):
    # Min-distances in HBM, per-block cached bests, window blocks only.  3-D
    # clouds take the kernel whose winners carry their coordinates; "stream0"
    # forces the any-dimension kernel on them too, "streambig" its
    # large-block instance (the planner's choice beyond about 805 M points).
    npdu_plan(plan)
    rng = np.random.default_rng(47)
    for name, pc, w in (("uniform3", rng.random((20000, 3)).astype(np.float32), 40),
                        ("dims5", rng.standard_normal((15000, 5)).astype(np.float32), 3000),
                        ("lattice_dups", np.repeat(rng.integers(0, 6, (4000, 2)), 4, axis=0).astype(np.float32), 17),
                        ("tiny_window", np.sort(rng.random((13 * 1024 + 3, 3)).astype(np.float32), axis=0), 0)):
        log(f"npdu stream {plan} {name} plan={gpu.npdu_plan_info(len(pc), pc.shape[1])}")
        for s0 in (0, len(pc) - 1, 5000):
            np.testing.assert_array_equal(run(gpu, pc, 400, w, s0), O.npdu(pc, 400, w, s0),
                                          err_msg=f"{plan} {name} s={s0}")


def test_npdu_beyond_register_file(gpu):
    # 9 M points: more than any register-resident shape holds, so the planner
    # takes the streaming form; the oracle checks a prefix.
    rng = np.random.default_rng(48)
    n = 9 * (1 << 20)
    pc = rng.random((n, 3)).astype(np.float32)
    info = gpu.npdu_plan_info(n, 3)
    log(f"npdu 9M plan {info}")
    assert info["ppt"] == 0 and info["kernel_dim"] == 3
    w = int(n / 4096 * 16)
    np.testing.assert_array_equal(run(gpu, pc, 40, w, 123), O.npdu(pc, 40, w, 123))


def test_npdu_pair_random_configurations(gpu, npdu_plan):
    # Randomised: the paired windowed-update kernel (the planner's choice for
    # 2..64 workgroups) against the one-pick kernel on random sizes, sample
    # counts, windows and starts.
    rng = np.random.default_rng(77)
    for case in range(12):
        n = int(rng.integers(20_000, 1 << 20))
        pc = (np.sort(rng.random((n, 3)), axis=0) if case % 2 else rng.random((n, 3))).astype(np.float32)
        m = int(rng.integers(500, 6000))
        w = int(rng.choice([0, 1, 7, 33, int(n / m * 16), 5 * n]))
        s = int(rng.integers(n))
        npdu_plan("3:1024:8" if n <= 8192 * 1024 else None)
        one = run(gpu, pc, m, w, s)
        npdu_plan(None)
        log(f"npdu random case {case}: n={n} m={m} w={w} s={s} plan={gpu.npdu_plan_info(n, 3)}")
        np.testing.assert_array_equal(run(gpu, pc, m, w, s), one, err_msg=f"case {case}")


@pytest.mark.parametrize("plan", ["stream", "stream0:3", "pair:1024:2"])
def test_npdu_fused_gather_plans(gpu, npdu_plan, plan):
    # The 3-D streaming kernel writes the picks' coordinates from the
    # exchanged winner words, the any-dimension one from its pivot in LDS,
    # the paired kernel for both picks of a round.
    from fpsample_amd import device as D

    npdu_plan(plan)
    rng = np.random.default_rng(50)
    n, m = 9000, 701
    q = rng.standard_normal((n, 3)).astype(np.float32)
    rows = D.DeviceArray.from_numpy(q)
    ds = D.DeviceArray.from_numpy(np.array([77], np.uint64))
    out = D.DeviceArray((m,), np.uint64)
    opts = D.DeviceArray((m, 3), np.float32)
    log(f"npdu fused gather {plan}")
    D.fps_npdu_sampling_device(rows, 3, 1, n, 3, m, 40, ds, out, opts)
    D.synchronize()
    D.device_status()
    want = O.npdu(q, m, 40, 77)
    np.testing.assert_array_equal(out.numpy(), want, err_msg=plan)
    np.testing.assert_array_equal(opts.numpy(), q[want.astype(np.int64)], err_msg=plan)

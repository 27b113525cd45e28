"""GPU parity of the batched path: every cloud of a batch equals the oracle's
fps_sampling of that cloud alone, for every batch-kernel shape, with teams of
several workgroups, through HIP-graph replays and through the Python API."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def ragged(seed, k, lo, hi, dim=3):
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(lo), np.log(hi), k)).astype(np.int64)
    clouds = []
    for i, n in enumerate(ns):
        if i % 3 == 0:
            pc = rng.random((n, dim))
        elif i % 3 == 1:
            v = rng.standard_normal((n, dim))
            pc = v / np.linalg.norm(v, axis=1, keepdims=True) + 0.01 * rng.standard_normal((n, dim))
        else:
            pc = rng.integers(0, 4, (n, dim))  # ties
        clouds.append(pc.astype(np.float32))
    return clouds


def oracle_all(clouds, ms, starts):
    return [O.fps(pc, m, s) for pc, m, s in zip(clouds, ms, starts)]


@pytest.mark.parametrize("shape", ["1024:16", "1024:8", "256:32", "256:16", "256:8", "256:4"])
def test_batch_shapes(gpu, monkeypatch, shape):
    import fpsample_amd as F

    monkeypatch.setenv("FPSX_BATCH_SHAPE", shape)
    clouds = ragged(1, 24, 50, 20000)
    ms = [max(1, len(pc) // 8) for pc in clouds]
    starts = [len(pc) // 3 for pc in clouds]
    log(f"shape {shape}: sizes {[len(c) for c in clouds]}")
    got = F.fps_sampling_batch(clouds, ms, starts)
    for i, (g, w) in enumerate(zip(got, oracle_all(clouds, ms, starts))):
        np.testing.assert_array_equal(g, w, err_msg=f"shape {shape} cloud {i} n={len(clouds[i])}")


def test_batch_teams_large_clouds(gpu, monkeypatch):
    import fpsample_amd as F

    # 256:16 holds 4096 points per workgroup: clouds of 3k..60k need teams of 1..15
    monkeypatch.setenv("FPSX_BATCH_SHAPE", "256:16")
    clouds = ragged(2, 10, 3000, 60000)
    ms = [min(len(pc), 600) for pc in clouds]
    got = F.fps_sampling_batch(clouds, ms, 0)
    for i, (g, w) in enumerate(zip(got, oracle_all(clouds, ms, [0] * len(clouds)))):
        np.testing.assert_array_equal(g, w, err_msg=f"cloud {i} n={len(clouds[i])}")


def test_batch_other_dims_fall_back(gpu):
    import fpsample_amd as F

    for d in (1, 2, 5):
        clouds = ragged(3 + d, 6, 20, 3000, dim=d)
        ms = [max(1, len(pc) // 4) for pc in clouds]
        got = F.fps_sampling_batch(clouds, ms, 1)
        for g, w in zip(got, oracle_all(clouds, ms, [1] * len(clouds))):
            np.testing.assert_array_equal(g, w)


def test_batch_graph_replays(gpu):
    from fpsample_amd import device as D
    from fpsample_amd import batch as Bt

    clouds = ragged(4, 40, 100, 30000)
    ms = [max(1, len(pc) // 4) for pc in clouds]
    starts = [0] * len(clouds)
    want = np.concatenate(oracle_all(clouds, ms, starts))
    b = Bt.DeviceBatch(clouds, ms, starts, graph=True)
    for rep in range(4):
        b.run()
        b.status()
        np.testing.assert_array_equal(b.picks_flat(), want, err_msg=f"replay {rep}")
        # eager work between replays: a single-cloud run on the same device
        rc_pc = clouds[rep]
        got1, _ = D.fps_sampling_tensor(D.DeviceArray.from_numpy(rc_pc), ms[rep], 0)
        np.testing.assert_array_equal(got1.numpy(), O.fps(rc_pc, ms[rep], 0))
    b.close()


def test_batch_oversize_cloud_falls_back(gpu):
    # A cloud larger than any batch team (64 workgroups) sends the batch to
    # one planned launch per cloud; every result still equals the oracle's.
    import fpsample_amd as F
    from fpsample_amd import batch as Bt

    rng = np.random.default_rng(77)
    big = rng.random((1_200_000, 3)).astype(np.float32)
    clouds = [rng.random((5000, 3)).astype(np.float32), big, rng.random((300, 3)).astype(np.float32)]
    ms = [500, 64, 100]
    got = F.fps_sampling_batch(clouds, ms, 2)
    for g, w in zip(got, oracle_all(clouds, ms, [2, 2, 2])):
        np.testing.assert_array_equal(g, w)
    b = Bt.DeviceBatch(clouds, ms, [2, 2, 2])
    assert b.info()["block"] == 0  # the per-cloud form


def test_batch_paired_teams_match_one_pick(gpu, monkeypatch):
    # Teams settle up to two picks per exchange; FPSX_BATCH_NO_PAIR=1 runs the
    # one-pick team loop.  256:4 holds 1024 points per workgroup, so clouds of
    # 2k..60k form teams of 2..59: random clouds (uniform, sphere, integer
    # ties, all points equal, nan/inf/huge coordinates), odd and even sample
    # counts, random starts; both loops and the oracle must agree.
    import fpsample_amd as F

    monkeypatch.setenv("FPSX_BATCH_SHAPE", "256:4")
    rng = np.random.default_rng(31)
    clouds, ms, starts = [], [], []
    for i in range(18):
        n = int(np.exp(rng.uniform(np.log(2000), np.log(60000))))
        kind = i % 6
        if kind == 0:
            pc = rng.random((n, 3))
        elif kind == 1:
            v = rng.standard_normal((n, 3))
            pc = v / np.linalg.norm(v, axis=1, keepdims=True)
        elif kind == 2:
            pc = rng.integers(0, 4, (n, 3))
        elif kind == 3:
            pc = np.tile(rng.random((1, 3)), (n, 1))
        elif kind == 4:
            pc = rng.random((n, 3)) * 4e19
            pc[rng.integers(0, n, 3)] = np.nan
            pc[rng.integers(0, n), 1] = np.inf
        else:
            pc = rng.standard_normal((n, 3)) * np.array([100.0, 1.0, 1e-3])
        clouds.append(pc.astype(np.float32))
        ms.append(int(rng.integers(2, min(n, 3000))) | (i & 1))
        starts.append(int(rng.integers(0, n)))
    log(f"paired teams: sizes {[len(c) for c in clouds]} picks {ms}")
    paired = F.fps_sampling_batch(clouds, ms, starts)
    monkeypatch.setenv("FPSX_BATCH_NO_PAIR", "1")
    one = F.fps_sampling_batch(clouds, ms, starts)
    for i, (p, o, w) in enumerate(zip(paired, one, oracle_all(clouds, ms, starts))):
        np.testing.assert_array_equal(o, w, err_msg=f"one-pick cloud {i} n={len(clouds[i])}")
        np.testing.assert_array_equal(p, w, err_msg=f"paired cloud {i} n={len(clouds[i])}")


@pytest.mark.parametrize("shape", ["pp:1024:8", "pp:1024:16", "pp:1024:4"])
def test_batch_pruned_kernels(gpu, monkeypatch, shape):
    # The pruned paired batch kernels (spatial order per cloud, wave boxes,
    # two picks per team exchange, work-queue claims): clouds of 50..70 K
    # points form teams of 1..18 on these shapes; uniform, sphere, integer
    # ties, all points equal and special values; random starts.
    import fpsample_amd as F
    from fpsample_amd import batch as Bt

    monkeypatch.setenv("FPSX_BATCH_SHAPE", shape)
    rng = np.random.default_rng(41)
    clouds, ms, starts = [], [], []
    for i in range(30):
        n = int(np.exp(rng.uniform(np.log(50), np.log(70000))))
        kind = i % 5
        if kind == 0:
            pc = rng.random((n, 3))
        elif kind == 1:
            v = rng.standard_normal((n, 3))
            pc = v / np.linalg.norm(v, axis=1, keepdims=True)
        elif kind == 2:
            pc = rng.integers(0, 5, (n, 3))
        elif kind == 3:
            pc = np.tile(rng.random((1, 3)), (n, 1))
        else:
            pc = rng.random((n, 3)) * 4e19
            pc[rng.integers(0, n, 2)] = np.nan
        clouds.append(pc.astype(np.float32))
        ms.append(int(rng.integers(1, min(n, 2500) + 1)))
        starts.append(int(rng.integers(0, n)))
    log(f"{shape}: sizes {[len(c) for c in clouds]}")
    want = oracle_all(clouds, ms, starts)
    got = F.fps_sampling_batch(clouds, ms, starts)
    for i, (g, w) in enumerate(zip(got, want)):
        np.testing.assert_array_equal(g, w, err_msg=f"{shape} cloud {i} n={len(clouds[i])}")
    b = Bt.DeviceBatch(clouds, ms, starts, graph=True)
    assert b.info()["ppt"] == int(shape.split(":")[2])
    for rep in range(3):
        b.run()
        b.status()
        np.testing.assert_array_equal(b.picks_flat(), np.concatenate(want), err_msg=f"replay {rep}")
    b.close()

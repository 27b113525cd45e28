"""GPU parity of the serving loop: requests of mixed sizes submitted from
several threads, admitted by HBM budget and batched continuously, each equal
to the oracle's fps_sampling of its cloud alone."""
import threading

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def requests(seed, k, lo, hi):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        pc = rng.random((n, 3)).astype(np.float32) if i % 4 else rng.integers(0, 5, (n, 3)).astype(np.float32)
        m = max(1, int(n * rng.choice([1 / 8, 1 / 4, 1 / 2])))
        out.append((pc, m, int(rng.integers(0, n))))
    return out


def test_serve_mixed_from_threads(gpu):
    from fpsample_amd.serve import Server

    reqs = requests(7, 24, 64, 40000)
    results = {}
    with Server(hbm_budget_bytes=256 << 20, chunk_passes=128) as srv:
        def client(ids):
            for i in ids:
                pc, m, s = reqs[i]
                rid = srv.submit(pc, m, s)
                results[i] = srv.wait(rid)

        ts = [threading.Thread(target=client, args=(list(range(t, len(reqs), 4)),)) for t in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["error"] == 0
    for i, (pc, m, s) in enumerate(reqs):
        np.testing.assert_array_equal(results[i][0], O.fps(pc, m, s), err_msg=f"request {i}")


def test_serve_tight_budget_queues(gpu):
    from fpsample_amd.serve import BudgetError, Server

    reqs = requests(8, 10, 20000, 60000)
    # room for about one request at a time (at most 60 K points: 32 B per
    # point with the spatial order, 8 B per pick): the rest wait for memory
    with Server(hbm_budget_bytes=3 << 20, chunk_passes=64) as srv:
        ids = [srv.submit(pc, m, s) for pc, m, s in reqs]
        outs = [srv.wait(r)[0] for r in ids]
        with pytest.raises(BudgetError):
            srv.submit(np.zeros((300000, 3), np.float32), 10, 0)
    for (pc, m, s), got in zip(reqs, outs):
        np.testing.assert_array_equal(got, O.fps(pc, m, s))


def test_serve_same_alone_and_in_mix(gpu):
    from fpsample_amd.serve import Server

    reqs = requests(9, 12, 500, 30000)
    with Server(hbm_budget_bytes=512 << 20, chunk_passes=32) as srv:
        alone = []
        for pc, m, s in reqs:
            alone.append(srv.wait(srv.submit(pc, m, s))[0])
        ids = [srv.submit(pc, m, s) for pc, m, s in reqs]
        mixed = [srv.wait(r)[0] for r in ids]
    for a, b in zip(alone, mixed):
        np.testing.assert_array_equal(a, b)


def test_serve_short_chunks_paired_teams(gpu):
    # Chunks of 7 passes: paired rounds end on both sides of chunk boundaries
    # (a chunk ending on a pair saves its state with the pair's first pick
    # applied).  The 256:8 kernel (2048 points per workgroup; the default
    # 1024:16 one keeps the one-pick loop) runs 10k..60k-point requests in
    # teams of 5..30.
    from fpsample_amd.serve import Server

    reqs = requests(10, 8, 10000, 60000)
    with Server(hbm_budget_bytes=512 << 20, chunk_passes=7, block=256, ppt=8) as srv:
        ids = [srv.submit(pc, min(m, 700), s) for pc, m, s in reqs]
        outs = [srv.wait(r)[0] for r in ids]
    for i, ((pc, m, s), got) in enumerate(zip(reqs, outs)):
        np.testing.assert_array_equal(got, O.fps(pc, min(m, 700), s), err_msg=f"request {i}")


@pytest.mark.parametrize("plain", [False, True], ids=["pruned", "plain"])
def test_serve_default_shape_short_chunks(gpu, monkeypatch, plain):
    # The default 1024 x 16 shape: the pruned paired kernel (spatial order per
    # request, teams of ceil(n / 16384), one workgroup per member) and, with
    # FPSX_SERVE_PLAIN=1, the plain batch kernel.  Chunks of 5 passes end on
    # both sides of paired rounds; multi-start-free requests of 300 to 70 K
    # points, ties included.
    from fpsample_amd.serve import Server

    if plain:
        monkeypatch.setenv("FPSX_SERVE_PLAIN", "1")
    else:
        monkeypatch.delenv("FPSX_SERVE_PLAIN", raising=False)
    reqs = requests(11, 10, 300, 70000)
    with Server(hbm_budget_bytes=256 << 20, chunk_passes=5) as srv:
        ids = [srv.submit(pc, min(m, 400), s) for pc, m, s in reqs]
        outs = [srv.wait(r)[0] for r in ids]
    for i, ((pc, m, s), got) in enumerate(zip(reqs, outs)):
        np.testing.assert_array_equal(got, O.fps(pc, min(m, 400), s), err_msg=f"request {i} n={len(pc)}")


def test_serve_pruned_whole_requests_against_oracle(gpu):
    # Long chunks (whole requests in one launch) and full-length picks on
    # clustered and lattice clouds of one to five members.
    from fpsample_amd.serve import Server

    rng = np.random.default_rng(12)
    clouds = []
    for n in (5000, 16384, 16385, 40000, 81920):
        centers = rng.random((8, 3))
        clouds.append((centers[rng.integers(0, 8, n)] + 0.01 * rng.standard_normal((n, 3))).astype(np.float32))
    clouds.append(rng.integers(0, 12, (30000, 3)).astype(np.float32))
    ms = [len(c) // 4 for c in clouds]
    with Server(hbm_budget_bytes=256 << 20, chunk_passes=1 << 20) as srv:
        ids = [srv.submit(pc, m, 3) for pc, m in zip(clouds, ms)]
        outs = [srv.wait(r)[0] for r in ids]
    for i, (pc, m, got) in enumerate(zip(clouds, ms, outs)):
        np.testing.assert_array_equal(got, O.fps(pc, m, 3), err_msg=f"cloud {i} n={len(pc)}")

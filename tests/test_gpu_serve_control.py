"""Request control of the serving loop on the GPU: cancellation (queued and
running), deadlines (queued and running), waits with a time limit,
priorities that pause lower-priority requests between chunks, and eviction of
lower-priority requests to host memory when HBM is short.

What must hold: a request that is stopped ends with its own error and frees
its HBM range; every request that is not stopped -- including ones paused
and resumed, and ones placed in ranges freed by stopped requests -- returns
exactly the reference's picks (the oracle's fps_sampling, or the full-length
fixture written by the reference for the 2^20-point Gaussian cloud)."""
import time

import numpy as np
import pytest

import bench
from oracle import oracle as O
from tests.golden import full_length as FL

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def cloud(seed, n):
    return np.random.default_rng(seed).random((n, 3)).astype(np.float32)


def wait_until(pred, limit=30.0):
    t0 = time.monotonic()
    while not pred():
        assert time.monotonic() - t0 < limit, "condition not reached"
        time.sleep(0.001)


def test_cancel_queued_and_running(gpu):
    from fpsample_amd.serve import CancelledError, Server

    long_pc, other_pc, late_pc = cloud(1, 60000), cloud(2, 30000), cloud(3, 40000)
    # Room for the long request and one short one (c queues for memory); 32
    # passes per launch, so the long request runs for about a thousand launches.
    with Server(hbm_budget_bytes=3 << 20, chunk_passes=32) as srv:
        log("cancel: submitting")
        a = srv.submit(long_pc, 30000, 5)
        b = srv.submit(other_pc, 3000, 7)
        c = srv.submit(other_pc, 3000, 9)
        assert srv.cancel(c)                       # still queued behind a
        with pytest.raises(CancelledError):
            srv.wait(c)
        wait_until(lambda: srv.stats()["launches"] >= 1)
        assert srv.cancel(a)                       # running: stops at a chunk boundary
        with pytest.raises(CancelledError):
            srv.wait(a)
        got_b, _ = srv.wait(b)
        # A request placed after the stopped ones (reusing their ranges).
        d = srv.submit(late_pc, 4000, 11)
        got_d, _ = srv.wait(d)
        with pytest.raises(RuntimeError, match="unknown request id"):
            srv.cancel(d)                          # already waited for
        st = srv.stats()
    log(f"cancel: stats {st}")
    assert st["cancelled"] == 2 and st["error"] == 0 and st["hbm_in_use"] == 0
    np.testing.assert_array_equal(got_b, O.fps(other_pc, 3000, 7))
    np.testing.assert_array_equal(got_d, O.fps(late_pc, 4000, 11))


def test_cancel_after_finish_returns_picks(gpu):
    from fpsample_amd.serve import Server

    pc = cloud(4, 5000)
    with Server(hbm_budget_bytes=64 << 20, chunk_passes=256) as srv:
        r = srv.submit(pc, 500, 0)
        wait_until(lambda: srv.stats()["completed"] >= 1)
        assert not srv.cancel(r)                   # finished: nothing to stop
        got, _ = srv.wait(r)
    np.testing.assert_array_equal(got, O.fps(pc, 500, 0))


def test_deadlines_queued_and_running(gpu):
    from fpsample_amd.serve import DeadlineError, Server

    big, small = cloud(5, 200000), cloud(6, 20000)
    with Server(hbm_budget_bytes=12 << 20, chunk_passes=32) as srv:
        log("deadline: submitting")
        a = srv.submit(big, 100000, 0, deadline=0.02)   # ~3000 launches: cannot make it
        b = srv.submit(big, 1000, 3, deadline=0.005)    # queued behind a (budget): expires there
        c = srv.submit(small, 2000, 1, deadline=120.0)  # generous: finishes
        with pytest.raises(DeadlineError):
            srv.wait(a)
        with pytest.raises(DeadlineError):
            srv.wait(b)
        got_c, _ = srv.wait(c)
        st = srv.stats()
    log(f"deadline: stats {st}")
    assert st["expired"] == 2 and st["error"] == 0 and st["hbm_in_use"] == 0
    np.testing.assert_array_equal(got_c, O.fps(small, 2000, 1))


def test_wait_with_time_limit(gpu):
    from fpsample_amd.serve import Server

    pc = cloud(7, 100000)
    with Server(hbm_budget_bytes=64 << 20, chunk_passes=16) as srv:
        r = srv.submit(pc, 20000, 2)
        with pytest.raises(TimeoutError):
            srv.wait(r, timeout=0.001)
        got, lat = srv.wait(r, timeout=120.0)
    assert lat > 0.001
    np.testing.assert_array_equal(got, O.fps(pc, 20000, 2))


@pytest.mark.skipif(not FL.available("gauss1m"), reason="full-length fixture missing")
def test_priority_pauses_lower_requests(gpu):
    """Low-priority 2^20-point requests (teams of 64 workgroups) are submitted
    until some of them queue for want of grid room; a high-priority request
    submitted then is admitted ahead of the queue and runs at once, pausing
    low-priority work between chunks.  Every request still returns the
    reference's picks."""
    from fpsample_amd.serve import Server

    d = FL.load("gauss1m")
    pc = bench.gen_gauss_mixture(1 << 20)
    assert FL.sha256(pc) == d["cloud_sha256"]
    m = int(d["m"])
    hi_pc = cloud(8, 30000)
    # Two passes per launch: a 2^20-point request then lasts thousands of
    # launches, much longer than submitting the next one takes.
    with Server(hbm_budget_bytes=2 << 30, chunk_passes=2) as srv:
        log("priority: submitting low-priority requests until two queue")
        lows = []
        while True:
            assert len(lows) < 48, "the grid never filled"
            lows.append((time.monotonic(), srv.submit(pc, m, 0, priority=0)))
            time.sleep(0.005)  # the scheduler admits what fits; the rest stays queued
            if srv.stats()["pending"] >= 2:
                break
        log(f"priority: {len(lows)} low-priority requests submitted")
        t_hi = time.monotonic()
        hi = srv.submit(hi_pc, 3000, 4, priority=5)
        got_hi, lat_hi = srv.wait(hi)
        done_hi = t_hi + lat_hi
        outs = [(t, srv.wait(r)) for t, r in lows]
        st = srv.stats()
    log(f"priority: high latency {lat_hi:.4f} s, stats {st}")
    np.testing.assert_array_equal(got_hi, O.fps(hi_pc, 3000, 4))
    for i, (_, (got, _)) in enumerate(outs):
        np.testing.assert_array_equal(got, d["out"], err_msg=f"low-priority request {i}")
    done_low = sorted(t + lat for t, (_, lat) in outs)
    assert st["paused"] > 0
    # The high-priority request did not wait for the queue of low ones.
    assert done_hi < done_low[-1]


@pytest.mark.skipif(not FL.available("gauss1m"), reason="full-length fixture missing")
def test_eviction_under_hbm_pressure(gpu):
    """A budget that holds the 2^20-point low-priority request or the
    high-priority one, not both: the high one evicts the low one at a chunk
    boundary (its whole range copied to host memory), runs, and the low one is
    copied back and resumes.  Both return the reference's picks."""
    from fpsample_amd.serve import Server

    d = FL.load("gauss1m")
    pc = bench.gen_gauss_mixture(1 << 20)
    assert FL.sha256(pc) == d["cloud_sha256"]
    m = int(d["m"])
    hi_pc = cloud(9, 600000)
    # 48 MiB: the low request needs ~33.7 MB, the high one ~19.3 MB.
    with Server(hbm_budget_bytes=48 << 20, chunk_passes=8) as srv:
        t_low = time.monotonic()
        low = srv.submit(pc, m, 0, priority=0)
        wait_until(lambda: srv.stats()["launches"] >= 3)
        t_hi = time.monotonic()
        hi = srv.submit(hi_pc, 2000, 6, priority=3)
        got_hi, lat_hi = srv.wait(hi)
        got_low, lat_low = srv.wait(low)
        st = srv.stats()
    log(f"evict: high {lat_hi:.4f} s, low {lat_low:.4f} s, stats {st}")
    assert st["evicted"] >= 1 and st["error"] == 0 and st["hbm_in_use"] == 0
    assert t_hi + lat_hi < t_low + lat_low
    np.testing.assert_array_equal(got_low, d["out"])
    np.testing.assert_array_equal(got_hi, O.fps(hi_pc, 2000, 6))

"""Pins the CPU oracle (oracle/fps_oracle.c, oracle/oracle.py) -- CPU only.

Each oracle function is checked against something it does not share code with:
golden vectors written by the compiled reference, hand-derived known answers,
the definition of FPS evaluated by brute force, the exactness of bucket
pruning, and (where oracle/_ref is built) the reference itself on random
inputs.  A planted mistake (the opposite tie rule) must fail the known answers.
"""
import numpy as np
import pytest

from oracle import oracle as O


def test_golden_fps(golden):
    n = 0
    for name, g in golden.items():
        if g["kind"] != "fps":
            continue
        out = O.fps(g["pc"], g["m"], g["start"])
        np.testing.assert_array_equal(out, g["out"], err_msg=name)
        n += 1
    assert n >= 15


def test_golden_kdline(golden):
    n = 0
    for name, g in golden.items():
        if g["kind"] != "kdline":
            continue
        s = int(g["start"][0])
        np.testing.assert_array_equal(O.kdline_quickfps(g["pc"], g["m"], g["h"], s), g["out"],
                                      err_msg=name)
        np.testing.assert_array_equal(O.kdline_exhaustive(g["pc"], g["m"], g["h"], s), g["out"],
                                      err_msg=name)
        n += 1
    assert n >= 15


# Known answers derived by hand from src/lib.cpp:211-238 (`>=` argmax).
KATS = [
    (np.arange(5, dtype=np.float32)[:, None], 5, 0, [0, 4, 2, 3, 1]),
    (np.array([[0], [0], [1], [1]], np.float32), 4, 0, [0, 3, 3, 3]),
    (np.zeros((6, 3), np.float32), 6, 2, [2, 5, 5, 5, 5, 5]),
    (np.array([[0, 0], [3, 4], [0, 4], [3, 0]], np.float32), 4, 0, [0, 1, 3, 2]),
]


@pytest.mark.parametrize("pc,m,s,want", KATS)
def test_known_answers(pc, m, s, want):
    assert O.fps(pc, m, s).tolist() == want
    assert O.fps_numpy(pc, m, s).tolist() == want


def test_planted_tie_mistake_is_caught():
    # The lowest-index rule is a plausible mistake; the known answers reject it.
    pc, m, s, want = KATS[1]
    assert O.fps_ex(pc, m, s, lowest=True).tolist() != want
    pc, m, s, want = KATS[0]
    assert O.fps_ex(pc, m, s, lowest=True).tolist() != want


def _brute_force(pc, m, starts, lowest=False):
    """FPS by definition: pick = argmax_i min_{p in picked} d(i, p), float32 d."""
    pc = np.asarray(pc, np.float32)
    n, d = pc.shape
    picked = [int(s) for s in starts[:1]]
    rest = [int(s) for s in starts[1:]]
    while len(picked) < m:
        if rest:
            picked.append(rest.pop(0))
            continue
        best, bi = None, None
        for i in range(n):
            md = np.float32(np.inf)
            for p in picked:
                acc = np.float32(0)
                for j in range(d):
                    t = np.float32(pc[i, j] - pc[p, j])
                    acc = np.float32(acc + np.float32(t * t))
                if acc < md:
                    md = acc
            if best is None or md > best or (md == best and not lowest):
                best, bi = md, i
        picked.append(bi)
    return picked


@pytest.mark.parametrize("seed", range(12))
def test_brute_force_tiny(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 9))
    d = int(rng.integers(1, 4))
    # small integer grid: many exact ties
    pc = rng.integers(0, 3, (n, d)).astype(np.float32)
    m = int(rng.integers(1, n + 1))
    k = int(rng.integers(1, m + 1))
    starts = rng.integers(0, n, k).tolist()
    assert O.fps(pc, m, starts).tolist() == _brute_force(pc, m, starts)
    assert O.fps_ex(pc, m, starts, lowest=True).tolist() == _brute_force(pc, m, starts, True)


def test_full_length_uniform50k():
    # Every pick of the reference's own 50k bench case (bench/test_bench.py:89-93,
    # N=50000, M=4096, start 0), written by the compiled reference.
    import bench
    from tests.golden import full_length as FL

    d = FL.load("uniform50k")
    pc = bench.gen_uniform_n(50_000)
    assert FL.sha256(pc) == d["cloud_sha256"]
    np.testing.assert_array_equal(O.fps(pc, int(d["m"]), 0), d["out"])


def test_full_length_kdline50k():
    # Every pick of the reference's bucket_fps_kdline (src/wrapper.hpp:118-132) on
    # its own 50k h=7 bench case (bench/test_bench.py:131-135), start position 0.
    import bench
    from tests.golden import full_length as FL

    d = FL.load("kdline50k")
    pc = bench.gen_uniform_n(50_000)
    assert FL.sha256(pc) == d["cloud_sha256"]
    np.testing.assert_array_equal(O.kdline_quickfps(pc, int(d["m"]), 7, 0), d["out"])
    np.testing.assert_array_equal(O.kdline_exhaustive(pc, int(d["m"]), 7, 0), d["out"])


def test_numpy_and_c_agree_on_ties():
    rng = np.random.default_rng(7)
    pc = rng.integers(0, 4, (400, 3)).astype(np.float32)
    for s in (0, 399, [3, 5, 7]):
        np.testing.assert_array_equal(O.fps(pc, 200, s), O.fps_numpy(pc, 200, s))
        np.testing.assert_array_equal(O.fps_ex(pc, 200, s, lowest=True),
                                      O.fps_numpy(pc, 200, s, lowest=True))


def test_resume_from_prefix():
    # Forced picks replay a prefix: the continuation equals the uninterrupted run.
    rng = np.random.default_rng(3)
    pc = rng.random((3000, 3)).astype(np.float32)
    full = O.fps(pc, 500, 11)
    for k in (1, 2, 37, 499, 500):
        np.testing.assert_array_equal(O.fps(pc, 500, full[:k]), full)


def test_picked_distances_non_increasing():
    rng = np.random.default_rng(4)
    pc = rng.standard_normal((2000, 4)).astype(np.float32)
    out, dist = O.fps_ex(pc, 400, 0, with_dist=True)
    assert np.all(np.diff(dist[1:]) <= 0)
    assert len(set(out.tolist())) == 400


def test_kdline_perm_invariants():
    rng = np.random.default_rng(6)
    pc = rng.random((1000, 3)).astype(np.float32)
    for h in (1, 3, 6):
        perm, lo, hi = O.kdline_perm(pc, h)
        assert sorted(perm.tolist()) == list(range(1000))
        assert len(lo) == 2 ** h
        assert lo[0] == 0 and hi[-1] == 1000 and np.all(lo[1:] == hi[:-1])


@pytest.mark.parametrize("h", [1, 2, 4, 6])
def test_kdline_pruning_is_exact(h):
    # Lazy bucket updates == exhaustive search over the tree order (ties included).
    rng = np.random.default_rng(h)
    pcs = [rng.integers(0, 5, (1500, 3)).astype(np.float32),
           rng.random((1200, 2)).astype(np.float32),
           np.repeat(rng.integers(0, 3, (200, 3)), 5, axis=0).astype(np.float32)]
    for pc in pcs:
        for s in (0, 5, len(pc) - 1):
            np.testing.assert_array_equal(O.kdline_quickfps(pc, 600, h, s),
                                          O.kdline_exhaustive(pc, 600, h, s))


def test_kdline_tie_free_equals_vanilla_tail():
    # On tie-free data only the start and the initial distance differ.
    rng = np.random.default_rng(8)
    pc = rng.random((2000, 3)).astype(np.float32)
    kd = O.kdline_quickfps(pc, 300, 5, 9)
    van = O.fps(pc, 300, int(kd[0]))
    np.testing.assert_array_equal(kd, van)


@pytest.mark.parametrize("seed", range(6))
def test_against_reference_random(reference, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(50, 3000))
    d = int(rng.integers(1, 9))
    pc = (rng.integers(0, 4, (n, d)) if seed % 2 else rng.standard_normal((n, d)))
    pc = pc.astype(np.float32)
    m = int(rng.integers(1, min(n, 400) + 1))
    s = int(rng.integers(0, n))
    np.testing.assert_array_equal(O.fps(pc, m, s), reference.fps_sampling(pc, m, s))
    h = int(rng.integers(1, max(2, int(np.log2(n)))))
    np.testing.assert_array_equal(O.kdline_quickfps(pc, m, h, s),
                                  reference.bucket_fps_kdline(pc, m, h, s))
    starts = rng.integers(0, n, int(rng.integers(1, m + 1))).tolist()
    np.testing.assert_array_equal(O.fps(pc, m, starts), reference.fps_sampling(pc, m, starts))


def test_rng_draw_is_numpy_legacy():
    # start_idx=None draws np.random.randint(0, n) from the global legacy RNG
    # (src/fpsample/__init__.py:22); the legacy MT19937 stream is frozen across
    # numpy versions, so the value after seed(42) is fixed.
    np.random.seed(42)
    assert np.random.randint(low=0, high=4096) == 3174


# --- windowed-update FPS (NPDU), src/lib.cpp:272-340 -------------------------

def test_golden_npdu(golden):
    n = 0
    for name, g in golden.items():
        if g["kind"] != "npdu":
            continue
        out = O.npdu(g["pc"], g["m"], g["h"], int(g["start"][0]))
        np.testing.assert_array_equal(out, g["out"], err_msg=name)
        n += 1
    assert n >= 6


# Known answers derived by hand from src/lib.cpp:294-316 (1-D clouds).
#  * w = 0: each pick updates only itself, so pick 2 is index 1 (d=100 kept),
#    where full-update FPS would pick index 2.
#  * w = 2 at the right edge: the window of pick 3 is shifted to [1, 3]
#    (src/lib.cpp:300), which lowers index 1 to 1 and leaves index 2 (d=9) the
#    farthest; a plain clamp to [2, 3] would pick index 1.
#  * the mirror image at the left edge ([0, 2] rather than [0, 1]).
NPDU_KATS = [
    (np.array([[0], [10], [2], [11]], np.float32), 4, 0, 0, [0, 3, 1, 2]),
    (np.array([[0], [19], [3], [20]], np.float32), 3, 2, 0, [0, 3, 2]),
    (np.array([[20], [3], [19], [0]], np.float32), 3, 2, 3, [3, 0, 1]),
]


@pytest.mark.parametrize("pc,m,w,s,want", NPDU_KATS)
def test_npdu_known_answers(pc, m, w, s, want):
    assert O.npdu(pc, m, w, s).tolist() == want


def test_npdu_known_answers_against_reference(reference):
    for pc, m, w, s, want in NPDU_KATS:
        assert reference.fps_npdu(pc, m, w, s).tolist() == want


def test_npdu_planted_mistakes_are_caught():
    # Full updates (vanilla FPS with lowest-index ties) and the highest-index
    # tie rule are the plausible mistakes; the known answers reject both.
    pc, m, w, s, want = NPDU_KATS[0]
    assert O.fps_ex(pc, m, s, lowest=True).tolist() != want
    pc = np.array([[0], [1], [1]], np.float32)
    assert O.npdu(pc, 2, 0, 0).tolist() == [0, 1]
    assert O.fps_ex(pc, 2, 0).tolist() == [0, 2]


def test_npdu_full_window_is_lowest_tie_fps():
    # A window covering the cloud makes every update global: vanilla FPS with
    # the argmax's strict `>` (lowest index among equal maxima).
    rng = np.random.default_rng(5)
    for pc in (rng.random((700, 3)).astype(np.float32),
               rng.integers(0, 4, (500, 2)).astype(np.float32)):
        n = len(pc)
        for w in (2 * n, 2 * n + 1, 10 ** 9):
            np.testing.assert_array_equal(O.npdu(pc, 200, w, 7), O.fps_ex(pc, 200, 7, lowest=True))


def test_npdu_against_reference_random(reference):
    rng = np.random.default_rng(6)
    for n, d, w in ((300, 3, 0), (1000, 3, 17), (800, 5, 64), (257, 1, 256), (600, 2, 5000)):
        pc = rng.standard_normal((n, d)).astype(np.float32)
        s = int(rng.integers(0, n))
        np.testing.assert_array_equal(O.npdu(pc, n // 2, w, s), reference.fps_npdu(pc, n // 2, w, s))


# --- double-precision yardstick (accuracy report) -----------------------------

def test_fps_f64_equals_f32_on_exact_lattices():
    # Integer coordinates below 2^7: every difference, square and sum is exact
    # in float and in double, so the double loop must reproduce the float
    # oracle (itself pinned by the golden vectors) pick for pick.
    rng = np.random.default_rng(21)
    for d in (1, 2, 3, 5):
        pc = rng.integers(0, 100, (3000, d)).astype(np.float32)
        got, dist = O.fps_f64(pc, 700, 11)
        np.testing.assert_array_equal(got, O.fps(pc, 700, 11))
        assert np.isinf(dist[0]) and np.all(np.diff(dist[1:]) <= 0)


def test_fps_f64_brute_force_tiny():
    rng = np.random.default_rng(22)
    for trial in range(20):
        n = int(rng.integers(2, 40))
        pc = rng.standard_normal((n, 3)).astype(np.float32)
        m = int(rng.integers(1, n + 1))
        x = pc.astype(np.float64)
        picks = [0]
        for _ in range(1, m):
            dmin = np.min(((x[:, None, :] - x[None, picks, :]) ** 2).sum(-1), axis=1)
            picks.append(int(np.flatnonzero(dmin == dmin.max())[-1]))
        np.testing.assert_array_equal(O.fps_f64(pc, m, 0)[0], picks)


def test_coverage_radius_definition_and_consistency():
    rng = np.random.default_rng(23)
    pc = rng.standard_normal((500, 3)).astype(np.float32)
    picks = rng.choice(500, 37, replace=False).astype(np.uint64)
    x = pc.astype(np.float64)
    want = np.sqrt(((x[:, None, :] - x[None, picks.astype(np.int64), :]) ** 2).sum(-1).min(axis=1).max())
    assert abs(O.coverage_radius(pc, picks) - want) <= 1e-12 * want
    # The greedy run's next pick distance is the radius of its prefix.
    ref, dist = O.fps_f64(pc, 60, 4)
    for k in (1, 7, 59):
        assert abs(O.coverage_radius(pc, ref[:k]) - np.sqrt(dist[k])) <= 1e-12
    acc = O.accuracy_vs_f64(pc, O.fps(pc, 59, 4), 4)
    assert acc["index_agreement"] == 1.0 and acc["coverage_radius_rel_err"] == 0.0


def test_reference_bench_workloads_have_full_length_fixtures():
    # Every bench workload taken from the reference's own bench (those that
    # carry its published time) has a full-length fixture of the reference's
    # picks, made from the same generator, so its bench line reports
    # identical_to_reference; the window of an NPDU case is the wrapper's
    # default int(N / M * 16) (src/fpsample/__init__.py:97).
    import bench
    from tests.golden import full_length as FL

    names = [k for k, wl in bench.WORKLOADS.items() if "published" in wl]
    assert len(names) == 9
    for name in names:
        wl = bench.WORKLOADS[name]
        d = FL.load(name)
        pc = wl["gen"]()
        assert FL.sha256(pc) == d["cloud_sha256"], name
        assert int(d["m"]) == wl["m"] and int(d["n"]) == pc.shape[0], name
        assert len(d["out"]) == wl["m"], name
        if wl.get("kind") == "npdu":
            assert f"w={int(pc.shape[0] / wl['m'] * 16)}," in wl["desc"], name

"""The drop-in boundary without a GPU: exports, validation, start resolution.

Everything here returns or raises before any device work, so it runs on the
CPU.  Reference behaviour followed: src/fpsample/__init__.py:19-65, 174-206
(asserts, dtype/layout coercion, RNG draw), src/lib.cpp:52-109, 249-270,
522-579 (pybind11 validation and exception types), src/wrapper.hpp:118-132
(C return codes).
"""
import ctypes
import os

import numpy as np
import pytest

import fpsample_amd as F
from oracle import oracle as O
from tests.capi import CApi, header_functions

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_function(capi):
    names = header_functions()
    assert {"fpsx_fps_sampling", "bucket_fps_kdline", "fpsx_fps_sampling_device"} <= set(names)
    for n in names:
        assert hasattr(capi.L, n), n


def test_version(capi):
    assert capi.L.fpsx_version().decode() == F.__version__


def _compile_abi_check(*extra):
    import subprocess

    src = os.path.join(ROOT, "tests", "abi_signature_check.cpp")
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), *extra, src]
    return subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def test_bucket_signature_matches_reference_abi():
    # include/fpsx.h's declaration has the type of src/wrapper.hpp:117-119 (a
    # static_assert on the function pointer type, checked by the compiler).
    p = _compile_abi_check()
    assert p.returncode == 0, p.stdout


def test_bucket_signature_agrees_with_reference_header():
    # With the reference's own header in the same translation unit, a
    # different C signature would be a conflicting declaration.
    ref_src = "/root/reference/src"
    if not os.path.exists(os.path.join(ref_src, "wrapper.hpp")):
        pytest.skip("reference sources not present")
    p = _compile_abi_check("-DWITH_REFERENCE_HEADER", "-I", ref_src, "-w")
    assert p.returncode == 0, p.stdout


def test_abi_check_catches_a_changed_signature(tmp_path):
    # The check itself is live: a header whose declaration differs fails it.
    hdr = open(os.path.join(ROOT, "include", "fpsx.h")).read()
    bad = hdr.replace("size_t start_idx, size_t height, size_t *sampled_point_indices);",
                      "size_t start_idx, int height, size_t *sampled_point_indices);")
    assert bad != hdr
    (tmp_path / "fpsx.h").write_text(bad)
    import subprocess

    src = os.path.join(ROOT, "tests", "abi_signature_check.cpp")
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", str(tmp_path), src],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode != 0


def test_capi_validation_codes(capi):
    pc = np.zeros((10, 3), np.float32)
    assert capi.fps(pc, 3, 0, dim=0)[0] == 1            # no columns
    assert capi.fps(pc, 3, 10)[0] == 2                  # start >= n
    assert capi.fps(pc, 3, [0, 11])[0] == 2
    assert capi.fps(pc, 0, 0)[0] == 3                   # no samples
    assert capi.fps(pc, 11, 0)[0] == 3                  # more samples than points
    assert capi.fps(pc, 2, [0, 1, 2])[0] == 3           # more starts than samples
    assert capi.fps(pc, 3, 0, layout=7)[0] == 3
    assert capi.kdline(np.zeros((10, 9), np.float32), 3, 0, 2)[0] == 1
    assert capi.kdline(pc, 3, 0, 2, dim=0)[0] == 1
    assert capi.kdline(pc, 3, 10, 2)[0] == 2
    assert capi.kdline(pc, 0, 0, 2)[0] == 3
    assert capi.last_error() != ""


@pytest.mark.parametrize("d,h", [(1, 3), (2, 5), (3, 1), (3, 7), (3, 9), (8, 4)])
def test_kdline_permutation_matches_oracle(capi, d, h):
    rng = np.random.default_rng(d * 10 + h)
    for pc in (rng.random((3000, d)), rng.integers(0, 4, (2000, d)), np.ones((300, d))):
        pc = pc.astype(np.float32)
        a = capi.kdline_perm(pc, h)
        b = O.kdline_perm(pc, h)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_kdline_permutation_golden_first_pick(capi, golden):
    # The reference's first pick is the point at position start_idx of the tree order.
    for name, g in golden.items():
        if g["kind"] != "kdline":
            continue
        perm, _, _ = capi.kdline_perm(g["pc"], g["h"])
        assert perm[int(g["start"][0])] == g["out"][0], name


# ---------------------------------------------------------------- Python layer
def test_asserts_match_reference():
    pc = np.random.rand(10, 3)
    cases = [
        (lambda: F.fps_sampling(pc, 0), "n_samples should be >= 1"),
        (lambda: F.fps_sampling(pc, 11), "n_pts should be >= n_samples"),
        (lambda: F.fps_sampling(pc, 3, start_idx=10), "start_idx should be None or 0 <= start_idx < n_pts"),
        (lambda: F.fps_sampling(pc, 3, start_idx=-1), "start_idx should be None or 0 <= start_idx < n_pts"),
        (lambda: F.fps_sampling(pc, 2, start_idx=[0, 1, 2]), "len(start_idx) should be <= n_samples"),
        (lambda: F.fps_sampling(pc, 3, start_idx=[0, 10]), "start_idx should be None or 0 <= start_idx < n_pts"),
        (lambda: F.bucket_fps_kdline_sampling(pc, 3, 0), "h should be >= 1"),
        (lambda: F.bucket_fps_kdline_sampling(pc, 3, 4), "2**h should be <= n_pts"),
        (lambda: F.bucket_fps_kdline_sampling(pc, 3, 2, start_idx=10), "start_idx should be None or 0 <= start_idx < n_pts"),
    ]
    for fn, msg in cases:
        with pytest.raises(AssertionError, match=msg.replace("*", r"\*").replace("(", r"\(").replace(")", r"\)")):
            fn()
    with pytest.raises(AssertionError):
        F.fps_sampling(np.random.rand(10), 3)


def test_start_idx_types():
    assert F.get_start_idx(10, 3) == 3
    a = F.get_start_idx(10, [1, 2])
    assert isinstance(a, np.ndarray) and a.dtype == np.uint64 and a.tolist() == [1, 2]
    with pytest.raises(ValueError, match="start_idx should be None, int or list"):
        F.get_start_idx(10, (1, 2))
    with pytest.raises(ValueError, match="start_idx should be None, int or list"):
        F.fps_sampling(np.random.rand(10, 3), 3, start_idx=np.int64(2))
    np.random.seed(42)
    want = np.random.randint(low=0, high=4096)
    np.random.seed(42)
    assert F.get_start_idx(4096, None) == want


def test_wrapper_hands_fortran_float32_and_one_rng_draw(monkeypatch):
    seen = {}

    def fake(pc, n, start):
        seen.update(pc=pc, n=n, start=start)
        return np.zeros(n, np.uint64)

    monkeypatch.setattr(F, "_fps_sampling", fake)
    pc = np.random.rand(50, 3)
    np.random.seed(7)
    F.fps_sampling(pc, 5)
    np.random.seed(7)
    assert seen["start"] == np.random.randint(0, 50)
    assert seen["pc"].dtype == np.float32 and seen["pc"].flags.f_contiguous
    np.testing.assert_array_equal(seen["pc"], pc.astype(np.float32))


def test_pybind_validation_before_device():
    from fpsample_amd import _fpsample as X

    pc = np.asfortranarray(np.random.rand(10, 3).astype(np.float32))
    with pytest.raises(ValueError, match="points must be a 2D array, but got shape 1"):
        X._fps_sampling(np.zeros(10, np.float32), 3, 0)
    with pytest.raises(ValueError, match="points must have at least one column"):
        X._fps_sampling(np.zeros((10, 0), np.float32), 3, 0)
    with pytest.raises(ValueError, match="n_samples must be less than the number of points: n_samples=11, P=10"):
        X._fps_sampling(pc, 11, 0)
    with pytest.raises(ValueError, match="start_idx must be less than the number of points: start_idx=10, P=10"):
        X._fps_sampling(pc, 3, 10)
    with pytest.raises(ValueError, match="The number of start indices must be less than or equal to n_samples: 3, n_samples=2"):
        X._fps_sampling(pc, 2, np.array([0, 1, 2], np.uint64))
    with pytest.raises(ValueError, match="All indices in start_idx must be less than the number of points: 12, P=10"):
        X._fps_sampling(pc, 3, np.array([0, 12], np.uint64))
    with pytest.raises(TypeError, match="start_idx must be int or 1D numpy array of size_t"):
        X._fps_sampling(pc, 3, 1.5)
    with pytest.raises(TypeError, match="start_idx must be int or 1D numpy array of size_t"):
        X._fps_sampling(pc, 3, np.array([0, 1], np.int64))
    with pytest.raises(ValueError, match="start_idx must not be an empty array"):
        X._fps_sampling(pc, 3, np.array([], np.uint64))
    assert X._fps_sampling(pc, 0, 0).shape == (0,)
    with pytest.raises(NotImplementedError, match="Array of start indices not implemented yet"):
        X._bucket_fps_kdline_sampling(pc, 3, 2, np.array([0], np.uint64))
    with pytest.raises(ValueError, match="points must be a 2D float32 array"):
        X._bucket_fps_kdline_sampling(np.zeros(10, np.float32), 3, 2, 0)
    with pytest.raises(ValueError, match="start_idx out of range"):
        X._bucket_fps_kdline_sampling(pc, 3, 2, 10)
    with pytest.raises(ValueError, match=r"n_samples must be in \[1, num_points\]"):
        X._bucket_fps_kdline_sampling(pc, 0, 2, 0)
    with pytest.raises(ValueError, match="height must be >= 1"):
        X._bucket_fps_kdline_sampling(pc, 3, 0, 0)


def test_pybind_validation_matches_reference(reference):
    """Same exception type and message as the compiled reference module."""
    from fpsample_amd import _fpsample as X

    pc = np.asfortranarray(np.random.rand(10, 3).astype(np.float32))
    bad = [
        (np.zeros(10, np.float32), 3, 0),
        (np.zeros((10, 0), np.float32), 3, 0),
        (pc, 11, 0),
        (pc, 3, 10),
        (pc, 2, np.array([0, 1, 2], np.uint64)),
        (pc, 3, np.array([0, 12], np.uint64)),
        (pc, 3, 1.5),
    ]
    for args in bad:
        with pytest.raises(Exception) as ours:
            X._fps_sampling(*args)
        with pytest.raises(Exception) as theirs:
            reference.module._fps_sampling(*args)
        assert type(ours.value) is type(theirs.value)
        assert str(ours.value) == str(theirs.value)


def test_npdu_validation_matches_reference(reference):
    """_fps_npdu_sampling: same exception type and message as the reference."""
    from fpsample_amd import _fpsample as X

    pc = np.random.rand(10, 3).astype(np.float32)
    bad = [
        (np.zeros(10, np.float32), 3, 4, 0),
        (np.zeros((10, 0), np.float32), 3, 4, 0),
        (pc, 11, 4, 0),
        (pc, 3, 4, 10),
        (pc, 3, 4, np.array([0, 1], np.uint64)),
        (pc, 2, 4, np.array([0, 1, 2], np.uint64)),
        (pc, 3, 4, 1.5),
    ]
    for args in bad:
        with pytest.raises(Exception) as ours:
            X._fps_npdu_sampling(*args)
        with pytest.raises(Exception) as theirs:
            reference.module._fps_npdu_sampling(*args)
        assert type(ours.value) is type(theirs.value), args
        assert str(ours.value) == str(theirs.value)
    assert X._fps_npdu_sampling(pc, 0, 4, 0).shape == (0,)


def test_npdu_wrapper_window_default_and_asserts(monkeypatch):
    seen = {}

    def fake(pc, n, w, start):
        seen.update(pc=pc, n=n, w=w, start=start)
        return np.zeros(n, np.uint64)

    monkeypatch.setattr(F, "_fps_npdu_sampling", fake)
    pc = np.random.rand(1000, 3)
    np.random.seed(9)
    F.fps_npdu_sampling(pc, 100)
    np.random.seed(9)
    assert seen["start"] == np.random.randint(0, 1000)
    assert seen["w"] == int(1000 / 100 * 16)
    assert seen["pc"].dtype == np.float32 and seen["pc"].flags.c_contiguous
    with pytest.warns(UserWarning, match="k is too large, set to 999"):
        F.fps_npdu_sampling(pc, 5, start_idx=0)
    assert seen["w"] == 999
    F.fps_npdu_sampling(pc, 5, w=3, start_idx=2)
    assert (seen["w"], seen["start"]) == (3, 2)
    with pytest.raises(AssertionError, match="n_pts should be >= n_samples"):
        F.fps_npdu_sampling(pc, 1001)
    with pytest.raises(AssertionError, match="start_idx should be None or 0 <= start_idx < n_pts"):
        F.fps_npdu_sampling(pc, 3, start_idx=1000)
    with pytest.raises(TypeError):
        F.fps_npdu_sampling(pc, 3, start_idx=[0, 1])   # `0 <= list` in the reference's assert


def test_kdline_permutation_special_values_fuzz(capi):
    """The permutation (columns, branch-free partition, threaded subtrees)
    against the oracle's restatement of src/_ext/KDTreeBase.h:84-207 on
    clouds with NaN, +-inf, -0, duplicates and all-NaN columns."""
    from oracle import oracle as O

    rng = np.random.default_rng(3)
    for trial in range(120):
        n = int(rng.integers(1, 3000))
        d = int(rng.integers(1, 9))
        q = (rng.integers(0, 4, (n, d)) if trial % 3 == 0 else rng.standard_normal((n, d))).astype(np.float32)
        if trial % 2 == 0 and n > 5:
            q[rng.integers(0, n, 3), rng.integers(0, d, 3)] = np.nan
        if trial % 5 == 0 and n > 5:
            q[rng.integers(0, n, 2)] = np.inf
            q[rng.integers(0, n, 2), 0] = -0.0
        if trial % 14 == 0:
            q[:, 0] = np.nan
        h = int(rng.integers(1, 14))
        p1, l1, h1 = capi.kdline_perm(q, h)
        p2, l2, h2 = O.kdline_perm(q, h)
        np.testing.assert_array_equal(p1, p2, err_msg=f"trial {trial} n={n} d={d} h={h}")
        np.testing.assert_array_equal(l1, l2)
        np.testing.assert_array_equal(h1, h2)
    big = np.random.default_rng(9).standard_normal((300_000, 3)).astype(np.float32)
    big[::1000] = np.nan
    for h in (5, 11):  # large enough for the threaded top levels
        np.testing.assert_array_equal(capi.kdline_perm(big, h)[0], O.kdline_perm(big, h)[0])


# ---------------------------------------------------------------- device-tensor entry
def test_tensor_entry_refuses_host_tensors_before_any_launch():
    # A host address handed to the kernels would be a GPU memory fault; the
    # tensor entry must raise first, without touching the device.
    from fpsample_amd import device as D

    torch = pytest.importorskip("torch")
    with pytest.raises(TypeError, match="GPU tensor"):
        D.fps_sampling_tensor(torch.zeros((16, 3), dtype=torch.float32), 4)
    # numpy exports DLPack on the CPU device: refused by device type
    with pytest.raises(TypeError, match="ROCm device memory"):
        D.fps_sampling_tensor(np.zeros((16, 3), np.float32), 4)


def test_tensor_entry_checks_rank_before_strides():
    from fpsample_amd import device as D

    class OneD:
        __cuda_array_interface__ = {"shape": (16,), "typestr": "<f4", "data": (4096, False),
                                    "strides": None, "version": 3}

    class Wrong:
        __cuda_array_interface__ = {"shape": (16, 3), "typestr": "<f8", "data": (4096, False),
                                    "strides": None, "version": 3}

    with pytest.raises(ValueError, match="2-D"):
        D.fps_sampling_tensor(OneD(), 4)
    with pytest.raises(TypeError, match="float32"):
        D.fps_sampling_tensor(Wrong(), 4)
    with pytest.raises(TypeError, match="expected a GPU tensor"):
        D.fps_sampling_tensor([[0.0, 1.0, 2.0]], 1)

"""GPU test of the allocation-free entry (fpsx_fps_sampling_device_noalloc)
through the C ABI: its picks equal those of the entry it varies
(fpsx_fps_sampling_device_strided) bit for bit, on two calls in a row, the
second with a smaller input (the shared scratch keeps the first call's
contents); sizes beyond its limit are refused before anything is enqueued."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_sz, _vp = ctypes.c_size_t, ctypes.c_void_p


def bind(L):
    f = L.fpsx_fps_sampling_device_noalloc
    f.argtypes = [_vp, _sz, _sz, _sz, _sz, _sz, _vp, _sz, _vp, _vp, _vp, ctypes.c_int,
                  ctypes.c_float, _vp]
    f.restype = ctypes.c_int
    return f


def gauss(n, d, seed):
    rng = np.random.default_rng(seed)
    centers = rng.random((32, d))
    return (centers[rng.integers(0, 32, n)] + 0.02 * rng.standard_normal((n, d))).astype(np.float32)


def run_both(D, noalloc, pc, m, starts, soa):
    n, d = pc.shape
    if soa:
        arr = D.DeviceArray.from_numpy(np.ascontiguousarray(pc.T))
        sp, sd = 1, n
    else:
        arr = D.DeviceArray.from_numpy(np.ascontiguousarray(pc))
        sp, sd = d, 1
    st = np.asarray(starts, np.uint64)
    d_st = D.DeviceArray.from_numpy(st)
    out_a = D.DeviceArray((m,), np.uint64)
    pts_a = D.DeviceArray((m, d), np.float32)
    status = D.DeviceArray((1,), np.uint32)
    rc = noalloc(arr.ptr, sp, sd, n, d, m, d_st.ptr, st.size, out_a.ptr, pts_a.ptr, status.ptr,
                 0, float("inf"), None)
    assert rc == 0, D.lib().fpsx_last_error().decode()
    D.synchronize()
    assert status.numpy()[0] == 0
    out_b = D.DeviceArray((m,), np.uint64)
    pts_b = D.DeviceArray((m, d), np.float32)
    D.fps_sampling_device_strided(arr, sp, sd, n, d, m, d_st, st.size, out_b, pts_b)
    D.device_status()
    return out_a.numpy(), out_b.numpy(), pts_a.numpy(), pts_b.numpy()


def test_noalloc_equals_strided_entry_twice(gpu):
    from fpsample_amd import device as D

    noalloc = bind(D.lib())
    # first call: 300 K clustered 3-D points, SoA, several start indices
    a, b, pa, pb = run_both(D, noalloc, gauss(300_000, 3, 1), 3000, [5, 299_999], soa=True)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(pa, pb)
    # second call, smaller input, row-major tensor layout: nothing of the
    # first call's scratch may leak into it
    a, b, pa, pb = run_both(D, noalloc, gauss(40_000, 3, 2), 900, [123], soa=False)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(pa, pb)


def test_noalloc_other_dimension_and_one_workgroup(gpu):
    from fpsample_amd import device as D

    noalloc = bind(D.lib())
    for n, d, m in ((70_000, 5, 500), (3000, 3, 3000)):
        pc = np.random.default_rng(n).integers(0, 6, (n, d)).astype(np.float32)  # ties
        a, b, _, _ = run_both(D, noalloc, pc, m, [0], soa=True)
        np.testing.assert_array_equal(a, b, err_msg=f"n={n} d={d}")


def test_noalloc_refuses_beyond_the_register_file(gpu):
    from fpsample_amd import device as D

    noalloc = bind(D.lib())
    small = D.DeviceArray((16,), np.float32)  # never read: the call returns first
    st = D.DeviceArray.from_numpy(np.zeros(1, np.uint64))
    out = D.DeviceArray((4,), np.uint64)
    rc = noalloc(small.ptr, 1, 1 << 25, 1 << 25, 3, 4, st.ptr, 1, out.ptr, None, None, 0,
                 float("inf"), None)
    assert rc == 4  # FPSX_ETOOBIG
    assert "register-resident" in D.lib().fpsx_last_error().decode()

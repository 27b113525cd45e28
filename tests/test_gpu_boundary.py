"""GPU tests of the boundary around the hot path: device tensors used in place
(torch on the caller's stream, DLPack), kernel status reported through every
entry, and the autotuners' choices.

Status tests shorten the exchange wait bound from the host
(fpsx_exchange_timeout) so that a multi-workgroup kernel gives up at once and
reports FPSX_ETIMEOUT; every kernel still exits normally.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


@pytest.fixture
def torch_gpu(gpu):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    return torch


# ---------------------------------------------------------------- torch / DLPack in place
def test_torch_tensor_written_just_before(torch_gpu):
    # A torch op writes the cloud on torch's current stream immediately
    # before the call; the FPS must read the finished values.
    from fpsample_amd import device as D

    torch = torch_gpu
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand((200_000, 3), device="cuda", generator=g)
    for _ in range(4):
        x = torch.sin(x * 7.0) * 0.5 + x  # several kernels still in flight
    idx, pts = D.fps_sampling_tensor(x, 1500, start_idx=11, return_points=True)
    host = x.cpu().numpy()
    assert idx.dtype == torch.int64 and idx.device == x.device
    want = O.fps(host, 1500, 11)
    np.testing.assert_array_equal(idx.cpu().numpy().astype(np.uint64), want)
    np.testing.assert_array_equal(pts.cpu().numpy(), host[want.astype(np.int64)])


def test_torch_soa_view_and_caller_stream(torch_gpu):
    # A (3, N) tensor's transpose is an SoA (N, 3) view with strides (1, N);
    # the work runs on a side stream the caller passes.
    from fpsample_amd import device as D

    torch = torch_gpu
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        soa = torch.randn((3, 70_001), device="cuda") * 100.0
        view = soa.t()
        assert view.stride() == (1, 70_001)
        idx, _ = D.fps_sampling_tensor(view, 900, start_idx=[5, 70_000])
    s.synchronize()
    host = view.cpu().numpy()
    np.testing.assert_array_equal(idx.cpu().numpy().astype(np.uint64), O.fps(host, 900, [5, 70_000]))


class _DLPackOnly:
    """Exports a torch tensor through DLPack alone (not recognised as torch)."""

    def __init__(self, t):
        self.t = t

    def __dlpack__(self, stream=None):
        return self.t.__dlpack__(stream=stream)

    def __dlpack_device__(self):
        return self.t.__dlpack_device__()


def test_dlpack_export(torch_gpu):
    from fpsample_amd import device as D

    torch = torch_gpu
    x = torch.rand((30_000, 3), device="cuda")
    x.mul_(3.0)
    idx, pts = D.fps_sampling_tensor(_DLPackOnly(x), 700, start_idx=0, return_points=True)
    host = x.cpu().numpy()
    want = O.fps(host, 700, 0)
    np.testing.assert_array_equal(idx.numpy(), want)
    np.testing.assert_array_equal(pts.numpy(), host[want.astype(np.int64)])


def test_device_array_and_two_streams(gpu):
    # The same cloud sampled on the library's stream and on a stream of the
    # caller's, alternately: the library orders its shared scratch between them.
    import ctypes

    from fpsample_amd import device as D

    rng = np.random.default_rng(8)
    pc = rng.random((120_000, 3)).astype(np.float32)
    want = O.fps(pc, 1200, 0)
    h = D.hip()
    s = ctypes.c_void_p()
    h.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    assert h.hipStreamCreateWithFlags(ctypes.byref(s), 1) == 0
    try:
        arr = D.DeviceArray.from_numpy(pc)
        outs = []
        for k in range(4):
            stream = s.value if k % 2 else None
            idx, _ = D.fps_sampling_tensor(arr, 1200, 0, stream=stream, sync=False)
            outs.append(idx)
        D.synchronize()
        for k, o in enumerate(outs):
            np.testing.assert_array_equal(o.numpy(), want, err_msg=f"call {k}")
    finally:
        D.synchronize()
        h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        h.hipStreamDestroy(s)


# ---------------------------------------------------------------- status through every entry
@pytest.fixture
def short_exchange_wait(gpu):
    from fpsample_amd import device as D

    D.exchange_timeout(1e-9)
    yield
    D.exchange_timeout(0)


def test_tensor_entry_raises_on_kernel_timeout(gpu, monkeypatch):
    from fpsample_amd import device as D

    monkeypatch.setenv("FPSX_PLAN", "resident:1024:4")  # 10 workgroups exchanging
    pc = np.random.default_rng(1).random((40_000, 3)).astype(np.float32)
    arr = D.DeviceArray.from_numpy(pc)
    D.exchange_timeout(1e-9)
    try:
        with pytest.raises(RuntimeError, match="error code 7"):
            D.fps_sampling_tensor(arr, 500, 0)
    finally:
        D.exchange_timeout(0)
    idx, _ = D.fps_sampling_tensor(arr, 500, 0)
    np.testing.assert_array_equal(idx.numpy(), O.fps(pc, 500, 0))


def test_batch_fallback_reports_an_earlier_clouds_failure(gpu):
    # 5-D clouds have no batch kernel: one launch per cloud.  The first cloud
    # needs several workgroups (and times out under the short bound), the last
    # one a single workgroup (no exchange, succeeds): the batch must still
    # report the first cloud's failure.
    from fpsample_amd import device as D
    from fpsample_amd.batch import DeviceBatch

    rng = np.random.default_rng(4)
    clouds = [rng.random((50_000, 5)).astype(np.float32), rng.random((300, 5)).astype(np.float32)]
    ms = [400, 100]
    b = DeviceBatch(clouds, ms, [0, 0], graph=False)
    assert b.info()["block"] == 0  # one launch per cloud
    D.exchange_timeout(1e-9)
    try:
        b.run()
        with pytest.raises(RuntimeError, match="error code 7"):
            b.status()
    finally:
        D.exchange_timeout(0)
    b.run()
    b.status()
    for i, (g, pc, m) in enumerate(zip(b.picks(), clouds, ms)):
        np.testing.assert_array_equal(g, O.fps(pc, m, 0), err_msg=f"cloud {i}")
    b.close()


# ---------------------------------------------------------------- autotuners
def test_autotune_choice_is_used_and_exact(gpu):
    from fpsample_amd import device as D

    rng = np.random.default_rng(12)
    centers = rng.random((16, 3))
    pc = (centers[rng.integers(0, 16, 200_000)] + 0.03 * rng.standard_normal((200_000, 3))).astype(np.float32)
    n = pc.shape[0]
    soa = D.DeviceArray.from_numpy(np.ascontiguousarray(pc.T))
    before = D.plan_info(n, 3, n_samples=2048)
    tuned = D.autotune(soa, 1, n, n, 3, passes=2048)
    after = D.plan_info(n, 3, n_samples=2048)
    log(f"planner {before}, tuned {tuned}, plan after tuning {after}")
    assert (after["block"], after["ppt"], after["grid"]) == (tuned["block"], tuned["ppt"], tuned["grid"])
    # the tuned plan serves later calls of any length on clouds of this size
    assert D.plan_info(n, 3, n_samples=100) == after
    d_start = D.DeviceArray.from_numpy(np.array([7], np.uint64))
    for m in (2048, 333):
        out = D.DeviceArray((m,), np.uint64)
        D.fps_sampling_device(soa, n, 3, m, d_start, 1, out)
        D.device_status()
        np.testing.assert_array_equal(out.numpy(), O.fps(pc, m, 7), err_msg=f"m={m}")


def test_batch_autotune_exact(gpu):
    from fpsample_amd.batch import DeviceBatch

    rng = np.random.default_rng(13)
    clouds = [rng.random((int(n), 3)).astype(np.float32)
              for n in np.exp(rng.uniform(np.log(1000), np.log(40_000), 20))]
    ms = [len(c) // 4 for c in clouds]
    b = DeviceBatch(clouds, ms, [0] * len(clouds), graph=True, autotune=True)
    log(f"tuned batch: {b.info()}")
    b.run()
    b.status()
    for i, (g, pc, m) in enumerate(zip(b.picks(), clouds, ms)):
        np.testing.assert_array_equal(g, O.fps(pc, m, 0), err_msg=f"cloud {i}")
    b.close()

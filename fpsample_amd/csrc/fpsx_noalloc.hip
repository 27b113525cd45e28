// fpsx_noalloc.hip -- the allocation-free form of fpsx_fps_sampling_device_strided.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The only
// scratch the register-resident kernels need -- the cross-workgroup exchange
// area and the status word -- lives in `scratch`, a fixed-size array in the
// code object: every GPU the library runs on holds its own copy from module
// load until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol); it uses the
// per-device address from hipGetSymbolAddress, taken once per device on the
// entry's first call there.
//
// Plans: the register-resident kernels only (points and min-distances in
// VGPRs), chosen by the same rule as the main planner's resident path.  The
// spatially ordered (pruned, paired) and streaming plans need per-point
// scratch and are not available here; a cloud beyond the register file is
// refused with FPSX_ETOOBIG.  Calls that overlap on one GPU share the array
// and race; the caller serialises them (one stream, or its own ordering).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "fps_device.hpp"
#include "fps_kernels.hpp"
#include "fpsx.h"
#include "fpsx_internal.hpp"

#define FPSX_EXPORT extern "C" __attribute__((visibility("default")))

namespace fpsx {
namespace {

constexpr uint32_t kNaReplicas = kMaxReplicas;  // 8 exchange replicas
constexpr size_t kNaXchgWords = (size_t)2 * kNaReplicas * kSlotWords * kPageWords;  // 131072
constexpr size_t kNaScratchWords = kNaXchgWords + 32;  // + status word (256-byte aligned tail)
constexpr int kNaMaxDevices = 64;
constexpr int kNaMaxKernels = 128;

}  // namespace

// Exchange area (kNaXchgWords words) followed by the status word: 1 MiB + 256 B.
static __device__ __align__(256) uint64_t scratch[kNaScratchWords];

namespace {

struct NaDevice {
  bool ready = false;
  uint64_t *scratch = nullptr;  // this device's address of `scratch`
  int cus = 0;
  uint64_t spin_ticks = 0;
  int occ[kNaMaxKernels];       // co-resident workgroups per resident-table entry (-1: unknown)
};

std::mutex g_na_mu;
NaDevice g_na[kNaMaxDevices];

// Once per device, with that device current: the array's device address and
// the launch limits.  Caller holds g_na_mu.
int na_setup(int dev, NaDevice &d) {
  if (d.ready) return FPSX_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(FPSX_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  int cus = 0, khz = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  if (e != hipSuccess) return fail(FPSX_EHIP, std::string("hipDeviceGetAttribute: ") + hipGetErrorString(e));
  if (khz <= 0) khz = 100000;
  const char *sec = std::getenv("FPSX_EXCHANGE_TIMEOUT_S");
  double secs = sec ? std::atof(sec) : 2.0;
  if (!(secs > 0)) secs = 2.0;
  d.scratch = (uint64_t *)ptr;
  d.cus = cus;
  d.spin_ticks = (uint64_t)((double)khz * 1000.0 * secs);
  for (int i = 0; i < kNaMaxKernels; ++i) d.occ[i] = -1;
  d.ready = true;
  return FPSX_OK;
}

}  // namespace
}  // namespace fpsx

using namespace fpsx;

FPSX_EXPORT int fpsx_fps_sampling_device_noalloc(const float *d_points, size_t stride_point,
                                                 size_t stride_dim, size_t n_points, size_t dim,
                                                 size_t n_samples, const uint64_t *d_start,
                                                 size_t n_start, uint64_t *d_out,
                                                 float *d_out_points, uint32_t *d_status, int tie,
                                                 float init, void *stream) {
  if (dim == 0) return fail(FPSX_EBADDIM, "dim must be >= 1");
  if (!d_points || !d_start || !d_out) return fail(FPSX_EINVAL, "null pointer argument");
  if (stride_point == 0 || (dim > 1 && stride_dim == 0)) return fail(FPSX_EINVAL, "zero stride");
  if (n_points == 0 || n_samples == 0) return fail(FPSX_EINVAL, "empty cloud or no samples");
  if (n_start == 0 || n_start > n_samples)
    return fail(FPSX_EINVAL, "the number of start indices must be in [1, n_samples]");
  if (tie != FPSX_TIE_HIGHEST && tie != FPSX_TIE_LOWEST) return fail(FPSX_EINVAL, "bad tie rule");
  if (n_points >= (size_t(1) << 31) || n_samples >= (size_t(1) << 32))
    return fail(FPSX_ETOOBIG, "n_points must be < 2^31");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(FPSX_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kNaMaxDevices) return fail(FPSX_EINVAL, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lk(g_na_mu);
  NaDevice &d = g_na[dev];
  int rc = na_setup(dev, d);
  if (rc) return rc;

  size_t cnt = 0;
  const KernelEntry *table = resident_table(&cnt);
  auto max_blocks = [&](const KernelEntry *k) -> int {
    const size_t i = (size_t)(k - table);
    if (i >= (size_t)kNaMaxKernels) return 0;
    if (d.occ[i] < 0) {
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k->fn, k->block, 0) != hipSuccess) nb = 0;
      d.occ[i] = nb * d.cus;
    }
    return d.occ[i];
  };
  Plan p;
  if (!plan_resident_only(n_points, dim, tie, max_blocks, p))
    return fail(FPSX_ETOOBIG, "the allocation-free entry runs the register-resident kernels only: "
                              "no resident shape holds " + std::to_string(n_points) +
                              " points of dimension " + std::to_string(dim) +
                              " on this device (use fpsx_fps_sampling_device_strided)");

  hipStream_t st = (hipStream_t)stream;  // NULL: the device's null stream
  uint64_t *xchg = d.scratch;
  uint32_t *status = d_status ? d_status : (uint32_t *)(d.scratch + kNaXchgWords);
  // Cleared by fill kernels, not hipMemsetAsync, so the call can be captured.
  // (The caller's status word is 4 bytes: cleared as one float of all-zero bits.)
  e = launch_fill_u64(xchg, xchg_words(kNaReplicas), 0, st);
  if (e == hipSuccess)
    e = d_status ? launch_fill_f32((float *)d_status, 1, 0.0f, st)
                 : launch_fill_u64(d.scratch + kNaXchgWords, 1, 0, st);
  if (e != hipSuccess) return fail(FPSX_EHIP, std::string("scratch clear: ") + hipGetErrorString(e));
  FpsParams a;
  std::memset(&a, 0, sizeof(a));
  a.pts = d_points;
  a.ld = stride_dim;
  a.sp = stride_point;
  a.out_pts = d_out_points;
  a.n = (uint32_t)n_points;
  a.m = (uint32_t)n_samples;
  a.starts = d_start;
  a.n_starts = (uint32_t)n_start;
  a.dim = (uint32_t)dim;
  a.out = d_out;
  a.xchg = xchg;
  a.status = status;
  a.init = init;
  a.spin_ticks = d.spin_ticks;
  a.replicas = kNaReplicas;
  a.backoff = 0;
  a.chunk = p.chunk;
  void *args[] = {&a};
  e = hipLaunchKernel(p.k->fn, dim3(p.grid), dim3(p.block), args, 0, st);
  if (e != hipSuccess) return fail(FPSX_EHIP, std::string("hipLaunchKernel: ") + hipGetErrorString(e));
  return FPSX_OK;
}

// fpsx_probe.hip -- the cross-workgroup step of the FPS kernels alone.
//
// exchange_only<NW> runs `rounds` rounds of what every multi-workgroup FPS
// pass ends with, and nothing else: the workgroup's candidate (NW words) is
// staged in LDS behind a barrier, wave 0 publishes it to the replicas of the
// exchange area, ceil(G/64) waves poll every workgroup's words until they
// carry the round's tag, and after one more barrier every wave holds the
// winner.  No point is updated.  The time per round is therefore a floor for
// one pass (one pick, or two for the paired kernel) at the same grid and
// candidate size: the kernels cannot go faster than their exchange.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "fps_device.hpp"
#include "fpsx.h"
#include "fpsx_internal.hpp"

#define FPSX_EXPORT extern "C" __attribute__((visibility("default")))

namespace fpsx {

struct ProbeParams {
  uint64_t *xchg;       // exchange area (cleared before the launch)
  uint32_t *status;     // status word (cleared before the launch)
  int *sink;            // one word, keeps the winners live
  uint32_t rounds;
  uint32_t replicas;
  uint64_t spin_ticks;
};

template <int NW>
__global__ void __launch_bounds__(1024) exchange_only(ProbeParams a) {
  constexpr int W = 16;
  __shared__ int s_cand[2][NW][W];
  __shared__ int s_xw[2][NW + 1][W];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  int sink = 0;
  for (uint32_t r = 0; r < a.rounds; ++r) {
    const uint32_t e = r + 1;
    const int par = (int)(e & 1u);
    // A candidate that differs per workgroup and round, so the winner moves.
    if (tid < NW) {
      const uint32_t h = (g * 2654435761u) ^ (e * 40503u) ^ (uint32_t)tid;
      s_cand[par][tid][0] = (tid == 1) ? (int)g : (int)(h & 0x7fffffffu);
    }
    lds_barrier();
    int win[NW];
    if (!exchange<kTieHighest, NW, W>(a.xchg, a.replicas, kPageWords, gridDim.x, g, e, 0u,
                                      a.spin_ticks, s_cand[par], s_xw[par], 0, win)) {
      if (tid == 0) atomicOr(a.status, kStatusTimeout);
      return;
    }
    sink ^= win[0] + win[NW - 1];
  }
  if (g == 0 && tid == 0) a.sink[0] = sink;
}

}  // namespace fpsx

using namespace fpsx;

FPSX_EXPORT int fpsx_probe_exchange(unsigned grid, unsigned words, unsigned rounds, void *stream,
                                    double *us_per_round) {
  if (grid < 2 || grid > 512) return fail(FPSX_EINVAL, "probe: grid must be in [2, 512]");
  if (words != 5 && words != 10) return fail(FPSX_EINVAL, "probe: words must be 5 or 10");
  if (rounds == 0 || !us_per_round) return fail(FPSX_EINVAL, "probe: bad arguments");
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  Ctx &c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  hipStream_t st = stream ? (hipStream_t)stream : c.stream;
  const void *fn = (words == 5) ? (const void *)&exchange_only<5> : (const void *)&exchange_only<10>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 1024, 0) != hipSuccess ||
      (size_t)per_cu * c.cus < grid)
    return fail(FPSX_ETOOBIG, "probe: grid exceeds co-residency");
  rc = scratch_acquire(c, st);
  if (rc) return rc;
  // Replicas as the kernels use them: up to 8 for one-word-group candidates,
  // 4 for the paired kernels' two (A and its lookahead).
  const uint32_t reps = (uint32_t)std::min(c.replicas, words == 10 ? 4 : kMaxReplicas);
  const size_t xwords = xchg_words(reps);
  if (c.xchg.reserve(xwords * sizeof(uint64_t)) != hipSuccess ||
      c.status.reserve(sizeof(uint64_t)) != hipSuccess || c.out.reserve(sizeof(uint64_t)) != hipSuccess)
    return fail(FPSX_ENOMEM, "probe: scratch allocation failed");
  if (launch_fill_u64((uint64_t *)c.xchg.p, xwords, 0, st) != hipSuccess ||
      launch_fill_u64((uint64_t *)c.status.p, 1, 0, st) != hipSuccess)
    return fail(FPSX_EHIP, "probe: fill failed");
  ProbeParams a;
  a.xchg = (uint64_t *)c.xchg.p;
  a.status = (uint32_t *)c.status.p;
  a.sink = (int *)c.out.p;
  a.rounds = rounds;
  a.replicas = reps;
  a.spin_ticks = c.spin_ticks;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return fail(FPSX_EHIP, "probe: hipEventCreate");
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(FPSX_EHIP, "probe: hipEventCreate");
  }
  void *args[] = {&a};
  hipError_t err = hipEventRecord(e0, st);
  if (err == hipSuccess) err = hipLaunchKernel(fn, dim3(grid), dim3(1024), args, 0, st);
  if (err == hipSuccess) err = hipEventRecord(e1, st);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err != hipSuccess) return fail(FPSX_EHIP, std::string("probe: ") + hipGetErrorString(err));
  rc = scratch_release(c, st);
  if (rc) return rc;
  if (hipMemcpy(c.host_status, c.status.p, sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(FPSX_EHIP, "probe: status read failed");
  rc = status_code(*c.host_status);
  if (rc) return rc;
  *us_per_round = (double)ms * 1000.0 / (double)rounds;
  return FPSX_OK;
}

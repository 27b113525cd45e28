// fps_batch.hip -- many clouds in one persistent launch (work queue of teams).
//
// Each cloud of the batch is a job: vanilla FPS (reference src/lib.cpp:
// 188-246) over that cloud alone, run by a team of `team` workgroups that hold
// its points and min-distances in VGPRs.  The host orders jobs (multi-workgroup
// teams first, then single-workgroup jobs, each group longest first) and
// flattens them into a list of team-member slots; every workgroup of the
// persistent grid repeatedly claims the next slot from an atomic counter and
// runs that member's share of the job.  Members of one team are consecutive
// slots, so a team is always completed by the next workgroups to come free:
// with every team no larger than the grid, nothing can wait forever.  Team
// members agree on each pick through the same tag-validated exchange as the
// single-cloud kernels, in an exchange area of the team's own.
//
// A job may cover only passes [s_begin, s_end) of its cloud: min-distances are
// then restored from / saved to a.state, and the pivot of the first pass is
// the pick the previous chunk wrote.  The serving loop uses this to advance
// many requests a chunk at a time and admit new ones between launches.
#include "fps_batch.hpp"
#include "fps_device.hpp"

namespace fpsx {

template <int D, int B, int PPT, int TIE>
__device__ __forceinline__ bool run_job(const BatchParams &a, const BatchJob &J, uint32_t j,
                                        uint32_t r, int (*s_cand)[2 + D][B / 64],
                                        int (*s_xw)[3 + D][B / 64]) {
  constexpr int W = B / 64;
  constexpr int NW = 2 + D;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t lo = r * J.share;
  const uint32_t hi = min(J.n, lo + J.share);
  const int kmax = (int)((hi - lo + B - 1) / B);  // live slots per thread (uniform, <= PPT)
  const float *pts = a.pts + J.pt_off;

  float c[PPT][D];
  float dm[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t i = lo + (uint32_t)(k * B + tid);
    const bool live = k < kmax && i < hi;
#pragma unroll
    for (int d = 0; d < D; ++d) c[k][d] = live ? pts[(size_t)d * J.ld + i] : 0.0f;
    dm[k] = live ? (J.s_begin == 0 ? a.init : a.state[J.state_off + i]) : -1.0f;
  }
  uint32_t pidx = (J.s_begin == 0) ? (uint32_t)J.start : (uint32_t)a.out[J.out_off + J.s_begin - 1];
  if (pidx >= J.n) {  // never read outside the cloud
    if (tid == 0) atomicOr(a.status, kStatusBadStart);
    pidx = 0;
  }
  float pv[D];
#pragma unroll
  for (int d = 0; d < D; ++d) pv[d] = pts[(size_t)d * J.ld + pidx];
#pragma unroll
  for (int k = 0; k < PPT; ++k) settle<D>(c[k]);
  settle<PPT>(dm);
  settle<D>(pv);
  if (J.s_begin == 0 && r == 0 && tid == B - 64) a.out[J.out_off] = pidx;

  uint64_t *area = a.xchg + (size_t)J.area * kTeamXchgWords;
  uint32_t e = 0;
  for (uint32_t s = (J.s_begin == 0 ? 1u : J.s_begin); s < J.s_end; ++s) {
    int bv = INT_MIN;
    int bk = 0;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      if (k < kmax) {
        dm[k] = min_keep(sqdist<D>(c[k], pv), dm[k]);
        const int b = __float_as_int(dm[k]);
        const bool take = (TIE == kTieHighest) ? (b >= bv) : (b > bv);
        bv = take ? b : bv;
        bk = take ? k : bk;
      }
    }
    ++e;
    const int par = (int)(e & 1u);
    const uint32_t bidx = lo + (uint32_t)(bk * B + tid);
    int wv;
    const int wl = wave_argmax_lane<TIE>(bv, bidx, wv);
    const int kw = readlane_i(bk, wl);
    const int widx = readlane_i((int)bidx, wl);
    float wc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float v = c[0][d];
#pragma unroll
      for (int q = 1; q < PPT; ++q) v = (kw == q) ? c[q][d] : v;
      wc[d] = readlane_f(v, wl);
    }
    if (lane == 0) {
      s_cand[par][0][wave] = wv;
      s_cand[par][1][wave] = widx;
#pragma unroll
      for (int d = 0; d < D; ++d) s_cand[par][2 + d][wave] = __float_as_int(wc[d]);
    }
    lds_barrier();
    const int v2 = (lane < W) ? s_cand[par][0][lane] : INT_MIN;
    const uint32_t i2 = (lane < W) ? (uint32_t)s_cand[par][1][lane] : 0u;
    int gv;
    const int gl = wave_argmax_lane<TIE>(v2, i2, gv);
    if (J.team == 1) {
      pidx = (uint32_t)readlane_i((int)i2, gl);
#pragma unroll
      for (int d = 0; d < D; ++d) pv[d] = __int_as_float(s_cand[par][2 + d][gl]);
    } else {
      int win[NW];
      if (!exchange<TIE, NW, W>(area, 1, kMaxTeam, J.team, r, e, 0, a.spin_ticks, s_cand[par],
                                s_xw[par], gl, win))
        return false;
      pidx = (uint32_t)win[1];
#pragma unroll
      for (int d = 0; d < D; ++d) pv[d] = __int_as_float(win[2 + d]);
    }
    if (r == 0 && tid == B - 64) a.out[J.out_off + s] = pidx;
  }
  if (J.s_end < J.m) {
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const uint32_t i = lo + (uint32_t)(k * B + tid);
      if (k < kmax && i < hi) a.state[J.state_off + i] = dm[k];
    }
  }
  return true;
}

template <int D, int B, int PPT, int TIE>
__global__ void __launch_bounds__(B) fps_batch(BatchParams a) {
  constexpr int W = B / 64;
  constexpr int NW = 2 + D;
  __shared__ int s_cand[2][NW][W];
  __shared__ int s_xw[2][NW + 1][W];
  __shared__ uint32_t s_claim;
  __shared__ int s_ok;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) s_claim = atomicAdd(a.claim, 1u);
    lds_barrier();
    const uint32_t k = s_claim;
    if (k >= a.total_members) return;
    const uint32_t j = a.member_job[k];
    const BatchJob J = a.jobs[j];
    const uint32_t r = k - J.member_base;
    if (J.team > 1) {
      // Team formation: wait until every member has claimed its slot.
      if (tid == 0) {
        atomicAdd(&a.ready[j], 1u);
        const uint64_t deadline = now_ticks() + a.team_ticks;
        int ok = 1;
        while (__hip_atomic_load(&a.ready[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
               J.team) {
          if (now_ticks() > deadline) {
            ok = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
        s_ok = ok;
      }
      lds_barrier();
      if (!s_ok) {
        if (tid == 0) atomicOr(a.status, kStatusTimeout);
        return;
      }
    }
    if (!run_job<D, B, PPT, TIE>(a, J, j, r, s_cand, s_xw)) {
      if (tid == 0) atomicOr(a.status, kStatusTimeout);
      return;
    }
    lds_barrier();  // every lane has read s_claim / s_ok before they are rewritten
  }
}

#define FPSX_BATCH(D, B, P) \
  {D, B, P, kTieHighest, (const void *)&fps_batch<D, B, P, kTieHighest>}

static const KernelEntry kBatch[] = {
    FPSX_BATCH(3, 1024, 16), FPSX_BATCH(3, 1024, 8), FPSX_BATCH(3, 256, 32),
    FPSX_BATCH(3, 256, 16), FPSX_BATCH(3, 256, 8), FPSX_BATCH(3, 256, 4),
};
#undef FPSX_BATCH

const KernelEntry *batch_table(size_t *count) {
  *count = sizeof(kBatch) / sizeof(kBatch[0]);
  return kBatch;
}

}  // namespace fpsx

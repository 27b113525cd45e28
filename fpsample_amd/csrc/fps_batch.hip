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

// Paired team loop: one exchange settles up to two picks.  Every member
// publishes its best point A and its lookahead A' (its best after a
// hypothetical update with A as the pivot, nothing stored), and one wave of
// every member evaluates the team's pairs:
//   p1 = the best A (member w1);
//   w1's best after p1 is A'_w1; a member whose A keeps its value under p1
//   (min(dmin, dist(A, p1)) == dmin, the update's own arithmetic) still has
//   A as its best; any other member is unknown, every point of it now ranking
//   below its A;
//   p2 = the best known best when it ranks above every unknown member's A.
// The picks are those of the one-pick loop; only the number of exchanges
// changes.  No spatial order is needed: every member sweeps all its points.
template <int D, int B, int PPT, int TIE>
__device__ __forceinline__ bool run_job_pair(const BatchParams &a, const BatchJob &J, uint32_t r,
                                             int (*s_cand)[2 + D][B / 64],
                                             int (*s_la)[2 + D][B / 64],
                                             int (*s_res)[2 * (2 + D) + 2]) {
  constexpr int W = B / 64;
  constexpr int NW = 2 + D;
  constexpr int NP = 2 * NW;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t lo = r * J.share;
  const uint32_t hi = min(J.n, lo + J.share);
  const int kmax = (int)((hi - lo + B - 1) / B);
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
  if (pidx >= J.n) {
    if (tid == 0) atomicOr(a.status, kStatusBadStart);
    pidx = 0;
  }
  float pv[D], pv2[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    pv[d] = pts[(size_t)d * J.ld + pidx];
    pv2[d] = 0.0f;
  }
#pragma unroll
  for (int k = 0; k < PPT; ++k) settle<D>(c[k]);
  settle<PPT>(dm);
  settle<D>(pv);
  if (J.s_begin == 0 && r == 0 && tid == B - 64) a.out[J.out_off] = pidx;

  uint64_t *area = a.xchg + (size_t)J.area * kTeamXchgWords;
  const uint32_t T = J.team;
  bool have2 = false;  // the last round gave two picks: pv, then pv2
  uint32_t e = 0;
  uint32_t s = (J.s_begin == 0 ? 1u : J.s_begin);
  while (s < J.s_end) {
    // Apply the last round's picks (in pick order) and take the lane's best.
    int bv = INT_MIN;
    int bk = 0;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      if (k < kmax) {
        float v = min_keep(sqdist<D>(c[k], pv), dm[k]);
        if (have2) v = min_keep(sqdist<D>(c[k], pv2), v);
        dm[k] = v;
        const int b = __float_as_int(v);
        const bool take = (TIE == kTieHighest) ? (b >= bv) : (b > bv);
        bv = take ? b : bv;
        bk = take ? k : bk;
      }
    }
    ++e;
    const int par = (int)(e & 1u);
    {
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
    }
    lds_barrier();
    // The member's best A, then the lookahead: the best of
    // min(dmin, dist(., A)) over the member's points.
    const int v2 = (lane < W) ? s_cand[par][0][lane] : INT_MIN;
    const uint32_t i2 = (lane < W) ? (uint32_t)s_cand[par][1][lane] : 0u;
    int gv;
    const int gl = wave_argmax_lane<TIE>(v2, i2, gv);
    float qa[D];
#pragma unroll
    for (int d = 0; d < D; ++d) qa[d] = __int_as_float(s_cand[par][2 + d][gl]);
    {
      int lv = INT_MIN;
      int lk = 0;
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        if (k < kmax) {
          const int b = __float_as_int(min_keep(sqdist<D>(c[k], qa), dm[k]));
          const bool take = (TIE == kTieHighest) ? (b >= lv) : (b > lv);
          lv = take ? b : lv;
          lk = take ? k : lk;
        }
      }
      const uint32_t lidx = lo + (uint32_t)(lk * B + tid);
      int wv;
      const int wl = wave_argmax_lane<TIE>(lv, lidx, wv);
      const int kw = readlane_i(lk, wl);
      const int widx = readlane_i((int)lidx, wl);
      float wc[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        float v = c[0][d];
#pragma unroll
        for (int q = 1; q < PPT; ++q) v = (kw == q) ? c[q][d] : v;
        wc[d] = readlane_f(v, wl);
      }
      if (lane == 0) {
        s_la[par][0][wave] = wv;
        s_la[par][1][wave] = widx;
#pragma unroll
        for (int d = 0; d < D; ++d) s_la[par][2 + d][wave] = __float_as_int(wc[d]);
      }
    }
    lds_barrier();
    if (wave == 0) {
      // Publish (A, A'): word k < NW of A, word NW + k of A'.  Member r's
      // word k of buffer e & 1 lives at area[((e & 1) * kSlotWords + k) *
      // kMaxTeam + r].
      const int u2 = (lane < W) ? s_la[par][0][lane] : INT_MIN;
      const uint32_t j2 = (lane < W) ? (uint32_t)s_la[par][1][lane] : 0u;
      int lw;
      const int ll = wave_argmax_lane<TIE>(u2, j2, lw);
      uint64_t *buf = area + (size_t)(e & 1u) * kSlotWords * kMaxTeam;
      if (lane < NP) {
        const int word = (lane < NW) ? s_cand[par][lane][gl] : s_la[par][lane - NW][ll];
        xchg_store(buf + (size_t)lane * kMaxTeam + r, ((uint64_t)(uint32_t)word << 32) | e);
      }
      // Poll: lane q reads member q's pair.
      int x[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) x[k] = 0;
      bool ok = true;
      const bool live = (uint32_t)lane < T;
      if (live) {
        uint64_t w[NP];
        const uint64_t deadline = now_ticks() + a.spin_ticks;
        for (;;) {
          bool all = true;
#pragma unroll
          for (int k = 0; k < NP; ++k) w[k] = xchg_load(buf + (size_t)k * kMaxTeam + lane);
#pragma unroll
          for (int k = 0; k < NP; ++k) all = all && ((uint32_t)w[k] == e);
          if (all) break;
          if (now_ticks() > deadline) {
            ok = false;
            break;
          }
        }
#pragma unroll
        for (int k = 0; k < NP; ++k) x[k] = (int)(w[k] >> 32);
      }
      const bool valid = live && ok;
      // p1: the best A.
      int v1;
      const int l1 = wave_argmax_lane<TIE>(valid ? x[0] : INT_MIN, (uint32_t)x[1], v1);
      int r1[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) r1[k] = readlane_i(x[k], l1);
      float q1[D], qx[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        q1[d] = __int_as_float(r1[2 + d]);
        qx[d] = __int_as_float(x[2 + d]);
      }
      // Each A after p1, in the update's own arithmetic.
      const float post = min_keep(sqdist<D>(qx, q1), __int_as_float(x[0]));
      const bool moved = valid && lane != l1 && __float_as_int(post) != x[0];
      const bool mine = lane == l1;
      const int tv = mine ? x[NW] : ((valid && !moved) ? x[0] : INT_MIN);
      const uint32_t ti = (uint32_t)(mine ? x[NW + 1] : x[1]);
      int vb;
      const int lb = wave_argmax_lane<TIE>(tv, ti, vb);
      const uint32_t ib = (uint32_t)readlane_i((int)ti, lb);
      const bool beaten = moved && !beats<TIE>(vb, ib, x[0], (uint32_t)x[1]);
      const bool two = __ballot(beaten) == 0;
      const bool all_ok = __ballot(live && !ok) == 0;
      int rb[D];
#pragma unroll
      for (int d = 0; d < D; ++d) rb[d] = readlane_i(mine ? x[NW + 2 + d] : x[2 + d], lb);
      if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NW; ++k) s_res[par][k] = r1[k];
        s_res[par][NW] = vb;
        s_res[par][NW + 1] = (int)ib;
#pragma unroll
        for (int d = 0; d < D; ++d) s_res[par][NW + 2 + d] = rb[d];
        s_res[par][NP] = two ? 1 : 0;
        s_res[par][NP + 1] = all_ok ? 1 : 0;
      }
    }
    lds_barrier();
    if (s_res[par][NP + 1] == 0) return false;
    pidx = (uint32_t)__builtin_amdgcn_readfirstlane(s_res[par][1]);
    const uint32_t i2nd = (uint32_t)__builtin_amdgcn_readfirstlane(s_res[par][NW + 1]);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      pv[d] = __int_as_float(__builtin_amdgcn_readfirstlane(s_res[par][2 + d]));
      pv2[d] = __int_as_float(__builtin_amdgcn_readfirstlane(s_res[par][NW + 2 + d]));
    }
    have2 = s_res[par][NP] != 0 && s + 1 < J.s_end;
    if (r == 0 && tid == B - 64) {
      a.out[J.out_off + s] = pidx;
      if (have2) a.out[J.out_off + s + 1] = i2nd;
    }
    s += have2 ? 2u : 1u;
  }
  if (J.s_end < J.m) {
    // The saved state holds every pick but the last applied (the next chunk
    // starts by applying the last one): a final pair leaves its first pick
    // to apply here.
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const uint32_t i = lo + (uint32_t)(k * B + tid);
      if (k < kmax && i < hi) {
        const float v = have2 ? min_keep(sqdist<D>(c[k], pv), dm[k]) : dm[k];
        a.state[J.state_off + i] = v;
      }
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
    // The paired loop holds a second pivot and the lookahead's candidate:
    // 1024-thread workgroups (128 VGPRs) fit it up to 8 points per thread.
    // Other shapes compile the one-pick loop alone (its code and LDS exactly
    // as without the paired loop).
    if constexpr (B < 1024 || PPT <= 8) {
      __shared__ int s_la[2][NW][W];
      __shared__ int s_res[2][2 * NW + 2];
      const bool done = (a.pair && J.team > 1)
                            ? run_job_pair<D, B, PPT, TIE>(a, J, r, s_cand, s_la, s_res)
                            : run_job<D, B, PPT, TIE>(a, J, j, r, s_cand, s_xw);
      if (!done) {
        if (tid == 0) atomicOr(a.status, kStatusTimeout);
        return;
      }
    } else {
      if (!run_job<D, B, PPT, TIE>(a, J, j, r, s_cand, s_xw)) {
        if (tid == 0) atomicOr(a.status, kStatusTimeout);
        return;
      }
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

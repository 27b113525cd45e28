// fps_batch_pruned.hip -- many clouds in one persistent launch, each run with
// the spatial order, per-wave box pruning and paired picks of fps_pruned_pair.
//
// The work queue is the batch kernel's (fps_batch.hip): jobs flattened into
// team-member slots, claimed from an atomic counter, teams formed by the next
// workgroups to come free.  Each job's cloud has been put in spatial order on
// the device beforehand (fpsx_spatial.hip: Hilbert order cut into blocks of
// 64 * PPT positions, each block re-sorted by original index), so member r of
// the job's team owns sorted positions [r * B * PPT, (r + 1) * B * PPT) and
// each wave one block.  A round of a member is the paired kernel's:
//   apply the last round's picks to the waves whose box lies closer than
//   their largest min-distance; form the member's best A (and, when the
//   member could hold the next pick, its lookahead A': its best after a
//   hypothetical update with A as the pivot); publish (A, A') to the team's
//   exchange area and poll every member's pair; derive p1 (the best A) and,
//   when nothing unknown can beat it, p2 -- the one-pick loop's next two
//   picks (reference src/lib.cpp:211-238) -- identically in every member.
// A team of one workgroup skips the exchange: its own pair is the result.
//
// A job may cover only picks [s_begin, s_end): min-distances are restored
// from / saved to state[state_off + position], and the last pick's
// coordinates travel in state[state_off + n .. + 3); the next chunk starts
// by applying that pick.  The serving loop advances its requests this way.
#include "fps_batch.hpp"
#include "fps_device.hpp"

namespace fpsx {

// VEC: the hand-offs as 16-byte records moved with 128-bit LDS accesses
// (claim-loop batch launches: 29.15 -> 28.4 ms on the 256-cloud batch), or
// word by word (the serving loop's one-slot 1024 x 16 kernel, where the
// records cost one more spilled VGPR and 1.8 % of busy time).
template <int B, int PPT, int TIE, bool VEC>
struct PPShared {
  static constexpr int D = 3;
  static constexpr int W = B / 64;
  static constexpr int NW = 2 + D;
  static constexpr int NP = 2 * NW;
  // Hand-offs by exchange parity: per wave its best and "changed" flag, its
  // lookahead best; the round's result (p1, p2, two-pick flag, all-arrived
  // flag); M.
  int4 top4[VEC ? 2 : 1][VEC ? W : 1][2];
  int4 la4[VEC ? 2 : 1][VEC ? W : 1][2];
  int4 res4[VEC ? 2 : 1][4];
  int top[VEC ? 1 : 2][VEC ? 1 : NW + 1][VEC ? 1 : W];
  int la[VEC ? 1 : 2][VEC ? 1 : NW][VEC ? 1 : W];
  int res[VEC ? 1 : 2][VEC ? 1 : NP + 2];
  int2 mu[2];
  float gbox[2 * D][64];   // every member's bounding box
  float wbox[2 * D][W];    // this member's wave boxes
  int ok;
  BatchJob job;            // the claimed job (read from LDS where used: short register lives)
  uint32_t perm[B * PPT];  // original index of every position this member owns
};

// One member's share of job J (member r of J.team).  Returns false when an
// exchange wait ran out.
template <int B, int PPT, int TIE, bool VEC>
__device__ __forceinline__ bool run_job_pp(const BatchPPParams &ap, uint32_t r,
                                           PPShared<B, PPT, TIE, VEC> &sh) {
  constexpr int D = 3;
  constexpr int W = B / 64;
  constexpr int NW = 2 + D;
  constexpr int NP = 2 * NW;
  constexpr int PER_WG = B * PPT;
  constexpr uint32_t R = kPPReplicas;
  constexpr uint32_t P = kPPPitch;  // pitch of a word array in the team area (>= kMaxTeam)
  static_assert(kPPPitch >= kMaxTeam, "a team area holds a word per member");
  const BatchParams &a = ap.b;
  // Job fields are read from LDS where they are used (J is the LDS copy),
  // not held in registers through the loop: the paired loop needs every
  // scalar and vector register the 1024 x 16 shape has.
  const BatchJob &J = sh.job;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t n = J.n;
  const uint32_t base = r * (uint32_t)PER_WG;
  const uint32_t wbase = base + (uint32_t)(wave * 64 * PPT);
  const bool recorder = (r == 0 && tid >= B - 64);
  const float *sorted = ap.sorted + J.sorted_off;
  const uint32_t *perm = ap.perm + J.perm_off;

  for (uint32_t q = tid; q < PER_WG; q += B) sh.perm[q] = (base + q < n) ? perm[base + q] : 0u;

  float c[PPT][D];
  float dm[PPT];
  float lo[D], hi[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    lo[j] = INFINITY;
    hi[j] = -INFINITY;
  }
  const bool fresh = J.s_begin == 0;
  // Unconditional loads of a clamped position; padding slots are then overwritten.
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t i = wbase + (uint32_t)(k * 64 + lane);
    const bool live = i < n;
    const uint32_t ic = live ? i : n - 1u;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const float v = sorted[(size_t)j * n + ic];
      c[k][j] = live ? v : 0.0f;
      lo[j] = live ? fminf(lo[j], v) : lo[j];
      hi[j] = live ? fmaxf(hi[j], v) : hi[j];
    }
    dm[k] = live ? a.init : -1.0f;
  }
  if (!fresh) {  // a later chunk: the saved min-distances (uniform branch)
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const uint32_t i = wbase + (uint32_t)(k * 64 + lane);
      const float sv = a.state[J.state_off + (i < n ? i : n - 1u)];
      dm[k] = (i < n) ? sv : dm[k];
    }
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    lo[j] = wave_min_f(lo[j]);
    hi[j] = wave_max_f(hi[j]);
  }
  // The pending pick: the start (coordinates from the original cloud), or
  // the last pick of the previous chunk.
  float p1[D], p2[D];
  uint32_t i1 = fresh ? (uint32_t)J.start : 0u;
  if (i1 >= n) {  // never read outside the cloud
    if (tid == 0) atomicOr(a.status, kStatusBadStart);
    i1 = 0;
  }
  const float *orig = a.pts + J.pt_off;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    p1[j] = fresh ? orig[(size_t)j * J.ld + i1] : a.state[J.state_off + n + j];
    p2[j] = 0.0f;
  }
#pragma unroll
  for (int k = 0; k < PPT; ++k) settle<D>(c[k]);
  settle<PPT>(dm);
  settle<D>(p1);
  if (fresh && recorder && lane == 0) a.out[J.out_off] = i1;
  lds_barrier();  // perm complete

  // Best (value, original index, coordinates) of this wave over val(k);
  // each lane carries its best slot's coordinates through the scan.
  auto wave_best = [&](auto val, int &ov, int &oi, float(&oc)[D]) {
    int bv = INT_MIN;
    int bk = 0;
    float bc[D];
#pragma unroll
    for (int j = 0; j < D; ++j) bc[j] = 0.0f;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int b = __float_as_int(val(k));
      const bool take = (TIE == kTieHighest) ? (b >= bv) : (b > bv);
      bv = take ? b : bv;
      bk = take ? k : bk;
#pragma unroll
      for (int j = 0; j < D; ++j) bc[j] = take ? c[k][j] : bc[j];
    }
    const uint32_t bpos = (uint32_t)(bk * 64 + lane);
    int wv;
    const int wl = wave_argmax_lane<TIE>(bv, bpos, wv);
    const uint32_t wpos = (uint32_t)readlane_i((int)bpos, wl);
#pragma unroll
    for (int j = 0; j < D; ++j) oc[j] = readlane_f(bc[j], wl);
    ov = wv;
    oi = (int)sh.perm[wave * 64 * PPT + wpos];
  };

  int cv = INT_MIN;  // this wave's best (uniform)
  int ci = 0;
  float cc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) cc[j] = 0.0f;
  bool cached = false;

  int ga[NW], gl_[NW];  // the member's A and A' (uniform), kept while unchanged
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    ga[k] = 0;
    gl_[k] = 0;
  }
  bool have2 = false;
  bool la_valid = false;
  int mu_v = INT_MIN;  // M: best A of the certainly-unchanged other members
  uint32_t mu_i = 0;
  uint32_t e = 0;
  constexpr int K = (int)R < 64 / NP ? (int)R : 64 / NP;  // replicas a pair's words allow

  if (tid < 2) {
    sh.mu[tid] = make_int2(INT_MIN, 0);
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      sh.wbox[j][wave] = lo[j];
      sh.wbox[D + j][wave] = hi[j];
    }
  }
  lds_barrier();
  ++e;
  if (wave == 0) {
    // The members' bounding boxes (exchange 1 of 2*D words; a team of one
    // keeps its own).
    float bw[2 * D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      bw[j] = (lane < W) ? sh.wbox[j][lane] : INFINITY;
      bw[D + j] = (lane < W) ? sh.wbox[D + j][lane] : -INFINITY;
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      bw[j] = wave_min_f(bw[j]);
      bw[D + j] = wave_max_f(bw[D + j]);
    }
    bool ok = true;
    const uint32_t T = J.team;
    if (T == 1) {
      if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 2 * D; ++k) sh.gbox[k][0] = bw[k];
      }
    } else {
      uint64_t *bufs = a.xchg + (size_t)J.area * kPPTeamWords + (size_t)(e & 1u) * R * kSlotWords * P;
      if (lane < K * 2 * D) {
        const int rep = lane / (2 * D);
        const int k = lane - rep * (2 * D);
        float word = bw[0];
#pragma unroll
        for (int q = 1; q < 2 * D; ++q) {
          float wq = bw[q];
          asm volatile("" : "+v"(wq));
          word = (k == q) ? wq : word;
        }
        xchg_store(bufs + ((size_t)rep * kSlotWords + k) * P + r,
                   ((uint64_t)(uint32_t)__float_as_int(word) << 32) | e);
      }
      const uint64_t *buf = bufs + (size_t)(r % (uint32_t)K) * kSlotWords * P;
      if ((uint32_t)lane < T) {
        uint64_t w[2 * D];
        const uint64_t deadline = now_ticks() + a.spin_ticks;
        for (;;) {
          bool all = true;
#pragma unroll
          for (int k = 0; k < 2 * D; ++k) w[k] = xchg_load(buf + (size_t)k * P + lane);
#pragma unroll
          for (int k = 0; k < 2 * D; ++k) all = all && ((uint32_t)w[k] == e);
          if (all) break;
          if (now_ticks() > deadline) {
            ok = false;
            break;
          }
        }
#pragma unroll
        for (int k = 0; k < 2 * D; ++k) sh.gbox[k][lane] = __int_as_float((int)(w[k] >> 32));
      }
    }
    const bool all_ok = __ballot((uint32_t)lane < T && !ok) == 0;
    if (lane == 0) sh.ok = all_ok ? 1 : 0;
  }
  lds_barrier();
  if (sh.ok == 0) return false;

  uint32_t s = fresh ? 1u : J.s_begin;
  while (s < sh.job.s_end) {
    // Apply the pending picks where a wave's box is closer than its largest
    // min-distance.
    const bool need1 = !cached || box_bound<D>(p1, lo, hi) < __int_as_float(cv);
    const bool need2 = have2 && box_bound<D>(p2, lo, hi) < __int_as_float(cv);
    if (need1) {
#pragma unroll
      for (int k = 0; k < PPT; ++k) dm[k] = min_keep(sqdist<D>(c[k], p1), dm[k]);
    }
    if (need2) {
#pragma unroll
      for (int k = 0; k < PPT; ++k) dm[k] = min_keep(sqdist<D>(c[k], p2), dm[k]);
    }
    if (need1 || need2) {
      wave_best([&](int k) { return dm[k]; }, cv, ci, cc);
      cached = true;
    }

    ++e;
    const int par = (int)(e & 1u);
    int tw[NW + 1];
    if constexpr (VEC) {
      if (lane == 0) {
        sh.top4[par][wave][0] = make_int4(cv, ci, __float_as_int(cc[0]), __float_as_int(cc[1]));
        sh.top4[par][wave][1] = make_int4(__float_as_int(cc[2]), (need1 || need2) ? 1 : 0, 0, 0);
      }
      lds_barrier();
      const int4 t0 = (lane < W) ? sh.top4[par][lane][0] : make_int4(INT_MIN, 0, 0, 0);
      const int4 t1 = (lane < W) ? sh.top4[par][lane][1] : make_int4(0, 0, 0, 0);
      tw[0] = t0.x;
      tw[1] = t0.y;
      tw[2] = t0.z;
      tw[3] = t0.w;
      tw[4] = t1.x;
      tw[5] = t1.y;
    } else {
      if (lane == 0) {
        sh.top[par][0][wave] = cv;
        sh.top[par][1][wave] = ci;
#pragma unroll
        for (int j = 0; j < D; ++j) sh.top[par][2 + j][wave] = __float_as_int(cc[j]);
        sh.top[par][NW][wave] = (need1 || need2) ? 1 : 0;
      }
      lds_barrier();
#pragma unroll
      for (int k = 0; k <= NW; ++k) tw[k] = (lane < W) ? sh.top[par][k][lane] : 0;
      if (lane >= W) tw[0] = INT_MIN;
    }
    {
      const int2 mu = sh.mu[par ^ 1];
      mu_v = mu.x;
      mu_i = (uint32_t)mu.y;
    }
    // The chunk's first round forms A everywhere.
    const bool changed = e == 2u || __ballot(lane < W && tw[NW] != 0) != 0;
    if (changed) {
      int gv;
      const int wl = wave_argmax_lane<TIE>(tw[0], (uint32_t)tw[1], gv);
#pragma unroll
      for (int k = 0; k < NW; ++k) ga[k] = readlane_i(tw[k], wl);
    }
    const bool needL = (changed || !la_valid) && beats<TIE>(ga[0], (uint32_t)ga[1], mu_v, mu_i);
    if (changed) la_valid = false;
    if (needL) {
      // Lookahead: only waves whose best moves under A and could still beat
      // the best unmoved wave best sweep their points.
      la_valid = true;
      float pa[D], tc[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        pa[j] = __int_as_float(ga[2 + j]);
        tc[j] = __int_as_float(tw[2 + j]);
      }
      const float tpost = min_keep(sqdist<D>(tc, pa), __int_as_float(tw[0]));
      const bool texact = lane < W && __float_as_int(tpost) == tw[0];
      int lbv;
      const int lbl = wave_argmax_lane<TIE>(texact ? tw[0] : INT_MIN, (uint32_t)tw[1], lbv);
      const uint32_t lbi = (uint32_t)readlane_i(tw[1], lbl);
      const uint64_t sweepers =
          __ballot(lane < W && !texact && beats<TIE>(tw[0], (uint32_t)tw[1], lbv, lbi));
      int lv = cv, li = ci;
      float lc[D];
#pragma unroll
      for (int j = 0; j < D; ++j) lc[j] = cc[j];
      if ((sweepers >> wave) & 1ull)
        wave_best([&](int k) { return min_keep(sqdist<D>(c[k], pa), dm[k]); }, lv, li, lc);
      int lw[NW];
      if constexpr (VEC) {
        if (lane == 0) {
          sh.la4[par][wave][0] = make_int4(lv, li, __float_as_int(lc[0]), __float_as_int(lc[1]));
          sh.la4[par][wave][1] = make_int4(__float_as_int(lc[2]), 0, 0, 0);
        }
        lds_barrier();
        const int4 t0 = (lane < W) ? sh.la4[par][lane][0] : make_int4(INT_MIN, 0, 0, 0);
        const int t1 = (lane < W) ? sh.la4[par][lane][1].x : 0;
        lw[0] = t0.x;
        lw[1] = t0.y;
        lw[2] = t0.z;
        lw[3] = t0.w;
        lw[4] = t1;
      } else {
        if (lane == 0) {
          sh.la[par][0][wave] = lv;
          sh.la[par][1][wave] = li;
#pragma unroll
          for (int j = 0; j < D; ++j) sh.la[par][2 + j][wave] = __float_as_int(lc[j]);
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < NW; ++k) lw[k] = (lane < W) ? sh.la[par][k][lane] : 0;
        if (lane >= W) lw[0] = INT_MIN;
      }
      int lgv;
      const int wl2 = wave_argmax_lane<TIE>(lw[0], (uint32_t)lw[1], lgv);
#pragma unroll
      for (int k = 0; k < NW; ++k) gl_[k] = readlane_i(lw[k], wl2);
    }

    int sv = INT_MIN, si = 0;  // wave 0: this lane's member's A (value, index)
    if (wave == 0) {
      int x[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) x[k] = 0;
      bool ok = true;
      const uint32_t T = J.team;
      const bool live = (uint32_t)lane < T;
      if (T == 1) {
        // A team of one: its own pair, no exchange.
#pragma unroll
        for (int k = 0; k < NW; ++k) {
          x[k] = (lane == 0) ? ga[k] : 0;
          x[NW + k] = (lane == 0) ? gl_[k] : 0;
        }
      } else {
        uint64_t *bufs = a.xchg + (size_t)J.area * kPPTeamWords + (size_t)(e & 1u) * R * kSlotWords * P;
        if (lane < K * NP) {
          const int rep = lane / NP;
          const int k = lane - rep * NP;
          int word = 0;
#pragma unroll
          for (int q = 0; q < NW; ++q) {
            int wa = ga[q], wb = gl_[q];
            asm volatile("" : "+v"(wa), "+v"(wb));
            word = (k == q) ? wa : word;
            word = (k == NW + q) ? wb : word;
          }
          uint32_t poff = (uint32_t)(((rep * kSlotWords + k) * (int)P + (int)r) * sizeof(uint64_t));
          asm volatile("" : "+v"(poff));
          xchg_store(reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(bufs) + poff),
                     ((uint64_t)(uint32_t)word << 32) | e);
        }
        const uint64_t *buf = bufs + (size_t)(r % (uint32_t)K) * kSlotWords * P;
        // The lane's byte offset, formed here from the lane id (a value kept
        // from the kernel's entry would be spilled and reloaded every round).
        uint32_t loff = (uint32_t)__lane_id() * (uint32_t)sizeof(uint64_t);
        if (live) {
          uint64_t w[NP];
          const uint64_t deadline = now_ticks() + a.spin_ticks;
          for (;;) {
            bool all = true;
            asm volatile("" : "+v"(loff));
#pragma unroll
            for (int k = 0; k < NP; ++k)
              w[k] = xchg_load(reinterpret_cast<const uint64_t *>(
                  reinterpret_cast<const char *>(buf + (size_t)k * P) + loff));
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
      }
      const bool valid = live && ok;
      // p1: the best A.
      int v1;
      const int l1 = wave_argmax_lane<TIE>(valid ? x[0] : INT_MIN, (uint32_t)x[1], v1);
      int r1[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) r1[k] = readlane_i(x[k], l1);
      float q1[D], qa[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        q1[j] = __int_as_float(r1[2 + j]);
        qa[j] = __int_as_float(x[2 + j]);
      }
      // Each A after p1, in the update's own arithmetic.
      const float post = min_keep(sqdist<D>(qa, q1), __int_as_float(x[0]));
      const bool moved = valid && lane != l1 && __float_as_int(post) != x[0];
      const bool mine = lane == l1;
      const int tv = mine ? x[NW] : ((valid && !moved) ? x[0] : INT_MIN);
      const uint32_t ti = (uint32_t)(mine ? x[NW + 1] : x[1]);
      int v2;
      const int l2 = wave_argmax_lane<TIE>(tv, ti, v2);
      const uint32_t i2 = (uint32_t)readlane_i((int)ti, l2);
      const bool beaten = moved && !beats<TIE>(v2, i2, x[0], (uint32_t)x[1]);
      const bool two = __ballot(beaten) == 0;
      const bool all_ok = __ballot(live && !ok) == 0;
      sv = valid ? x[0] : INT_MIN;
      si = x[1];
      int r2[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const int cw = mine ? x[NW + 2 + j] : x[2 + j];
        r2[j] = readlane_i(cw, l2);
      }
      if (lane == 0) {
        if constexpr (VEC) {
          sh.res4[par][0] = make_int4(r1[0], r1[1], r1[2], r1[3]);
          sh.res4[par][1] = make_int4(r1[4], v2, (int)i2, r2[0]);
          sh.res4[par][2] = make_int4(r2[1], r2[2], two ? 1 : 0, all_ok ? 1 : 0);
        } else {
#pragma unroll
          for (int k = 0; k < NW; ++k) sh.res[par][k] = r1[k];
          sh.res[par][NW] = v2;
          sh.res[par][NW + 1] = (int)i2;
#pragma unroll
          for (int j = 0; j < D; ++j) sh.res[par][NW + 2 + j] = r2[j];
          sh.res[par][NP] = two ? 1 : 0;
          sh.res[par][NP + 1] = all_ok ? 1 : 0;
        }
      }
    }
    lds_barrier();
    int res[NP + 2];
    if constexpr (VEC) {
      const int4 q0 = sh.res4[par][0], q1 = sh.res4[par][1], q2 = sh.res4[par][2];
      res[0] = q0.x;
      res[1] = q0.y;
      res[2] = q0.z;
      res[3] = q0.w;
      res[4] = q1.x;
      res[5] = q1.y;
      res[6] = q1.z;
      res[7] = q1.w;
      res[8] = q2.x;
      res[9] = q2.y;
      res[10] = q2.z;
      res[11] = q2.w;
    } else {
#pragma unroll
      for (int k = 0; k < NP + 2; ++k) res[k] = sh.res[par][k];
    }
    if (res[NP + 1] == 0) return false;
    i1 = (uint32_t)__builtin_amdgcn_readfirstlane(res[1]);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      p1[j] = __int_as_float(__builtin_amdgcn_readfirstlane(res[2 + j]));
      p2[j] = __int_as_float(__builtin_amdgcn_readfirstlane(res[NW + 2 + j]));
    }
    const uint32_t i2 = (uint32_t)__builtin_amdgcn_readfirstlane(res[NW + 1]);
    const bool two = res[NP] != 0;
    have2 = two && s + 1 < J.s_end;
    if (wave == 0 && J.team > 1) {
      // Members the round's picks certainly leave unchanged keep their A;
      // M = the best of those A's other than this member's.
      float bl[D], bh[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        bl[j] = sh.gbox[j][lane];
        bh[j] = sh.gbox[D + j][lane];
      }
      const float avf = __int_as_float(sv);
      const bool keep = sv != INT_MIN && !(box_bound<D>(p1, bl, bh) < avf) &&
                        !(two && box_bound<D>(p2, bl, bh) < avf);
      int mv;
      const int ml = wave_argmax_lane<TIE>((keep && (uint32_t)lane != r) ? sv : INT_MIN,
                                           (uint32_t)si, mv);
      const int mi = readlane_i(si, ml);
      if (lane == 0) sh.mu[par] = make_int2(mv, mi);
    }
    if (recorder && lane == 0) {
      a.out[J.out_off + s] = i1;
      if (have2) a.out[J.out_off + s + 1] = i2;
    }
    s += have2 ? 2u : 1u;
  }
  if (J.s_end < J.m) {
    // The saved state holds every pick but the last applied (the next chunk
    // starts by applying the last one): a final pair applies its first pick
    // here, and the last pick's coordinates travel with the state.
    // Re-derived here rather than kept live through the loop.
    float *state = a.state + J.state_off;
    const uint32_t nn = J.n;
    const uint32_t wb = r * (uint32_t)PER_WG + (uint32_t)(wave * 64 * PPT);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const uint32_t i = wb + (uint32_t)(k * 64 + lane);
      if (i < nn) state[i] = have2 ? min_keep(sqdist<D>(c[k], p1), dm[k]) : dm[k];
    }
    if (r == 0 && tid < D) {
      const float last = have2 ? (tid == 0 ? p2[0] : tid == 1 ? p2[1] : p2[2])
                               : (tid == 0 ? p1[0] : tid == 1 ? p1[1] : p1[2]);
      state[nn + tid] = last;
    }
  }
  return true;
}

// ONE: one member slot per workgroup (slot = workgroup index; the host
// launches exactly one workgroup per member and all are co-resident, as the
// serving loop's admission guarantees), no claim loop.  Otherwise the work
// queue of fps_batch: workgroups claim member slots until none are left.
// The claim loop keeps its own values live around the paired loop, which then
// no longer fits the 1024 x 16 shape's registers without spilling; the
// serving loop therefore runs the one-slot form.
template <int B, int PPT, int TIE, bool ONE>
__global__ void __launch_bounds__(B) fps_batch_pp(BatchPPParams ap) {
  __shared__ PPShared<B, PPT, TIE, !ONE> sh;
  __shared__ uint32_t s_claim;
  __shared__ int s_ok;
  const BatchParams &a = ap.b;
  const int tid = threadIdx.x;
  for (;;) {
    uint32_t k;
    if (ONE) {
      k = blockIdx.x;
    } else {
      if (tid == 0) s_claim = atomicAdd(a.claim, 1u);
      lds_barrier();
      k = s_claim;
    }
    if (k >= a.total_members) return;
    const uint32_t j = a.member_job[k];
    if (tid == 0) sh.job = a.jobs[j];
    lds_barrier();
    const BatchJob &J = sh.job;
    const uint32_t r = k - J.member_base;
    if (!ONE && J.team > 1) {
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
    if (!run_job_pp<B, PPT, TIE, !ONE>(ap, r, sh)) {
      if (tid == 0) atomicOr(a.status, kStatusTimeout);
      return;
    }
    if (ONE) return;
    lds_barrier();  // every lane has read s_claim / s_ok and the shared state before reuse
  }
}

// Table order: the one-slot kernels (serving), then the claim-loop kernels (batches).
#define FPSX_BATCH_PP(B, P, ONE) \
  {3, B, P, kTieHighest, (const void *)&fps_batch_pp<B, P, kTieHighest, ONE>}

static const KernelEntry kBatchPP[] = {
    FPSX_BATCH_PP(1024, 16, true), FPSX_BATCH_PP(1024, 8, true),
    FPSX_BATCH_PP(1024, 16, false), FPSX_BATCH_PP(1024, 8, false), FPSX_BATCH_PP(1024, 4, false),
};
#undef FPSX_BATCH_PP

const KernelEntry *batch_pp_table(size_t *count) {
  *count = sizeof(kBatchPP) / sizeof(kBatchPP[0]);
  return kBatchPP;
}

const KernelEntry *batch_pp_kernel(int block, int ppt, bool one_slot) {
  const size_t n_one = 2;  // the one-slot kernels lead the table
  const size_t lo = one_slot ? 0 : n_one;
  const size_t hi = one_slot ? n_one : sizeof(kBatchPP) / sizeof(kBatchPP[0]);
  for (size_t i = lo; i < hi; ++i)
    if (kBatchPP[i].block == block && kBatchPP[i].ppt == ppt) return &kBatchPP[i];
  return nullptr;
}

}  // namespace fpsx

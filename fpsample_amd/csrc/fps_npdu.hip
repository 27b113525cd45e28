// fps_npdu.hip -- FPS with nearest-point-distance-updating (NPDU).
//
// Reference: fps_npdu_sampling, src/lib.cpp:272-340.  The start updates every
// point (src/lib.cpp:319-330); every later pick p updates only the window of
// indices [p - w/2, p + w/2], shifted to stay inside [0, n-1]
// (src/lib.cpp:296-309); the next pick is the argmax over ALL points with a
// strict `>` from -1.0f (src/lib.cpp:311-315), i.e. the lowest index among
// equal maxima.  Only a single start exists (src/lib.cpp:360-365).
//
// Layout: every wave owns 64*PPT consecutive indices (lane l holds indices
// wbase + k*64 + l) and keeps their min-distances in registers, so a window
// of w + 1 indices touches at most ceil((w + 1) / (64*PPT)) + 1 waves of the
// whole grid.  A wave outside the window keeps its cached candidate (value,
// index, coordinates) and publishes it again; only the waves under the
// window update their points and reduce anew.
//   D > 0: the coordinates live in registers too (loaded once).
//   D == 0 (any dimension): the coordinates stay in memory and are read only
//   for the points under the window and for the pivot.
#include "fps_device.hpp"
#include "fps_kernels.hpp"
#include "fps_npdu.hpp"

namespace fpsx {

template <int D, int B, int PPT>
__global__ void __launch_bounds__(B) fps_npdu(NpduParams a) {
  constexpr int W = B / 64;
  constexpr int NW = 2 + D;          // value, index, coordinates (D > 0)
  constexpr int DC = D > 0 ? D : 1;  // register arrays of the any-dimension kernel
  __shared__ int s_cand[2][NW][W];
  __shared__ int s_xw[2][NW + 1][W];

  const FpsParams &f = a.fps;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t wbase =
      __builtin_amdgcn_readfirstlane(g * (uint32_t)(B * PPT) + (uint32_t)(wave * 64 * PPT));
  const uint32_t wlast = wbase + (uint32_t)(64 * PPT - 1);
  const uint32_t n = f.n;

  float c[PPT][DC];
  float dm[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t i = wbase + (uint32_t)(k * 64 + lane);
    const bool live = i < n;
#pragma unroll
    for (int j = 0; j < DC; ++j)
      c[k][j] = (D > 0 && live) ? f.pts[(size_t)i * f.sp + (size_t)j * f.ld] : 0.0f;
    dm[k] = live ? f.init : -1.0f;
  }

  uint64_t st = f.starts[0];
  if (st >= n) {
    if (tid == 0) atomicOr(f.status, kStatusBadStart);
    st = 0;
  }
  uint32_t pidx = (uint32_t)st;
  float pv[DC];
#pragma unroll
  for (int j = 0; j < DC; ++j)
    pv[j] = (D > 0) ? f.pts[(size_t)pidx * f.sp + (size_t)j * f.ld] : 0.0f;
  if constexpr (D > 0) {
#pragma unroll
    for (int k = 0; k < PPT; ++k) settle<DC>(c[k]);
    settle<DC>(pv);
  }
  const bool recorder = (g == 0 && tid >= B - 64);
  auto record = [&](uint32_t s) {
    if constexpr (D > 0) {
      record_pick<D>(f, s, pidx, pv);
    } else {
      if (lane == 0) f.out[s] = pidx;
      if (f.out_pts != nullptr)
        for (uint32_t j = lane; j < f.dim; j += 64)
          f.out_pts[(size_t)s * f.dim + j] = f.pts[(size_t)pidx * f.sp + (size_t)j * f.ld];
    }
  };
  if (recorder) record(0);

  // Cached wave candidate (uniform): value bits, index, coordinates.
  int cv = INT_MIN;
  int ci = 0;
  float cc[DC];
#pragma unroll
  for (int j = 0; j < DC; ++j) cc[j] = 0.0f;

  const int64_t hw = (int64_t)(a.window / 2);
  uint32_t e = 0;
  for (uint32_t s = 1; s < f.m; ++s) {
    // Window of the previous pick; the start's update covers every point.
    uint32_t lo = 0, hi = n - 1;
    if (s > 1) {
      int64_t l = (int64_t)pidx - hw;
      int64_t h = (int64_t)pidx + hw;
      if (l < 0) {
        h -= l;
        l = 0;
      }
      if (h >= (int64_t)n) {
        l = l - (h - (int64_t)n + 1);
        if (l < 0) l = 0;
        h = (int64_t)n - 1;
      }
      lo = (uint32_t)l;
      hi = (uint32_t)h;
    }
    if (wbase <= hi && wlast >= lo) {
      // This lane's first index, opaque to the compiler so that the PPT
      // per-slot indices below are formed here and not kept live in
      // registers across the whole pass loop.
      uint32_t ib = wbase + (uint32_t)lane;
      asm volatile("" : "+v"(ib));
      // Slots the window reaches (uniform): k_lo .. k_hi.
      const int k_lo = (int)((lo > wbase ? lo - wbase : 0u) >> 6);
      const int k_hi = (int)((hi - wbase) >> 6);
      if constexpr (D > 0) {
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
          if (k < k_lo || k > k_hi) continue;
          const uint32_t i = ib + (uint32_t)(k * 64);
          const float d = sqdist<D>(c[k], pv);
          dm[k] = (i >= lo && i <= hi) ? min_keep(d, dm[k]) : dm[k];
        }
      } else {
        // Coordinates from memory, only for the slots the window reaches,
        // dimension by dimension in the reference's order (dist = 0;
        // dist += d*d; and 0 + x == x for every x >= 0 and for NaN).
        const float *pp = f.pts + (size_t)pidx * f.sp;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
          if (k < k_lo || k > k_hi) continue;
          const uint32_t i = ib + (uint32_t)(k * 64);
          if (i >= lo && i <= hi) {
            const float *q = f.pts + (size_t)i * f.sp;
            float acc = 0.0f;
#pragma unroll 4
            for (uint32_t j = 0; j < f.dim; ++j) {
              const float dj = q[(size_t)j * f.ld] - pp[(size_t)j * f.ld];
              acc = (j == 0) ? dj * dj : acc + dj * dj;
            }
            dm[k] = min_keep(acc, dm[k]);
          }
        }
      }
      // Lane best with a strict `>`: the lowest k, i.e. the lowest index.
      int bv = INT_MIN;
      int bk = 0;
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        const int b = __float_as_int(dm[k]);
        const bool take = b > bv;
        bv = take ? b : bv;
        bk = take ? k : bk;
      }
      const uint32_t bidx = ib + (uint32_t)(bk * 64);
      int wv;
      const int wl = wave_argmax_lane<kTieLowest>(bv, bidx, wv);
      if constexpr (D > 0) {
        const int kw = readlane_i(bk, wl);
#pragma unroll
        for (int j = 0; j < D; ++j) {
          float r = c[0][j];
#pragma unroll
          for (int q = 1; q < PPT; ++q) r = (kw == q) ? c[q][j] : r;
          cc[j] = readlane_f(r, wl);
        }
      }
      cv = wv;
      ci = readlane_i((int)bidx, wl);
    }
    ++e;
    const int par = (int)(e & 1u);
    if (lane == 0) {
      s_cand[par][0][wave] = cv;
      s_cand[par][1][wave] = ci;
#pragma unroll
      for (int j = 0; j < D; ++j) s_cand[par][2 + j][wave] = __float_as_int(cc[j]);
    }
    lds_barrier();
    const int v2 = (lane < W) ? s_cand[par][0][lane] : INT_MIN;
    const uint32_t i2 = (lane < W) ? (uint32_t)s_cand[par][1][lane] : 0u;
    int gv;
    const int gl = wave_argmax_lane<kTieLowest>(v2, i2, gv);
    if (G == 1) {
      pidx = (uint32_t)readlane_i((int)i2, gl);
#pragma unroll
      for (int j = 0; j < D; ++j) pv[j] = __int_as_float(s_cand[par][2 + j][gl]);
    } else {
      int win[NW];
      if (!exchange<kTieLowest, NW, W>(f.xchg, f.replicas, kPageWords, G, g, e, f.backoff,
                                       f.spin_ticks, s_cand[par], s_xw[par], gl, win)) {
        if (tid == 0) atomicOr(f.status, kStatusTimeout);
        return;
      }
      pidx = (uint32_t)win[1];
#pragma unroll
      for (int j = 0; j < D; ++j) pv[j] = __int_as_float(win[2 + j]);
    }
    if (recorder) record(s);
  }
}

// ---------------------------------------------------------------------------
// Paired form (3-D, 2 <= G <= 64): one cross-workgroup exchange settles up to
// two picks, by the rule of the paired pruned kernel (fps_pruned.hip) with the
// windowed update in place of the full one:
//   every workgroup publishes its best A and its lookahead A' = its best after
//   a hypothetical window update around A (only the waves under that window
//   can change, and only a wave whose best point moves sweeps);
//   p1 = the best A (workgroup w1); w1's best after p1 is A' of w1; another
//   workgroup keeps A unless A lies in p1's window and its value drops there,
//   else it is known only as "below A"; p2 = the best known best when it beats
//   every such bound, else the round yields p1 alone;
//   a workgroup forms A' only when it could be w1: its A beats M, the best A
//   of the other workgroups whose index range misses the last round's
//   windows (those certainly kept their A; the test is exact here).
// Ties go to the lowest index (strict `>`), everywhere.
// ---------------------------------------------------------------------------
template <int B, int PPT>
__global__ void __launch_bounds__(B) fps_npdu_pair(NpduParams a) {
  constexpr int D = 3;
  constexpr int TIE = kTieLowest;
  constexpr int W = B / 64;
  constexpr int NW = 2 + D;   // value, index, coordinates
  constexpr int NP = 2 * NW;  // words published per workgroup: A, then A'
  // Hand-offs as 16-byte records, moved with 128-bit LDS accesses, by
  // exchange parity: per wave its best and "changed" flag, its lookahead
  // best; the round's result (p1, p2, two-pick flag, all-arrived flag); M.
  __shared__ int4 s_top[2][W][2];
  __shared__ int4 s_la[2][W][2];
  __shared__ int4 s_res[2][4];
  __shared__ int2 s_mu[2];

  const FpsParams &f = a.fps;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t n = f.n;
  const uint32_t wbase =
      __builtin_amdgcn_readfirstlane(g * (uint32_t)(B * PPT) + (uint32_t)(wave * 64 * PPT));
  const uint32_t wlast = wbase + (uint32_t)(64 * PPT - 1);
  const bool recorder = (g == 0 && tid >= B - 64);

  // Unconditional loads of a clamped index; padding slots are then overwritten.
  float c[PPT][D];
  float dm[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t i = wbase + (uint32_t)(k * 64 + lane);
    const bool live = i < n;
    const uint32_t ic = live ? i : n - 1u;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const float v = f.pts[(size_t)ic * f.sp + (size_t)j * f.ld];
      c[k][j] = live ? v : 0.0f;
    }
    dm[k] = live ? f.init : -1.0f;
  }
#pragma unroll
  for (int k = 0; k < PPT; ++k) settle<D>(c[k]);

  uint64_t st = f.starts[0];
  if (st >= n) {
    if (tid == 0) atomicOr(f.status, kStatusBadStart);
    st = 0;
  }
  uint32_t i1 = (uint32_t)st;
  float p1[D], p2[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    p1[j] = f.pts[(size_t)i1 * f.sp + (size_t)j * f.ld];
    p2[j] = 0.0f;
  }
  settle<D>(p1);
  if (recorder) record_pick<D>(f, 0, i1, p1);

  // Window of a pick: [idx - w/2, idx + w/2] shifted into [0, n-1]
  // (src/lib.cpp:296-309).
  const int64_t hw = (int64_t)(a.window / 2);
  auto window = [&](uint32_t idx, uint32_t &wl, uint32_t &wh) {
    int64_t l = (int64_t)idx - hw;
    int64_t h = (int64_t)idx + hw;
    if (l < 0) {
      h -= l;
      l = 0;
    }
    if (h >= (int64_t)n) {
      l = l - (h - (int64_t)n + 1);
      if (l < 0) l = 0;
      h = (int64_t)n - 1;
    }
    wl = (uint32_t)l;
    wh = (uint32_t)h;
  };
  // The window's update of this wave's slots; false when it does not reach the wave.
  auto apply = [&](const float(&p)[D], uint32_t wl, uint32_t wh) -> bool {
    if (!(wbase <= wh && wlast >= wl)) return false;
    uint32_t ib = wbase + (uint32_t)lane;
    asm volatile("" : "+v"(ib));
    const int k_lo = (int)((wl > wbase ? wl - wbase : 0u) >> 6);
    const int k_hi = (int)((wh - wbase) >> 6);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      if (k < k_lo || k > k_hi) continue;
      const uint32_t i = ib + (uint32_t)(k * 64);
      const float d = sqdist<D>(c[k], p);
      dm[k] = (i >= wl && i <= wh) ? min_keep(d, dm[k]) : dm[k];
    }
    return true;
  };
  // Best (value, index, coordinates) of this wave over val(k): the lowest
  // slot, then the lowest lane, among equal values, i.e. the lowest index.
  auto wave_best = [&](auto val, int &ov, int &oi, float(&oc)[D]) {
    int bv = INT_MIN;
    int bk = 0;
    float bc[D];
#pragma unroll
    for (int j = 0; j < D; ++j) bc[j] = 0.0f;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int b = __float_as_int(val(k));
      const bool take = b > bv;
      bv = take ? b : bv;
      bk = take ? k : bk;
#pragma unroll
      for (int j = 0; j < D; ++j) bc[j] = take ? c[k][j] : bc[j];
    }
    const uint32_t bidx = wbase + (uint32_t)(bk * 64 + lane);
    int wv;
    const int wl = wave_argmax_lane<TIE>(bv, bidx, wv);
#pragma unroll
    for (int j = 0; j < D; ++j) oc[j] = readlane_f(bc[j], wl);
    ov = wv;
    oi = readlane_i((int)bidx, wl);
  };

  int cv = INT_MIN;  // this wave's best (uniform)
  int ci = 0;
  float cc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) cc[j] = 0.0f;
  bool cached = false;

  int ga[NW], gl_[NW];  // the workgroup's A and A' (uniform), kept while unchanged
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    ga[k] = 0;
    gl_[k] = 0;
  }
  bool la_valid = false;
  int mu_v = INT_MIN;
  uint32_t mu_i = 0;
  bool have2 = false;
  uint32_t w1l = 0, w1h = n - 1u, w2l = 0, w2h = 0;  // the start updates every point
  uint32_t e = 0;
  const int K = min(min((int)f.replicas, 64 / NP), 4);
  if (tid < 2) s_mu[tid] = make_int2(INT_MIN, 0);
  lds_barrier();

  uint32_t s = 1;
  while (s < f.m) {
    bool upd = apply(p1, w1l, w1h);
    if (have2) upd = apply(p2, w2l, w2h) || upd;
    if (upd || !cached) {
      wave_best([&](int k) { return dm[k]; }, cv, ci, cc);
      cached = true;
    }

    ++e;
    const int par = (int)(e & 1u);
    if (lane == 0) {
      s_top[par][wave][0] = make_int4(cv, ci, __float_as_int(cc[0]), __float_as_int(cc[1]));
      s_top[par][wave][1] = make_int4(__float_as_int(cc[2]), upd ? 1 : 0, 0, 0);
    }
    lds_barrier();
    int tw[NW + 1];
    {
      const int4 t0 = (lane < W) ? s_top[par][lane][0] : make_int4(INT_MIN, 0, 0, 0);
      const int4 t1 = (lane < W) ? s_top[par][lane][1] : make_int4(0, 0, 0, 0);
      tw[0] = t0.x;
      tw[1] = t0.y;
      tw[2] = t0.z;
      tw[3] = t0.w;
      tw[4] = t1.x;
      tw[5] = t1.y;
      const int2 mu = s_mu[par ^ 1];
      mu_v = mu.x;
      mu_i = (uint32_t)mu.y;
    }
    const bool changed = e == 1u || __ballot(lane < W && tw[NW] != 0) != 0;
    if (changed) {
      int gv;
      const int wl = wave_argmax_lane<TIE>(tw[0], (uint32_t)tw[1], gv);
#pragma unroll
      for (int k = 0; k < NW; ++k) ga[k] = readlane_i(tw[k], wl);
    }
    const bool needL = (changed || !la_valid) && beats<TIE>(ga[0], (uint32_t)ga[1], mu_v, mu_i);
    if (changed) la_valid = false;
    if (needL) {
      // Lookahead: the window around A.  A wave whose best point lies outside
      // it, or keeps its value under it, still has that best; L = the best of
      // those; only a wave whose best moves and could beat L sweeps.
      la_valid = true;
      float pa[D], tc[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        pa[j] = __int_as_float(ga[2 + j]);
        tc[j] = __int_as_float(tw[2 + j]);
      }
      uint32_t al, ah;
      window((uint32_t)ga[1], al, ah);
      const uint32_t ti = (uint32_t)tw[1];
      const bool inw = ti >= al && ti <= ah;
      const float tpost = min_keep(sqdist<D>(tc, pa), __int_as_float(tw[0]));
      const bool texact = lane < W && (!inw || __float_as_int(tpost) == tw[0]);
      int lbv;
      const int lbl = wave_argmax_lane<TIE>(texact ? tw[0] : INT_MIN, ti, lbv);
      const uint32_t lbi = (uint32_t)readlane_i(tw[1], lbl);
      const uint64_t sweepers = __ballot(lane < W && !texact && beats<TIE>(tw[0], ti, lbv, lbi));
      int lv = cv, li = ci;
      float lc[D];
#pragma unroll
      for (int j = 0; j < D; ++j) lc[j] = cc[j];
      if ((sweepers >> wave) & 1ull) {
        uint32_t ib = wbase + (uint32_t)lane;
        asm volatile("" : "+v"(ib));
        wave_best(
            [&](int k) {
              const uint32_t i = ib + (uint32_t)(k * 64);
              return (i >= al && i <= ah) ? min_keep(sqdist<D>(c[k], pa), dm[k]) : dm[k];
            },
            lv, li, lc);
      }
      if (lane == 0) {
        s_la[par][wave][0] = make_int4(lv, li, __float_as_int(lc[0]), __float_as_int(lc[1]));
        s_la[par][wave][1] = make_int4(__float_as_int(lc[2]), 0, 0, 0);
      }
      lds_barrier();
      int lw[NW];
      {
        const int4 t0 = (lane < W) ? s_la[par][lane][0] : make_int4(INT_MIN, 0, 0, 0);
        const int t1 = (lane < W) ? s_la[par][lane][1].x : 0;
        lw[0] = t0.x;
        lw[1] = t0.y;
        lw[2] = t0.z;
        lw[3] = t0.w;
        lw[4] = t1;
      }
      int lgv;
      const int wl2 = wave_argmax_lane<TIE>(lw[0], (uint32_t)lw[1], lgv);
#pragma unroll
      for (int k = 0; k < NW; ++k) gl_[k] = readlane_i(lw[k], wl2);
    }

    int sv = INT_MIN, si = 0;  // wave 0: this lane's workgroup's A (value, index)
    if (wave == 0) {
      uint64_t *bufs = f.xchg + (size_t)(e & 1u) * f.replicas * kSlotWords * kPageWords;
      if (lane < K * NP) {
        const int rep = lane / NP;
        const int k = lane - rep * NP;
        // Each word goes through an empty asm: a select chain over a local
        // array indexed by the lane would otherwise become a dynamic index.
        int word = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
          int wa = ga[q], wb = gl_[q];
          asm volatile("" : "+v"(wa), "+v"(wb));
          word = (k == q) ? wa : word;
          word = (k == NW + q) ? wb : word;
        }
        uint32_t poff = (uint32_t)(((rep * kSlotWords + k) * kPageWords + (int)g) * sizeof(uint64_t));
        asm volatile("" : "+v"(poff));
        xchg_store(reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(bufs) + poff),
                   ((uint64_t)(uint32_t)word << 32) | e);
      }
      const uint64_t *buf = bufs + (size_t)(g % (uint32_t)K) * kSlotWords * kPageWords;
      uint32_t loff = (uint32_t)lane * (uint32_t)sizeof(uint64_t);
      int x[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) x[k] = 0;
      bool ok = true;
      const bool live = (uint32_t)lane < G;
      if (live) {
        uint64_t w[NP];
        const uint64_t deadline = now_ticks() + f.spin_ticks;
        for (;;) {
          bool all = true;
          asm volatile("" : "+v"(loff));
#pragma unroll
          for (int k = 0; k < NP; ++k)
            w[k] = xchg_load(reinterpret_cast<const uint64_t *>(
                reinterpret_cast<const char *>(buf + (size_t)k * kPageWords) + loff));
#pragma unroll
          for (int k = 0; k < NP; ++k) all = all && ((uint32_t)w[k] == e);
          if (all) break;
          if (now_ticks() > deadline) {
            ok = false;
            break;
          }
          for (uint32_t z = 0; z < f.backoff; ++z) __builtin_amdgcn_s_sleep(1);
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
      float q1[D], qa[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        q1[j] = __int_as_float(r1[2 + j]);
        qa[j] = __int_as_float(x[2 + j]);
      }
      uint32_t ql, qh;
      window((uint32_t)r1[1], ql, qh);
      // Each A after p1's window update, in the update's own arithmetic.
      const uint32_t xi = (uint32_t)x[1];
      const float post = (xi >= ql && xi <= qh) ? min_keep(sqdist<D>(qa, q1), __int_as_float(x[0]))
                                                : __int_as_float(x[0]);
      const bool moved = valid && lane != l1 && __float_as_int(post) != x[0];
      const bool mine = lane == l1;
      const int tv = mine ? x[NW] : ((valid && !moved) ? x[0] : INT_MIN);
      const uint32_t tix = (uint32_t)(mine ? x[NW + 1] : x[1]);
      int v2;
      const int l2 = wave_argmax_lane<TIE>(tv, tix, v2);
      const uint32_t i2 = (uint32_t)readlane_i((int)tix, l2);
      const bool beaten = moved && !beats<TIE>(v2, i2, x[0], xi);
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
        s_res[par][0] = make_int4(r1[0], r1[1], r1[2], r1[3]);
        s_res[par][1] = make_int4(r1[4], v2, (int)i2, r2[0]);
        s_res[par][2] = make_int4(r2[1], r2[2], two ? 1 : 0, all_ok ? 1 : 0);
      }
    }
    lds_barrier();
    int res[NP + 2];
    {
      const int4 q0 = s_res[par][0], q1 = s_res[par][1], q2 = s_res[par][2];
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
    }
    if (res[NP + 1] == 0) {
      if (tid == 0) atomicOr(f.status, kStatusTimeout);
      return;
    }
    i1 = (uint32_t)__builtin_amdgcn_readfirstlane(res[1]);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      p1[j] = __int_as_float(__builtin_amdgcn_readfirstlane(res[2 + j]));
      p2[j] = __int_as_float(__builtin_amdgcn_readfirstlane(res[NW + 2 + j]));
    }
    const uint32_t i2 = (uint32_t)__builtin_amdgcn_readfirstlane(res[NW + 1]);
    const bool two = res[NP] != 0;
    have2 = two && s + 1 < f.m;
    window(i1, w1l, w1h);
    if (two) window(i2, w2l, w2h);
    if (wave == 0) {
      // M: the best A of the other workgroups whose index range misses the
      // windows of this round's picks (formed after the hand-off).
      const uint32_t qlo = (uint32_t)lane * (uint32_t)(B * PPT);
      const uint32_t qhi = qlo + (uint32_t)(B * PPT - 1);
      const bool keep = sv != INT_MIN && (uint32_t)lane != g && !(qlo <= w1h && qhi >= w1l) &&
                        !(two && qlo <= w2h && qhi >= w2l);
      int mv;
      const int ml = wave_argmax_lane<TIE>(keep ? sv : INT_MIN, (uint32_t)si, mv);
      const int mi = readlane_i(si, ml);
      if (lane == 0) s_mu[par] = make_int2(mv, mi);
    }
    if (recorder) {
      record_pick<D>(f, s, i1, p1);
      if (have2) record_pick<D>(f, s + 1, i2, p2);
    }
    s += have2 ? 2u : 1u;
  }
  // The launch's exchange count (one per round), read by fpsx_launch_exchanges.
  if (recorder && lane == 0) f.status[1] = e;
}

// ---------------------------------------------------------------------------
// Streaming form for clouds beyond the register file, any dimension.  The
// min-distances live in HBM (f.dmin, by index) and the cloud is cut into
// blocks of BLK consecutive indices (kNpduBlock, or kNpduBigBlock where a
// workgroup's cache would not fit LDS); block b belongs to workgroup
// b mod G, which caches in LDS each of its blocks' best (value, index).  A
// pass re-reads only the blocks the window reaches (every block for the
// start): in-window min-distances are updated, and the block's best is
// re-taken over all its min-distances.  The winner travels as (value,
// index); every workgroup then re-reads the pivot's coordinates.
// ---------------------------------------------------------------------------
template <int BLK>
__global__ void __launch_bounds__(1024) fps_npdu_stream(NpduParams a) {
  constexpr int B = 1024;
  constexpr int W = B / 64;
  constexpr int NW = 2;
  constexpr int SL = BLK / 64;
  __shared__ int s_cand[2][NW][W];
  __shared__ int s_xw[2][NW + 1][W];
  extern __shared__ int s_dyn[];  // pivot coordinates (dim), block values (nbl), block indices (nbl)

  const FpsParams &f = a.fps;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t n = f.n;
  const uint32_t dim = f.dim;
  const uint32_t nblocks = (n + BLK - 1) / BLK;
  const uint32_t nbl = (nblocks > g) ? (nblocks - g + G - 1) / G : 0u;
  float *s_pv = reinterpret_cast<float *>(s_dyn);
  int *m_v = s_dyn + dim;
  int *m_i = m_v + nbl;

  uint64_t st = f.starts[0];
  if (st >= n) {
    if (tid == 0) atomicOr(f.status, kStatusBadStart);
    st = 0;
  }
  uint32_t pidx = (uint32_t)st;
  for (uint32_t j = tid; j < dim; j += B) s_pv[j] = f.pts[(size_t)pidx * f.sp + (size_t)j * f.ld];
  lds_barrier();
  const bool recorder = (g == 0 && tid >= B - 64);
  auto record = [&](uint32_t s) {
    if (lane == 0) f.out[s] = pidx;
    if (f.out_pts != nullptr)
      for (uint32_t j = lane; j < dim; j += 64) f.out_pts[(size_t)s * dim + j] = s_pv[j];
  };
  if (recorder) record(0);

  const int64_t hw = (int64_t)(a.window / 2);
  const size_t lane_off = (size_t)lane * f.sp;
  uint32_t e = 0;
  for (uint32_t s = 1; s < f.m; ++s) {
    uint32_t lo = 0, hi = n - 1;
    if (s > 1) {
      int64_t l = (int64_t)pidx - hw;
      int64_t h = (int64_t)pidx + hw;
      if (l < 0) {
        h -= l;
        l = 0;
      }
      if (h >= (int64_t)n) {
        l = l - (h - (int64_t)n + 1);
        if (l < 0) l = 0;
        h = (int64_t)n - 1;
      }
      lo = (uint32_t)l;
      hi = (uint32_t)h;
    }
    // Local slots j whose block g + j*G meets [lo, hi].
    const uint32_t blo = lo / BLK, bhi = hi / BLK;
    const uint32_t j_first = (blo > g) ? (blo - g + G - 1) / G : 0u;
    const uint32_t j_end = (bhi >= g) ? min(nbl, (bhi - g) / G + 1) : 0u;
    for (uint32_t j = j_first + (uint32_t)wave; j < j_end; j += W) {
      const uint32_t base = (g + j * G) * (uint32_t)BLK;
      // Every load of the block before any update or store: a store between
      // two slots' loads would hold the later loads behind it (the compiler
      // cannot rule out aliasing), one memory round trip per slot and
      // dimension.  Here a block costs one round trip for its min-distances
      // and one per dimension, each of SL independent loads.
      // In the window: lo <= p <= hi (hi < n, so p is a point of the cloud).
      const uint32_t pl = base + (uint32_t)lane;
      const uint32_t span = hi - lo;
      auto in_window = [&](int k) { return pl + (uint32_t)(k * 64) - lo <= span; };
      float old[SL], acc[SL];
#pragma unroll
      for (int k = 0; k < SL; ++k) {
        const uint32_t p = pl + (uint32_t)(k * 64);
        old[k] = (s == 1) ? f.init : f.dmin[p < n ? p : n - 1];
        acc[k] = 0.0f;
      }
      // dimension by dimension in the reference's order (dist = 0;
      // dist += d*d; and 0 + x == x for every x >= 0 and for NaN); only
      // in-window positions read their coordinates.
      for (uint32_t c = 0; c < dim; ++c) {
        const float pvc = s_pv[c];
        // A running pointer over the slots' rows (base + 64 k + lane) and
        // the window test formed afresh per dimension: both are opaque to
        // the compiler, so that neither 16 per-slot addresses nor 16 masks
        // stay live across the loop (they spill, and address temporaries
        // that reuse the registers of loads in flight add waits).
        const float *rows = f.pts + (size_t)base * f.sp + (size_t)c * f.ld;
        const size_t step = (size_t)64 * f.sp;
        size_t off = lane_off;
        uint32_t plc = pl;
        asm volatile("" : "+v"(plc));
        float x[SL];
#pragma unroll
        for (int k = 0; k < SL; ++k) {
          x[k] = (plc + (uint32_t)(k * 64) - lo <= span) ? rows[off] : 0.0f;
          off += step;
          asm volatile("" : "+v"(off));
        }
#pragma unroll
        for (int k = 0; k < SL; ++k) {
          const float dc = x[k] - pvc;
          acc[k] = (c == 0) ? dc * dc : acc[k] + dc * dc;
        }
      }
      // One wait for the min-distance loads here; otherwise every slot's
      // update below waits for all memory operations, the previous slot's
      // store included.
      settle<SL>(old);
      int bv = INT_MIN;
      uint32_t bi = 0;
#pragma unroll
      for (int k = 0; k < SL; ++k) {
        const uint32_t p = pl + (uint32_t)(k * 64);
        const bool live = p < n;
        float nm = old[k];
        if (in_window(k)) {
          nm = min_keep(acc[k], old[k]);
          if (s == 1 || __float_as_int(nm) != __float_as_int(old[k])) f.dmin[p] = nm;
        }
        // Strict `>` over ascending indices: the lowest index among equals.
        const int b = __float_as_int(nm);
        const bool take = live && b > bv;
        bv = take ? b : bv;
        bi = take ? p : bi;
      }
      int wv;
      const int wl = wave_argmax_lane<kTieLowest>(bv, bi, wv);
      const int wi = readlane_i((int)bi, wl);
      if (lane == 0) {
        m_v[j] = wv;
        m_i[j] = wi;
      }
    }
    lds_barrier();  // the cached bests are current; every read of s_pv is done
    ++e;
    const int par = (int)(e & 1u);
    int bv = INT_MIN;
    uint32_t bi = 0;
    for (uint32_t j = (uint32_t)lane; j < nbl; j += 64u) {
      const int v = m_v[j];
      const uint32_t id = (uint32_t)m_i[j];
      if (beats<kTieLowest>(v, id, bv, bi)) {
        bv = v;
        bi = id;
      }
    }
    int gv;
    const int gl = wave_argmax_lane<kTieLowest>(bv, bi, gv);
    const uint32_t gi = (uint32_t)readlane_i((int)bi, gl);
    if (G == 1) {
      pidx = gi;
    } else {
      if (wave == 0 && lane < NW) s_cand[par][lane][0] = (lane == 0) ? gv : (int)gi;
      // The exchange publishes row 0 from wave 0 (which just wrote it) and
      // ends with a workgroup barrier.
      int win[NW];
      if (!exchange<kTieLowest, NW, W>(f.xchg, f.replicas, kPageWords, G, g, e, f.backoff,
                                       f.spin_ticks, s_cand[par], s_xw[par], 0, win)) {
        if (tid == 0) atomicOr(f.status, kStatusTimeout);
        return;
      }
      pidx = (uint32_t)win[1];
    }
    lds_barrier();  // every wave is done with this pass's cache reads
    for (uint32_t j = tid; j < dim; j += B) s_pv[j] = f.pts[(size_t)pidx * f.sp + (size_t)j * f.ld];
    lds_barrier();
    if (recorder) record(s);
  }
}

// ---------------------------------------------------------------------------
// 3-D streaming form: fps_npdu_stream with each block's best coordinates
// cached in LDS beside its (value, index).  The workgroup's winner then
// travels with its coordinates through the exchange, and no workgroup reads
// the pivot back from memory: one memory round trip and one barrier less per
// pass.  A block update loads the coordinates of every position of the block
// (in the same batch as its min-distances), since its best may lie outside
// the window.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) fps_npdu_stream3(NpduParams a) {
  constexpr int B = 1024;
  constexpr int W = B / 64;
  constexpr int D = 3;
  constexpr int NW = 2 + D;
  constexpr int SL = kNpduBlock / 64;
  __shared__ int s_cand[2][NW][W];
  __shared__ int s_xw[2][NW + 1][W];
  extern __shared__ int s_dyn[];  // block values (nbl), indices (nbl), coordinates (3 nbl)

  const FpsParams &f = a.fps;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t n = f.n;
  const uint32_t nblocks = (n + kNpduBlock - 1) / kNpduBlock;
  const uint32_t nbl = (nblocks > g) ? (nblocks - g + G - 1) / G : 0u;
  int *m_v = s_dyn;
  int *m_i = m_v + nbl;
  int *m_c = m_i + nbl;  // m_c[c * nbl + j]: coordinate bits of block j's best

  uint64_t st = f.starts[0];
  if (st >= n) {
    if (tid == 0) atomicOr(f.status, kStatusBadStart);
    st = 0;
  }
  uint32_t pidx = (uint32_t)st;
  float pv[D];
#pragma unroll
  for (int c = 0; c < D; ++c) pv[c] = f.pts[(size_t)pidx * f.sp + (size_t)c * f.ld];
  settle<D>(pv);
  const bool recorder = (g == 0 && tid >= B - 64);
  if (recorder) record_pick<D>(f, 0, pidx, pv);

  const int64_t hw = (int64_t)(a.window / 2);
  uint32_t e = 0;
  for (uint32_t s = 1; s < f.m; ++s) {
    uint32_t lo = 0, hi = n - 1;
    if (s > 1) {
      int64_t l = (int64_t)pidx - hw;
      int64_t h = (int64_t)pidx + hw;
      if (l < 0) {
        h -= l;
        l = 0;
      }
      if (h >= (int64_t)n) {
        l = l - (h - (int64_t)n + 1);
        if (l < 0) l = 0;
        h = (int64_t)n - 1;
      }
      lo = (uint32_t)l;
      hi = (uint32_t)h;
    }
    const uint32_t blo = lo / kNpduBlock, bhi = hi / kNpduBlock;
    const uint32_t j_first = (blo > g) ? (blo - g + G - 1) / G : 0u;
    const uint32_t j_end = (bhi >= g) ? min(nbl, (bhi - g) / G + 1) : 0u;
    for (uint32_t j = j_first + (uint32_t)wave; j < j_end; j += W) {
      const uint32_t base = (g + j * G) * (uint32_t)kNpduBlock;
      // Every load of the block (min-distances and coordinates, positions
      // past the cloud clamped to the last point and masked out) before any
      // update or store: one memory round trip per block.
      const uint32_t pl = base + (uint32_t)lane;
      const uint32_t span = hi - lo;  // in the window: p - lo <= span (hi < n)
      float old[SL], x[SL][D];
#pragma unroll
      for (int k = 0; k < SL; ++k) {
        const uint32_t p = pl + (uint32_t)(k * 64);
        const uint32_t q = p < n ? p : n - 1;
        old[k] = (s == 1) ? f.init : f.dmin[q];
        const float *pp = f.pts + (size_t)q * f.sp;
#pragma unroll
        for (int c = 0; c < D; ++c) x[k][c] = pp[(size_t)c * f.ld];
      }
      int bv = INT_MIN;
      uint32_t bi = 0;
      float bc[D] = {0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int k = 0; k < SL; ++k) {
        const uint32_t p = pl + (uint32_t)(k * 64);
        const bool live = p < n;
        float nm = old[k];
        if (p - lo <= span) {
          nm = min_keep(sqdist<D>(x[k], pv), old[k]);
          if (s == 1 || __float_as_int(nm) != __float_as_int(old[k])) f.dmin[p] = nm;
        }
        // Strict `>` over ascending indices: the lowest index among equals.
        const int b = __float_as_int(nm);
        const bool take = live && b > bv;
        bv = take ? b : bv;
        bi = take ? p : bi;
#pragma unroll
        for (int c = 0; c < D; ++c) bc[c] = take ? x[k][c] : bc[c];
      }
      int wv;
      const int wl = wave_argmax_lane<kTieLowest>(bv, bi, wv);
      const int wi = readlane_i((int)bi, wl);
      float wc[D];
#pragma unroll
      for (int c = 0; c < D; ++c) wc[c] = readlane_f(bc[c], wl);
      if (lane == 0) {
        m_v[j] = wv;
        m_i[j] = wi;
#pragma unroll
        for (int c = 0; c < D; ++c) m_c[c * nbl + j] = __float_as_int(wc[c]);
      }
    }
    lds_barrier();  // the cached bests are current
    ++e;
    const int par = (int)(e & 1u);
    int bv = INT_MIN;
    uint32_t bi = 0;
    uint32_t bj = 0;
    for (uint32_t j = (uint32_t)lane; j < nbl; j += 64u) {
      const int v = m_v[j];
      const uint32_t id = (uint32_t)m_i[j];
      if (beats<kTieLowest>(v, id, bv, bi)) {
        bv = v;
        bi = id;
        bj = j;
      }
    }
    int gv;
    const int gl = wave_argmax_lane<kTieLowest>(bv, bi, gv);
    const uint32_t wj = (uint32_t)readlane_i((int)bj, gl);
    int win[NW];
    win[0] = gv;
    win[1] = readlane_i((int)bi, gl);
#pragma unroll
    for (int c = 0; c < D; ++c) win[2 + c] = m_c[c * nbl + wj];
    if (G == 1) {
      lds_barrier();  // every wave is done reading the cache before the next update
    } else {
      if (wave == 0 && lane < NW) {
        int w = win[0];
#pragma unroll
        for (int q = 1; q < NW; ++q) w = (lane == q) ? win[q] : w;
        s_cand[par][lane][0] = w;
      }
      // The exchange publishes row 0 of s_cand from wave 0 (which just wrote
      // it) and ends with a workgroup barrier.
      if (!exchange<kTieLowest, NW, W>(f.xchg, f.replicas, kPageWords, G, g, e, f.backoff,
                                       f.spin_ticks, s_cand[par], s_xw[par], 0, win)) {
        if (tid == 0) atomicOr(f.status, kStatusTimeout);
        return;
      }
    }
    pidx = (uint32_t)win[1];
#pragma unroll
    for (int c = 0; c < D; ++c) pv[c] = __int_as_float(win[2 + c]);
    if (recorder) record_pick<D>(f, s, pidx, pv);
  }
}

static const KernelEntry kNpduStream[] = {
    {0, 1024, 0, kTieLowest, (const void *)&fps_npdu_stream<kNpduBlock>},
    {3, 1024, 0, kTieLowest, (const void *)&fps_npdu_stream3},
    {0, 1024, 0, kTieLowest, (const void *)&fps_npdu_stream<kNpduBigBlock>},
};

const KernelEntry *npdu_stream_kernel(bool three_d) { return &kNpduStream[three_d ? 1 : 0]; }
const KernelEntry *npdu_stream_kernel_big() { return &kNpduStream[2]; }

#define FPSX_NPDU(D, B, P) {D, B, P, kTieLowest, (const void *)&fps_npdu<D, B, P>}

static const KernelEntry kNpdu[] = {
    FPSX_NPDU(3, 256, 1),   FPSX_NPDU(3, 256, 2),   FPSX_NPDU(3, 256, 4),   FPSX_NPDU(3, 256, 16),
    FPSX_NPDU(3, 1024, 1),  FPSX_NPDU(3, 1024, 2),  FPSX_NPDU(3, 1024, 4),
    FPSX_NPDU(3, 1024, 8),  FPSX_NPDU(3, 1024, 16), FPSX_NPDU(0, 256, 1),
    FPSX_NPDU(0, 256, 4),   FPSX_NPDU(0, 1024, 1),  FPSX_NPDU(0, 1024, 4),
    FPSX_NPDU(0, 1024, 16), FPSX_NPDU(0, 1024, 32),
};
#undef FPSX_NPDU

const KernelEntry *npdu_table(size_t *count) {
  *count = sizeof(kNpdu) / sizeof(kNpdu[0]);
  return kNpdu;
}

#define FPSX_NPDU_PAIR(B, P) {3, B, P, kTieLowest, (const void *)&fps_npdu_pair<B, P>}

static const KernelEntry kNpduPair[] = {
    FPSX_NPDU_PAIR(1024, 2), FPSX_NPDU_PAIR(1024, 4), FPSX_NPDU_PAIR(1024, 8),
    FPSX_NPDU_PAIR(1024, 16),
};
#undef FPSX_NPDU_PAIR

const KernelEntry *npdu_pair_table(size_t *count) {
  *count = sizeof(kNpduPair) / sizeof(kNpduPair[0]);
  return kNpduPair;
}

}  // namespace fpsx

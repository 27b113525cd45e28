// fps_pruned.hip -- register-resident FPS with per-wave bounding-box pruning.
//
// Same picks as fps_resident (vanilla FPS, reference src/lib.cpp:188-246),
// less work: the points are first put in spatial order (Hilbert indices,
// radix sort), and every wave owns one block of 64*PPT consecutive points of that
// order, re-sorted inside the block by original index.  A wave keeps the
// bounding box of its block.  For a pivot p, the squared distance of every
// point x of the box is >= L(p, box) = sum_j gap_j^2 with gap_j the distance
// from p_j to [lo_j, hi_j], computed in the same float operations and order
// as the point distance; float subtraction, squaring and addition are
// monotone, so the bound holds after rounding.  When L >= the wave's largest
// min-distance no point of the wave can take the `dist < dist_min` branch:
// the wave skips the update and re-publishes its cached candidate.  This is
// the bucket test of the reference's QuickFPS (src/_ext/KDNode.h:105-130),
// applied per wave.
//
// Ties: inside a block, position order is original-index order, so the
// position tie-break of the wave reduction equals the index tie-break; the
// wave's winner position is mapped to its original index through LDS, and
// everything above the wave compares original indices.
#include "fps_device.hpp"
#include "fps_kernels.hpp"
#include "fps_pruned.hpp"

namespace fpsx {

template <int D, int B, int PPT, int TIE>
__global__ void __launch_bounds__(B) fps_pruned(PrunedParams a) {
  constexpr int W = B / 64;
  constexpr int NW = 2 + D;
  constexpr int PER_WG = B * PPT;
  __shared__ int s_cand[2][NW][W];
  __shared__ int s_xw[2][NW + 1][W];
  __shared__ uint32_t s_perm[PER_WG];  // original index of every position of this workgroup

  const FpsParams &f = a.fps;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t base = g * (uint32_t)PER_WG;
  const uint32_t wbase = base + (uint32_t)(wave * 64 * PPT);

  for (uint32_t q = tid; q < PER_WG; q += B)
    s_perm[q] = (base + q < f.n) ? a.perm[base + q] : 0u;

  // This lane's points: positions wbase + k*64 + lane (a contiguous block per wave).
  float c[PPT][D];
  float dm[PPT];
  float lo[D], hi[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    lo[j] = INFINITY;
    hi[j] = -INFINITY;
  }
  // Unconditional loads of a clamped position (a load under a branch would
  // wait for each one in turn); padding slots are then overwritten.
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t i = wbase + (uint32_t)(k * 64 + lane);
    const bool live = i < f.n;
    const uint32_t ic = live ? i : f.n - 1u;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const float v = a.sorted[(size_t)j * a.sorted_ld + ic];
      c[k][j] = live ? v : 0.0f;
      lo[j] = live ? fminf(lo[j], v) : lo[j];
      hi[j] = live ? fmaxf(hi[j], v) : hi[j];
    }
    dm[k] = live ? f.init : -1.0f;
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    lo[j] = wave_min_f(lo[j]);
    hi[j] = wave_max_f(hi[j]);
  }
  lds_barrier();  // s_perm complete

  float pv[D];
  uint32_t pidx = load_forced<D>(f, 0, pv);
  if (g == 0 && tid >= B - 64) record_pick<D>(f, 0, pidx, pv);

  // Cached wave candidate (uniform): value bits, original index, coordinates.
  int cv = INT_MIN;
  int ci = 0;
  float cc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) cc[j] = 0.0f;
  bool cached = false;

  uint32_t e = 0;
  for (uint32_t s = 1; s < f.m; ++s) {
    // Skip test: no point of this wave can move when the box is at least as
    // far from the pivot as the wave's largest min-distance.
    const bool skip = cached && !(box_bound<D>(pv, lo, hi) < __int_as_float(cv));
    if (!skip) {
      int bv = INT_MIN;
      int bk = 0;
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        dm[k] = min_keep(sqdist<D>(c[k], pv), dm[k]);
        const int b = __float_as_int(dm[k]);
        const bool take = (TIE == kTieHighest) ? (b >= bv) : (b > bv);
        bv = take ? b : bv;
        bk = take ? k : bk;
      }
      // Position within the block orders like the original index.
      const uint32_t bpos = (uint32_t)(bk * 64 + lane);
      int wv;
      const int wl = wave_argmax_lane<TIE>(bv, bpos, wv);
      const int kw = readlane_i(bk, wl);
      const uint32_t wpos = (uint32_t)readlane_i((int)bpos, wl);
#pragma unroll
      for (int j = 0; j < D; ++j) {
        float r = c[0][j];
#pragma unroll
        for (int q = 1; q < PPT; ++q) r = (kw == q) ? c[q][j] : r;
        cc[j] = readlane_f(r, wl);
      }
      cv = wv;
      ci = (int)s_perm[wave * 64 * PPT + wpos];
      cached = true;
    }
    if (s < f.n_starts) {
      pidx = load_forced<D>(f, s, pv);
      if (g == 0 && tid >= B - 64) record_pick<D>(f, s, pidx, pv);
      continue;
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
    const int gl = wave_argmax_lane<TIE>(v2, i2, gv);
    if (G == 1) {
      pidx = (uint32_t)readlane_i((int)i2, gl);
#pragma unroll
      for (int j = 0; j < D; ++j) pv[j] = __int_as_float(s_cand[par][2 + j][gl]);
    } else {
      int win[NW];
      if (!exchange<TIE, NW, W>(f.xchg, f.replicas, kPageWords, G, g, e, f.backoff, f.spin_ticks,
                                s_cand[par], s_xw[par], gl, win)) {
        if (tid == 0) atomicOr(f.status, kStatusTimeout);
        return;
      }
      pidx = (uint32_t)win[1];
#pragma unroll
      for (int j = 0; j < D; ++j) pv[j] = __int_as_float(win[2 + j]);
    }
    if (g == 0 && tid >= B - 64) record_pick<D>(f, s, pidx, pv);
  }
}

// ---------------------------------------------------------------------------
// One workgroup, paired at the wave level: the decision of fps_pruned_pair
// with waves in place of workgroups and an LDS barrier in place of the
// exchange.  Every wave keeps its best (value, original index, coordinates)
// and its self-lookahead -- its best after a hypothetical update by its own
// best -- the best recomputed when a pick reaches the wave's box, the
// lookahead too (for 8 and 16 points per thread only while the wave could
// hold the next p1: while its best beats M, the best of the other waves'
// bests the last picks left unmoved).  After one
// barrier every wave reads the W pairs and derives, identically: p1 the best
// wave best; p2 the best of p1's wave's lookahead and the wave bests p1
// leaves unmoved, unless a moved wave best could still beat it (then p1
// alone).  One barrier per pair instead of one per pick; the picks are those
// of the one-pick loop.  G must be 1.
// ---------------------------------------------------------------------------
template <int D, int B, int PPT, int TIE>
__global__ void __launch_bounds__(B) fps_pruned_solo(PrunedParams a) {
  constexpr int W = B / 64;
  constexpr int PER_WG = B * PPT;
  // Skipping lookaheads pays where a wave's sweep is long: 16 K points at 16
  // per thread 4.55 -> 4.25 ms, against 0.74 -> 0.80 ms for 4 K points at 4,
  // where M's reduction costs more than the sweeps it saves.
  constexpr bool kSkip = PPT >= 8;
  static_assert(D == 3, "the records below hold three coordinates");
  static_assert(W <= 64, "one lane per wave in the decision");
  // Per wave and parity: (best value, index, x, y), (z, lookahead value,
  // index, x), (y, z, -, -).
  __shared__ int4 s_w[2][W][3];
  __shared__ uint32_t s_perm[PER_WG];

  const FpsParams &f = a.fps;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t wbase = (uint32_t)(wave * 64 * PPT);
  const bool recorder = tid >= B - 64;

  for (uint32_t q = tid; q < PER_WG; q += B) s_perm[q] = (q < f.n) ? a.perm[q] : 0u;

  float c[PPT][D];
  float dm[PPT];
  float lo[D], hi[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    lo[j] = INFINITY;
    hi[j] = -INFINITY;
  }
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t i = wbase + (uint32_t)(k * 64 + lane);
    const bool live = i < f.n;
    const uint32_t ic = live ? i : f.n - 1u;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const float v = a.sorted[(size_t)j * a.sorted_ld + ic];
      c[k][j] = live ? v : 0.0f;
      lo[j] = live ? fminf(lo[j], v) : lo[j];
      hi[j] = live ? fmaxf(hi[j], v) : hi[j];
    }
    dm[k] = live ? f.init : -1.0f;
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    lo[j] = wave_min_f(lo[j]);
    hi[j] = wave_max_f(hi[j]);
  }
  lds_barrier();  // s_perm complete

  // Best of this wave over val(k) (value, original index, coordinates).
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
    // Position within the block orders like the original index.
    const uint32_t bpos = (uint32_t)(bk * 64 + lane);
    int wv;
    const int wl = wave_argmax_lane<TIE>(bv, bpos, wv);
    const uint32_t wpos = (uint32_t)readlane_i((int)bpos, wl);
#pragma unroll
    for (int j = 0; j < D; ++j) oc[j] = readlane_f(bc[j], wl);
    ov = wv;
    oi = (int)s_perm[wave * 64 * PPT + wpos];
  };

  float p1[D], p2[D];
  uint32_t i1 = load_forced<D>(f, 0, p1);
#pragma unroll
  for (int j = 0; j < D; ++j) p2[j] = 0.0f;
  if (recorder) record_pick<D>(f, 0, i1, p1);

  int cv = INT_MIN, ci = 0;  // this wave's best (uniform)
  float cc[D];
  int lv = INT_MIN, li = 0;  // and its self-lookahead
  float lc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    cc[j] = 0.0f;
    lc[j] = 0.0f;
  }
  bool cached = false;
  bool la_valid = false;  // lv/li/lc is the lookahead of the current state
  int mu_v = INT_MIN;     // M: best of the other waves' bests the last picks left unmoved
  uint32_t mu_i = 0;
  bool have2 = false;
  uint32_t s = 1, e = 0;
  while (s < f.m) {
    // Apply the pending picks where they can reach.
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
      la_valid = false;
    }
    // The lookahead only while this wave could hold the next p1: a wave
    // whose best does not beat M cannot (M's wave keeps its best).
    if (!la_valid && s >= f.n_starts && (!kSkip || beats<TIE>(cv, (uint32_t)ci, mu_v, mu_i))) {
      wave_best([&](int k) { return min_keep(sqdist<D>(c[k], cc), dm[k]); }, lv, li, lc);
      la_valid = true;
    }
    if (s < f.n_starts) {
      // Forced prefix (multi-start): one pick per pass, no decision.
      i1 = load_forced<D>(f, s, p1);
      have2 = false;
      if (recorder) record_pick<D>(f, s, i1, p1);
      ++s;
      continue;
    }
    ++e;
    const int par = (int)(e & 1u);
    if (lane == 0) {
      s_w[par][wave][0] = make_int4(cv, ci, __float_as_int(cc[0]), __float_as_int(cc[1]));
      s_w[par][wave][1] = make_int4(__float_as_int(cc[2]), lv, li, __float_as_int(lc[0]));
      s_w[par][wave][2] = make_int4(__float_as_int(lc[1]), __float_as_int(lc[2]), 0, 0);
    }
    lds_barrier();
    // Every wave: one wave per lane.
    const bool live = lane < W;
    const int r = live ? lane : 0;
    const int4 w0 = s_w[par][r][0], w1 = s_w[par][r][1], w2 = s_w[par][r][2];
    const int xv = live ? w0.x : INT_MIN;
    const uint32_t xi = (uint32_t)w0.y;
    int v1;
    const int l1 = wave_argmax_lane<TIE>(xv, xi, v1);
    i1 = (uint32_t)readlane_i((int)xi, l1);
    const float q1[D] = {__int_as_float(readlane_i(w0.z, l1)), __int_as_float(readlane_i(w0.w, l1)),
                         __int_as_float(readlane_i(w1.x, l1))};
    const float qa[D] = {__int_as_float(w0.z), __int_as_float(w0.w), __int_as_float(w1.x)};
    // Each wave best after p1, in the update's own arithmetic.
    const float post = min_keep(sqdist<D>(qa, q1), __int_as_float(xv));
    const bool moved = live && lane != l1 && __float_as_int(post) != xv;
    const bool mine = lane == l1;
    const int tv = mine ? w1.y : ((live && !moved) ? xv : INT_MIN);
    const uint32_t ti = (uint32_t)(mine ? w1.z : w0.y);
    int v2;
    const int l2 = wave_argmax_lane<TIE>(tv, ti, v2);
    const uint32_t i2 = (uint32_t)readlane_i((int)ti, l2);
    const bool beaten = moved && !beats<TIE>(v2, i2, xv, xi);
    const bool two = __ballot(beaten) == 0;
    const float q2[D] = {__int_as_float(readlane_i(mine ? w1.w : w0.z, l2)),
                         __int_as_float(readlane_i(mine ? w2.x : w0.w, l2)),
                         __int_as_float(readlane_i(mine ? w2.y : w1.x, l2))};
#pragma unroll
    for (int j = 0; j < D; ++j) {
      p1[j] = q1[j];
      p2[j] = q2[j];
    }
    if constexpr (kSkip) {
      // M for the next round, this wave's own: the best of the other waves'
      // bests that keep their value under the picks (so stay their bests).
      const float post2 = min_keep(sqdist<D>(qa, q2), __int_as_float(xv));
      const bool keep = live && __float_as_int(post) == xv && !(two && __float_as_int(post2) != xv);
      int mv;
      const int ml = wave_argmax_lane<TIE>((keep && lane != wave) ? xv : INT_MIN, xi, mv);
      mu_v = mv;
      mu_i = (uint32_t)readlane_i((int)xi, ml);
    }
    have2 = two && s + 1 < f.m;
    if (recorder) {
      record_pick<D>(f, s, i1, p1);
      if (have2) record_pick<D>(f, s + 1, i2, p2);
    }
    s += have2 ? 2u : 1u;
  }
  // The launch's round count in the second status word.
  if (recorder && lane == 0) f.status[1] = e;
}

// ---------------------------------------------------------------------------
// Paired form: one cross-workgroup exchange settles up to two picks.
//
// Every workgroup publishes two candidates: its best point A (as above) and
// its lookahead A' -- its best point after a hypothetical update with A as
// the pivot, computed without storing anything.  From the G pairs every
// workgroup derives, identically:
//   p1 = the best A (workgroup w1), the next pick exactly as before;
//   the state of each workgroup after p1:
//     w1: its best is A'_w1 exactly (its lookahead pivot was p1 itself);
//     w != w1 whose A keeps its value under p1 (min(dmin, dist(A, p1)) ==
//       dmin, the kernel's own arithmetic): its best is still A_w, since
//       every other point of w ranked below A_w and can only go down;
//     any other w: unknown, but every point of w now ranks strictly below
//       A_w (its pre-update key);
//   C = the best of the known bests.  When C ranks above A_w of every
//   unknown workgroup, nothing can beat it: p2 = C, the pick after p1.
//   Otherwise the round yields p1 alone.
// Both picks are then applied (box-pruned) before the next pair is formed.
// The picks are those of the one-pick loop; only the number of exchanges
// changes.  A workgroup none of whose waves changed keeps A and A' from
// the previous round.
// Only w1's lookahead is ever used, so a workgroup computes it only when it
// could be w1.  The workgroups exchange their bounding boxes once; after each
// round every workgroup finds, from the picks and the published A's, the
// workgroups that certainly keep their A (the picks lie at least as far from
// their box as their A's value) and the best such A among the others, M.
// In the next round a workgroup whose A does not beat M cannot be w1 and
// skips the lookahead (its published A' is then stale and marked so; it is
// computed again before it can be used).  The grid must satisfy 2 <= G <= 64
// (one polling wave evaluates all pairs).
// ---------------------------------------------------------------------------
template <int D, int B, int PPT, int TIE, int E>
__global__ void __launch_bounds__(B) fps_pruned_pair(PrunedParams a) {
  constexpr int W = B / 64;
  constexpr int NW = 2 + D;      // value, original index, coordinates
  constexpr int NP = 2 * NW;     // words published per workgroup: A, then A'
  constexpr int R = PPT + E;     // rows per wave: PPT in registers, E in LDS
  constexpr int PER_WG = B * R;
  // LDS hand-offs as 16-byte-aligned records, moved with 128-bit accesses:
  // per wave its best and "changed" flag, its lookahead best; the round's
  // result (p1, p2, two-pick flag, all-arrived flag).  By exchange parity.
  static_assert(D == 3, "the records below hold three coordinates");
  __shared__ int4 s_top[2][W][2];
  __shared__ int4 s_la[2][W][2];
  __shared__ int4 s_res[2][4];
  __shared__ int2 s_mu[2];             // M of the round of this parity
  __shared__ float s_gbox[2 * D][64];  // every workgroup's bounding box
  __shared__ float s_wbox[2 * D][W];   // this workgroup's wave boxes
  __shared__ int s_ok;
  __shared__ uint32_t s_perm[PER_WG];
  // Rows PPT .. R-1 of every wave (E > 0): coordinates and min-distance of a
  // slot in one 16-byte record (x, y, z, min-distance).
  __shared__ float4 s_ex[E > 0 ? E : 1][E > 0 ? W : 1][64];

  const FpsParams &f = a.fps;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t base = g * (uint32_t)PER_WG;
  const uint32_t wbase = base + (uint32_t)(wave * 64 * R);
  const bool recorder = (g == 0 && tid >= B - 64);

  for (uint32_t q = tid; q < PER_WG; q += B)
    s_perm[q] = (base + q < f.n) ? a.perm[base + q] : 0u;

  float c[PPT][D];
  float dm[PPT];
  float lo[D], hi[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    lo[j] = INFINITY;
    hi[j] = -INFINITY;
  }
  // Unconditional loads of a clamped position (a load under a branch would
  // wait for each one in turn); padding slots are then overwritten.
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t i = wbase + (uint32_t)(k * 64 + lane);
    const bool live = i < f.n;
    const uint32_t ic = live ? i : f.n - 1u;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const float v = a.sorted[(size_t)j * a.sorted_ld + ic];
      c[k][j] = live ? v : 0.0f;
      lo[j] = live ? fminf(lo[j], v) : lo[j];
      hi[j] = live ? fmaxf(hi[j], v) : hi[j];
    }
    dm[k] = live ? f.init : -1.0f;
  }
  if constexpr (E > 0) {
#pragma unroll
    for (int x = 0; x < E; ++x) {
      const uint32_t i = wbase + (uint32_t)((PPT + x) * 64 + lane);
      const bool live = i < f.n;
      const uint32_t ic = live ? i : f.n - 1u;
      float v[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        v[j] = a.sorted[(size_t)j * a.sorted_ld + ic];
        lo[j] = live ? fminf(lo[j], v[j]) : lo[j];
        hi[j] = live ? fmaxf(hi[j], v[j]) : hi[j];
        v[j] = live ? v[j] : 0.0f;
      }
      s_ex[x][wave][lane] = make_float4(v[0], v[1], v[2], live ? f.init : -1.0f);
    }
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    lo[j] = wave_min_f(lo[j]);
    hi[j] = wave_max_f(hi[j]);
  }
  lds_barrier();  // s_perm (and the LDS rows) complete

  // LDS rows (E > 0): min-distances under pivot q1 (when u1) and q2 (when u2).
  auto update_ex = [&](bool u1, const float(&q1)[D], bool u2, const float(&q2)[D]) {
    if constexpr (E > 0) {
#pragma unroll
      for (int x = 0; x < E; ++x) {
        const float4 r = s_ex[x][wave][lane];
        const float ce[D] = {r.x, r.y, r.z};
        float d = r.w;
        if (u1) d = min_keep(sqdist<D>(ce, q1), d);
        if (u2) d = min_keep(sqdist<D>(ce, q2), d);
        s_ex[x][wave][lane].w = d;
      }
    }
  };

  // Best (value, original index, coordinates) of this wave over val(k) for
  // the register rows and xval(x, coordinates, stored min-distance) for the
  // LDS rows (after them in position order).  Each lane carries its best slot's coordinates
  // through the scan (static indices only: selecting c[slot] afterwards by a
  // uniform slot number is turned by the compiler into a dynamic index,
  // which puts c in scratch).
  auto wave_best = [&](auto val, auto xval, int &ov, int &oi, float(&oc)[D]) {
    float xc[E > 0 ? E : 1][D];
    float xv[E > 0 ? E : 1];
    if constexpr (E > 0) {
#pragma unroll
      for (int x = 0; x < E; ++x) {
        const float4 r = s_ex[x][wave][lane];
        xc[x][0] = r.x;
        xc[x][1] = r.y;
        xc[x][2] = r.z;
        xv[x] = xval(x, xc[x], r.w);
      }
    }
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
    if constexpr (E > 0) {
#pragma unroll
      for (int x = 0; x < E; ++x) {
        const int b = __float_as_int(xv[x]);
        const bool take = (TIE == kTieHighest) ? (b >= bv) : (b > bv);
        bv = take ? b : bv;
        bk = take ? PPT + x : bk;
#pragma unroll
        for (int j = 0; j < D; ++j) bc[j] = take ? xc[x][j] : bc[j];
      }
    }
    const uint32_t bpos = (uint32_t)(bk * 64 + lane);
    int wv;
    const int wl = wave_argmax_lane<TIE>(bv, bpos, wv);
    const uint32_t wpos = (uint32_t)readlane_i((int)bpos, wl);
#pragma unroll
    for (int j = 0; j < D; ++j) oc[j] = readlane_f(bc[j], wl);
    ov = wv;
    oi = (int)s_perm[wave * 64 * R + wpos];
  };
  // The LDS rows' stored min-distances, for wave_best after an update.
  auto xstored = [&](int, const float(&)[D], float dmx) { return dmx; };

  float p1[D], p2[D];
  uint32_t i1 = load_forced<D>(f, 0, p1);
#pragma unroll
  for (int j = 0; j < D; ++j) p2[j] = 0.0f;
  if (recorder) record_pick<D>(f, 0, i1, p1);

  int cv = INT_MIN;  // this wave's best (uniform)
  int ci = 0;
  float cc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) cc[j] = 0.0f;
  bool cached = false;

  // Forced prefix (multi-start): one pick per pass, no exchange.
  uint32_t s = 1;
  for (; s < f.m && s < f.n_starts; ++s) {
    if (!cached || box_bound<D>(p1, lo, hi) < __int_as_float(cv)) {
#pragma unroll
      for (int k = 0; k < PPT; ++k) dm[k] = min_keep(sqdist<D>(c[k], p1), dm[k]);
      update_ex(true, p1, false, p1);
      wave_best([&](int k) { return dm[k]; }, xstored, cv, ci, cc);
      cached = true;
    }
    i1 = load_forced<D>(f, s, p1);
    if (recorder) record_pick<D>(f, s, i1, p1);
  }

  int ga[NW], gl_[NW];  // the workgroup's A and A' (uniform), kept while unchanged
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    ga[k] = 0;
    gl_[k] = 0;
  }
  bool have2 = false;
  bool la_valid = false;    // gl_ is the lookahead of the current state
  int mu_v = INT_MIN;       // M: best A of the certainly-unchanged others
  uint32_t mu_i = 0;
  uint32_t e = 0;
  // Replicas: at most 4 for pairs (measured on gauss1m: 4 replicas 31.7 ms
  // per launch, 6 replicas 32.4 ms; fewer polls per page do not pay for the
  // extra publishes).
  const int K = min(min((int)f.replicas, 64 / NP), 4);

  // Bounding boxes of all workgroups: one exchange (number 1) of 2*D words.
  if (tid < 2) s_mu[tid] = make_int2(INT_MIN, 0);
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      s_wbox[j][wave] = lo[j];
      s_wbox[D + j][wave] = hi[j];
    }
  }
  lds_barrier();
  ++e;
  if (wave == 0) {
    float bw[2 * D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      bw[j] = (lane < W) ? s_wbox[j][lane] : INFINITY;
      bw[D + j] = (lane < W) ? s_wbox[D + j][lane] : -INFINITY;
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      bw[j] = wave_min_f(bw[j]);
      bw[D + j] = wave_max_f(bw[D + j]);
    }
    uint64_t *bufs = f.xchg + (size_t)(e & 1u) * f.replicas * kSlotWords * kPageWords;
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
      xchg_store(bufs + ((size_t)rep * kSlotWords + k) * kPageWords + g,
                 ((uint64_t)(uint32_t)__float_as_int(word) << 32) | e);
    }
    const uint64_t *buf = bufs + (size_t)(g % (uint32_t)K) * kSlotWords * kPageWords;
    bool ok = true;
    if ((uint32_t)lane < G) {
      uint64_t w[2 * D];
      const uint64_t deadline = now_ticks() + f.spin_ticks;
      for (;;) {
        bool all = true;
#pragma unroll
        for (int k = 0; k < 2 * D; ++k) w[k] = xchg_load(buf + (size_t)k * kPageWords + lane);
#pragma unroll
        for (int k = 0; k < 2 * D; ++k) all = all && ((uint32_t)w[k] == e);
        if (all) break;
        if (now_ticks() > deadline) {
          ok = false;
          break;
        }
        for (uint32_t z = 0; z < f.backoff; ++z) __builtin_amdgcn_s_sleep(1);
      }
#pragma unroll
      for (int k = 0; k < 2 * D; ++k) s_gbox[k][lane] = __int_as_float((int)(w[k] >> 32));
    }
    const bool all_ok = __ballot((uint32_t)lane < G && !ok) == 0;
    if (lane == 0) s_ok = all_ok ? 1 : 0;
  }
  lds_barrier();
  if (s_ok == 0) {
    if (tid == 0) atomicOr(f.status, kStatusTimeout);
    return;
  }

  while (s < f.m) {
    // Apply the pending picks.  A wave whose box is at least as far from a
    // pivot as its largest min-distance cannot change under that pivot.
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
      update_ex(need1, p1, need2, p2);
      wave_best([&](int k) { return dm[k]; }, xstored, cv, ci, cc);
      cached = true;
    }

    ++e;
    const int par = (int)(e & 1u);
    if (lane == 0) {
      s_top[par][wave][0] = make_int4(cv, ci, __float_as_int(cc[0]), __float_as_int(cc[1]));
      s_top[par][wave][1] = make_int4(__float_as_int(cc[2]), (need1 || need2) ? 1 : 0, 0, 0);
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
    }
    // M of the previous round (written by wave 0 before this barrier).
    {
      const int2 mu = s_mu[par ^ 1];
      mu_v = mu.x;
      mu_i = (uint32_t)mu.y;
    }
    // The first round computes A everywhere: after a forced prefix a wave's
    // candidate is cached, but the workgroup's A has never been formed.
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
      // This workgroup could hold the next pick: its lookahead.  A wave
      // whose best point keeps its value under A (the update's own
      // arithmetic) still has that best after the hypothetical update; L =
      // the best of those.  Only a wave whose best moves and could still beat
      // L sweeps its points -- normally just the wave holding A.  Every wave
      // derives the same set from the wave bests in LDS.
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
        wave_best([&](int k) { return min_keep(sqdist<D>(c[k], pa), dm[k]); },
                  [&](int, const float(&ce)[D], float dmx) { return min_keep(sqdist<D>(ce, pa), dmx); },
                  lv, li, lc);
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
    // Exchange: wave 0 publishes the pair to K replicas and polls replica
    // g mod K, one workgroup per lane, all NP words at once.
    if (wave == 0) {
      uint64_t *bufs = f.xchg + (size_t)(e & 1u) * f.replicas * kSlotWords * kPageWords;
      if (lane < K * NP) {
        const int rep = lane / NP;
        const int k = lane - rep * NP;
        // Each word goes through an empty asm: a select chain over a local
        // array indexed by the lane would otherwise become a dynamic index
        // and move the array to scratch memory.
        int word = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
          int wa = ga[q], wb = gl_[q];
          asm volatile("" : "+v"(wa), "+v"(wb));
          word = (k == q) ? wa : word;
          word = (k == NW + q) ? wb : word;
        }
        // Byte offset formed per round (a hoisted 64-bit address would be
        // spilled and reloaded on this path).
        uint32_t poff = (uint32_t)(((rep * kSlotWords + k) * kPageWords + (int)g) * sizeof(uint64_t));
        asm volatile("" : "+v"(poff));
        xchg_store(reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(bufs) + poff),
                   ((uint64_t)(uint32_t)word << 32) | e);
      }
      const uint64_t *buf = bufs + (size_t)(g % (uint32_t)K) * kSlotWords * kPageWords;
      // The lane offset is re-laundered on every poll so the NP per-lane
      // addresses are formed just before their loads instead of being
      // hoisted out of the retry loop (20 registers held across it).
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
    // Uniform: keep the picks in scalar registers.
    i1 = (uint32_t)__builtin_amdgcn_readfirstlane(res[1]);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      p1[j] = __int_as_float(__builtin_amdgcn_readfirstlane(res[2 + j]));
      p2[j] = __int_as_float(__builtin_amdgcn_readfirstlane(res[NW + 2 + j]));
    }
    const uint32_t i2 = (uint32_t)__builtin_amdgcn_readfirstlane(res[NW + 1]);
    const bool two = res[NP] != 0;
    have2 = two && s + 1 < f.m;
    if (wave == 0) {
      // Workgroups the round's picks certainly leave unchanged keep their A;
      // M = the best of those A's other than this workgroup's.  Formed after
      // the hand-off, while the other waves start the next round, and read
      // by every wave after the next round's first barrier.
      float bl[D], bh[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        bl[j] = s_gbox[j][lane];
        bh[j] = s_gbox[D + j][lane];
      }
      const float avf = __int_as_float(sv);
      const bool keep = sv != INT_MIN && !(box_bound<D>(p1, bl, bh) < avf) &&
                        !(two && box_bound<D>(p2, bl, bh) < avf);
      int mv;
      const int ml = wave_argmax_lane<TIE>((keep && (uint32_t)lane != g) ? sv : INT_MIN,
                                           (uint32_t)si, mv);
      const int mi = readlane_i(si, ml);
      if (lane == 0) s_mu[par] = make_int2(mv, mi);
    }
    if (recorder) {
      record_pick<D>(f, s, i1, p1);
      if (have2) record_pick<D>(f, s + 1, i2, p2);
    }
    s += have2 ? 2u : 1u;
  }
  // The launch's exchange count (the box exchange included) in the second
  // status word: how many rounds the pairing saved is data-dependent.
  if (recorder && lane == 0) f.status[1] = e;
}

// Paired form beyond 64 workgroups: PW waves poll the exchange, 64
// workgroups each (G <= 64 * PW).  Each evaluates its workgroups' pairs and
// the waves combine partial results through LDS in three steps (the best A;
// the best pick known after p1; whether any workgroup whose A moved could
// beat it) -- two LDS barriers more per round than fps_pruned_pair, whose
// single polling wave evaluates the round alone.  Otherwise the same kernel.
template <int D, int B, int PPT, int TIE, int PW>
__global__ void __launch_bounds__(B) fps_pruned_pair_wide(PrunedParams a) {
  constexpr int W = B / 64;
  constexpr int NW = 2 + D;      // value, original index, coordinates
  constexpr int NP = 2 * NW;     // words published per workgroup: A, then A'
  constexpr int PER_WG = B * PPT;
  static_assert(PW >= 2 && PW <= W, "polling waves");
  __shared__ int s_top[2][NW + 1][W];  // wave bests + "changed" flags, by exchange parity
  __shared__ int s_la[2][NW][W];       // wave lookahead bests
  __shared__ int s_mu[2][PW][2];       // M of the round of this parity (partial per polling wave)
  __shared__ float s_gbox[2 * D][64 * PW];  // every workgroup's bounding box
  __shared__ float s_wbox[2 * D][W];   // this workgroup's wave boxes
  __shared__ int s_okw[PW];            // box exchange: every workgroup arrived, per polling wave
  // The polling waves' partial results of a round.
  __shared__ int s_pa[2][PW][NW + 1];  // best A: words, then its workgroup
  __shared__ int s_pb[2][PW][NW + 2];  // best known pick after p1 (value, index, coordinates),
                                       // then the best A that moved under p1 (value, index)
  __shared__ int s_pf[2][PW];          // every candidate arrived
  __shared__ uint32_t s_perm[PER_WG];

  const FpsParams &f = a.fps;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // Wave-uniform in a scalar register: LDS addresses indexed by the wave are
  // then scalar too, rather than vector values kept (and spilled) across the loop.
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t base = g * (uint32_t)PER_WG;
  const uint32_t wbase = base + (uint32_t)(wave * 64 * PPT);
  const bool recorder = (g == 0 && tid >= B - 64);

  for (uint32_t q = tid; q < PER_WG; q += B)
    s_perm[q] = (base + q < f.n) ? a.perm[base + q] : 0u;

  float c[PPT][D];
  float dm[PPT];
  float lo[D], hi[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    lo[j] = INFINITY;
    hi[j] = -INFINITY;
  }
  // Unconditional loads of a clamped position (a load under a branch would
  // wait for each one in turn); padding slots are then overwritten.
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t i = wbase + (uint32_t)(k * 64 + lane);
    const bool live = i < f.n;
    const uint32_t ic = live ? i : f.n - 1u;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const float v = a.sorted[(size_t)j * a.sorted_ld + ic];
      c[k][j] = live ? v : 0.0f;
      lo[j] = live ? fminf(lo[j], v) : lo[j];
      hi[j] = live ? fmaxf(hi[j], v) : hi[j];
    }
    dm[k] = live ? f.init : -1.0f;
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    lo[j] = wave_min_f(lo[j]);
    hi[j] = wave_max_f(hi[j]);
  }
  lds_barrier();  // s_perm complete

  // Best (value, original index, coordinates) of this wave over val(k).
  // Each lane carries its best slot's coordinates through the scan (static
  // indices only: selecting c[slot] afterwards by a uniform slot number is
  // turned by the compiler into a dynamic index, which puts c in scratch).
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
    oi = (int)s_perm[wave * 64 * PPT + wpos];
  };

  float p1[D], p2[D];
  uint32_t i1 = load_forced<D>(f, 0, p1);
#pragma unroll
  for (int j = 0; j < D; ++j) p2[j] = 0.0f;
  if (recorder) record_pick<D>(f, 0, i1, p1);

  int cv = INT_MIN;  // this wave's best (uniform)
  int ci = 0;
  float cc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) cc[j] = 0.0f;
  bool cached = false;

  // Forced prefix (multi-start): one pick per pass, no exchange.
  uint32_t s = 1;
  for (; s < f.m && s < f.n_starts; ++s) {
    if (!cached || box_bound<D>(p1, lo, hi) < __int_as_float(cv)) {
#pragma unroll
      for (int k = 0; k < PPT; ++k) dm[k] = min_keep(sqdist<D>(c[k], p1), dm[k]);
      wave_best([&](int k) { return dm[k]; }, cv, ci, cc);
      cached = true;
    }
    i1 = load_forced<D>(f, s, p1);
    if (recorder) record_pick<D>(f, s, i1, p1);
  }

  int ga[NW], gl_[NW];  // the workgroup's A and A' (uniform), kept while unchanged
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    ga[k] = 0;
    gl_[k] = 0;
  }
  bool have2 = false;
  bool la_valid = false;    // gl_ is the lookahead of the current state
  int mu_v = INT_MIN;       // M: best A of the certainly-unchanged others
  uint32_t mu_i = 0;
  uint32_t e = 0;
  // Replicas: at most 4 for pairs (measured on gauss1m: 4 replicas 31.7 ms
  // per launch, 6 replicas 32.4 ms; fewer polls per page do not pay for the
  // extra publishes).
  const int K = min(min((int)f.replicas, 64 / NP), 4);

  // Bounding boxes of all workgroups: one exchange (number 1) of 2*D words.
  if (tid < 2 * PW) {
    s_mu[tid / PW][tid % PW][0] = INT_MIN;
    s_mu[tid / PW][tid % PW][1] = 0;
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      s_wbox[j][wave] = lo[j];
      s_wbox[D + j][wave] = hi[j];
    }
  }
  // Several polling waves: wave 0 publishes, wave p polls workgroups
  // 64p .. 64p + 63.
  lds_barrier();
  ++e;
  uint64_t *bufs = f.xchg + (size_t)(e & 1u) * f.replicas * kSlotWords * kPageWords;
  if (wave == 0) {
    float bw[2 * D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      bw[j] = (lane < W) ? s_wbox[j][lane] : INFINITY;
      bw[D + j] = (lane < W) ? s_wbox[D + j][lane] : -INFINITY;
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      bw[j] = wave_min_f(bw[j]);
      bw[D + j] = wave_max_f(bw[D + j]);
    }
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
      xchg_store(bufs + ((size_t)rep * kSlotWords + k) * kPageWords + g,
                 ((uint64_t)(uint32_t)__float_as_int(word) << 32) | e);
    }
  }
  if (wave < PW) {
    const uint64_t *buf = bufs + (size_t)(g % (uint32_t)K) * kSlotWords * kPageWords;
    const uint32_t q = (uint32_t)(wave * 64 + lane);
    bool ok = true;
    if (q < G) {
      uint64_t w[2 * D];
      const uint64_t deadline = now_ticks() + f.spin_ticks;
      for (;;) {
        bool all = true;
#pragma unroll
        for (int k = 0; k < 2 * D; ++k) w[k] = xchg_load(buf + (size_t)k * kPageWords + q);
#pragma unroll
        for (int k = 0; k < 2 * D; ++k) all = all && ((uint32_t)w[k] == e);
        if (all) break;
        if (now_ticks() > deadline) {
          ok = false;
          break;
        }
        for (uint32_t z = 0; z < f.backoff; ++z) __builtin_amdgcn_s_sleep(1);
      }
#pragma unroll
      for (int k = 0; k < 2 * D; ++k) s_gbox[k][q] = __int_as_float((int)(w[k] >> 32));
    }
    const bool all_ok = __ballot(q < G && !ok) == 0;
    if (lane == 0) s_okw[wave] = all_ok ? 1 : 0;
  }
  lds_barrier();
  bool all_arrived = true;
#pragma unroll
  for (int p = 0; p < PW; ++p) all_arrived = all_arrived && s_okw[p] != 0;
  if (!all_arrived) {
    if (tid == 0) atomicOr(f.status, kStatusTimeout);
    return;
  }

  while (s < f.m) {
    // Apply the pending picks.  A wave whose box is at least as far from a
    // pivot as its largest min-distance cannot change under that pivot.
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
    if (lane == 0) {
      s_top[par][0][wave] = cv;
      s_top[par][1][wave] = ci;
#pragma unroll
      for (int j = 0; j < D; ++j) s_top[par][2 + j][wave] = __float_as_int(cc[j]);
      s_top[par][NW][wave] = (need1 || need2) ? 1 : 0;
    }
    lds_barrier();
    int tw[NW + 1];
#pragma unroll
    for (int k = 0; k <= NW; ++k) tw[k] = (lane < W) ? s_top[par][k][lane] : 0;
    if (lane >= W) tw[0] = INT_MIN;
    // M of the previous round (written by the polling waves before this barrier).
    const int pv = (lane < PW) ? s_mu[par ^ 1][lane % PW][0] : INT_MIN;
    const int pi = (lane < PW) ? s_mu[par ^ 1][lane % PW][1] : 0;
    int mv;
    const int ml = wave_argmax_lane<TIE>(pv, (uint32_t)pi, mv);
    mu_v = mv;
    mu_i = (uint32_t)readlane_i(pi, ml);
    // The first round computes A everywhere: after a forced prefix a wave's
    // candidate is cached, but the workgroup's A has never been formed.
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
      // This workgroup could hold the next pick: its lookahead.  A wave
      // whose best point keeps its value under A (the update's own
      // arithmetic) still has that best after the hypothetical update; L =
      // the best of those.  Only a wave whose best moves and could still beat
      // L sweeps its points -- normally just the wave holding A.  Every wave
      // derives the same set from the wave bests in LDS.
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
      if (lane == 0) {
        s_la[par][0][wave] = lv;
        s_la[par][1][wave] = li;
#pragma unroll
        for (int j = 0; j < D; ++j) s_la[par][2 + j][wave] = __float_as_int(lc[j]);
      }
      lds_barrier();
      int lw[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) lw[k] = (lane < W) ? s_la[par][k][lane] : 0;
      if (lane >= W) lw[0] = INT_MIN;
      int lgv;
      const int wl2 = wave_argmax_lane<TIE>(lw[0], (uint32_t)lw[1], lgv);
#pragma unroll
      for (int k = 0; k < NW; ++k) gl_[k] = readlane_i(lw[k], wl2);
    }

    int sv = INT_MIN, si = 0;  // polling waves: this lane's workgroup's A (value, index)
    uint32_t i2 = 0;  // the round's second pick (valid when two)
    bool two = false;
    // Several polling waves.  Wave 0 publishes; wave p polls workgroups
    // 64p .. 64p + 63 (lane -> workgroup q) and the waves combine their
    // partial results through LDS.
    int x[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) x[k] = 0;
    const uint32_t q = (uint32_t)(wave * 64 + lane);
    const bool live = wave < PW && q < G;
    bool valid = false;
    uint64_t *bufs = f.xchg + (size_t)(e & 1u) * f.replicas * kSlotWords * kPageWords;
    if (wave == 0 && lane < K * NP) {
      const int rep = lane / NP;
      const int k = lane - rep * NP;
      int word = 0;
#pragma unroll
      for (int qq = 0; qq < NW; ++qq) {
        int wa = ga[qq], wb = gl_[qq];
        asm volatile("" : "+v"(wa), "+v"(wb));
        word = (k == qq) ? wa : word;
        word = (k == NW + qq) ? wb : word;
      }
      uint32_t poff = (uint32_t)(((rep * kSlotWords + k) * kPageWords + (int)g) * sizeof(uint64_t));
      asm volatile("" : "+v"(poff));
      xchg_store(reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(bufs) + poff),
                 ((uint64_t)(uint32_t)word << 32) | e);
    }
    if (wave < PW) {
      // Step 1: poll, then this wave's best A.
      const uint64_t *buf = bufs + (size_t)(g % (uint32_t)K) * kSlotWords * kPageWords;
      uint32_t loff = q * (uint32_t)sizeof(uint64_t);
      bool ok = true;
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
      valid = live && ok;
      int v1;
      const int l1 = wave_argmax_lane<TIE>(valid ? x[0] : INT_MIN, (uint32_t)x[1], v1);
      int r1[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) r1[k] = readlane_i(x[k], l1);
      const bool wave_ok = __ballot(live && !ok) == 0;
      if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NW; ++k) s_pa[par][wave][k] = r1[k];
        s_pa[par][wave][0] = v1;  // INT_MIN when the wave polled no workgroup
        s_pa[par][wave][NW] = wave * 64 + l1;
        s_pf[par][wave] = wave_ok ? 1 : 0;
      }
    }
    lds_barrier();
    if (wave < PW) {
      // Step 2: p1 over the partial bests, then each A after p1 (the
      // update's own arithmetic), this wave's best known pick after p1 and
      // the best of its A's that moved (an unknown workgroup's best: the
      // round yields two picks only when no such A beats the known best).
      // (Lane p holds partial p's words: all loads issued at once, the
      // winner's words then taken by readlane.)
      int pa[NW + 1];
#pragma unroll
      for (int k = 0; k <= NW; ++k) pa[k] = s_pa[par][lane % PW][k];
      if (lane >= PW) pa[0] = INT_MIN;
      int t1;
      const int lw = wave_argmax_lane<TIE>(pa[0], (uint32_t)pa[1], t1);
      const uint32_t w1 = (uint32_t)readlane_i(pa[NW], lw);
      float q1[D], qa[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        q1[j] = __int_as_float(readlane_i(pa[2 + j], lw));
        qa[j] = __int_as_float(x[2 + j]);
      }
      const float post = min_keep(sqdist<D>(qa, q1), __int_as_float(x[0]));
      const bool moved = valid && q != w1 && __float_as_int(post) != x[0];
      const bool mine = valid && q == w1;
      const int tv = mine ? x[NW] : ((valid && !moved) ? x[0] : INT_MIN);
      const uint32_t ti = (uint32_t)(mine ? x[NW + 1] : x[1]);
      int v2;
      const int l2 = wave_argmax_lane<TIE>(tv, ti, v2);
      const int i2w = readlane_i((int)ti, l2);
      int r2[D];
#pragma unroll
      for (int j = 0; j < D; ++j) r2[j] = readlane_i(mine ? x[NW + 2 + j] : x[2 + j], l2);
      int vm;
      const int lm = wave_argmax_lane<TIE>(moved ? x[0] : INT_MIN, (uint32_t)x[1], vm);
      const int im = readlane_i(x[1], lm);
      if (lane == 0) {
        s_pb[par][wave][0] = v2;
        s_pb[par][wave][1] = i2w;
#pragma unroll
        for (int j = 0; j < D; ++j) s_pb[par][wave][2 + j] = r2[j];
        s_pb[par][wave][NW] = vm;
        s_pb[par][wave][NW + 1] = im;
      }
      sv = valid ? x[0] : INT_MIN;
      si = x[1];
    }
    lds_barrier();
    // Every wave: the round's result from the partials.
    bool all_ok = true;
#pragma unroll
    for (int p = 0; p < PW; ++p) all_ok = all_ok && s_pf[par][p] != 0;
    if (!all_ok) {
      if (tid == 0) atomicOr(f.status, kStatusTimeout);
      return;
    }
    {
      // Lane p holds partial p's words (loads issued at once).
      int pa[NW], pb[NW + 2];
#pragma unroll
      for (int k = 0; k < NW; ++k) pa[k] = s_pa[par][lane % PW][k];
#pragma unroll
      for (int k = 0; k < NW + 2; ++k) pb[k] = s_pb[par][lane % PW][k];
      if (lane >= PW) {
        pa[0] = INT_MIN;
        pb[0] = INT_MIN;
        pb[NW] = INT_MIN;
      }
      int t1;
      const int lw = wave_argmax_lane<TIE>(pa[0], (uint32_t)pa[1], t1);
      i1 = (uint32_t)readlane_i(pa[1], lw);
#pragma unroll
      for (int j = 0; j < D; ++j) p1[j] = __int_as_float(readlane_i(pa[2 + j], lw));
      int v2;
      const int lb = wave_argmax_lane<TIE>(pb[0], (uint32_t)pb[1], v2);
      i2 = (uint32_t)readlane_i(pb[1], lb);
#pragma unroll
      for (int j = 0; j < D; ++j) p2[j] = __int_as_float(readlane_i(pb[2 + j], lb));
      int vmax;
      const int lmax = wave_argmax_lane<TIE>(pb[NW], (uint32_t)pb[NW + 1], vmax);
      const uint32_t imax = (uint32_t)readlane_i(pb[NW + 1], lmax);
      // p2 is the next pick unless some moved A beats it (an unknown
      // workgroup may then hold a better point).
      two = vmax == INT_MIN || beats<TIE>(v2, i2, vmax, imax);
    }
    have2 = two && s + 1 < f.m;
    if (wave < PW) {
      // This wave's part of M (workgroups 64p .. 64p + 63), read after the
      // next round's first barrier.
      float bl[D], bh[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        bl[j] = s_gbox[j][q];
        bh[j] = s_gbox[D + j][q];
      }
      const float avf = __int_as_float(sv);
      const bool keep = sv != INT_MIN && !(box_bound<D>(p1, bl, bh) < avf) &&
                        !(two && box_bound<D>(p2, bl, bh) < avf);
      int mv;
      const int ml = wave_argmax_lane<TIE>((keep && q != g) ? sv : INT_MIN, (uint32_t)si, mv);
      const int mi = readlane_i(si, ml);
      if (lane == 0) {
        s_mu[par][wave][0] = mv;
        s_mu[par][wave][1] = mi;
      }
    }
    if (recorder) {
      record_pick<D>(f, s, i1, p1);
      if (have2) record_pick<D>(f, s + 1, i2, p2);
    }
    s += have2 ? 2u : 1u;
  }
  // The launch's exchange count (the box exchange included) in the second
  // status word: how many rounds the pairing saved is data-dependent.
  if (recorder && lane == 0) f.status[1] = e;
}

// ---------------------------------------------------------------------------
// Streaming form for clouds beyond the register file.  The points stay in
// HBM in the same spatial order, cut into blocks of kSBlock positions (each
// block re-sorted by original index); block b belongs to workgroup b mod G,
// local slot j = b / G.  Per local block the workgroup keeps in LDS its
// bounding box and its cached best candidate (value, original index,
// coordinates).  A pass tests the boxes (one block per lane, 64 consecutive
// slots per wave), lists the blocks whose box lies closer to the pivot than
// their cached best, and its waves take the listed blocks in turn: for each,
// coordinates, original indices and min-distances are read (all loads of a
// block at once), changed min-distances written back, and the cached
// candidate recomputed.  Every other block cannot change.  The first pass streams
// every block and records the boxes.  Ties: within a block, position order
// is original-index order; across blocks the original index decides.
// ---------------------------------------------------------------------------
template <int TIE>
__global__ void __launch_bounds__(1024) fps_stream_pruned(PrunedParams a) {
  constexpr int B = 1024;
  constexpr int W = B / 64;
  constexpr int D = 3;
  constexpr int NW = 2 + D;
  constexpr int SL = kSBlock / 64;  // slots per lane in a block
  __shared__ int s_cand[2][NW][W];
  __shared__ int s_xw[2][NW + 1][W];
  extern __shared__ int s_meta[];  // kSMetaWords rows of nbl words, then the block list
  __shared__ int s_count[2];       // listed blocks, by pass parity

  const FpsParams &f = a.fps;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t n = f.n;
  const uint32_t nblocks = (n + kSBlock - 1) / kSBlock;
  const uint32_t nbl = (nblocks > g) ? (nblocks - g + G - 1) / G : 0u;
  float *m_lo = reinterpret_cast<float *>(s_meta);  // m_lo[c * nbl + j]
  float *m_hi = m_lo + 3 * nbl;
  int *m_v = reinterpret_cast<int *>(m_hi + 3 * nbl);
  int *m_i = m_v + nbl;
  int *m_c = m_i + nbl;  // m_c[c * nbl + j]: coordinate bits of the block's best
  int *s_list = m_c + 3 * nbl;  // this pass's blocks to stream
  if (tid < 2) s_count[tid] = 0;
  lds_barrier();
  const float *xs = a.sorted;
  const float *ys = a.sorted + a.sorted_ld;
  const float *zs = a.sorted + 2 * a.sorted_ld;

  float pv[D];
  uint32_t pidx = load_forced<D>(f, 0, pv);
  if (g == 0 && tid >= B - 64) record_pick<D>(f, 0, pidx, pv);

  uint32_t e = 0;
  for (uint32_t s = 1; s < f.m; ++s) {
    const int sp = (int)(s & 1u);
    // This wave's min-distance stores of the previous pass complete before
    // the barrier below, after which any wave may read those blocks.
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    // Box tests: wave w takes slots [64 w, 64 w + 64), [64 w + 64 W, ...).
    for (uint32_t j0 = (uint32_t)(wave * 64); j0 < nbl; j0 += 64u * W) {
      const uint32_t jl = j0 + (uint32_t)lane;
      bool need = false;
      if (jl < nbl) {
        if (s == 1) {
          need = true;
        } else {
          const float lo[D] = {m_lo[jl], m_lo[nbl + jl], m_lo[2 * nbl + jl]};
          const float hi[D] = {m_hi[jl], m_hi[nbl + jl], m_hi[2 * nbl + jl]};
          need = box_bound<D>(pv, lo, hi) < __int_as_float(m_v[jl]);
        }
      }
      const uint64_t mask = __ballot(need);
      if (mask != 0) {
        int at = 0;
        if (lane == 0) at = atomicAdd(&s_count[sp], __popcll(mask));
        at = readlane_i(at, 0);
        if (need) s_list[at + __popcll(mask & ((1ull << lane) - 1))] = (int)jl;
      }
    }
    lds_barrier();  // the list is complete
    if (tid == 0) s_count[sp ^ 1] = 0;  // next pass's counter; nobody reads it before then
    const int listed = s_count[sp];
    // Update: the waves take the listed blocks in turn.
    for (int t = wave; t < listed; t += W) {
      {
        const uint32_t j = (uint32_t)s_list[t];
        const uint32_t base = (g + j * G) * (uint32_t)kSBlock;
        int bv = INT_MIN;
        uint32_t bi = 0;
        float bc[D] = {0.0f, 0.0f, 0.0f};
        float lo[D] = {INFINITY, INFINITY, INFINITY};
        float hi[D] = {-INFINITY, -INFINITY, -INFINITY};
        // Every load of the block first, then the updates and their stores:
        // a store between two slots' loads would keep the later loads behind
        // it (the compiler cannot rule out aliasing), one memory round trip
        // per slot.  Loads are unconditional (a position past the cloud reads
        // the last point and is masked out).
        float bx[SL][D], bo[SL];
        uint32_t bid[SL];
#pragma unroll
        for (int k = 0; k < SL; ++k) {
          const uint32_t p = base + (uint32_t)(k * 64 + lane);
          const uint32_t q = p < n ? p : n - 1;
          bx[k][0] = xs[q];
          bx[k][1] = ys[q];
          bx[k][2] = zs[q];
          bid[k] = a.perm[q];
          bo[k] = (s == 1) ? f.init : f.dmin[q];
        }
#pragma unroll
        for (int k = 0; k < SL; ++k) {
          const uint32_t p = base + (uint32_t)(k * 64 + lane);
          const bool live = p < n;
          const float(&c)[D] = bx[k];
          const uint32_t id = bid[k];
          const float old = bo[k];
          const float nm = min_keep(sqdist<D>(c, pv), old);
          if (live && (s == 1 || __float_as_int(nm) != __float_as_int(old))) f.dmin[p] = nm;
          const int b = __float_as_int(nm);
          const bool take = live && ((TIE == kTieHighest) ? (b >= bv) : (b > bv));
          bv = take ? b : bv;
          bi = take ? id : bi;
#pragma unroll
          for (int r = 0; r < D; ++r) bc[r] = take ? c[r] : bc[r];
          if (s == 1 && live) {
#pragma unroll
            for (int r = 0; r < D; ++r) {
              lo[r] = fminf(lo[r], c[r]);
              hi[r] = fmaxf(hi[r], c[r]);
            }
          }
        }
        int wv;
        const int wl = wave_argmax_lane<TIE>(bv, bi, wv);
        const int wi = readlane_i((int)bi, wl);
        float wc[D];
#pragma unroll
        for (int q = 0; q < D; ++q) wc[q] = readlane_f(bc[q], wl);
        if (s == 1) {
#pragma unroll
          for (int q = 0; q < D; ++q) {
            lo[q] = wave_min_f(lo[q]);
            hi[q] = wave_max_f(hi[q]);
          }
        }
        if (lane == 0) {
          m_v[j] = wv;
          m_i[j] = wi;
#pragma unroll
          for (int q = 0; q < D; ++q) m_c[q * nbl + j] = __float_as_int(wc[q]);
          if (s == 1) {
#pragma unroll
            for (int q = 0; q < D; ++q) {
              m_lo[q * nbl + j] = lo[q];
              m_hi[q * nbl + j] = hi[q];
            }
          }
        }
      }
    }
    lds_barrier();  // every block's cached candidate is current
    if (s < f.n_starts) {
      pidx = load_forced<D>(f, s, pv);
      if (g == 0 && tid >= B - 64) record_pick<D>(f, s, pidx, pv);
      continue;
    }
    ++e;
    const int par = (int)(e & 1u);
    // Workgroup argmax over the cached candidates, eight slots per lane
    // loaded at once: by wave 0 alone, which publishes it, unless there is no
    // exchange (one workgroup: every wave).
    int win[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) win[q] = 0;
    if (G == 1 || wave == 0) {
      int bv = INT_MIN;
      uint32_t bi = 0;
      uint32_t bj = 0;
      for (uint32_t j0 = 0; j0 < nbl; j0 += 64u * 8u) {
        int v[8];
        uint32_t id[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const uint32_t jl = j0 + (uint32_t)(u * 64 + lane);
          const uint32_t jc = jl < nbl ? jl : nbl - 1u;
          v[u] = m_v[jc];
          id[u] = (uint32_t)m_i[jc];
          v[u] = jl < nbl ? v[u] : INT_MIN;  // padding never beats a real slot
          id[u] = jl < nbl ? id[u] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const bool take = beats<TIE>(v[u], id[u], bv, bi);
          bv = take ? v[u] : bv;
          bi = take ? id[u] : bi;
          bj = take ? j0 + (uint32_t)(u * 64 + lane) : bj;
        }
      }
      int gv;
      const int gl = wave_argmax_lane<TIE>(bv, bi, gv);
      const uint32_t wj = (uint32_t)readlane_i((int)bj, gl);
      win[0] = gv;
      win[1] = readlane_i((int)bi, gl);
#pragma unroll
      for (int q = 0; q < D; ++q) win[2 + q] = m_c[q * nbl + wj];
    }
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
      if (!exchange<TIE, NW, W>(f.xchg, f.replicas, kPageWords, G, g, e, f.backoff, f.spin_ticks,
                                s_cand[par], s_xw[par], 0, win)) {
        if (tid == 0) atomicOr(f.status, kStatusTimeout);
        return;
      }
    }
    pidx = (uint32_t)win[1];
#pragma unroll
    for (int q = 0; q < D; ++q) pv[q] = __int_as_float(win[2 + q]);
    if (g == 0 && tid >= B - 64) record_pick<D>(f, s, pidx, pv);
  }
}

// ---------------------------------------------------------------------------
// Paired streaming form: the blocks and per-block caches of fps_stream_pruned,
// settling up to two picks per exchange as fps_pruned_pair does.  Every
// workgroup publishes its best A and a lookahead A' (its best after a
// hypothetical update by A); the polled pairs give p1 and, unless some
// workgroup whose A moves under p1 could still beat it, p2.  While waves
// 1..14 update the listed blocks, the idle last wave speculates on A (the
// best cached candidate of the blocks not updated) and stages that block in
// LDS; A is then formed from its candidate and the updated blocks.  For the
// lookahead, waves 0..7 test the blocks: a block whose box lies out of A's
// reach, or whose cached best keeps its value under A, keeps its best; the
// others are listed and streamed (their min-distances min'ed again with this
// pass's pivots where written this pass -- a no-op once the stores have
// landed -- and with A; nothing written back).  Waves 8..15 meanwhile take
// A's own block one row each, from the staged copy when the guess held.
// Wave 0 takes the best of the waves' results.  A workgroup skips the
// lookahead while the best A of the workgroups the last picks certainly left
// unchanged (their A keeps its value, so it stays their best) beats its own:
// it cannot then hold the next p1, the only workgroup whose A' is read.
// 2 <= G <= 64.
// ---------------------------------------------------------------------------
template <int TIE>
__global__ void __launch_bounds__(1024) fps_stream_pair(PrunedParams a) {
  constexpr int B = 1024;
  constexpr int W = B / 64;
  constexpr int D = 3;
  constexpr int NW = 2 + D;
  constexpr int NP = 2 * NW;
  constexpr int SL = kSBlock / 64;  // slots per lane in a block
  __shared__ int4 s_res[2][3];      // p1, p2, two-pick flag, all-arrived flag
  extern __shared__ __align__(16) int s_meta[];  // kSPairMetaWords words per block
  __shared__ int s_count[2];        // listed blocks, by pass parity
  __shared__ int4 s_a[2];           // A (value, index, coordinates), lookahead flag
  __shared__ int s_cnt2;            // blocks the lookahead streams
  __shared__ int4 s_lp[W][2];       // per wave: best kept candidate (L)
  __shared__ int4 s_hp[W][2];       // per wave: best of its streamed rows or blocks
  __shared__ float4 s_stg[SL][64];  // the staged block: x, y, z, min-distance
  __shared__ uint32_t s_stg_id[SL][64];
  __shared__ int s_spec;            // slot of the staged block, or -1
  __shared__ int4 s_sb[2];          // best of the blocks not updated (valid when s_spec >= 0)
  static_assert(SL == W / 2, "the second half of the waves streams A's block, one row each");

  const FpsParams &f = a.fps;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t n = f.n;
  const uint32_t nblocks = (n + kSBlock - 1) / kSBlock;
  const uint32_t nbl = (nblocks > g) ? (nblocks - g + G - 1) / G : 0u;
  // Per block j, 16-byte records (read with one LDS instruction each): box
  // low corner and cached best value bits, box high corner and cached best
  // index, cached best coordinates; the value and index again by themselves
  // for the scans that need only them.
  float4 *m_b0 = reinterpret_cast<float4 *>(s_meta);  // lo.x, lo.y, lo.z, value bits
  float4 *m_b1 = m_b0 + nbl;                          // hi.x, hi.y, hi.z, index bits
  float4 *m_cb = m_b1 + nbl;                          // best x, y, z, -
  int *m_v = reinterpret_cast<int *>(m_cb + nbl);
  int *m_i = m_v + nbl;
  int *s_list = m_i + nbl;  // blocks to stream (update, then the lookahead)
  uint32_t *m_mark = reinterpret_cast<uint32_t *>(s_list + nbl);  // pass that last listed the slot
  for (uint32_t j = (uint32_t)tid; j < nbl; j += B) m_mark[j] = 0u;
  if (tid < 2) s_count[tid] = 0;
  lds_barrier();
  const float *xs = a.sorted;
  const float *ys = a.sorted + a.sorted_ld;
  const float *zs = a.sorted + 2 * a.sorted_ld;
  const bool recorder = (g == 0 && tid >= B - 64);

  float p1[D], p2[D];
  uint32_t i1 = load_forced<D>(f, 0, p1);
#pragma unroll
  for (int j = 0; j < D; ++j) p2[j] = 0.0f;
  if (recorder) record_pick<D>(f, 0, i1, p1);

  // Wave 0's state: the workgroup's A and A', M (best A among the other
  // workgroups the last picks left unchanged).
  int ga[NW], gl_[NW];
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    ga[k] = 0;
    gl_[k] = 0;
  }
  uint32_t ga_j = 0;  // A's block slot
  bool la_valid = false;
  int mu_v = INT_MIN;
  uint32_t mu_i = 0;
  const int K = min(min((int)f.replicas, 64 / NP), 4);

  bool have2 = false;
  bool first = true;
  uint32_t e = 0;
  uint32_t t = 0;
  uint32_t s = 1;
  while (s < f.m) {
    const int sp = (int)(t & 1u);
    ++t;
    // This wave's min-distance stores of the previous pass complete before
    // the barrier below, after which any wave may read those blocks (wave 0
    // writes none).
    if (wave != 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    // Box tests against the pending pivots: wave w takes slots from
    // 64 ((w + W - 1) mod W) on, so that wave 0, which is forming M, has
    // none for up to 960 blocks.
    for (uint32_t j0 = (uint32_t)(((wave + W - 1) % W) * 64); j0 < nbl; j0 += 64u * W) {
      const uint32_t jl = j0 + (uint32_t)lane;
      bool need = false;
      if (jl < nbl) {
        if (first) {
          need = true;
        } else {
          const float4 b0 = m_b0[jl], b1 = m_b1[jl];
          const float lo[D] = {b0.x, b0.y, b0.z};
          const float hi[D] = {b1.x, b1.y, b1.z};
          const float bvf = b0.w;
          need = box_bound<D>(p1, lo, hi) < bvf || (have2 && box_bound<D>(p2, lo, hi) < bvf);
        }
      }
      const uint64_t mask = __ballot(need);
      if (mask != 0) {
        int at = 0;
        if (lane == 0) at = atomicAdd(&s_count[sp], __popcll(mask));
        at = readlane_i(at, 0);
        if (need) {
          s_list[at + __popcll(mask & ((1ull << lane) - 1))] = (int)jl;
          m_mark[jl] = t;
        }
      }
    }
    lds_barrier();  // the list is complete
    if (tid == 0) s_count[sp ^ 1] = 0;  // next pass's counter; nobody reads it before then
    const int listed = s_count[sp];
    // Update: waves 1..W-2 take the listed blocks in turn (both pivots);
    // wave 0 stores no min-distance, so its next pass need not wait for
    // its stores (its exchange publish among them), and the last wave is
    // left to the speculation.
    for (int u = wave - 1; wave >= 1 && wave <= W - 2 && u < listed; u += W - 2) {
      const uint32_t j = (uint32_t)s_list[u];
      const uint32_t base = (g + j * G) * (uint32_t)kSBlock;
      int bv = INT_MIN;
      uint32_t bi = 0;
      float bc[D] = {0.0f, 0.0f, 0.0f};
      float lo[D] = {INFINITY, INFINITY, INFINITY};
      float hi[D] = {-INFINITY, -INFINITY, -INFINITY};
      float bx[SL][D], bo[SL];
      uint32_t bid[SL];
#pragma unroll
      for (int k = 0; k < SL; ++k) {
        const uint32_t p = base + (uint32_t)(k * 64 + lane);
        const uint32_t q = p < n ? p : n - 1;
        bx[k][0] = xs[q];
        bx[k][1] = ys[q];
        bx[k][2] = zs[q];
        bid[k] = a.perm[q];
        bo[k] = first ? f.init : f.dmin[q];
      }
#pragma unroll
      for (int k = 0; k < SL; ++k) {
        const uint32_t p = base + (uint32_t)(k * 64 + lane);
        const bool live = p < n;
        const float(&c)[D] = bx[k];
        const float old = bo[k];
        float nm = min_keep(sqdist<D>(c, p1), old);
        if (have2) nm = min_keep(sqdist<D>(c, p2), nm);
        if (live && (first || __float_as_int(nm) != __float_as_int(old))) f.dmin[p] = nm;
        const int b = __float_as_int(nm);
        const bool take = live && ((TIE == kTieHighest) ? (b >= bv) : (b > bv));
        bv = take ? b : bv;
        bi = take ? bid[k] : bi;
#pragma unroll
        for (int r = 0; r < D; ++r) bc[r] = take ? c[r] : bc[r];
        if (first && live) {
#pragma unroll
          for (int r = 0; r < D; ++r) {
            lo[r] = fminf(lo[r], c[r]);
            hi[r] = fmaxf(hi[r], c[r]);
          }
        }
      }
      int wv;
      const int wl = wave_argmax_lane<TIE>(bv, bi, wv);
      const int wi = readlane_i((int)bi, wl);
      float wc[D];
#pragma unroll
      for (int q = 0; q < D; ++q) wc[q] = readlane_f(bc[q], wl);
      if (first) {
#pragma unroll
        for (int q = 0; q < D; ++q) {
          lo[q] = wave_min_f(lo[q]);
          hi[q] = wave_max_f(hi[q]);
        }
      }
      if (lane == 0) {
        m_v[j] = wv;
        m_i[j] = wi;
        m_cb[j] = make_float4(wc[0], wc[1], wc[2], 0.0f);
        if (first) {
          m_b0[j] = make_float4(lo[0], lo[1], lo[2], __int_as_float(wv));
          m_b1[j] = make_float4(hi[0], hi[1], hi[2], __int_as_float(wi));
        } else {
          m_b0[j].w = __int_as_float(wv);
          m_b1[j].w = __int_as_float(wi);
        }
      }
    }
    // Speculation (the last wave, when the update leaves it idle): the best
    // cached candidate of the blocks this pass does not update.  It is the
    // workgroup's A after the update unless an updated block's new best beats
    // it, and a lookahead streams A's own block, so that block's points and
    // min-distances (current: it is not written this pass) are staged in LDS
    // while the update's memory round trip runs.  The lookahead uses them
    // only if A's block is the staged one.
    if (wave == W - 1) {
      int sj = -1;
      if (!first && s >= f.n_starts && listed > 0 && listed < 64) {
        int bv = INT_MIN;
        uint32_t bi = 0, bj = 0;
        for (uint32_t j0 = 0; j0 < nbl; j0 += 64u * 8u) {
          int v[8];
          uint32_t id[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const uint32_t jl = j0 + (uint32_t)(u * 64 + lane);
            const uint32_t jc = jl < nbl ? jl : nbl - 1u;
            v[u] = m_v[jc];
            id[u] = (uint32_t)m_i[jc];
            const bool skip = jl >= nbl || m_mark[jc] == t;  // padding, or updated this pass
            v[u] = skip ? INT_MIN : v[u];
            id[u] = skip ? 0u : id[u];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const bool take = beats<TIE>(v[u], id[u], bv, bi);
            bv = take ? v[u] : bv;
            bi = take ? id[u] : bi;
            bj = take ? j0 + (uint32_t)(u * 64 + lane) : bj;
          }
        }
        int gv;
        const int gl = wave_argmax_lane<TIE>(bv, bi, gv);
        if (gv != INT_MIN) {
          const uint32_t tj = (uint32_t)readlane_i((int)bj, gl);
          sj = (int)tj;
          if (lane == 0) {
            // The best cached candidate of the blocks not updated this pass:
            // wave 0 then forms A from it and the updated blocks alone.
            const float4 cb = m_cb[tj];
            s_sb[0] = make_int4(gv, readlane_i((int)bi, gl), (int)tj, __float_as_int(cb.x));
            s_sb[1] = make_int4(__float_as_int(cb.y), __float_as_int(cb.z), 0, 0);
          }
          const uint32_t base = (g + tj * G) * (uint32_t)kSBlock;
          float4 r[SL];
          uint32_t rid[SL];
#pragma unroll
          for (int k = 0; k < SL; ++k) {
            const uint32_t p = base + (uint32_t)(k * 64 + lane);
            const uint32_t q = p < n ? p : n - 1;
            r[k] = make_float4(xs[q], ys[q], zs[q], f.dmin[q]);
            rid[k] = a.perm[q];
          }
#pragma unroll
          for (int k = 0; k < SL; ++k) {
            s_stg[k][lane] = r[k];
            s_stg_id[k][lane] = rid[k];
          }
        }
      }
      if (lane == 0) s_spec = sj;
    }
    lds_barrier();  // every block's cached candidate is current
    if (s < f.n_starts) {
      // Forced prefix (multi-start): one pick per pass, no exchange.
      i1 = load_forced<D>(f, s, p1);
      have2 = false;
      if (recorder) record_pick<D>(f, s, i1, p1);
      ++s;
      first = false;
      continue;
    }
    ++e;
    const int par = (int)(e & 1u);
    if (wave == 0) {
      if (e == 1u || listed > 0) {
        if (s_spec >= 0) {
          // A from the speculation's best of the blocks not updated (lane 0)
          // and the updated blocks (lanes 1..listed, listed < 64).
          const int4 b0 = s_sb[0];
          const uint32_t jl = lane >= 1 && lane <= listed ? (uint32_t)s_list[lane - 1] : 0u;
          int v = INT_MIN;
          uint32_t id = 0, js = 0;
          if (lane == 0) {
            v = b0.x;
            id = (uint32_t)b0.y;
            js = (uint32_t)b0.z;
          } else if (lane <= listed) {
            v = m_v[jl];
            id = (uint32_t)m_i[jl];
            js = jl;
          }
          int gv;
          const int gl = wave_argmax_lane<TIE>(v, id, gv);
          const uint32_t wj = (uint32_t)readlane_i((int)js, gl);
          ga[0] = gv;
          ga[1] = readlane_i((int)id, gl);
          if (gl == 0) {
            const int4 b1 = s_sb[1];
            ga[2] = b0.w;
            ga[3] = b1.x;
            ga[4] = b1.y;
          } else {
            const float4 cb = m_cb[wj];
            ga[2] = __float_as_int(cb.x);
            ga[3] = __float_as_int(cb.y);
            ga[4] = __float_as_int(cb.z);
          }
          ga_j = wj;
          la_valid = false;
        } else {
          // The workgroup's A: argmax over the cached candidates, eight slots
          // per lane loaded at once.
          int bv = INT_MIN;
          uint32_t bi = 0, bj = 0;
          for (uint32_t j0 = 0; j0 < nbl; j0 += 64u * 8u) {
            int v[8];
            uint32_t id[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const uint32_t jl = j0 + (uint32_t)(u * 64 + lane);
              const uint32_t jc = jl < nbl ? jl : nbl - 1u;
              v[u] = m_v[jc];
              id[u] = (uint32_t)m_i[jc];
              v[u] = jl < nbl ? v[u] : INT_MIN;  // padding never beats a real slot
              id[u] = jl < nbl ? id[u] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const bool take = beats<TIE>(v[u], id[u], bv, bi);
              bv = take ? v[u] : bv;
              bi = take ? id[u] : bi;
              bj = take ? j0 + (uint32_t)(u * 64 + lane) : bj;
            }
          }
          int gv;
          const int gl = wave_argmax_lane<TIE>(bv, bi, gv);
          const uint32_t wj = (uint32_t)readlane_i((int)bj, gl);
          ga[0] = gv;
          ga[1] = readlane_i((int)bi, gl);
          const float4 cb = m_cb[wj];
          ga[2] = __float_as_int(cb.x);
          ga[3] = __float_as_int(cb.y);
          ga[4] = __float_as_int(cb.z);
          ga_j = wj;
          la_valid = false;
        }
      }
      // Lookahead only while this workgroup could hold the next p1.
      const bool needL = !la_valid && beats<TIE>(ga[0], (uint32_t)ga[1], mu_v, mu_i);
      if (needL) la_valid = true;
      if (lane == 0) {
        s_a[0] = make_int4(ga[0], ga[1], ga[2], ga[3]);
        s_a[1] = make_int4(ga[4], needL ? 1 : 0, (int)ga_j, 0);
        s_cnt2 = 0;
      }
    }
    lds_barrier();  // A and the lookahead decision
    if (s_a[1].y != 0) {
      // Lookahead.  L: the best cached candidate of the blocks A leaves
      // unchanged at their best (out of reach, or their best keeps its
      // value): waves 0..7 test the blocks and list the others.  A's own
      // block always moves (its best is A) but for degenerate values: waves
      // 8..15 stream it meanwhile, one row of 64 slots each, from the staged
      // copy when the speculation found it (a kept one is also in L, as the
      // same candidate).  The listed blocks, if any, are streamed after that.
      // A block written this pass gets this pass's pivots again (its stores
      // may not have landed; min'ing twice changes nothing), and in the first
      // pass every min-distance starts from the initial value.
      const int4 sa0 = s_a[0], sa1 = s_a[1];
      const float pa[D] = {__int_as_float(__builtin_amdgcn_readfirstlane(sa0.z)),
                           __int_as_float(__builtin_amdgcn_readfirstlane(sa0.w)),
                           __int_as_float(__builtin_amdgcn_readfirstlane(sa1.x))};
      const uint32_t ja = (uint32_t)__builtin_amdgcn_readfirstlane(sa1.z);
      int hv = INT_MIN;  // this lane's best of the streamed slots
      uint32_t hid = 0;
      float hc[D] = {0.0f, 0.0f, 0.0f};
      // Folds slot (c, min-distance d0, original index id) under the pivots.
      auto sweep_slot = [&](bool live, const float(&c)[D], float d0, uint32_t id, bool again) {
        float h = d0;
        if (again) {
          h = min_keep(sqdist<D>(c, p1), h);
          if (have2) h = min_keep(sqdist<D>(c, p2), h);
        }
        h = min_keep(sqdist<D>(c, pa), h);
        const int b = __float_as_int(h);
        const bool take = live && beats<TIE>(b, id, hv, hid);
        hv = take ? b : hv;
        hid = take ? id : hid;
#pragma unroll
        for (int r = 0; r < D; ++r) hc[r] = take ? c[r] : hc[r];
      };
      auto publish_h = [&]() {
        int Hv;
        const int hl = wave_argmax_lane<TIE>(hv, hid, Hv);
        const int Hi = readlane_i((int)hid, hl);
        int Hc[D];
#pragma unroll
        for (int q = 0; q < D; ++q) Hc[q] = __float_as_int(readlane_f(hc[q], hl));
        if (lane == 0) {
          s_hp[wave][0] = make_int4(Hv, Hi, Hc[0], Hc[1]);
          s_hp[wave][1] = make_int4(Hc[2], 0, 0, 0);
        }
      };
      if (wave >= W / 2) {
        // Row (wave - W/2) of A's block.
        const int k = wave - W / 2;
        const uint32_t p = (g + ja * G) * (uint32_t)kSBlock + (uint32_t)(k * 64 + lane);
        float c[D], d0;
        uint32_t id;
        bool again;
        if (s_spec == (int)ja) {
          const float4 r = s_stg[k][lane];
          c[0] = r.x;
          c[1] = r.y;
          c[2] = r.z;
          d0 = r.w;
          id = s_stg_id[k][lane];
          again = false;  // staged after this pass's stores were ordered; not written since
        } else {
          const uint32_t q = p < n ? p : n - 1;
          c[0] = xs[q];
          c[1] = ys[q];
          c[2] = zs[q];
          id = a.perm[q];
          d0 = first ? f.init : f.dmin[q];
          again = first || m_mark[ja] == t;
        }
        sweep_slot(p < n, c, d0, id, again);
        publish_h();
      } else {
        // Box tests of slots [64 w, 64 w + 64), [64 w + 64 W/2, ...).
        int lv = INT_MIN;
        uint32_t li = 0;
        int lc[D] = {0, 0, 0};
        for (uint32_t j0 = (uint32_t)(wave * 64); j0 < nbl; j0 += 64u * (W / 2)) {
          const uint32_t jl = j0 + (uint32_t)lane;
          bool sweep = false;
          if (jl < nbl) {
            const float4 b0 = m_b0[jl], b1 = m_b1[jl], c4 = m_cb[jl];
            const int v = __float_as_int(b0.w);
            const uint32_t id = (uint32_t)__float_as_int(b1.w);
            const float lo[D] = {b0.x, b0.y, b0.z};
            const float hi[D] = {b1.x, b1.y, b1.z};
            const int cb[D] = {__float_as_int(c4.x), __float_as_int(c4.y), __float_as_int(c4.z)};
            const float cf[D] = {__int_as_float(cb[0]), __int_as_float(cb[1]), __int_as_float(cb[2])};
            const float vf = __int_as_float(v);
            const bool kept = !(box_bound<D>(pa, lo, hi) < vf) ||
                              __float_as_int(min_keep(sqdist<D>(cf, pa), vf)) == v;
            if (kept && beats<TIE>(v, id, lv, li)) {
              lv = v;
              li = id;
#pragma unroll
              for (int q = 0; q < D; ++q) lc[q] = cb[q];
            }
            sweep = !kept && jl != ja;
          }
          const uint64_t mask = __ballot(sweep);
          if (mask != 0) {
            int at = 0;
            if (lane == 0) at = atomicAdd(&s_cnt2, __popcll(mask));
            at = readlane_i(at, 0);
            if (sweep) s_list[at + __popcll(mask & ((1ull << lane) - 1))] = (int)jl;
          }
        }
        int Lv;
        const int ll = wave_argmax_lane<TIE>(lv, li, Lv);
        const int Li = readlane_i((int)li, ll);
        int Lc[D];
#pragma unroll
        for (int q = 0; q < D; ++q) Lc[q] = readlane_i(lc[q], ll);
        if (lane == 0) {
          s_lp[wave][0] = make_int4(Lv, Li, Lc[0], Lc[1]);
          s_lp[wave][1] = make_int4(Lc[2], 0, 0, 0);
          s_hp[wave][0] = make_int4(INT_MIN, 0, 0, 0);  // no streamed block yet
          s_hp[wave][1] = make_int4(0, 0, 0, 0);
        }
      }
      lds_barrier();  // every wave's L and A-row best; the list
      const int n2 = s_cnt2;
      if (n2 > 0) {
        // The listed blocks, waves 0..7 in turn.
        if (wave < W / 2) {
          for (int u = wave; u < n2; u += W / 2) {
            const uint32_t j = (uint32_t)s_list[u];
            const uint32_t base = (g + j * G) * (uint32_t)kSBlock;
            const bool again = first || m_mark[j] == t;
            float bx[SL][D], bo[SL];
            uint32_t bid[SL];
#pragma unroll
            for (int k = 0; k < SL; ++k) {
              const uint32_t p = base + (uint32_t)(k * 64 + lane);
              const uint32_t q = p < n ? p : n - 1;
              bx[k][0] = xs[q];
              bx[k][1] = ys[q];
              bx[k][2] = zs[q];
              bid[k] = a.perm[q];
              bo[k] = first ? f.init : f.dmin[q];
            }
#pragma unroll
            for (int k = 0; k < SL; ++k)
              sweep_slot(base + (uint32_t)(k * 64 + lane) < n, bx[k], bo[k], bid[k], again);
          }
          publish_h();
        }
        lds_barrier();  // every wave's streamed best
      }
      if (wave == 0) {
        // A' = the best of the waves' L's and streamed bests, one per lane.
        const bool has = lane < W / 2 + W;
        const bool isl = lane < W / 2;
        const int r = has ? (isl ? lane : lane - W / 2) : 0;
        const int4 c0 = isl ? s_lp[r][0] : s_hp[r][0];
        const int c4 = isl ? s_lp[r][1].x : s_hp[r][1].x;
        int av;
        const int al = wave_argmax_lane<TIE>(has ? c0.x : INT_MIN, (uint32_t)c0.y, av);
        gl_[0] = av;
        gl_[1] = readlane_i(c0.y, al);
        gl_[2] = readlane_i(c0.z, al);
        gl_[3] = readlane_i(c0.w, al);
        gl_[4] = readlane_i(c4, al);
      }
    }
    int mc_v = INT_MIN;  // wave 0: this lane's workgroup's A if it counts for M
    uint32_t mc_i = 0;
    if (wave == 0) {
      // Exchange: publish the pair to K replicas, poll replica g mod K, one
      // workgroup per lane, all NP words at once.
      uint64_t *bufs = f.xchg + (size_t)(e & 1u) * f.replicas * kSlotWords * kPageWords;
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
      int r2[D];
      float q2[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const int cw = mine ? x[NW + 2 + j] : x[2 + j];
        r2[j] = readlane_i(cw, l2);
        q2[j] = __int_as_float(r2[j]);
      }
      if (lane == 0) {
        s_res[par][0] = make_int4(r1[0], r1[1], r1[2], r1[3]);
        s_res[par][1] = make_int4(r1[4], v2, (int)i2, r2[0]);
        s_res[par][2] = make_int4(r2[1], r2[2], two ? 1 : 0, all_ok ? 1 : 0);
      }
      // M for the next pass: the best A among the other workgroups whose A
      // keeps its value under this pass's picks (so stays their best);
      // reduced after the hand-off, while the other waves start the next
      // pass's box tests (wave 0 has none of its own there).
      const bool keep1 = __float_as_int(post) == x[0];
      const float post2 = min_keep(sqdist<D>(qa, q2), __int_as_float(x[0]));
      const bool keep = valid && x[0] != INT_MIN && keep1 &&
                        !(two && __float_as_int(post2) != x[0]);
      mc_v = (keep && (uint32_t)lane != g) ? x[0] : INT_MIN;
      mc_i = (uint32_t)x[1];
    }
    lds_barrier();
    const int4 q0 = s_res[par][0], qq1 = s_res[par][1], qq2 = s_res[par][2];
    if (qq2.w == 0) {
      if (tid == 0) atomicOr(f.status, kStatusTimeout);
      return;
    }
    if (wave == 0) {
      int mv;
      const int ml = wave_argmax_lane<TIE>(mc_v, mc_i, mv);
      mu_v = mv;
      mu_i = (uint32_t)readlane_i((int)mc_i, ml);
    }
    i1 = (uint32_t)__builtin_amdgcn_readfirstlane(q0.y);
    p1[0] = __int_as_float(__builtin_amdgcn_readfirstlane(q0.z));
    p1[1] = __int_as_float(__builtin_amdgcn_readfirstlane(q0.w));
    p1[2] = __int_as_float(__builtin_amdgcn_readfirstlane(qq1.x));
    const uint32_t i2 = (uint32_t)__builtin_amdgcn_readfirstlane(qq1.z);
    p2[0] = __int_as_float(__builtin_amdgcn_readfirstlane(qq1.w));
    p2[1] = __int_as_float(__builtin_amdgcn_readfirstlane(qq2.x));
    p2[2] = __int_as_float(__builtin_amdgcn_readfirstlane(qq2.y));
    have2 = qq2.z != 0 && s + 1 < f.m;
    if (recorder) {
      record_pick<D>(f, s, i1, p1);
      if (have2) record_pick<D>(f, s + 1, i2, p2);
    }
    s += have2 ? 2u : 1u;
    first = false;
  }
  // The launch's exchange count in the second status word.
  if (recorder && lane == 0) f.status[1] = e;
}

static const KernelEntry kStreamPruned[] = {
    {3, 1024, 0, kTieHighest, (const void *)&fps_stream_pruned<kTieHighest>},
    {3, 1024, 0, kTieLowest, (const void *)&fps_stream_pruned<kTieLowest>},
};

const KernelEntry *stream_pruned_table(size_t *count) {
  *count = sizeof(kStreamPruned) / sizeof(kStreamPruned[0]);
  return kStreamPruned;
}

static const KernelEntry kStreamPair[] = {
    {3, 1024, 0, kTieHighest, (const void *)&fps_stream_pair<kTieHighest>},
    {3, 1024, 0, kTieLowest, (const void *)&fps_stream_pair<kTieLowest>},
};

const KernelEntry *stream_pair_table(size_t *count) {
  *count = sizeof(kStreamPair) / sizeof(kStreamPair[0]);
  return kStreamPair;
}

#define FPSX_PRUNED(D, B, P)                                                   \
  {D, B, P, kTieHighest, (const void *)&fps_pruned<D, B, P, kTieHighest>}, \
  {D, B, P, kTieLowest, (const void *)&fps_pruned<D, B, P, kTieLowest>}

static const KernelEntry kPruned[] = {
    FPSX_PRUNED(3, 1024, 16), FPSX_PRUNED(3, 1024, 8), FPSX_PRUNED(3, 1024, 4),
    FPSX_PRUNED(3, 256, 16),  FPSX_PRUNED(3, 256, 32),
};
#undef FPSX_PRUNED

const KernelEntry *pruned_table(size_t *count) {
  *count = sizeof(kPruned) / sizeof(kPruned[0]);
  return kPruned;
}

#define FPSX_SOLO(D, B, P)                                                        \
  {D, B, P, kTieHighest, (const void *)&fps_pruned_solo<D, B, P, kTieHighest>}, \
  {D, B, P, kTieLowest, (const void *)&fps_pruned_solo<D, B, P, kTieLowest>}

static const KernelEntry kSolo[] = {
    FPSX_SOLO(3, 1024, 16), FPSX_SOLO(3, 1024, 8), FPSX_SOLO(3, 1024, 4),
};
#undef FPSX_SOLO

const KernelEntry *solo_table(size_t *count) {
  *count = sizeof(kSolo) / sizeof(kSolo[0]);
  return kSolo;
}

// (points per thread P = registers PPT + LDS E)
#define FPSX_PAIR(D, B, PPT, E)                                                                 \
  {D, B, PPT + E, kTieHighest, (const void *)&fps_pruned_pair<D, B, PPT, kTieHighest, E>}, \
  {D, B, PPT + E, kTieLowest, (const void *)&fps_pruned_pair<D, B, PPT, kTieLowest, E>}

static const KernelEntry kPair[] = {
    FPSX_PAIR(3, 1024, 16, 0), FPSX_PAIR(3, 1024, 8, 0), FPSX_PAIR(3, 1024, 4, 0),
    FPSX_PAIR(3, 1024, 16, 1),
};
#undef FPSX_PAIR

#define FPSX_PAIR_WIDE(D, B, P, PW)                                                          \
  {D, B, P, kTieHighest, (const void *)&fps_pruned_pair_wide<D, B, P, kTieHighest, PW>}, \
  {D, B, P, kTieLowest, (const void *)&fps_pruned_pair_wide<D, B, P, kTieLowest, PW>}

static const KernelEntry kPairWide[] = {
    FPSX_PAIR_WIDE(3, 1024, 16, 2), FPSX_PAIR_WIDE(3, 1024, 16, 4),
};
#undef FPSX_PAIR_WIDE

const KernelEntry *pair_table(size_t *count) {
  *count = sizeof(kPair) / sizeof(kPair[0]);
  return kPair;
}

const KernelEntry *pair_wide_kernel(int block, int ppt, int tie, int polling_waves) {
  const int idx = polling_waves == 2 ? 0 : polling_waves == 4 ? 2 : -1;
  if (idx < 0 || block != 1024 || ppt != 16) return nullptr;
  return &kPairWide[idx + (tie == kTieLowest ? 1 : 0)];
}

}  // namespace fpsx

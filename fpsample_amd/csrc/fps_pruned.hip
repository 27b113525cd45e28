// fps_pruned.hip -- register-resident FPS with per-wave bounding-box pruning.
//
// Same picks as fps_resident (vanilla FPS, reference src/lib.cpp:188-246),
// less work: the points are first put in spatial order (Morton codes, radix
// sort), and every wave owns one block of 64*PPT consecutive points of that
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

// Wave-wide min / max of a float (NaN lanes ignored: v_min/v_max return the
// other operand), result uniform.
__device__ __forceinline__ float wave_min_f(float v) {
  int x = __float_as_int(v);
  const int id = __float_as_int(INFINITY);
#define STEP(ctrl, rm)                                                                     \
  x = __float_as_int(fminf(__int_as_float(x),                                              \
                           __int_as_float(__builtin_amdgcn_update_dpp(id, x, ctrl, rm, 0xf, false))))
  STEP(0xb1, 0xf); STEP(0x4e, 0xf); STEP(0x114, 0xf); STEP(0x118, 0xf); STEP(0x142, 0xa);
  STEP(0x143, 0xc);
#undef STEP
  return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}
__device__ __forceinline__ float wave_max_f(float v) {
  int x = __float_as_int(v);
  const int id = __float_as_int(-INFINITY);
#define STEP(ctrl, rm)                                                                     \
  x = __float_as_int(fmaxf(__int_as_float(x),                                              \
                           __int_as_float(__builtin_amdgcn_update_dpp(id, x, ctrl, rm, 0xf, false))))
  STEP(0xb1, 0xf); STEP(0x4e, 0xf); STEP(0x114, 0xf); STEP(0x118, 0xf); STEP(0x142, 0xa);
  STEP(0x143, 0xc);
#undef STEP
  return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}

// Lower bound of the squared distance from p to the box [lo, hi], in the
// point distance's operand order (src/_ext/KDNode.h:105-118 does the same).
template <int D>
__device__ __forceinline__ float box_bound(const float (&p)[D], const float (&lo)[D],
                                           const float (&hi)[D]) {
  float acc = 0.0f;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    float g = 0.0f;
    if (p[j] > hi[j]) g = p[j] - hi[j];
    else if (p[j] < lo[j]) g = lo[j] - p[j];
    acc = (j == 0) ? g * g : acc + g * g;
  }
  return acc;
}

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
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t i = wbase + (uint32_t)(k * 64 + lane);
    const bool live = i < f.n;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      c[k][j] = live ? a.sorted[(size_t)j * a.sorted_ld + i] : 0.0f;
      if (live) {
        lo[j] = fminf(lo[j], c[k][j]);
        hi[j] = fmaxf(hi[j], c[k][j]);
      }
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
        // Fully unrolled, loads unconditional (a position past the cloud
        // reads the last point and is masked out), so every load of the
        // block is in flight at once: one memory round trip per block.
#pragma unroll
        for (int k = 0; k < SL; ++k) {
          const uint32_t p = base + (uint32_t)(k * 64 + lane);
          const bool live = p < n;
          const uint32_t q = live ? p : n - 1;
          const float c[D] = {xs[q], ys[q], zs[q]};
          const uint32_t id = a.perm[q];
          const float old = (s == 1) ? f.init : f.dmin[q];
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
    // Workgroup argmax over the cached candidates (every wave).
    int bv = INT_MIN;
    uint32_t bi = 0;
    uint32_t bj = 0;
    for (uint32_t j = (uint32_t)lane; j < nbl; j += 64u) {
      const int v = m_v[j];
      const uint32_t id = (uint32_t)m_i[j];
      if (beats<TIE>(v, id, bv, bi)) {
        bv = v;
        bi = id;
        bj = j;
      }
    }
    int gv;
    const int gl = wave_argmax_lane<TIE>(bv, bi, gv);
    const uint32_t wj = (uint32_t)readlane_i((int)bj, gl);
    int win[NW];
    win[0] = gv;
    win[1] = readlane_i((int)bi, gl);
#pragma unroll
    for (int q = 0; q < D; ++q) win[2 + q] = m_c[q * nbl + wj];
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

static const KernelEntry kStreamPruned[] = {
    {3, 1024, 0, kTieHighest, (const void *)&fps_stream_pruned<kTieHighest>},
    {3, 1024, 0, kTieLowest, (const void *)&fps_stream_pruned<kTieLowest>},
};

const KernelEntry *stream_pruned_table(size_t *count) {
  *count = sizeof(kStreamPruned) / sizeof(kStreamPruned[0]);
  return kStreamPruned;
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

}  // namespace fpsx

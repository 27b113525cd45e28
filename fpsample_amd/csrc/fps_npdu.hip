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
// Streaming form for clouds beyond the register file, any dimension.  The
// min-distances live in HBM (f.dmin, by index) and the cloud is cut into
// blocks of kNpduBlock consecutive indices; block b belongs to workgroup
// b mod G, which caches in LDS each of its blocks' best (value, index).  A
// pass re-reads only the blocks the window reaches (every block for the
// start): in-window min-distances are updated, and the block's best is
// re-taken over all its min-distances.  The winner travels as (value,
// index); every workgroup then re-reads the pivot's coordinates.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) fps_npdu_stream(NpduParams a) {
  constexpr int B = 1024;
  constexpr int W = B / 64;
  constexpr int NW = 2;
  constexpr int SL = kNpduBlock / 64;
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
  const uint32_t nblocks = (n + kNpduBlock - 1) / kNpduBlock;
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
    const uint32_t blo = lo / kNpduBlock, bhi = hi / kNpduBlock;
    const uint32_t j_first = (blo > g) ? (blo - g + G - 1) / G : 0u;
    const uint32_t j_end = (bhi >= g) ? min(nbl, (bhi - g) / G + 1) : 0u;
    for (uint32_t j = j_first + (uint32_t)wave; j < j_end; j += W) {
      const uint32_t base = (g + j * G) * (uint32_t)kNpduBlock;
      int bv = INT_MIN;
      uint32_t bi = 0;
#pragma unroll
      for (int k = 0; k < SL; ++k) {
        const uint32_t p = base + (uint32_t)(k * 64 + lane);
        const bool live = p < n;
        const uint32_t q = live ? p : n - 1;
        const float old = (s == 1) ? f.init : f.dmin[q];
        float nm = old;
        if (live && p >= lo && p <= hi) {
          // dimension by dimension in the reference's order (dist = 0;
          // dist += d*d; and 0 + x == x for every x >= 0 and for NaN)
          const float *pp = f.pts + (size_t)p * f.sp;
          float acc = 0.0f;
          for (uint32_t c = 0; c < dim; ++c) {
            const float dc = pp[(size_t)c * f.ld] - s_pv[c];
            acc = (c == 0) ? dc * dc : acc + dc * dc;
          }
          nm = min_keep(acc, old);
          if (s == 1 || __float_as_int(nm) != __float_as_int(old)) f.dmin[p] = nm;
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

static const KernelEntry kNpduStream[] = {
    {0, 1024, 0, kTieLowest, (const void *)&fps_npdu_stream},
};

const KernelEntry *npdu_stream_kernel() { return &kNpduStream[0]; }

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

}  // namespace fpsx

// fps_kernels.hip -- persistent farthest-point-sampling kernels for gfx950.
//
// The N x M loop of fps_sampling (reference src/lib.cpp:188-246 and the
// multi-start form :111-186): M-1 passes, each updating every point's
// min-distance against the last pick and taking the global argmax.
//
// One launch runs the whole loop.  Workgroup g owns a fixed set of points for
// all M-1 passes; one pass per sample is
//   update  -> per-lane best -> DPP wave argmax -> LDS workgroup argmax
//           -> (G > 1) cross-workgroup exchange of one candidate per workgroup
// and every workgroup reduces the G candidates identically, so the whole grid
// agrees on the next pivot (index and coordinates) without a second round trip.
//
// Kernels:
//   fps_resident<D, B, PPT, TIE>  points and min-distances live in VGPRs
//                                 (PPT points per thread, B threads per WG);
//                                 state never touches HBM after the first load.
//   fps_stream<D, B, TIE>         points and min-distances stream from HBM every
//                                 pass (coalesced SoA loads), for clouds larger
//                                 than the register file; D == 0 is the generic
//                                 any-dimension form (pivot re-read from HBM).
// Both need all G workgroups co-resident (cooperative launch when G > 1),
// G <= B (the exchange polls with one lane per workgroup over ceil(G/64)
// waves) and G <= 512 (one exchange page per word array).
#include "fps_device.hpp"
#include "fps_kernels.hpp"

namespace fpsx {

// The whole grid as one set of participants (single-cloud kernels).
template <int TIE, int NW, int W>
__device__ __forceinline__ bool grid_exchange(const FpsParams &a, int (*cand)[W], int (*xw)[W],
                                              int gl, uint32_t e, int (&win)[NW]) {
  return exchange<TIE, NW, W>(a.xchg, a.replicas, kPageWords, gridDim.x, blockIdx.x, e, a.backoff,
                              a.spin_ticks, cand, xw, gl, win);
}

// ---------------------------------------------------------------------------
// Register-resident kernel.
// Thread t of workgroup g owns points i = g*B*PPT + k*B + t, k = 0..PPT-1
// (coalesced first load; ascending i within a thread as k ascends).
// ---------------------------------------------------------------------------
template <int D, int B, int PPT, int TIE>
__global__ void __launch_bounds__(B) fps_resident(FpsParams a) {
  constexpr int W = B / 64;
  constexpr int NW = 2 + D;
  __shared__ int s_cand[2][NW][W];
  __shared__ int s_xw[2][NW + 1][W];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t base = g * (uint32_t)(B * PPT);

  float c[PPT][D];
  float dm[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t i = base + (uint32_t)(k * B + tid);
    const bool live = i < a.n;
#pragma unroll
    for (int j = 0; j < D; ++j) c[k][j] = live ? a.pts[(size_t)i * a.sp + (size_t)j * a.ld] : 0.0f;
    dm[k] = live ? a.init : -1.0f;
  }
#pragma unroll
  for (int k = 0; k < PPT; ++k) settle<D>(c[k]);

  float pv[D];
  uint32_t pidx = load_forced<D>(a, 0, pv);
  if (g == 0 && tid >= B - 64) record_pick<D>(a, 0, pidx, pv);

  uint32_t e = 0;
  for (uint32_t s = 1; s < a.m; ++s) {
    // UPDATE against the last pick, tracking this lane's best (value, slot).
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
    if (s < a.n_starts) {
      pidx = load_forced<D>(a, s, pv);
      if (g == 0 && tid >= B - 64) record_pick<D>(a, s, pidx, pv);
      continue;
    }
    ++e;
    const int par = (int)(e & 1u);

    // Wave argmax.
    const uint32_t bidx = base + (uint32_t)(bk * B + tid);
    int wv;
    const int wl = wave_argmax_lane<TIE>(bv, bidx, wv);
    const int kw = readlane_i(bk, wl);
    const int widx = readlane_i((int)bidx, wl);
    float wc[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      float r = c[0][j];
#pragma unroll
      for (int q = 1; q < PPT; ++q) r = (kw == q) ? c[q][j] : r;
      wc[j] = readlane_f(r, wl);
    }
    if (lane == 0) {
      s_cand[par][0][wave] = wv;
      s_cand[par][1][wave] = widx;
#pragma unroll
      for (int j = 0; j < D; ++j) s_cand[par][2 + j][wave] = __float_as_int(wc[j]);
    }
    lds_barrier();

    // Workgroup argmax (every wave reduces the W wave candidates).
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
      if (!grid_exchange<TIE, NW, W>(a, s_cand[par], s_xw[par], gl, e, win)) {
        if (tid == 0) atomicOr(a.status, kStatusTimeout);
        return;
      }
      pidx = (uint32_t)win[1];
#pragma unroll
      for (int j = 0; j < D; ++j) pv[j] = __int_as_float(win[2 + j]);
    }
    // The last wave records the pick: it issues no global loads in this
    // loop, so the store's completion is never waited for.
    if (g == 0 && tid >= B - 64) record_pick<D>(a, s, pidx, pv);
  }
}

// ---------------------------------------------------------------------------
// HBM-streaming kernel.  Workgroup g owns the contiguous range
// [g*chunk, min(n, (g+1)*chunk)); min-distances persist in a.dmin (filled with
// a.init before the launch).  D == 0: any dimension a.dim, the pivot's
// coordinates are re-read from HBM into LDS after each exchange.
// ---------------------------------------------------------------------------
template <int D, int B, int TIE>
__global__ void __launch_bounds__(B) fps_stream(FpsParams a) {
  constexpr int W = B / 64;
  constexpr int DC = (D > 0) ? D : 1;       // coordinates carried in registers
  constexpr int NW = (D > 0) ? 2 + D : 2;   // words exchanged per candidate
  __shared__ int s_cand[2][NW][W];
  __shared__ int s_xw[2][NW + 1][W];
  extern __shared__ float s_pv_dyn[];       // D == 0: pivot coordinates (a.dim)

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint32_t G = gridDim.x;
  const uint32_t g = blockIdx.x;
  const uint32_t lo = g * a.chunk;
  const uint32_t hi = min(a.n, lo + a.chunk);
  const uint32_t dim = (D > 0) ? (uint32_t)D : a.dim;

  float pv[DC];
  uint32_t pidx;
  {
    uint64_t idx = a.starts[0];
    if (idx >= a.n) {
      if (tid == 0) atomicOr(a.status, kStatusBadStart);
      idx = 0;
    }
    pidx = (uint32_t)idx;
  }
  if (D > 0) {
#pragma unroll
    for (int j = 0; j < DC; ++j) pv[j] = a.pts[(size_t)pidx * a.sp + (size_t)j * a.ld];
    if (g == 0 && tid >= B - 64) record_pick<DC>(a, 0, pidx, pv);
  } else {
    for (uint32_t j = tid; j < dim; j += B) s_pv_dyn[j] = a.pts[(size_t)pidx * a.sp + (size_t)j * a.ld];
    lds_barrier();
    if (g == 0 && tid >= B - 64) record_pick_dyn(a, 0, pidx, s_pv_dyn, dim);
  }

  uint32_t e = 0;
  for (uint32_t s = 1; s < a.m; ++s) {
    int bv = INT_MIN;
    uint32_t bi = 0;
    float bc[DC];
#pragma unroll
    for (int j = 0; j < DC; ++j) bc[j] = 0.0f;
    for (uint32_t i = lo + tid; i < hi; i += B) {
      float d;
      float p[DC];
      if (D > 0) {
#pragma unroll
        for (int j = 0; j < DC; ++j) p[j] = a.pts[(size_t)i * a.sp + (size_t)j * a.ld];
        d = sqdist<DC>(p, pv);
      } else {
        const float *pi = a.pts + (size_t)i * a.sp;
        float df = pi[0] - s_pv_dyn[0];
        d = df * df;
        for (uint32_t j = 1; j < dim; ++j) {
          float dj = pi[(size_t)j * a.ld] - s_pv_dyn[j];
          d = d + dj * dj;
        }
      }
      const float old = a.dmin[i];
      const float nm = min_keep(d, old);
      if (__float_as_int(nm) != __float_as_int(old)) a.dmin[i] = nm;
      const int b = __float_as_int(nm);
      const bool take = (TIE == kTieHighest) ? (b >= bv) : (b > bv);
      if (take) {
        bv = b;
        bi = i;
        if (D > 0) {
#pragma unroll
          for (int j = 0; j < DC; ++j) bc[j] = p[j];
        }
      }
    }
    if (s < a.n_starts) {
      uint64_t idx = a.starts[s];
      if (idx >= a.n) {
        if (tid == 0) atomicOr(a.status, kStatusBadStart);
        idx = 0;
      }
      pidx = (uint32_t)idx;
      if (D > 0) {
#pragma unroll
        for (int j = 0; j < DC; ++j) pv[j] = a.pts[(size_t)pidx * a.sp + (size_t)j * a.ld];
      } else {
        lds_barrier();  // every lane is done reading the old pivot
        for (uint32_t j = tid; j < dim; j += B) s_pv_dyn[j] = a.pts[(size_t)pidx * a.sp + (size_t)j * a.ld];
        lds_barrier();
      }
      if (g == 0 && tid >= B - 64) {
        if (D > 0) record_pick<DC>(a, s, pidx, pv);
        else record_pick_dyn(a, s, pidx, s_pv_dyn, dim);
      }
      continue;
    }
    ++e;
    const int par = (int)(e & 1u);

    int wv;
    const int wl = wave_argmax_lane<TIE>(bv, bi, wv);
    const int widx = readlane_i((int)bi, wl);
    float wc[DC];
#pragma unroll
    for (int j = 0; j < DC; ++j) wc[j] = readlane_f(bc[j], wl);
    if (lane == 0) {
      s_cand[par][0][wave] = wv;
      s_cand[par][1][wave] = widx;
      if (D > 0) {
#pragma unroll
        for (int j = 0; j < DC; ++j) s_cand[par][2 + j][wave] = __float_as_int(wc[j]);
      }
    }
    lds_barrier();

    const int v2 = (lane < W) ? s_cand[par][0][lane] : INT_MIN;
    const uint32_t i2 = (lane < W) ? (uint32_t)s_cand[par][1][lane] : 0u;
    int gv;
    const int gl = wave_argmax_lane<TIE>(v2, i2, gv);

    if (G == 1) {
      pidx = (uint32_t)readlane_i((int)i2, gl);
      if (D > 0) {
#pragma unroll
        for (int j = 0; j < DC; ++j) pv[j] = __int_as_float(s_cand[par][2 + j][gl]);
      }
    } else {
      int win[NW];
      if (!grid_exchange<TIE, NW, W>(a, s_cand[par], s_xw[par], gl, e, win)) {
        if (tid == 0) atomicOr(a.status, kStatusTimeout);
        return;
      }
      pidx = (uint32_t)win[1];
      if (D > 0) {
#pragma unroll
        for (int j = 0; j < DC; ++j) pv[j] = __int_as_float(win[2 + j]);
      }
    }
    if (D == 0) {
      // Re-read the pivot's coordinates (uniform index, every lane agrees).
      // Every lane finished this pass's reads of s_pv_dyn before the
      // candidate barrier above, so the overwrite is safe.
      for (uint32_t j = tid; j < dim; j += B) s_pv_dyn[j] = a.pts[(size_t)pidx * a.sp + (size_t)j * a.ld];
      lds_barrier();
    }
    if (g == 0 && tid >= B - 64) {
      if (D > 0) record_pick<DC>(a, s, pidx, pv);
      else record_pick_dyn(a, s, pidx, s_pv_dyn, dim);
    }
  }
}

// Fill kernel: clears the exchange area / status and fills min-distances.
// (A kernel rather than hipMemsetAsync so callers may capture it in graphs.)
__global__ void fill_u64(uint64_t *p, size_t n, uint64_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_f32(float *p, size_t n, float v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// acc |= src: folds one launch's status word into a word that outlives it.
__global__ void or_u32(const uint32_t *src, uint32_t *acc) {
  if (threadIdx.x == 0) acc[0] |= src[0];
}

// Strided (point stride sp, coordinate stride sd) -> SoA (stride ld) copy.
__global__ void strided_to_soa(const float *__restrict__ src, uint64_t sp, uint64_t sd,
                               float *__restrict__ soa, uint32_t n, uint32_t dim, uint64_t ld) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < (size_t)n * dim;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t j = t / n;
    const size_t i = t - j * n;
    soa[j * ld + i] = src[i * sp + j * sd];
  }
}

// ---------------------------------------------------------------------------
// Instantiation tables.
// ---------------------------------------------------------------------------
#define FPSX_RES(D, B, P)                                   \
  {D, B, P, kTieHighest, (const void *)&fps_resident<D, B, P, kTieHighest>}, \
  {D, B, P, kTieLowest, (const void *)&fps_resident<D, B, P, kTieLowest>}

static const KernelEntry kResident[] = {
    // D = 3: the full set of shapes the tuner chooses from
    FPSX_RES(3, 256, 1), FPSX_RES(3, 256, 2), FPSX_RES(3, 256, 4), FPSX_RES(3, 256, 8),
    FPSX_RES(3, 256, 16), FPSX_RES(3, 256, 32), FPSX_RES(3, 512, 4), FPSX_RES(3, 512, 8),
    FPSX_RES(3, 1024, 1), FPSX_RES(3, 1024, 2), FPSX_RES(3, 1024, 4), FPSX_RES(3, 1024, 8),
    FPSX_RES(3, 1024, 16),
    // other small dimensions (shapes that fit 128 VGPRs without spilling)
    FPSX_RES(1, 1024, 4), FPSX_RES(1, 1024, 16), FPSX_RES(2, 1024, 4), FPSX_RES(2, 1024, 16),
    FPSX_RES(4, 1024, 4), FPSX_RES(4, 1024, 8), FPSX_RES(5, 1024, 4), FPSX_RES(5, 1024, 8),
    FPSX_RES(6, 1024, 4), FPSX_RES(6, 1024, 8), FPSX_RES(7, 1024, 4), FPSX_RES(8, 1024, 4),
};
#undef FPSX_RES

#define FPSX_STR(D, B) \
  {D, B, 0, kTieHighest, (const void *)&fps_stream<D, B, kTieHighest>}, \
  {D, B, 0, kTieLowest, (const void *)&fps_stream<D, B, kTieLowest>}

static const KernelEntry kStream[] = {
    FPSX_STR(0, 256), FPSX_STR(0, 1024), FPSX_STR(1, 1024), FPSX_STR(2, 1024),
    FPSX_STR(3, 256), FPSX_STR(3, 1024), FPSX_STR(4, 1024), FPSX_STR(5, 1024),
    FPSX_STR(6, 1024), FPSX_STR(7, 1024), FPSX_STR(8, 1024),
};
#undef FPSX_STR

const KernelEntry *resident_table(size_t *count) {
  *count = sizeof(kResident) / sizeof(kResident[0]);
  return kResident;
}
const KernelEntry *stream_table(size_t *count) {
  *count = sizeof(kStream) / sizeof(kStream[0]);
  return kStream;
}

hipError_t launch_fill_u64(uint64_t *p, size_t n, uint64_t v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  unsigned blocks = (unsigned)min((size_t)1024, (n + 255) / 256);
  hipLaunchKernelGGL(fill_u64, dim3(blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_f32(float *p, size_t n, float v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  unsigned blocks = (unsigned)min((size_t)2048, (n + 255) / 256);
  hipLaunchKernelGGL(fill_f32, dim3(blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}
hipError_t launch_or_u32(const uint32_t *src, uint32_t *acc, hipStream_t st) {
  hipLaunchKernelGGL(or_u32, dim3(1), dim3(64), 0, st, src, acc);
  return hipGetLastError();
}
hipError_t launch_strided_to_soa(const float *src, uint64_t sp, uint64_t sd, float *soa,
                                 uint32_t n, uint32_t dim, uint64_t ld, hipStream_t st) {
  size_t total = (size_t)n * dim;
  if (total == 0) return hipSuccess;
  unsigned blocks = (unsigned)min((size_t)4096, (total + 255) / 256);
  hipLaunchKernelGGL(strided_to_soa, dim3(blocks), dim3(256), 0, st, src, sp, sd, soa, n, dim, ld);
  return hipGetLastError();
}

}  // namespace fpsx

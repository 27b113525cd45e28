// fpsx_spatial.hip -- spatial order of a 3-D cloud for the pruned FPS kernel.
//
//   1. bounding box of the cloud (two-stage min/max, NaN ignored);
//   2. 30-bit Hilbert index of every point (10 bits per axis over the box;
//      Hilbert rather than Morton: the curve has no long jumps, so blocks of
//      consecutive points have smaller boxes and prune better);
//   3. radix sort of (code, original index)  -> nearby points become neighbours;
//   4. radix sort of (block of the position, original index): inside every
//      block of `block` positions the points are re-ordered by original index,
//      which keeps the kernel's tie-break (it compares positions inside a block);
//   5. gather of the coordinates into SoA in that order, and of the original
//      indices (perm).
// The order only decides which points share a wave: the picks never depend
// on it (ties are broken by original index everywhere).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "fps_pruned.hpp"
#include "fpsx.h"
#include "fpsx_curve.hpp"
#include "fpsx_internal.hpp"

namespace fpsx {

struct SpatialScratch {
  DevBuf part, box, keys, vals, keys_alt, vals_alt, keys64, keys64_alt, tmp;
};

SpatialScratch *new_spatial_scratch() { return new SpatialScratch(); }

void delete_spatial_scratch(SpatialScratch *S) {
  if (!S) return;
  for (DevBuf *b : {&S->part, &S->box, &S->keys, &S->vals, &S->keys_alt, &S->vals_alt, &S->keys64,
                    &S->keys64_alt, &S->tmp})
    b->free();
  delete S;
}

namespace {

constexpr int kBoxThreads = 256;
constexpr int kBoxBlocks = 512;

__global__ void __launch_bounds__(kBoxThreads)
    bbox_partial(const float *__restrict__ pts, uint64_t sp, uint64_t ld, uint32_t n,
                 float *__restrict__ part) {
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = blockIdx.x * kBoxThreads + threadIdx.x; i < n; i += gridDim.x * kBoxThreads) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float v = pts[(size_t)i * sp + (size_t)j * ld];
      lo[j] = fminf(lo[j], v);
      hi[j] = fmaxf(hi[j], v);
    }
  }
  __shared__ float s[6][kBoxThreads];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    s[j][threadIdx.x] = lo[j];
    s[3 + j][threadIdx.x] = hi[j];
  }
  __syncthreads();
  for (int w = kBoxThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        s[j][threadIdx.x] = fminf(s[j][threadIdx.x], s[j][threadIdx.x + w]);
        s[3 + j][threadIdx.x] = fmaxf(s[3 + j][threadIdx.x], s[3 + j][threadIdx.x + w]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// One block of kBoxThreads: thread t folds partials t, t + kBoxThreads, ...,
// then a tree reduction in LDS (the partials are few, the latency matters).
__global__ void __launch_bounds__(kBoxThreads)
    bbox_final(const float *__restrict__ part, uint32_t nparts, float *__restrict__ box) {
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t b = threadIdx.x; b < nparts; b += kBoxThreads) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      lo[j] = fminf(lo[j], part[b * 6 + j]);
      hi[j] = fmaxf(hi[j], part[b * 6 + 3 + j]);
    }
  }
  __shared__ float s[6][kBoxThreads];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    s[j][threadIdx.x] = lo[j];
    s[3 + j][threadIdx.x] = hi[j];
  }
  __syncthreads();
  for (int w = kBoxThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        s[j][threadIdx.x] = fminf(s[j][threadIdx.x], s[j][threadIdx.x + w]);
        s[3 + j][threadIdx.x] = fmaxf(s[3 + j][threadIdx.x], s[3 + j][threadIdx.x + w]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) box[threadIdx.x] = s[threadIdx.x][0];
}

__global__ void curve_codes(const float *__restrict__ pts, uint64_t sp, uint64_t ld, uint32_t n,
                             const float *__restrict__ box, uint32_t *__restrict__ keys,
                             uint32_t *__restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t qc[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float lo = box[j], hi = box[3 + j];
    const float span = hi - lo;
    const float scale = (span > 0.0f && span < INFINITY) ? 1023.0f / span : 0.0f;
    float q = (pts[(size_t)i * sp + (size_t)j * ld] - lo) * scale;
    if (!(q >= 0.0f)) q = 0.0f;  // NaN and anything below the box
    if (q > 1023.0f) q = 1023.0f;
    qc[j] = (uint32_t)q;
  }
  keys[i] = hilbert10(qc[0], qc[1], qc[2]);
  vals[i] = i;
}

__global__ void block_keys(const uint32_t *__restrict__ order, uint32_t n, uint32_t block,
                           uint64_t *__restrict__ keys64) {
  const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= n) return;
  keys64[pos] = ((uint64_t)(pos / block) << 32) | order[pos];
}

__global__ void gather_sorted(const float *__restrict__ pts, uint64_t sp, uint64_t ld, uint32_t n,
                              const uint64_t *__restrict__ keys64, float *__restrict__ sorted,
                              uint32_t *__restrict__ perm) {
  const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= n) return;
  const uint32_t i = (uint32_t)keys64[pos];
  perm[pos] = i;
#pragma unroll
  for (int j = 0; j < 3; ++j) sorted[(size_t)j * n + pos] = pts[(size_t)i * sp + (size_t)j * ld];
}

}  // namespace

#define HIP_OR_FAIL(expr)                                                            \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess)                                                            \
      return fail(_e == hipErrorOutOfMemory ? FPSX_ENOMEM : FPSX_EHIP,               \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                \
  } while (0)

int spatial_reserve(SpatialScratch &S, uint32_t n, uint32_t block) {
  if (n == 0 || block == 0) return fail(FPSX_EINVAL, "spatial_reserve: empty input");
  HIP_OR_FAIL(S.part.reserve(kBoxBlocks * 6 * sizeof(float)));
  HIP_OR_FAIL(S.box.reserve(6 * sizeof(float)));
  HIP_OR_FAIL(S.keys.reserve((size_t)n * 4));
  HIP_OR_FAIL(S.vals.reserve((size_t)n * 4));
  HIP_OR_FAIL(S.keys_alt.reserve((size_t)n * 4));
  HIP_OR_FAIL(S.vals_alt.reserve((size_t)n * 4));
  HIP_OR_FAIL(S.keys64.reserve((size_t)n * 8));
  HIP_OR_FAIL(S.keys64_alt.reserve((size_t)n * 8));
  // Temporary storage of the two sorts at the largest size (the smallest
  // block gives the most block-key bits).
  int hi_bit = 32;
  const uint32_t nblk = (n + block - 1) / block;
  while (hi_bit < 64 && (((uint64_t)1 << (hi_bit - 32)) < (uint64_t)nblk)) ++hi_bit;
  size_t t1 = 0, t2 = 0;
  HIP_OR_FAIL(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, (uint32_t *)S.keys.p,
                                                 (uint32_t *)S.keys_alt.p, (uint32_t *)S.vals.p,
                                                 (uint32_t *)S.vals_alt.p, (int)n, 0, 30));
  HIP_OR_FAIL(hipcub::DeviceRadixSort::SortKeys(nullptr, t2, (uint64_t *)S.keys64.p,
                                                (uint64_t *)S.keys64_alt.p, (int)n, 0, hi_bit));
  // With headroom; spatial_order still grows it should some size ask for more.
  HIP_OR_FAIL(S.tmp.reserve(std::max(t1, t2) * 2));
  return FPSX_OK;
}

int spatial_order(const float *pts, uint64_t sp, uint64_t ld, uint32_t n, uint32_t block,
                  float *sorted, uint32_t *perm, SpatialScratch &S, hipStream_t st) {
  if (n == 0 || block == 0) return fail(FPSX_EINVAL, "spatial_order: empty input");
  HIP_OR_FAIL(S.part.reserve(kBoxBlocks * 6 * sizeof(float)));
  HIP_OR_FAIL(S.box.reserve(6 * sizeof(float)));
  HIP_OR_FAIL(S.keys.reserve((size_t)n * 4));
  HIP_OR_FAIL(S.vals.reserve((size_t)n * 4));
  HIP_OR_FAIL(S.keys_alt.reserve((size_t)n * 4));
  HIP_OR_FAIL(S.vals_alt.reserve((size_t)n * 4));
  HIP_OR_FAIL(S.keys64.reserve((size_t)n * 8));
  HIP_OR_FAIL(S.keys64_alt.reserve((size_t)n * 8));
  const uint32_t nblk = (n + block - 1) / block;
  int hi_bit = 32;
  while (hi_bit < 64 && (((uint64_t)1 << (hi_bit - 32)) < (uint64_t)nblk)) ++hi_bit;
  size_t t1 = 0, t2 = 0;
  HIP_OR_FAIL(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, (uint32_t *)S.keys.p,
                                                 (uint32_t *)S.keys_alt.p, (uint32_t *)S.vals.p,
                                                 (uint32_t *)S.vals_alt.p, (int)n, 0, 30, st));
  HIP_OR_FAIL(hipcub::DeviceRadixSort::SortKeys(nullptr, t2, (uint64_t *)S.keys64.p,
                                                (uint64_t *)S.keys64_alt.p, (int)n, 0, hi_bit, st));
  HIP_OR_FAIL(S.tmp.reserve(std::max(t1, t2)));
  const unsigned nb = (unsigned)std::min<uint64_t>(kBoxBlocks, (n + kBoxThreads - 1) / kBoxThreads);
  hipLaunchKernelGGL(bbox_partial, dim3(nb), dim3(kBoxThreads), 0, st, pts, sp, ld, n,
                     (float *)S.part.p);
  HIP_OR_FAIL(hipGetLastError());
  hipLaunchKernelGGL(bbox_final, dim3(1), dim3(kBoxThreads), 0, st, (const float *)S.part.p, nb,
                     (float *)S.box.p);
  HIP_OR_FAIL(hipGetLastError());
  const unsigned g = (n + 255) / 256;
  hipLaunchKernelGGL(curve_codes, dim3(g), dim3(256), 0, st, pts, sp, ld, n,
                     (const float *)S.box.p, (uint32_t *)S.keys.p, (uint32_t *)S.vals.p);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipcub::DeviceRadixSort::SortPairs(S.tmp.p, t1, (uint32_t *)S.keys.p,
                                                 (uint32_t *)S.keys_alt.p, (uint32_t *)S.vals.p,
                                                 (uint32_t *)S.vals_alt.p, (int)n, 0, 30, st));
  hipLaunchKernelGGL(block_keys, dim3(g), dim3(256), 0, st, (const uint32_t *)S.vals_alt.p, n,
                     block, (uint64_t *)S.keys64.p);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipcub::DeviceRadixSort::SortKeys(S.tmp.p, t2, (uint64_t *)S.keys64.p,
                                                (uint64_t *)S.keys64_alt.p, (int)n, 0, hi_bit, st));
  hipLaunchKernelGGL(gather_sorted, dim3(g), dim3(256), 0, st, pts, sp, ld, n,
                     (const uint64_t *)S.keys64_alt.p, sorted, perm);
  HIP_OR_FAIL(hipGetLastError());
  return FPSX_OK;
}

}  // namespace fpsx

// fps_device.hpp -- device-side building blocks of the FPS kernels (gfx950).
//
// Values: a min-distance is a float32 in [+0, +inf] (never NaN, never -0), so
// its bit pattern read as int32 orders exactly like the float.  Padding lanes
// carry -1.0f, whose int32 pattern is negative and loses to every real value.
//
// Exchange area (device memory, uint64 words).  Word k of workgroup g's
// candidate for exchange buffer buf, replica r, lives at
//   xchg[((buf * R + r) * kSlotWords + k) * kPageWords + g]
// so every word array sits in its own 4 KiB page: the polls of all
// workgroups spread over many memory channels instead of hammering the one
// or two pages a packed layout would occupy.  A workgroup writes its
// candidate to every replica and reads replica (g mod R).
//   k = 0 value bits, k = 1 point index, k = 2 + j coordinate j bits.
// Each word is (payload << 32) | tag with tag = exchange number (1, 2, ...).
// Every 64-bit word carries its own tag, so a reader never needs ordering
// between words: it polls until every word it needs shows the tag it expects.
// Exchange e uses buffer e & 1.  A workgroup writes exchange e only after it
// has read every candidate of exchange e - 1, which every workgroup publishes
// only after finishing its reads of exchange e - 2: the buffer being
// overwritten is never still being read.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fps_kernels.hpp"

namespace fpsx {

constexpr int kTieHighest = 0;
constexpr int kTieLowest = 1;
constexpr int kSlotWords = 16;  // value, index, up to 14 coordinates
constexpr int kMaxXchgDim = kSlotWords - 2;
constexpr int kPageWords = 512;  // one 4 KiB page per word array: G <= 512
constexpr int kMaxReplicas = 8;

// Status bits written by kernels (device word, cleared before each launch).
constexpr uint32_t kStatusTimeout = 1u;

// ---------------------------------------------------------------------------
// DPP wave reductions (64-lane wavefront).  Result is read from lane 63,
// so every lane returns the same (uniform) value.
//   quad_perm[1,0,3,2], quad_perm[2,3,0,1]  -> max within each quad
//   row_shr:4, row_shr:8                    -> lane 15 of a row holds the row max
//   row_bcast:15 (rows 1,3), row_bcast:31 (rows 2,3) -> lane 63 holds the wave max
// ---------------------------------------------------------------------------
// old = the identity of max, so lanes a DPP move cannot source from (and rows
// the row mask leaves out) contribute nothing; this form lets the compiler
// fold each move into the max instruction (v_max_*_dpp).
#define FPSX_DPP(id, v, ctrl, rmask) __builtin_amdgcn_update_dpp((id), (v), (ctrl), (rmask), 0xf, false)

__device__ __forceinline__ int wave_max_i32(int v) {
  constexpr int kId = INT_MIN;
  v = max(v, FPSX_DPP(kId, v, 0xb1, 0xf));   // quad_perm [1,0,3,2]
  v = max(v, FPSX_DPP(kId, v, 0x4e, 0xf));   // quad_perm [2,3,0,1]
  v = max(v, FPSX_DPP(kId, v, 0x114, 0xf));  // row_shr:4
  v = max(v, FPSX_DPP(kId, v, 0x118, 0xf));  // row_shr:8
  v = max(v, FPSX_DPP(kId, v, 0x142, 0xa));  // row_bcast:15
  v = max(v, FPSX_DPP(kId, v, 0x143, 0xc));  // row_bcast:31
  return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t u) {
  uint32_t v = u;
  v = max(v, (uint32_t)FPSX_DPP(0, (int)v, 0xb1, 0xf));
  v = max(v, (uint32_t)FPSX_DPP(0, (int)v, 0x4e, 0xf));
  v = max(v, (uint32_t)FPSX_DPP(0, (int)v, 0x114, 0xf));
  v = max(v, (uint32_t)FPSX_DPP(0, (int)v, 0x118, 0xf));
  v = max(v, (uint32_t)FPSX_DPP(0, (int)v, 0x142, 0xa));
  v = max(v, (uint32_t)FPSX_DPP(0, (int)v, 0x143, 0xc));
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

#undef FPSX_DPP

__device__ __forceinline__ int readlane_i(int v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}
__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// True when candidate (v, i) beats (bv, bi) under the tie rule.
template <int TIE>
__device__ __forceinline__ bool beats(int v, uint32_t i, int bv, uint32_t bi) {
  if (TIE == kTieHighest)
    return v > bv || (v == bv && i > bi);
  return v > bv || (v == bv && i < bi);
}

// Lane holding the wave's best (value, index) under the tie rule; the wave's
// best value is returned through wv.  Lanes without a candidate pass INT_MIN.
// The common case -- one lane holds the maximum -- costs one value reduction
// and a ballot; only real ties pay for an index reduction.
template <int TIE>
__device__ __forceinline__ int wave_argmax_lane(int v, uint32_t idx, int &wv) {
  wv = wave_max_i32(v);
  const uint64_t tied = __ballot(v == wv);
  if (__popcll(tied) == 1)
    return __ffsll((unsigned long long)tied) - 1;
  const uint32_t key = (v == wv) ? (TIE == kTieHighest ? idx : ~idx) : 0u;
  const uint32_t km = wave_max_u32(key);
  const uint32_t widx = (TIE == kTieHighest) ? km : ~km;
  return __ffsll((unsigned long long)__ballot(v == wv && idx == widx)) - 1;
}

// Exact `d < m ? d : m` for d in [+0, +inf] or NaN and m never NaN.
// v_min_f32 returns the non-NaN operand for a quiet NaN (IEEE mode), and a
// NaN distance produced by arithmetic is always quiet, so one instruction
// reproduces the reference's `if (dist < dist_min[i]) dist_min[i] = dist`.
__device__ __forceinline__ float min_keep(float d, float m) {
  float r;
  asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(d), "v"(m));
  return r;
}

// Squared distance in the reference's operand order: ((d0*d0 + d1*d1) + d2*d2)...,
// every product and sum separately rounded (compiled with -ffp-contract=off).
template <int D>
__device__ __forceinline__ float sqdist(const float (&p)[D], const float (&q)[D]) {
  float df = p[0] - q[0];
  float acc = df * df;
#pragma unroll
  for (int j = 1; j < D; ++j) {
    float dj = p[j] - q[j];
    acc = acc + dj * dj;
  }
  return acc;
}

// Select arr[k] for a wave-uniform k.
template <int PPT>
__device__ __forceinline__ float pick_uniform(const float (&arr)[PPT], int k) {
  float r = arr[0];
#pragma unroll
  for (int q = 1; q < PPT; ++q)
    r = (k == q) ? arr[q] : r;
  return r;
}

// Agent-scope (cross-XCD coherent) relaxed 64-bit word access.
__device__ __forceinline__ void xchg_store(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t xchg_load(const uint64_t *p) {
  return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t now_ticks() { return wall_clock64(); }

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for
// every outstanding global store of the wave (vmcnt counts stores on gfx9),
// which would put a memory round trip on every pass; the kernels never need
// global memory ordered across this barrier.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// ---------------------------------------------------------------------------
// Cross-workgroup step among G participants (a whole grid, or a team of
// workgroups working on one cloud), entered by every wave of participant g
// once the workgroup's candidate is known (word k of it in cand[k][gl],
// k < NW).  Word k of participant g's candidate for buffer buf, replica r
// lives at base[((buf * R + r) * kSlotWords + k) * pitch + g] (pitch >= G).
// Wave 0 publishes to every replica the candidate size allows; waves
// 0..ceil(G/64)-1 (<= W, guaranteed by the host) each poll 64 slots of replica
// g mod R, one slot per lane with all NW words loaded at once, reduce them and
// leave their winner in xw[k][wave] (xw[NW][wave] = arrived in time); after
// one LDS barrier every wave reduces those winners.  On return win[k] holds
// word k of the winner in every lane.  Returns false when some candidate did
// not arrive within spin_ticks.
// ---------------------------------------------------------------------------
template <int TIE, int NW, int W>
__device__ __forceinline__ bool exchange(uint64_t *base, uint32_t R, uint32_t pitch, uint32_t G,
                                         uint32_t g, uint32_t e, uint32_t backoff,
                                         uint64_t spin_ticks, int (*cand)[W], int (*xw)[W],
                                         int gl, int (&win)[NW]) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int P = (int)((G + 63u) >> 6);
  const int K = min((int)R, 64 / NW);  // replicas this candidate size allows
  uint64_t *bufs = base + (size_t)(e & 1u) * R * kSlotWords * pitch;
  if (wave == 0 && lane < K * NW) {
    const int rep = lane / NW;
    const int k = lane - rep * NW;
    xchg_store(bufs + ((size_t)rep * kSlotWords + k) * pitch + g,
               ((uint64_t)(uint32_t)cand[k][gl] << 32) | e);
  }
  uint64_t *buf = bufs + (size_t)(g % (uint32_t)K) * kSlotWords * pitch;
  if (wave < P) {
    const uint32_t q = (uint32_t)(wave * 64 + lane);
    int cv = INT_MIN;
    uint32_t ci = 0;
    int cw[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) cw[k] = 0;
    bool ok = true;
    if (q < G) {
      uint64_t w[NW];
      const uint64_t deadline = now_ticks() + spin_ticks;
      for (;;) {
        bool all = true;
#pragma unroll
        for (int k = 0; k < NW; ++k) w[k] = xchg_load(buf + (size_t)k * pitch + q);
#pragma unroll
        for (int k = 0; k < NW; ++k) all = all && ((uint32_t)w[k] == e);
        if (all) break;
        if (now_ticks() > deadline) {
          ok = false;
          break;
        }
        for (uint32_t z = 0; z < backoff; ++z) __builtin_amdgcn_s_sleep(1);
      }
      if (ok) {
        cv = (int)(w[0] >> 32);
        ci = (uint32_t)(w[1] >> 32);
#pragma unroll
        for (int k = 0; k < NW; ++k) cw[k] = (int)(w[k] >> 32);
      }
    }
    int xv;
    const int xl = wave_argmax_lane<TIE>(cv, ci, xv);
    const bool wave_ok = __ballot(!ok) == 0;
    int word = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const int x = readlane_i(cw[k], xl);
      if (lane == k) word = x;
    }
    if (lane < NW) xw[lane][wave] = word;
    if (lane == NW) xw[NW][wave] = wave_ok ? 1 : 0;
  }
  lds_barrier();
  if (P == 1) {
    // One polling wave: its winner is the result; every word is read at once.
    int x[NW + 1];
#pragma unroll
    for (int k = 0; k <= NW; ++k) x[k] = xw[k][0];
#pragma unroll
    for (int k = 0; k < NW; ++k) win[k] = x[k];
    return x[NW] != 0;
  }
  bool ok_all = true;
  for (int r = 0; r < P; ++r) ok_all = ok_all && (xw[NW][r] != 0);
  const int v3 = (lane < P) ? xw[0][lane] : INT_MIN;
  const uint32_t i3 = (lane < P) ? (uint32_t)xw[1][lane] : 0u;
  int gv;
  const int wl = wave_argmax_lane<TIE>(v3, i3, gv);
#pragma unroll
  for (int k = 0; k < NW; ++k) win[k] = xw[k][wl];
  return ok_all;
}

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
// The gap along j is max(p - hi, lo - p, 0), written without branches: it is
// p - hi when p > hi, lo - p when p < lo and 0 otherwise, as in the branchy
// form, including NaN p (both differences NaN, max returns 0) and an empty
// box (lo = +inf, hi = -inf: gap +inf).
template <int D>
__device__ __forceinline__ float box_bound(const float (&p)[D], const float (&lo)[D],
                                           const float (&hi)[D]) {
  float acc = 0.0f;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const float g = fmaxf(fmaxf(p[j] - hi[j], lo[j] - p[j]), 0.0f);
    acc = (j == 0) ? g * g : acc + g * g;
  }
  return acc;
}

// Records pick s: lane 0 of the calling wave writes the index and, when the
// caller asked for them, the coordinates (the fused `pc[idx]` gather).  The
// coordinates are stored one by one from lane 0: selecting pv[j] by lane
// number would index pv dynamically, and the compiler would then move pv
// out of registers into LDS for the whole kernel.
template <int D>
__device__ __forceinline__ void record_pick(const FpsParams &a, uint32_t s, uint32_t pidx,
                                            const float (&pv)[D]) {
  if ((threadIdx.x & 63) == 0) {
    a.out[s] = pidx;
    if (a.out_pts != nullptr) {
#pragma unroll
      for (int j = 0; j < D; ++j) a.out_pts[(size_t)s * D + j] = pv[j];
    }
  }
}

// Same for the any-dimension kernel, coordinates read from LDS.
__device__ __forceinline__ void record_pick_dyn(const FpsParams &a, uint32_t s, uint32_t pidx,
                                                const float *pv, uint32_t dim) {
  const int lane = threadIdx.x & 63;
  if (lane == 0) a.out[s] = pidx;
  if (a.out_pts != nullptr)
    for (uint32_t j = lane; j < dim; j += 64) a.out_pts[(size_t)s * dim + j] = pv[j];
}

// Marks registers filled by global loads as settled.  The compiler's wait
// counters are merged over a loop's entry and back edge, so a register still
// awaiting a load when a pass loop is entered makes the loop header wait for
// every outstanding memory operation on every pass -- including the store
// that recorded the previous pick, a full memory round trip on the critical
// path of the recording wave.  An empty asm that reads and rewrites the
// register moves that wait to this point, once.
template <int N>
__device__ __forceinline__ void settle(float (&v)[N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) asm volatile("" : "+v"(v[j]));
}

// Coordinates of a forced pick (multi-start prefix).  An index outside the
// cloud is reported through the status word and replaced by 0, so no access
// ever leaves the arrays.
template <int D>
__device__ __forceinline__ uint32_t load_forced(const FpsParams &a, uint32_t s, float (&pv)[D]) {
  uint64_t idx = a.starts[s];
  if (idx >= a.n) {
    if (threadIdx.x == 0) atomicOr(a.status, kStatusBadStart);
    idx = 0;
  }
#pragma unroll
  for (int j = 0; j < D; ++j) pv[j] = a.pts[(size_t)idx * a.sp + (size_t)j * a.ld];
  settle<D>(pv);
  return (uint32_t)idx;
}

}  // namespace fpsx

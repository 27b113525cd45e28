// fps_kernels.hpp -- host-visible description of the FPS kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fpsx {

// Status bits a kernel may set in FpsParams::status.
constexpr uint32_t kStatusBadStart = 2u;  // a start index >= n (replaced by 0)

// Kernel arguments (passed by value).
struct FpsParams {
  const float *pts;       // coordinate j of point i at pts[i * sp + j * ld]
  uint64_t ld;            // stride between coordinates (SoA: >= n; row-major: 1)
  uint64_t sp;            // stride between points (SoA: 1; row-major: dim)
  uint32_t n;             // number of points (< 2^31)
  uint32_t m;             // number of samples (1..n)
  const uint64_t *starts; // n_starts forced picks (device)
  uint32_t n_starts;      // >= 1
  uint32_t dim;           // runtime dimension (generic streaming kernel)
  uint64_t *out;          // m indices (device)
  float *out_pts;         // optional m x dim row-major coordinates of the picks (or null)
  uint64_t *xchg;         // exchange area, 2 * kSlotWords * G words, zeroed before launch
  uint32_t *status;       // status word, zeroed before launch
  float *dmin;            // streaming kernels: n min-distances, filled with init
  uint32_t chunk;         // streaming kernels: points per workgroup
  float init;             // initial min-distance
  uint64_t spin_ticks;    // realtime ticks an exchange may wait before giving up
  uint32_t replicas;      // exchange replicas (1..8); each candidate is written to all
  uint32_t backoff;       // s_sleep rounds between unanswered polls (0 = spin)
};

// Exchange-area words the host must provide for a grid of G workgroups.
inline size_t xchg_words(uint32_t replicas) {
  return (size_t)2 * replicas * 16 /*kSlotWords*/ * 512 /*kPageWords*/;
}

struct KernelEntry {
  int dim;    // 0 = any dimension
  int block;  // threads per workgroup
  int ppt;    // points per thread (resident kernels), 0 for streaming
  int tie;    // 0 highest, 1 lowest
  const void *fn;
};

const KernelEntry *resident_table(size_t *count);
const KernelEntry *stream_table(size_t *count);

// Enqueue helpers; each returns the launch status.
hipError_t launch_fill_u64(uint64_t *p, size_t n, uint64_t v, hipStream_t st);
hipError_t launch_fill_f32(float *p, size_t n, float v, hipStream_t st);
// acc[0] |= src[0] on the stream (one status word folded into another).
hipError_t launch_or_u32(const uint32_t *src, uint32_t *acc, hipStream_t st);
// Copy a strided point array (coordinate j of point i at src[i*sp + j*sd])
// into SoA layout (soa[j*ld + i]).
hipError_t launch_strided_to_soa(const float *src, uint64_t sp, uint64_t sd, float *soa,
                                 uint32_t n, uint32_t dim, uint64_t ld, hipStream_t st);

}  // namespace fpsx

// fpsx_curve.hpp -- the 3-D Hilbert index of the spatial ordering, usable on
// the device (fpsx_spatial.hip) and on the host (the CPU test that checks it).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define FPSX_HD __host__ __device__ __forceinline__
#else
#define FPSX_HD inline
#endif

namespace fpsx {

// Hilbert index of (x, y, z), 10 bits each, by Skilling's transform
// ("Programming the Hilbert curve", AIP Conf. Proc. 707, 2004): undo the
// excess work axis by axis, Gray-encode, then interleave the bits with x the
// most significant of every triple.  A bijection of the 2^30 grid on
// [0, 2^30) in which consecutive indices are unit steps.
FPSX_HD uint32_t hilbert10(uint32_t x, uint32_t y, uint32_t z) {
  uint32_t X[3] = {x & 0x3ffu, y & 0x3ffu, z & 0x3ffu};
  for (uint32_t Q = 1u << 9; Q > 1u; Q >>= 1) {
    const uint32_t P = Q - 1u;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  uint32_t t = 0;
  for (uint32_t Q = 1u << 9; Q > 1u; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1u;
  X[0] ^= t;
  X[1] ^= t;
  X[2] ^= t;
  uint32_t h = 0;
  for (int b = 9; b >= 0; --b)
    for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
  return h;
}

}  // namespace fpsx

// kdline.hpp -- KD-line permutation of the bucket FPS entry (host code).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace fpsx {

// Rebuilds the point order of the reference's KD-line tree of height
// `height` over row-major n x d float32 `rows`.  perm[pos] = original index of
// the point at position pos.  Optionally returns the permuted points as
// columns (coordinate j of position p at [j * n + p]) and the bucket ranges
// [lo, hi) in bucket order.  Returns 0, or -1 on bad sizes.
int kdline_permutation(const float *rows, size_t n, size_t d, size_t height,
                       std::vector<uint64_t> &perm, std::vector<float> *permuted_cols,
                       std::vector<uint64_t> *bucket_lo, std::vector<uint64_t> *bucket_hi);

}  // namespace fpsx

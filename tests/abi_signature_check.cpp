// Compile-time check that include/fpsx.h declares bucket_fps_kdline with the
// type of the reference's C entry (src/wrapper.hpp:117-119):
//   int bucket_fps_kdline(const float *raw_data, size_t n_points, size_t dim,
//                         size_t n_samples, size_t start_idx, size_t height,
//                         size_t *sampled_point_indices);
// Built by tests/test_boundary_cpu.py with -fsyntax-only.  With
// -DWITH_REFERENCE_HEADER it also includes the reference's own wrapper.hpp,
// whose extern "C" definition must then agree with this declaration, or the
// translation unit does not compile (conflicting C linkage declarations).
#include <cstddef>
#include <type_traits>

#include "fpsx.h"

#ifdef WITH_REFERENCE_HEADER
#include "wrapper.hpp"
#endif

static_assert(std::is_same<decltype(&bucket_fps_kdline),
                           int (*)(const float *, size_t, size_t, size_t, size_t, size_t,
                                   size_t *)>::value,
              "bucket_fps_kdline must keep the reference's signature");
static_assert(sizeof(size_t) == sizeof(uint64_t),
              "the C ABI hands size_t picks to callers as uint64");

int main() { return 0; }

// fps_pruned.hpp -- host-visible description of the pruned FPS kernel and of
// the spatial ordering it needs.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fps_kernels.hpp"

namespace fpsx {

struct PrunedParams {
  FpsParams fps;          // the cloud (original order), picks, exchange, status
  const float *sorted;    // SoA coordinates in spatial order: sorted[j * sorted_ld + pos]
  uint64_t sorted_ld;
  const uint32_t *perm;   // original index of every position of the spatial order
};

const KernelEntry *pruned_table(size_t *count);
// One workgroup paired at the wave level (same parameters, G = 1): up to
// two picks per LDS barrier.
const KernelEntry *solo_table(size_t *count);
// Paired form (same parameters): up to two picks per exchange; 2 <= G <= 64.
const KernelEntry *pair_table(size_t *count);
constexpr uint32_t kPairMaxGrid = 64;
// Paired form with 2 or 4 polling waves (G <= 64 * polling_waves), 1024 x 16
// only; null for other shapes.
const KernelEntry *pair_wide_kernel(int block, int ppt, int tie, int polling_waves);
constexpr uint32_t kPairWideMaxGrid = 256;

// Streaming pruned kernel: blocks of kSBlock positions in spatial order,
// kSMetaWords LDS words per block (box, cached candidate, list entry).
// 512 positions (lattice16m at 64 workgroups: 256, 512 and 1024 take 21.3,
// 20.3 and 23.9 ms).
constexpr int kSBlock = 512;
constexpr int kSMetaWords = 12;
const KernelEntry *stream_pruned_table(size_t *count);
// Paired streaming form (same parameters; 16 LDS words per block: 16-byte
// records of the box and cached candidate, the candidate's value and index
// again, a list entry and the pass that last listed it): up to two picks per
// exchange; 2 <= G <= kPairMaxGrid.
constexpr int kSPairMetaWords = 16;
const KernelEntry *stream_pair_table(size_t *count);

// Device scratch of the spatial ordering (grow-only, owned by the caller).
struct SpatialScratch;

// Enqueue on `st`: the spatial order of the n points of `pts` (coordinate j of
// point i at pts[i*sp + j*ld], dim == 3), grouped in blocks of `block` positions
// whose members are sorted by original index; writes sorted (SoA, stride n)
// and perm.  Returns a FPSX_* code.
int spatial_order(const float *pts, uint64_t sp, uint64_t ld, uint32_t n, uint32_t block,
                  float *sorted, uint32_t *perm, SpatialScratch &scratch, hipStream_t st);
// Size the scratch for clouds of up to n_max points (blocks of `block`), so
// later spatial_order calls up to that size allocate nothing.
int spatial_reserve(SpatialScratch &scratch, uint32_t n_max, uint32_t block);

}  // namespace fpsx

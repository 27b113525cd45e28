// fps_npdu.hpp -- host-visible description of the windowed-update (NPDU) FPS kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fps_kernels.hpp"

namespace fpsx {

struct NpduParams {
  FpsParams fps;    // the cloud, the single start (starts[0]), picks, exchange, status
  uint64_t window;  // update window w: pick p updates indices p - w/2 .. p + w/2 (shifted into range)
};

// Entries: dim 3 (coordinates in registers) and dim 0 (any dimension,
// coordinates read from memory for the window only); tie is always lowest.
const KernelEntry *npdu_table(size_t *count);
// Paired form (dim 3): up to two picks per exchange; 2 <= G <= 64.
const KernelEntry *npdu_pair_table(size_t *count);
constexpr uint32_t kNpduPairMaxGrid = 64;

// Streaming forms (ppt 0): blocks of kNpduBlock indices; LDS holds 2 words
// per block of the workgroup plus dim floats of pivot (any dimension), or 5
// words per block (3-D).  256 indices (4 M 3-D points, 4096 samples, 64
// workgroups: blocks of 1024, 512 and 256 take 19.3, 15.0 and 13.5 ms).
constexpr int kNpduBlock = 256;
// Streaming windowed-update kernel: any dimension, or the 3-D form that
// carries the winner's coordinates through the exchange (both with blocks of
// kNpduBlock indices); the any-dimension kernel with blocks of kNpduBigBlock
// for clouds whose per-workgroup cache would not fit LDS otherwise.
constexpr int kNpduBigBlock = 1024;
const KernelEntry *npdu_stream_kernel(bool three_d);
const KernelEntry *npdu_stream_kernel_big();

}  // namespace fpsx

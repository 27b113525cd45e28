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

// Streaming form (any dimension, ppt 0): blocks of kNpduBlock indices; LDS
// holds dim floats of pivot plus 2 words per block of the workgroup.
constexpr int kNpduBlock = 1024;
// Streaming windowed-update kernel: any dimension, or the 3-D form that
// carries the winner's coordinates through the exchange.
const KernelEntry *npdu_stream_kernel(bool three_d);

}  // namespace fpsx

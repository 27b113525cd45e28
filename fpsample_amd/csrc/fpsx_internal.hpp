// fpsx_internal.hpp -- host runtime internals shared by the C-ABI sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "fps_kernels.hpp"

#define FPSX_VERSION_STRING "0.1.0"

namespace fpsx {

// Grow-only device buffer.  Buffers of the per-device contexts live until the
// process ends; owners with a shorter life call free().
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t bytes);
  void free();
  ~DevBuf();
};

enum class PlanKind : int { Resident = 0, Stream = 1, Pruned = 2, Npdu = 3, StreamPruned = 4, Pair = 5 };

struct Plan {
  PlanKind kind = PlanKind::Resident;
  const KernelEntry *k = nullptr;
  uint32_t block = 0;
  uint32_t grid = 0;
  uint32_t chunk = 0;  // points per workgroup
  size_t shmem = 0;    // dynamic LDS bytes
};

// One per device, created on first use and kept until the process ends.
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  int clock_khz = 100000;  // realtime clock of s_memrealtime
  uint64_t spin_ticks = 0;
  int replicas = 8;  // exchange replicas (FPSX_XCHG_REPLICAS)
  int backoff = 0;   // poll backoff rounds (FPSX_XCHG_BACKOFF)
  uint32_t *host_status = nullptr;  // pinned
  std::mutex mu;
  DevBuf pts, rows, starts, out, xchg, status, dmin, sorted, perm;
  struct SpatialScratch *spatial = nullptr;
  std::map<std::pair<const void *, size_t>, int> occ;  // co-resident workgroups per kernel
  int max_blocks(const KernelEntry *k, size_t shmem);
  // Autotuned plans: (n_points, dim, tie) -> the plan the tuner chose.
  std::map<std::tuple<size_t, size_t, int>, Plan> tuned;
  // Scratch hand-over between streams: recorded on the stream of the last
  // call that used the scratch buffers above; a call on another stream waits
  // for it first (scratch_acquire / scratch_release).
  hipEvent_t scratch_ev = nullptr;
  hipStream_t scratch_owner = nullptr;
  bool scratch_recorded = false;
  // The last FPS launch (for fpsx_launch_exchanges).
  Plan last_plan;
  size_t last_m = 0, last_starts = 0;
  // Optional per-launch timing of the FPS kernel (HIP events on its stream).
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  size_t events_used = 0;
};

int fail(int code, const std::string &msg);
int get_ctx(Ctx **out);
// Kernel and grid for a cloud of n points of dimension dim and m samples.
int plan(Ctx &c, size_t n, size_t dim, int tie, size_t m, Plan &p);
// Enqueue one FPS run on device data: coordinate j of point i at
// d_pts[i*sp + j*ld].  d_out_pts (may be null) receives the picks' coordinates.
int launch_fps(Ctx &c, const Plan &p, const float *d_pts, size_t sp, size_t ld, size_t n,
               size_t dim, size_t m, const uint64_t *d_starts, size_t n_starts, uint64_t *d_out,
               float *d_out_pts, float init, hipStream_t st);
// Windowed-update (NPDU) runs: plan, and enqueue on device data (start d_start[0]).
int plan_npdu(Ctx &c, size_t n, size_t dim, Plan &p);
int launch_npdu(Ctx &c, const Plan &p, const float *d_pts, size_t sp, size_t ld, size_t n,
                size_t dim, size_t m, size_t window, const uint64_t *d_start, uint64_t *d_out,
                float *d_out_pts, hipStream_t st);
int status_code(uint32_t st);
// Order a call on stream `st` after the previous user of the context's
// scratch buffers (another stream), and mark `st` as their user once the
// call's work is enqueued.  Both are no-ops on a stream under graph capture.
int scratch_acquire(Ctx &c, hipStream_t st);
int scratch_release(Ctx &c, hipStream_t st);
struct SpatialScratch *new_spatial_scratch();
int fps_host(const float *points, size_t n, size_t dim, int layout, size_t m,
             const uint64_t *starts, size_t n_starts, uint64_t *out, int tie, float init);

}  // namespace fpsx

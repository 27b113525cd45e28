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

enum class PlanKind : int {
  Resident = 0, Stream = 1, Pruned = 2, Npdu = 3, StreamPruned = 4, Pair = 5, StreamPair = 6
};

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

// Per-pass time model of the resident kernel fitted to MI355X measurements
// (microseconds): reduction chain ~0.55, update 12 VALU ops per point over 64
// lanes/clock at 2.4 GHz, exchange ~1.15 + 0.0125 * G (1 M points: 3.39 /
// 3.78 / 4.38 us measured at G = 64 / 128 / 256; the model keeps that order).
inline double resident_model_us(size_t n, size_t g) {
  const double per_wg = (double)n / (double)g;
  return 0.55 + per_wg * 12.0 / (64.0 * 2400.0) + (g > 1 ? 1.15 + 0.0125 * (double)g : 0.0);
}

// Grid of a resident-shaped kernel k (points in registers) for n points, or 0
// when it cannot hold them: more workgroups than co-resident (max_blocks),
// than the exchange polls (one per lane of B), or than exchange pages.
inline uint32_t resident_grid(size_t n, const KernelEntry *k, size_t max_blocks) {
  const size_t per = (size_t)k->block * k->ppt;
  const size_t g = (n + per - 1) / per;
  if (g > 1 && g > max_blocks) return 0;
  if (g > (size_t)k->block || g > 512 /* kPageWords */) return 0;
  return (uint32_t)g;
}

// The register-resident kernel for a cloud, without spatial order or HBM
// state: up to 16 K points one workgroup with the smallest shape that holds
// the cloud; beyond, the feasible shape the per-pass model prefers.
// max_blocks(k) gives kernel k's co-resident workgroup count.  Returns false
// when no resident shape holds the cloud (beyond the register file).
template <class MaxBlocks>
bool plan_resident_only(size_t n, size_t dim, int tie, MaxBlocks max_blocks, Plan &p) {
  size_t cnt;
  const KernelEntry *t = resident_table(&cnt);
  auto take = [&](const KernelEntry *k, uint32_t g) {
    p.kind = PlanKind::Resident;
    p.k = k;
    p.block = (uint32_t)k->block;
    p.grid = g;
    p.chunk = (uint32_t)(k->block * k->ppt);
    p.shmem = 0;
  };
  static const int kSmall[][2] = {{256, 1}, {256, 2}, {256, 4}, {1024, 1}, {1024, 2},
                                  {1024, 4}, {1024, 8}, {1024, 16}};
  if (n <= 16384) {
    for (auto &sh : kSmall) {
      if ((size_t)sh[0] * sh[1] < n) continue;
      for (size_t i = 0; i < cnt; ++i) {
        const KernelEntry *k = &t[i];
        if (k->dim != (int)dim || k->tie != tie || k->block != sh[0] || k->ppt != sh[1]) continue;
        const uint32_t g = resident_grid(n, k, (size_t)max_blocks(k));
        if (g) {
          take(k, g);
          return true;
        }
      }
    }
  }
  double best = 1e30;
  const KernelEntry *bk = nullptr;
  uint32_t bg = 0;
  for (size_t i = 0; i < cnt; ++i) {
    const KernelEntry *k = &t[i];
    if (k->dim != (int)dim || k->tie != tie) continue;
    const uint32_t g = resident_grid(n, k, (size_t)max_blocks(k));
    if (!g) continue;
    const double us = resident_model_us(n, g) + (k->block == 1024 ? 0.0 : 0.05);
    if (us < best) {
      best = us;
      bk = k;
      bg = g;
    }
  }
  if (!bk) return false;
  take(bk, bg);
  return true;
}

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
// Frees the scratch's device buffers (owners with a shorter life than the process).
void delete_spatial_scratch(struct SpatialScratch *s);
int fps_host(const float *points, size_t n, size_t dim, int layout, size_t m,
             const uint64_t *starts, size_t n_starts, uint64_t *out, int tie, float init);

}  // namespace fpsx

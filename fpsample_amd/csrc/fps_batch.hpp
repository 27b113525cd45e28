// fps_batch.hpp -- host-visible description of the batched FPS kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fps_kernels.hpp"

namespace fpsx {

constexpr uint32_t kMaxTeam = 64;                     // workgroups per cloud at most
constexpr size_t kTeamXchgWords = 2 * 16 * kMaxTeam;  // 2 buffers x kSlotWords x kMaxTeam

// One cloud of a batch (and the passes this launch makes on it).
struct BatchJob {
  uint64_t pt_off;     // the cloud's SoA block starts at pts + pt_off
  uint64_t ld;         // coordinate d of point i at pts[pt_off + d * ld + i]
  uint64_t out_off;    // first pick of the cloud in out[]
  uint64_t state_off;  // first saved min-distance of the cloud in state[] (chunked runs)
  uint64_t start;      // pick 0
  uint32_t n;          // points
  uint32_t m;          // samples
  uint32_t s_begin;    // this launch computes picks s_begin .. s_end-1 (0 = fresh start)
  uint32_t s_end;
  uint32_t team;       // workgroups working on the cloud (1..kMaxTeam)
  uint32_t member_base;// claim slot of the team's first member
  uint32_t share;      // points per member: ceil(n / team) <= B * PPT (B * PPT for the pruned kernel)
  uint32_t area;       // exchange area of the team (teams only)
  uint64_t sorted_off; // pruned kernel: the cloud in spatial order, SoA, at sorted + sorted_off (ld = n)
  uint64_t perm_off;   // pruned kernel: original index of every sorted position, perm + perm_off
};

struct BatchParams {
  const float *pts;          // base of every job's SoA block (BatchJob::pt_off, ::ld)
  const BatchJob *jobs;
  const uint32_t *member_job;  // claim slot -> job
  uint32_t total_members;
  uint32_t *claim;           // claim counter, zeroed before the launch
  uint32_t *ready;           // per-job team-formation counters, zeroed
  uint64_t *out;             // concatenated picks
  float *state;              // saved min-distances (chunked runs; may be null otherwise)
  uint64_t *xchg;            // team exchange areas (kTeamXchgWords each), zeroed
  uint32_t *status;          // status word, zeroed
  float init;                // initial min-distance
  uint32_t pair;             // 1: teams settle up to two picks per exchange
  uint64_t spin_ticks;       // bound on one exchange wait
  uint64_t team_ticks;       // bound on waiting for a team to form
};

const KernelEntry *batch_table(size_t *count);

// Batched FPS with the spatial order, per-wave box pruning and paired picks
// (fps_batch_pp): every job is a team of ceil(n / (B * PPT)) members, member r
// owning sorted positions [r * B * PPT, (r + 1) * B * PPT).  Chunked jobs save
// their min-distances by sorted position in state[state_off .. + n) and the
// last pick's coordinates in state[state_off + n .. + 3).
constexpr uint32_t kPPReplicas = 4;  // exchange replicas of a team
// Words between a team area's word arrays.  64 (one word per member); a
// 4 KiB page per array as the single-cloud kernels use was slower here
// (largest serving request alone, 256-pass chunks: 1033 against 995 ms).
constexpr uint32_t kPPPitch = kMaxTeam;
constexpr size_t kPPTeamWords = 2 * kPPReplicas * 16 * kPPPitch;  // 2 buffers x R x kSlotWords x pitch

struct BatchPPParams {
  BatchParams b;         // jobs, claims, picks, state, status, exchange areas (kPPTeamWords each)
  const float *sorted;   // spatially ordered clouds (BatchJob::sorted_off)
  const uint32_t *perm;  // their original indices (BatchJob::perm_off)
};

const KernelEntry *batch_pp_table(size_t *count);
// The pruned batch kernel of a shape: one_slot (one workgroup per member slot,
// all co-resident) or the claim-loop form; null when not instantiated.
const KernelEntry *batch_pp_kernel(int block, int ppt, bool one_slot);

}  // namespace fpsx

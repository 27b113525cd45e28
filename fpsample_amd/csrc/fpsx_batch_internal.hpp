// fpsx_batch_internal.hpp -- job tables shared by the batch object and the
// serving loop.
#pragma once
#include <stdint.h>

#include <vector>

#include "fps_batch.hpp"
#include "fpsx_internal.hpp"

namespace fpsx {

// One cloud's work for a launch, before team sizes are chosen.
struct JobSpec {
  uint64_t pt_off, ld, out_off, state_off, start;
  uint32_t n, m, s_begin, s_end;
  uint64_t sorted_off = 0, perm_off = 0;  // pruned kernel: the spatial order
};

// Team sizes (ceil(n / min(B * PPT, share_cap))), claim order (teams first,
// longest first) and exchange-area numbering for one launch.  aligned: every
// member's share is exactly B * PPT (the last one takes the rest), as the
// pruned kernel's wave blocks need.
int build_jobs(const JobSpec *specs, size_t n_specs, const KernelEntry *k, uint32_t share_cap,
               std::vector<BatchJob> &jobs, std::vector<uint32_t> &member_job, uint32_t *n_areas,
               bool aligned = false);

// Device tables and scratch of one batch launch.
struct BatchRun {
  const KernelEntry *k = nullptr;
  std::vector<BatchJob> jobs;
  std::vector<uint32_t> member_job;
  uint32_t members = 0;
  uint32_t grid = 0;
  size_t counter_words = 0, xchg_words = 0;
  DevBuf d_jobs, d_member_job, d_counters, d_xchg, d_status;
  BatchParams params;
  // Pruned kernel (fps_batch_pp): its parameters wrap params; one_slot: one
  // workgroup per member slot (grid = members, which must all be co-resident).
  bool pruned = false, one_slot = false;
  BatchPPParams pp_params;
  // Upload tables for this job list (host-synchronous) and fill params.
  // For a pruned kernel, sorted / perm are the bases of BatchJob::sorted_off
  // and ::perm_off.
  int prepare(Ctx &c, const KernelEntry *kernel, const std::vector<BatchJob> &js,
              const std::vector<uint32_t> &mj, uint32_t n_areas, const float *d_pts,
              uint64_t *d_out, float *d_state, const float *sorted = nullptr,
              const uint32_t *perm = nullptr, bool one_slot = false);
  // Clear counters / exchange areas / status and launch (graph-capturable).
  int enqueue(hipStream_t st);
  ~BatchRun() {
    d_jobs.free();
    d_member_job.free();
    d_counters.free();
    d_xchg.free();
    d_status.free();
  }
};

}  // namespace fpsx

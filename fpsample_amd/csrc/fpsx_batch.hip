// fpsx_batch.hip -- host side of batched FPS: job tables, the batch object
// (device tables + a captured HIP graph of its launch sequence) and its C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "fps_batch.hpp"
#include "fps_device.hpp"
#include "fpsx.h"
#include "fpsx_batch_internal.hpp"
#include "fpsx_internal.hpp"
#include "fps_pruned.hpp"

#define FPSX_EXPORT extern "C" __attribute__((visibility("default")))

namespace fpsx {

static int hip_err(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? FPSX_ENOMEM : FPSX_EHIP,
              std::string(what) + ": " + hipGetErrorString(e));
}
#define HIP_TRY(expr)                                 \
  do {                                                \
    hipError_t _e = (expr);                           \
    if (_e != hipSuccess) return hip_err(_e, #expr);  \
  } while (0)

static const KernelEntry *find_batch(int dim, int B, int ppt) {
  size_t cnt;
  const KernelEntry *t = batch_table(&cnt);
  for (size_t i = 0; i < cnt; ++i)
    if (t[i].dim == dim && t[i].block == B && t[i].ppt == ppt) return &t[i];
  return nullptr;
}

int build_jobs(const JobSpec *specs, size_t n_specs, const KernelEntry *k, uint32_t share_cap,
               std::vector<BatchJob> &jobs, std::vector<uint32_t> &member_job,
               uint32_t *n_areas, bool aligned) {
  const uint32_t cap = std::min<uint32_t>((uint32_t)(k->block * k->ppt), share_cap ? share_cap : ~0u);
  std::vector<BatchJob> js(n_specs);
  for (size_t c = 0; c < n_specs; ++c) {
    const JobSpec &sp = specs[c];
    BatchJob &J = js[c];
    std::memset(&J, 0, sizeof(J));
    J.pt_off = sp.pt_off;
    J.ld = sp.ld;
    J.out_off = sp.out_off;
    J.state_off = sp.state_off;
    J.start = sp.start;
    J.n = sp.n;
    J.m = sp.m;
    J.s_begin = sp.s_begin;
    J.s_end = sp.s_end;
    J.sorted_off = sp.sorted_off;
    J.perm_off = sp.perm_off;
    uint32_t team = (sp.n + cap - 1) / cap;
    if (team < 1) team = 1;
    if (team > kMaxTeam || team > (uint32_t)k->block)
      return fail(FPSX_ETOOBIG, "cloud of " + std::to_string(sp.n) +
                                    " points exceeds the batch kernel's team capacity");
    J.share = aligned ? cap : (sp.n + team - 1) / team;
    J.team = (sp.n + J.share - 1) / J.share;  // no empty member
  }
  // Teams first (they form at once while the grid is free), then single
  // workgroup jobs; longest first within each group.
  std::vector<size_t> order(n_specs);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    const bool ta = js[a].team > 1, tb = js[b].team > 1;
    if (ta != tb) return ta;
    const double wa = (double)js[a].share * (double)(js[a].s_end - js[a].s_begin);
    const double wb = (double)js[b].share * (double)(js[b].s_end - js[b].s_begin);
    return wa > wb;
  });
  jobs.clear();
  member_job.clear();
  uint32_t areas = 0;
  for (size_t o = 0; o < n_specs; ++o) {
    BatchJob J = js[order[o]];
    J.member_base = (uint32_t)member_job.size();
    J.area = (J.team > 1) ? areas++ : 0;
    for (uint32_t r = 0; r < J.team; ++r) member_job.push_back((uint32_t)jobs.size());
    jobs.push_back(J);
  }
  *n_areas = areas;
  return FPSX_OK;
}

int BatchRun::prepare(Ctx &c, const KernelEntry *kernel, const std::vector<BatchJob> &js,
                      const std::vector<uint32_t> &mj, uint32_t n_areas, const float *d_pts,
                      uint64_t *d_out, float *d_state, const float *sorted,
                      const uint32_t *perm, bool one) {
  k = kernel;
  jobs = js;
  member_job = mj;
  members = (uint32_t)mj.size();
  pruned = sorted != nullptr;
  one_slot = one;
  int gmax = c.max_blocks(k, 0);
  if (gmax <= 0) return fail(FPSX_EHIP, "batch kernel cannot be resident");
  if (one_slot && members > (uint32_t)gmax)
    return fail(FPSX_ETOOBIG, "one-slot batch launch: more members than co-resident workgroups");
  grid = std::min<uint32_t>(members, (uint32_t)gmax);
  uint32_t max_team = 1;
  for (auto &J : jobs) max_team = std::max(max_team, J.team);
  if (max_team > grid) return fail(FPSX_ETOOBIG, "a team is larger than the co-resident grid");
  HIP_TRY(d_jobs.reserve(jobs.size() * sizeof(BatchJob)));
  HIP_TRY(d_member_job.reserve(members * sizeof(uint32_t)));
  counter_words = (1 + jobs.size() + 1) / 2 + 1;  // claim + per-job ready counters, in u64 words
  HIP_TRY(d_counters.reserve(counter_words * sizeof(uint64_t)));
  xchg_words = std::max<size_t>(1, (size_t)n_areas * (pruned ? kPPTeamWords : kTeamXchgWords));
  HIP_TRY(d_xchg.reserve(xchg_words * sizeof(uint64_t)));
  HIP_TRY(d_status.reserve(sizeof(uint64_t)));
  HIP_TRY(hipMemcpy(d_jobs.p, jobs.data(), jobs.size() * sizeof(BatchJob),
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_member_job.p, member_job.data(), members * sizeof(uint32_t),
                    hipMemcpyHostToDevice));
  std::memset(&params, 0, sizeof(params));
  params.pts = d_pts;
  params.jobs = (const BatchJob *)d_jobs.p;
  params.member_job = (const uint32_t *)d_member_job.p;
  params.total_members = members;
  params.claim = (uint32_t *)d_counters.p;
  params.ready = (uint32_t *)d_counters.p + 2;
  params.out = d_out;
  params.state = d_state;
  params.xchg = (uint64_t *)d_xchg.p;
  params.status = (uint32_t *)d_status.p;
  params.init = INFINITY;
  params.spin_ticks = c.spin_ticks;
  params.team_ticks = c.spin_ticks * 5;
  // Teams settle up to two picks per exchange unless FPSX_BATCH_NO_PAIR=1
  // (the one-pick team loop; same picks).
  const char *np = std::getenv("FPSX_BATCH_NO_PAIR");
  params.pair = (np && np[0] == '1') ? 0u : 1u;
  pp_params.b = params;
  pp_params.sorted = sorted;
  pp_params.perm = perm;
  return FPSX_OK;
}

int BatchRun::enqueue(hipStream_t st) {
  HIP_TRY(launch_fill_u64((uint64_t *)d_counters.p, counter_words, 0, st));
  HIP_TRY(launch_fill_u64((uint64_t *)d_xchg.p, xchg_words, 0, st));
  HIP_TRY(launch_fill_u64((uint64_t *)d_status.p, 1, 0, st));
  void *args[] = {pruned ? (void *)&pp_params : (void *)&params};
  HIP_TRY(hipLaunchKernel(k->fn, dim3(grid), dim3(k->block), args, 0, st));
  return FPSX_OK;
}

}  // namespace fpsx

using namespace fpsx;

// ---------------------------------------------------------------------------
// Batch object
// ---------------------------------------------------------------------------
struct fpsx_batch {
  Ctx *ctx = nullptr;
  int dim = 0;
  size_t n_clouds = 0;
  BatchRun run;
  // Pruned kernel (fps_batch_pp): every cloud in spatial order, cloud c's
  // coordinates at sorted + 3 * offsets[c] (SoA, ld = its size) and original
  // indices at perm + offsets[c]; blocks of `order_block` positions.
  DevBuf sorted, perm;
  SpatialScratch *spatial = nullptr;
  uint32_t order_block = 0;
  std::vector<uint64_t> offsets;
  bool fallback = false;  // no batch kernel for this dimension or cloud size: one launch per cloud
  std::vector<JobSpec> specs;
  const float *d_pts = nullptr;
  uint64_t ld = 0;
  uint64_t *d_out = nullptr;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int validate_batch(const uint64_t *offsets, size_t n_clouds, const uint64_t *n_samples,
                          const uint64_t *start_idx, uint64_t ld, std::vector<JobSpec> &specs) {
  if (!offsets || !n_samples || !start_idx || n_clouds == 0)
    return fail(FPSX_EINVAL, "empty batch or null pointer");
  specs.resize(n_clouds);
  uint64_t out_off = 0;
  for (size_t c = 0; c < n_clouds; ++c) {
    if (offsets[c + 1] < offsets[c]) return fail(FPSX_EINVAL, "offsets must not decrease");
    const uint64_t n = offsets[c + 1] - offsets[c];
    if (n == 0) return fail(FPSX_EINVAL, "cloud " + std::to_string(c) + " is empty");
    if (n >= (uint64_t(1) << 31)) return fail(FPSX_ETOOBIG, "cloud too large");
    if (n_samples[c] == 0 || n_samples[c] > n)
      return fail(FPSX_EINVAL, "cloud " + std::to_string(c) + ": n_samples must be in [1, n]");
    if (start_idx[c] >= n)
      return fail(FPSX_EBADSTART, "cloud " + std::to_string(c) + ": start index out of range");
    JobSpec &s = specs[c];
    s.pt_off = offsets[c];
    s.ld = ld;
    s.out_off = out_off;
    s.state_off = 0;
    s.start = start_idx[c];
    s.n = (uint32_t)n;
    s.m = (uint32_t)n_samples[c];
    s.s_begin = 0;
    s.s_end = s.m;
    out_off += n_samples[c];
  }
  return FPSX_OK;
}

// Put every cloud in spatial order with blocks of `block` positions (once
// per block size), on the context's stream.  Caller holds the context lock.
static int batch_order(fpsx_batch *b, uint32_t block) {
  if (b->order_block == block) return FPSX_OK;
  const uint64_t total = b->offsets.back();
  uint32_t nmax = 0;
  for (auto &s : b->specs) nmax = std::max(nmax, s.n);
  HIP_TRY(b->sorted.reserve(total * 3 * sizeof(float)));
  HIP_TRY(b->perm.reserve(total * sizeof(uint32_t)));
  if (!b->spatial) b->spatial = new_spatial_scratch();
  int rc = spatial_reserve(*b->spatial, nmax, block);
  if (rc) return rc;
  hipStream_t st = b->ctx->stream;
  for (size_t c = 0; c < b->n_clouds; ++c) {
    const uint64_t off = b->offsets[c];
    rc = spatial_order(b->d_pts + off, 1, b->ld, b->specs[c].n, block,
                       (float *)b->sorted.p + 3 * off, (uint32_t *)b->perm.p + off, *b->spatial, st);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(st));
  b->order_block = block;
  return FPSX_OK;
}

static int batch_prepare_kernel(fpsx_batch *b, const KernelEntry *k, bool pruned) {
  std::vector<BatchJob> jobs;
  std::vector<uint32_t> mj;
  uint32_t areas = 0;
  if (pruned) {
    for (size_t c = 0; c < b->n_clouds; ++c) {
      b->specs[c].sorted_off = 3 * b->offsets[c];
      b->specs[c].perm_off = b->offsets[c];
    }
  }
  int rc = build_jobs(b->specs.data(), b->specs.size(), k, 0, jobs, mj, &areas, pruned);
  if (rc) return rc;
  if (pruned) {
    rc = batch_order(b, (uint32_t)(64 * k->ppt));
    if (rc) return rc;
    return b->run.prepare(*b->ctx, k, jobs, mj, areas, b->d_pts, b->d_out, nullptr,
                          (const float *)b->sorted.p, (const uint32_t *)b->perm.p, false);
  }
  return b->run.prepare(*b->ctx, k, jobs, mj, areas, b->d_pts, b->d_out, nullptr);
}

FPSX_EXPORT int fpsx_batch_create(const float *d_points, size_t ld, size_t dim,
                                  const uint64_t *offsets, size_t n_clouds,
                                  const uint64_t *n_samples, const uint64_t *start_idx,
                                  uint64_t *d_out, int flags, fpsx_batch **out) {
  if (!out) return fail(FPSX_EINVAL, "null output handle");
  *out = nullptr;
  if (dim == 0) return fail(FPSX_EBADDIM, "dim must be >= 1");
  if (!d_points || !d_out) return fail(FPSX_EINVAL, "null pointer argument");
  std::unique_ptr<fpsx_batch> b(new fpsx_batch());
  int rc = validate_batch(offsets, n_clouds, n_samples, start_idx, ld, b->specs);
  if (rc) return rc;
  if (ld < offsets[n_clouds]) return fail(FPSX_EINVAL, "ld must be >= the total point count");
  rc = get_ctx(&b->ctx);
  if (rc) return rc;
  b->dim = (int)dim;
  b->n_clouds = n_clouds;
  b->d_pts = d_points;
  b->ld = ld;
  b->d_out = d_out;
  b->offsets.assign(offsets, offsets + n_clouds + 1);
  Ctx &c = *b->ctx;
  std::lock_guard<std::mutex> lk(c.mu);

  // Candidates: 3-D batches first the pruned paired kernels (spatial order,
  // wave boxes, two picks per team exchange; FPSX_BATCH_PLAIN=1 leaves them
  // out), then the plain batch kernels of the dimension.
  struct Cand {
    const KernelEntry *k;
    bool pruned;
  };
  std::vector<Cand> cands;
  const char *plain = std::getenv("FPSX_BATCH_PLAIN");
  if (dim == 3 && !(plain && plain[0] == '1')) {
    for (int ppt : {8, 16, 4}) {
      const KernelEntry *k = batch_pp_kernel(1024, ppt, false);
      if (k) cands.push_back({k, true});
    }
  }
  size_t cnt;
  const KernelEntry *t = batch_table(&cnt);
  for (size_t i = 0; i < cnt; ++i)
    if (t[i].dim == (int)dim) cands.push_back({&t[i], false});
  if (cands.empty()) {
    b->fallback = true;
    *out = b.release();
    return FPSX_OK;
  }
  const KernelEntry *chosen = nullptr;
  bool chosen_pruned = false;
  const char *env = std::getenv("FPSX_BATCH_SHAPE");
  int eb = 0, ep = 0;
  if (env && std::sscanf(env, "pp:%d:%d", &eb, &ep) == 2) {
    chosen = batch_pp_kernel(eb, ep, false);
    chosen_pruned = chosen != nullptr && dim == 3;
    if (!chosen_pruned) chosen = nullptr;
  } else if (env && std::sscanf(env, "%d:%d", &eb, &ep) == 2) {
    chosen = find_batch((int)dim, eb, ep);
  }
  if (!chosen && (flags & FPSX_BATCH_AUTOTUNE)) {
    // Time one launch of every shape that can hold the batch; keep the fastest.
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    for (const Cand &cd : cands) {
      const KernelEntry *k = cd.k;
      if (batch_prepare_kernel(b.get(), k, cd.pruned) != FPSX_OK) continue;
      HIP_TRY(hipEventRecord(e0, c.stream));
      if (b->run.enqueue(c.stream) != FPSX_OK) continue;
      HIP_TRY(hipEventRecord(e1, c.stream));
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      HIP_TRY(hipMemcpy(c.host_status, b->run.d_status.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
      if (*c.host_status == 0 && ms < best) {
        best = ms;
        chosen = k;
        chosen_pruned = cd.pruned;
      }
    }
    (void)hipEventDestroy(e0);  // timing events only
    (void)hipEventDestroy(e1);
  }
  if (!chosen) {
    // Heuristic: the first pruned kernel whose teams hold every cloud; else
    // the smallest plain per-workgroup capacity that still fits every cloud
    // in a team the grid can hold (short reduction chains for small clouds),
    // falling back to larger shapes.
    uint32_t nmax = 0;
    for (auto &s : b->specs) nmax = std::max(nmax, s.n);
    for (const Cand &cd : cands) {
      if (!cd.pruned) continue;
      const uint32_t cap = (uint32_t)(cd.k->block * cd.k->ppt);
      if ((nmax + cap - 1) / cap <= kMaxTeam) {
        chosen = cd.k;
        chosen_pruned = true;
        break;
      }
    }
    static const int kOrder[][2] = {{256, 16}, {1024, 16}, {256, 32}, {1024, 8}, {256, 8}, {256, 4}};
    for (auto &sh : kOrder) {
      if (chosen) break;
      const KernelEntry *k = find_batch((int)dim, sh[0], sh[1]);
      if (!k) continue;
      if ((nmax + (uint32_t)(sh[0] * sh[1]) - 1) / (uint32_t)(sh[0] * sh[1]) > 8 && sh[0] * sh[1] < 16384)
        continue;
      chosen = k;
    }
    if (!chosen) {
      chosen = cands.back().k;
      chosen_pruned = cands.back().pruned;
    }
  }
  rc = batch_prepare_kernel(b.get(), chosen, chosen_pruned);
  if (rc == FPSX_ETOOBIG) {
    // Some cloud needs more workgroups than a team may have: every cloud
    // gets a planned launch of its own (the single-cloud kernels have no
    // such limit).
    b->fallback = true;
    *out = b.release();
    return FPSX_OK;
  }
  if (rc) return rc;
  if (flags & FPSX_BATCH_GRAPH) {
    HIP_TRY(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
    rc = b->run.enqueue(b->capture_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return hip_err(e, "hipStreamEndCapture");
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  }
  *out = b.release();
  return FPSX_OK;
}

FPSX_EXPORT int fpsx_batch_run(fpsx_batch *b, void *stream) {
  if (!b) return fail(FPSX_EINVAL, "null batch");
  Ctx &c = *b->ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  hipStream_t st = stream ? (hipStream_t)stream : c.stream;
  if (b->fallback) {
    // No batch kernel for this dimension or cloud size: one planned launch
    // per cloud.  Every launch clears the context's status word, so each
    // one's status is folded into the batch's own word, which the batch
    // status then reports (a failure in any cloud, not only the last).
    HIP_TRY(b->run.d_status.reserve(sizeof(uint64_t)));
    int rc = scratch_acquire(c, st);
    if (rc) return rc;
    HIP_TRY(launch_fill_u64((uint64_t *)b->run.d_status.p, 1, 0, st));
    HIP_TRY(c.starts.reserve(b->specs.size() * sizeof(uint64_t)));
    std::vector<uint64_t> starts(b->specs.size());
    for (size_t i = 0; i < starts.size(); ++i) starts[i] = b->specs[i].start;
    HIP_TRY(hipMemcpyAsync(c.starts.p, starts.data(), starts.size() * sizeof(uint64_t),
                           hipMemcpyHostToDevice, st));
    for (size_t i = 0; i < b->specs.size(); ++i) {
      const JobSpec &s = b->specs[i];
      Plan p;
      rc = plan(c, s.n, b->dim, kTieHighest, s.m, p);
      if (rc) return rc;
      rc = launch_fps(c, p, b->d_pts + s.pt_off, 1, b->ld, s.n, b->dim, s.m,
                      (const uint64_t *)c.starts.p + i, 1, b->d_out + s.out_off, nullptr,
                      INFINITY, st);
      if (rc) return rc;
      HIP_TRY(launch_or_u32((const uint32_t *)c.status.p, (uint32_t *)b->run.d_status.p, st));
    }
    return scratch_release(c, st);
  }
  if (b->exec) {
    HIP_TRY(hipGraphLaunch(b->exec, st));
    return FPSX_OK;
  }
  return b->run.enqueue(st);
}

FPSX_EXPORT int fpsx_batch_status(fpsx_batch *b, void *stream) {
  if (!b) return fail(FPSX_EINVAL, "null batch");
  Ctx &c = *b->ctx;
  std::lock_guard<std::mutex> lk(c.mu);
  hipStream_t st = stream ? (hipStream_t)stream : c.stream;
  HIP_TRY(hipStreamSynchronize(st));
  if (!b->run.d_status.p) return FPSX_OK;  // never run
  HIP_TRY(hipMemcpy(c.host_status, b->run.d_status.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
  return status_code(*c.host_status);
}

FPSX_EXPORT int fpsx_batch_info(fpsx_batch *b, int *block, int *ppt, int *grid, int *members,
                                int *graph) {
  if (!b) return fail(FPSX_EINVAL, "null batch");
  if (block) *block = b->fallback ? 0 : b->run.k->block;
  if (ppt) *ppt = b->fallback ? 0 : b->run.k->ppt;
  if (grid) *grid = b->fallback ? 0 : (int)b->run.grid;
  if (members) *members = b->fallback ? 0 : (int)b->run.members;
  if (graph) *graph = b->exec != nullptr;
  return FPSX_OK;
}

FPSX_EXPORT void fpsx_batch_destroy(fpsx_batch *b) {
  if (!b) return;
  // Release path: a failed destroy leaves nothing for the caller to do.
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  b->sorted.free();
  b->perm.free();
  delete_spatial_scratch(b->spatial);
  delete b;
}

FPSX_EXPORT int fpsx_fps_batch(const float *points, size_t dim, const uint64_t *offsets,
                               size_t n_clouds, const uint64_t *n_samples,
                               const uint64_t *start_idx, uint64_t *out) {
  if (dim == 0) return fail(FPSX_EBADDIM, "dim must be >= 1");
  if (!points || !out) return fail(FPSX_EINVAL, "null pointer argument");
  std::vector<JobSpec> specs;
  int rc = validate_batch(offsets, n_clouds, n_samples, start_idx, offsets[n_clouds], specs);
  if (rc) return rc;
  const size_t total = offsets[n_clouds];
  size_t total_m = 0;
  for (size_t c = 0; c < n_clouds; ++c) total_m += n_samples[c];
  Ctx *cp = nullptr;
  rc = get_ctx(&cp);
  if (rc) return rc;
  DevBuf rows, soa, dout;
  HIP_TRY(rows.reserve(total * dim * sizeof(float)));
  HIP_TRY(soa.reserve(total * dim * sizeof(float)));
  HIP_TRY(dout.reserve(total_m * sizeof(uint64_t)));
  {
    std::lock_guard<std::mutex> lk(cp->mu);
    HIP_TRY(hipMemcpyAsync(rows.p, points, total * dim * sizeof(float), hipMemcpyHostToDevice,
                           cp->stream));
    HIP_TRY(launch_strided_to_soa((const float *)rows.p, dim, 1, (float *)soa.p, (uint32_t)total,
                                  (uint32_t)dim, total, cp->stream));
    HIP_TRY(hipStreamSynchronize(cp->stream));
  }
  fpsx_batch *b = nullptr;
  rc = fpsx_batch_create((const float *)soa.p, total, dim, offsets, n_clouds, n_samples, start_idx,
                         (uint64_t *)dout.p, 0, &b);
  if (rc == FPSX_OK) rc = fpsx_batch_run(b, nullptr);
  if (rc == FPSX_OK) rc = fpsx_batch_status(b, nullptr);
  if (rc == FPSX_OK) {
    hipError_t e = hipMemcpy(out, dout.p, total_m * sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_err(e, "hipMemcpy");
  }
  fpsx_batch_destroy(b);
  rows.free();
  soa.free();
  dout.free();
  return rc;
}

// fpsx_server.cpp -- FPS serving loop with HBM-budget admission and
// continuous batching (host runtime, HIP API only; the kernel is fps_batch).
//
// Requests (one cloud each, any size) are submitted from any thread.  A
// scheduler thread owns the device:
//   admit    -- pending requests enter, oldest first, while their device
//               footprint fits the HBM budget (a device arena allocated once;
//               first-fit ranges, no per-request hipMalloc) and their team of
//               workgroups fits the co-resident grid; their points are copied
//               on a separate stream, overlapping the running kernel;
//   launch   -- one batch-kernel launch advances every admitted, uploaded
//               request by up to `chunk` passes (min-distances saved in the
//               arena between launches), so requests that arrive while others
//               run join at a following launch.  Launches are pipelined: the
//               next one is prepared and queued on the compute stream while the
//               current one runs (two launch slots), so the GPU does not wait
//               for the host between chunks;
//   retire   -- once a launch completes, the requests it finished copy their
//               picks back, free their range and wake their waiters.
// Request control: pending requests are admitted by priority (FIFO among
// equals); each launch takes the admitted requests in priority order while
// their workgroups fit the grid, so a lower-priority request is paused
// between chunks (its state stays in its range) when higher ones need the
// room.  When a request's HBM is missing, lower-priority requests are evicted
// at their next chunk boundary: the whole range (points, spatial order,
// min-distances, picks so far) is copied to host memory and freed, and the
// request waits in the queue to be copied back and resumed where it stopped.
// A cancelled request, or one past its deadline, leaves the queue at once or,
// when running, stops at its next chunk boundary: its range is freed once no
// queued launch still writes it.
// Each request's picks equal fps_sampling on its cloud alone: a job's result
// never depends on what else shares the launch (exact min and argmax, ties to
// the highest index), and a chunk boundary only saves and restores state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fps_batch.hpp"
#include "fpsx.h"
#include "fpsx_batch_internal.hpp"
#include "fpsx_internal.hpp"
#include "fps_pruned.hpp"

#define FPSX_EXPORT extern "C" __attribute__((visibility("default")))

namespace fpsx {

namespace {

using Clock = std::chrono::steady_clock;

double secs_since(Clock::time_point t0, Clock::time_point t) {
  return std::chrono::duration<double>(t - t0).count();
}

// First-fit allocator over [0, size) in 256-byte units.
class Arena {
 public:
  explicit Arena(size_t bytes) { free_[0] = bytes / kAlign; units_ = bytes / kAlign; }
  bool alloc(size_t bytes, size_t *off) {
    const size_t need = (bytes + kAlign - 1) / kAlign;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      *off = it->first * kAlign;
      const size_t start = it->first + need, left = it->second - need;
      free_.erase(it);
      if (left) free_[start] = left;
      used_ += need;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    const size_t u = off / kAlign, n = (bytes + kAlign - 1) / kAlign;
    used_ -= n;
    auto it = free_.emplace(u, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used_bytes() const { return used_ * kAlign; }
  size_t free_bytes() const { return (units_ - used_) * kAlign; }
  static constexpr size_t kAlign = 256;

 private:
  std::map<size_t, size_t> free_;  // start unit -> length in units
  size_t units_ = 0, used_ = 0;
};

size_t align_up(size_t v) { return (v + Arena::kAlign - 1) / Arena::kAlign * Arena::kAlign; }

struct Request {
  uint64_t id = 0;
  std::vector<float> soa;  // host copy of the cloud, SoA (dim x n)
  uint32_t n = 0, m = 0, dim = 0;
  uint64_t start = 0;
  // device placement (byte offsets into the arena)
  size_t off = 0, bytes = 0, pts_off = 0, state_off = 0, out_off = 0;
  size_t sorted_off = 0, perm_off = 0;  // pruned kernel: the spatial order
  uint32_t team = 1;
  uint32_t next = 0;  // next pick to compute (0 = not started)
  hipEvent_t uploaded = nullptr;
  bool ready = false, done = false;
  bool finishing = false;  // its last chunk is queued: in no further launch
  uint32_t inflight = 0;   // queued launches that include it
  int priority = 0;
  bool has_deadline = false;
  Clock::time_point deadline;
  bool stop = false;        // cancelled or past its deadline: no further chunk
  int stop_code = FPSX_OK;  // FPSX_ECANCELED or FPSX_EDEADLINE
  bool evict = false;       // leaves the device at its next chunk boundary
  std::vector<char> saved;  // evicted: the device range, restored on re-admission
  int error = FPSX_OK;
  std::vector<uint64_t> picks;
  Clock::time_point t_submit, t_admit, t_done;
};

}  // namespace

}  // namespace fpsx

using namespace fpsx;

struct fpsx_server {
  Ctx *ctx = nullptr;
  const KernelEntry *k = nullptr;
  // The pruned paired kernel (fps_batch_pp, one slot per member): every
  // request is put in spatial order on the copy stream after its upload.
  // FPSX_SERVE_PLAIN=1 at creation keeps the plain batch kernel.
  bool pruned = true;
  SpatialScratch *spatial = nullptr;
  uint32_t chunk = 256;
  size_t budget = 0;
  void *arena_base = nullptr;
  std::unique_ptr<Arena> arena;
  uint32_t grid_cap = 0;       // co-resident workgroups of the batch kernel
  hipStream_t compute = nullptr, copy = nullptr;
  // Two launch slots: the launch running on the compute stream and the next
  // one, prepared and queued behind it.
  struct Launch {
    BatchRun run;
    hipEvent_t t0 = nullptr, t1 = nullptr;  // around the kernel, for busy_s
    std::vector<std::shared_ptr<Request>> reqs;
    std::vector<uint32_t> s_end;
    bool busy = false;
  };
  Launch slot[2];
  int cur = 0;  // slot of the oldest queued launch

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::vector<std::shared_ptr<Request>> active;
  std::map<uint64_t, std::shared_ptr<Request>> all;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread thread;
  Clock::time_point t0 = Clock::now();
  // counters
  uint64_t launches = 0, completed = 0, samples_done = 0;
  uint64_t n_cancelled = 0, n_expired = 0, n_paused = 0, n_evicted = 0;
  double busy_s = 0;
  int fatal = FPSX_OK;
  std::string fatal_msg;

  uint32_t team_of(uint32_t n) const {
    const uint32_t cap = (uint32_t)(k->block * k->ppt);
    const uint32_t t = (n + cap - 1) / cap;
    return std::max(1u, t);
  }
  // Points, min-distances (+ the last pick's coordinates), picks; for the
  // pruned kernel also the spatially ordered points and their original indices.
  size_t footprint(const Request &r) const {
    size_t b = align_up((size_t)r.n * r.dim * sizeof(float)) +
               align_up(((size_t)r.n + 4) * sizeof(float)) + align_up((size_t)r.m * sizeof(uint64_t));
    if (pruned) b += align_up((size_t)r.n * r.dim * sizeof(float)) + align_up((size_t)r.n * sizeof(uint32_t));
    return b;
  }
  // Ends a request that will not finish (caller holds mu; wakes no one).  Its
  // host copy is no longer read by any upload.
  void end_request(Request &r, int code, Clock::time_point now) {
    r.soa.clear();
    r.soa.shrink_to_fit();
    r.saved.clear();
    r.saved.shrink_to_fit();
    r.done = true;
    r.error = code;
    r.t_done = now;
    if (code == FPSX_ECANCELED) ++n_cancelled;
    if (code == FPSX_EDEADLINE) ++n_expired;
  }
  // r's HBM is missing: mark lower-priority requests for eviction (lowest
  // priority first, then the most recently admitted) until the free bytes
  // plus theirs would hold it.  Caller holds mu.
  void choose_victims(const Request &r) {
    size_t coming = 0;
    for (auto &a : active)
      if (a->evict) coming += a->bytes;
    if (coming) return;  // evictions under way: retry once they are done
    std::vector<std::shared_ptr<Request>> cand;
    for (auto &a : active)
      if (a->priority < r.priority && !a->finishing && !a->stop && a->ready) cand.push_back(a);
    std::sort(cand.begin(), cand.end(),
              [](const std::shared_ptr<Request> &x, const std::shared_ptr<Request> &y) {
                if (x->priority != y->priority) return x->priority < y->priority;
                return x->t_admit > y->t_admit;
              });
    // (at least one victim: with enough free bytes, the range is fragmented)
    for (auto &a : cand) {
      a->evict = true;
      coming += a->bytes;
      if (arena->free_bytes() + coming >= r.bytes) break;
    }
  }
  // Insert r into the queue behind the requests of its priority or higher.
  void enqueue(const std::shared_ptr<Request> &r) {
    auto pos = std::find_if(pending.begin(), pending.end(),
                            [&](const std::shared_ptr<Request> &q) { return q->priority < r->priority; });
    pending.insert(pos, r);
  }
  void loop();
  int step();  // one admit / launch / retire round; caller does not hold mu
  int finish(Launch &L);  // wait for a queued launch, check it, retire what it finished
};

int fpsx_server::step() {
  std::vector<std::shared_ptr<Request>> admitted;
  bool ended = false;  // a request ended without finishing: wake waiters
  {
    std::unique_lock<std::mutex> lk(mu);
    const auto now = Clock::now();
    // Deadlines: a queued request leaves the queue, a running one stops at
    // this chunk boundary unless its last chunk is already queued.
    for (auto it = pending.begin(); it != pending.end();) {
      if ((*it)->has_deadline && now >= (*it)->deadline) {
        end_request(**it, FPSX_EDEADLINE, now);
        it = pending.erase(it);
        ended = true;
      } else {
        ++it;
      }
    }
    for (auto &r : active)
      if (!r->stop && !r->finishing && r->has_deadline && now >= r->deadline) {
        r->stop = true;
        r->stop_code = FPSX_EDEADLINE;
      }
    while (!pending.empty()) {
      auto &r = pending.front();
      // Launches run one after another, so only the requests still to be
      // launched count against the co-resident grid, and of those only the
      // ones this request cannot preempt (its priority or higher).
      uint32_t members = 0;
      for (auto &a : active)
        if (!a->finishing && !a->stop && a->priority >= r->priority) members += a->team;
      if (members + r->team > grid_cap) break;
      size_t off;
      if (!arena->alloc(r->bytes, &off)) {
        choose_victims(*r);
        break;  // wait for memory to free up
      }
      r->off = off;
      r->pts_off = off;
      r->state_off = off + align_up((size_t)r->n * r->dim * sizeof(float));
      r->out_off = r->state_off + align_up(((size_t)r->n + 4) * sizeof(float));
      r->sorted_off = r->out_off + align_up((size_t)r->m * sizeof(uint64_t));
      r->perm_off = r->sorted_off + align_up((size_t)r->n * r->dim * sizeof(float));
      r->t_admit = now;
      admitted.push_back(r);
      active.push_back(r);
      pending.pop_front();
    }
  }
  // Upload newly admitted clouds on the copy stream (overlaps the kernel); an
  // evicted request gets its saved range back as it was.
  for (auto &r : admitted) {
    char *base = (char *)arena_base;
    if (!r->saved.empty()) {
      hipError_t e = hipMemcpyAsync(base + r->off, r->saved.data(), r->bytes,
                                    hipMemcpyHostToDevice, copy);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&r->uploaded, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(r->uploaded, copy);
      if (e != hipSuccess) return fail(FPSX_EHIP, std::string("restore: ") + hipGetErrorString(e));
      continue;
    }
    hipError_t e = hipMemcpyAsync(base + r->pts_off, r->soa.data(), r->soa.size() * sizeof(float),
                                  hipMemcpyHostToDevice, copy);
    if (e == hipSuccess && pruned) {
      const int rc = spatial_order((const float *)(base + r->pts_off), 1, r->n, r->n,
                                   (uint32_t)(64 * k->ppt), (float *)(base + r->sorted_off),
                                   (uint32_t *)(base + r->perm_off), *spatial, copy);
      if (rc) return rc;
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->uploaded, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(r->uploaded, copy);
    if (e != hipSuccess) return fail(FPSX_EHIP, std::string("upload: ") + hipGetErrorString(e));
  }
  // Jobs of the next launch: the uploaded requests not yet in their last
  // chunk, highest priority first, while their teams fit the grid (the rest
  // are paused for this launch), each from where the queued launches leave it.
  std::vector<std::shared_ptr<Request>> in_launch, leaving;
  std::vector<JobSpec> specs;
  bool waiting_upload = false;
  {
    std::lock_guard<std::mutex> lk(mu);
    std::vector<std::shared_ptr<Request>> order(active);
    std::stable_sort(order.begin(), order.end(),
                     [](const std::shared_ptr<Request> &a, const std::shared_ptr<Request> &b) {
                       return a->priority > b->priority;
                     });
    uint32_t members = 0;
    for (auto &r : order) {
      if (r->finishing) continue;
      if (r->stop) {
        // Retire once no queued launch writes its range and its upload
        // (points and spatial order on the copy stream) is complete.
        if (r->inflight) continue;
        if (!r->ready && hipEventQuery(r->uploaded) != hipSuccess) {
          waiting_upload = true;
          continue;
        }
        arena->release(r->off, r->bytes);
        if (r->uploaded) (void)hipEventDestroy(r->uploaded);
        r->uploaded = nullptr;
        active.erase(std::find(active.begin(), active.end(), r));
        end_request(*r, r->stop_code, Clock::now());
        ended = true;
        continue;
      }
      if (r->evict) {
        if (!r->inflight) leaving.push_back(r);  // no queued launch writes its range
        continue;
      }
      if (!r->ready) {
        if (hipEventQuery(r->uploaded) != hipSuccess) {
          waiting_upload = true;
          continue;
        }
        r->ready = true;
        r->soa.clear();
        r->soa.shrink_to_fit();
        r->saved.clear();
        r->saved.shrink_to_fit();
      }
      JobSpec s;
      s.pt_off = r->pts_off / sizeof(float);
      s.ld = r->n;
      s.out_off = r->out_off / sizeof(uint64_t);
      s.state_off = r->state_off / sizeof(float);
      s.start = r->start;
      s.n = r->n;
      s.m = r->m;
      s.s_begin = r->next;
      s.s_end = std::min<uint32_t>(r->m, (r->next == 0 ? 1u : r->next) + chunk);
      s.sorted_off = r->sorted_off / sizeof(float);
      s.perm_off = r->perm_off / sizeof(uint32_t);
      if (members + r->team > grid_cap) {
        ++n_paused;  // preempted by higher-priority requests for this launch
        continue;
      }
      members += r->team;
      // Committed to this launch (a failure to launch stops the server).
      r->next = s.s_end;
      if (s.s_end >= r->m) r->finishing = true;
      ++r->inflight;
      specs.push_back(s);
      in_launch.push_back(r);
    }
  }
  // Evictions: copy each range to host memory (nothing in flight writes it),
  // free it and queue the request again at its priority.
  for (auto &r : leaving) {
    r->saved.resize(r->bytes);
    hipError_t e = hipMemcpy(r->saved.data(), (char *)arena_base + r->off, r->bytes,
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(FPSX_EHIP, std::string("evict: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(mu);
    arena->release(r->off, r->bytes);
    if (r->uploaded) (void)hipEventDestroy(r->uploaded);
    r->uploaded = nullptr;
    r->ready = false;
    r->evict = false;
    active.erase(std::find(active.begin(), active.end(), r));
    ++n_evicted;
    if (r->stop) {  // cancelled or expired meanwhile
      end_request(*r, r->stop_code, Clock::now());
      ended = true;
    } else {
      // ahead of the requests of its priority that never ran
      auto pos = std::find_if(pending.begin(), pending.end(), [&](const std::shared_ptr<Request> &q) {
        return q->priority < r->priority || (q->priority == r->priority && q->saved.empty());
      });
      pending.insert(pos, r);
    }
  }
  if (ended) cv_done.notify_all();
  Launch &head = slot[cur];
  if (specs.empty()) {
    if (head.busy) return finish(head);
    if (waiting_upload) {
      // Everything admitted is still uploading: wait for the copies instead
      // of spinning.
      hipError_t e = hipStreamSynchronize(copy);
      if (e != hipSuccess) return fail(FPSX_EHIP, std::string("upload: ") + hipGetErrorString(e));
    }
    return FPSX_OK;
  }
  Launch &L = head.busy ? slot[cur ^ 1] : head;
  std::vector<BatchJob> jobs;
  std::vector<uint32_t> mj;
  uint32_t areas = 0;
  int rc = build_jobs(specs.data(), specs.size(), k, 0, jobs, mj, &areas, pruned);
  if (rc) return rc;
  // pt/out/state/sorted/perm offsets are element offsets from the arena base for every view
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    rc = pruned ? L.run.prepare(*ctx, k, jobs, mj, areas, (const float *)arena_base,
                                (uint64_t *)arena_base, (float *)arena_base,
                                (const float *)arena_base, (const uint32_t *)arena_base, true)
                : L.run.prepare(*ctx, k, jobs, mj, areas, (const float *)arena_base,
                                (uint64_t *)arena_base, (float *)arena_base);
  }
  if (rc) return rc;
  hipError_t e = hipEventRecord(L.t0, compute);
  if (e != hipSuccess) return fail(FPSX_EHIP, std::string("serve launch: ") + hipGetErrorString(e));
  rc = L.run.enqueue(compute);
  if (rc) return rc;
  e = hipEventRecord(L.t1, compute);
  if (e != hipSuccess) return fail(FPSX_EHIP, std::string("serve launch: ") + hipGetErrorString(e));
  L.reqs = in_launch;
  L.s_end.resize(specs.size());
  for (size_t i = 0; i < specs.size(); ++i) L.s_end[i] = specs[i].s_end;
  L.busy = true;
  // With a launch queued behind it, wait for the older one and retire.
  if (&L != &head) return finish(head);
  return FPSX_OK;
}

int fpsx_server::finish(Launch &L) {
  hipError_t e = hipEventSynchronize(L.t1);
  if (e != hipSuccess) return fail(FPSX_EHIP, std::string("serve launch: ") + hipGetErrorString(e));
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, L.t0, L.t1) == hipSuccess) busy_s += ms * 1e-3;
  L.busy = false;
  cur ^= 1;  // the other slot (if queued) is now the oldest
  uint32_t st = 0;
  e = hipMemcpy(&st, L.run.d_status.p, sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(FPSX_EHIP, std::string("status: ") + hipGetErrorString(e));
  if (st) return status_code(st);
  ++launches;
  // Retire the requests this launch finished.
  std::vector<std::shared_ptr<Request>> finished;
  {
    std::lock_guard<std::mutex> lk(mu);
    for (auto &r : L.reqs) --r->inflight;
  }
  for (size_t i = 0; i < L.reqs.size(); ++i)
    if (L.s_end[i] >= L.reqs[i]->m) finished.push_back(L.reqs[i]);
  L.reqs.clear();
  for (auto &r : finished) {
    r->picks.resize(r->m);
    e = hipMemcpy(r->picks.data(), (char *)arena_base + r->out_off, r->m * sizeof(uint64_t),
                  hipMemcpyDeviceToHost);
    if (e != hipSuccess) r->error = FPSX_EHIP;
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    const auto now = Clock::now();
    for (auto &r : finished) {
      r->done = true;
      r->t_done = now;
      arena->release(r->off, r->bytes);
      if (r->uploaded) (void)hipEventDestroy(r->uploaded);
      r->uploaded = nullptr;
      active.erase(std::find(active.begin(), active.end(), r));
      ++completed;
      samples_done += r->m;
    }
  }
  if (!finished.empty()) cv_done.notify_all();
  return FPSX_OK;
}

void fpsx_server::loop() {
  // Without its device this thread can run nothing: every request fails.
  int rc = hipSetDevice(ctx->device) == hipSuccess
               ? FPSX_OK
               : fail(FPSX_EHIP, "server thread: hipSetDevice failed");
  for (;;) {
    if (rc == FPSX_OK) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
        if (stop && pending.empty() && active.empty()) return;
      }
      rc = step();
    }
    if (rc != FPSX_OK) {
      std::lock_guard<std::mutex> lk(mu);
      fatal = rc;
      fatal_msg = fpsx_last_error();
      for (auto &r : active) {
        r->done = true;
        r->error = rc;
        if (r->uploaded) (void)hipEventDestroy(r->uploaded);
        r->uploaded = nullptr;
      }
      for (auto &r : pending) {
        r->done = true;
        r->error = rc;
      }
      active.clear();
      pending.clear();
      cv_done.notify_all();
      return;
    }
  }
}

FPSX_EXPORT int fpsx_server_create(size_t hbm_budget_bytes, unsigned chunk_passes, int block,
                                   int ppt, fpsx_server **out) {
  if (!out) return fail(FPSX_EINVAL, "null output handle");
  *out = nullptr;
  if (hbm_budget_bytes < (1u << 20)) return fail(FPSX_EINVAL, "budget must be >= 1 MiB");
  std::unique_ptr<fpsx_server> s(new fpsx_server());
  int rc = get_ctx(&s->ctx);
  if (rc) return rc;
  const int B = block ? block : 1024, P = ppt ? ppt : 16;
  const char *plain = std::getenv("FPSX_SERVE_PLAIN");
  s->pruned = !(plain && plain[0] == '1');
  if (s->pruned) {
    s->k = batch_pp_kernel(B, P, true);
    if (!s->k) s->pruned = false;  // no pruned kernel of that shape: the plain one
  }
  if (!s->pruned) {
    size_t cnt;
    const KernelEntry *t = batch_table(&cnt);
    for (size_t i = 0; i < cnt; ++i)
      if (t[i].dim == 3 && t[i].block == B && t[i].ppt == P) s->k = &t[i];
  }
  if (!s->k) return fail(FPSX_EINVAL, "no batch kernel with that shape");
  {
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    s->grid_cap = (uint32_t)s->ctx->max_blocks(s->k, 0);
  }
  if (s->grid_cap == 0) return fail(FPSX_EHIP, "batch kernel cannot be resident");
  s->chunk = chunk_passes ? chunk_passes : 256;
  s->budget = hbm_budget_bytes / Arena::kAlign * Arena::kAlign;
  hipError_t e = hipMalloc(&s->arena_base, s->budget);
  if (e != hipSuccess) return fail(FPSX_ENOMEM, "arena: " + std::string(hipGetErrorString(e)));
  s->arena.reset(new Arena(s->budget));
  if (s->pruned) {
    // Scratch of the spatial order sized once for the largest request a team
    // can hold, so admissions allocate nothing.
    s->spatial = new_spatial_scratch();
    rc = spatial_reserve(*s->spatial, std::min<uint32_t>(s->grid_cap, kMaxTeam) * (uint32_t)(B * P),
                         (uint32_t)(64 * P));
    if (rc) {
      delete_spatial_scratch(s->spatial);
      (void)hipFree(s->arena_base);
      return rc;
    }
  }
  e = hipStreamCreateWithFlags(&s->compute, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking);
  if (e != hipSuccess) {
    const std::string msg = "streams: " + std::string(hipGetErrorString(e));
    if (s->compute) (void)hipStreamDestroy(s->compute);
    delete_spatial_scratch(s->spatial);
    (void)hipFree(s->arena_base);
    return fail(FPSX_EHIP, msg);
  }
  for (auto &L : s->slot) {
    if (e == hipSuccess) e = hipEventCreate(&L.t0);
    if (e == hipSuccess) e = hipEventCreate(&L.t1);
  }
  if (e != hipSuccess) {
    const std::string msg = "events: " + std::string(hipGetErrorString(e));
    (void)hipStreamDestroy(s->compute);
    (void)hipStreamDestroy(s->copy);
    delete_spatial_scratch(s->spatial);
    (void)hipFree(s->arena_base);
    return fail(FPSX_EHIP, msg);
  }
  fpsx_server *raw = s.get();
  raw->thread = std::thread([raw] { raw->loop(); });
  *out = s.release();
  return FPSX_OK;
}

FPSX_EXPORT int fpsx_server_submit_ex(fpsx_server *s, const float *points, size_t n_points,
                                      size_t dim, size_t n_samples, uint64_t start_idx,
                                      int priority, double deadline_s, uint64_t *request_id) {
  if (!s || !points || !request_id) return fail(FPSX_EINVAL, "null argument");
  if (std::isnan(deadline_s)) return fail(FPSX_EINVAL, "deadline is NaN");
  if (dim != 3) return fail(FPSX_EBADDIM, "the serving loop takes 3-D clouds");
  if (n_points == 0 || n_points >= (size_t(1) << 31)) return fail(FPSX_ETOOBIG, "bad n_points");
  if (n_samples == 0 || n_samples > n_points) return fail(FPSX_EINVAL, "n_samples must be in [1, n]");
  if (start_idx >= n_points) return fail(FPSX_EBADSTART, "start index out of range");
  auto r = std::make_shared<Request>();
  r->n = (uint32_t)n_points;
  r->m = (uint32_t)n_samples;
  r->dim = (uint32_t)dim;
  r->start = start_idx;
  r->team = s->team_of(r->n);
  r->bytes = s->footprint(*r);
  if (r->bytes > s->budget)
    return fail(FPSX_EBUDGET, "request needs " + std::to_string(r->bytes) +
                                  " bytes of HBM, more than the whole budget");
  if (r->team > s->grid_cap || r->team > kMaxTeam || r->team > (uint32_t)s->k->block)
    return fail(FPSX_ETOOBIG, "cloud too large for the serving kernel's team limit");
  r->soa.resize(n_points * dim);
  for (size_t i = 0; i < n_points; ++i)
    for (size_t d = 0; d < dim; ++d) r->soa[d * n_points + i] = points[i * dim + d];
  r->priority = priority;
  r->t_submit = Clock::now();
  if (deadline_s > 0) {
    r->has_deadline = true;
    // (capped at about 30 years so the time point cannot overflow)
    r->deadline = r->t_submit + std::chrono::duration_cast<Clock::duration>(
                                    std::chrono::duration<double>(std::min(deadline_s, 1e9)));
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal) return fail(s->fatal, "server stopped: " + s->fatal_msg);
    r->id = s->next_id++;
    s->all[r->id] = r;
    s->enqueue(r);  // behind every pending request of its priority or higher
  }
  s->cv_work.notify_one();
  *request_id = r->id;
  return FPSX_OK;
}

FPSX_EXPORT int fpsx_server_submit(fpsx_server *s, const float *points, size_t n_points, size_t dim,
                                   size_t n_samples, uint64_t start_idx, uint64_t *request_id) {
  return fpsx_server_submit_ex(s, points, n_points, dim, n_samples, start_idx, 0, 0.0, request_id);
}

FPSX_EXPORT int fpsx_server_wait_for(fpsx_server *s, uint64_t request_id, double timeout_s,
                                     uint64_t *out, double *latency_s) {
  if (!s) return fail(FPSX_EINVAL, "null server");
  if (std::isnan(timeout_s)) return fail(FPSX_EINVAL, "timeout is NaN");
  std::shared_ptr<Request> r;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(request_id);
  if (it == s->all.end()) return fail(FPSX_EINVAL, "unknown request id");
  r = it->second;
  auto done = [&] { return r->done; };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, done);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(std::min(timeout_s, 1e9)),
                                  done)) {
    return fail(FPSX_EPENDING, "request " + std::to_string(request_id) + " still running");
  }
  s->all.erase(request_id);
  if (r->error == FPSX_ECANCELED) return fail(r->error, "request was cancelled");
  if (r->error == FPSX_EDEADLINE) return fail(r->error, "request's deadline passed");
  if (r->error) return fail(r->error, "request failed: " + s->fatal_msg);
  if (out) std::memcpy(out, r->picks.data(), r->m * sizeof(uint64_t));
  if (latency_s) *latency_s = secs_since(r->t_submit, r->t_done);
  return FPSX_OK;
}

FPSX_EXPORT int fpsx_server_wait(fpsx_server *s, uint64_t request_id, uint64_t *out,
                                 double *latency_s) {
  return fpsx_server_wait_for(s, request_id, -1.0, out, latency_s);
}

FPSX_EXPORT int fpsx_server_cancel(fpsx_server *s, uint64_t request_id, int *cancelled) {
  if (cancelled) *cancelled = 0;
  if (!s) return fail(FPSX_EINVAL, "null server");
  bool wake_waiters = false, wake_loop = false;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(request_id);
    if (it == s->all.end()) return fail(FPSX_EINVAL, "unknown request id");
    Request &r = *it->second;
    if (r.done) {
      // finished (or already ended): nothing to stop
      if (cancelled) *cancelled = (r.error != FPSX_OK);
      return FPSX_OK;
    }
    auto q = std::find(s->pending.begin(), s->pending.end(), it->second);
    if (q != s->pending.end()) {
      s->pending.erase(q);
      s->end_request(r, FPSX_ECANCELED, Clock::now());
      wake_waiters = true;
    } else if (r.finishing) {
      return FPSX_OK;  // its last chunk is queued: it finishes
    } else {
      if (!r.stop) {
        r.stop = true;
        r.stop_code = FPSX_ECANCELED;
      }
      wake_loop = true;
    }
    if (cancelled) *cancelled = 1;
  }
  if (wake_waiters) s->cv_done.notify_all();
  if (wake_loop) s->cv_work.notify_one();
  return FPSX_OK;
}

FPSX_EXPORT int fpsx_server_counters(fpsx_server *s, uint64_t *cancelled, uint64_t *expired,
                                     uint64_t *paused, uint64_t *evicted) {
  if (!s) return fail(FPSX_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (cancelled) *cancelled = s->n_cancelled;
  if (expired) *expired = s->n_expired;
  if (paused) *paused = s->n_paused;
  if (evicted) *evicted = s->n_evicted;
  return s->fatal;
}

FPSX_EXPORT int fpsx_server_stats(fpsx_server *s, uint64_t *launches, uint64_t *completed,
                                  uint64_t *samples, double *busy_s, size_t *hbm_in_use,
                                  size_t *pending, size_t *active) {
  if (!s) return fail(FPSX_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (launches) *launches = s->launches;
  if (completed) *completed = s->completed;
  if (samples) *samples = s->samples_done;
  if (busy_s) *busy_s = s->busy_s;
  if (hbm_in_use) *hbm_in_use = s->arena->used_bytes();
  if (pending) *pending = s->pending.size();
  if (active) *active = s->active.size();
  return s->fatal;
}

FPSX_EXPORT void fpsx_server_destroy(fpsx_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->thread.joinable()) s->thread.join();
  // Destroy path (no error can be returned): the streams drain before their
  // memory is released.
  (void)hipStreamSynchronize(s->compute);
  (void)hipStreamSynchronize(s->copy);
  for (auto &L : s->slot) {
    L.run.d_jobs.free();
    L.run.d_member_job.free();
    L.run.d_counters.free();
    L.run.d_xchg.free();
    L.run.d_status.free();
    if (L.t0) (void)hipEventDestroy(L.t0);
    if (L.t1) (void)hipEventDestroy(L.t1);
  }
  if (s->arena_base) (void)hipFree(s->arena_base);
  delete_spatial_scratch(s->spatial);
  (void)hipStreamDestroy(s->compute);
  (void)hipStreamDestroy(s->copy);
  delete s;
}

// fpsx_api.hip -- host runtime and C ABI of libfpsx.so (declared in include/fpsx.h).
//
// Per device the library keeps one context: a non-blocking stream at the
// default priority, grow-only device buffers (points, starts, output,
// exchange area, status word, streamed min-distances) and the occupancy of
// every kernel.  A call plans a launch (kernel shape and grid, see plan()),
// uploads, clears the exchange area and status word with fill kernels,
// launches once, and copies the picks back.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fps_device.hpp"
#include "fps_kernels.hpp"
#include "fps_npdu.hpp"
#include "fps_pruned.hpp"
#include "fpsx.h"
#include "fpsx_internal.hpp"
#include "kdline.hpp"

#define FPSX_EXPORT extern "C" __attribute__((visibility("default")))

namespace fpsx {

static thread_local std::string t_err;
constexpr int kDefaultReplicas = 8;

int fail(int code, const std::string &msg) {
  t_err = msg;
  return code;
}

static int hip_fail(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? FPSX_ENOMEM : FPSX_EHIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                  \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_fail(_e, #expr);  \
  } while (0)

hipError_t DevBuf::reserve(size_t bytes) {
  if (bytes <= cap) return hipSuccess;
  if (p) {
    hipError_t e = hipFree(p);
    p = nullptr;
    cap = 0;
    if (e != hipSuccess) return e;
  }
  size_t want = bytes < 4096 ? 4096 : bytes;
  hipError_t e = hipMalloc(&p, want);
  if (e == hipSuccess) cap = want;
  return e;
}

void DevBuf::free() {
  if (p) (void)hipFree(p);  // release path: nothing to report to
  p = nullptr;
  cap = 0;
}

DevBuf::~DevBuf() {
  // Device memory is released with the process; contexts live until exit.
}

// Realtime-clock ticks an exchange may wait: FPSX_EXCHANGE_TIMEOUT_S, else 2 s.
static uint64_t default_spin_ticks(int khz) {
  const char *sec = std::getenv("FPSX_EXCHANGE_TIMEOUT_S");
  double secs = sec ? std::atof(sec) : 2.0;
  if (!(secs > 0)) secs = 2.0;
  return (uint64_t)((double)khz * 1000.0 * secs);
}

static std::mutex g_ctx_mu;
static std::map<int, std::unique_ptr<Ctx>> g_ctx;

int get_ctx(Ctx **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  auto it = g_ctx.find(dev);
  if (it != g_ctx.end()) {
    *out = it->second.get();
    return FPSX_OK;
  }
  std::unique_ptr<Ctx> c(new Ctx());
  c->device = dev;
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIP_TRY(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev));
  int khz = 0;
  HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (khz <= 0) khz = 100000;
  c->clock_khz = khz;
  c->spin_ticks = default_spin_ticks(khz);

  HIP_TRY(hipHostMalloc((void **)&c->host_status, sizeof(uint32_t), hipHostMallocDefault));
  *out = c.get();
  g_ctx[dev] = std::move(c);
  return FPSX_OK;
}

int Ctx::max_blocks(const KernelEntry *k, size_t shmem) {
  auto key = std::make_pair(k->fn, shmem);
  auto it = occ.find(key);
  if (it != occ.end()) return it->second;
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k->fn, k->block, shmem);
  if (e != hipSuccess) nb = 0;
  int total = nb * cus;
  occ[key] = total;
  return total;
}

// ---------------------------------------------------------------------------
// Planner.  FPSX_PLAN overrides it (tests use this to reach every kernel):
//   "resident:B:PPT"   register-resident kernel, grid = ceil(n / (B*PPT))
//   "stream:B[:G]"     streaming kernel of the cloud's dimension (or generic)
//   "generic:B[:G]"    generic any-dimension streaming kernel
// Automatic choice (dimensions with resident shapes):
//   n <= 16384: one workgroup, the smallest shape holding the cloud;
//   otherwise: as many 4096-point workgroups as the cloud needs, growing the
//   per-workgroup share (PPT 8, 16) when that exceeds the co-resident limit;
//   beyond the register file: the streaming kernel.
// ---------------------------------------------------------------------------
static const KernelEntry *find_resident(int dim, int B, int ppt, int tie) {
  size_t cnt;
  const KernelEntry *t = resident_table(&cnt);
  for (size_t i = 0; i < cnt; ++i)
    if (t[i].dim == dim && t[i].block == B && t[i].ppt == ppt && t[i].tie == tie) return &t[i];
  return nullptr;
}

static const KernelEntry *find_pruned(int dim, int B, int ppt, int tie) {
  size_t cnt;
  const KernelEntry *t = pruned_table(&cnt);
  for (size_t i = 0; i < cnt; ++i)
    if (t[i].dim == dim && t[i].block == B && t[i].ppt == ppt && t[i].tie == tie) return &t[i];
  return nullptr;
}

static const KernelEntry *find_solo(int dim, int B, int ppt, int tie) {
  size_t cnt;
  const KernelEntry *t = solo_table(&cnt);
  for (size_t i = 0; i < cnt; ++i)
    if (t[i].dim == dim && t[i].block == B && t[i].ppt == ppt && t[i].tie == tie) return &t[i];
  return nullptr;
}

static const KernelEntry *find_pair(int dim, int B, int ppt, int tie) {
  size_t cnt;
  const KernelEntry *t = pair_table(&cnt);
  for (size_t i = 0; i < cnt; ++i)
    if (t[i].dim == dim && t[i].block == B && t[i].ppt == ppt && t[i].tie == tie) return &t[i];
  return nullptr;
}

static const KernelEntry *find_stream(int dim, int B, int tie) {
  size_t cnt;
  const KernelEntry *t = stream_table(&cnt);
  for (size_t i = 0; i < cnt; ++i)
    if (t[i].dim == dim && t[i].block == B && t[i].tie == tie) return &t[i];
  return nullptr;
}

static int plan_stream(Ctx &c, size_t n, size_t dim, int tie, int B, size_t want_g, bool generic,
                       Plan &p) {
  const KernelEntry *k = generic ? nullptr : find_stream((int)dim, B, tie);
  if (!k) k = find_stream(0, B, tie);
  if (!k) return fail(FPSX_EINVAL, "no streaming kernel with block " + std::to_string(B));
  p.kind = PlanKind::Stream;
  p.k = k;
  p.block = (uint32_t)B;
  p.shmem = (k->dim == 0) ? dim * sizeof(float) : 0;
  const size_t gmax = (size_t)c.max_blocks(k, p.shmem);
  if (gmax == 0) return fail(FPSX_EHIP, "streaming kernel cannot be resident on this device");
  size_t g = want_g ? want_g : (n + B - 1) / B;
  if (g > gmax) g = gmax;
  if (g > (size_t)B) g = (size_t)B;  // exchange polls one workgroup per lane
  if (g > (size_t)kPageWords) g = kPageWords;
  if (g < 1) g = 1;
  size_t chunk = (n + g - 1) / g;
  g = (n + chunk - 1) / chunk;  // no empty workgroup
  p.grid = (uint32_t)g;
  p.chunk = (uint32_t)chunk;
  return FPSX_OK;
}

// Streaming pruned plan: 1024-thread workgroups; G = 64 (lattice16m at 32,
// 64, 128 and 256 workgroups: 23.2, 20.3, 22.2 and 26.2 ms), more where a
// workgroup would cache over 2048 blocks (96 KB of LDS metadata), bounded by
// co-residency and the exchange limits.  want_g > 0 forces G.  pair: the
// paired form, which needs 2 <= G <= 64.
static int plan_stream_pruned(Ctx &c, size_t n, int tie, size_t want_g, Plan &p,
                              bool pair = false) {
  size_t cnt;
  const KernelEntry *t = pair ? stream_pair_table(&cnt) : stream_pruned_table(&cnt);
  const KernelEntry *k = nullptr;
  for (size_t i = 0; i < cnt; ++i)
    if (t[i].tie == tie) k = &t[i];
  if (!k) return fail(FPSX_EINVAL, "no streaming pruned kernel");
  const size_t nblocks = (n + kSBlock - 1) / kSBlock;
  size_t g = want_g ? want_g
                    : std::max<size_t>(std::min<size_t>(64, nblocks), (nblocks + 2047) / 2048);
  g = std::min<size_t>(g, std::min<size_t>(nblocks, 256));
  if (pair && (g < 2 || g > kPairMaxGrid))
    return fail(FPSX_ETOOBIG, "paired streaming plan: needs 2..64 workgroups");
  size_t nbl = (nblocks + g - 1) / g;
  size_t shmem = (size_t)(pair ? kSPairMetaWords : kSMetaWords) * nbl * sizeof(int);
  if (shmem > (pair ? 128 : 96) * 1024)
    return fail(FPSX_ETOOBIG, "streaming pruned plan: too many blocks per workgroup");
  const size_t gmax = (size_t)c.max_blocks(k, shmem);
  if (g > gmax) return fail(FPSX_ETOOBIG, "streaming pruned plan: grid exceeds co-residency");
  p.kind = pair ? PlanKind::StreamPair : PlanKind::StreamPruned;
  p.k = k;
  p.block = 1024;
  p.grid = (uint32_t)g;
  p.chunk = (uint32_t)(nbl * kSBlock);
  p.shmem = shmem;
  return FPSX_OK;
}

static int plan_resident_shape(Ctx &c, size_t n, const KernelEntry *k, Plan &p,
                               PlanKind kind = PlanKind::Resident) {
  const uint32_t g = resident_grid(n, k, (size_t)c.max_blocks(k, 0));
  if (!g) return -1;
  p.kind = kind;
  p.k = k;
  p.block = (uint32_t)k->block;
  p.grid = g;
  p.chunk = (uint32_t)(k->block * k->ppt);
  p.shmem = 0;
  return FPSX_OK;
}

static int plan_from_tuned(Ctx &c, size_t n, size_t dim, int tie, Plan &p) {
  auto it = c.tuned.find(std::make_tuple(n, dim, tie));
  if (it == c.tuned.end()) return -1;
  p = it->second;
  return FPSX_OK;
}

int plan(Ctx &c, size_t n, size_t dim, int tie, size_t m, Plan &p) {
  const char *env = std::getenv("FPSX_PLAN");
  if (env && *env) {
    std::string s(env);
    int a = 0, b = 0, g = 0;
    if (s.rfind("spruned", 0) == 0) {
      if (dim != 3) return fail(FPSX_EINVAL, "FPSX_PLAN: the streaming pruned kernel is 3-D only");
      int got = std::sscanf(s.c_str(), "spruned:%d", &g);
      return plan_stream_pruned(c, n, tie, got == 1 ? (size_t)g : 0, p);
    }
    if (s.rfind("spair", 0) == 0) {
      if (dim != 3) return fail(FPSX_EINVAL, "FPSX_PLAN: the streaming pruned kernel is 3-D only");
      int got = std::sscanf(s.c_str(), "spair:%d", &g);
      return plan_stream_pruned(c, n, tie, got == 1 ? (size_t)g : 0, p, true);
    }
    int pw = 0;
    if (std::sscanf(s.c_str(), "pairw%d:%d:%d", &pw, &a, &b) == 3) {
      // The paired kernel with pw polling waves, whatever the grid.
      const KernelEntry *k = dim == 3 ? pair_wide_kernel(a, b, tie, pw) : nullptr;
      if (!k) return fail(FPSX_EINVAL, "FPSX_PLAN: no wide paired kernel " + s);
      if (plan_resident_shape(c, n, k, p, PlanKind::Pair) != FPSX_OK || p.grid < 2 ||
          p.grid > (uint32_t)(64 * pw))
        return fail(FPSX_ETOOBIG, "FPSX_PLAN: the grid does not fit " + s);
      return FPSX_OK;
    }
    if (std::sscanf(s.c_str(), "pair:%d:%d", &a, &b) == 2) {
      const KernelEntry *k = find_pair((int)dim, a, b, tie);
      if (!k) return fail(FPSX_EINVAL, "FPSX_PLAN: no paired kernel for dim " +
                                           std::to_string(dim) + " shape " + s);
      if (plan_resident_shape(c, n, k, p, PlanKind::Pair) == FPSX_OK && p.grid > kPairMaxGrid) {
        // Beyond one polling wave's 64 workgroups: the form with 2 or 4.
        const KernelEntry *kw = pair_wide_kernel(a, b, tie, p.grid <= 128 ? 2 : 4);
        if (kw && plan_resident_shape(c, n, kw, p, PlanKind::Pair) == FPSX_OK &&
            p.grid <= kPairWideMaxGrid)
          return FPSX_OK;
        return fail(FPSX_ETOOBIG, "FPSX_PLAN: the paired kernel needs 2..256 workgroups: " + s);
      }
      if (p.k != k || p.grid < 2)
        return fail(FPSX_ETOOBIG, "FPSX_PLAN: the paired kernel needs 2..256 workgroups: " + s);
      return FPSX_OK;
    }
    if (std::sscanf(s.c_str(), "solo:%d:%d", &a, &b) == 2) {
      // One workgroup paired at the wave level.
      const KernelEntry *k = find_solo((int)dim, a, b, tie);
      if (!k) return fail(FPSX_EINVAL, "FPSX_PLAN: no one-workgroup paired kernel for dim " +
                                           std::to_string(dim) + " shape " + s);
      if (plan_resident_shape(c, n, k, p, PlanKind::Pair) != FPSX_OK || p.grid != 1)
        return fail(FPSX_ETOOBIG, "FPSX_PLAN: the cloud does not fit one workgroup of " + s);
      return FPSX_OK;
    }
    if (std::sscanf(s.c_str(), "pruned:%d:%d", &a, &b) == 2) {
      const KernelEntry *k = find_pruned((int)dim, a, b, tie);
      if (!k) return fail(FPSX_EINVAL, "FPSX_PLAN: no pruned kernel for dim " +
                                           std::to_string(dim) + " shape " + s);
      if (plan_resident_shape(c, n, k, p, PlanKind::Pruned) != FPSX_OK)
        return fail(FPSX_ETOOBIG, "FPSX_PLAN: cloud exceeds the co-resident grid of " + s);
      return FPSX_OK;
    }
    if (std::sscanf(s.c_str(), "resident:%d:%d", &a, &b) == 2) {
      const KernelEntry *k = find_resident((int)dim, a, b, tie);
      if (!k) return fail(FPSX_EINVAL, "FPSX_PLAN: no resident kernel for dim " +
                                           std::to_string(dim) + " shape " + s);
      if (plan_resident_shape(c, n, k, p) != FPSX_OK)
        return fail(FPSX_ETOOBIG, "FPSX_PLAN: cloud exceeds the co-resident grid of " + s);
      return FPSX_OK;
    }
    if (s.rfind("stream:", 0) == 0 || s.rfind("generic:", 0) == 0) {
      bool generic = s[0] == 'g';
      int got = std::sscanf(s.c_str() + (generic ? 8 : 7), "%d:%d", &a, &g);
      if (got < 1) return fail(FPSX_EINVAL, "FPSX_PLAN: bad value " + s);
      return plan_stream(c, n, dim, tie, a, got == 2 ? (size_t)g : 0, generic, p);
    }
    return fail(FPSX_EINVAL, "FPSX_PLAN: bad value " + s);
  }

  if (plan_from_tuned(c, n, dim, tie, p) == FPSX_OK) return FPSX_OK;

  // 3-D clouds with enough samples: spatial order + per-wave pruning.  The
  // ordering costs a few radix-sort passes over the cloud, repaid once most
  // passes touch few waves.  Measured on MI355X (uniform and clustered
  // clouds, whole call): up to 16 K points one pruned workgroup (no
  // exchange) with the smallest share holding the cloud beats the resident
  // kernel from 4 K points and 1024 samples (16 K points, 4096 samples:
  // 4.6 ms against 7.2 ms); beyond, the paired form with the smallest share
  // its 64-workgroup limit allows (64 K points, 16384 samples: 24.3 ms at
  // 4 points per thread against 29.1 ms at 16).
  const char *np = std::getenv("FPSX_NO_PRUNE");
  const char *npair = std::getenv("FPSX_NO_PAIR");
  const bool prune = dim == 3 && !(np && np[0] == '1');
  if (prune && n <= 16384 && ((n >= 4096 && m >= 1024) || (n > 4096 && m >= 512))) {
    // One workgroup, paired at the wave level (up to two picks per barrier)
    // unless FPSX_NO_PAIR=1.
    const bool solo = !(npair && npair[0] == '1');
    for (int ppt : {4, 8, 16}) {
      if ((size_t)1024 * ppt < n) continue;
      const KernelEntry *k = solo ? find_solo(3, 1024, ppt, tie) : find_pruned(3, 1024, ppt, tie);
      if (k && plan_resident_shape(c, n, k, p, solo ? PlanKind::Pair : PlanKind::Pruned) == FPSX_OK)
        return FPSX_OK;
    }
  }
  if (prune && n > 16384 && m >= 1024) {
    // Paired form first (half the exchanges) where its grid limit allows.
    if (!(npair && npair[0] == '1')) {
      // 1024 x 17: 16 points per thread in registers and one in LDS, so 64
      // workgroups hold up to 1,114,112 points without the wider exchange.
      static const int kPa[][2] = {{1024, 4}, {1024, 8}, {1024, 16}, {1024, 17}};
      for (auto &sh : kPa) {
        const KernelEntry *k = find_pair(3, sh[0], sh[1], tie);
        if (k && plan_resident_shape(c, n, k, p, PlanKind::Pair) == FPSX_OK && p.grid >= 2 &&
            p.grid <= kPairMaxGrid)
          return FPSX_OK;
      }
      // Beyond 64 workgroups (more than 1,114,112 points): 1024 x 16 with two or
      // four polling waves, up to 256 workgroups.
      for (int pw : {2, 4}) {
        const KernelEntry *k = pair_wide_kernel(1024, 16, tie, pw);
        if (k && plan_resident_shape(c, n, k, p, PlanKind::Pair) == FPSX_OK &&
            p.grid > kPairMaxGrid && p.grid <= (uint32_t)(64 * pw))
          return FPSX_OK;
      }
    }
    // One pick per exchange (beyond 256 workgroups, or FPSX_NO_PAIR=1): from
    // 64 K points, where it beats the resident kernel.
    static const int kPr[][2] = {{1024, 16}, {256, 32}, {1024, 8}, {1024, 4}};
    for (auto &sh : kPr) {
      const KernelEntry *k = find_pruned(3, sh[0], sh[1], tie);
      if (n >= 65536 && k && plan_resident_shape(c, n, k, p, PlanKind::Pruned) == FPSX_OK)
        return FPSX_OK;
    }
  }

  // Register-resident: one workgroup up to 16 K points, else the feasible
  // shape the per-pass time model prefers.
  if (plan_resident_only(n, dim, tie, [&](const KernelEntry *k) { return c.max_blocks(k, 0); }, p))
    return FPSX_OK;
  // Beyond the register file: 3-D clouds stream only the blocks a pick can
  // change; other dimensions stream everything.
  // The paired form first (up to 64 workgroups, i.e. 64 M points).
  if (dim == 3 && m >= 64 && !(np && np[0] == '1')) {
    if (!(npair && npair[0] == '1') && plan_stream_pruned(c, n, tie, 0, p, true) == FPSX_OK)
      return FPSX_OK;
    if (plan_stream_pruned(c, n, tie, 0, p) == FPSX_OK) return FPSX_OK;
  }
  return plan_stream(c, n, dim, tie, 1024, 0, false, p);
}

static const KernelEntry *find_npdu(int dim, int B, int ppt) {
  size_t cnt;
  const KernelEntry *t = npdu_table(&cnt);
  for (size_t i = 0; i < cnt; ++i)
    if (t[i].dim == dim && t[i].block == B && t[i].ppt == ppt) return &t[i];
  return nullptr;
}

// Windowed-update plans.  FPSX_NPDU_PLAN="D:B:PPT" (D = 3 or 0, the
// any-dimension kernel) overrides.  Otherwise 3-D clouds keep their
// coordinates in registers and other dimensions use the any-dimension
// kernel: one workgroup, the smallest shape holding the cloud, while one
// suffices (16384 / 32768 points); beyond, 1024-thread workgroups with the
// share per thread that minimises 0.012 us per slot + 0.0125 us per
// workgroup of exchange (fitted to MI355X sweeps at 50 K and 100 K points).
// Streaming windowed-update plan: G = 64 (4 M and 9 M points at 32, 64 and
// 128 workgroups: 14.1 / 13.5 / 16.8 and 18.4 / 15.3 / 19.0 ms), more where a
// workgroup would cache over 4096 blocks; want_g > 0 forces G.  3-D clouds
// take the kernel that carries the winner's coordinates (5 LDS words per
// block) unless any_dim is set; the any-dimension kernel keeps the pivot
// (dim words) and 2 words per block.
static int plan_npdu_stream(Ctx &c, size_t n, size_t dim, size_t want_g, Plan &p,
                            bool any_dim = false, bool big_only = false) {
  // Candidates in order: the 3-D kernel, the any-dimension kernel, the
  // any-dimension kernel with 4x larger blocks (clouds whose per-workgroup
  // cache of small blocks would not fit 96 KB of LDS: beyond about 322 M 3-D
  // points, or 805 M with the any-dimension kernel).
  struct Cand {
    const KernelEntry *k;
    size_t blk;
  };
  Cand cands[3];
  int nc = 0;
  if (dim == 3 && !any_dim && !big_only) cands[nc++] = {npdu_stream_kernel(true), (size_t)kNpduBlock};
  if (!big_only) cands[nc++] = {npdu_stream_kernel(false), (size_t)kNpduBlock};
  cands[nc++] = {npdu_stream_kernel_big(), (size_t)kNpduBigBlock};
  for (int i = 0; i < nc; ++i) {
    const KernelEntry *k = cands[i].k;
    const size_t blk = cands[i].blk;
    const size_t nblocks = (n + blk - 1) / blk;
    size_t g = want_g ? want_g
                      : std::max<size_t>(std::min<size_t>(64, nblocks), (nblocks + 4095) / 4096);
    g = std::min<size_t>(g, std::min<size_t>(nblocks, 256));
    const size_t nbl = (nblocks + g - 1) / g;
    const size_t shmem = (k->dim == 3) ? 5 * nbl * sizeof(int)
                                       : dim * sizeof(float) + 2 * nbl * sizeof(int);
    if (shmem > 96 * 1024) continue;
    const size_t gmax = (size_t)c.max_blocks(k, shmem);
    if (g > gmax)
      return fail(FPSX_ETOOBIG, "windowed-update streaming plan: grid exceeds co-residency");
    p.kind = PlanKind::Npdu;
    p.k = k;
    p.block = 1024;
    p.grid = (uint32_t)g;
    p.chunk = (uint32_t)(nbl * blk);
    p.shmem = shmem;
    return FPSX_OK;
  }
  return fail(FPSX_ETOOBIG, "windowed-update streaming plan: too many blocks per workgroup");
}

static const KernelEntry *find_npdu_pair(int B, int ppt) {
  size_t cnt;
  const KernelEntry *t = npdu_pair_table(&cnt);
  for (size_t i = 0; i < cnt; ++i)
    if (t[i].block == B && t[i].ppt == ppt) return &t[i];
  return nullptr;
}

int plan_npdu(Ctx &c, size_t n, size_t dim, Plan &p) {
  const char *env = std::getenv("FPSX_NPDU_PLAN");
  if (env && std::strncmp(env, "stream", 6) == 0) {
    // "stream[:G]"; "stream0[:G]" forces the any-dimension kernel,
    // "streambig[:G]" its large-block instance.
    const bool any = std::strncmp(env, "stream0", 7) == 0;
    const bool big = std::strncmp(env, "streambig", 9) == 0;
    int g = 0;
    const int got = std::sscanf(env + (big ? 9 : any ? 7 : 6), ":%d", &g);
    return plan_npdu_stream(c, n, dim, got == 1 ? (size_t)g : 0, p, any || big, big);
  }
  if (env && std::strncmp(env, "pair:", 5) == 0) {
    int b = 0, q = 0;
    if (dim != 3 || std::sscanf(env, "pair:%d:%d", &b, &q) != 2)
      return fail(FPSX_EINVAL, std::string("FPSX_NPDU_PLAN: bad value (pairs are 3-D) ") + env);
    const KernelEntry *k = find_npdu_pair(b, q);
    if (!k) return fail(FPSX_EINVAL, std::string("FPSX_NPDU_PLAN: no kernel ") + env);
    if (plan_resident_shape(c, n, k, p, PlanKind::Npdu) != FPSX_OK || p.grid < 2 ||
        p.grid > kNpduPairMaxGrid)
      return fail(FPSX_ETOOBIG, std::string("FPSX_NPDU_PLAN: the paired kernel needs 2..64 workgroups: ") + env);
    return FPSX_OK;
  }
  if (env && *env) {
    int d = 0, b = 0, q = 0;
    if (std::sscanf(env, "%d:%d:%d", &d, &b, &q) != 3 || (d != 0 && d != (int)dim))
      return fail(FPSX_EINVAL, std::string("FPSX_NPDU_PLAN: bad value ") + env);
    const KernelEntry *k = find_npdu(d, b, q);
    if (!k) return fail(FPSX_EINVAL, std::string("FPSX_NPDU_PLAN: no kernel ") + env);
    if (plan_resident_shape(c, n, k, p, PlanKind::Npdu) != FPSX_OK)
      return fail(FPSX_ETOOBIG, std::string("FPSX_NPDU_PLAN: cloud exceeds the grid of ") + env);
    return FPSX_OK;
  }
  static const int k3[][2] = {{256, 1}, {256, 2}, {256, 4}, {1024, 1}, {1024, 2}, {1024, 4},
                               {1024, 8}, {1024, 16}};
  static const int k0[][2] = {{256, 1}, {256, 4}, {1024, 4}, {1024, 16}, {1024, 32}};
  const bool d3 = dim == 3;
  const int(*shapes)[2] = d3 ? k3 : k0;
  const size_t ns = d3 ? sizeof(k3) / sizeof(k3[0]) : sizeof(k0) / sizeof(k0[0]);
  for (size_t i = 0; i < ns; ++i) {
    if ((size_t)shapes[i][0] * shapes[i][1] < n) continue;
    const KernelEntry *k = find_npdu(d3 ? 3 : 0, shapes[i][0], shapes[i][1]);
    if (k && plan_resident_shape(c, n, k, p, PlanKind::Npdu) == FPSX_OK) return FPSX_OK;
  }
  // Several workgroups: a pass costs the reduction chain, the window's slots
  // and a lane scan over PPT slots, plus an exchange growing with G.  3-D
  // clouds whose grid fits 2..64 workgroups take the paired form (one
  // exchange per two picks).  Its round costs about 0.06 us per slot and
  // barely depends on G (100 K points: 83.9 / 90.2 / 102.3 ms at 4 / 8 / 16
  // points per thread), so it takes the smallest share that fits 64
  // workgroups.
  double best = 1e30;
  Plan bp;
  if (d3 && !(std::getenv("FPSX_NO_PAIR") && std::getenv("FPSX_NO_PAIR")[0] == '1')) {
    for (int ppt : {2, 4, 8, 16}) {
      const KernelEntry *k = find_npdu_pair(1024, ppt);
      Plan q;
      if (!k || plan_resident_shape(c, n, k, q, PlanKind::Npdu) != FPSX_OK || q.grid < 2 ||
          q.grid > kNpduPairMaxGrid)
        continue;
      p = q;
      return FPSX_OK;
    }
  }
  for (int ppt : {4, 8, 16, 32}) {
    const KernelEntry *k = find_npdu(d3 ? 3 : 0, 1024, ppt);
    Plan q;
    if (!k || plan_resident_shape(c, n, k, q, PlanKind::Npdu) != FPSX_OK) continue;
    const double us = 0.012 * ppt + 0.0125 * (double)q.grid;
    if (us < best) {
      best = us;
      bp = q;
    }
  }
  if (bp.k) {
    p = bp;
    return FPSX_OK;
  }
  if (d3) {
    const KernelEntry *k = find_npdu(0, 1024, 32);
    if (k && plan_resident_shape(c, n, k, p, PlanKind::Npdu) == FPSX_OK) return FPSX_OK;
  }
  // Beyond the register file: min-distances in HBM, window blocks only.
  return plan_npdu_stream(c, n, dim, 0, p);
}

// Candidate plans the tuner times for a cloud of n points and m samples:
// the planner's own choice, every feasible resident shape of this dimension,
// for 3-D clouds every feasible pruned and paired shape and the streaming
// pruned plan, and the streaming kernel at two block sizes.
static std::vector<Plan> candidates(Ctx &c, size_t n, size_t dim, int tie, size_t m) {
  std::vector<Plan> out;
  Plan q;
  if (plan(c, n, dim, tie, m, q) == FPSX_OK) out.push_back(q);
  size_t cnt;
  const KernelEntry *t = resident_table(&cnt);
  for (size_t i = 0; i < cnt; ++i) {
    if (t[i].dim != (int)dim || t[i].tie != tie) continue;
    if (plan_resident_shape(c, n, &t[i], q) == FPSX_OK) out.push_back(q);
  }
  if (dim == 3) {
    t = pruned_table(&cnt);
    for (size_t i = 0; i < cnt; ++i)
      if (t[i].dim == 3 && t[i].tie == tie &&
          plan_resident_shape(c, n, &t[i], q, PlanKind::Pruned) == FPSX_OK)
        out.push_back(q);
    t = pair_table(&cnt);
    for (size_t i = 0; i < cnt; ++i)
      if (t[i].dim == 3 && t[i].tie == tie &&
          plan_resident_shape(c, n, &t[i], q, PlanKind::Pair) == FPSX_OK && q.grid >= 2 &&
          q.grid <= kPairMaxGrid)
        out.push_back(q);
    t = solo_table(&cnt);
    for (size_t i = 0; i < cnt; ++i)
      if (t[i].dim == 3 && t[i].tie == tie &&
          plan_resident_shape(c, n, &t[i], q, PlanKind::Pair) == FPSX_OK && q.grid == 1)
        out.push_back(q);
    for (int pw : {2, 4}) {
      const KernelEntry *k = pair_wide_kernel(1024, 16, tie, pw);
      if (k && plan_resident_shape(c, n, k, q, PlanKind::Pair) == FPSX_OK && q.grid > kPairMaxGrid &&
          q.grid <= (uint32_t)(64 * pw))
        out.push_back(q);
    }
    if (plan_stream_pruned(c, n, tie, 0, q) == FPSX_OK) out.push_back(q);
    if (plan_stream_pruned(c, n, tie, 0, q, true) == FPSX_OK) out.push_back(q);
  }
  for (int B : {256, 1024}) {
    if (plan_stream(c, n, dim, tie, B, 0, false, q) == FPSX_OK) out.push_back(q);
  }
  t_err.clear();  // infeasible candidates are skipped, not reported
  return out;
}

// ---------------------------------------------------------------------------
// Launch on device-resident data.  Caller holds c.mu.
// ---------------------------------------------------------------------------
// Kernel arguments common to every FPS kernel; clears the exchange area and
// the status word on `st`.
static int common_params(Ctx &c, const float *d_pts, size_t sp, size_t ld, size_t n, size_t dim,
                         size_t m, const uint64_t *d_starts, size_t n_starts, uint64_t *d_out,
                         float *d_out_pts, float init, uint32_t chunk, hipStream_t st,
                         FpsParams &a) {
  // Exchange tuning knobs (read per launch so a tuner can vary them).
  const char *rep_env = std::getenv("FPSX_XCHG_REPLICAS");
  const char *bo_env = std::getenv("FPSX_XCHG_BACKOFF");
  c.replicas = rep_env ? std::atoi(rep_env) : kDefaultReplicas;
  if (c.replicas < 1) c.replicas = 1;
  if (c.replicas > kMaxReplicas) c.replicas = kMaxReplicas;
  c.backoff = bo_env ? std::atoi(bo_env) : 0;
  if (c.backoff < 0) c.backoff = 0;
  const uint32_t reps = (uint32_t)c.replicas;
  const size_t xwords = xchg_words(reps);
  HIP_TRY(c.xchg.reserve(xwords * sizeof(uint64_t)));
  HIP_TRY(c.status.reserve(sizeof(uint64_t)));
  HIP_TRY(launch_fill_u64((uint64_t *)c.xchg.p, xwords, 0, st));
  HIP_TRY(launch_fill_u64((uint64_t *)c.status.p, 1, 0, st));
  std::memset(&a, 0, sizeof(a));
  a.pts = d_pts;
  a.ld = ld;
  a.sp = sp;
  a.out_pts = d_out_pts;
  a.n = (uint32_t)n;
  a.m = (uint32_t)m;
  a.starts = d_starts;
  a.n_starts = (uint32_t)n_starts;
  a.dim = (uint32_t)dim;
  a.out = d_out;
  a.xchg = (uint64_t *)c.xchg.p;
  a.status = (uint32_t *)c.status.p;
  a.init = init;
  a.spin_ticks = c.spin_ticks;
  a.replicas = reps;
  a.backoff = (uint32_t)c.backoff;
  a.chunk = chunk;
  return FPSX_OK;
}

// Enqueue the planned kernel with its argument block (timed when asked).
static int dispatch(Ctx &c, const Plan &p, void *arg, hipStream_t st) {
  void *args[] = {arg};
  std::pair<hipEvent_t, hipEvent_t> *ev = nullptr;
  if (c.timing) {
    if (c.events_used == c.events.size()) {
      std::pair<hipEvent_t, hipEvent_t> e2;
      HIP_TRY(hipEventCreate(&e2.first));
      HIP_TRY(hipEventCreate(&e2.second));
      c.events.push_back(e2);
    }
    ev = &c.events[c.events_used++];
    HIP_TRY(hipEventRecord(ev->first, st));
  }
  // The planner bounds the grid by the co-resident capacity and every
  // exchange wait is bounded, so a plain launch is used: were some workgroup
  // not resident, the kernel would report FPSX_ETIMEOUT rather than hang.
  // FPSX_COOPERATIVE=1 asks the runtime to validate co-residency instead
  // (same speed; some profilers crash on cooperative dispatches).
  const char *coop_env = std::getenv("FPSX_COOPERATIVE");
  const bool coop = coop_env && coop_env[0] == '1';
  if (p.grid > 1 && coop) {
    HIP_TRY(hipLaunchCooperativeKernel(p.k->fn, dim3(p.grid), dim3(p.block), args,
                                       (unsigned)p.shmem, st));
  } else {
    HIP_TRY(hipLaunchKernel(p.k->fn, dim3(p.grid), dim3(p.block), args, p.shmem, st));
  }
  if (ev) HIP_TRY(hipEventRecord(ev->second, st));
  return FPSX_OK;
}

int launch_fps(Ctx &c, const Plan &p, const float *d_pts, size_t sp, size_t ld, size_t n,
               size_t dim, size_t m, const uint64_t *d_starts, size_t n_starts, uint64_t *d_out,
               float *d_out_pts, float init, hipStream_t st) {
  if (p.kind == PlanKind::Npdu) return fail(FPSX_EINVAL, "launch_fps: windowed-update plan");
  int rc = scratch_acquire(c, st);
  if (rc) return rc;
  if (p.kind == PlanKind::Stream && sp != 1) {
    // The streaming kernel re-reads every point on every pass: give it SoA.
    HIP_TRY(c.pts.reserve(n * dim * sizeof(float)));
    HIP_TRY(launch_strided_to_soa(d_pts, sp, ld, (float *)c.pts.p, (uint32_t)n, (uint32_t)dim, n, st));
    d_pts = (const float *)c.pts.p;
    sp = 1;
    ld = n;
  }
  FpsParams a;
  rc = common_params(c, d_pts, sp, ld, n, dim, m, d_starts, n_starts, d_out, d_out_pts, init,
                     p.chunk, st, a);
  if (rc) return rc;
  if (p.kind == PlanKind::Stream) {
    HIP_TRY(c.dmin.reserve(n * sizeof(float)));
    a.dmin = (float *)c.dmin.p;
    HIP_TRY(launch_fill_f32(a.dmin, n, init, st));
  }
  PrunedParams pa;
  void *arg = &a;
  const bool streamed = p.kind == PlanKind::StreamPruned || p.kind == PlanKind::StreamPair;
  if (p.kind == PlanKind::Pruned || p.kind == PlanKind::Pair || streamed) {
    HIP_TRY(c.sorted.reserve(n * 3 * sizeof(float)));
    HIP_TRY(c.perm.reserve(n * sizeof(uint32_t)));
    if (!c.spatial) c.spatial = new_spatial_scratch();
    const uint32_t block = streamed ? (uint32_t)kSBlock : (uint32_t)(64 * p.k->ppt);
    rc = spatial_order(d_pts, sp, ld, (uint32_t)n, block, (float *)c.sorted.p,
                       (uint32_t *)c.perm.p, *c.spatial, st);
    if (rc) return rc;
    if (streamed) {
      // Min-distances by position; the first pass writes every one.
      HIP_TRY(c.dmin.reserve(n * sizeof(float)));
      a.dmin = (float *)c.dmin.p;
    }
    pa.fps = a;
    pa.sorted = (const float *)c.sorted.p;
    pa.sorted_ld = n;
    pa.perm = (const uint32_t *)c.perm.p;
    arg = &pa;
  }
  rc = dispatch(c, p, arg, st);
  if (rc) return rc;
  c.last_plan = p;
  c.last_m = m;
  c.last_starts = n_starts;
  return scratch_release(c, st);
}

// Enqueue one windowed-update run (single start d_start[0]).  Caller holds c.mu.
int launch_npdu(Ctx &c, const Plan &p, const float *d_pts, size_t sp, size_t ld, size_t n,
                size_t dim, size_t m, size_t window, const uint64_t *d_start, uint64_t *d_out,
                float *d_out_pts, hipStream_t st) {
  if (p.kind != PlanKind::Npdu) return fail(FPSX_EINVAL, "launch_npdu: not a windowed-update plan");
  int rc = scratch_acquire(c, st);
  if (rc) return rc;
  NpduParams a;
  rc = common_params(c, d_pts, sp, ld, n, dim, m, d_start, 1, d_out, d_out_pts, INFINITY,
                         p.chunk, st, a.fps);
  if (rc) return rc;
  // Any window of 2n or more covers the whole cloud; the clamp keeps the
  // kernel's index arithmetic far from overflow.
  a.window = std::min<uint64_t>((uint64_t)window, (uint64_t)4 * n + 4);
  if (p.k->ppt == 0) {
    // Streaming form: min-distances by index; the start pass writes them all.
    HIP_TRY(c.dmin.reserve(n * sizeof(float)));
    a.fps.dmin = (float *)c.dmin.p;
  }
  rc = dispatch(c, p, &a, st);
  if (rc) return rc;
  c.last_plan = p;
  c.last_m = m;
  c.last_starts = 1;
  return scratch_release(c, st);
}

static bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

int scratch_acquire(Ctx &c, hipStream_t st) {
  if (!c.scratch_recorded || c.scratch_owner == st || capturing(st)) return FPSX_OK;
  HIP_TRY(hipStreamWaitEvent(st, c.scratch_ev, 0));
  return FPSX_OK;
}

int scratch_release(Ctx &c, hipStream_t st) {
  if (capturing(st)) return FPSX_OK;
  if (!c.scratch_ev) HIP_TRY(hipEventCreateWithFlags(&c.scratch_ev, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(c.scratch_ev, st));
  c.scratch_owner = st;
  c.scratch_recorded = true;
  return FPSX_OK;
}

int status_code(uint32_t st) {
  if (st & kStatusTimeout)
    return fail(FPSX_ETIMEOUT, "cross-workgroup exchange timed out (workgroups not co-resident?)");
  if (st & kStatusBadStart) return fail(FPSX_EBADSTART, "a start index is >= n_points");
  return FPSX_OK;
}

// Host-array entry: upload, plan, launch, read back.  layout: FPSX_LAYOUT_*.
int fps_host(const float *points, size_t n, size_t dim, int layout, size_t m,
             const uint64_t *starts, size_t n_starts, uint64_t *out, int tie, float init) {
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  Ctx &c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  Plan p;
  rc = plan(c, n, dim, tie, m, p);
  if (rc) return rc;
  hipStream_t st = c.stream;
  rc = scratch_acquire(c, st);
  if (rc) return rc;
  const size_t bytes = n * dim * sizeof(float);
  HIP_TRY(c.pts.reserve(bytes));
  HIP_TRY(c.starts.reserve(n_starts * sizeof(uint64_t)));
  HIP_TRY(c.out.reserve(m * sizeof(uint64_t)));
  if (layout == FPSX_LAYOUT_COLS) {
    HIP_TRY(hipMemcpyAsync(c.pts.p, points, bytes, hipMemcpyHostToDevice, st));
  } else {
    HIP_TRY(c.rows.reserve(bytes));
    HIP_TRY(hipMemcpyAsync(c.rows.p, points, bytes, hipMemcpyHostToDevice, st));
    HIP_TRY(launch_strided_to_soa((const float *)c.rows.p, dim, 1, (float *)c.pts.p, (uint32_t)n,
                                  (uint32_t)dim, n, st));
  }
  HIP_TRY(hipMemcpyAsync(c.starts.p, starts, n_starts * sizeof(uint64_t), hipMemcpyHostToDevice,
                         st));
  rc = launch_fps(c, p, (const float *)c.pts.p, 1, n, n, dim, m, (const uint64_t *)c.starts.p,
                  n_starts, (uint64_t *)c.out.p, nullptr, init, st);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(out, c.out.p, m * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(c.host_status, c.status.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return status_code(*c.host_status);
}

}  // namespace fpsx

using namespace fpsx;

FPSX_EXPORT const char *fpsx_last_error(void) { return t_err.c_str(); }

FPSX_EXPORT const char *fpsx_version(void) { return FPSX_VERSION_STRING; }

FPSX_EXPORT int fpsx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

FPSX_EXPORT int fpsx_fps_sampling(const float *points, size_t n_points, size_t dim, int layout,
                                  size_t n_samples, const uint64_t *start_idx, size_t n_start,
                                  uint64_t *out) {
  t_err.clear();
  if (dim == 0) return fail(FPSX_EBADDIM, "points must have at least one column");
  if (!points || !start_idx || !out) return fail(FPSX_EINVAL, "null pointer argument");
  if (layout != FPSX_LAYOUT_ROWS && layout != FPSX_LAYOUT_COLS)
    return fail(FPSX_EINVAL, "unknown layout");
  if (n_samples == 0 || n_samples > n_points)
    return fail(FPSX_EINVAL, "n_samples must be in [1, n_points]: n_samples=" +
                                 std::to_string(n_samples) + ", P=" + std::to_string(n_points));
  if (n_start == 0 || n_start > n_samples)
    return fail(FPSX_EINVAL, "the number of start indices must be in [1, n_samples]");
  if (n_points >= (size_t(1) << 31)) return fail(FPSX_ETOOBIG, "n_points must be < 2^31");
  for (size_t i = 0; i < n_start; ++i)
    if (start_idx[i] >= n_points)
      return fail(FPSX_EBADSTART, "start index " + std::to_string(start_idx[i]) +
                                      " >= n_points " + std::to_string(n_points));
  return fps_host(points, n_points, dim, layout, n_samples, start_idx, n_start, out,
                  kTieHighest, INFINITY);
}

FPSX_EXPORT int fpsx_fps_sampling_device_strided(const float *d_points, size_t stride_point,
                                                 size_t stride_dim, size_t n_points, size_t dim,
                                                 size_t n_samples, const uint64_t *d_start,
                                                 size_t n_start, uint64_t *d_out,
                                                 float *d_out_points, int tie, float init,
                                                 void *stream) {
  t_err.clear();
  if (dim == 0) return fail(FPSX_EBADDIM, "dim must be >= 1");
  if (!d_points || !d_start || !d_out) return fail(FPSX_EINVAL, "null pointer argument");
  if (stride_point == 0 || (dim > 1 && stride_dim == 0))
    return fail(FPSX_EINVAL, "zero stride");
  if (n_points == 0 || n_samples == 0) return fail(FPSX_EINVAL, "empty cloud or no samples");
  if (n_start == 0 || n_start > n_samples)
    return fail(FPSX_EINVAL, "the number of start indices must be in [1, n_samples]");
  if (tie != FPSX_TIE_HIGHEST && tie != FPSX_TIE_LOWEST) return fail(FPSX_EINVAL, "bad tie rule");
  if (n_points >= (size_t(1) << 31) || n_samples >= (size_t(1) << 32))
    return fail(FPSX_ETOOBIG, "n_points must be < 2^31");
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  Ctx &c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  Plan p;
  rc = plan(c, n_points, dim, tie, n_samples, p);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : c.stream;
  return launch_fps(c, p, d_points, stride_point, stride_dim, n_points, dim, n_samples, d_start,
                    n_start, d_out, d_out_points, init, st);
}

FPSX_EXPORT int fpsx_fps_sampling_device(const float *d_points, size_t ld, size_t n_points,
                                         size_t dim, size_t n_samples, const uint64_t *d_start,
                                         size_t n_start, uint64_t *d_out, int tie, float init,
                                         void *stream) {
  if (ld < n_points) {
    t_err.clear();
    return fail(FPSX_EINVAL, "ld must be >= n_points");
  }
  return fpsx_fps_sampling_device_strided(d_points, 1, ld, n_points, dim, n_samples, d_start,
                                          n_start, d_out, nullptr, tie, init, stream);
}

FPSX_EXPORT int fpsx_fps_npdu_sampling_device(const float *d_points, size_t stride_point,
                                              size_t stride_dim, size_t n_points, size_t dim,
                                              size_t n_samples, size_t w, const uint64_t *d_start,
                                              uint64_t *d_out, float *d_out_points, void *stream) {
  t_err.clear();
  if (dim == 0) return fail(FPSX_EBADDIM, "dim must be >= 1");
  if (!d_points || !d_start || !d_out) return fail(FPSX_EINVAL, "null pointer argument");
  if (stride_point == 0 || (dim > 1 && stride_dim == 0))
    return fail(FPSX_EINVAL, "zero stride");
  if (n_points == 0 || n_samples == 0) return fail(FPSX_EINVAL, "empty cloud or no samples");
  if (n_points >= (size_t(1) << 31) || n_samples >= (size_t(1) << 32))
    return fail(FPSX_ETOOBIG, "n_points must be < 2^31");
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  Ctx &c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  Plan p;
  rc = plan_npdu(c, n_points, dim, p);
  if (rc) return rc;
  hipStream_t st = stream ? (hipStream_t)stream : c.stream;
  return launch_npdu(c, p, d_points, stride_point, stride_dim, n_points, dim, n_samples, w,
                     d_start, d_out, d_out_points, st);
}

FPSX_EXPORT int fpsx_fps_npdu_sampling(const float *points, size_t n_points, size_t dim,
                                       size_t n_samples, size_t w, size_t start_idx,
                                       uint64_t *out) {
  t_err.clear();
  if (dim == 0) return fail(FPSX_EBADDIM, "points must have at least one column");
  if (!points || !out) return fail(FPSX_EINVAL, "null pointer argument");
  if (n_samples == 0 || n_samples > n_points)
    return fail(FPSX_EINVAL, "n_samples must be in [1, n_points]: n_samples=" +
                                 std::to_string(n_samples) + ", P=" + std::to_string(n_points));
  if (start_idx >= n_points) return fail(FPSX_EBADSTART, "start_idx out of range");
  if (n_points >= (size_t(1) << 31)) return fail(FPSX_ETOOBIG, "n_points must be < 2^31");
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  Ctx &c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  Plan p;
  rc = plan_npdu(c, n_points, dim, p);
  if (rc) return rc;
  hipStream_t st = c.stream;
  rc = scratch_acquire(c, st);
  if (rc) return rc;
  const size_t bytes = n_points * dim * sizeof(float);
  HIP_TRY(c.pts.reserve(bytes));
  HIP_TRY(c.starts.reserve(sizeof(uint64_t)));
  HIP_TRY(c.out.reserve(n_samples * sizeof(uint64_t)));
  // Row-major points as given: the 3-D kernel reads them once into
  // registers, the any-dimension kernel reads only the windows.
  HIP_TRY(hipMemcpyAsync(c.pts.p, points, bytes, hipMemcpyHostToDevice, st));
  const uint64_t s64 = start_idx;
  HIP_TRY(hipMemcpyAsync(c.starts.p, &s64, sizeof(uint64_t), hipMemcpyHostToDevice, st));
  rc = launch_npdu(c, p, (const float *)c.pts.p, dim, 1, n_points, dim, n_samples, w,
                   (const uint64_t *)c.starts.p, (uint64_t *)c.out.p, nullptr, st);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(out, c.out.p, n_samples * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(c.host_status, c.status.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return status_code(*c.host_status);
}

FPSX_EXPORT int fpsx_npdu_plan_info(size_t n_points, size_t dim, int *kernel_dim, int *block,
                                    int *ppt, int *grid) {
  t_err.clear();
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  Ctx &c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  Plan p;
  rc = plan_npdu(c, n_points, dim, p);
  if (rc) return rc;
  if (kernel_dim) *kernel_dim = p.k->dim;
  if (block) *block = (int)p.block;
  if (ppt) *ppt = p.k->ppt;
  if (grid) *grid = (int)p.grid;
  return FPSX_OK;
}

FPSX_EXPORT int fpsx_launch_exchanges(void *stream, uint64_t *exchanges) {
  t_err.clear();
  if (!exchanges) return fail(FPSX_EINVAL, "null pointer argument");
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  Ctx &c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  *exchanges = 0;
  if (!c.last_plan.k) return FPSX_OK;
  if (c.last_plan.grid < 2 && c.last_plan.kind != PlanKind::Pair) return FPSX_OK;
  size_t npdu_pairs = 0;
  const KernelEntry *npdu_pair = npdu_pair_table(&npdu_pairs);
  const bool paired_npdu = c.last_plan.kind == PlanKind::Npdu && c.last_plan.k >= npdu_pair &&
                           c.last_plan.k < npdu_pair + npdu_pairs;
  if (c.last_plan.kind != PlanKind::Pair && c.last_plan.kind != PlanKind::StreamPair &&
      !paired_npdu) {
    // One exchange per pick after the forced prefix.
    *exchanges = c.last_m > c.last_starts ? c.last_m - c.last_starts : 0;
    return FPSX_OK;
  }
  hipStream_t st = stream ? (hipStream_t)stream : c.stream;
  rc = scratch_acquire(c, st);
  if (rc) return rc;
  uint32_t words[2] = {0, 0};
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipMemcpy(words, c.status.p, sizeof(words), hipMemcpyDeviceToHost));
  *exchanges = words[1];
  return FPSX_OK;
}

FPSX_EXPORT void *fpsx_stream(void) {
  t_err.clear();
  Ctx *cp = nullptr;
  if (get_ctx(&cp) != FPSX_OK) return nullptr;
  return (void *)cp->stream;
}

FPSX_EXPORT int fpsx_device_status(void *stream) {
  t_err.clear();
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  Ctx &c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  hipStream_t st = stream ? (hipStream_t)stream : c.stream;
  if (!c.status.p) return FPSX_OK;
  rc = scratch_acquire(c, st);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(c.host_status, c.status.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return status_code(*c.host_status);
}

FPSX_EXPORT int fpsx_plan_info(size_t n_points, size_t dim, int tie, size_t n_samples, int *kind,
                               int *block, int *ppt, int *grid, int *chunk) {
  t_err.clear();
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  Ctx &c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  Plan p;
  rc = plan(c, n_points, dim, tie, n_samples, p);
  if (rc) return rc;
  if (kind) *kind = (int)p.kind;
  if (block) *block = (int)p.block;
  if (ppt) *ppt = p.k->ppt;
  if (grid) *grid = (int)p.grid;
  if (chunk) *chunk = (int)p.chunk;
  return FPSX_OK;
}

FPSX_EXPORT int fpsx_autotune(const float *d_points, size_t stride_point, size_t stride_dim,
                              size_t n_points, size_t dim, int tie, size_t passes, void *stream,
                              int *chosen_block, int *chosen_ppt, int *chosen_grid) {
  t_err.clear();
  if (!d_points || n_points == 0 || dim == 0) return fail(FPSX_EINVAL, "bad arguments");
  if (n_points >= (size_t(1) << 31)) return fail(FPSX_ETOOBIG, "n_points must be < 2^31");
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  Ctx &c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  hipStream_t st = stream ? (hipStream_t)stream : c.stream;
  const size_t m = std::max<size_t>(2, std::min(passes ? passes : 256, n_points));
  rc = scratch_acquire(c, st);
  if (rc) return rc;
  HIP_TRY(c.starts.reserve(sizeof(uint64_t)));
  HIP_TRY(c.out.reserve(m * sizeof(uint64_t)));
  HIP_TRY(launch_fill_u64((uint64_t *)c.starts.p, 1, 0, st));
  c.tuned.erase(std::make_tuple(n_points, dim, tie));  // the planner's own choice is a candidate
  std::vector<Plan> cands = candidates(c, n_points, dim, tie, m);
  if (cands.empty()) return fail(FPSX_EINVAL, "no feasible plan");
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  std::vector<double> times(cands.size(), 1e30);
  for (size_t ci = 0; ci < cands.size(); ++ci) {
    const Plan &q = cands[ci];
    float tmin = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      HIP_TRY(hipEventRecord(e0, st));
      rc = launch_fps(c, q, d_points, stride_point, stride_dim, n_points, dim, m,
                      (const uint64_t *)c.starts.p, 1, (uint64_t *)c.out.p, nullptr, INFINITY, st);
      if (rc) return rc;
      HIP_TRY(hipEventRecord(e1, st));
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      tmin = std::min(tmin, ms);
    }
    HIP_TRY(hipMemcpyAsync(c.host_status, c.status.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (*c.host_status != 0) continue;  // a candidate that failed is never chosen
    times[ci] = tmin;
  }
  (void)hipEventDestroy(e0);  // timing events only
  (void)hipEventDestroy(e1);
  // Within 3% of the fastest, prefer the fewest workgroups: the choice is
  // then stable from run to run, and fewer workgroups exchange less.
  const double best = *std::min_element(times.begin(), times.end());
  const Plan *bp = nullptr;
  for (size_t ci = 0; ci < cands.size(); ++ci) {
    if (times[ci] > best * 1.03) continue;
    if (!bp || cands[ci].grid < bp->grid) bp = &cands[ci];
  }
  if (!bp) return fail(FPSX_EHIP, "every candidate plan failed");
  c.tuned[std::make_tuple(n_points, dim, tie)] = *bp;
  if (chosen_block) *chosen_block = (int)bp->block;
  if (chosen_ppt) *chosen_ppt = bp->k->ppt;
  if (chosen_grid) *chosen_grid = (int)bp->grid;
  return FPSX_OK;
}

FPSX_EXPORT int fpsx_exchange_timeout(double seconds) {
  t_err.clear();
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(cp->mu);
  cp->spin_ticks = seconds > 0 ? std::max<uint64_t>(1, (uint64_t)((double)cp->clock_khz * 1000.0 * seconds))
                               : default_spin_ticks(cp->clock_khz);
  return FPSX_OK;
}

FPSX_EXPORT int fpsx_timing(int enable) {
  t_err.clear();
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(cp->mu);
  cp->timing = enable != 0;
  cp->events_used = 0;
  return FPSX_OK;
}

FPSX_EXPORT int fpsx_timing_read(double *total_ms, int *launches) {
  t_err.clear();
  Ctx *cp = nullptr;
  int rc = get_ctx(&cp);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(cp->mu);
  double sum = 0.0;
  for (size_t i = 0; i < cp->events_used; ++i) {
    HIP_TRY(hipEventSynchronize(cp->events[i].second));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, cp->events[i].first, cp->events[i].second));
    sum += ms;
  }
  if (total_ms) *total_ms = sum;
  if (launches) *launches = (int)cp->events_used;
  cp->events_used = 0;
  return FPSX_OK;
}

FPSX_EXPORT long fpsx_kdline_permutation(const float *raw_data, size_t n_points, size_t dim,
                                         size_t height, uint64_t *perm, uint64_t *leaf_lo,
                                         uint64_t *leaf_hi) {
  t_err.clear();
  if (dim == 0) return -(long)fail(FPSX_EBADDIM, "dim must be >= 1");
  if (!raw_data || !perm || n_points == 0) return -(long)fail(FPSX_EINVAL, "bad arguments");
  if (n_points >= (size_t(1) << 31)) return -(long)fail(FPSX_ETOOBIG, "n_points must be < 2^31");
  std::vector<uint64_t> pv, lo, hi;
  if (kdline_permutation(raw_data, n_points, dim, height, pv, nullptr, &lo, &hi) != 0)
    return -(long)fail(FPSX_EINVAL, "kdline permutation failed");
  std::memcpy(perm, pv.data(), n_points * sizeof(uint64_t));
  for (size_t k = 0; k < lo.size(); ++k) {
    if (leaf_lo) leaf_lo[k] = lo[k];
    if (leaf_hi) leaf_hi[k] = hi[k];
  }
  return (long)lo.size();
}

FPSX_EXPORT int bucket_fps_kdline(const float *raw_data, size_t n_points, size_t dim,
                                  size_t n_samples, size_t start_idx, size_t height,
                                  size_t *sampled_point_indices) {
  t_err.clear();
  if (dim == 0 || dim > FPSX_BUCKET_MAX_DIM)
    return fail(FPSX_EBADDIM, "dim must be in [1, 8]");
  if (start_idx >= n_points) return fail(FPSX_EBADSTART, "start_idx must be < n_points");
  if (!raw_data || !sampled_point_indices) return fail(FPSX_EINVAL, "null pointer argument");
  if (n_samples == 0) return fail(FPSX_EINVAL, "n_samples must be >= 1");
  if (n_points >= (size_t(1) << 31)) return fail(FPSX_ETOOBIG, "n_points must be < 2^31");
  std::vector<uint64_t> perm;
  std::vector<float> cols;
  if (kdline_permutation(raw_data, n_points, dim, height, perm, &cols, nullptr, nullptr) != 0)
    return fail(FPSX_EINVAL, "kdline permutation failed");
  std::vector<uint64_t> pos(n_samples);
  uint64_t s0 = start_idx;
  int rc = fps_host(cols.data(), n_points, dim, FPSX_LAYOUT_COLS, n_samples, &s0, 1, pos.data(),
                    kTieLowest, FLT_MAX);
  if (rc) return rc;
  for (size_t i = 0; i < n_samples; ++i) sampled_point_indices[i] = (size_t)perm[pos[i]];
  return FPSX_OK;
}

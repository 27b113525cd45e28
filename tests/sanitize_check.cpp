// sanitize_check.cpp -- host code under the compiler's sanitizers.
//
// Drives the library's host-side KD-line permutation (fpsample_amd/csrc/
// kdline.cpp) and every entry of the oracle's C restatement (oracle/
// fps_oracle.c) over random, lattice (many ties), duplicate and NaN/inf
// clouds, including sizes large enough for kdline's threaded top levels.  It
// is built with g++/gcc only (no device code) under AddressSanitizer +
// UndefinedBehaviorSanitizer and, separately, ThreadSanitizer; a sanitizer
// report aborts the run (halt_on_error), so a clean run prints "OK".
//
// The permutation is also compared with the oracle's (the same rule as
// src/_ext/KDTreeBase.h:84-207), and the bucket sampler with the exhaustive
// form, so a run that is clean but wrong fails as well.
#include <stdint.h>
#include <stdio.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "kdline.hpp"

extern "C" {
int oracle_fps(const float *pts, size_t n, size_t d, size_t m, const uint64_t *starts,
               size_t n_starts, uint64_t *out);
int oracle_fps_ex(const float *pts, size_t n, size_t d, size_t m, const uint64_t *starts,
                  size_t n_starts, int lowest, float init, uint64_t *out, float *picked_dist);
int oracle_npdu(const float *pts, size_t n, size_t d, size_t m, size_t w, size_t start,
                uint64_t *out);
int oracle_fps_f64(const float *pts, size_t n, size_t d, size_t m, size_t start, uint64_t *out,
                   double *picked_dist);
double oracle_coverage_f64(const float *pts, size_t n, size_t d, const uint64_t *picks, size_t k);
long oracle_kdline_perm(const float *pts, size_t n, size_t d, size_t h, uint64_t *perm,
                        uint64_t *leaf_lo, uint64_t *leaf_hi);
int oracle_kdline_quickfps(const float *pts, size_t n, size_t d, size_t m, size_t start, size_t h,
                           uint64_t *out);
int oracle_kdline_exhaustive(const float *pts, size_t n, size_t d, size_t m, size_t start,
                             size_t h, uint64_t *out);
}

namespace {

int g_fail = 0;

#define CHECK(cond, ...)                  \
  do {                                    \
    if (!(cond)) {                        \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);       \
      fprintf(stderr, "\n");              \
      ++g_fail;                           \
    }                                     \
  } while (0)

struct Rng {  // splitmix64
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  float unit() { return (float)((next() >> 40) * (1.0 / 16777216.0)); }  // [0, 1)
  size_t below(size_t k) { return (size_t)(next() % k); }
};

enum Kind { kUniform, kLattice, kDuplicate, kSpecial };

std::vector<float> cloud(size_t n, size_t d, Kind kind, uint64_t seed) {
  Rng r{seed};
  std::vector<float> p(n * d);
  for (size_t i = 0; i < n * d; ++i) {
    switch (kind) {
      case kUniform: p[i] = r.unit() * 2.0f - 1.0f; break;
      case kLattice: p[i] = (float)r.below(4); break;
      case kDuplicate: p[i] = 0.25f * (float)(i % d); break;
      case kSpecial: p[i] = r.unit(); break;
    }
  }
  if (kind == kSpecial && n > 8) {
    const float nan = std::numeric_limits<float>::quiet_NaN();
    const float inf = std::numeric_limits<float>::infinity();
    for (size_t k = 0; k < n / 997 + 3; ++k) p[r.below(n) * d + r.below(d)] = nan;
    p[r.below(n) * d] = inf;
    p[r.below(n) * d] = -inf;
    p[r.below(n) * d] = -0.0f;
  }
  return p;
}

// kdline's permutation against the oracle's, plus its permuted columns.
void check_kdline(size_t n, size_t d, size_t h, Kind kind, uint64_t seed) {
  const std::vector<float> p = cloud(n, d, kind, seed);
  std::vector<uint64_t> perm, lo, hi;
  std::vector<float> cols;
  int rc = fpsx::kdline_permutation(p.data(), n, d, h, perm, &cols, &lo, &hi);
  CHECK(rc == 0, "kdline_permutation(n=%zu d=%zu h=%zu) -> %d", n, d, h, rc);
  if (rc) return;
  std::vector<uint64_t> operm(n), olo(n + 1), ohi(n + 1);
  long nl = oracle_kdline_perm(p.data(), n, d, h, operm.data(), olo.data(), ohi.data());
  CHECK(nl > 0, "oracle_kdline_perm(n=%zu d=%zu h=%zu) -> %ld", n, d, h, nl);
  if (nl <= 0) return;
  CHECK(perm == operm, "permutation differs (n=%zu d=%zu h=%zu kind=%d)", n, d, h, (int)kind);
  CHECK(lo.size() == (size_t)nl && hi.size() == (size_t)nl, "bucket count %zu vs %ld", lo.size(), nl);
  for (size_t b = 0; b < lo.size() && b < (size_t)nl; ++b)
    CHECK(lo[b] == olo[b] && hi[b] == ohi[b], "bucket %zu differs", b);
  CHECK(cols.size() == n * d, "permuted columns have %zu values", cols.size());
  for (size_t pos = 0; pos < n && cols.size() == n * d; ++pos)
    for (size_t j = 0; j < d; ++j) {
      const float a = cols[j * n + pos], b = p[perm[pos] * d + j];
      if (std::memcmp(&a, &b, sizeof(float)) != 0) {
        CHECK(false, "permuted column %zu at %zu differs", j, pos);
        return;
      }
    }
  // Error path: no points.
  std::vector<uint64_t> none;
  CHECK(fpsx::kdline_permutation(p.data(), 0, d, h, none, nullptr, nullptr, nullptr) == -1,
        "n=0 accepted");
}

void check_oracle(size_t n, size_t d, size_t m, Kind kind, uint64_t seed) {
  const std::vector<float> p = cloud(n, d, kind, seed);
  const size_t cap = m < 2 ? 2 : m;  // room for the two-pick rejection calls
  std::vector<uint64_t> a(cap), b(cap), c(cap);
  std::vector<float> dist(m);
  const uint64_t starts[3] = {0, n / 2, n - 1};
  const size_t ns = m >= 3 ? 3 : 1;
  CHECK(oracle_fps(p.data(), n, d, m, starts, ns, a.data()) == 0, "oracle_fps");
  CHECK(oracle_fps_ex(p.data(), n, d, m, starts, ns, 0, INFINITY, b.data(), dist.data()) == 0,
        "oracle_fps_ex");
  CHECK(a == b, "oracle_fps_ex(highest, inf) differs from oracle_fps");
  CHECK(oracle_fps_ex(p.data(), n, d, m, starts, 1, 1, 3.4e38f, c.data(), dist.data()) == 0,
        "oracle_fps_ex(lowest)");
  CHECK(oracle_npdu(p.data(), n, d, m, 5, 0, c.data()) == 0, "oracle_npdu");
  CHECK(oracle_npdu(p.data(), n, d, m, n, 0, b.data()) == 0, "oracle_npdu(full window)");
  std::vector<double> dd(m);
  CHECK(oracle_fps_f64(p.data(), n, d, m, n - 1, c.data(), dd.data()) == 0, "oracle_fps_f64");
  const double cov = oracle_coverage_f64(p.data(), n, d, c.data(), m);
  CHECK(cov >= 0.0 || std::isnan(cov), "coverage %g", cov);
  // Indices outside the cloud are rejected, not read.
  const uint64_t bad[2] = {0, n};
  CHECK(oracle_fps(p.data(), n, d, 2, bad, 2, a.data()) == -1, "start n accepted");
  CHECK(oracle_npdu(p.data(), n, d, 2, 5, n, a.data()) == -1, "npdu start n accepted");
  CHECK(oracle_fps_f64(p.data(), n, d, 2, n, a.data(), nullptr) == -1, "f64 start n accepted");
  CHECK(oracle_coverage_f64(p.data(), n, d, bad, 2) == -1.0, "coverage pick n accepted");
  CHECK(oracle_kdline_quickfps(p.data(), n, d, 2, n, 3, a.data()) == -1, "quickfps start n accepted");
  CHECK(oracle_kdline_exhaustive(p.data(), n, d, 2, n, 3, a.data()) == -1,
        "exhaustive start n accepted");
  CHECK(oracle_kdline_perm(p.data(), 0, d, 3, a.data(), nullptr, nullptr) == -1, "n=0 accepted");
  for (size_t h : {(size_t)1, (size_t)4, (size_t)9}) {
    if (kind == kSpecial) break;  // the bucket sampler's bounds assume finite coordinates
    CHECK(oracle_kdline_quickfps(p.data(), n, d, m, 0, h, b.data()) == 0, "quickfps h=%zu", h);
    CHECK(oracle_kdline_exhaustive(p.data(), n, d, m, 0, h, c.data()) == 0, "exhaustive h=%zu", h);
    CHECK(b == c, "quickfps differs from exhaustive (n=%zu d=%zu h=%zu kind=%d)", n, d, h, (int)kind);
  }
}

}  // namespace

// Planted mistakes, run only on request, that show the build's sanitizer is
// live: a read one past a heap array (ASan) and an unsynchronised counter
// shared by two threads (TSan).
int planted(const char *what) {
  if (std::strcmp(what, "--plant-overflow") == 0) {
    std::vector<float> v(3, 1.0f);
    volatile const float *q = v.data();
    return (int)q[3];
  }
  if (std::strcmp(what, "--plant-race") == 0) {
    int counter = 0;
    auto bump = [&counter] {
      for (int i = 0; i < 100000; ++i) ++*(volatile int *)&counter;
    };
    std::thread t1(bump), t2(bump);
    t1.join();
    t2.join();
    return counter == 0;
  }
  return -1;
}

int main(int argc, char **argv) {
  if (argc > 1 && std::strncmp(argv[1], "--plant", 7) == 0) return planted(argv[1]);
  const bool large = !(argc > 1 && std::strcmp(argv[1], "--small") == 0);
  uint64_t seed = 1;
  for (Kind kind : {kUniform, kLattice, kDuplicate, kSpecial}) {
    for (size_t d : {(size_t)1, (size_t)2, (size_t)3, (size_t)4, (size_t)5, (size_t)8}) {
      check_kdline(1, d, 3, kind, ++seed);
      check_kdline(7, d, 0, kind, ++seed);
      check_kdline(2000, d, 6, kind, ++seed);
      check_kdline(2000, d, 13, kind, ++seed);
    }
    // Threaded top levels: subtrees of >= 2^15 points get their own thread.
    if (large) {
      check_kdline(300000, 3, 11, kind, ++seed);
      check_kdline(140000, 5, 8, kind, ++seed);
    }
  }
  for (Kind kind : {kUniform, kLattice, kDuplicate, kSpecial}) {
    check_oracle(1, 3, 1, kind, ++seed);
    check_oracle(64, 3, 64, kind, ++seed);
    check_oracle(3000, 3, 200, kind, ++seed);
    check_oracle(1500, 7, 100, kind, ++seed);
  }
  if (g_fail) {
    fprintf(stderr, "%d check(s) failed\n", g_fail);
    return 1;
  }
  printf("OK\n");
  return 0;
}

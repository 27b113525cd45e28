// kdline.cpp -- host-side KD-line permutation used by the bucket FPS entry.
//
// The reference's bucket sampler (src/wrapper.hpp:45-60) builds a KD tree of
// height h over the points, permuting them in place (src/_ext/KDTreeBase.h:
// 84-207), starts at the point that ends up at position start_idx
// (src/wrapper.hpp:54-55) and resolves ties to the lowest position.  Its
// picks depend on that permutation, so this file rebuilds the same order:
//   * split dimension = widest bounding-box span, first on equal spans
//     (src/_ext/KDTreeBase.h:160-179);
//   * split value = float-accumulated mean divided by the count converted to
//     float (src/_ext/KDTreeBase.h:151-158);
//   * two-pointer partition with the same swap sequence and the same fix-ups
//     for an empty side (src/_ext/KDTreeBase.h:123-149);
//   * children's boxes recomputed from their points (:181-207); a node is a
//     bucket at depth h or when it holds one point (src/_ext/KDLineTree.h:37-39).
// Arithmetic is float32 with no contraction (built with -ffp-contract=off),
// matching the reference's default x86-64 build.
//
// Layout and parallelism (the result does not depend on either):
//   * the points are kept as columns (one array per coordinate plus the
//     original ids), so the mean and the partition scan one contiguous array;
//     a swap moves the same elements the reference's row swap moves;
//   * once a node is split, its two subtrees touch disjoint position ranges,
//     so large subtrees near the root are built on their own threads; the
//     buckets are concatenated left subtree first, the reference's order;
//   * below the threaded levels the walk is iterative, so a degenerate split
//     sequence (duplicates: one point split off per level) cannot exhaust the
//     stack.
#include "kdline.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cstring>
#include <thread>
#include <utility>
#include <vector>

namespace fpsx {

namespace {

// Subtree size worth a thread: with 4 K, a 50 K-point cloud at h = 7 builds
// in 1.4 ms instead of 2.1 ms on 8 cores (100 K points: 1.9 against 2.3 ms);
// 4 K-point clouds stay on the calling thread.
constexpr size_t kThreadMinPoints = size_t(1) << 12;
constexpr size_t kThreadMaxDepth = 6;

using Buckets = std::vector<std::pair<size_t, size_t>>;

// D > 0: dimension known at compile time; D == 0: runtime dimension.
template <int D>
struct Builder {
  size_t n = 0, d = 0, height = 0;
  std::vector<float> cols;  // coordinate j of position p at cols[j * n + p]
  std::vector<uint64_t> id;
  std::atomic<int> spare_threads{0};

  size_t dims() const { return D > 0 ? (size_t)D : d; }
  float *col(size_t j) { return cols.data() + j * n; }
  const float *col(size_t j) const { return cols.data() + j * n; }

  void box(size_t lo, size_t hi, float *bmin, float *bmax) const {
    const size_t dd = dims();
    for (size_t j = 0; j < dd; ++j) {
      // Four independent folds; no NaN ever enters one (`v < m` is false for
      // NaN), so combining them gives the sequential result.
      const float *c = col(j);
      float m0 = FLT_MAX, m1 = FLT_MAX, m2 = FLT_MAX, m3 = FLT_MAX;
      float x0 = -FLT_MAX, x1 = -FLT_MAX, x2 = -FLT_MAX, x3 = -FLT_MAX;
      size_t i = lo;
      for (; i + 4 <= hi; i += 4) {
        const float v0 = c[i], v1 = c[i + 1], v2 = c[i + 2], v3 = c[i + 3];
        m0 = (v0 < m0) ? v0 : m0;
        m1 = (v1 < m1) ? v1 : m1;
        m2 = (v2 < m2) ? v2 : m2;
        m3 = (v3 < m3) ? v3 : m3;
        x0 = (x0 < v0) ? v0 : x0;
        x1 = (x1 < v1) ? v1 : x1;
        x2 = (x2 < v2) ? v2 : x2;
        x3 = (x3 < v3) ? v3 : x3;
      }
      for (; i < hi; ++i) {
        const float v = c[i];
        m0 = (v < m0) ? v : m0;
        x0 = (x0 < v) ? v : x0;
      }
      m0 = (m1 < m0) ? m1 : m0;
      m2 = (m3 < m2) ? m3 : m2;
      bmin[j] = (m2 < m0) ? m2 : m0;
      x0 = (x0 < x1) ? x1 : x0;
      x2 = (x2 < x3) ? x3 : x2;
      bmax[j] = (x0 < x2) ? x2 : x0;
    }
  }

  size_t widest(const float *bmin, const float *bmax) const {
    float span = 0.0f;
    size_t best = 0;
    for (size_t j = 0; j < dims(); ++j) {
      const float s = bmax[j] - bmin[j];
      if (s > span) {
        span = s;
        best = j;
      }
    }
    return best;
  }

  // Sequential float sum in position order, as the reference accumulates it.
  float mean(size_t dim, size_t lo, size_t hi) const {
    const float *c = col(dim);
    float acc = 0.0f;
    for (size_t i = lo; i < hi; ++i) acc = acc + c[i];
    return acc / static_cast<float>(hi - lo);
  }

  void swap_pos(size_t a, size_t b) {
    for (size_t j = 0; j < dims(); ++j) std::swap(col(j)[a], col(j)[b]);
    std::swap(id[a], id[b]);
  }

  // The reference's two-pointer partition (a scans up past c < split, b
  // scans down past c >= split, then they swap), computed without
  // data-dependent branches.  Positions in [a, b] are untouched until a and b
  // reach them, so a's k-th stop is up[k], the k-th position (ascending)
  // with !(c < split), and b's k-th stop is dn[k], the k-th position
  // (descending) with !(c >= split); a NaN stops both and is in both lists.
  // Swap k happens while up[k] <= dn[k]; equality (a NaN where a and b meet)
  // is a self-swap that ends the loop.  After K swaps a stops at
  // min(up[K], dn[K-1]), but never below up[K-1] + 1, where the last swap
  // left it (the self-swap case); with no swap at all, at up[0] or hi.
  size_t partition(size_t lo, size_t hi, size_t dim, float split, std::vector<uint32_t> &up,
                   std::vector<uint32_t> &dn) {
    const float *c = col(dim);
    const size_t m = hi - lo;
    if (up.size() < m + 1) up.resize(m + 1);
    if (dn.size() < m + 1) dn.resize(m + 1);
    size_t nu = 0, nd = 0;
    for (size_t i = lo; i < hi; ++i) {
      up[nu] = (uint32_t)i;
      nu += !(c[i] < split);
    }
    for (size_t i = hi; i-- > lo;) {
      dn[nd] = (uint32_t)i;
      nd += !(c[i] >= split);
    }
    size_t k = 0;
    const size_t kmax = std::min(nu, nd);
    while (k < kmax && up[k] <= dn[k]) {
      if (up[k] != dn[k]) swap_pos(up[k], dn[k]);
      ++k;
    }
    size_t a = (k == 0) ? hi : (size_t)dn[k - 1];
    if (k < nu && (size_t)up[k] < a) a = up[k];
    if (k > 0 && a < (size_t)up[k - 1] + 1) a = (size_t)up[k - 1] + 1;
    size_t left = a - lo;
    if (a == lo) left = 1;
    if (a == hi) left = m - 1;
    return left;
  }

  // Splits [lo, hi) (box given); returns the left size and the child boxes.
  // up / dn: the calling thread's partition scratch.
  size_t split(size_t lo, size_t hi, const float *bmin, const float *bmax, float *lmin,
               float *lmax, float *rmin, float *rmax, std::vector<uint32_t> &up,
               std::vector<uint32_t> &dn) {
    const size_t dim = widest(bmin, bmax);
    const float s = mean(dim, lo, hi);
    const size_t left = partition(lo, hi, dim, s, up, dn);
    // The reference computes the left box, builds the left subtree, then the
    // right box; the left subtree only permutes positions inside its own
    // range, so computing both boxes now gives the same values.
    box(lo, lo + left, lmin, lmax);
    box(lo + left, hi, rmin, rmax);
    return left;
  }

  // Iterative depth-first walk, left child first.
  void walk(size_t lo0, size_t hi0, size_t depth0, const float *bmin0, const float *bmax0,
            Buckets &out) {
    struct Node {
      size_t lo, hi, depth;
      std::vector<float> bmin, bmax;
    };
    const size_t dd = dims();
    std::vector<uint32_t> up, dn;
    std::vector<Node> stack;
    stack.push_back(Node{lo0, hi0, depth0, std::vector<float>(bmin0, bmin0 + dd),
                         std::vector<float>(bmax0, bmax0 + dd)});
    while (!stack.empty()) {
      Node nd = std::move(stack.back());
      stack.pop_back();
      if (nd.depth == height || nd.hi - nd.lo == 1) {
        out.emplace_back(nd.lo, nd.hi);
        continue;
      }
      Node l{nd.lo, 0, nd.depth + 1, std::vector<float>(dd), std::vector<float>(dd)};
      Node r{0, nd.hi, nd.depth + 1, std::vector<float>(dd), std::vector<float>(dd)};
      const size_t left = split(nd.lo, nd.hi, nd.bmin.data(), nd.bmax.data(), l.bmin.data(),
                                l.bmax.data(), r.bmin.data(), r.bmax.data(), up, dn);
      l.hi = nd.lo + left;
      r.lo = nd.lo + left;
      stack.push_back(std::move(r));
      stack.push_back(std::move(l));
    }
  }

  // Top levels: split here, hand a large right subtree to another thread.
  void build(size_t lo, size_t hi, size_t depth, const float *bmin, const float *bmax,
             Buckets &out) {
    if (depth >= kThreadMaxDepth || hi - lo < 2 * kThreadMinPoints || depth == height ||
        hi - lo == 1) {
      walk(lo, hi, depth, bmin, bmax, out);
      return;
    }
    const size_t dd = dims();
    std::vector<float> lmin(dd), lmax(dd), rmin(dd), rmax(dd);
    size_t left;
    {
      std::vector<uint32_t> up, dn;
      left = split(lo, hi, bmin, bmax, lmin.data(), lmax.data(), rmin.data(), rmax.data(), up, dn);
    }
    const size_t mid = lo + left;
    Buckets right;
    std::thread helper;
    bool threaded = false;
    if (hi - mid >= kThreadMinPoints) {
      if (spare_threads.fetch_sub(1) > 0) {
        try {
          helper = std::thread([&] { build(mid, hi, depth + 1, rmin.data(), rmax.data(), right); });
          threaded = true;
        } catch (...) {
          // no thread available: build the right subtree here
        }
      }
      if (!threaded) spare_threads.fetch_add(1);
    }
    build(lo, mid, depth + 1, lmin.data(), lmax.data(), out);
    if (threaded) {
      helper.join();
      spare_threads.fetch_add(1);
    } else {
      build(mid, hi, depth + 1, rmin.data(), rmax.data(), right);
    }
    out.insert(out.end(), right.begin(), right.end());
  }

  void run(const float *rows, size_t n_, size_t d_, size_t height_, Buckets &out) {
    n = n_;
    d = d_;
    height = height_;
    const size_t dd = dims();
    cols.resize(n * dd);
    id.resize(n);
    for (size_t i = 0; i < n; ++i) {
      id[i] = i;
      for (size_t j = 0; j < dd; ++j) cols[j * n + i] = rows[i * dd + j];
    }
    const unsigned hw = std::thread::hardware_concurrency();
    spare_threads = (int)std::min<unsigned>(hw > 1 ? hw - 1 : 0, 15);
    std::vector<float> bmin(dd), bmax(dd);
    box(0, n, bmin.data(), bmax.data());
    build(0, n, 0, bmin.data(), bmax.data(), out);
  }
};

template <int D>
int permute(const float *rows, size_t n, size_t d, size_t height, std::vector<uint64_t> &perm,
            std::vector<float> *permuted_cols, std::vector<uint64_t> *bucket_lo,
            std::vector<uint64_t> *bucket_hi) {
  Builder<D> b;
  Buckets buckets;
  b.run(rows, n, d, height, buckets);
  perm.swap(b.id);
  if (permuted_cols) permuted_cols->swap(b.cols);
  if (bucket_lo) {
    bucket_lo->clear();
    for (auto &x : buckets) bucket_lo->push_back(x.first);
  }
  if (bucket_hi) {
    bucket_hi->clear();
    for (auto &x : buckets) bucket_hi->push_back(x.second);
  }
  return 0;
}

}  // namespace

int kdline_permutation(const float *rows, size_t n, size_t d, size_t height,
                       std::vector<uint64_t> &perm, std::vector<float> *permuted_cols,
                       std::vector<uint64_t> *bucket_lo, std::vector<uint64_t> *bucket_hi) {
  if (n == 0 || d == 0) return -1;
  switch (d) {
    case 1: return permute<1>(rows, n, d, height, perm, permuted_cols, bucket_lo, bucket_hi);
    case 2: return permute<2>(rows, n, d, height, perm, permuted_cols, bucket_lo, bucket_hi);
    case 3: return permute<3>(rows, n, d, height, perm, permuted_cols, bucket_lo, bucket_hi);
    case 4: return permute<4>(rows, n, d, height, perm, permuted_cols, bucket_lo, bucket_hi);
    default: return permute<0>(rows, n, d, height, perm, permuted_cols, bucket_lo, bucket_hi);
  }
}

}  // namespace fpsx

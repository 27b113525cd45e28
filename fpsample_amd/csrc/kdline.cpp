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
//   * children's boxes recomputed from their points (:181-207), left subtree
//     first; a node is a bucket at depth h or when it holds one point
//     (src/_ext/KDLineTree.h:37-39).
// Arithmetic is float32 with no contraction (built with -ffp-contract=off),
// matching the reference's default x86-64 build.  The walk is iterative so a
// degenerate split sequence cannot exhaust the stack.
#include "kdline.hpp"

#include <cfloat>
#include <cstring>
#include <vector>

namespace fpsx {

namespace {

struct Work {
  std::vector<float> xyz;  // n x d, row-major, permuted in place
  std::vector<uint64_t> id;
  size_t d = 0;

  void swap_rows(size_t a, size_t b) {
    float *pa = &xyz[a * d];
    float *pb = &xyz[b * d];
    for (size_t j = 0; j < d; ++j) {
      float t = pa[j];
      pa[j] = pb[j];
      pb[j] = t;
    }
    uint64_t t = id[a];
    id[a] = id[b];
    id[b] = t;
  }

  void box(size_t lo, size_t hi, float *bmin, float *bmax) const {
    for (size_t j = 0; j < d; ++j) {
      bmin[j] = FLT_MAX;
      bmax[j] = -FLT_MAX;
    }
    for (size_t i = lo; i < hi; ++i) {
      const float *p = &xyz[i * d];
      for (size_t j = 0; j < d; ++j) {
        const float v = p[j];
        bmin[j] = (v < bmin[j]) ? v : bmin[j];
        bmax[j] = (bmax[j] < v) ? v : bmax[j];
      }
    }
  }

  size_t widest(const float *bmin, const float *bmax) const {
    float span = 0.0f;
    size_t best = 0;
    for (size_t j = 0; j < d; ++j) {
      const float s = bmax[j] - bmin[j];
      if (s > span) {
        span = s;
        best = j;
      }
    }
    return best;
  }

  float mean(size_t dim, size_t lo, size_t hi) const {
    float acc = 0.0f;
    for (size_t i = lo; i < hi; ++i) acc = acc + xyz[i * d + dim];
    return acc / static_cast<float>(hi - lo);
  }

  size_t partition(size_t lo, size_t hi, size_t dim, float split) {
    int64_t a = static_cast<int64_t>(lo);
    int64_t b = static_cast<int64_t>(hi) - 1;
    for (;;) {
      while (a <= b && xyz[static_cast<size_t>(a) * d + dim] < split) ++a;
      while (a <= b && xyz[static_cast<size_t>(b) * d + dim] >= split) --b;
      if (a > b) break;
      swap_rows(static_cast<size_t>(a), static_cast<size_t>(b));
      ++a;
      --b;
    }
    int64_t left = a - static_cast<int64_t>(lo);
    if (a == static_cast<int64_t>(lo)) left = 1;
    if (a == static_cast<int64_t>(hi)) left = static_cast<int64_t>(hi - lo) - 1;
    return static_cast<size_t>(left);
  }
};

struct Node {
  size_t lo, hi, depth;
  std::vector<float> bmin, bmax;
};

}  // namespace

int kdline_permutation(const float *rows, size_t n, size_t d, size_t height,
                       std::vector<uint64_t> &perm, std::vector<float> *permuted_rows,
                       std::vector<uint64_t> *bucket_lo, std::vector<uint64_t> *bucket_hi) {
  if (n == 0 || d == 0) return -1;
  Work w;
  w.d = d;
  w.xyz.assign(rows, rows + n * d);
  w.id.resize(n);
  for (size_t i = 0; i < n; ++i) w.id[i] = i;
  if (bucket_lo) bucket_lo->clear();
  if (bucket_hi) bucket_hi->clear();

  // Depth-first, left child first: the stack holds nodes still to visit; a
  // right child is pushed before its left sibling so the left one pops first.
  std::vector<Node> stack;
  Node root{0, n, 0, std::vector<float>(d), std::vector<float>(d)};
  w.box(0, n, root.bmin.data(), root.bmax.data());
  stack.push_back(std::move(root));
  while (!stack.empty()) {
    Node nd = std::move(stack.back());
    stack.pop_back();
    if (nd.depth == height || nd.hi - nd.lo == 1) {
      if (bucket_lo) bucket_lo->push_back(nd.lo);
      if (bucket_hi) bucket_hi->push_back(nd.hi);
      continue;
    }
    const size_t dim = w.widest(nd.bmin.data(), nd.bmax.data());
    const float split = w.mean(dim, nd.lo, nd.hi);
    const size_t left = w.partition(nd.lo, nd.hi, dim, split);
    Node l{nd.lo, nd.lo + left, nd.depth + 1, std::vector<float>(d), std::vector<float>(d)};
    Node r{nd.lo + left, nd.hi, nd.depth + 1, std::vector<float>(d), std::vector<float>(d)};
    // The reference computes the left box, builds the whole left subtree,
    // then computes the right box.  Building the left subtree only permutes
    // positions inside [lo, lo + left), so the right box is the same whether
    // it is computed before or after.
    w.box(l.lo, l.hi, l.bmin.data(), l.bmax.data());
    w.box(r.lo, r.hi, r.bmin.data(), r.bmax.data());
    stack.push_back(std::move(r));
    stack.push_back(std::move(l));
  }
  perm.swap(w.id);
  if (permuted_rows) permuted_rows->swap(w.xyz);
  return 0;
}

}  // namespace fpsx

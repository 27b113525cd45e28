/*
 * fps_oracle.c -- CPU restatement of the reference farthest-point-sampling
 * loops.  TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library as the checker.  Nothing in
 * the fpsample_amd package links or calls it.
 *
 * Every function below restates one reference routine and cites the
 * reference file:line it follows.  Build flags (oracle/Makefile): -O2, no
 * -march, -ffp-contract=off, so every float operation is a separately rounded
 * IEEE binary32 SSE operation, as in the reference's default build
 * (CMakeLists.txt:1-19 sets no -march and no fast-math).
 *
 * Pinning (how each function is checked, tests/test_oracle.py):
 *   oracle_fps            golden vectors written by the compiled reference
 *                         (tests/golden/), hand-derived known answers, brute
 *                         force on tiny inputs, prefix-resume property.
 *   oracle_fps_ex         = oracle_fps when lowest=0, init=+inf; the
 *                         lowest/FLT_MAX form is pinned through
 *                         oracle_kdline_* against the reference golden vectors.
 *   oracle_kdline_perm    golden permutation-dependent outputs of the
 *                         reference bucket_fps_kdline (first pick =
 *                         perm[start_idx]) and the permutation invariants
 *                         (it is a permutation; buckets follow the split).
 *   oracle_kdline_quickfps golden vectors of the reference bucket_fps_kdline.
 *   oracle_kdline_exhaustive equality with oracle_kdline_quickfps and the
 *                         golden vectors (the exactness argument of bucket
 *                         pruning, checked on tie-heavy lattices).
 *   oracle_fps_f64,       (accuracy yardstick, double precision) brute force of
 *   oracle_coverage_f64   the definitions on tiny clouds; the f64 run's pick
 *                         distances equal the coverage radius of its prefixes;
 *                         on integer lattices (every value exact in float and
 *                         double) the f64 picks equal oracle_fps's.
 *   oracle_npdu           golden vectors of the reference _fps_npdu_sampling;
 *                         with a window covering the cloud it must equal
 *                         oracle_fps_ex(lowest) (the full update case).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Vanilla FPS                                                               */
/* ------------------------------------------------------------------------ */

/*
 * Squared distance of row i to pivot row p, summed in dimension order from
 * 0.0f with separately rounded products: src/lib.cpp:215-219 (and :141-145).
 */
static inline float sqdist_row(const float *pts, size_t d, size_t i, size_t p)
{
    const float *a = pts + i * d;
    const float *b = pts + p * d;
    float acc = 0.0f;
    for (size_t j = 0; j < d; ++j) {
        float diff = a[j] - b[j];
        float sq = diff * diff;
        acc = acc + sq;
    }
    return acc;
}

/*
 * General FPS over a row-major n x d float32 array.
 *   starts[0..n_starts) are forced picks (n_starts >= 1): src/lib.cpp:111-186.
 *   With n_starts == 1 this is fps_sampling: src/lib.cpp:188-246.
 *   lowest == 0: argmax keeps the HIGHEST index among equal maxima, the
 *   reference's `dist_min[i] >= max_val` scan from max_val = -1
 *   (src/lib.cpp:223-230, :157-165).
 *   lowest != 0: keeps the LOWEST index (strict `>`), the tie rule of the
 *   bucket scan (src/_ext/KDLineTree.h:56-67, src/_ext/KDNode.h:95-99).
 *   init: initial min-distance, +inf in the reference (src/lib.cpp:206),
 *   FLT_MAX in the bucket variant (src/_ext/Point.h:61-65).
 *   picked_dist (may be NULL): min-distance of each pick at the time it was
 *   picked (out[0] gets init).
 * Returns 0, or -1 when memory cannot be allocated.
 */
ORACLE_API int oracle_fps_ex(const float *pts, size_t n, size_t d, size_t m,
                             const uint64_t *starts, size_t n_starts,
                             int lowest, float init, uint64_t *out,
                             float *picked_dist)
{
    if (m == 0)
        return 0;
    if (!pts || !out || !starts || n == 0 || d == 0 || n_starts == 0)
        return -1;
    for (size_t k = 0; k < n_starts && k < m; ++k)
        if (starts[k] >= n)
            return -1; /* a start index outside the cloud */
    float *dmin = (float *)malloc(n * sizeof(float));
    if (!dmin)
        return -1;
    for (size_t i = 0; i < n; ++i)
        dmin[i] = init;
    size_t last = (size_t)starts[0];
    size_t used = 1;
    out[0] = last;
    if (picked_dist)
        picked_dist[0] = init;
    for (size_t s = 1; s < m; ++s) {
        /* UPDATE against the previous pick (src/lib.cpp:214-221). */
        for (size_t i = 0; i < n; ++i) {
            float dist = sqdist_row(pts, d, i, last);
            if (dist < dmin[i])
                dmin[i] = dist;
        }
        if (used < n_starts) {
            /* Forced pick from the start list (src/lib.cpp:151-155). */
            last = (size_t)starts[used++];
            if (picked_dist)
                picked_dist[s] = dmin[last];
        } else {
            /* ARGMAX (src/lib.cpp:223-230). */
            size_t best = 0;
            float bestv = -1.0f;
            if (!lowest) {
                for (size_t i = 0; i < n; ++i)
                    if (dmin[i] >= bestv) {
                        bestv = dmin[i];
                        best = i;
                    }
            } else {
                bestv = -FLT_MAX;
                for (size_t i = 0; i < n; ++i)
                    if (dmin[i] > bestv) {
                        bestv = dmin[i];
                        best = i;
                    }
            }
            last = best;
            if (picked_dist)
                picked_dist[s] = bestv;
        }
        out[s] = last;
    }
    free(dmin);
    return 0;
}

/*
 * FPS with nearest-point-distance-updating (NPDU), src/lib.cpp:272-340: the
 * first pick updates every point; every later pick p updates only the window
 * of `w + 1` indices centred on p (src/lib.cpp:296-300: start = p - w/2,
 * end = p + w/2, shifted to stay inside [0, n-1]); the argmax is over all
 * points with a strict `>` (lowest index among equal maxima, src/lib.cpp:310-315).
 */
ORACLE_API int oracle_npdu(const float *pts, size_t n, size_t d, size_t m, size_t w,
                           size_t start, uint64_t *out)
{
    if (m == 0)
        return 0;
    if (!pts || !out || n == 0 || d == 0 || start >= n)
        return -1;
    float *dmin = (float *)malloc(n * sizeof(float));
    if (!dmin)
        return -1;
    for (size_t i = 0; i < n; ++i)
        dmin[i] = INFINITY;
    for (size_t i = 0; i < n; ++i) {
        float dist = sqdist_row(pts, d, i, start);
        if (dist < dmin[i])
            dmin[i] = dist;
    }
    out[0] = start;
    size_t last = start;
    for (size_t s = 1; s < m; ++s) {
        int64_t hw = (int64_t)(w / 2);
        int64_t lo = (int64_t)last - hw;
        int64_t hi = (int64_t)last + hw;
        if (lo < 0) {
            hi -= lo;
            lo = 0;
        }
        if (hi >= (int64_t)n) {
            lo = lo - (hi - (int64_t)n + 1);
            if (lo < 0)
                lo = 0;
            hi = (int64_t)n - 1;
        }
        for (int64_t i = lo; i <= hi; ++i) {
            float dist = sqdist_row(pts, d, (size_t)i, last);
            if (dist < dmin[i])
                dmin[i] = dist;
        }
        size_t best = 0;
        float bestv = -1.0f;
        for (size_t i = 0; i < n; ++i)
            if (dmin[i] > bestv) {
                bestv = dmin[i];
                best = i;
            }
        last = best;
        out[s] = last;
    }
    free(dmin);
    return 0;
}

static inline double sqdist_f64(const float *pts, size_t d, size_t i, size_t p)
{
    const float *a = pts + i * d, *b = pts + p * d;
    if (d == 3) {
        double x = (double)a[0] - (double)b[0];
        double y = (double)a[1] - (double)b[1];
        double z = (double)a[2] - (double)b[2];
        return (x * x + y * y) + z * z;
    }
    double acc = 0.0;
    for (size_t j = 0; j < d; ++j) {
        double diff = (double)a[j] - (double)b[j];
        acc += diff * diff;
    }
    return acc;
}

/*
 * Accuracy yardstick, not a restatement of the reference: the same greedy
 * loop as src/lib.cpp:211-238 (start, update with `<`, argmax with `>=`),
 * with every coordinate difference, square and sum in double.  Where the
 * float32 loop and this one agree on every pick, the float32 picks are the
 * exact farthest points; where they part, the coverage radius below measures
 * how much it matters.  picked_dist (optional) gets the squared distance of
 * each pick to the picks before it (+inf for pick 0).
 */
ORACLE_API int oracle_fps_f64(const float *pts, size_t n, size_t d, size_t m, size_t start,
                              uint64_t *out, double *picked_dist)
{
    if (m == 0)
        return 0;
    if (!pts || !out || n == 0 || d == 0 || start >= n)
        return -1;
    double *dmin = (double *)malloc(n * sizeof(double));
    if (!dmin)
        return -1;
    for (size_t i = 0; i < n; ++i)
        dmin[i] = INFINITY;
    size_t last = start;
    out[0] = last;
    if (picked_dist)
        picked_dist[0] = INFINITY;
    for (size_t s = 1; s < m; ++s) {
        size_t best = 0;
        double bestv = -1.0;
        for (size_t i = 0; i < n; ++i) {
            double acc = sqdist_f64(pts, d, i, last);
            if (acc < dmin[i])
                dmin[i] = acc;
            if (dmin[i] >= bestv) {
                bestv = dmin[i];
                best = i;
            }
        }
        last = best;
        out[s] = last;
        if (picked_dist)
            picked_dist[s] = bestv;
    }
    free(dmin);
    return 0;
}

/*
 * Coverage radius of a pick set, in double: max over points of the Euclidean
 * distance to the nearest pick (the quantity farthest point sampling
 * greedily shrinks).  Returns -1.0 for a pick outside the cloud or on
 * allocation failure.
 */
ORACLE_API double oracle_coverage_f64(const float *pts, size_t n, size_t d, const uint64_t *picks,
                                      size_t k)
{
    if (!pts || n == 0 || d == 0 || (k && !picks))
        return -1.0;
    for (size_t s = 0; s < k; ++s)
        if (picks[s] >= n)
            return -1.0; /* a pick outside the cloud */
    double *dmin = (double *)malloc(n * sizeof(double));
    if (!dmin)
        return -1.0;
    for (size_t i = 0; i < n; ++i)
        dmin[i] = INFINITY;
    for (size_t s = 0; s < k; ++s) {
        const size_t p = (size_t)picks[s];
        for (size_t i = 0; i < n; ++i) {
            double acc = sqdist_f64(pts, d, i, p);
            if (acc < dmin[i])
                dmin[i] = acc;
        }
    }
    double r = 0.0;
    for (size_t i = 0; i < n; ++i)
        if (dmin[i] > r)
            r = dmin[i];
    free(dmin);
    return sqrt(r);
}

/* fps_sampling / fps_sampling_multi_start_index semantics (src/lib.cpp:111-246). */
ORACLE_API int oracle_fps(const float *pts, size_t n, size_t d, size_t m,
                          const uint64_t *starts, size_t n_starts, uint64_t *out)
{
    return oracle_fps_ex(pts, n, d, m, starts, n_starts, 0, INFINITY, out, NULL);
}

/* ------------------------------------------------------------------------ */
/* KD-line tree (QuickFPS bucket FPS)                                        */
/* ------------------------------------------------------------------------ */

/*
 * Working set of the KD build: coordinates kept per point (row-major, d
 * floats) together with the original id, permuted in place exactly as the
 * reference swaps whole Point objects (src/_ext/KDTreeBase.h:137).
 */
typedef struct {
    float *xyz; /* n x d, permuted */
    uint64_t *id;
    size_t d;
    size_t h;
    /* leaves in creation order: [lo, hi) ranges and bounding boxes */
    size_t n_leaves, cap_leaves;
    size_t *leaf_lo, *leaf_hi;
    float *leaf_bmin, *leaf_bmax; /* n_leaves x d */
} kd_build;

static void kd_swap(kd_build *kb, size_t a, size_t b)
{
    size_t d = kb->d;
    for (size_t j = 0; j < d; ++j) {
        float t = kb->xyz[a * d + j];
        kb->xyz[a * d + j] = kb->xyz[b * d + j];
        kb->xyz[b * d + j] = t;
    }
    uint64_t ti = kb->id[a];
    kb->id[a] = kb->id[b];
    kb->id[b] = ti;
}

/* Bounding box of positions [lo, hi): src/_ext/KDTreeBase.h:181-207
 * (min starts at FLT_MAX, max at -FLT_MAX; std::min/std::max keep the
 * running value when the comparison is false, so NaN never enters). */
static void kd_bbox(const kd_build *kb, size_t lo, size_t hi, float *bmin, float *bmax)
{
    size_t d = kb->d;
    for (size_t j = 0; j < d; ++j) {
        bmin[j] = FLT_MAX;
        bmax[j] = -FLT_MAX;
    }
    for (size_t i = lo; i < hi; ++i)
        for (size_t j = 0; j < d; ++j) {
            float v = kb->xyz[i * d + j];
            if (v < bmin[j])
                bmin[j] = v;
            if (bmax[j] < v)
                bmax[j] = v;
        }
}

/* Widest-span dimension, first one on equal spans: src/_ext/KDTreeBase.h:160-179. */
static size_t kd_split_dim(const float *bmin, const float *bmax, size_t d)
{
    float span = 0.0f;
    size_t best = 0;
    for (size_t j = 0; j < d; ++j) {
        float s = bmax[j] - bmin[j];
        if (s > span) {
            best = j;
            span = s;
        }
    }
    return best;
}

/* Split value = float-accumulated mean of the coordinate, divided by the
 * count converted to float: src/_ext/KDTreeBase.h:151-158 (the accumulator
 * is handed to the lambda as `const float&`, so every partial sum is a float). */
static float kd_mean(const kd_build *kb, size_t dim, size_t lo, size_t hi)
{
    float acc = 0.0f;
    for (size_t i = lo; i < hi; ++i)
        acc = acc + kb->xyz[i * kb->d + dim];
    return acc / (float)(hi - lo);
}

/* Two-pointer partition around split_val and the size fix-ups for empty
 * sides: src/_ext/KDTreeBase.h:123-149. Returns the left-side count. */
static size_t kd_partition(kd_build *kb, size_t lo, size_t hi, size_t dim, float split_val)
{
    int64_t a = (int64_t)lo;
    int64_t b = (int64_t)hi - 1;
    size_t d = kb->d;
    for (;;) {
        while (a <= b && kb->xyz[(size_t)a * d + dim] < split_val)
            ++a;
        while (a <= b && kb->xyz[(size_t)b * d + dim] >= split_val)
            --b;
        if (a > b)
            break;
        kd_swap(kb, (size_t)a, (size_t)b);
        ++a;
        --b;
    }
    int64_t left = a - (int64_t)lo;
    if (a == (int64_t)lo)
        left = 1;
    if (a == (int64_t)hi)
        left = (int64_t)(hi - lo) - 1;
    return (size_t)left;
}

static int kd_add_leaf(kd_build *kb, size_t lo, size_t hi, const float *bmin, const float *bmax)
{
    if (kb->n_leaves == kb->cap_leaves) {
        size_t nc = kb->cap_leaves ? kb->cap_leaves * 2 : 64;
        size_t *nlo = (size_t *)realloc(kb->leaf_lo, nc * sizeof(size_t));
        if (!nlo)
            return -1;
        kb->leaf_lo = nlo;
        size_t *nhi = (size_t *)realloc(kb->leaf_hi, nc * sizeof(size_t));
        if (!nhi)
            return -1;
        kb->leaf_hi = nhi;
        float *nbmin = (float *)realloc(kb->leaf_bmin, nc * kb->d * sizeof(float));
        if (!nbmin)
            return -1;
        kb->leaf_bmin = nbmin;
        float *nbmax = (float *)realloc(kb->leaf_bmax, nc * kb->d * sizeof(float));
        if (!nbmax)
            return -1;
        kb->leaf_bmax = nbmax;
        kb->cap_leaves = nc;
    }
    size_t k = kb->n_leaves++;
    kb->leaf_lo[k] = lo;
    kb->leaf_hi[k] = hi;
    memcpy(kb->leaf_bmin + k * kb->d, bmin, kb->d * sizeof(float));
    memcpy(kb->leaf_bmax + k * kb->d, bmax, kb->d * sizeof(float));
    return 0;
}

/* Recursive split: src/_ext/KDTreeBase.h:92-121; a node is a leaf at depth h
 * or when it holds one point (src/_ext/KDLineTree.h:37-39); leaves are
 * appended left subtree first (src/_ext/KDLineTree.h:87-92). */
static int kd_divide(kd_build *kb, size_t lo, size_t hi, const float *bmin,
                     const float *bmax, size_t depth)
{
    size_t d = kb->d;
    if (depth == kb->h || hi - lo == 1)
        return kd_add_leaf(kb, lo, hi, bmin, bmax);
    size_t dim = kd_split_dim(bmin, bmax, d);
    float split_val = kd_mean(kb, dim, lo, hi);
    size_t left = kd_partition(kb, lo, hi, dim, split_val);
    float *cb = (float *)malloc(2 * d * sizeof(float));
    if (!cb)
        return -1;
    kd_bbox(kb, lo, lo + left, cb, cb + d);
    int rc = kd_divide(kb, lo, lo + left, cb, cb + d, depth + 1);
    if (rc == 0) {
        kd_bbox(kb, lo + left, hi, cb, cb + d);
        rc = kd_divide(kb, lo + left, hi, cb, cb + d, depth + 1);
    }
    free(cb);
    return rc;
}

static void kd_free(kd_build *kb)
{
    free(kb->xyz);
    free(kb->id);
    free(kb->leaf_lo);
    free(kb->leaf_hi);
    free(kb->leaf_bmin);
    free(kb->leaf_bmax);
}

/* Build the KD-line tree over row-major pts: src/wrapper.hpp:49-54 and
 * src/_ext/KDTreeBase.h:84-90. */
static int kd_build_tree(kd_build *kb, const float *pts, size_t n, size_t d, size_t h)
{
    memset(kb, 0, sizeof(*kb));
    kb->d = d;
    kb->h = h;
    kb->xyz = (float *)malloc(n * d * sizeof(float));
    kb->id = (uint64_t *)malloc(n * sizeof(uint64_t));
    if (!kb->xyz || !kb->id)
        return -1;
    memcpy(kb->xyz, pts, n * d * sizeof(float));
    for (size_t i = 0; i < n; ++i)
        kb->id[i] = i;
    float *rb = (float *)malloc(2 * d * sizeof(float));
    if (!rb)
        return -1;
    kd_bbox(kb, 0, n, rb, rb + d);
    int rc = kd_divide(kb, 0, n, rb, rb + d, 0);
    free(rb);
    return rc;
}

/*
 * Permutation produced by the KD-line build: perm[pos] = original id of the
 * point at position pos after the build; leaf_lo/leaf_hi (optional, capacity
 * 2^h each) receive the bucket ranges.  Returns the number of buckets, or -1.
 */
ORACLE_API long oracle_kdline_perm(const float *pts, size_t n, size_t d, size_t h,
                                   uint64_t *perm, uint64_t *leaf_lo, uint64_t *leaf_hi)
{
    if (!pts || !perm || n == 0 || d == 0)
        return -1;
    kd_build kb;
    if (kd_build_tree(&kb, pts, n, d, h) != 0) {
        kd_free(&kb);
        return -1;
    }
    memcpy(perm, kb.id, n * sizeof(uint64_t));
    for (size_t k = 0; k < kb.n_leaves; ++k) {
        if (leaf_lo)
            leaf_lo[k] = kb.leaf_lo[k];
        if (leaf_hi)
            leaf_hi[k] = kb.leaf_hi[k];
    }
    long nl = (long)kb.n_leaves;
    kd_free(&kb);
    return nl;
}

/* Squared distance as the bucket code computes it: the recursion
 * sq(D-1) + (sq(D-2) + (... + (sq(0) + 0))) evaluates dimension 0 innermost,
 * so the partial sums run first dimension first; each square is exactly
 * rounded (float times the double value pow(x, 1) == x, rounded back to
 * float) (src/_ext/Point.h:48-53, src/_ext/utils.h:12-16). */
static inline float sqdist_bucket(const float *a, const float *b, size_t d)
{
    float acc = 0.0f;
    for (size_t j = 0; j < d; ++j) {
        float diff = a[j] - b[j];
        acc = (diff * diff) + acc;
    }
    return acc;
}

/* Lower bound of the squared distance from ref to a bucket's box:
 * src/_ext/KDNode.h:105-118. */
static inline float bucket_bound(const float *ref, const float *bmin, const float *bmax, size_t d)
{
    float acc = 0.0f;
    for (size_t j = 0; j < d; ++j) {
        float g = 0.0f;
        if (ref[j] > bmax[j])
            g = ref[j] - bmax[j];
        else if (ref[j] < bmin[j])
            g = bmin[j] - ref[j];
        acc = acc + g * g;
    }
    return acc;
}

/*
 * QuickFPS bucket sampling with lazy per-bucket updates, restated from
 * src/wrapper.hpp:45-60, src/_ext/KDLineTree.h:56-85 and
 * src/_ext/KDNode.h:84-166 (leaf paths only: the line tree only ever asks
 * leaves to update, src/_ext/KDLineTree.h:69-75).
 */
ORACLE_API int oracle_kdline_quickfps(const float *pts, size_t n, size_t d, size_t m,
                                      size_t start, size_t h, uint64_t *out)
{
    if (!pts || !out || n == 0 || d == 0 || start >= n)
        return -1;
    kd_build kb;
    if (kd_build_tree(&kb, pts, n, d, h) != 0) {
        kd_free(&kb);
        return -1;
    }
    size_t nb = kb.n_leaves;
    float *dis = (float *)malloc(n * sizeof(float));
    size_t *maxpos = (size_t *)malloc(nb * sizeof(size_t)); /* position of bucket max */
    float *maxval = (float *)malloc(nb * sizeof(float));    /* copy of its dis */
    size_t *dcount = (size_t *)calloc(nb, sizeof(size_t));  /* deferred refs per bucket */
    size_t *dcap = (size_t *)calloc(nb, sizeof(size_t));
    float **dref = (float **)calloc(nb, sizeof(float *));   /* deferred ref coords */
    float *ref = (float *)malloc(d * sizeof(float));
    int rc = 0;
    if (!dis || !maxpos || !maxval || !dcount || !dcap || !dref || !ref) {
        rc = -1;
        goto done;
    }
    for (size_t i = 0; i < n; ++i)
        dis[i] = FLT_MAX; /* src/_ext/Point.h:61-65 */

    /* init(points[start]): the point now at position `start` (wrapper.hpp:55). */
    memcpy(ref, kb.xyz + start * d, d * sizeof(float));
    out[0] = kb.id[start];
    for (size_t b = 0; b < nb; ++b) {
        float mx = -FLT_MAX;
        size_t mp = kb.leaf_lo[b];
        for (size_t i = kb.leaf_lo[b]; i < kb.leaf_hi[b]; ++i) {
            float dd = sqdist_bucket(kb.xyz + i * d, ref, d);
            if (dd < dis[i])
                dis[i] = dd;
            if (dis[i] > mx) {
                mx = dis[i];
                mp = i;
            }
        }
        maxpos[b] = mp;
        maxval[b] = mx;
    }
    for (size_t s = 1; s < m; ++s) {
        /* max_point(): first bucket with a strictly larger max (KDLineTree.h:56-67). */
        float best = -FLT_MAX;
        size_t bp = 0;
        for (size_t b = 0; b < nb; ++b)
            if (maxval[b] > best) {
                best = maxval[b];
                bp = maxpos[b];
            }
        out[s] = kb.id[bp];
        memcpy(ref, kb.xyz + bp * d, d * sizeof(float));
        /* update_distance(ref) for every bucket (KDNode.h:120-166). */
        for (size_t b = 0; b < nb; ++b) {
            float last = maxval[b];
            float cur = sqdist_bucket(kb.xyz + maxpos[b] * d, ref, d);
            if (cur > last) {
                float bound = bucket_bound(ref, kb.leaf_bmin + b * d, kb.leaf_bmax + b * d, d);
                if (bound < last) {
                    if (dcount[b] == dcap[b]) {
                        size_t nc = dcap[b] ? dcap[b] * 2 : 4;
                        float *nr = (float *)realloc(dref[b], nc * d * sizeof(float));
                        if (!nr) {
                            rc = -1;
                            goto done;
                        }
                        dref[b] = nr;
                        dcap[b] = nc;
                    }
                    memcpy(dref[b] + dcount[b] * d, ref, d * sizeof(float));
                    dcount[b]++;
                }
            } else {
                /* apply every deferred ref plus this one, rescanning the bucket */
                size_t total = dcount[b] + 1;
                float mx = -FLT_MAX;
                size_t mp = maxpos[b];
                for (size_t r = 0; r < total; ++r) {
                    const float *rr = (r < dcount[b]) ? dref[b] + r * d : ref;
                    mx = -FLT_MAX;
                    for (size_t i = kb.leaf_lo[b]; i < kb.leaf_hi[b]; ++i) {
                        float dd = sqdist_bucket(kb.xyz + i * d, rr, d);
                        if (dd < dis[i])
                            dis[i] = dd;
                        if (dis[i] > mx) {
                            mx = dis[i];
                            mp = i;
                        }
                    }
                }
                maxpos[b] = mp;
                maxval[b] = mx;
                dcount[b] = 0;
            }
        }
    }
done:
    if (dref)
        for (size_t b = 0; b < nb; ++b)
            free(dref[b]);
    free(dref);
    free(dcap);
    free(dcount);
    free(maxval);
    free(maxpos);
    free(dis);
    free(ref);
    kd_free(&kb);
    return rc;
}

/*
 * The same selection computed exhaustively: FPS over the KD-permuted order
 * with the lowest-position tie rule and FLT_MAX initial distances, starting
 * from position `start`, positions mapped back to original ids.  Equal to
 * oracle_kdline_quickfps because bucket pruning only skips updates that
 * cannot lower any distance (src/_ext/KDNode.h:127-130).
 */
ORACLE_API int oracle_kdline_exhaustive(const float *pts, size_t n, size_t d, size_t m,
                                        size_t start, size_t h, uint64_t *out)
{
    if (!pts || !out || n == 0 || d == 0 || start >= n)
        return -1;
    kd_build kb;
    if (kd_build_tree(&kb, pts, n, d, h) != 0) {
        kd_free(&kb);
        return -1;
    }
    uint64_t s0 = start;
    int rc = oracle_fps_ex(kb.xyz, n, d, m, &s0, 1, 1, FLT_MAX, out, NULL);
    if (rc == 0)
        for (size_t s = 0; s < m; ++s)
            out[s] = kb.id[out[s]];
    kd_free(&kb);
    return rc;
}

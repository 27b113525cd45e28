/*
 * fpsx.h -- C ABI of the MI355X farthest-point-sampling library (libfpsx.so).
 *
 * Plain pointers and sizes only.  Every call is synchronous with respect to
 * host memory unless its name ends in _device; those enqueue work on a HIP
 * stream and return immediately.  All entry points are thread-safe: calls on
 * one device are serialised by the library.
 *
 * Return codes follow the reference's bucket C API (src/wrapper.hpp:121-126):
 * 0 success, 1 bad dimension, 2 bad start index; the codes from 3 up are new.
 * fpsx_last_error() gives a message for the calling thread's last failure.
 */
#ifndef FPSX_H
#define FPSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FPSX_OK 0
#define FPSX_EBADDIM 1    /* dim == 0, or dim > FPSX_BUCKET_MAX_DIM for the bucket entry */
#define FPSX_EBADSTART 2  /* a start index >= n_points */
#define FPSX_EINVAL 3     /* other invalid argument (null pointer, n_samples, layout, ...) */
#define FPSX_ETOOBIG 4    /* n_points >= 2^31 or a buffer larger than the device allows */
#define FPSX_EHIP 5       /* HIP runtime error (no device, launch failure, ...) */
#define FPSX_ENOMEM 6     /* device or host allocation failed */
#define FPSX_ETIMEOUT 7   /* a cross-workgroup exchange did not complete (kernel gave up) */
#define FPSX_EBUDGET 8    /* serving: request refused by the HBM budget */
#define FPSX_ECANCELED 9  /* serving: the request was cancelled before it finished */
#define FPSX_EPENDING 10  /* serving: fpsx_server_wait_for timed out; the request still runs */
#define FPSX_EDEADLINE 11 /* serving: the request's deadline passed before it finished */

/* Point layouts accepted by the host entry points. */
#define FPSX_LAYOUT_ROWS 0 /* row-major n x dim (C order): point i at p[i*dim + j] */
#define FPSX_LAYOUT_COLS 1 /* column-major n x dim (Fortran order, SoA): p[j*n + i] */

/* Tie rule among equal maximal distances. */
#define FPSX_TIE_HIGHEST 0 /* highest index wins: src/lib.cpp:223-230 (`>=` scan) */
#define FPSX_TIE_LOWEST 1  /* lowest index wins: src/_ext/KDLineTree.h:56-67 (`>`) */

/* Largest dimension accepted by bucket_fps_kdline (src/wrapper.hpp:8-11). */
#define FPSX_BUCKET_MAX_DIM 8

/* Message describing the calling thread's last failure ("" if none). */
const char *fpsx_last_error(void);

/* Library version string. */
const char *fpsx_version(void);

/* Number of visible HIP devices (0 when none or the runtime is unavailable). */
int fpsx_device_count(void);

/*
 * Synchronise `stream` (NULL: the library's stream for the current device) and
 * report the status word of the last kernel enqueued by
 * fpsx_fps_sampling_device: FPSX_OK, FPSX_ETIMEOUT or FPSX_EBADSTART.
 */
int fpsx_device_status(void *stream);

/*
 * The library's own stream for the current device (a hipStream_t; NULL when
 * no device is usable): what a NULL `stream` argument means.  A caller
 * handing the library a tensor still being written on another stream orders
 * that work before this one (e.g. the `stream` argument of DLPack's
 * __dlpack__).  Calls on different streams are ordered by the library
 * wherever they share its scratch buffers.
 */
void *fpsx_stream(void);

/*
 * The launch the planner picks for a cloud of n_points x dim and n_samples
 * picks under `tie`: kind 0 = register-resident, 1 = HBM-streaming,
 * 2 = register-resident with spatial order and per-wave pruning,
 * 4 = HBM-streaming with spatial order and per-block pruning,
 * 5 = kind 2 settling up to two picks per cross-workgroup exchange (2..256
 * workgroups; ppt 17 = 1024 x 16 points in registers plus one per thread in
 * LDS, taken for 2^20 < n_points <= 1,114,112; beyond 64 workgroups, 1024 x 16
 * with 2 or 4 waves polling the exchange), or with grid 1 up to two picks per
 * workgroup barrier (the waves paired instead of workgroups), 6 = kind 4
 * settling up to two picks per exchange (2..64 workgroups);
 * threads per workgroup, points per thread (0 when streaming), workgroups,
 * points per workgroup.  FPSX_PLAN in the environment forces a plan:
 * "resident:B:PPT", "pruned:B:PPT", "pair:B:PPT", "pairwPW:B:PPT" (the paired
 * kernel with PW = 2 or 4 polling waves at any grid it holds), "solo:B:PPT" (one
 * workgroup paired at the wave level), "stream:B[:G]",
 * "generic:B[:G]", "spruned[:G]", "spair[:G]";
 * FPSX_NO_PRUNE=1 keeps the planner off the pruned kinds, FPSX_NO_PAIR=1 off
 * kinds 5 and 6.
 */
int fpsx_plan_info(size_t n_points, size_t dim, int tie, size_t n_samples, int *kind, int *block,
                   int *ppt, int *grid, int *chunk);

/*
 * Autotuner: times every feasible plan for `passes` passes (0: 256; pass the
 * number of samples later calls will draw) on the given device cloud -- the
 * planner's own choice, every resident shape (threads per workgroup, points
 * per thread, workgroups), for 3-D clouds every pruned and paired shape and
 * the streaming pruned plan, and the streaming kernel -- and keeps, among the
 * plans within 3% of the fastest, the one with the fewest workgroups, for
 * clouds of exactly this (n_points, dim, tie) on the current device; later
 * calls use it whatever their sample count.  The chosen shape is returned
 * through the optional pointers (fpsx_plan_info reports its kind).  Without
 * tuning the planner uses its per-pass time model.
 */
int fpsx_autotune(const float *d_points, size_t stride_point, size_t stride_dim, size_t n_points,
                  size_t dim, int tie, size_t passes, void *stream, int *chosen_block,
                  int *chosen_ppt, int *chosen_grid);

/*
 * Bound on one cross-workgroup exchange wait for later launches on the
 * current device (FPS, windowed-update and batch objects created afterwards):
 * a kernel whose exchange waits longer gives up and reports FPSX_ETIMEOUT.
 * seconds <= 0 restores the default (FPSX_EXCHANGE_TIMEOUT_S, else 2 s).
 * No counterpart in the reference (its loop is single-threaded).
 */
int fpsx_exchange_timeout(double seconds);

/*
 * Number of cross-workgroup exchanges the last FPS or windowed-update launch
 * on the current device made (after `stream`, NULL: the library's, reaches
 * it): 0 for a one-workgroup plan (but for kind 5 with one workgroup: its
 * rounds, one barrier each), one per pick after the forced prefix for the
 * one-pick kernels, counted by the kernel for the paired kinds and the paired
 * windowed-update kernel (up to two picks per exchange; kind 5 also exchanges
 * workgroup boxes once).  Measurement aid; no counterpart in the reference.
 */
int fpsx_launch_exchanges(void *stream, uint64_t *exchanges);

/*
 * Exchange-only probe: `rounds` rounds of the kernels' cross-workgroup step
 * (candidate staged in LDS, published to the exchange area, every
 * workgroup's words polled, winner handed to every wave) among `grid`
 * workgroups of 1024 threads, with candidates of `words` 64-bit words (5: a
 * 3-D pick; 10: a paired kernel's two), and no point work.  Returns the
 * average microseconds per round: the floor of one FPS pass at that grid.
 * Measurement aid; no counterpart in the reference.
 */
int fpsx_probe_exchange(unsigned grid, unsigned words, unsigned rounds, void *stream,
                        double *us_per_round);

/*
 * Kernel timing on the current device: fpsx_timing(1) makes every later FPS
 * launch record a HIP event pair around the FPS kernel on the stream it is
 * launched on (and forgets earlier records); fpsx_timing_read waits for the
 * recorded launches, returns their summed kernel time and count, and forgets
 * them.  fpsx_timing(0) stops recording.
 */
int fpsx_timing(int enable);
int fpsx_timing_read(double *total_ms, int *launches);

/*
 * Vanilla farthest point sampling.
 * Replaces fps_sampling (src/lib.cpp:188-246) when n_start == 1 and
 * fps_sampling_multi_start_index (src/lib.cpp:111-186) when n_start > 1:
 * start_idx[0..n_start) are picked first, in order, then each further pick is
 * the point with the largest squared distance to the picked set, highest
 * index on ties.  Squared distances are float32, summed in dimension order,
 * no fused multiply-add.
 *   points        host array, n_points x dim float32 in the given layout
 *   start_idx     host array of n_start indices (1 <= n_start <= n_samples)
 *   out           host array of n_samples indices (size_t in the reference)
 * Errors: FPSX_EBADDIM (dim == 0), FPSX_EBADSTART, FPSX_EINVAL (n_samples == 0
 * or > n_points, n_start == 0 or > n_samples, null pointer, bad layout).
 */
int fpsx_fps_sampling(const float *points, size_t n_points, size_t dim, int layout,
                      size_t n_samples, const uint64_t *start_idx, size_t n_start,
                      uint64_t *out);

/*
 * Same computation on device-resident data, enqueued on `stream`
 * (a hipStream_t; NULL uses the library's own stream for the current device).
 *   d_points  device SoA array: coordinate j of point i at d_points[j*ld + i], ld >= n_points
 *   d_start   device array of n_start uint64 indices (values are not re-validated)
 *   d_out     device array of n_samples uint64 indices
 *   tie       FPSX_TIE_HIGHEST or FPSX_TIE_LOWEST
 *   init      initial min-distance (+inf for fps_sampling, FLT_MAX for bucket FPS)
 * The call returns once the work is enqueued.  The library's per-device
 * scratch buffers stay in use until the stream reaches that point; a later
 * call on another stream waits for them on the device (an event), so calls
 * on different streams are correct but do not overlap.
 */
int fpsx_fps_sampling_device(const float *d_points, size_t ld, size_t n_points, size_t dim,
                             size_t n_samples, const uint64_t *d_start, size_t n_start,
                             uint64_t *d_out, int tie, float init, void *stream);

/*
 * General device form: coordinate j of point i at d_points[i*stride_point +
 * j*stride_dim] (SoA: 1, ld; a row-major n x dim tensor: dim, 1), so device
 * tensors of either layout are used in place.  d_out_points (may be NULL)
 * receives the picks' coordinates as an n_samples x dim row-major float32
 * array, written by the kernel as it picks (the `pc[idx]` gather fused in).
 * Streaming plans copy a non-SoA input into the library's SoA scratch first.
 */
int fpsx_fps_sampling_device_strided(const float *d_points, size_t stride_point, size_t stride_dim,
                                     size_t n_points, size_t dim, size_t n_samples,
                                     const uint64_t *d_start, size_t n_start, uint64_t *d_out,
                                     float *d_out_points, int tie, float init, void *stream);

/*
 * Allocation-free form of fpsx_fps_sampling_device_strided, for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  No counterpart in the reference.
 * The call allocates nothing (no hipMalloc, hipFree or allocating library
 * routine): its only scratch, the cross-workgroup exchange area and status
 * word (1 MiB + 256 B), is a fixed array in the library's code object that
 * every GPU the library runs on holds until the process ends.  It runs the
 * register-resident kernels only (the spatially ordered and streaming plans
 * need per-point scratch), so it is slower than the strided entry on large
 * 3-D clouds, and a cloud beyond the register file (about 4 M 3-D points on
 * MI355X) is refused with FPSX_ETOOBIG before anything is enqueued.  The
 * picks are the same as every other entry's.
 *   d_status  optional device uint32: cleared by the call, then receives the
 *             kernel's status bits once the stream reaches the end of the
 *             call (bit 0: an exchange timed out, bit 1: a start index >=
 *             n_points); NULL keeps them in the library's scratch, unread
 *   stream    NULL means the device's null stream (not the library's stream)
 * Every call on one GPU shares that GPU's scratch, which keeps the last call's
 * contents: calls that overlap on one GPU (other streams, other threads)
 * race, so the caller orders them.  The exchange wait bound is
 * FPSX_EXCHANGE_TIMEOUT_S (default 2 s), read on the first call per device.
 */
int fpsx_fps_sampling_device_noalloc(const float *d_points, size_t stride_point,
                                     size_t stride_dim, size_t n_points, size_t dim,
                                     size_t n_samples, const uint64_t *d_start, size_t n_start,
                                     uint64_t *d_out, float *d_out_points, uint32_t *d_status,
                                     int tie, float init, void *stream);

/*
 * FPS with nearest-point-distance-updating (NPDU).
 * Replaces fps_npdu_sampling(points, n_samples, k, start_idx)
 * (src/lib.cpp:272-340): the start updates the min-distance of every point;
 * every later pick p updates only indices p - w/2 .. p + w/2 (the window
 * shifted to stay inside [0, n_points-1], src/lib.cpp:296-300); each pick is
 * the largest min-distance over all points, lowest index on ties.
 *   points    host array, n_points x dim float32, row-major (C order)
 *   w         window (the reference's k; any w >= 2*n_points covers the cloud)
 *   out       host array of n_samples indices
 * Errors: FPSX_EBADDIM (dim == 0), FPSX_EBADSTART (start_idx >= n_points),
 * FPSX_EINVAL (n_samples == 0 or > n_points, null pointer).
 */
int fpsx_fps_npdu_sampling(const float *points, size_t n_points, size_t dim, size_t n_samples,
                           size_t w, size_t start_idx, uint64_t *out);

/*
 * Same computation on device data, enqueued on `stream` (NULL: the library's
 * stream).  Coordinate j of point i at d_points[i*stride_point + j*stride_dim];
 * d_start is a device array holding the one start index; d_out_points (may be
 * NULL) receives the picks' coordinates, n_samples x dim row-major.
 */
int fpsx_fps_npdu_sampling_device(const float *d_points, size_t stride_point, size_t stride_dim,
                                  size_t n_points, size_t dim, size_t n_samples, size_t w,
                                  const uint64_t *d_start, uint64_t *d_out, float *d_out_points,
                                  void *stream);

/*
 * Kernel the windowed-update entries would use for a cloud: kernel_dim 3
 * (coordinates in registers) or 0 (any dimension), workgroup size, points per
 * thread and grid.  FPSX_NPDU_PLAN="D:B:PPT" in the environment forces a shape,
 * "pair:B:PPT" the paired 3-D form (up to two picks per cross-workgroup
 * exchange, 2..64 workgroups, used automatically where it fits; reported as
 * kernel_dim 3; FPSX_NO_PAIR=1 keeps the planner off it), "stream[:G]" the
 * streaming form (min-distances in HBM, used automatically beyond the
 * register file; ppt 0, kernel_dim 3 for 3-D clouds, whose winners carry
 * their coordinates through the exchange, else 0), "stream0[:G]" its
 * any-dimension kernel on any cloud, "streambig[:G]" that kernel with 4x
 * larger blocks (taken automatically where the per-workgroup cache of small
 * blocks would not fit LDS).
 */
int fpsx_npdu_plan_info(size_t n_points, size_t dim, int *kernel_dim, int *block, int *ppt,
                        int *grid);

/*
 * Batched FPS: many independent clouds sampled in one persistent launch.
 * The clouds are concatenated in one SoA device array (coordinate j of point p
 * at d_points[j*ld + p]); cloud c holds points offsets[c] .. offsets[c+1]-1
 * (n_clouds + 1 host offsets) and gets n_samples[c] picks starting at its own
 * index start_idx[c] (host arrays).  Picks land in d_out, cloud after cloud
 * (cloud c at sum of earlier n_samples), each equal to fpsx_fps_sampling on
 * that cloud alone.  A batch object owns its device tables; flags:
 *   FPSX_BATCH_GRAPH     capture the launch sequence in a HIP graph, replayed by run
 *   FPSX_BATCH_AUTOTUNE  time every batch-kernel shape once at creation, keep the fastest
 * Clouds of a dimension with no batch kernel run one launch per cloud.
 * Clouds that need several workgroups settle up to two picks per exchange
 * (same picks); FPSX_BATCH_NO_PAIR=1, read at creation, keeps one per exchange.
 */
#define FPSX_BATCH_GRAPH 1
#define FPSX_BATCH_AUTOTUNE 2
typedef struct fpsx_batch fpsx_batch;
int fpsx_batch_create(const float *d_points, size_t ld, size_t dim, const uint64_t *offsets,
                      size_t n_clouds, const uint64_t *n_samples, const uint64_t *start_idx,
                      uint64_t *d_out, int flags, fpsx_batch **out);
int fpsx_batch_run(fpsx_batch *b, void *stream);    /* enqueue one run (graph replay if captured) */
int fpsx_batch_status(fpsx_batch *b, void *stream); /* wait, report the kernel status */
int fpsx_batch_info(fpsx_batch *b, int *block, int *ppt, int *grid, int *members, int *graph);
void fpsx_batch_destroy(fpsx_batch *b);

/* Host convenience: row-major concatenated clouds in, picks out (synchronous). */
int fpsx_fps_batch(const float *points, size_t dim, const uint64_t *offsets, size_t n_clouds,
                   const uint64_t *n_samples, const uint64_t *start_idx, uint64_t *out);

/*
 * Serving loop: requests (one 3-D cloud each) are submitted from any thread
 * and sampled by a scheduler thread that
 *   - admits pending requests, oldest first, while their device footprint
 *     (points + min-distances + picks) fits `hbm_budget_bytes` of HBM (one
 *     arena allocated at creation) and their workgroups fit the GPU;
 *   - advances every admitted request by up to `chunk_passes` passes per
 *     launch of the batch kernel (continuous batching: requests that arrive
 *     while others run join at the next launch; uploads overlap the kernel);
 *   - retires finished requests and wakes their waiters.
 * Each request's picks equal fpsx_fps_sampling on its cloud alone.
 * block/ppt select the batch-kernel shape (0, 0: 1024 x 16).
 * fpsx_server_submit returns FPSX_EBUDGET for a request larger than the whole
 * budget.  fpsx_server_wait blocks until the request is done, copies its
 * picks and its latency (submit to completion, seconds).  Destroy drains.
 */
typedef struct fpsx_server fpsx_server;
int fpsx_server_create(size_t hbm_budget_bytes, unsigned chunk_passes, int block, int ppt,
                       fpsx_server **out);
int fpsx_server_submit(fpsx_server *s, const float *points, size_t n_points, size_t dim,
                       size_t n_samples, uint64_t start_idx, uint64_t *request_id);
int fpsx_server_wait(fpsx_server *s, uint64_t request_id, uint64_t *out, double *latency_s);
int fpsx_server_stats(fpsx_server *s, uint64_t *launches, uint64_t *completed, uint64_t *samples,
                      double *busy_s, size_t *hbm_in_use, size_t *pending, size_t *active);
void fpsx_server_destroy(fpsx_server *s);

/*
 * Request control (no counterpart in the reference, which has no serving loop).
 *
 * fpsx_server_submit_ex: fpsx_server_submit with a priority and a deadline.
 *   Pending requests are admitted highest priority first (FIFO among equals;
 *   fpsx_server_submit is priority 0).  Priorities also preempt: each launch
 *   takes the admitted requests in priority order while their workgroups fit
 *   the co-resident grid, and a lower-priority request that no longer fits is
 *   paused between two chunks -- its min-distances stay in its HBM range --
 *   and resumes, bit-identically, once the grid has room.  A request is
 *   admitted while the requests of its priority or higher leave room in the
 *   grid.  When its HBM is missing, lower-priority requests are evicted at
 *   their next chunk boundary: each one's whole device range is copied to
 *   host memory and freed, and it waits in the queue (ahead of the requests
 *   of its priority that never ran) to be copied back and resumed where it
 *   stopped, with the same picks.
 *   deadline_s > 0: seconds after submission by which the request must have
 *   finished; past it, a request not yet in its last chunk is stopped at the
 *   next chunk boundary (or dropped from the queue) and its wait returns
 *   FPSX_EDEADLINE.  deadline_s <= 0: no deadline.
 * fpsx_server_cancel: stops a request at its next chunk boundary (or drops it
 *   from the queue) and frees its HBM; its wait returns FPSX_ECANCELED
 *   (FPSX_EDEADLINE if its deadline had already stopped it).  *cancelled
 *   (optional) is 1 if the request ends without picks, 0 if it had already
 *   finished or its last chunk was already queued (its wait returns the
 *   picks).  FPSX_EINVAL for an id that is unknown or already waited for.
 * fpsx_server_wait_for: fpsx_server_wait with a time limit in seconds (< 0:
 *   none, 0: poll).  FPSX_EPENDING when the limit passes first; the request
 *   stays registered and can be waited for or cancelled again.
 * fpsx_server_counters: requests cancelled, requests stopped by their
 *   deadline, (request, launch) pairs in which a request was paused, and
 *   evictions to host memory.
 */
int fpsx_server_submit_ex(fpsx_server *s, const float *points, size_t n_points, size_t dim,
                          size_t n_samples, uint64_t start_idx, int priority, double deadline_s,
                          uint64_t *request_id);
int fpsx_server_cancel(fpsx_server *s, uint64_t request_id, int *cancelled);
int fpsx_server_wait_for(fpsx_server *s, uint64_t request_id, double timeout_s, uint64_t *out,
                         double *latency_s);
int fpsx_server_counters(fpsx_server *s, uint64_t *cancelled, uint64_t *expired,
                         uint64_t *paused, uint64_t *evicted);

/*
 * Bucket FPS entry with the exact signature and return codes of the
 * reference's C API bucket_fps_kdline (src/wrapper.hpp:118-132).
 * Same picks as the reference: the KD-line tree of height `height` is built on
 * the host exactly as src/_ext/KDTreeBase.h:84-207 builds it, the first pick
 * is the point at position start_idx of the permuted order
 * (src/wrapper.hpp:54-55), and exhaustive FPS runs on the GPU over the
 * permuted order with FLT_MAX initial distances and lowest-position ties,
 * which equals the reference's pruned bucket search.
 *   raw_data  host row-major n_points x dim float32
 *   sampled_point_indices  host array of n_samples size_t
 * Returns 0, 1 (dim not in [1, 8]), 2 (start_idx >= n_points), or the codes
 * from 3 up (n_samples == 0, GPU failures).  As in the reference, n_samples
 * may exceed n_points; later picks then repeat the lowest tree position.
 */
int bucket_fps_kdline(const float *raw_data, size_t n_points, size_t dim, size_t n_samples,
                      size_t start_idx, size_t height, size_t *sampled_point_indices);

/*
 * KD-line permutation alone (host): perm[pos] = original index of the point
 * at position pos after the build of src/_ext/KDTreeBase.h:84-207 with leaves
 * at depth `height` (src/_ext/KDLineTree.h:37-39).  leaf_lo/leaf_hi (may be
 * NULL; capacity >= min(n_points, 2^height)) receive the bucket ranges.
 * Returns the number of buckets (>= 1) or a negative FPSX_* code.
 */
long fpsx_kdline_permutation(const float *raw_data, size_t n_points, size_t dim, size_t height,
                             uint64_t *perm, uint64_t *leaf_lo, uint64_t *leaf_hi);

#ifdef __cplusplus
}
#endif

#endif /* FPSX_H */

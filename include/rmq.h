/*
 * rmq.h -- C ABI of the MI355X batch range-minimum engine (librtxrmq.so).
 *
 * Replaces the reference's RTX_blocks launcher
 *     template <typename T> T* rtx_rmq(int alg, int n, int bs, int q,
 *                                      float* darray, int2* dquery, CmdArgs args)
 * (src/rtx.h:24-177, called from src/main.cu:132-137 for alg 5 with T=float
 * and alg 105 with T=int).  The reference fuses build and query in one call
 * and returns a malloc'd host array; here build and query are separate,
 * stream-ordered and write caller-owned device buffers.
 *
 * Semantics (identical to the reference CPU BASE, algs 0/100,
 * src/cpu_methods.h:7-53):
 *   value  = min A[l..r], the exact float32 bits of the minimum element;
 *   index  = the smallest i in [l, r] with A[i] equal to that minimum.
 * Queries are int32 pairs {l, r}, inclusive, 0 <= l <= r < n, the layout of
 * the reference's int2 queries (src/rand.cuh:149-150).  Inputs must not
 * contain NaN (the reference's scan order is undefined for NaN as well).
 * Invalid queries never fault: they produce value NaN and index -1
 * (rmq_query_check counts them).
 *
 * Errors: every call returns RMQ_OK (0) or a negative RMQ_ERR_* code; no
 * exception or abort crosses the ABI (the reference's CUDA_CHECK prints and
 * continues, common/common.h:57-69).
 */
#ifndef RTXRMQ_AMD_RMQ_H
#define RTXRMQ_AMD_RMQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* HIP's stream handle (the same typedef hip_runtime_api.h makes). */
typedef struct ihipStream_t *hipStream_t;

typedef struct rmq_handle rmq_handle_t;

#define RMQ_OK 0
#define RMQ_ERR_INVALID_ARG (-1)
#define RMQ_ERR_OOM (-2)
#define RMQ_ERR_HIP (-3)
#define RMQ_ERR_DEVICE (-4)    /* handle used on a device other than the one it was built on */
#define RMQ_ERR_BAD_QUERY (-5) /* returned by rmq_query_check when a query had l > r or r >= n */
#define RMQ_ERR_LIMIT (-6)     /* beyond the static scratch of rmq_plan_launch_static */

/* Array kinds for the generators (the reference draws U(0,1] with cuRAND,
 * src/rand.cuh:15-21; these draw exact grid values from a counter-based
 * stream so that host and device produce the same array). */
/* rmq_build flags */
#define RMQ_BUILD_EDGES 1u /* precompute the edge-block searches (32 B per element): see rmq_build */
#define RMQ_BUILD_MID 2u   /* with RMQ_BUILD_EDGES: a coarse row for an LDS superset test (64 KB at most) */

#define RMQ_GEN_UNIFORM 0 /* k * 2^-24, k uniform in [0, 2^24): U[0,1) exactly      */
#define RMQ_GEN_TIES16 1  /* integers 0..15 as float (heavy ties)                  */
#define RMQ_GEN_DESC 2    /* bits 0x3F800000 - 1 - i: strictly descending, positive */
#define RMQ_GEN_CONST 3   /* all 1.0f                                               */

/* Query tuning; zero fields take the defaults (rmq_query_opts_default). */
typedef struct rmq_query_opts {
    int32_t short_max;   /* ranges of at most this many elements go to the group scan;
                            with edge tables so does every range inside one 32-element
                            sub-block (one block when bs <= 32), all others are O(1)
                            (default 128)                                              */
    int32_t group_lanes; /* lanes per short query: 4, 8, 16 or 64 (default 8)           */
    int32_t long_wgs;    /* persistent workgroups for the long-range kernel (0 = auto)  */
    int32_t long_lanes;  /* lanes per long query: 4, 8, 16 or 64 (default 4)            */
} rmq_query_opts_t;

/* Per-kernel device time of one rmq_query_timed call, milliseconds. */
typedef struct rmq_query_times {
    float short_ms; /* binning pass fused with the short-range group scans;
                       0 with edge tables (one pass answers every range)   */
    float long_ms;  /* long-range wave scans (edge blocks + table); with
                       edge tables the single pass over all queries        */
    float total_ms; /* from the first launch to the last                   */
} rmq_query_times_t;

/* ---- index build: replaces gen_vertices_blocks_dev + gen_triangles_dev +
 *      buildASFromDeviceData (src/rtx.h:40-60,96; src/device_tools.cuh:243-287;
 *      src/rtx_functions.h:289-379) and compute_min_blocks for alg 105
 *      (src/rtx.h:3-22).  Builds the per-block leftmost min/argmin over blocks
 *      of bs elements (nb = ceil(n/bs), src/rtx.h:42) and a sparse table over
 *      them.  d_A must stay alive and unchanged while the handle is used: the
 *      edge-block scans read it.  n <= 2^31-1, 1 <= bs, d_A 16-byte aligned.
 *      flags: 0, or RMQ_BUILD_EDGES to also precompute, per element, the
 *      leftmost (min, argmin) from its block start and to its block end, and
 *      the same inside 32-element sub-blocks with a sparse table over them
 *      (32 bytes per element): every range not inside one 32-element
 *      sub-block (one block when bs <= 32) is then answered by one lane with
 *      a few O(1) lookups instead of edge scans.  RMQ_BUILD_MID (only with
 *      RMQ_BUILD_EDGES) also keeps one row over coarse blocks of 2^j blocks,
 *      the fewest that make at most 8192 of them, whose entry i is the
 *      leftmost minimum of coarse blocks i and i+1; a query pass whose block
 *      table does not fit in LDS copies that row into LDS and answers a range
 *      from one or two of its entries when their window (the range's at
 *      most 4 coarse blocks) has its leftmost minimum inside the range,
 *      before any table read in global memory.  It pays for ranges a
 *      sizeable fraction of a coarse block long (the autotuner decides).
 *      Stream-ordered; the handle is usable on stream s at once, on other
 *      streams after s is synchronised. ---- */
int rmq_build(const float *d_A, int64_t n, int32_t bs, uint32_t flags, hipStream_t s, rmq_handle_t **out);

/* ---- batched query: replaces the optixLaunch loop running
 *      __raygen__rmq_blocks / __raygen__rmq_blocks_idx (src/rtx.h:150-155,
 *      src/rtx_kernels.cu:85-180,350-449).  d_Q holds q pairs {l, r};
 *      d_val (alg 5) and/or d_idx (alg 105) receive q results; either may be
 *      NULL but not both.  Stream-ordered and asynchronous; allocates no
 *      device memory; the handle is read-only, so concurrent queries on
 *      different streams are safe. ---- */
int rmq_query(const rmq_handle_t *h, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx, hipStream_t s);

/* rmq_query with tuning options (NULL = defaults). */
int rmq_query_ex(const rmq_handle_t *h, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx,
                 const rmq_query_opts_t *opts, hipStream_t s);

/* rmq_query_ex that also times its kernels with HIP events recorded on s and
 * waits for them (a synchronising call, for measurement). */
int rmq_query_timed(const rmq_handle_t *h, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx,
                    const rmq_query_opts_t *opts, hipStream_t s, rmq_query_times_t *times);

/* Synchronising validation: counts queries with l < 0, l > r or r >= n.
 * Returns RMQ_OK when none, RMQ_ERR_BAD_QUERY otherwise (count in *n_bad). */
int rmq_query_check(const int32_t *d_Q, int64_t q, int64_t n, hipStream_t s, int64_t *n_bad);

/* Releases everything the handle owns (not d_A).  NULL is a no-op. */
void rmq_free(rmq_handle_t *h);

/* Handle geometry: number of blocks, sparse-table levels, device bytes owned. */
int rmq_handle_info(const rmq_handle_t *h, int64_t *n, int32_t *bs, int64_t *nb, int32_t *levels,
                    int64_t *bytes);

/* Device bytes rmq_build(n, bs, flags) allocates (what rmq_handle_info
 * reports for the built handle), computed on the host without building.
 * The serving loop admits new arrays by this figure.  No counterpart in the
 * reference, which reports its BVH's output and temp buffer sizes only after
 * building (src/rtx_functions.h:289-379, src/tools.h:316-336). */
int rmq_handle_bytes(int64_t n, int32_t bs, uint32_t flags, int64_t *bytes);

/* Names of the kernels rmq_query launches on this handle with these options
 * (NULL = defaults) for a query that asks for both outputs, separated by ';'
 * in launch order, as a profiler reports them: "k_query_edges_lds<G, 3>"
 * (edge tables, block sparse table held in LDS; a query asking for values
 * only runs <G, 1>, for indexes only <G, 2>), "k_query_edges<G, TWO>" (edge tables; TWO = true when the table is
 * large enough that failed superset tests are retried) or
 * "k_query_tiles<G, false>;k_query_tiles<GL, true>".  Writes a NUL-terminated string of at most len
 * bytes into buf; RMQ_ERR_INVALID_ARG when it does not fit.  No counterpart
 * in the reference: for matching kernel profiles to queries. */
int rmq_query_kernels(const rmq_handle_t *h, const rmq_query_opts_t *opts, char *buf, int64_t len);

/* Default tuning for a handle (what a NULL opts means). */
int rmq_query_opts_default(const rmq_handle_t *h, rmq_query_opts_t *opts);

/* Autotuner pick: block size and query options for one (n, length
 * distribution).  key is the distribution signature the cache is keyed by. */
typedef struct rmq_tune_result {
    int32_t bs;
    uint32_t build_flags; /* 0, RMQ_BUILD_EDGES, or RMQ_BUILD_EDGES | RMQ_BUILD_MID */
    rmq_query_opts_t opts;
    float best_ms;      /* one query pass over the tuning sample */
    int32_t n_tried;    /* configurations timed (0 when taken from the cache) */
    int32_t from_cache;
    char key[96];
} rmq_tune_result_t;

/* ---- autotuner (no counterpart in the reference, whose sweeps are shell
 *      scripts over --nb, scripts/perf-fulltest.sh:45-58): times query passes
 *      over the sample d_Q[0..q) for each block size in bs_list (NULL/0: 2^3 ..
 *      2^18), with and without edge tables (RMQ_BUILD_EDGES), and with edge
 *      tables and coarse rows (RMQ_BUILD_MID) where the block table does not
 *      fit in LDS, then, for the
 *      four fastest of those, the short-range threshold and lanes per query,
 *      and returns the fastest configuration (build it with rmq_build(..., bs,
 *      build_flags, ...)).  Results are cached in the text file cache_path
 *      (NULL: no cache) under a key of: the tuner's version, log2 n and the
 *      10/50/90th percentiles of log2(range length) of the sample, the
 *      array's value signature (mean position of the leftmost minimum in
 *      sampled 64-element windows, and a tie-density class), the outputs
 *      timed ("vi", "v" or "i") and, when bs_list is given, that list.  d_val / d_idx are scratch outputs of q
 *      entries (one may be NULL: pass the outputs the real queries will
 *      ask for).  Synchronising; builds and frees its own handles. ---- */
int rmq_tune(const float *d_A, int64_t n, const int32_t *d_Q, int64_t q, const int32_t *bs_list, int32_t n_bs,
             float *d_val, int32_t *d_idx, const char *cache_path, hipStream_t s, rmq_tune_result_t *out);
/* rmq_tune choosing only among build flags within flags_allowed (0: plain
 * handles only; RMQ_BUILD_EDGES: with or without edge tables; RMQ_BUILD_EDGES
 * | RMQ_BUILD_MID: everything, as rmq_tune).  For handles that go into plans
 * or a server, whose kernel reads no coarse rows, pass RMQ_BUILD_EDGES.  A
 * restricted set is part of the cache key ("/f<flags_allowed>"). */
int rmq_tune_ex(const float *d_A, int64_t n, const int32_t *d_Q, int64_t q, const int32_t *bs_list, int32_t n_bs,
                uint32_t flags_allowed, float *d_val, int32_t *d_idx, const char *cache_path, hipStream_t s,
                rmq_tune_result_t *out);

/* ---- multi-array plans: many arrays (the batched mode) or many requests (the
 *      serving loop) answered by ONE persistent launch.  No counterpart in the
 *      reference, which answers one array per run (src/main.cu:51-171).
 *      A plan holds, on the device, one descriptor per handle and a work list
 *      of chunks of at most item_queries queries (0: sized so the launch has
 *      about four items per resident wave, 64..4096); array i's d_Qs[i]
 *      holds qs[i] pairs {l, r} (0 <= l <= r < n_i) and its answers go to
 *      d_vals[i] and/or d_idxs[i] (per-array NULL allowed; d_vals or d_idxs
 *      may be NULL as a whole).  opts applies to every array (NULL =
 *      defaults).  The pointers are captured at creation: the buffers must
 *      outlive the plan, their contents may change between launches.
 *      Creation is synchronising; launches are stream-ordered.  The plan's
 *      kernel reads each handle's block table and edge tables; coarse rows
 *      (RMQ_BUILD_MID) are not used. ---- */
typedef struct rmq_plan rmq_plan_t;
int rmq_plan_create(const rmq_handle_t *const *handles, int32_t m, const int32_t *const *d_Qs, const int64_t *qs,
                    float *const *d_vals, int32_t *const *d_idxs, const rmq_query_opts_t *opts, int32_t item_queries,
                    hipStream_t s, rmq_plan_t **out);
/* Enqueues the plan: a one-thread kernel resets the device queue head, then
 * the persistent kernel drains the work list (no memset; graph-capturable). */
int rmq_plan_launch(const rmq_plan_t *p, hipStream_t s);
/* Same work through a HIP graph captured on the first call and replayed after. */
int rmq_plan_graph_launch(rmq_plan_t *p, hipStream_t s);
int rmq_plan_info(const rmq_plan_t *p, int32_t *n_arrays, int32_t *n_items, int64_t *n_queries);
void rmq_plan_free(rmq_plan_t *p);
/* ---- allocation-free form of rmq_plan_create + rmq_plan_launch (no
 *      counterpart in the reference): for callers that cannot allocate
 *      device memory (tight serving loops, code that must not call hipMalloc).
 *      Same arguments and answers as a plan over the same handles and
 *      buffers, launched on s; nothing is allocated or freed, on the device
 *      or the host.  The plan's device state (queue head, one descriptor per
 *      handle, the work list) lives in a static device array of the library,
 *      about 2.9 MB, which each GPU holds until the process ends: at most
 *      8192 arrays and 65536 work items (item_queries 0 sizes items as the
 *      plan does and enlarges them to stay within the limit; an explicit
 *      item_queries that passes it, or more arrays, return RMQ_ERR_LIMIT).
 *      The work list is uploaded from a static host image and the call waits
 *      for that upload on s (work already queued on s runs first); the query
 *      kernel itself is asynchronous.  Every call on one GPU shares that
 *      GPU's array: calls on the same stream are ordered, calls that overlap
 *      on different streams of one GPU race.  Not capturable in a HIP graph
 *      (it synchronises s). ---- */
int rmq_plan_launch_static(const rmq_handle_t *const *handles, int32_t m, const int32_t *const *d_Qs,
                           const int64_t *qs, float *const *d_vals, int32_t *const *d_idxs,
                           const rmq_query_opts_t *opts, int32_t item_queries, hipStream_t s);
/* The per-array launch loop the plan is measured against: rmq_query_ex on
 * each array in turn, same arguments as rmq_plan_create. */
int rmq_query_loop(const rmq_handle_t *const *handles, int32_t m, const int32_t *const *d_Qs, const int64_t *qs,
                   float *const *d_vals, int32_t *const *d_idxs, const rmq_query_opts_t *opts, hipStream_t s);

/* ---- serving loop (no counterpart in the reference): continuous batching of
 *      query requests over arrays kept resident in HBM.
 *      - rmq_server_add_array builds a handle for a caller-owned device array
 *        (kept alive by the caller) after checking that its block table fits
 *        in free HBM beside the request budget; returns its id.  The build
 *        runs on the server's own stream, unordered with the caller's: the
 *        array's values must be complete when the call is made.  Launches
 *        and completions of requests go on while it builds; adds run one at
 *        a time.
 *      - rmq_server_submit queues a request: q host query pairs for one array
 *        and host output buffers; it returns at once with a ticket.  The
 *        buffers must stay valid, and the queries unchanged, until
 *        rmq_server_wait returns for that ticket.  Pinned buffers
 *        (hipHostMalloc, torch's pin_memory) are read and written by the
 *        kernel directly over PCIe: no staging copies and no device buffer.
 *        Pageable buffers are staged through device memory from the budget.
 *      - A scheduler thread admits pending requests FIFO while their device
 *        buffers fit in the memory budget and coalesces them, up to
 *        max_batch_queries, into one plan launch on a free slot (stream):
 *        upload, one persistent kernel over all their work items, download.
 *        Up to `slots` launches are in flight; requests arriving meanwhile go
 *        to the next free slot, so their work overlaps on the device.
 *      - Latency is measured from submit to the answers in host memory.
 *      - Memory budget: mem_budget_bytes when set; its unspent part stays
 *        free, so an array whose tables do not fit beside it is refused
 *        (RMQ_ERR_OOM).  Unset, the budget is free HBM less a runtime reserve
 *        (a quarter of free HBM, at most 8 GiB), and it shrinks as resident
 *        arrays are added, so that the arrays' tables take precedence.
 *      Answers are the same as rmq_query's for the same array and queries,
 *      whatever else shares the launch. ---- */
typedef struct rmq_server rmq_server_t;
#define RMQ_SERVER_STAGE_COPIES 1u
typedef struct rmq_server_config {
    int32_t slots;             /* launches in flight at once (0: 3)                    */
    int32_t max_arrays;        /* resident arrays (0: 1024)                            */
    int64_t mem_budget_bytes;  /* cap on request device memory in flight (0: from HBM) */
    int64_t max_batch_queries; /* queries coalesced into one launch (0: 2^24)          */
    int32_t item_queries;      /* queries per work item (0: sized per launch, 64..4096) */
    int32_t flags;             /* RMQ_SERVER_STAGE_COPIES: stage pinned buffers too     */
    rmq_query_opts_t opts;     /* applied to every array (zeros: defaults)             */
} rmq_server_config_t;
typedef struct rmq_server_stats {
    int64_t requests_done, queries_done, launches, failed;
    double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> answers in host memory */
    double wall_ms;                         /* first submit -> last completion  */
    double mean_requests_per_launch;
    int64_t budget_bytes, peak_inflight_bytes;
    double mean_enqueue_ms;     /* host time to enqueue one launch            */
    double mean_device_ms;      /* device time of one launch (HIP events)     */
    double mean_launch_wait_ms; /* end of enqueue -> completion noticed       */
} rmq_server_stats_t;
int rmq_server_create(const rmq_server_config_t *cfg, rmq_server_t **out);
int rmq_server_add_array(rmq_server_t *sv, const float *d_A, int64_t n, int32_t bs, int32_t *array_id);
/* rmq_server_add_array with rmq_build flags (RMQ_BUILD_EDGES). */
int rmq_server_add_array_ex(rmq_server_t *sv, const float *d_A, int64_t n, int32_t bs, uint32_t flags,
                            int32_t *array_id);
int rmq_server_submit(rmq_server_t *sv, int32_t array_id, const int32_t *h_Q, int64_t q, int32_t *h_idx,
                      float *h_val, uint64_t *ticket);
/* Blocks until the request is answered; returns its status (RMQ_ERR_OOM if it
 * can never fit the budget).  Several threads may wait on one ticket: each
 * gets its status, and the ticket is forgotten when the last of them returns
 * (a later wait on it returns RMQ_ERR_INVALID_ARG). */
int rmq_server_wait(rmq_server_t *sv, uint64_t ticket, double *latency_ms);
/* Latency percentiles and throughput since the last reset (reset != 0
 * clears).  p50/p99/mean/max are taken over the latest 2^20 requests at most;
 * the counts and wall_ms cover every request since the reset. */
int rmq_server_stats(rmq_server_t *sv, rmq_server_stats_t *st, int32_t reset);
/* Finishes the launches in flight, then frees everything (not the caller's
 * arrays).  Requests not yet launched are dropped; no thread may submit to or
 * wait on the server during or after this call. */
void rmq_server_destroy(rmq_server_t *sv);

/* ---- host HRMQ, algs 1/101: replaces rmq_rmm_par / rmq_rmm_par_idx
 *      (src/cpu_methods.h:80-159), which build the authors' RMQRMM64
 *      (hrmq/RMQRMM64.cpp:196-251) single-threaded and answer queryRMQ
 *      (hrmq/RMQRMM64.cpp:1463-1465) in an OpenMP loop.  A restatement of
 *      the same published structure (Ferrada & Navarro 2017: balanced
 *      parentheses of the 2d-min-heap, rank/select, a range min-tree over
 *      the excess).  A holds n values compared as int32 (the reference reads
 *      the float array as int*, src/main.cu:117-125) and must outlive the
 *      handle.  Answers: idx = the RIGHTMOST minimum of A[l..r] (HRMQ's tie
 *      rule), val = A[idx] as int32 bits; invalid ranges give -1 and the
 *      bits of NaN.  rmq_hrmq_query runs on nthreads host threads (0: all
 *      cores) in static chunks, as the reference's loop; either output may
 *      be NULL but not both.  No GPU is involved. ---- */
typedef struct rmq_hrmq rmq_hrmq_t;
int rmq_hrmq_build(const int32_t *A, int64_t n, rmq_hrmq_t **out);
int rmq_hrmq_query(const rmq_hrmq_t *h, const int32_t *Q, int64_t q, int32_t *idx, int32_t *val, int32_t nthreads);
/* Bytes the structure holds (the reference prints RMQRMM64::getSize). */
int rmq_hrmq_bytes(const rmq_hrmq_t *h, int64_t *bytes);
void rmq_hrmq_free(rmq_hrmq_t *h);

/* ---- brute force on the device, algs 2/102 and the --check oracle:
 *      replaces kernel_rmq_basic / kernel_rmq_basic_idx / gpu_rmq_basic
 *      (src/cuda_methods.cuh:3-169): one thread per query, linear scan from
 *      A[l] with strict '<'.  Either output may be NULL but not both. ---- */
int rmq_gpu_base(const float *d_A, int64_t n, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx,
                 hipStream_t s);

/* ---- input generators: replace create_random_array_dev (src/rand.cuh:23-43)
 *      and random_queries_par_cpu (src/rand.cuh:179-201).  Streams depend on
 *      (seed, index) only, never on the thread count (unlike the reference's
 *      mt19937(seed*tid), src/rand.cuh:120,140,160).
 *      lr > 0: fixed length lr; -1 uniform length in [1, n-l]; -2 / -3
 *      lognormal round(exp(N(ln 2^14 | ln 2^6, 1))) clamped to [1, n-l];
 *      -4 / -5 lognormal with the reference's scales n/2^8, n/2^15 and
 *      sigma 0.3 (src/rand.cuh:194-198), clamped; -6 r = n-1; -7 / -8 the
 *      -2 / -3 lognormals with medians scaled to n (n^(14/log2 1e8),
 *      n^(6/log2 1e8)).  rmq_gen_queries runs the same stream on the device;
 *      its lognormal lengths can differ from the host's where the device and
 *      host libm round differently (integer-only codes are identical). ---- */
int rmq_gen_array(float *d_A, int64_t n, uint64_t seed, int32_t kind, hipStream_t s);
int rmq_gen_array_host(float *A, int64_t n, uint64_t seed, int32_t kind);
int rmq_gen_queries_host(int32_t *Q, int64_t q, int32_t n, int32_t lr, uint64_t seed, int32_t nthreads);
int rmq_gen_queries(int32_t *d_Q, int64_t q, int32_t n, int32_t lr, uint64_t seed, hipStream_t s);
/* The reference's own query stream (random_queries_par_cpu, src/rand.cuh:116-201),
 * call for call: mt19937(seed*tid) per chunk of ceil(q/nt) queries, the
 * standard library's uniform_int / lognormal distributions, lr > 0 and
 * -1..-5 with the reference's meanings (lognormal sigma 0.3, scales n^0.6,
 * n^0.3, n/2^8, n/2^15).  Depends on nt, as the reference's does. */
int rmq_gen_queries_ref(int32_t *Q, int64_t q, int32_t n, int32_t lr, int32_t seed, int32_t nt);

/* Static description of a return code. */
/* Measurement (no counterpart in the reference): the random-request ceiling
 * that bounds the edge-table query kernel.  count threads each read two
 * 8-byte entries at hashed positions of d_t0 and d_t1 (entries each, a power
 * of two) and write 8 bytes to d_out (count entries); one warm-up launch,
 * then reps timed launches on s; *ms_per_launch is their mean.  Random
 * gathers per second = 2 * count / (ms_per_launch / 1e3). */
int rmq_probe_gather(const void *d_t0, const void *d_t1, int64_t entries, int64_t count, void *d_out, int32_t reps,
                     hipStream_t s, float *ms_per_launch);

const char *rmq_strerror(int code);

/* Library version string. */
const char *rmq_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RTXRMQ_AMD_RMQ_H */

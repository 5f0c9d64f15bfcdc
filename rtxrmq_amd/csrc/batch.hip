// batch.hip -- many arrays (or many requests) in one persistent launch.
//
// No counterpart in the reference, whose driver answers one array per run
// (src/main.cu:51-171).  A plan is a table of array descriptors (the built
// handles) plus a list of work items, each a chunk of up to item_queries
// queries of one array with its own query and output pointers.  One
// persistent kernel drains the item list through a device-side queue head
// (one atomic per item per wave); each item runs the same tile binning and
// answer_class code as the single-array path, so answers are identical.
// rmq_plan_launch enqueues only kernels (the head is reset by a one-thread
// kernel, never a memset), so it can be captured into a HIP graph; a plan's
// graph is captured once and replayed by rmq_plan_graph_launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "rmq_common.h"

namespace rmqx {

struct ArrayDesc {
    const float *A;
    const Pair *st;
    uint64_t magic;  // block_magic(bs)
    int32_t n, bs, nb, pad;
};

struct WorkItem {
    const int2 *Q;
    float *out_v;
    int32_t *out_i;
    int32_t count;
    int32_t array;
};

__global__ void k_reset_head(unsigned *head) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *head = 0u;
}

// Persistent: each wave takes the next item from the queue head until the
// list is exhausted (every wave reaches an index >= n_items and leaves).
template <int GS, int GL>
__global__ __launch_bounds__(kWG) void k_plan_query(const ArrayDesc *__restrict__ descs,
                                                    const WorkItem *__restrict__ items, int n_items,
                                                    unsigned *__restrict__ head, int short_max) {
    __shared__ TileStage stage[kWavesPerWG];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    for (;;) {
        unsigned it = 0;
        if (lane == 0) it = atomicAdd(head, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= (unsigned)n_items) break;
        const WorkItem w = items[it];
        const ArrayDesc d = descs[w.array];
        const int64_t tiles = ((int64_t)w.count + kWave - 1) / kWave;
        for (int64_t t = 0; t < tiles; ++t) {
            const TileMasks m = load_tile(w.Q, w.count, d.n, short_max, t, lane, true, w.out_v, w.out_i);
            answer_class<GS, false>(d.A, d.bs, d.magic, d.nb, d.st, m.shorts, m, lane, t * kWave, w.out_v,
                                    w.out_i, stage[wv]);
            answer_class<GL, true>(d.A, d.bs, d.magic, d.nb, d.st, m.longs, m, lane, t * kWave, w.out_v,
                                   w.out_i, stage[wv]);
        }
    }
}

using PlanKernel = void (*)(const ArrayDesc *, const WorkItem *, int, unsigned *, int);

template <int GS>
PlanKernel pick_long(int gl) {
    switch (gl) {
    case 4: return k_plan_query<GS, 4>;
    case 8: return k_plan_query<GS, 8>;
    case 16: return k_plan_query<GS, 16>;
    default: return k_plan_query<GS, 64>;
    }
}

PlanKernel pick_kernel(int gs, int gl) {
    switch (gs) {
    case 4: return pick_long<4>(gl);
    case 8: return pick_long<8>(gl);
    case 16: return pick_long<16>(gl);
    default: return pick_long<64>(gl);
    }
}

}  // namespace rmqx

using namespace rmqx;

struct rmq_plan {
    int device = -1;
    ArrayDesc *d_descs = nullptr;
    WorkItem *d_items = nullptr;
    unsigned *d_head = nullptr;
    int32_t n_items = 0;
    int32_t n_arrays = 0;
    int64_t n_queries = 0;
    rmq_query_opts_t opts{};
    PlanKernel kernel = nullptr;
    int grid = 1;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void plan_release(rmq_plan *p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->capture_stream) (void)hipStreamDestroy(p->capture_stream);
    if (p->d_descs) (void)hipFree(p->d_descs);
    if (p->d_items) (void)hipFree(p->d_items);
    if (p->d_head) (void)hipFree(p->d_head);
    delete p;
}

int enqueue(const rmq_plan *p, hipStream_t s) {
    k_reset_head<<<1, 64, 0, s>>>(p->d_head);
    int rc = hip_rc(hipGetLastError());
    if (rc != RMQ_OK || p->n_items == 0) return rc;
    hipLaunchKernelGGL(p->kernel, dim3(p->grid), dim3(kWG), 0, s, p->d_descs, p->d_items, p->n_items, p->d_head,
                       p->opts.short_max);
    return hip_rc(hipGetLastError());
}

}  // namespace

RMQ_API int rmq_plan_create(const rmq_handle_t *const *handles, int32_t m, const int32_t *const *d_Qs,
                            const int64_t *qs, float *const *d_vals, int32_t *const *d_idxs,
                            const rmq_query_opts_t *opts, int32_t item_queries, hipStream_t s, rmq_plan_t **out) {
    if (!out) return RMQ_ERR_INVALID_ARG;
    *out = nullptr;
    if (m < 0 || (m > 0 && (!handles || !d_Qs || !qs)) || item_queries < 0) return RMQ_ERR_INVALID_ARG;
    if (item_queries == 0) item_queries = 4096;
    rmq_plan *p = new (std::nothrow) rmq_plan{};
    if (!p) return RMQ_ERR_OOM;
    int rc = hip_rc(hipGetDevice(&p->device));
    if (rc != RMQ_OK) {
        plan_release(p);
        return rc;
    }
    const rmq_handle *first = m > 0 ? handles[0] : nullptr;
    if (first) {
        rc = resolve_opts(first, opts, &p->opts);
    } else {
        rmq_handle dummy{};
        rc = resolve_opts(&dummy, opts, &p->opts);
    }
    if (rc != RMQ_OK) {
        plan_release(p);
        return rc;
    }
    std::vector<ArrayDesc> descs((size_t)m);
    std::vector<WorkItem> items;
    std::vector<double> cost;
    for (int32_t i = 0; i < m; ++i) {
        const rmq_handle *h = handles[i];
        const bool has_out = (d_vals && d_vals[i]) || (d_idxs && d_idxs[i]);
        if (!h || qs[i] < 0 || qs[i] > INT32_MAX || (qs[i] > 0 && (!d_Qs[i] || !has_out)) || h->device != p->device) {
            plan_release(p);
            return h && h->device != p->device ? RMQ_ERR_DEVICE : RMQ_ERR_INVALID_ARG;
        }
        descs[i] = ArrayDesc{h->A, h->st, block_magic(h->bs), (int32_t)h->n, h->bs, (int32_t)h->nb, 0};
        for (int64_t off = 0; off < qs[i]; off += item_queries) {
            const int32_t c = (int32_t)std::min<int64_t>(item_queries, qs[i] - off);
            items.push_back(WorkItem{reinterpret_cast<const int2 *>(d_Qs[i]) + off,
                                     (d_vals && d_vals[i]) ? d_vals[i] + off : nullptr,
                                     (d_idxs && d_idxs[i]) ? d_idxs[i] + off : nullptr, c, i});
            cost.push_back((double)c * std::log2(2.0 + (double)h->n));
        }
        p->n_queries += qs[i];
    }
    // Longest items first, so the queue's tail holds the cheap ones.
    std::vector<size_t> order(items.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<WorkItem> sorted;
    sorted.reserve(items.size());
    for (size_t k : order) sorted.push_back(items[k]);
    if (sorted.size() > (size_t)INT32_MAX) {
        plan_release(p);
        return RMQ_ERR_INVALID_ARG;
    }
    p->n_items = (int32_t)sorted.size();
    p->n_arrays = m;
    p->kernel = pick_kernel(p->opts.group_lanes, p->opts.long_lanes);
    p->grid = (int)std::max<int64_t>(1, std::min<int64_t>((p->n_items + kWavesPerWG - 1) / kWavesPerWG,
                                                          (int64_t)device_cus() * 8));
    if (p->opts.long_wgs > 0) p->grid = std::min(p->grid, p->opts.long_wgs);
    rc = hip_rc(hipMalloc(&p->d_head, 256));
    if (rc == RMQ_OK && m > 0) rc = hip_rc(hipMalloc(&p->d_descs, sizeof(ArrayDesc) * descs.size()));
    if (rc == RMQ_OK && !sorted.empty()) rc = hip_rc(hipMalloc(&p->d_items, sizeof(WorkItem) * sorted.size()));
    if (rc == RMQ_OK && m > 0)
        rc = hip_rc(hipMemcpyAsync(p->d_descs, descs.data(), sizeof(ArrayDesc) * descs.size(),
                                   hipMemcpyHostToDevice, s));
    if (rc == RMQ_OK && !sorted.empty())
        rc = hip_rc(hipMemcpyAsync(p->d_items, sorted.data(), sizeof(WorkItem) * sorted.size(),
                                   hipMemcpyHostToDevice, s));
    if (rc == RMQ_OK) rc = hip_rc(hipStreamSynchronize(s));  // the host vectors die here
    if (rc != RMQ_OK) {
        plan_release(p);
        return rc;
    }
    *out = p;
    return RMQ_OK;
}

RMQ_API int rmq_plan_launch(const rmq_plan_t *p, hipStream_t s) {
    if (!p) return RMQ_ERR_INVALID_ARG;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return RMQ_ERR_HIP;
    if (dev != p->device) return RMQ_ERR_DEVICE;
    return enqueue(p, s);
}

RMQ_API int rmq_plan_graph_launch(rmq_plan_t *p, hipStream_t s) {
    if (!p) return RMQ_ERR_INVALID_ARG;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return RMQ_ERR_HIP;
    if (dev != p->device) return RMQ_ERR_DEVICE;
    if (!p->exec) {
        int rc = RMQ_OK;
        if (!p->capture_stream) rc = hip_rc(hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking));
        if (rc == RMQ_OK) rc = hip_rc(hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal));
        if (rc == RMQ_OK) {
            const int rc_enq = enqueue(p, p->capture_stream);
            const int rc_end = hip_rc(hipStreamEndCapture(p->capture_stream, &p->graph));
            rc = rc_enq != RMQ_OK ? rc_enq : rc_end;
        }
        if (rc == RMQ_OK) rc = hip_rc(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
        if (rc != RMQ_OK) {
            if (p->graph) (void)hipGraphDestroy(p->graph);
            p->graph = nullptr;
            p->exec = nullptr;
            return rc;
        }
    }
    return hip_rc(hipGraphLaunch(p->exec, s));
}

RMQ_API int rmq_plan_info(const rmq_plan_t *p, int32_t *n_arrays, int32_t *n_items, int64_t *n_queries) {
    if (!p) return RMQ_ERR_INVALID_ARG;
    if (n_arrays) *n_arrays = p->n_arrays;
    if (n_items) *n_items = p->n_items;
    if (n_queries) *n_queries = p->n_queries;
    return RMQ_OK;
}

RMQ_API void rmq_plan_free(rmq_plan_t *p) { plan_release(p); }

RMQ_API int rmq_query_loop(const rmq_handle_t *const *handles, int32_t m, const int32_t *const *d_Qs,
                           const int64_t *qs, float *const *d_vals, int32_t *const *d_idxs,
                           const rmq_query_opts_t *opts, hipStream_t s) {
    if (m < 0 || (m > 0 && (!handles || !d_Qs || !qs))) return RMQ_ERR_INVALID_ARG;
    for (int32_t i = 0; i < m; ++i) {
        const int rc = rmq_query_ex(handles[i], d_Qs[i], qs[i], d_vals ? d_vals[i] : nullptr,
                                    d_idxs ? d_idxs[i] : nullptr, opts, s);
        if (rc != RMQ_OK) return rc;
    }
    return RMQ_OK;
}

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
    Blocks k;
    EdgeTables edges;  // pre_b == nullptr: no edge tables
};

inline ArrayDesc desc_of(const rmq_handle *h) {
    ArrayDesc d{blocks_of(h), h->edges};
    return d;
}

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
        const bool edges = d.edges.pre_b != nullptr;  // uniform across the wave
        const int64_t tiles = ((int64_t)w.count + kWave - 1) / kWave;
        for (int64_t t = 0; t < tiles; ++t) {
            // ordinary loads and stores: a server's work items may point at pinned host memory
            const TileMasks m = edges ? load_tile_edges<false>(w.Q, w.count, d.k, short_max, t, lane, w.out_v, w.out_i)
                                      : load_tile<false>(w.Q, w.count, d.k.n, short_max, t, lane, true, w.out_v, w.out_i);
            answer_class<GS, false, false>(d.k, m.shorts, m, lane, t * kWave, w.out_v, w.out_i, stage[wv]);
            if (edges) {
                if ((m.longs >> lane) & 1) {
                    const Pair b = answer_edges(d.edges, d.k, m.l, m.r);
                    write_result<false>(w.out_v, w.out_i, t * kWave + lane, b.v, b.i);
                }
            } else {
                answer_class<GL, true, false>(d.k, m.longs, m, lane, t * kWave, w.out_v, w.out_i, stage[wv]);
            }
        }
    }
}

using PlanKernel = void (*)(const ArrayDesc *, const WorkItem *, int, unsigned *, int);

// Queries per work item when the caller leaves it to us: about four items per
// resident wave (8 workgroups of 4 waves per CU), so a small launch still
// fills the GPU, clamped to [64, 4096] and a multiple of the 64-query tile.
inline int32_t auto_item_queries(int64_t total_queries) {
    const int64_t target_items = (int64_t)device_cus() * 8 * kWavesPerWG * 4;
    int64_t per = (total_queries + target_items - 1) / target_items;
    per = (per + kWave - 1) / kWave * kWave;
    return (int32_t)std::max<int64_t>(kWave, std::min<int64_t>(4096, per));
}

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
    if (item_queries == 0) {
        int64_t total = 0;
        for (int32_t i = 0; i < m; ++i) total += std::max<int64_t>(0, qs[i]);
        item_queries = auto_item_queries(total);
    }
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
        descs[i] = desc_of(h);
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

// ===========================================================================
// allocation-free batched launch: the plan's device state (queue head, array
// descriptors, work items) in a static device array instead of allocations
// ===========================================================================
#include <mutex>

namespace rmqx {
constexpr int32_t kStaticMaxArrays = 8192;
constexpr int32_t kStaticMaxItems = 1 << 16;
constexpr size_t kStaticDescOff = 256;  // bytes [0, 256): the queue head
constexpr size_t kStaticItemOff = kStaticDescOff + (sizeof(ArrayDesc) * kStaticMaxArrays + 255) / 256 * 256;
constexpr size_t kStaticBytes = kStaticItemOff + sizeof(WorkItem) * kStaticMaxItems;  // about 2.9 MB
}  // namespace rmqx

// The device state of rmq_plan_launch_static.  Each GPU the library runs on
// holds its own copy until the process ends.  In host code this name is
// not a device address: its address on a device comes from
// hipGetSymbolAddress (static_scratch below).
static __device__ __align__(256) unsigned char scratch[rmqx::kStaticBytes];

namespace {

constexpr int kMaxStaticDevices = 64;
std::mutex g_static_mu;                                  // one host staging buffer, one caller at a time
unsigned char *g_static_dev[kMaxStaticDevices];          // device address of `scratch`, per device
alignas(256) unsigned char g_static_stage[rmqx::kStaticBytes];  // host image of the descriptors and items

// This device's address of `scratch`, looked up on the device's first call
// (with that device current) and kept.  Caller holds g_static_mu.
int static_scratch(int dev, unsigned char **out) {
    if (dev < 0 || dev >= kMaxStaticDevices) return RMQ_ERR_DEVICE;
    if (!g_static_dev[dev]) {
        void *p = nullptr;
        const int rc = hip_rc(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (rc != RMQ_OK) return rc;
        if (!p) return RMQ_ERR_HIP;
        g_static_dev[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_static_dev[dev];
    return RMQ_OK;
}

}  // namespace

RMQ_API int rmq_plan_launch_static(const rmq_handle_t *const *handles, int32_t m, const int32_t *const *d_Qs,
                                   const int64_t *qs, float *const *d_vals, int32_t *const *d_idxs,
                                   const rmq_query_opts_t *opts, int32_t item_queries, hipStream_t s) {
    if (m > kStaticMaxArrays) return RMQ_ERR_LIMIT;
    if (m < 0 || (m > 0 && (!handles || !d_Qs || !qs)) || item_queries < 0) return RMQ_ERR_INVALID_ARG;
    int dev = -1;
    RMQ_HIP(hipGetDevice(&dev));
    rmq_query_opts_t o;
    rmq_handle dummy{};
    int rc = resolve_opts(m > 0 ? handles[0] : &dummy, opts, &o);
    if (rc != RMQ_OK) return rc;
    int64_t total = 0;
    for (int32_t i = 0; i < m; ++i) {
        const rmq_handle *h = handles[i];
        const bool has_out = (d_vals && d_vals[i]) || (d_idxs && d_idxs[i]);
        if (!h || qs[i] < 0 || qs[i] > INT32_MAX || (qs[i] > 0 && (!d_Qs[i] || !has_out))) return RMQ_ERR_INVALID_ARG;
        if (h->device != dev) return RMQ_ERR_DEVICE;
        total += qs[i];
    }
    if (total == 0) return RMQ_OK;
    // work items: as rmq_plan_create sizes them, made larger when that would
    // pass the static limit (an explicit size that passes it is refused)
    int64_t iq = item_queries > 0 ? item_queries : auto_item_queries(total);
    auto count_items = [&](int64_t per) {
        int64_t c = 0;
        for (int32_t i = 0; i < m; ++i) c += (qs[i] + per - 1) / per;
        return c;
    };
    int64_t n_items = count_items(iq);
    if (n_items > kStaticMaxItems) {
        if (item_queries > 0) return RMQ_ERR_LIMIT;
        while (n_items > kStaticMaxItems) n_items = count_items(iq *= 2);
    }
    std::lock_guard<std::mutex> lk(g_static_mu);
    unsigned char *d_scr = nullptr;
    if ((rc = static_scratch(dev, &d_scr)) != RMQ_OK) return rc;
    ArrayDesc *h_descs = reinterpret_cast<ArrayDesc *>(g_static_stage + kStaticDescOff);
    WorkItem *h_items = reinterpret_cast<WorkItem *>(g_static_stage + kStaticItemOff);
    size_t k = 0;
    for (int32_t i = 0; i < m; ++i) {
        h_descs[i] = desc_of(handles[i]);
        for (int64_t off = 0; off < qs[i]; off += iq)
            h_items[k++] = WorkItem{reinterpret_cast<const int2 *>(d_Qs[i]) + off,
                                    (d_vals && d_vals[i]) ? d_vals[i] + off : nullptr,
                                    (d_idxs && d_idxs[i]) ? d_idxs[i] + off : nullptr,
                                    (int32_t)std::min<int64_t>(iq, qs[i] - off), i};
    }
    // longest items first, as in rmq_plan_create (in place: no allocation)
    auto cost = [&](const WorkItem &w) { return (double)w.count * std::log2(2.0 + (double)handles[w.array]->n); };
    std::sort(h_items, h_items + k, [&](const WorkItem &a, const WorkItem &b) {
        const double ca = cost(a), cb = cost(b);
        if (ca != cb) return ca > cb;
        return a.array != b.array ? a.array < b.array : a.Q < b.Q;
    });
    // upload, then wait for it: the host image is reused by the next call
    rc = hip_rc(hipMemcpyAsync(d_scr + kStaticDescOff, h_descs, sizeof(ArrayDesc) * (size_t)m, hipMemcpyHostToDevice, s));
    if (rc == RMQ_OK)
        rc = hip_rc(hipMemcpyAsync(d_scr + kStaticItemOff, h_items, sizeof(WorkItem) * k, hipMemcpyHostToDevice, s));
    if (rc == RMQ_OK) rc = hip_rc(hipStreamSynchronize(s));
    if (rc != RMQ_OK) return rc;
    unsigned *d_head = reinterpret_cast<unsigned *>(d_scr);
    k_reset_head<<<1, 64, 0, s>>>(d_head);  // a kernel, not a memset
    rc = hip_rc(hipGetLastError());
    if (rc != RMQ_OK) return rc;
    int grid = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)k + kWavesPerWG - 1) / kWavesPerWG,
                                                           (int64_t)device_cus() * 8));
    if (o.long_wgs > 0) grid = std::min(grid, o.long_wgs);
    hipLaunchKernelGGL(pick_kernel(o.group_lanes, o.long_lanes), dim3(grid), dim3(kWG), 0, s,
                       reinterpret_cast<const ArrayDesc *>(d_scr + kStaticDescOff),
                       reinterpret_cast<const WorkItem *>(d_scr + kStaticItemOff), (int)k, d_head, o.short_max);
    return hip_rc(hipGetLastError());
}

// ===========================================================================
// serving loop: continuous batching over resident arrays
// ===========================================================================
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int32_t array = 0;
    const int32_t *h_Q = nullptr;
    int64_t q = 0;
    int32_t *h_idx = nullptr;
    float *h_val = nullptr;
    // Device-visible addresses of pinned host buffers: the kernel reads the
    // queries and writes the answers over PCIe, with no staging copies.
    bool zero_copy = false;
    const int2 *z_Q = nullptr;
    int32_t *z_idx = nullptr;
    float *z_val = nullptr;
    int64_t bytes = 0;  // device bytes while in flight (0 when zero_copy)
    double t_submit = 0, t_done = 0;
    int status = RMQ_OK;
    bool done = false;
    int waiters = 0;  // threads in rmq_server_wait for this ticket
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t start_ev = nullptr;
    hipEvent_t done_ev = nullptr;
    double t_enqueued = 0;  // host time launch_group returned
    WorkItem *h_items = nullptr;  // pinned staging of the work list
    WorkItem *d_items = nullptr;
    size_t items_cap = 0;
    unsigned *d_head = nullptr;
    void *d_buf = nullptr;
    int64_t bytes = 0;
    bool busy = false;      // a launch is in flight
    bool launched = false;  // its completion is awaited
    std::vector<uint64_t> tickets;
};

int64_t round256(int64_t b) { return (b + 255) / 256 * 256; }

// Latency percentiles are taken over at most this many of the latest
// requests, so a server that never resets its statistics holds 8 MiB of them.
constexpr size_t kLatencyWindow = (size_t)1 << 20;

// Device bytes a request holds while in flight: its queries and outputs,
// each rounded to 256 bytes (the carving launch_group does).
int64_t request_bytes(const Request &r) {
    if (r.zero_copy) return 0;
    return round256(r.q * 8) + (r.h_idx ? round256(r.q * 4) : 0) + (r.h_val ? round256(r.q * 4) : 0);
}

// Device-visible address of a pinned host buffer (hipHostMalloc, torch's
// pin_memory), or nullptr for pageable memory, which must be staged.
template <typename T>
T *device_view(T *p) {
    if (!p) return nullptr;
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // an unknown pointer leaves an error behind
        return nullptr;
    }
    if (a.type != hipMemoryTypeHost || !a.devicePointer || !a.hostPointer) return nullptr;
    char *d = static_cast<char *>(a.devicePointer) +
              (reinterpret_cast<const char *>(p) - static_cast<const char *>(a.hostPointer));
    return reinterpret_cast<T *>(d);
}

}  // namespace

struct rmq_server {
    rmq_server_config_t cfg{};
    int device = -1;
    std::vector<rmq_handle *> arrays;
    hipMemPool_t pool = nullptr;  // staging buffers of pageable requests (own pool, freed memory kept)
    ArrayDesc *d_descs = nullptr;
    int32_t descs_cap = 0;
    PlanKernel kernel = nullptr;
    std::vector<Slot> slots;
    hipStream_t build_stream = nullptr;  // builds of added arrays, apart from the slots' launches
    std::mutex add_mu;  // one rmq_server_add_array at a time; held without mu while a table builds
    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_slot;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    uint64_t next_ticket = 1;
    int64_t budget = 0, inflight = 0, peak_inflight = 0;
    int64_t reserve = 0;        // HBM kept free for the runtime
    bool budget_fixed = false;  // the caller set mem_budget_bytes
    bool stop = false;
    std::thread sched;
    std::vector<std::thread> completers;
    // statistics
    std::vector<double> lat_ms;  // the last kLatencyWindow latencies, as a ring from lat_next
    size_t lat_next = 0;
    int64_t requests_done = 0, queries_done = 0, launches = 0, failed = 0, batched_requests = 0;
    double first_submit = -1, last_done = -1;
    double sum_enqueue_ms = 0, sum_device_ms = 0, sum_wait_ms = 0;
};

namespace {

int launch_group(rmq_server *sv, Slot &sl, std::vector<Request *> &group) {
    int64_t total = 0, total_q = 0;
    for (Request *r : group) {
        total += r->bytes;
        total_q += r->q;
    }
    const int iq = sv->cfg.item_queries > 0 ? sv->cfg.item_queries : auto_item_queries(total_q);
    size_t n_items = 0;
    for (Request *r : group) n_items += (size_t)((r->q + iq - 1) / iq);
    if (n_items > sl.items_cap) {  // the slot is idle: its previous launch completed
        if (sl.h_items) (void)hipHostFree(sl.h_items);
        if (sl.d_items) (void)hipFree(sl.d_items);
        sl.h_items = nullptr;
        sl.d_items = nullptr;
        sl.items_cap = 0;
        const size_t cap = std::max<size_t>(n_items, 1024);
        int rc = hip_rc(hipHostMalloc((void **)&sl.h_items, sizeof(WorkItem) * cap, hipHostMallocDefault));
        if (rc == RMQ_OK) rc = hip_rc(hipMalloc(&sl.d_items, sizeof(WorkItem) * cap));
        if (rc != RMQ_OK) return rc;
        sl.items_cap = cap;
    }
    int rc = hip_rc(hipEventRecord(sl.start_ev, sl.stream));
    // staged requests (pageable host buffers) share one stream-ordered buffer
    if (rc == RMQ_OK && total > 0) rc = hip_rc(hipMallocFromPoolAsync(&sl.d_buf, (size_t)total, sv->pool, sl.stream));
    if (rc != RMQ_OK) return rc;
    char *p = static_cast<char *>(sl.d_buf);
    size_t k = 0;
    for (Request *r : group) {
        const int2 *d_Q = r->z_Q;
        int32_t *d_idx = r->z_idx;
        float *d_val = r->z_val;
        if (!r->zero_copy) {
            d_Q = reinterpret_cast<const int2 *>(p);
            p += round256(r->q * 8);
            d_idx = nullptr;
            d_val = nullptr;
            if (r->h_idx) {
                d_idx = reinterpret_cast<int32_t *>(p);
                p += round256(r->q * 4);
            }
            if (r->h_val) {
                d_val = reinterpret_cast<float *>(p);
                p += round256(r->q * 4);
            }
            if (rc == RMQ_OK)
                rc = hip_rc(hipMemcpyAsync(const_cast<int2 *>(d_Q), r->h_Q, (size_t)r->q * 8, hipMemcpyHostToDevice,
                                           sl.stream));
        }
        for (int64_t off = 0; off < r->q; off += iq) {
            sl.h_items[k++] = WorkItem{d_Q + off, d_val ? d_val + off : nullptr, d_idx ? d_idx + off : nullptr,
                                       (int32_t)std::min<int64_t>(iq, r->q - off), r->array};
        }
    }
    if (rc == RMQ_OK)
        rc = hip_rc(hipMemcpyAsync(sl.d_items, sl.h_items, sizeof(WorkItem) * n_items, hipMemcpyHostToDevice,
                                   sl.stream));
    if (rc == RMQ_OK) {
        k_reset_head<<<1, 64, 0, sl.stream>>>(sl.d_head);
        const int grid = (int)std::max<int64_t>(
            1, std::min<int64_t>(((int64_t)n_items + kWavesPerWG - 1) / kWavesPerWG, (int64_t)device_cus() * 8));
        hipLaunchKernelGGL(sv->kernel, dim3(grid), dim3(kWG), 0, sl.stream, sv->d_descs, sl.d_items, (int)n_items,
                           sl.d_head, sv->cfg.opts.short_max);
        rc = hip_rc(hipGetLastError());
    }
    // copy the staged answers back, request by request
    p = static_cast<char *>(sl.d_buf);
    for (Request *r : group) {
        if (r->zero_copy) continue;
        p += round256(r->q * 8);
        if (r->h_idx) {
            if (rc == RMQ_OK)
                rc = hip_rc(hipMemcpyAsync(r->h_idx, p, (size_t)r->q * 4, hipMemcpyDeviceToHost, sl.stream));
            p += round256(r->q * 4);
        }
        if (r->h_val) {
            if (rc == RMQ_OK)
                rc = hip_rc(hipMemcpyAsync(r->h_val, p, (size_t)r->q * 4, hipMemcpyDeviceToHost, sl.stream));
            p += round256(r->q * 4);
        }
    }
    int rc_free = RMQ_OK;
    if (sl.d_buf) rc_free = hip_rc(hipFreeAsync(sl.d_buf, sl.stream));
    sl.d_buf = nullptr;
    const int rc_ev = hip_rc(hipEventRecord(sl.done_ev, sl.stream));
    return rc != RMQ_OK ? rc : (rc_free != RMQ_OK ? rc_free : rc_ev);
}

void scheduler_loop(rmq_server *sv) {
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    const int64_t max_q = sv->cfg.max_batch_queries > 0 ? sv->cfg.max_batch_queries : (int64_t)1 << 24;
    for (;;) {
        Slot *free_slot = nullptr;
        sv->cv_sched.wait(lk, [&] {
            if (sv->stop) return true;
            if (sv->pending.empty()) return false;
            for (Slot &s : sv->slots)
                if (!s.busy) {
                    free_slot = &s;
                    return true;
                }
            return false;
        });
        if (sv->stop) return;
        // admission: FIFO, as many pending requests as the memory budget and
        // the per-launch query cap allow
        std::vector<Request *> group;
        std::vector<uint64_t> tickets;
        int64_t bytes = 0, nq = 0;
        while (!sv->pending.empty()) {
            Request &r = sv->reqs[sv->pending.front()];
            if (r.bytes > sv->budget) {  // can never fit: fail it
                r.status = RMQ_ERR_OOM;
                r.done = true;
                r.t_done = now_ms();
                ++sv->failed;
                sv->pending.pop_front();
                sv->cv_done.notify_all();
                continue;
            }
            if (sv->inflight + bytes + r.bytes > sv->budget) break;
            if (!group.empty() && nq + r.q > max_q) break;
            group.push_back(&r);
            tickets.push_back(sv->pending.front());
            bytes += r.bytes;
            nq += r.q;
            sv->pending.pop_front();
        }
        if (group.empty()) {
            // waiting for memory: sleep until a launch completes
            sv->cv_sched.wait(lk, [&] {
                if (sv->stop || sv->pending.empty()) return true;
                return sv->inflight + sv->reqs[sv->pending.front()].bytes <= sv->budget;
            });
            continue;
        }
        sv->inflight += bytes;
        sv->peak_inflight = std::max(sv->peak_inflight, sv->inflight);
        free_slot->busy = true;
        free_slot->bytes = bytes;
        free_slot->tickets = tickets;
        ++sv->launches;
        sv->batched_requests += (int64_t)group.size();
        lk.unlock();
        const double t0 = now_ms();
        const int rc = launch_group(sv, *free_slot, group);
        free_slot->t_enqueued = now_ms();
        lk.lock();
        sv->sum_enqueue_ms += free_slot->t_enqueued - t0;
        if (rc != RMQ_OK) {
            // whatever was enqueued before the failure may still read the
            // requests' host buffers: let it finish before they are released
            lk.unlock();
            (void)hipStreamSynchronize(free_slot->stream);
            lk.lock();
            for (Request *r : group) r->status = rc;
        }
        free_slot->launched = true;
        sv->cv_slot.notify_all();
    }
}

void completer_loop(rmq_server *sv, size_t si) {
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    Slot &sl = sv->slots[si];
    for (;;) {
        sv->cv_slot.wait(lk, [&] { return sv->stop || sl.launched; });
        if (sl.launched) {
            lk.unlock();
            // poll with a short sleep: completion is noticed within tens of
            // microseconds without a core spinning on the event
            hipError_t e;
            while ((e = hipEventQuery(sl.done_ev)) == hipErrorNotReady)
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            const double t = now_ms();
            float dev_ms = 0.f;
            if (e == hipSuccess) (void)hipEventElapsedTime(&dev_ms, sl.start_ev, sl.done_ev);
            lk.lock();
            sv->sum_device_ms += dev_ms;
            sv->sum_wait_ms += t - sl.t_enqueued;
            for (uint64_t tk : sl.tickets) {
                Request &r = sv->reqs[tk];
                if (e != hipSuccess && r.status == RMQ_OK) r.status = RMQ_ERR_HIP;
                r.done = true;
                r.t_done = t;
                if (r.status == RMQ_OK) {
                    ++sv->requests_done;
                    sv->queries_done += r.q;
                    if (sv->lat_ms.size() < kLatencyWindow) {
                        sv->lat_ms.push_back(t - r.t_submit);
                    } else {
                        sv->lat_ms[sv->lat_next] = t - r.t_submit;
                        sv->lat_next = (sv->lat_next + 1) % kLatencyWindow;
                    }
                } else {
                    ++sv->failed;
                }
            }
            sv->last_done = t;
            sv->inflight -= sl.bytes;
            sl.bytes = 0;
            sl.tickets.clear();
            sl.launched = false;
            sl.busy = false;
            sv->cv_done.notify_all();
            sv->cv_sched.notify_all();
            continue;
        }
        if (sv->stop) return;
    }
}

void server_release(rmq_server *sv) {
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_sched.notify_all();
    sv->cv_slot.notify_all();
    if (sv->sched.joinable()) sv->sched.join();
    for (auto &t : sv->completers)
        if (t.joinable()) t.join();
    for (Slot &s : sv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.done_ev) (void)hipEventDestroy(s.done_ev);
        if (s.start_ev) (void)hipEventDestroy(s.start_ev);
        if (s.stream) (void)hipStreamDestroy(s.stream);
        if (s.h_items) (void)hipHostFree(s.h_items);
        if (s.d_items) (void)hipFree(s.d_items);
        if (s.d_head) (void)hipFree(s.d_head);
    }
    if (sv->build_stream) {
        (void)hipStreamSynchronize(sv->build_stream);
        (void)hipStreamDestroy(sv->build_stream);
    }
    for (rmq_handle *h : sv->arrays) rmq_free(h);
    if (sv->d_descs) (void)hipFree(sv->d_descs);
    if (sv->pool) (void)hipMemPoolDestroy(sv->pool);  // every staging buffer was freed on a synchronised stream
    delete sv;
}

}  // namespace

RMQ_API int rmq_server_create(const rmq_server_config_t *cfg, rmq_server_t **out) {
    if (!out) return RMQ_ERR_INVALID_ARG;
    *out = nullptr;
    rmq_server_config_t c{};
    if (cfg) c = *cfg;
    if (c.slots < 0 || c.max_arrays < 0 || c.mem_budget_bytes < 0 || c.max_batch_queries < 0 || c.item_queries < 0 ||
        (c.flags & ~(int32_t)RMQ_SERVER_STAGE_COPIES))
        return RMQ_ERR_INVALID_ARG;
    if (c.slots == 0) c.slots = 3;
    if (c.max_arrays == 0) c.max_arrays = 1024;
    rmq_server *sv = new (std::nothrow) rmq_server{};
    if (!sv) return RMQ_ERR_OOM;
    int rc = hip_rc(hipGetDevice(&sv->device));
    rmq_handle dummy{};
    if (rc == RMQ_OK) rc = resolve_opts(&dummy, &c.opts, &c.opts);
    sv->cfg = c;
    if (rc == RMQ_OK) {
        // staging buffers of pageable requests come from a stream-ordered pool
        // of the server's own, which keeps what is freed cached instead of
        // returning it at every synchronisation (the device's default pool,
        // shared with the rest of the process, is left as it is)
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = sv->device;
        rc = hip_rc(hipMemPoolCreate(&sv->pool, &props));
        if (rc == RMQ_OK) {
            uint64_t keep = UINT64_MAX;
            rc = hip_rc(hipMemPoolSetAttribute(sv->pool, hipMemPoolAttrReleaseThreshold, &keep));
        }
    }
    size_t free_b = 0, total_b = 0;
    if (rc == RMQ_OK) rc = hip_rc(hipMemGetInfo(&free_b, &total_b));
    if (rc == RMQ_OK) {
        // admission budget for request buffers: free HBM less a reserve for
        // array tables and the runtime, or the configured cap if smaller
        sv->reserve = (int64_t)std::min<size_t>(free_b / 4, (size_t)8 << 30);
        sv->budget = (int64_t)free_b - sv->reserve;
        if (c.mem_budget_bytes > 0) sv->budget = std::min<int64_t>(sv->budget, c.mem_budget_bytes);
        sv->budget_fixed = c.mem_budget_bytes > 0;
        sv->kernel = pick_kernel(c.opts.group_lanes, c.opts.long_lanes);
        rc = hip_rc(hipMalloc(&sv->d_descs, sizeof(ArrayDesc) * (size_t)c.max_arrays));
        sv->descs_cap = c.max_arrays;
    }
    if (rc == RMQ_OK) rc = hip_rc(hipStreamCreateWithFlags(&sv->build_stream, hipStreamNonBlocking));
    sv->slots.resize((size_t)c.slots);
    for (Slot &s : sv->slots) {
        if (rc == RMQ_OK) rc = hip_rc(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
        if (rc == RMQ_OK) rc = hip_rc(hipEventCreate(&s.start_ev));
        if (rc == RMQ_OK) rc = hip_rc(hipEventCreate(&s.done_ev));
        if (rc == RMQ_OK) rc = hip_rc(hipMalloc(&s.d_head, 256));
    }
    if (rc != RMQ_OK) {
        server_release(sv);
        return rc;
    }
    sv->sched = std::thread(scheduler_loop, sv);
    for (size_t i = 0; i < sv->slots.size(); ++i) sv->completers.emplace_back(completer_loop, sv, i);
    *out = sv;
    return RMQ_OK;
}

RMQ_API int rmq_server_add_array_ex(rmq_server_t *sv, const float *d_A, int64_t n, int32_t bs, uint32_t flags,
                                    int32_t *array_id) {
    if (!sv || !array_id || !d_A || n <= 0 || n > INT32_MAX || bs <= 0 || !build_flags_ok(flags))
        return RMQ_ERR_INVALID_ARG;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return RMQ_ERR_HIP;
    if (dev != sv->device) return RMQ_ERR_DEVICE;
    // Adds are serialised by add_mu; the server lock is held only to admit
    // and to publish, so launches and completions go on while a table builds.
    std::lock_guard<std::mutex> add_lk(sv->add_mu);
    const int64_t table = handle_sizes(n, bs, flags).total();
    int32_t id = 0;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if ((int32_t)sv->arrays.size() >= sv->descs_cap) return RMQ_ERR_OOM;
        // admission of a new resident array: its tables (exactly what
        // rmq_build allocates) must fit in free HBM beside the request budget
        // still available.  A configured budget is a promise: its unspent part
        // stays free (requests launched during the build spend from that part
        // only).  A budget derived from free HBM yields to resident arrays
        // instead (it shrinks below, once the tables are built); only the
        // runtime reserve stays free.
        size_t free_b = 0, total_b = 0;
        const int rc = hip_rc(hipMemGetInfo(&free_b, &total_b));
        if (rc != RMQ_OK) return rc;
        const int64_t promised = sv->budget_fixed ? sv->budget - sv->inflight : sv->reserve;
        if (table + promised > (int64_t)free_b) return RMQ_ERR_OOM;
        id = (int32_t)sv->arrays.size();  // stays the next id: only an add (under add_mu) appends
    }
    rmq_handle_t *h = nullptr;
    int rc = rmq_build(d_A, n, bs, flags, sv->build_stream, &h);
    if (rc == RMQ_OK) rc = hip_rc(hipStreamSynchronize(sv->build_stream));
    // the entry is written beyond every entry an in-flight launch can read:
    // requests name it only once it is published below
    const ArrayDesc d = rc == RMQ_OK ? desc_of(h) : ArrayDesc{};
    if (rc == RMQ_OK) rc = hip_rc(hipMemcpy(sv->d_descs + id, &d, sizeof(d), hipMemcpyHostToDevice));
    if (rc != RMQ_OK) {
        rmq_free(h);
        return rc;
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->arrays.push_back(h);
    *array_id = id;
    size_t free_b = 0, total_b = 0;
    if (!sv->budget_fixed && hipMemGetInfo(&free_b, &total_b) == hipSuccess)
        sv->budget = std::max<int64_t>(0, std::min<int64_t>(sv->budget, sv->inflight + (int64_t)free_b - sv->reserve));
    return RMQ_OK;
}

RMQ_API int rmq_server_add_array(rmq_server_t *sv, const float *d_A, int64_t n, int32_t bs, int32_t *array_id) {
    return rmq_server_add_array_ex(sv, d_A, n, bs, 0, array_id);
}

RMQ_API int rmq_server_submit(rmq_server_t *sv, int32_t array_id, const int32_t *h_Q, int64_t q, int32_t *h_idx,
                              float *h_val, uint64_t *ticket) {
    if (!sv || !ticket || q <= 0 || q > INT32_MAX || !h_Q || (!h_idx && !h_val)) return RMQ_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (array_id < 0 || array_id >= (int32_t)sv->arrays.size()) return RMQ_ERR_INVALID_ARG;
    Request r;
    r.array = array_id;
    r.h_Q = h_Q;
    r.q = q;
    r.h_idx = h_idx;
    r.h_val = h_val;
    if (!(sv->cfg.flags & RMQ_SERVER_STAGE_COPIES)) {
        r.z_Q = device_view(reinterpret_cast<const int2 *>(h_Q));
        r.z_idx = device_view(h_idx);
        r.z_val = device_view(h_val);
        r.zero_copy = r.z_Q && (!h_idx || r.z_idx) && (!h_val || r.z_val);
    }
    r.bytes = request_bytes(r);
    r.t_submit = now_ms();
    if (sv->first_submit < 0) sv->first_submit = r.t_submit;
    const uint64_t tk = sv->next_ticket++;
    sv->reqs.emplace(tk, r);
    sv->pending.push_back(tk);
    *ticket = tk;
    sv->cv_sched.notify_one();
    return RMQ_OK;
}

RMQ_API int rmq_server_wait(rmq_server_t *sv, uint64_t ticket, double *latency_ms) {
    if (!sv) return RMQ_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return RMQ_ERR_INVALID_ARG;
    // a reference, not the iterator: submits during the wait may rehash the
    // map, which moves no element but invalidates iterators
    Request &r = it->second;
    ++r.waiters;
    sv->cv_done.wait(lk, [&] { return r.done; });
    const int status = r.status;
    if (latency_ms) *latency_ms = r.t_done - r.t_submit;
    if (--r.waiters == 0) sv->reqs.erase(ticket);  // the last waiter forgets the ticket
    return status;
}

RMQ_API int rmq_server_stats(rmq_server_t *sv, rmq_server_stats_t *st, int32_t reset) {
    if (!sv || !st) return RMQ_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    *st = rmq_server_stats_t{};
    st->requests_done = sv->requests_done;
    st->queries_done = sv->queries_done;
    st->launches = sv->launches;
    st->failed = sv->failed;
    st->budget_bytes = sv->budget;
    st->peak_inflight_bytes = sv->peak_inflight;
    st->mean_requests_per_launch = sv->launches ? (double)sv->batched_requests / sv->launches : 0.0;
    if (!sv->lat_ms.empty()) {
        std::vector<double> v = sv->lat_ms;
        std::sort(v.begin(), v.end());
        auto at = [&](double p) { return v[(size_t)std::min<double>(v.size() - 1, p * (v.size() - 1) + 0.5)]; };
        st->p50_ms = at(0.50);
        st->p99_ms = at(0.99);
        st->max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        st->mean_ms = sum / v.size();
    }
    st->wall_ms = (sv->first_submit >= 0 && sv->last_done >= 0) ? sv->last_done - sv->first_submit : 0.0;
    if (sv->launches) {
        st->mean_enqueue_ms = sv->sum_enqueue_ms / sv->launches;
        st->mean_device_ms = sv->sum_device_ms / sv->launches;
        st->mean_launch_wait_ms = sv->sum_wait_ms / sv->launches;
    }
    if (reset) {
        sv->lat_ms.clear();
        sv->lat_next = 0;
        sv->requests_done = sv->queries_done = sv->launches = sv->failed = sv->batched_requests = 0;
        sv->first_submit = sv->last_done = -1;
        sv->peak_inflight = sv->inflight;
        sv->sum_enqueue_ms = sv->sum_device_ms = sv->sum_wait_ms = 0;
    }
    return RMQ_OK;
}

RMQ_API void rmq_server_destroy(rmq_server_t *sv) {
    if (sv) server_release(sv);
}

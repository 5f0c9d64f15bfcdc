// rmq.hip -- CDNA4 (gfx950) kernels and C ABI of the batch range-minimum engine.
//
// The reference answers a query with up to three OptiX ray casts into a
// two-level triangle scene (src/rtx_kernels.cu:85-180): one over the block
// minima between the edge blocks, one into each edge block.  Here the two
// levels are plain arrays in HBM:
//   level 1: the leftmost (min, argmin) of every block of bs elements, with a
//            sparse table over them (O(1) lookup of any run of whole blocks);
//   level 0: the array itself, scanned directly for the edge blocks.
// A query pass is a binning kernel that sends short ranges to group scans
// (several queries per wave) and queues the rest for a persistent kernel that
// gives each long range a whole wave: two edge-block scans plus one table
// lookup, combined so that ties resolve to the leftmost index.
//
// All comparisons are compare-and-select on float32 ('<' and '=='), never
// fminf, so the result is the exact bits of the leftmost minimum element.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <string>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "rmq_common.h"

namespace rmqx {

// ---------------------------------------------------------------------------
// build kernels
// ---------------------------------------------------------------------------

// Leftmost (min, argmin) of every block [b*bs, min(n, (b+1)*bs)).
// Replaces kernel_min_blocks (src/device_tools.cuh:243-253, one thread per
// block) and compute_min_blocks (src/rtx.h:3-22, which also overreads the last
// partial block): here one workgroup per block, float4-coalesced.
__global__ __launch_bounds__(kWG) void k_block_minloc(const float *__restrict__ A, int n, int bs, int nb,
                                                      Pair *__restrict__ out) {
    __shared__ float sv[kWavesPerWG];
    __shared__ int si[kWavesPerWG];
    const int tid = threadIdx.x;
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int lo = (int)((int64_t)b * bs);
        const int hi = (int)min((int64_t)n - 1, (int64_t)lo + bs - 1);
        float bv = INFINITY;
        int bi = INT_MAX;
        scan_range<kWG, 2>(A, lo, hi, tid, bv, bi);
        wave_reduce(bv, bi);
        if ((tid & (kWave - 1)) == 0) {
            sv[tid / kWave] = bv;
            si[tid / kWave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
#pragma unroll
            for (int w = 1; w < kWavesPerWG; ++w) take_if_better(sv[w], si[w], bv, bi);
            if (bv == INFINITY) bi = lo;
            out[b] = Pair{bv, bi};
        }
        __syncthreads();
    }
}

// Small blocks (bs <= 64): one lane per block, strict '<' in index order
// (leftmost); adjacent lanes read adjacent blocks.
__global__ __launch_bounds__(kWG) void k_block_minloc_lane(const float *__restrict__ A, int n, int bs, int nb,
                                                           Pair *__restrict__ out) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int lo = (int)(b * bs);
    const int hi = (int)min((int64_t)n - 1, (int64_t)lo + bs - 1);
    float bv = A[lo];
    int bi = lo;
    for (int i = lo + 1; i <= hi; ++i) {
        const float x = A[i];
        if (x < bv) {
            bv = x;
            bi = i;
        }
    }
    out[b] = Pair{bv, bi};
}

// Medium blocks (bs <= 8192): one wave per block.
__global__ __launch_bounds__(kWG) void k_block_minloc_wave(const float *__restrict__ A, int n, int bs, int nb,
                                                           Pair *__restrict__ out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t waves = (int64_t)gridDim.x * kWavesPerWG;
    for (int64_t b = (int64_t)blockIdx.x * kWavesPerWG + threadIdx.x / kWave; b < nb; b += waves) {
        const int lo = (int)(b * bs);
        const int hi = (int)min((int64_t)n - 1, (int64_t)lo + bs - 1);
        float bv = INFINITY;
        int bi = INT_MAX;
        scan_range<kWave, 2>(A, lo, hi, lane, bv, bi);
        wave_reduce(bv, bi);
        if (lane == 0) {
            if (bv == INFINITY) bi = lo;
            out[b] = Pair{bv, bi};
        }
    }
}

// Sparse-table level: cur[i] = leftmost min of prev[i], prev[i+half].  The two
// windows are disjoint and the left one holds the smaller indices, so the
// left wins ties.
__global__ void k_sparse_level(const Pair *__restrict__ prev, Pair *__restrict__ cur, int count, int half) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const Pair x = prev[i];
    const Pair y = prev[i + half];
    cur[i] = (y.v < x.v) ? y : x;
}

// Edge tables, step 1: 32 lanes per sub-block (32 elements, aligned to its
// block's start), one lane per element, so loads and stores are contiguous.
// Segmented inclusive scans across the 32 lanes give the prefix and suffix
// (min, argmin) inside the sub-block; the lexicographic (value, index) order
// keeps the leftmost of equal values in both directions.  Lane 0's suffix is
// the sub-block minimum.  Sub-blocks past a partial last block are empty and
// get the sentinel (they are never read by a valid query).
__global__ __launch_bounds__(kWG) void k_sub_edges(const float *__restrict__ A, int n, int bs, int nsub,
                                                   int64_t subs_total, Pair *__restrict__ pre_s,
                                                   Pair *__restrict__ suf_s, Pair *__restrict__ submin) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t g = t >> kSubShift;  // the same for all 32 lanes of a segment
    if (g >= subs_total) return;
    const int o = (int)(t & (kSub - 1));
    const int64_t b = g / nsub, j = g - b * nsub;
    const int64_t start = b * bs + j * kSub;
    const int64_t end = min(min(start + kSub, (b + 1) * bs), (int64_t)n);
    const int64_t i = start + o;
    const bool valid = i < end;
    const float x = valid ? A[i] : INFINITY;
    const int xi = valid ? (int)i : INT_MAX;
    float pv = x, sv = x;
    int pi = xi, si = xi;
#pragma unroll
    for (int d = 1; d < kSub; d <<= 1) {
        const float uv = __shfl_up(pv, d, kSub);
        const int ui = __shfl_up(pi, d, kSub);
        const float dv = __shfl_down(sv, d, kSub);
        const int di = __shfl_down(si, d, kSub);
        if (o >= d) take_if_better(uv, ui, pv, pi);
        if (o + d < kSub) take_if_better(dv, di, sv, si);
    }
    if (valid) {
        if (pre_s) pre_s[i] = Pair{pv, pi};
        if (suf_s) suf_s[i] = Pair{sv, si};
    }
    if (o == 0) submin[g] = Pair{sv, si};
}

// Step 2: one wave per block: exclusive prefix and suffix minima over the
// block's sub-block minima, 64 sub-blocks per step with a carry across steps.
__global__ __launch_bounds__(kWG) void k_block_sub_scan(const Pair *__restrict__ submin, int nb, int nsub,
                                                        Pair *__restrict__ excl_pre, Pair *__restrict__ excl_suf) {
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // uniform across the wave
    if (w >= nb) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t base = w * nsub;
    float cv = INFINITY;  // carry: minimum of the sub-blocks already passed
    int ci = INT_MAX;
    for (int c = 0; c < nsub; c += kWave) {
        const int j = c + lane;
        const Pair x = j < nsub ? submin[base + j] : Pair{INFINITY, INT_MAX};
        float v = x.v;
        int i = x.i;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const float ov = __shfl_up(v, d);
            const int oi = __shfl_up(i, d);
            if (lane >= d) take_if_better(ov, oi, v, i);
        }
        float ev = __shfl_up(v, 1);
        int ei = __shfl_up(i, 1);
        if (lane == 0) {
            ev = INFINITY;
            ei = INT_MAX;
        }
        take_if_better(cv, ci, ev, ei);
        if (j < nsub) excl_pre[base + j] = Pair{ev, ei};
        take_if_better(__shfl(v, kWave - 1), __shfl(i, kWave - 1), cv, ci);
    }
    cv = INFINITY;
    ci = INT_MAX;
    for (int c = (nsub - 1) / kWave * kWave; c >= 0; c -= kWave) {
        const int j = c + lane;
        const Pair x = j < nsub ? submin[base + j] : Pair{INFINITY, INT_MAX};
        float v = x.v;
        int i = x.i;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const float ov = __shfl_down(v, d);
            const int oi = __shfl_down(i, d);
            if (lane + d < kWave) take_if_better(ov, oi, v, i);
        }
        float ev = __shfl_down(v, 1);
        int ei = __shfl_down(i, 1);
        if (lane == kWave - 1) {
            ev = INFINITY;
            ei = INT_MAX;
        }
        take_if_better(cv, ci, ev, ei);
        if (j < nsub) excl_suf[base + j] = Pair{ev, ei};
        take_if_better(__shfl(v, 0), __shfl(i, 0), cv, ci);
    }
}

// Step 3: one thread per element: block-level prefix = (sub-blocks before,
// prefix inside its sub-block); block-level suffix = (suffix inside its
// sub-block, sub-blocks after).  The earlier part wins ties.
__global__ void k_block_edges(int n, int bs, BlockDiv div, int nsub, const Pair *__restrict__ pre_s,
                              const Pair *__restrict__ suf_s, const Pair *__restrict__ excl_pre,
                              const Pair *__restrict__ excl_suf, Pair *__restrict__ pre_b, Pair *__restrict__ suf_b) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)t;
    const int b = block_of(i, div);
    const int64_t g = (int64_t)b * nsub + ((i - b * bs) >> kSubShift);
    Pair p = excl_pre[g];
    const Pair x = pre_s[i];
    take_if_better(x.v, x.i, p.v, p.i);
    pre_b[i] = p;
    Pair sf = suf_s[i];
    const Pair e = excl_suf[g];
    take_if_better(e.v, e.i, sf.v, sf.i);
    suf_b[i] = sf;
}

// Every query of a handle with edge tables, in one pass over the query
// array: each wave loads a tile of 64 queries once, answers the ranges that
// are not short (load_tile_edges) one lane per query by O(1) lookups
// (answer_edges: no scan, no reduction), then the short ranges with G-lane
// group scans (answer_class), and writes (NaN, -1) for invalid queries.
// TWO: the phased superset test and failure path for tables beyond the
// caches (answer_two_stage).
template <int G, bool TWO>
__global__ __launch_bounds__(kWG) RMQ_SGPR_BUDGET void k_query_edges(Blocks k, EdgeTables e, const int2 *__restrict__ Q, int64_t q,
                                                     int short_max, float *__restrict__ out_v,
                                                     int32_t *__restrict__ out_i) {
    __shared__ TileStage stage[kWavesPerWG];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int64_t n_tiles = (q + kWave - 1) / kWave;
    const int64_t wave_stride = (int64_t)gridDim.x * kWavesPerWG;
    for (int64_t t = (int64_t)blockIdx.x * kWavesPerWG + wv; t < n_tiles; t += wave_stride) {
        const TileMasks m = load_tile_edges(Q, q, k, short_max, t, lane, out_v, out_i);
        if ((m.longs >> lane) & 1) {
            const Pair best = answer_edges<true, false, TWO>(e, k, m.l, m.r);
            write_result(out_v, out_i, t * kWave + lane, best.v, best.i);
        }
        if (m.shorts) answer_class<G, false>(k, m.shorts, m, lane, t * kWave, out_v, out_i, stage[wv]);
    }
}

// k_query_edges for a handle with a coarse row (RMQ_BUILD_MID): workgroups of
// kLdsWaves waves, two per CU, first copy the coarse row into LDS; every
// long range then tries the coarse superset test (answer_edges_mid) before
// the global table.
template <int G, bool TWO>
__global__ __launch_bounds__(kLdsWaves *kWave) RMQ_SGPR_BUDGET void k_query_edges_mid(
    Blocks k, EdgeTables e, const Pair *__restrict__ mid_g, int mid_shift, int mid_nb, const int2 *__restrict__ Q,
    int64_t q, int short_max, float *__restrict__ out_v, int32_t *__restrict__ out_i) {
    __shared__ Pair mid[kMidMaxBlocks];
    __shared__ TileStage stage[kLdsWaves];
    for (int j = threadIdx.x; j < mid_nb; j += kLdsWaves * kWave) mid[j] = mid_g[j];
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int64_t n_tiles = (q + kWave - 1) / kWave;
    const int64_t wave_stride = (int64_t)gridDim.x * kLdsWaves;
    for (int64_t t = (int64_t)blockIdx.x * kLdsWaves + wv; t < n_tiles; t += wave_stride) {
        const TileMasks m = load_tile_edges(Q, q, k, short_max, t, lane, out_v, out_i);
        if ((m.longs >> lane) & 1) {
            const Pair best = answer_edges_mid<true, false, TWO>(e, k, mid, mid_shift, m.l, m.r);
            write_result(out_v, out_i, t * kWave + lane, best.v, best.i);
        }
        if (m.shorts) answer_class<G, false>(k, m.shorts, m, lane, t * kWave, out_v, out_i, stage[wv]);
    }
}

// The coarse row of a handle (RMQ_BUILD_MID): entry i is the leftmost minimum
// of coarse blocks i and i+1 (the last entry, of the last coarse block), that
// is of blocks [i << shift, min(nb, (i + 2) << shift)), from the block table.
__global__ void k_mid_rows(const Pair *__restrict__ st, int nb, int shift, int mid_nb, Pair *__restrict__ mid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mid_nb) return;
    const int64_t a = (int64_t)i << shift;
    const int64_t end = (int64_t)(i + 2) << shift;
    const int64_t b = end < nb ? end : nb;
    mid[i] = table_min(st, (uint32_t)nb, a, (int)(b - a));
}

// One tile of a handle with edge tables, its block sparse table in LDS: the
// answer of this lane's query in registers -- (NaN, -1) for an invalid query,
// the O(1) edge-table answer for a long range, and for a scanned range the
// result its lane group leaves in the wave's LDS slots.  Writes nothing to
// global memory.
template <int G>
__device__ __forceinline__ Pair answer_tile_lds(const Blocks &kl, const EdgeTables &e, int2 p, int64_t q, int64_t t,
                                                int short_max, int lane, TileStage &stage, TileResults &res) {
    const TileMasks m = classify_edges<true, false>(p, q, kl, short_max, t, lane, nullptr, nullptr);
    Pair a{__int_as_float(0x7FC00000), -1};
    if (m.own_long) a = answer_edges<false, true>(e, kl, m.l, m.r);
    if (m.shorts) {
        answer_class<G, false, true, true>(kl, m.shorts, m, lane, 0, nullptr, nullptr, stage, &res);
        if (m.own_short) a = Pair{res.v[lane], res.i[lane]};
    }
    return a;
}

// The edge-table pass with the block sparse table copied into LDS first, for
// a table that fits: the superset test and the edge blocks' (min, argmin)
// then read LDS, and the vector-memory path (whose address and L1 stages
// bound the global-table kernel: two divergent table reads per query)
// carries only the queries, the results, edge pairs and scans.  Workgroups of
// kLdsWaves waves, two resident per CU.  With its reads mostly from LDS the
// pass streams queries and results, so it is pipelined on memory latency:
// each wave keeps the queries of its next two tiles in flight, and every
// full tile ends in exactly one store per output (OUT: 1 values, 2 indexes,
// 3 both; scanned ranges' answers go through LDS to their own lanes), so the
// compiler's wait before a tile's queries are used leaves the later loads and
// stores outstanding.  The loop is unrolled three times with the three query
// registers in rotation: written as a copy (c0 = c1; c1 = c2) at the back
// edge, the copy of a register whose load is in flight waits for that load,
// which left only one tile's work to cover each load.  The partial last tile
// is answered after the loop.
template <int G, int OUT>
__global__ __launch_bounds__(kLdsWaves *kWave) void k_query_edges_lds(Blocks k, EdgeTables e, int table_pairs,
                                                                     const int2 *__restrict__ Q, int64_t q,
                                                                     int short_max, float *__restrict__ out_v,
                                                                     int32_t *__restrict__ out_i) {
    extern __shared__ Pair lds_table[];
    __shared__ TileStage stage[kLdsWaves];
    __shared__ TileResults res[kLdsWaves];
    {
        // four loads in flight per thread, then their LDS writes: one memory
        // round trip per 4 x 1024 pairs rather than per 1024
        constexpr int kThreads = kLdsWaves * kWave, kCopyU = 4;
        for (int j0 = threadIdx.x; j0 < table_pairs; j0 += kCopyU * kThreads) {
            Pair x[kCopyU];
#pragma unroll
            for (int u = 0; u < kCopyU; ++u)
                if (j0 + u * kThreads < table_pairs) x[u] = k.st[j0 + u * kThreads];
#pragma unroll
            for (int u = 0; u < kCopyU; ++u)
                if (j0 + u * kThreads < table_pairs) lds_table[j0 + u * kThreads] = x[u];
        }
    }
    __syncthreads();
    Blocks kl = k;
    kl.st = lds_table;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int64_t full = q / kWave;
    const int64_t stride = (int64_t)gridDim.x * kLdsWaves;
    int64_t t = (int64_t)blockIdx.x * kLdsWaves + wv;
    if (t < full) {
        // tiles t, t+stride, t+2*stride... ; loads past the last full tile
        // re-read that tile (always in bounds), so every step issues exactly
        // one load and its count never depends on a branch
        const int64_t last = full - 1;
        auto ahead = [&](int64_t u) { return load_query(Q, min(u, last) * kWave + lane); };
        auto tile = [&](int2 c, int64_t u) {
            const Pair a = answer_tile_lds<G>(kl, e, c, q, u, short_max, lane, stage[wv], res[wv]);
            if constexpr (OUT & 1) __builtin_nontemporal_store(a.v, out_v + u * kWave + lane);
            if constexpr (OUT & 2) __builtin_nontemporal_store(a.i, out_i + u * kWave + lane);
        };
        int2 c0 = ahead(t), c1 = ahead(t + stride), c2;
        do {
            // two steps ahead of the loop, so that it is entered, as at its
            // back edge, with a full step's stores and load behind each
            // query register: the wait before a register's first use is
            // then the same on both paths into the loop and keeps the next
            // two tiles' loads in flight
            c2 = ahead(t + 2 * stride);
            tile(c0, t);
            if ((t += stride) >= full) break;
            c0 = ahead(t + 2 * stride);
            tile(c1, t);
            if ((t += stride) >= full) break;
            for (;;) {
                c1 = ahead(t + 2 * stride);
                tile(c2, t);
                if ((t += stride) >= full) break;
                c2 = ahead(t + 2 * stride);
                tile(c0, t);
                if ((t += stride) >= full) break;
                c0 = ahead(t + 2 * stride);
                tile(c1, t);
                if ((t += stride) >= full) break;
            }
        } while (false);
    }
    if (t == full && (q & (kWave - 1))) {  // the partial last tile falls to exactly this wave
        const Pair a = answer_tile_lds<G>(kl, e, fetch_query(Q, q, t, lane), q, t, short_max, lane, stage[wv],
                                          res[wv]);
        if (t * kWave + lane < q) {
            if constexpr (OUT & 1) __builtin_nontemporal_store(a.v, out_v + t * kWave + lane);
            if constexpr (OUT & 2) __builtin_nontemporal_store(a.i, out_i + t * kWave + lane);
        }
    }
}

// Tile kernel: each wave takes tiles of 64 consecutive queries and answers
// the ones of its class (LONG = more than short_max elements) with
// answer_class.  The short-class kernel also writes (NaN, -1) for invalid
// queries.  Persistent grid, static stride over tiles.
template <int G, bool LONG>
__global__ __launch_bounds__(kWG) RMQ_SGPR_BUDGET void k_query_tiles(Blocks k, const int2 *__restrict__ Q, int64_t q, int short_max,
                                                     float *__restrict__ out_v, int32_t *__restrict__ out_i) {
    __shared__ TileStage stage[kWavesPerWG];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int64_t n_tiles = (q + kWave - 1) / kWave;
    const int64_t wave_stride = (int64_t)gridDim.x * kWavesPerWG;
    for (int64_t t = (int64_t)blockIdx.x * kWavesPerWG + wv; t < n_tiles; t += wave_stride) {
        const TileMasks m = load_tile(Q, q, k.n, short_max, t, lane, !LONG, out_v, out_i);
        // short ranges: up to six float4 loads per lane in flight, each under
        // its own bounds test, so a range of up to 6*4*G elements is one
        // round trip for its lane group
        answer_class<G, LONG, true, false, 6, true>(k, LONG ? m.longs : m.shorts, m, lane, t * kWave, out_v, out_i,
                                                    stage[wv]);
    }
}

// ---------------------------------------------------------------------------
// brute force on the device (algs 2/102, --check): src/cuda_methods.cuh:3-66
// ---------------------------------------------------------------------------
__global__ void k_gpu_base(const float *__restrict__ A, const int2 *__restrict__ Q, int64_t q, int n,
                           float *__restrict__ out_v, int32_t *__restrict__ out_i) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= q) return;
    const int2 p = Q[tid];
    if (p.x < 0 || p.x > p.y || p.y >= n) {
        write_result(out_v, out_i, tid, __int_as_float(0x7FC00000), -1);
        return;
    }
    float m = A[p.x];
    int mi = p.x;
    for (int i = p.x; i <= p.y; ++i) {
        const float v = A[i];
        if (v < m) {
            m = v;
            mi = i;
        }
    }
    write_result(out_v, out_i, tid, m, mi);
}

__global__ void k_gather_stride(const int2 *__restrict__ Q, int64_t step, int64_t m, int2 *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) out[k] = Q[k * step];
}

// Value signature of an array for the tuner's cache key: one wave per
// window of w <= 64 elements starting at k * step; writes the leftmost
// minimum's offset in the window and how many elements equal that minimum.
__global__ __launch_bounds__(kWG) void k_window_sig(const float *__restrict__ A, int w, int64_t step, int64_t m,
                                                    int2 *__restrict__ out) {
    const int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // uniform across the wave
    if (k >= m) return;
    const int lane = threadIdx.x & (kWave - 1);
    const float x = lane < w ? A[k * step + lane] : INFINITY;
    float v = x;
    int i = lane < w ? lane : INT_MAX;
    wave_reduce(v, i);
    const uint64_t eq = __ballot(lane < w && x == v);
    if (lane == 0) out[k] = make_int2(i, __popcll(eq));
}

__global__ void k_count_bad(const int2 *__restrict__ Q, int64_t q, int64_t n, unsigned long long *bad) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool b = false;
    if (tid < q) {
        const int2 p = Q[tid];
        b = p.x < 0 || p.x > p.y || (int64_t)p.y >= n;
    }
    const uint64_t m = __ballot(b);
    if (m && (threadIdx.x & (kWave - 1)) == 0) atomicAdd(bad, (unsigned long long)__popcll(m));
}

// ---------------------------------------------------------------------------
// generators (same definitions on host and device)
// ---------------------------------------------------------------------------
constexpr uint64_t kPhi = 0x9E3779B97F4A7C15ULL;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}
__host__ __device__ __forceinline__ uint64_t rng_key(uint64_t seed, uint64_t stream) { return mix64(seed * kPhi + stream); }
__host__ __device__ __forceinline__ uint64_t rng_at(uint64_t key, uint64_t i) { return mix64(key + (i + 1) * kPhi); }

__host__ __device__ __forceinline__ float gen_value(uint64_t key, int64_t i, int kind) {
    switch (kind) {
    case RMQ_GEN_UNIFORM: return (float)(rng_at(key, (uint64_t)i) >> 40) * 0x1p-24f;
    case RMQ_GEN_TIES16: return (float)(rng_at(key, (uint64_t)i) >> 60);
    case RMQ_GEN_DESC: {
        const uint32_t bits = 0x3F800000u - 1u - (uint32_t)i;
        float f;
        memcpy(&f, &bits, 4);
        return f;
    }
    default: return 1.0f;
    }
}

__global__ void k_gen_array(float *__restrict__ A, int64_t n, uint64_t key, int kind) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) A[i] = gen_value(key, i, kind);
}

}  // namespace rmqx

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

namespace rmqx {

int tiles_grid(int64_t q, int per_cu) {
    const int64_t tiles = (q + kWave - 1) / kWave;
    const int64_t want = (tiles + kWavesPerWG - 1) / kWavesPerWG;
    return (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)device_cus() * per_cu));
}

template <bool LONG>
int launch_tiles(const rmq_handle *h, const int2 *Q, int64_t q, int lanes, int grid, int short_max, float *val,
                 int32_t *idx, hipStream_t s) {
    const Blocks k = blocks_of(h);
    switch (lanes) {
    case 4: k_query_tiles<4, LONG><<<grid, kWG, 0, s>>>(k, Q, q, short_max, val, idx); break;
    case 8: k_query_tiles<8, LONG><<<grid, kWG, 0, s>>>(k, Q, q, short_max, val, idx); break;
    case 16: k_query_tiles<16, LONG><<<grid, kWG, 0, s>>>(k, Q, q, short_max, val, idx); break;
    default: k_query_tiles<64, LONG><<<grid, kWG, 0, s>>>(k, Q, q, short_max, val, idx); break;
    }
    return hip_rc(hipGetLastError());
}

// With edge tables the long pass answers the short ranges too (k_query_edges),
// so there is no separate short pass.
int launch_short(const rmq_handle *h, const int2 *Q, int64_t q, const rmq_query_opts_t &o, float *val, int32_t *idx,
                 hipStream_t s) {
    if (h->edges.pre_b) return RMQ_OK;
    return launch_tiles<false>(h, Q, q, o.group_lanes, tiles_grid(q, 16), o.short_max, val, idx, s);
}

// Whether one workgroup of k_query_edges_lds with `bytes` of LDS fits twice
// in a CU of the current device (full occupancy) and within the
// per-workgroup limit.  The limits are read once per device.
bool lds_table_fits(size_t bytes) {
    constexpr int kMaxDevices = 64;
    static std::atomic<int64_t> limits[kMaxDevices];  // (per_block << 32) | per_cu; 0 = not read yet
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return false;
    int64_t v = limits[dev].load(std::memory_order_relaxed);
    if (v == 0) {
        int per_block = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
            hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess)
            return false;
        v = ((int64_t)per_block << 32) | (uint32_t)per_cu;
        limits[dev].store(v, std::memory_order_relaxed);
    }
    const size_t per_block = (size_t)(v >> 32), per_cu = (size_t)(uint32_t)v;
    return bytes <= per_block && 2 * bytes <= per_cu;
}

// Which kernel answers a handle's long ranges (and, with edge tables, all of
// its ranges in one pass).
enum class LongPass { kTiles, kEdges, kEdgesLds };

size_t lds_table_bytes(const rmq_handle *h) { return (size_t)h->levels * (size_t)h->nb * sizeof(Pair); }

// Whether the global-table kernel takes the phased path of answer_two_stage:
// only for tables far beyond the caches, where its saved DRAM lines pay.
bool two_stage(const rmq_handle *h) { return h->st_bytes > kTwoStageTableBytes; }

LongPass long_pass(const rmq_handle *h) {
    if (!h->edges.pre_b) return LongPass::kTiles;
    return lds_table_fits(lds_table_bytes(h) + (sizeof(TileStage) + sizeof(TileResults)) * kLdsWaves) ? LongPass::kEdgesLds
                                                                                : LongPass::kEdges;
}

template <int G>
void launch_edges_lds(int out, int grid, size_t lds, hipStream_t s, const Blocks &k, const EdgeTables &e, int pairs,
                      const int2 *Q, int64_t q, int sm, float *val, int32_t *idx) {
    const dim3 wg(kLdsWaves * kWave);
    switch (out) {
    case 1: k_query_edges_lds<G, 1><<<grid, wg, lds, s>>>(k, e, pairs, Q, q, sm, val, idx); break;
    case 2: k_query_edges_lds<G, 2><<<grid, wg, lds, s>>>(k, e, pairs, Q, q, sm, val, idx); break;
    default: k_query_edges_lds<G, 3><<<grid, wg, lds, s>>>(k, e, pairs, Q, q, sm, val, idx); break;
    }
}

template <bool TWO>
void launch_edges_mid(int g, int grid, hipStream_t s, const rmq_handle *h, const Blocks &k, const int2 *Q, int64_t q,
                      int sm, float *val, int32_t *idx) {
    const dim3 wg(kLdsWaves * kWave);
    const Pair *m = h->mid;
    const int sh = h->mid_shift, mn = h->mid_nb;
    switch (g) {
    case 4: k_query_edges_mid<4, TWO><<<grid, wg, 0, s>>>(k, h->edges, m, sh, mn, Q, q, sm, val, idx); break;
    case 8: k_query_edges_mid<8, TWO><<<grid, wg, 0, s>>>(k, h->edges, m, sh, mn, Q, q, sm, val, idx); break;
    case 16: k_query_edges_mid<16, TWO><<<grid, wg, 0, s>>>(k, h->edges, m, sh, mn, Q, q, sm, val, idx); break;
    default: k_query_edges_mid<64, TWO><<<grid, wg, 0, s>>>(k, h->edges, m, sh, mn, Q, q, sm, val, idx); break;
    }
}

int launch_long(const rmq_handle *h, const int2 *Q, int64_t q, const rmq_query_opts_t &o, float *val, int32_t *idx,
                hipStream_t s) {
    int grid = tiles_grid(q, 8);
    if (o.long_wgs > 0) grid = (int)std::min<int64_t>(o.long_wgs, grid);
    if (h->edges.pre_b) {
        const Blocks k = blocks_of(h);
        const int sm = o.short_max;
        if (long_pass(h) == LongPass::kEdgesLds) {
            const size_t lds = lds_table_bytes(h);
            const int64_t tiles = (q + kWave - 1) / kWave;
            int g2 = (int)std::max<int64_t>(1, std::min<int64_t>((tiles + kLdsWaves - 1) / kLdsWaves,
                                                                 (int64_t)device_cus() * 2));
            if (o.long_wgs > 0) g2 = std::min(g2, o.long_wgs);
            const int p = (int)(lds / sizeof(Pair));
            const int out = (val ? 1 : 0) | (idx ? 2 : 0);
            switch (o.group_lanes) {
            case 4: launch_edges_lds<4>(out, g2, lds, s, k, h->edges, p, Q, q, sm, val, idx); break;
            case 8: launch_edges_lds<8>(out, g2, lds, s, k, h->edges, p, Q, q, sm, val, idx); break;
            case 16: launch_edges_lds<16>(out, g2, lds, s, k, h->edges, p, Q, q, sm, val, idx); break;
            default: launch_edges_lds<64>(out, g2, lds, s, k, h->edges, p, Q, q, sm, val, idx); break;
            }
            return hip_rc(hipGetLastError());
        }
        if (h->mid) {
            const int64_t tiles = (q + kWave - 1) / kWave;
            int g2 = (int)std::max<int64_t>(1, std::min<int64_t>((tiles + kLdsWaves - 1) / kLdsWaves,
                                                                 (int64_t)device_cus() * 2));
            if (o.long_wgs > 0) g2 = std::min(g2, o.long_wgs);
            if (two_stage(h))
                launch_edges_mid<true>(o.group_lanes, g2, s, h, k, Q, q, sm, val, idx);
            else
                launch_edges_mid<false>(o.group_lanes, g2, s, h, k, Q, q, sm, val, idx);
            return hip_rc(hipGetLastError());
        }
        if (two_stage(h)) {
            switch (o.group_lanes) {
            case 4: k_query_edges<4, true><<<grid, kWG, 0, s>>>(k, h->edges, Q, q, sm, val, idx); break;
            case 8: k_query_edges<8, true><<<grid, kWG, 0, s>>>(k, h->edges, Q, q, sm, val, idx); break;
            case 16: k_query_edges<16, true><<<grid, kWG, 0, s>>>(k, h->edges, Q, q, sm, val, idx); break;
            default: k_query_edges<64, true><<<grid, kWG, 0, s>>>(k, h->edges, Q, q, sm, val, idx); break;
            }
        } else {
            switch (o.group_lanes) {
            case 4: k_query_edges<4, false><<<grid, kWG, 0, s>>>(k, h->edges, Q, q, sm, val, idx); break;
            case 8: k_query_edges<8, false><<<grid, kWG, 0, s>>>(k, h->edges, Q, q, sm, val, idx); break;
            case 16: k_query_edges<16, false><<<grid, kWG, 0, s>>>(k, h->edges, Q, q, sm, val, idx); break;
            default: k_query_edges<64, false><<<grid, kWG, 0, s>>>(k, h->edges, Q, q, sm, val, idx); break;
            }
        }
        return hip_rc(hipGetLastError());
    }
    return launch_tiles<true>(h, Q, q, o.long_lanes, grid, o.short_max, val, idx, s);
}

int query_impl(const rmq_handle *h, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx,
               const rmq_query_opts_t *opts_in, hipStream_t s, rmq_query_times_t *times) {
    if (!h || q < 0 || q > INT32_MAX) return RMQ_ERR_INVALID_ARG;
    rmq_query_opts_t o;
    int rc = resolve_opts(h, opts_in, &o);
    if (rc != RMQ_OK) return rc;
    if (q == 0) {  // nothing to read or write: empty buffers may be NULL
        if (times) *times = rmq_query_times_t{0.f, 0.f, 0.f};
        return RMQ_OK;
    }
    if (!d_Q || (!d_val && !d_idx)) return RMQ_ERR_INVALID_ARG;
    rc = check_device(h);
    if (rc != RMQ_OK) return rc;
    const int2 *Q = reinterpret_cast<const int2 *>(d_Q);
    if (!times) {
        rc = launch_short(h, Q, q, o, d_val, d_idx, s);
        if (rc == RMQ_OK) rc = launch_long(h, Q, q, o, d_val, d_idx, s);
        return rc;
    }
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    for (auto &e : ev)
        if ((rc = hip_rc(hipEventCreate(&e))) != RMQ_OK) break;
    if (rc == RMQ_OK) rc = hip_rc(hipEventRecord(ev[0], s));
    if (rc == RMQ_OK) rc = launch_short(h, Q, q, o, d_val, d_idx, s);
    if (rc == RMQ_OK) rc = hip_rc(hipEventRecord(ev[1], s));
    if (rc == RMQ_OK) rc = launch_long(h, Q, q, o, d_val, d_idx, s);
    if (rc == RMQ_OK) rc = hip_rc(hipEventRecord(ev[2], s));
    if (rc == RMQ_OK) rc = hip_rc(hipEventSynchronize(ev[2]));
    if (rc == RMQ_OK) {
        float a = 0.f, b = 0.f;
        (void)hipEventElapsedTime(&a, ev[0], ev[1]);
        (void)hipEventElapsedTime(&b, ev[1], ev[2]);
        *times = rmq_query_times_t{a, b, a + b};
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

// Random-request ceiling: each thread reads two 8-byte entries at hashed
// positions of two tables (masked to the table size, so always in bounds)
// and writes one 8-byte result -- the access pattern of a long range answered
// from edge tables, with no RMQ logic around it.
__global__ void k_probe_gather(const uint2 *__restrict__ t0, const uint2 *__restrict__ t1, uint64_t mask, int64_t m,
                               uint2 *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t h0 = mix64((uint64_t)i), h1 = mix64(h0);
        const uint2 a = t0[h0 & mask], b = t1[h1 & mask];
        out[i] = make_uint2(a.x ^ b.x, a.y + b.y);
    }
}

}  // namespace rmqx

using namespace rmqx;

// ===========================================================================
// C ABI
// ===========================================================================

namespace {

// Rows 1 .. levels_s-1 of the sparse table over the subs sub-block minima in
// row 0 (one flat sequence: windows may cross block boundaries).
int build_sub_levels(Pair *st_s, int64_t subs, int levels_s, hipStream_t s) {
    int rc = RMQ_OK;
    for (int k = 1; rc == RMQ_OK && k < levels_s; ++k) {
        const int64_t count = subs - ((int64_t)1 << k) + 1;
        k_sparse_level<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(st_s + (int64_t)(k - 1) * subs,
                                                                      st_s + (int64_t)k * subs, (int)count,
                                                                      1 << (k - 1));
        rc = hip_rc(hipGetLastError());
    }
    return rc;
}

// Edge tables in one allocation (handle_sizes): pre_b, suf_b (8 B per
// element each), and when bs > 32 also pre_s, suf_s (the same inside
// sub-blocks), the sub-block sparse table and the two per-sub-block
// exclusive scans.
int build_edges(rmq_handle *h, const HandleSizes &z, hipStream_t s) {
    const int64_t n = h->n;
    const int bs = h->bs;
    const int nsub = z.nsub;
    const int64_t subs = z.subs;
    const int levels_s = z.levels_s;
    const bool two_level = bs > kSub;
    const int64_t per_elem = two_level ? 4 : 2;
    const int64_t bytes = z.edge_bytes;
    int rc = hip_rc(hipMalloc(&h->edge_mem, (size_t)bytes));
    if (rc != RMQ_OK) return rc;
    h->edge_bytes = bytes;
    Pair *p = static_cast<Pair *>(h->edge_mem);
    Pair *pre_b = p, *suf_b = p + n;
    Pair *pre_s = two_level ? p + 2 * n : pre_b;
    Pair *suf_s = two_level ? p + 3 * n : suf_b;
    Pair *st_s = p + per_elem * n;
    Pair *excl_pre = st_s + (int64_t)levels_s * subs;
    Pair *excl_suf = excl_pre + subs;
    k_sub_edges<<<(unsigned)((subs * kSub + kWG - 1) / kWG), kWG, 0, s>>>(h->A, (int)n, bs, nsub, subs, pre_s, suf_s,
                                                                         st_s);
    rc = hip_rc(hipGetLastError());
    if (rc == RMQ_OK) rc = build_sub_levels(st_s, subs, levels_s, s);
    if (rc == RMQ_OK && two_level) {
        k_block_sub_scan<<<(unsigned)((h->nb + kWavesPerWG - 1) / kWavesPerWG), kWG, 0, s>>>(st_s, (int)h->nb, nsub,
                                                                                           excl_pre, excl_suf);
        rc = hip_rc(hipGetLastError());
        if (rc == RMQ_OK) {
            k_block_edges<<<(unsigned)((n + 255) / 256), 256, 0, s>>>((int)n, bs, block_div(bs), nsub, pre_s, suf_s,
                                                                      excl_pre, excl_suf, pre_b, suf_b);
            rc = hip_rc(hipGetLastError());
        }
    }
    h->edges = EdgeTables{pre_b, suf_b, pre_s, suf_s, st_s, subs, nsub, levels_s};
    return rc;
}

}  // namespace

RMQ_API int rmq_build(const float *d_A, int64_t n, int32_t bs, uint32_t flags, hipStream_t s, rmq_handle_t **out) {
    if (!out) return RMQ_ERR_INVALID_ARG;
    *out = nullptr;
    if (!d_A || n <= 0 || n > INT32_MAX || bs <= 0 || !build_flags_ok(flags) || (reinterpret_cast<uintptr_t>(d_A) & 15))
        return RMQ_ERR_INVALID_ARG;
    const HandleSizes z = handle_sizes(n, bs, flags);
    const int64_t nb = z.nb;
    const int32_t levels = z.levels;

    rmq_handle *h = new (std::nothrow) rmq_handle{};
    if (!h) return RMQ_ERR_OOM;
    h->A = d_A;
    h->n = n;
    h->bs = bs;
    h->nb = nb;
    h->levels = levels;
    h->st_bytes = z.st_bytes;
    int rc = hip_rc(hipGetDevice(&h->device));
    if (rc == RMQ_OK) rc = hip_rc(hipMalloc(&h->st, (size_t)h->st_bytes));
    if (rc == RMQ_OK) {
        if (bs <= 64) {
            k_block_minloc_lane<<<(unsigned)((nb + kWG - 1) / kWG), kWG, 0, s>>>(d_A, (int)n, bs, (int)nb, h->st);
        } else if (bs <= 8192) {
            const int grid = (int)std::min<int64_t>((nb + kWavesPerWG - 1) / kWavesPerWG, (int64_t)device_cus() * 16);
            k_block_minloc_wave<<<grid, kWG, 0, s>>>(d_A, (int)n, bs, (int)nb, h->st);
        } else {
            const int grid = (int)std::min<int64_t>(nb, 65535);
            k_block_minloc<<<grid, kWG, 0, s>>>(d_A, (int)n, bs, (int)nb, h->st);
        }
        rc = hip_rc(hipGetLastError());
    }
    for (int32_t k = 1; rc == RMQ_OK && k < levels; ++k) {
        const int half = 1 << (k - 1);
        const int count = (int)(nb - ((int64_t)1 << k) + 1);
        k_sparse_level<<<(unsigned)(((int64_t)count + 255) / 256), 256, 0, s>>>(h->st + (int64_t)(k - 1) * nb, h->st + (int64_t)k * nb,
                                                           count, half);
        rc = hip_rc(hipGetLastError());
    }
    h->flags = flags;
    if (rc == RMQ_OK && (flags & RMQ_BUILD_EDGES)) rc = build_edges(h, z, s);
    if (rc == RMQ_OK && (flags & RMQ_BUILD_MID)) {
        h->mid_shift = z.mid_shift;
        h->mid_nb = z.mid_nb;
        h->mid_bytes = z.mid_bytes;
        rc = hip_rc(hipMalloc(&h->mid, (size_t)z.mid_bytes));
        if (rc == RMQ_OK) {
            k_mid_rows<<<(unsigned)((z.mid_nb + 255) / 256), 256, 0, s>>>(h->st, (int)nb, z.mid_shift, z.mid_nb,
                                                                        h->mid);
            rc = hip_rc(hipGetLastError());
        }
    }
    // the sparse table over sub-block minima (32 elements, restarting at
    // every block start; row 0 the minima) for long scans: the edge tables'
    // when built, else an allocation of its own, 8 bytes per 32 elements and
    // level
    if (rc == RMQ_OK && bs > kSub) {
        h->nsub = z.nsub;
        if (h->edges.st_s) {
            h->sub = const_cast<Pair *>(h->edges.st_s);
        } else {
            const int64_t subs = z.subs;
            const int levels_s = z.levels_s;
            h->sub_bytes = z.sub_bytes;
            rc = hip_rc(hipMalloc(&h->sub_mem, (size_t)h->sub_bytes));
            if (rc == RMQ_OK) {
                h->sub = static_cast<Pair *>(h->sub_mem);
                k_sub_edges<<<(unsigned)((subs * kSub + kWG - 1) / kWG), kWG, 0, s>>>(d_A, (int)n, bs, h->nsub, subs,
                                                                                     nullptr, nullptr, h->sub);
                rc = hip_rc(hipGetLastError());
            }
            if (rc == RMQ_OK) rc = build_sub_levels(h->sub, subs, levels_s, s);
        }
    }
    if (rc != RMQ_OK) {
        if (h->st) (void)hipFree(h->st);
        if (h->edge_mem) (void)hipFree(h->edge_mem);
        if (h->sub_mem) (void)hipFree(h->sub_mem);
        if (h->mid) (void)hipFree(h->mid);
        delete h;
        return rc;
    }
    *out = h;
    return RMQ_OK;
}

RMQ_API int rmq_query(const rmq_handle_t *h, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx,
                      hipStream_t s) {
    return query_impl(h, d_Q, q, d_val, d_idx, nullptr, s, nullptr);
}

RMQ_API int rmq_query_ex(const rmq_handle_t *h, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx,
                         const rmq_query_opts_t *opts, hipStream_t s) {
    return query_impl(h, d_Q, q, d_val, d_idx, opts, s, nullptr);
}

RMQ_API int rmq_query_timed(const rmq_handle_t *h, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx,
                            const rmq_query_opts_t *opts, hipStream_t s, rmq_query_times_t *times) {
    if (!times) return RMQ_ERR_INVALID_ARG;
    return query_impl(h, d_Q, q, d_val, d_idx, opts, s, times);
}

RMQ_API int rmq_query_check(const int32_t *d_Q, int64_t q, int64_t n, hipStream_t s, int64_t *n_bad) {
    if (!n_bad || q < 0 || (q > 0 && !d_Q) || n <= 0) return RMQ_ERR_INVALID_ARG;
    *n_bad = 0;
    if (q == 0) return RMQ_OK;
    unsigned long long *d_bad = nullptr;
    RMQ_HIP(hipMalloc(&d_bad, sizeof(unsigned long long)));
    unsigned long long hb = 0;
    int rc = hip_rc(hipMemcpyAsync(d_bad, &hb, sizeof(hb), hipMemcpyHostToDevice, s));
    if (rc == RMQ_OK) {
        k_count_bad<<<(unsigned)((q + 255) / 256), 256, 0, s>>>(reinterpret_cast<const int2 *>(d_Q), q, n, d_bad);
        rc = hip_rc(hipGetLastError());
    }
    if (rc == RMQ_OK) rc = hip_rc(hipMemcpyAsync(&hb, d_bad, sizeof(hb), hipMemcpyDeviceToHost, s));
    if (rc == RMQ_OK) rc = hip_rc(hipStreamSynchronize(s));
    (void)hipFree(d_bad);
    if (rc != RMQ_OK) return rc;
    *n_bad = (int64_t)hb;
    return hb ? RMQ_ERR_BAD_QUERY : RMQ_OK;
}

RMQ_API void rmq_free(rmq_handle_t *h) {
    if (!h) return;
    if (h->st) (void)hipFree(h->st);
    if (h->edge_mem) (void)hipFree(h->edge_mem);
    if (h->sub_mem) (void)hipFree(h->sub_mem);
    if (h->mid) (void)hipFree(h->mid);
    delete h;
}

RMQ_API int rmq_handle_info(const rmq_handle_t *h, int64_t *n, int32_t *bs, int64_t *nb, int32_t *levels,
                            int64_t *bytes) {
    if (!h) return RMQ_ERR_INVALID_ARG;
    if (n) *n = h->n;
    if (bs) *bs = h->bs;
    if (nb) *nb = h->nb;
    if (levels) *levels = h->levels;
    if (bytes) *bytes = h->st_bytes + h->edge_bytes + h->sub_bytes + h->mid_bytes;
    return RMQ_OK;
}

RMQ_API int rmq_handle_bytes(int64_t n, int32_t bs, uint32_t flags, int64_t *bytes) {
    if (!bytes || n <= 0 || n > INT32_MAX || bs <= 0 || !build_flags_ok(flags)) return RMQ_ERR_INVALID_ARG;
    *bytes = handle_sizes(n, bs, flags).total();
    return RMQ_OK;
}

RMQ_API int rmq_query_kernels(const rmq_handle_t *h, const rmq_query_opts_t *opts, char *buf, int64_t len) {
    if (!h || !buf || len <= 0) return RMQ_ERR_INVALID_ARG;
    rmq_query_opts_t o;
    int rc = resolve_opts(h, opts, &o);
    if (rc != RMQ_OK) return rc;
    if ((rc = check_device(h)) != RMQ_OK) return rc;
    int w = 0;
    switch (long_pass(h)) {
    case LongPass::kEdgesLds: w = snprintf(buf, (size_t)len, "k_query_edges_lds<%d, 3>", o.group_lanes); break;
    case LongPass::kEdges:
        w = snprintf(buf, (size_t)len, h->mid ? "k_query_edges_mid<%d, %s>" : "k_query_edges<%d, %s>", o.group_lanes,
                     two_stage(h) ? "true" : "false");
        break;
    default:
        w = snprintf(buf, (size_t)len, "k_query_tiles<%d, false>;k_query_tiles<%d, true>", o.group_lanes,
                     o.long_lanes);
        break;
    }
    return (w < 0 || w >= len) ? RMQ_ERR_INVALID_ARG : RMQ_OK;
}

RMQ_API int rmq_query_opts_default(const rmq_handle_t *h, rmq_query_opts_t *opts) {
    if (!h || !opts) return RMQ_ERR_INVALID_ARG;
    fill_default_opts(h, opts);
    return RMQ_OK;
}

RMQ_API int rmq_gpu_base(const float *d_A, int64_t n, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx,
                         hipStream_t s) {
    if (!d_A || n <= 0 || n > INT32_MAX || q < 0 || q > INT32_MAX) return RMQ_ERR_INVALID_ARG;
    if (q == 0) return RMQ_OK;
    if (!d_Q || (!d_val && !d_idx)) return RMQ_ERR_INVALID_ARG;
    const int bsz = 256;  // the reference uses BSIZE=1024 (src/main.cu:14); wave64 prefers 256
    k_gpu_base<<<(unsigned)((q + bsz - 1) / bsz), bsz, 0, s>>>(d_A, reinterpret_cast<const int2 *>(d_Q), q, (int)n,
                                                               d_val, d_idx);
    return hip_rc(hipGetLastError());
}

RMQ_API int rmq_gen_array(float *d_A, int64_t n, uint64_t seed, int32_t kind, hipStream_t s) {
    if (!d_A || n <= 0 || kind < 0 || kind > 3) return RMQ_ERR_INVALID_ARG;
    if (kind == RMQ_GEN_DESC && n > 0x3F7FFFFFLL) return RMQ_ERR_INVALID_ARG;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)device_cus() * 32);
    k_gen_array<<<grid, 256, 0, s>>>(d_A, n, rng_key(seed, 0), kind);
    return hip_rc(hipGetLastError());
}

RMQ_API int rmq_gen_array_host(float *A, int64_t n, uint64_t seed, int32_t kind) {
    if (!A || n <= 0 || kind < 0 || kind > 3) return RMQ_ERR_INVALID_ARG;
    if (kind == RMQ_GEN_DESC && n > 0x3F7FFFFFLL) return RMQ_ERR_INVALID_ARG;
    const uint64_t key = rng_key(seed, 0);
    for (int64_t i = 0; i < n; ++i) A[i] = gen_value(key, i, kind);
    return RMQ_OK;
}

namespace rmqx {

// One query stream: counter-based, the same definition on host and device.
struct QueryGen {
    uint64_t k1, k2, k3, k4;
    double mu, sigma;
    int32_t n, lr;
};

// lr codes (after src/rand.cuh:179-201): > 0 fixed length; -1 uniform length;
// -2 / -3 lognormal medians 2^14 / 2^6, sigma 1; -4 / -5 the reference's
// scales n/2^8, n/2^15 with sigma 0.3; -6 r = n-1; -7 / -8 lognormal with the
// -2 / -3 medians scaled to n (median n^(14/log2 1e8), n^(6/log2 1e8)).
inline bool make_query_gen(int32_t n, int32_t lr, uint64_t seed, QueryGen *g) {
    if (n <= 0 || lr == 0 || lr > n || lr < -8) return false;
    g->k1 = rng_key(seed, 1);
    g->k2 = rng_key(seed, 2);
    g->k3 = rng_key(seed, 3);
    g->k4 = rng_key(seed, 4);
    g->n = n;
    g->lr = lr;
    g->mu = 0.0;
    g->sigma = 1.0;
    if (lr == -2) g->mu = 14.0 * 0.69314718055994530942;
    if (lr == -3) g->mu = 6.0 * 0.69314718055994530942;
    if (lr == -4 || lr == -5) {
        const int64_t sc = n / (lr == -4 ? 256 : 32768);
        g->mu = std::log((double)(sc > 1 ? sc : 1));
        g->sigma = 0.3;
    }
    if (lr == -7) g->mu = 14.0 / std::log2(1e8) * std::log((double)n);
    if (lr == -8) g->mu = 6.0 / std::log2(1e8) * std::log((double)n);
    return true;
}

__host__ __device__ inline int64_t below(uint64_t r, uint64_t m) {
    return (int64_t)(((unsigned __int128)r * m) >> 64);
}

// Box-Muller in double with contraction off, so every build rounds alike
// (host builds bit for bit; the device's libm may differ in the last ulp).
__host__ __device__ inline double lognormal_len(uint64_t k3, uint64_t k4, uint64_t i, double mu, double sigma) {
#pragma clang fp contract(off)
    const double u1 = (double)((rng_at(k3, i) >> 11) + 1) * 0x1p-53;  // (0,1]
    const double u2 = (double)(rng_at(k4, i) >> 11) * 0x1p-53;        // [0,1)
    const double rad = sqrt(-2.0 * log(u1));
    const double ang = 6.283185307179586476925286766559 * u2;
    const double z = rad * cos(ang);
    const double sz = sigma * z;
    const double e = mu + sz;
    return exp(e);
}

__host__ __device__ inline void gen_query(const QueryGen &g, uint64_t i, int32_t *out) {
    int64_t l, len;
    if (g.lr > 0) {
        len = g.lr;
        l = below(rng_at(g.k1, i), (uint64_t)(g.n - g.lr + 1));
    } else {
        l = below(rng_at(g.k1, i), (uint64_t)g.n);
        const int64_t room = (int64_t)g.n - l;
        if (g.lr == -1) {
            len = 1 + below(rng_at(g.k2, i), (uint64_t)room);
        } else if (g.lr == -6) {
            len = room;
        } else {
            double x = lognormal_len(g.k3, g.k4, i, g.mu, g.sigma);
            if (!(x < 4.0e18)) x = 4.0e18;
            len = llround(x);
            if (len < 1) len = 1;
            if (len > room) len = room;
        }
    }
    out[0] = (int32_t)l;
    out[1] = (int32_t)(l + len - 1);
}

__global__ void k_gen_queries(int32_t *__restrict__ Q, int64_t q, QueryGen g) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < q; i += stride)
        gen_query(g, (uint64_t)i, Q + 2 * i);
}

}  // namespace rmqx

RMQ_API int rmq_gen_queries_host(int32_t *Q, int64_t q, int32_t n, int32_t lr, uint64_t seed, int32_t nthreads) {
    QueryGen g;
    if (!Q || q < 0 || !make_query_gen(n, lr, seed, &g)) return RMQ_ERR_INVALID_ARG;
    auto work = [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) gen_query(g, (uint64_t)i, Q + 2 * i);
    };
    int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    if (q < 65536) nt = 1;
    if (nt == 1) {
        work(0, q);
        return RMQ_OK;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (q + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int64_t lo = t * chunk, hi = std::min<int64_t>(q, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back(work, lo, hi);
    }
    for (auto &th : pool) th.join();
    return RMQ_OK;
}

RMQ_API int rmq_gen_queries(int32_t *d_Q, int64_t q, int32_t n, int32_t lr, uint64_t seed, hipStream_t s) {
    QueryGen g;
    if (q < 0 || (q > 0 && !d_Q) || !make_query_gen(n, lr, seed, &g)) return RMQ_ERR_INVALID_ARG;
    if (q == 0) return RMQ_OK;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((q + 255) / 256, (int64_t)device_cus() * 32));
    k_gen_queries<<<grid, 256, 0, s>>>(d_Q, q, g);
    return hip_rc(hipGetLastError());
}

// ===========================================================================
// autotuner
// ===========================================================================
namespace {

// Cache-key version: bump whenever the tuner's search or timing changes, so
// picks made by another tuner never match.
constexpr const char *kTuneVersion = "t3";
constexpr int kSigWindow = 64;     // elements per sampled window of the array
constexpr int kSigWindows = 4096;  // windows sampled, evenly spaced

// Value signature of an array: over kSigWindows evenly spaced windows of
// kSigWindow elements, the mean position of the leftmost minimum within its
// window (0 ascending, 0.5 i.i.d., 1 descending; in quarters) and
// floor(log2(1 + mean count of further elements equal to the window
// minimum)) (0 for distinct values, 2 for 16 distinct values).  Both move
// the superset test's success rate, so arrays that differ in them are tuned
// apart.
int value_signature(const float *d_A, int64_t n, hipStream_t s, char *out, size_t out_len) {
    const int w = (int)std::min<int64_t>(n, kSigWindow);
    const int64_t m = std::min<int64_t>(kSigWindows, n / w);
    const int64_t step = (n - w) / std::max<int64_t>(1, m - 1);  // window k starts at k * step, ends by n
    int2 *d_sig = nullptr;
    RMQ_HIP(hipMalloc(&d_sig, sizeof(int2) * (size_t)m));
    std::vector<int2> sig((size_t)m);
    k_window_sig<<<(unsigned)((m + kWavesPerWG - 1) / kWavesPerWG), kWG, 0, s>>>(d_A, w, step, m, d_sig);
    int rc = hip_rc(hipGetLastError());
    if (rc == RMQ_OK) rc = hip_rc(hipMemcpyAsync(sig.data(), d_sig, sizeof(int2) * (size_t)m, hipMemcpyDeviceToHost, s));
    if (rc == RMQ_OK) rc = hip_rc(hipStreamSynchronize(s));
    (void)hipFree(d_sig);
    if (rc != RMQ_OK) return rc;
    double pos = 0.0, extra = 0.0;
    for (const int2 &g : sig) {
        pos += w > 1 ? (double)g.x / (w - 1) : 0.5;
        extra += g.y - 1;
    }
    pos /= (double)m;
    extra /= (double)m;
    snprintf(out, out_len, "pos%.2f/tie%d", std::round(4.0 * pos) / 4.0, (int)std::floor(std::log2(1.0 + extra)));
    return RMQ_OK;
}

// Length-distribution signature of a query batch: log2 n and the 10th, 50th
// and 90th percentiles of log2(range length), in half steps, over up to 2^16
// evenly spaced queries.
int tune_signature(const int32_t *d_Q, int64_t q, int64_t n, hipStream_t s, char *key, size_t key_len) {
    const int64_t m = std::min<int64_t>(q, 1 << 16);
    const int64_t step = q / m;
    std::vector<int32_t> sample(2 * (size_t)m);
    int2 *d_s = nullptr;
    RMQ_HIP(hipMalloc(&d_s, sizeof(int2) * (size_t)m));
    k_gather_stride<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(reinterpret_cast<const int2 *>(d_Q), step, m, d_s);
    int rc = hip_rc(hipGetLastError());
    if (rc == RMQ_OK) rc = hip_rc(hipMemcpyAsync(sample.data(), d_s, sizeof(int2) * (size_t)m, hipMemcpyDeviceToHost, s));
    if (rc == RMQ_OK) rc = hip_rc(hipStreamSynchronize(s));
    (void)hipFree(d_s);
    if (rc != RMQ_OK) return rc;
    std::vector<double> lg((size_t)m);
    for (int64_t k = 0; k < m; ++k) lg[k] = std::log2(std::max(1.0, (double)sample[2 * k + 1] - sample[2 * k] + 1));
    std::sort(lg.begin(), lg.end());
    auto pct = [&](double p) { return std::round(2.0 * lg[(size_t)std::min<double>(m - 1, p * m)]) / 2.0; };
    snprintf(key, key_len, "n2^%.1f/p10=%.1f/p50=%.1f/p90=%.1f", std::round(2.0 * std::log2((double)n)) / 2.0,
             pct(0.10), pct(0.50), pct(0.90));
    return RMQ_OK;
}

bool cache_lookup(const char *path, const char *key, rmq_tune_result_t *out) {
    FILE *f = path ? fopen(path, "r") : nullptr;
    if (!f) return false;
    char line[512], k[256];
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        rmq_tune_result_t r{};
        unsigned flags = 0;
        const int got = sscanf(line, "%255s %d %d %d %d %f %u", k, &r.bs, &r.opts.short_max, &r.opts.group_lanes,
                               &r.opts.long_lanes, &r.best_ms, &flags);
        if (got >= 6 && strcmp(k, key) == 0) {
            r.build_flags = got == 7 ? flags : 0u;
            snprintf(r.key, sizeof(r.key), "%s", key);
            *out = r;  // last entry for a key wins
            found = true;
        }
    }
    fclose(f);
    return found;
}

void cache_store(const char *path, const char *key, const rmq_tune_result_t &r) {
    FILE *f = path ? fopen(path, "a") : nullptr;
    if (!f) return;
    fprintf(f, "%s %d %d %d %d %.6f %u\n", key, r.bs, r.opts.short_max, r.opts.group_lanes, r.opts.long_lanes,
            r.best_ms, r.build_flags);
    fclose(f);
}

// A device buffer larger than the last-level cache, written before each
// timed pass so that the pass starts from caches holding other data, as a
// batch does when other work ran before it.  Timed back to back on warm
// caches, the lognormal 2^6 batch's pick ran 1.82 ms in the tuner and 1.90 ms
// in a mixed stream; flushed, 1.88 against 1.91.
struct CacheFlush {
    void *buf = nullptr;
    size_t bytes = 0;
    explicit CacheFlush(size_t want) {
        if (hipMalloc(&buf, want) == hipSuccess) bytes = want;
        else (void)hipGetLastError();  // no room: time on warm caches
    }
    ~CacheFlush() {
        if (buf) (void)hipFree(buf);
    }
    int run(hipStream_t s) const { return buf ? hip_rc(hipMemsetAsync(buf, 0, bytes, s)) : RMQ_OK; }
};
constexpr size_t kFlushBytes = size_t(512) << 20;  // twice the 256 MB last-level cache

// Best of `reps` timed passes (after one warm-up pass), milliseconds; each
// timed pass follows a cache flush (outside the timed events).
int time_pass(const rmq_handle *h, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx,
              const rmq_query_opts_t &o, hipStream_t s, int reps, float *best, const CacheFlush *flush = nullptr) {
    rmq_query_times_t t;
    int rc = query_impl(h, d_Q, q, d_val, d_idx, &o, s, &t);
    *best = 1e30f;
    for (int k = 0; rc == RMQ_OK && k < reps; ++k) {
        if (flush) rc = flush->run(s);
        if (rc == RMQ_OK) rc = query_impl(h, d_Q, q, d_val, d_idx, &o, s, &t);
        if (rc == RMQ_OK) *best = std::min(*best, t.total_ms);
    }
    return rc;
}

}  // namespace

RMQ_API int rmq_tune_ex(const float *d_A, int64_t n, const int32_t *d_Q, int64_t q, const int32_t *bs_list,
                        int32_t n_bs, uint32_t flags_allowed, float *d_val, int32_t *d_idx, const char *cache_path,
                        hipStream_t s, rmq_tune_result_t *out) {
    if (!out || !d_A || n <= 0 || n > INT32_MAX || q <= 0 || !d_Q || (!d_val && !d_idx) || n_bs < 0 ||
        (n_bs > 0 && !bs_list) || !build_flags_ok(flags_allowed))
        return RMQ_ERR_INVALID_ARG;
    *out = rmq_tune_result_t{};
    // cache key: tuner version / n and length percentiles / value signature /
    // outputs timed (value+index, value or index)
    char lens[96], vals[32], key[160];
    int rc = tune_signature(d_Q, q, n, s, lens, sizeof(lens));
    if (rc == RMQ_OK) rc = value_signature(d_A, n, s, vals, sizeof(vals));
    if (rc != RMQ_OK) return rc;
    // a caller's candidate list is part of the key (its pick answers a
    // different question): the list itself when short, else its FNV-1a hash
    char cands_tag[40] = "";
    if (n_bs > 0) {
        std::string list;
        for (int32_t i = 0; i < n_bs; ++i) list += (i ? "," : "") + std::to_string(bs_list[i]);
        if (list.size() <= 24) {
            snprintf(cands_tag, sizeof(cands_tag), "/bs=%s", list.c_str());
        } else {
            uint32_t hsh = 2166136261u;
            for (char ch : list) hsh = (hsh ^ (uint8_t)ch) * 16777619u;
            snprintf(cands_tag, sizeof(cands_tag), "/bs#%08x", hsh);
        }
    }
    // so are the build flags it may choose from, when some are excluded
    char flags_tag[16] = "";
    if (flags_allowed != (RMQ_BUILD_EDGES | RMQ_BUILD_MID)) snprintf(flags_tag, sizeof(flags_tag), "/f%u", flags_allowed);
    snprintf(key, sizeof(key), "%s/%s/%s/out=%s%s%s", kTuneVersion, lens, vals,
             d_val && d_idx ? "vi" : (d_val ? "v" : "i"), cands_tag, flags_tag);
    if (strlen(key) >= sizeof(out->key)) return RMQ_ERR_INVALID_ARG;
    if (cache_lookup(cache_path, key, out)) {
        out->from_cache = 1;
        out->n_tried = 0;
        return RMQ_OK;
    }
    const CacheFlush flush(kFlushBytes);
    std::vector<int32_t> cands;
    if (n_bs > 0) {
        cands.assign(bs_list, bs_list + n_bs);
    } else {
        for (int b = 3; b <= 18; ++b) cands.push_back(1 << b);
    }
    // an array smaller than every candidate is one block
    if (std::none_of(cands.begin(), cands.end(), [&](int32_t b) { return b > 0 && b <= n; }))
        cands.push_back((int32_t)n);
    // stage 1's lanes and wide short-range threshold, fixed apart from the
    // query defaults so that one tuner version makes one set of picks
    constexpr int32_t kBaseShortMax = 1024;
    auto base_opts = [](int32_t) {
        rmq_query_opts_t o{kBaseShortMax, 16, 0, 4};
        return o;
    };
    constexpr int kReps = 5;     // timed passes per configuration (best kept)
    constexpr int kFinalists = 4;  // block-size candidates that get the threshold search
    constexpr int kRunoff = 3;     // fastest options per finalist timed again, interleaved
    constexpr int kRounds = 3;     // rounds of that runoff
    int tried = 0;
    // stage 1: block size, with and without edge tables, default lanes; each
    // at a small and at the default short-range threshold (with the default
    // alone, short ranges would all be scanned and hide the block size), and
    // with edge tables also at 1 (only ranges inside one sub-block scanned)
    struct Cand {
        float ms;
        int32_t bs;
        uint32_t flags;
    };
    std::vector<Cand> stage1;
    for (int32_t bs : cands) {
        if (bs <= 0 || bs > n) continue;
        // coarse rows only where the query pass reads the block table from
        // global memory (a table that fits in LDS is read from there whole)
        const HandleSizes zb = handle_sizes(n, bs, 0);
        const bool mid_useful =
            !lds_table_fits((size_t)zb.levels * (size_t)zb.nb * sizeof(Pair) +
                            (sizeof(TileStage) + sizeof(TileResults)) * kLdsWaves);
        for (uint32_t flags : {0u, (uint32_t)RMQ_BUILD_EDGES, (uint32_t)(RMQ_BUILD_EDGES | RMQ_BUILD_MID)}) {
            if ((flags & ~flags_allowed) || ((flags & RMQ_BUILD_MID) && !mid_useful)) continue;
            rmq_handle_t *h = nullptr;
            rc = rmq_build(d_A, n, bs, flags, s, &h);
            if (rc == RMQ_ERR_OOM && flags) continue;  // no room for edge tables: plain only
            if (rc != RMQ_OK) return rc;
            float ms = 1e30f;
            for (int32_t sm : {1, 32, kBaseShortMax}) {
                if (sm == 1 && !flags) continue;
                rmq_query_opts_t o = base_opts(bs);
                o.short_max = sm;
                float t = 0.f;
                if ((rc = time_pass(h, d_Q, q, d_val, d_idx, o, s, kReps, &t, &flush)) != RMQ_OK) break;
                ++tried;
                ms = std::min(ms, t);
            }
            rmq_free(h);
            if (rc != RMQ_OK) return rc;
            stage1.push_back(Cand{ms, bs, flags});
        }
    }
    if (stage1.empty()) return RMQ_ERR_INVALID_ARG;
    std::sort(stage1.begin(), stage1.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
    if (stage1.size() > (size_t)kFinalists) stage1.resize(kFinalists);
    // stages 2 and 3 for each finalist (defaults can favour a block size that
    // is not the best once the thresholds are tuned): short-range threshold
    // and lanes per short query, then lanes per long query
    float best = 1e30f;
    int32_t best_bs = 0;
    uint32_t best_flags = 0;
    rmq_query_opts_t best_o{};
    struct Timed {
        float ms;
        rmq_query_opts_t o;
    };
    auto same_opts = [](const rmq_query_opts_t &x, const rmq_query_opts_t &y) {
        return x.short_max == y.short_max && x.group_lanes == y.group_lanes && x.long_lanes == y.long_lanes;
    };
    for (const Cand &c : stage1) {
        rmq_handle_t *h = nullptr;
        if ((rc = rmq_build(d_A, n, c.bs, c.flags, s, &h)) != RMQ_OK) return rc;
        std::vector<Timed> seen;  // every option set timed for this finalist
        rmq_query_opts_t cand_o = base_opts(c.bs);
        float cand_ms = 1e30f;
        if ((rc = time_pass(h, d_Q, q, d_val, d_idx, cand_o, s, kReps, &cand_ms, &flush)) == RMQ_OK) {
            ++tried;
            seen.push_back(Timed{cand_ms, cand_o});
        }
        for (int32_t sm : {1, 8, 32, 128, 512, 1024, 4096}) {
            if (sm < 32 && !c.flags) continue;
            for (int32_t g : {4, 8, 16}) {
                if (rc != RMQ_OK) break;
                rmq_query_opts_t o = cand_o;
                o.short_max = sm;
                o.group_lanes = g;
                float ms = 0.f;
                if ((rc = time_pass(h, d_Q, q, d_val, d_idx, o, s, kReps, &ms, &flush)) != RMQ_OK) break;
                ++tried;
                seen.push_back(Timed{ms, o});
                if (ms < cand_ms) {
                    cand_ms = ms;
                    cand_o = o;
                }
            }
        }
        for (int32_t g : {4, 8, 16, 64}) {
            if (rc != RMQ_OK) break;
            rmq_query_opts_t o = cand_o;
            o.long_lanes = g;
            float ms = 0.f;
            if ((rc = time_pass(h, d_Q, q, d_val, d_idx, o, s, kReps, &ms, &flush)) != RMQ_OK) break;
            ++tried;
            seen.push_back(Timed{ms, o});
            if (ms < cand_ms) {
                cand_ms = ms;
                cand_o = o;
            }
        }
        // Runoff: near-equal option sets swap places between single timings,
        // so the fastest few are timed again, interleaved, and the lowest
        // total wins.
        if (rc == RMQ_OK) {
            std::sort(seen.begin(), seen.end(), [](const Timed &x, const Timed &y) { return x.ms < y.ms; });
            std::vector<Timed> top;
            for (const Timed &t : seen) {
                if (std::none_of(top.begin(), top.end(), [&](const Timed &u) { return same_opts(u.o, t.o); }))
                    top.push_back(Timed{0.f, t.o});
                if ((int)top.size() == kRunoff) break;
            }
            for (int round = 0; rc == RMQ_OK && round < kRounds; ++round) {
                for (Timed &t : top) {
                    float ms = 0.f;
                    if ((rc = time_pass(h, d_Q, q, d_val, d_idx, t.o, s, kReps, &ms, &flush)) != RMQ_OK) break;
                    ++tried;
                    t.ms += ms;
                }
            }
            if (rc == RMQ_OK && !top.empty()) {
                const Timed &w =
                    *std::min_element(top.begin(), top.end(), [](const Timed &x, const Timed &y) { return x.ms < y.ms; });
                cand_o = w.o;
                cand_ms = w.ms / kRounds;
            }
        }
        rmq_free(h);
        if (rc != RMQ_OK) return rc;
        if (cand_ms < best) {
            best = cand_ms;
            best_bs = c.bs;
            best_flags = c.flags;
            best_o = cand_o;
        }
    }
    out->bs = best_bs;
    out->build_flags = best_flags;
    out->opts = best_o;
    out->best_ms = best;
    out->n_tried = tried;
    out->from_cache = 0;
    snprintf(out->key, sizeof(out->key), "%s", key);
    cache_store(cache_path, key, *out);
    return RMQ_OK;
}

RMQ_API int rmq_tune(const float *d_A, int64_t n, const int32_t *d_Q, int64_t q, const int32_t *bs_list,
                     int32_t n_bs, float *d_val, int32_t *d_idx, const char *cache_path, hipStream_t s,
                     rmq_tune_result_t *out) {
    return rmq_tune_ex(d_A, n, d_Q, q, bs_list, n_bs, RMQ_BUILD_EDGES | RMQ_BUILD_MID, d_val, d_idx, cache_path, s,
                       out);
}

// ===========================================================================
// the reference's own query stream (random_queries_par_cpu, src/rand.cuh:116-201)
// ===========================================================================
// Reproduces the reference's generator call for call, for reruns of its
// sweeps on the same queries: per "thread" tid in [0, nt) an mt19937 seeded
// with seed*tid fills the chunk [tid*ceil(q/nt), ...) (src/rand.cuh:119-123);
// lr > 0 fixed lengths (:116-133), -1 uniform lengths in [1, n] (:136-154),
// -2..-5 lognormal lengths with sigma 0.3, redrawn until in [1, n], at the
// scales n^0.6, n^0.3, n/2^8, n/2^15 (:156-177, :188-199).  The distributions
// are the standard library's (mt19937, uniform_int_distribution,
// lognormal_distribution), so the stream is the reference's wherever it is
// built against the same C++ standard library.  The stream depends on nt,
// as in the reference.
RMQ_API int rmq_gen_queries_ref(int32_t *Q, int64_t q, int32_t n, int32_t lr, int32_t seed, int32_t nt) {
    if (!Q || q < 0 || q > INT32_MAX || n <= 0 || lr == 0 || lr > n || lr < -5 || nt <= 0) return RMQ_ERR_INVALID_ARG;
    const int qi = (int)q;
    const int chunk = (qi + nt - 1) / nt;
    auto fill = [&](int tid) {
        std::mt19937 gen((uint32_t)seed * (uint32_t)tid);  // the reference's int product, wrapped
        const int begin = chunk * tid;
        const int end = begin + chunk;
        if (lr > 0) {
            for (int i = begin; i < qi && i < end; ++i) {
                std::uniform_int_distribution<int> lrand(0, n - 1 - (lr - 1));
                const int l = lrand(gen);
                Q[2 * (int64_t)i] = l;
                Q[2 * (int64_t)i + 1] = l + (lr - 1);
            }
        } else if (lr == -1) {
            std::uniform_int_distribution<int> dist(1, n);
            for (int i = begin; i < qi && i < end; ++i) {
                const int len = dist(gen);
                std::uniform_int_distribution<int> lrand(0, n - 1 - (len - 1));
                const int l = lrand(gen);
                Q[2 * (int64_t)i] = l;
                Q[2 * (int64_t)i + 1] = l + (len - 1);
            }
        } else {
            int scale = 1;
            if (lr == -2) scale = (int)std::pow((double)n, 0.6);
            if (lr == -3) scale = (int)std::pow((double)n, 0.3);
            if (lr == -4) scale = std::max(1, n / (1 << 8));
            if (lr == -5) scale = std::max(1, n / (1 << 15));
            std::lognormal_distribution<double> dist(std::log((double)scale), 0.3);
            for (int i = begin; i < qi && i < end; ++i) {
                int len;
                do {
                    const double x = dist(gen);
                    len = (x >= 2147483648.0 || x != x) ? INT_MIN : (int)x;  // x86 conversion of out-of-range values
                } while (len <= 0 || len > n);
                std::uniform_int_distribution<int> lrand(0, n - 1 - (len - 1));
                const int l = lrand(gen);
                Q[2 * (int64_t)i] = l;
                Q[2 * (int64_t)i + 1] = l + (len - 1);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(fill, t);
    fill(0);
    for (auto &th : pool) th.join();
    return RMQ_OK;
}

RMQ_API int rmq_probe_gather(const void *d_t0, const void *d_t1, int64_t entries, int64_t count, void *d_out,
                             int32_t reps, hipStream_t s, float *ms_per_launch) {
    if (!d_t0 || !d_t1 || !d_out || !ms_per_launch || entries < 2 || (entries & (entries - 1)) || count <= 0 ||
        reps <= 0)
        return RMQ_ERR_INVALID_ARG;
    const int grid = device_cus() * 16;
    auto launch = [&]() {
        k_probe_gather<<<grid, kWG, 0, s>>>(static_cast<const uint2 *>(d_t0), static_cast<const uint2 *>(d_t1),
                                            (uint64_t)entries - 1, count, static_cast<uint2 *>(d_out));
        return hip_rc(hipGetLastError());
    };
    int rc = launch();  // warm-up
    hipEvent_t a = nullptr, b = nullptr;
    if (rc == RMQ_OK) rc = hip_rc(hipEventCreate(&a));
    if (rc == RMQ_OK) rc = hip_rc(hipEventCreate(&b));
    if (rc == RMQ_OK) rc = hip_rc(hipEventRecord(a, s));
    for (int r = 0; r < reps && rc == RMQ_OK; ++r) rc = launch();
    if (rc == RMQ_OK) rc = hip_rc(hipEventRecord(b, s));
    if (rc == RMQ_OK) rc = hip_rc(hipEventSynchronize(b));
    float ms = 0.f;
    if (rc == RMQ_OK) rc = hip_rc(hipEventElapsedTime(&ms, a, b));
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (rc == RMQ_OK) *ms_per_launch = ms / (float)reps;
    return rc;
}

RMQ_API const char *rmq_strerror(int code) {
    switch (code) {
    case RMQ_OK: return "ok";
    case RMQ_ERR_INVALID_ARG: return "invalid argument";
    case RMQ_ERR_OOM: return "out of device memory";
    case RMQ_ERR_HIP: return "HIP runtime error";
    case RMQ_ERR_DEVICE: return "handle used on a device other than the one it was built on";
    case RMQ_ERR_BAD_QUERY: return "query with l < 0, l > r or r >= n";
    case RMQ_ERR_LIMIT:
        return "more arrays or work items than the static scratch of rmq_plan_launch_static holds (8192 arrays, "
               "65536 work items)";
    default: return "unknown error";
    }
}

RMQ_API const char *rmq_version(void) { return "rtxrmq-amd 0.1 (gfx950)"; }

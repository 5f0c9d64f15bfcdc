// rmq_common.h -- device helpers and host plumbing shared by the engine's
// HIP sources (rmq.hip: single-array build and query; batch.hip: work-item
// plans, the persistent multi-array kernel and the serving loop).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>

#include "rmq.h"

#define RMQ_API extern "C" __attribute__((visibility("default")))

namespace rmqx {

constexpr int kWave = 64;
constexpr int kWG = 256;                 // threads per workgroup (4 waves)
constexpr int kWavesPerWG = kWG / kWave;
constexpr int kLdsWaves = 16;            // waves per workgroup of the LDS-table query kernel
constexpr int64_t kTwoStageTableBytes = int64_t(64) << 20;  // block tables beyond this take answer_two_stage
// SGPR budget for query kernels that would otherwise allocate more than 80
// SGPRs: on gfx950 an allocation of 96 or more cuts resident waves by 40-50%
// (measured with SQ_WAVE_CYCLES / SQ_BUSY_CYCLES) although the compiler still
// reports 8 waves per SIMD, while the few SGPRs spilled to VGPR lanes cost
// little in these memory-bound kernels.
#define RMQ_SGPR_BUDGET __attribute__((amdgpu_num_sgpr(72)))
constexpr int kLongUnroll = 4;           // float4 loads in flight per lane in the wave scan
// Default short-range threshold and lanes per short range: ranges up to 128
// elements are scanned by 8 lanes, longer ones inside a block read their whole
// sub-blocks as minima (plain handle at bs=2^15, n=1e8, 2^26 queries per batch,
// uniform / lognormal 2^14 / lognormal 2^6 lengths: 1.02 + 5.05 + 6.34 ms at
// 1024 and 16 lanes, 1.00 + 4.92 + 4.45 ms at 128 and 8).
constexpr int kDefaultShortMax = 128;
constexpr int kDefaultGroupLanes = 8;
// Lanes per long query by default: with the sparse table over sub-block
// minima a long in-block scan is two partial sub-blocks and two entries, so
// more queries in flight beat more lanes per query (literal bs=2^15,
// lognormal 2^14 lengths: 64 lanes 29.4 ms, 16 lanes 13.4, 4 lanes 6.5;
// uniform lengths 1.01 -> 1.07 ms).
constexpr int kDefaultLongLanes = 4;
constexpr int kSub = 32;  // sub-block size (elements); sub-blocks restart at every block start
constexpr int kSubShift = 5;
constexpr int kHierMin = 4;  // whole sub-blocks between the ends before a scan reads their minima instead
constexpr int kHierTable = 16;  // more whole sub-blocks than this: two sparse-table entries instead of the minima

struct __align__(8) Pair {
    float v;
    int32_t i;
};

// ---------------------------------------------------------------------------
// per-lane accumulation: lanes visit their elements in increasing index order,
// so strict '<' keeps the leftmost minimum within a lane.  The sentinel is
// (+inf, INT_MAX); an element equal to +inf never replaces it, which the
// final fix-up (value +inf => index l) corrects: with no NaN, a +inf minimum
// means every element of the range is +inf and the leftmost is l.
// ---------------------------------------------------------------------------
// A (min, argmin) pair loaded non-temporally.
__device__ __forceinline__ Pair nt_load_pair(const Pair *p) {
    const long long raw = __builtin_nontemporal_load(reinterpret_cast<const long long *>(p));
    return Pair{__int_as_float((int)(uint32_t)raw), (int)(uint32_t)((unsigned long long)raw >> 32)};
}

__device__ __forceinline__ void upd(float x, int e, float &bv, int &bi) {
    const bool t = x < bv;
    bv = t ? x : bv;
    bi = t ? e : bi;
}

__device__ __forceinline__ void upd4(const float4 &x, int e, float &bv, int &bi) {
    upd(x.x, e, bv, bi);
    upd(x.y, e + 1, bv, bi);
    upd(x.z, e + 2, bv, bi);
    upd(x.w, e + 3, bv, bi);
}

// (ov, oi) replaces (v, i) when smaller, or equal with a smaller index.
__device__ __forceinline__ void take_if_better(float ov, int oi, float &v, int &i) {
    const bool t = (ov < v) || (ov == v && oi < i);
    v = t ? ov : v;
    i = t ? oi : i;
}

// One DPP exchange step.  'old' is the lane's own value, so a lane whose
// source is out of the row keeps its own pair (idempotent for a minimum).
template <int CTRL>
__device__ __forceinline__ void dpp_step(float &v, int &i) {
    const int ov = __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false);
    const int oi = __builtin_amdgcn_update_dpp(i, i, CTRL, 0xF, 0xF, false);
    take_if_better(__int_as_float(ov), oi, v, i);
}

// DPP controls: quad_perm [1,0,3,2] (xor 1), quad_perm [2,3,0,1] (xor 2),
// row_half_mirror (pairs the two quads of each 8), row_mirror (pairs the two
// halves of each row of 16).  After G steps every lane of an aligned group of
// G lanes holds the group's minimum.
template <int G>
__device__ __forceinline__ void group_reduce(float &v, int &i) {
    static_assert(G == 4 || G == 8 || G == 16, "DPP group reduction works within a row of 16");
    dpp_step<0xB1>(v, i);
    dpp_step<0x4E>(v, i);
    if constexpr (G >= 8) dpp_step<0x141>(v, i);
    if constexpr (G >= 16) dpp_step<0x140>(v, i);
}

// Whole-wave reduction: rows of 16 by DPP, then the four row results are read
// into scalar registers.  Every lane returns the same pair.
__device__ __forceinline__ void wave_reduce(float &v, int &i) {
    group_reduce<16>(v, i);
    float rv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    int ri = __builtin_amdgcn_readlane(i, 0);
#pragma unroll
    for (int row = 1; row < 4; ++row) {
        const float ov = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), row * 16));
        const int oi = __builtin_amdgcn_readlane(i, row * 16);
        take_if_better(ov, oi, rv, ri);
    }
    v = rv;
    i = ri;
}

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint32_t mask_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Scans A[a..b] (a <= b) with STRIDE cooperating lanes, lane in [0, STRIDE).
// Head (up to 3 elements before the first 16-byte boundary) and tail (up to
// 3 after the last) are scalar loads; the aligned middle is float4 loads, U
// per lane in flight.  Each lane visits head, middle, tail in index order.
// Every load lies inside [a, b], so nothing past the array end is touched.
// PRED: the middle's loads go out U at a time, each under its own bounds
// test, so that a lane with between 1 and U float4s has them all in flight
// at once (without it, U loads go out only while U remain).
template <int STRIDE, int U, bool PRED = false>
__device__ __forceinline__ void scan_range(const float *__restrict__ A, int a, int b, int lane, float &bv, int &bi) {
    static_assert(STRIDE >= 4, "the scalar head and tail (up to 3 elements each) take one lane per element");
    const int ah = (int)min(((int64_t)a + 3) & ~(int64_t)3, (int64_t)INT_MAX);  // a + 3 can pass INT_MAX
    const int bt = (b + 1) & ~3;  // b + 1 <= INT_MAX: b < n <= INT_MAX
    const int h_end = min(ah, b + 1);
    const int t_beg = max(bt, h_end);
    if (lane < h_end - a) upd(A[a + lane], a + lane, bv, bi);
    const float4 *__restrict__ A4 = reinterpret_cast<const float4 *>(A);
    int v = (h_end >> 2) + lane;
    const int v_end = t_beg >> 2;  // exclusive; empty when h_end == t_beg
    if (h_end < t_beg) {
        if constexpr (PRED) {
            for (; v < v_end; v += U * STRIDE) {
                float4 x[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (v + u * STRIDE < v_end) x[u] = A4[v + u * STRIDE];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (v + u * STRIDE < v_end) upd4(x[u], (v + u * STRIDE) << 2, bv, bi);
            }
        } else {
            for (; v + (U - 1) * STRIDE < v_end; v += U * STRIDE) {
                float4 x[U];
#pragma unroll
                for (int u = 0; u < U; ++u) x[u] = A4[v + u * STRIDE];
#pragma unroll
                for (int u = 0; u < U; ++u) upd4(x[u], (v + u * STRIDE) << 2, bv, bi);
            }
            for (; v < v_end; v += STRIDE) upd4(A4[v], v << 2, bv, bi);
        }
    }
    if (lane < b + 1 - t_beg) upd(A[t_beg + lane], t_beg + lane, bv, bi);
}

// ---------------------------------------------------------------------------
// query kernels
//
// Binning is done per tile of 64 consecutive queries, in registers: lane j of
// a wave loads query t*64+j, and two ballots split the tile into short ranges
// (scanned by lane groups) and long ones.  Without edge tables two kernels
// run: the short one answers the short ranges of every tile, several per
// wave; the long one the long ranges, a lane group each.  With edge tables
// one kernel answers both classes of a tile (long ones one lane per query).
// Each kernel reads the query array once, so binning costs 8 bytes per query
// per kernel and no atomics, no queue and no workspace: each query writes
// only its own output slot.
// ---------------------------------------------------------------------------

// Results are written once and not read again by the kernels, so they are
// stored non-temporally (NT): they do not displace table lines from the
// caches (+2% on the three n=1e8 batches, most on the L2-resident lognormal
// 2^14 one).  The plan kernel uses ordinary stores (NT = false): a server
// hands it pinned host buffers, which it reads and writes over PCIe, where
// the hint has nothing to gain (its batched rate is the same either way).
template <bool NT = true>
__device__ __forceinline__ void write_result(float *out_v, int32_t *out_i, int64_t qi, float v, int i) {
    if constexpr (NT) {
        if (out_v) __builtin_nontemporal_store(v, out_v + qi);
        if (out_i) __builtin_nontemporal_store(i, out_i + qi);
    } else {
        if (out_v) out_v[qi] = v;
        if (out_i) out_i[qi] = i;
    }
}

// The query stream is read once per pass: a non-temporal load keeps its lines
// from displacing table lines in the caches (+2-3% on the n=1e8 batches);
// NT = false as for write_result.
template <bool NT = true>
__device__ __forceinline__ int2 load_query(const int2 *__restrict__ Q, int64_t qi) {
    if constexpr (NT) {
        const long long raw = __builtin_nontemporal_load(reinterpret_cast<const long long *>(Q) + qi);
        return make_int2((int)(uint32_t)raw, (int)(uint32_t)((unsigned long long)raw >> 32));
    } else {
        return Q[qi];
    }
}

struct TileMasks {
    int l, r;          // this lane's query
    uint64_t shorts;   // lanes holding a valid range of at most short_max elements
    uint64_t longs;    // lanes holding a valid longer range
    bool own_short;    // this lane's bit of shorts, without a 64-bit shift and test
    bool own_long;     // this lane's bit of longs
};

template <bool NT = true>
__device__ __forceinline__ TileMasks load_tile(const int2 *__restrict__ Q, int64_t q, int n, int short_max, int64_t t,
                                               int lane, bool flag_invalid, float *out_v, int32_t *out_i) {
    const int64_t qi = t * kWave + lane;
    const bool in = qi < q;
    int l = 0, r = -1;
    if (in) {
        const int2 p = load_query<NT>(Q, qi);
        l = p.x;
        r = p.y;
    }
    const bool ok = in && l >= 0 && l <= r && r < n;
    if (flag_invalid && in && !ok) write_result<NT>(out_v, out_i, qi, __int_as_float(0x7FC00000), -1);
    const bool is_short = ok && (r - l) < short_max;
    return TileMasks{l, r, __ballot(is_short), __ballot(ok && !is_short), is_short, ok && !is_short};
}

// Block index of element e (0 <= e < 2^31): e / bs as mul_hi(2e, magic) >>
// shift, one 32-bit multiply, with shift = ceil(log2 bs) and magic =
// ceil(2^(31+shift) / bs), which is below 2^32 for every bs < 2^31.  Exact:
// magic = 2^(31+shift)/bs + d with 0 <= d < 1, so e*magic / 2^(31+shift) =
// e/bs + e*d / 2^(31+shift), and the error e*d / 2^(31+shift) < 2^-shift <=
// 1/bs never reaches the next multiple of 1/bs above e/bs.  bs = 1 gives
// magic = 2^31, shift 0: the identity.
struct BlockDiv {
    uint32_t magic;
    int32_t shift;
};

__host__ __device__ __forceinline__ BlockDiv block_div(int bs) {
    int sh = 0;
    while ((int64_t)1 << sh < (int64_t)bs) ++sh;
    return BlockDiv{(uint32_t)((((uint64_t)1 << (31 + sh)) + (uint64_t)bs - 1) / (uint64_t)bs), sh};
}

__device__ __forceinline__ int block_of(int e, BlockDiv d) {
    return (int)(__umulhi((uint32_t)e << 1, d.magic) >> d.shift);
}

// Leftmost minimum of the c >= 1 entries [a, a+c) of a sparse table whose
// rows hold `stride` entries; the left window wins ties.  Row offsets are
// zero-extended 32 x 32-bit products (one multiply-add).
__device__ __forceinline__ Pair table_min(const Pair *__restrict__ st, uint32_t stride, int64_t a, int c) {
    const int lev = 31 - __clz(c);
    const Pair *row = st + (uint64_t)(uint32_t)lev * stride;
    const Pair x = row[a];
    const Pair y = row[a + c - (1 << lev)];
    return (y.v < x.v) ? y : x;
}

// Everything the query kernels read of one built array.
struct Blocks {
    const float *A;
    const Pair *st;   // levels rows of nb pairs; row 0 = each block's (min, leftmost argmin)
    const Pair *sub;  // sparse table over the nb * nsub sub-block minima (row 0: each 32-element sub-block's (min,
                      // leftmost argmin)); null when bs <= 32
    BlockDiv div;     // block_div(bs)
    int n, bs, nb, nsub;
};

// With edge tables every range is answered in O(1) (answer_edges) except
// one inside a single 32-element sub-block (a single block when bs <= 32):
// those, and ranges of at most short_max elements, go to the group scan.
// Otherwise as load_tile, with invalid queries flagged.
// Query t*64+lane, or the empty range (0, -1) past the end.
template <bool NT = true>
__device__ __forceinline__ int2 fetch_query(const int2 *__restrict__ Q, int64_t q, int64_t t, int lane) {
    const int64_t qi = t * kWave + lane;
    return qi < q ? load_query<NT>(Q, qi) : make_int2(0, -1);
}

// load_tile_edges on a query already loaded (p = fetch_query(Q, q, t, lane)).
template <bool NT = true, bool WRITE_INVALID = true>
__device__ __forceinline__ TileMasks classify_edges(int2 p, int64_t q, const Blocks &k, int short_max, int64_t t,
                                                    int lane, float *out_v, int32_t *out_i) {
    const int64_t qi = t * kWave + lane;
    const bool in = qi < q;
    const int l = p.x, r = p.y;
    const bool ok = in && l >= 0 && l <= r && r < k.n;
    if (WRITE_INVALID && in && !ok) write_result<NT>(out_v, out_i, qi, __int_as_float(0x7FC00000), -1);
    bool is_short = ok && (r - l) < short_max;
    if (ok && !is_short && r - l < kSub) {
        const int bl = block_of(l, k.div);
        const int base = bl * k.bs;  // <= l
        is_short = block_of(r, k.div) == bl && ((l - base) >> kSubShift) == ((r - base) >> kSubShift);
    }
    return TileMasks{l, r, __ballot(is_short), __ballot(ok && !is_short), is_short, ok && !is_short};
}

template <bool NT = true>
__device__ __forceinline__ TileMasks load_tile_edges(const int2 *__restrict__ Q, int64_t q, const Blocks &k,
                                                     int short_max, int64_t t, int lane, float *out_v,
                                                     int32_t *out_i) {
    return classify_edges<NT>(fetch_query<NT>(Q, q, t, lane), q, k, short_max, t, lane, out_v, out_i);
}

// Scans pairs P[a..b] (each a (min, leftmost argmin) of consecutive elements,
// in index order) with STRIDE lanes, U loads per lane in flight.
template <int STRIDE, int U>
__device__ __forceinline__ void scan_pairs(const Pair *__restrict__ P, int64_t a, int64_t b, int lane, float &bv,
                                           int &bi) {
    int64_t j = a + lane;
    for (; j + (U - 1) * STRIDE <= b; j += U * STRIDE) {
        Pair x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = P[j + u * STRIDE];
#pragma unroll
        for (int u = 0; u < U; ++u) take_if_better(x[u].v, x[u].i, bv, bi);
    }
    for (; j <= b; j += STRIDE) take_if_better(P[j].v, P[j].i, bv, bi);
}

// Scans [lo, hi] inside block b.  With a sub-block level and at least
// kHierMin whole sub-blocks between the ends, the partial sub-blocks at both
// ends are scanned element by element and the whole ones between read as
// their (min, argmin) pairs, 8 bytes per 32 elements, or, beyond kHierTable
// of them, as two entries of the sparse table over sub-block minima (its
// rows follow row 0 at a stride of nb * nsub).  Otherwise one direct scan.
// The split scan_range makes of [a, b]: scalar head [a, h_end), float4s
// [h_end/4, t_beg/4), scalar tail [t_beg, b].
struct ScanSplit {
    int a, b, h_end, t_beg;
};

__device__ __forceinline__ ScanSplit split_scan(int a, int b) {
    const int ah = (int)min(((int64_t)a + 3) & ~(int64_t)3, (int64_t)INT_MAX);
    const int h_end = min(ah, b + 1);
    return ScanSplit{a, b, h_end, max((b + 1) & ~3, h_end)};
}

// A scan of at most 32 elements by G lanes with every load issued before the
// first use (at most 8 float4s: max(1, 8/G) per lane), then combined in index
// order, as scan_range visits them.
template <int G>
struct ShortScan {
    static_assert(G >= 4, "the scalar head and tail (up to 3 elements each) take one lane per element");
    static constexpr int UM = G >= 8 ? 1 : 8 / G;
    float h, t;
    float4 m[UM];
    __device__ __forceinline__ void issue(const float *__restrict__ A, const ScanSplit &p, int lane) {
        if (lane < p.h_end - p.a) h = A[p.a + lane];
        const float4 *__restrict__ A4 = reinterpret_cast<const float4 *>(A);
        const int v0 = (p.h_end >> 2) + lane, v_end = p.t_beg >> 2;
#pragma unroll
        for (int u = 0; u < UM; ++u)
            if (v0 + u * G < v_end) m[u] = A4[v0 + u * G];
        if (lane < p.b + 1 - p.t_beg) t = A[p.t_beg + lane];
    }
    __device__ __forceinline__ void combine(const ScanSplit &p, int lane, float &bv, int &bi) const {
        if (lane < p.h_end - p.a) upd(h, p.a + lane, bv, bi);
        const int v0 = (p.h_end >> 2) + lane, v_end = p.t_beg >> 2;
#pragma unroll
        for (int u = 0; u < UM; ++u)
            if (v0 + u * G < v_end) upd4(m[u], (v0 + u * G) << 2, bv, bi);
        if (lane < p.b + 1 - p.t_beg) upd(t, p.t_beg + lane, bv, bi);
    }
};

template <int G, int U>
__device__ __forceinline__ void scan_hier(const Blocks &k, int b, int lo, int hi, int gl, float &bv, int &bi) {
    if (k.sub) {
        const int base = b * k.bs;
        const int jl = (lo - base) >> kSubShift, jr = (hi - base) >> kSubShift;
        const int whole = jr - jl - 1;
        if (whole >= kHierMin) {
            // Every load goes out before the first combine: the whole
            // sub-blocks' two table entries, then both partial sub-blocks'
            // elements (ShortScan), so that the three parts share one round
            // trip.  The combines run in index order (left partial, middle,
            // right partial), as the strict '<' of upd needs.
            const int64_t row = (int64_t)b * k.nsub;
            const ScanSplit p1 = split_scan(lo, base + (jl << kSubShift) + kSub - 1);
            const ScanSplit p2 = split_scan(base + (jr << kSubShift), hi);
            ShortScan<G> s1, s2;
            if (whole > kHierTable) {
                const Pair m = table_min(k.sub, (uint32_t)(k.nb * k.nsub), row + jl + 1, whole);
                s1.issue(k.A, p1, gl);
                s2.issue(k.A, p2, gl);
                s1.combine(p1, gl, bv, bi);
                take_if_better(m.v, m.i, bv, bi);
                s2.combine(p2, gl, bv, bi);
            } else {
                s1.issue(k.A, p1, gl);
                s2.issue(k.A, p2, gl);
                s1.combine(p1, gl, bv, bi);
                scan_pairs<G, U>(k.sub, row + jl + 1, row + jr - 1, gl, bv, bi);
                s2.combine(p2, gl, bv, bi);
            }
            return;
        }
    }
    scan_range<G, U>(k.A, lo, hi, gl, bv, bi);
}

// One query [l, r] answered by the G lanes of a group (gl = lane in group).
// Short class: one direct scan.  Long class: lB == rB is one in-block scan
// (src/rtx_kernels.cu:116-135), through the sub-block minima when it is long;
// otherwise the blocks between give their leftmost minimum in O(1) from the
// sparse table over block minima (:138-149), and the edge blocks (:151-177)
// are taken from their minima or scanned only when they can still win.  The
// group's partial results are combined with (mv, mi) with the index as
// tie-break, so the leftmost minimum wins.
template <int G, int U, bool LONG, bool PRED = false>
__device__ __forceinline__ void answer_partial(const Blocks &k, int l, int r, int gl, float &bv, int &bi, float &mv,
                                               int &mi) {
    if constexpr (!LONG) {
        scan_range<G, U, PRED>(k.A, l, r, gl, bv, bi);
    } else {
        const int lB = block_of(l, k.div), rB = block_of(r, k.div);
        if (lB == rB) {
            scan_hier<G, U>(k, lB, l, r, gl, bv, bi);
        } else {
            // The superset test (minimum of blocks lB..rB) has already been
            // tried one lane per query by answer_class; what remains is exact
            // without it.  The middle's two table entries and both edge
            // blocks' row-0 entries depend only on l and r: all four reads are
            // issued together (with no whole block between, the middle reads
            // repeat row 0 at lB and are not used).
            const bool has_mid = rB - lB > 1;
            const int c = has_mid ? rB - lB - 1 : 1;
            const int lev = 31 - __clz(c);
            const Pair *row = k.st + (int64_t)lev * k.nb;
            const Pair x = row[has_mid ? lB + 1 : lB];
            const Pair y = row[has_mid ? rB - (1 << lev) : lB];
            const Pair eb = k.st[lB];
            const Pair er = k.st[rB];
            if (has_mid) {
                const Pair m = (y.v < x.v) ? y : x;
                mv = m.v;
                mi = m.i;
            }
            // Edge blocks, from their (min, leftmost argmin) in table row 0:
            // exact when that argmin lies inside the edge part; skipped when
            // the block minimum, a lower bound for the part, cannot win; only
            // otherwise scanned.  (mv, mi) is known to every lane of the group,
            // so the group takes the same branch.
            if (eb.i >= l) take_if_better(eb.v, eb.i, mv, mi);
            else if (eb.v <= mv) scan_hier<G, U>(k, lB, l, lB * k.bs + k.bs - 1, gl, bv, bi);  // earlier: wins ties
            if (er.i <= r) take_if_better(er.v, er.i, mv, mi);
            else if (er.v < mv) scan_hier<G, U>(k, rB, rB * k.bs, r, gl, bv, bi);  // later: must be smaller
        }
    }
}

// ---------------------------------------------------------------------------
// edge tables (RMQ_BUILD_EDGES): precomputed edge-block searches
// ---------------------------------------------------------------------------

struct EdgeTables {
    const Pair *pre_b;  // per element i: leftmost min of [start of i's block, i]
    const Pair *suf_b;  // per element i: leftmost min of [i, end of i's block]
    const Pair *pre_s;  // the same within i's 32-element sub-block (sub-blocks start
    const Pair *suf_s;  //   at the block start; equal to pre_b / suf_b when bs <= 32)
    const Pair *st_s;   // sparse table over sub-block minima: levels_s rows of subs_total
    int64_t subs_total; // nb * nsub
    int32_t nsub;       // sub-blocks per block, ceil(bs / 32)
    int32_t levels_s;
};


// The superset window's minimum for a range across blocks bl < br: blocks
// bl..br from two level-lev entries, or, with LINES, when those two entries
// lie in different 128-byte lines, the one level-(lev+1) entry at bl (blocks
// bl..bl+2^(lev+1)-1, up to twice as many): one line instead of two.
template <bool LINES>
__device__ __forceinline__ Pair superset_min(const Pair *__restrict__ st, int nb, int bl, int br) {
    const int c = br - bl + 1;
    const int lev = 31 - __clz(c);
    const Pair *row = st + (int64_t)lev * nb;
    const int a2 = br - (1 << lev) + 1;
    if (LINES && ((reinterpret_cast<uintptr_t>(row + bl) ^ reinterpret_cast<uintptr_t>(row + a2)) >> 7) != 0 &&
        (int64_t)bl + ((int64_t)2 << lev) <= nb) {
        const Pair w = row[nb + bl];  // level lev + 1 (it exists: 2^(lev+1) <= nb)
        return w;
    }
    const Pair x = row[bl];
    const Pair y = row[a2];
    return (y.v < x.v) ? y : x;
}

// A range across blocks bl < br whose superset window minimum has its argmin
// outside [l, r]: whole blocks between, then each edge from its block's
// (min, leftmost argmin) in table row 0: exact when the argmin lies in the
// edge part, skipped when the block minimum cannot win, and only otherwise
// read from the edge table.
template <bool NT_PAIRS>
__device__ __forceinline__ Pair answer_failed(const EdgeTables &e, const Blocks &k, int l, int r, int bl, int br) {
    const Pair *__restrict__ st = k.st;
    const int nb = k.nb;
    Pair best = br - bl > 1 ? table_min(st, nb, bl + 1, br - bl - 1) : Pair{INFINITY, INT_MAX};
    const Pair el = st[bl];
    if (el.i >= l) {
        if (el.v <= best.v) best = el;  // earlier indices win ties
    } else if (el.v <= best.v) {
        Pair x;
        if constexpr (NT_PAIRS) x = nt_load_pair(e.suf_b + l); else x = e.suf_b[l];
        if (x.v <= best.v) best = x;
    }
    const Pair er = st[br];
    if (er.i <= r) {
        if (er.v < best.v) best = er;  // later indices must be smaller
    } else if (er.v < best.v) {
        Pair x;
        if constexpr (NT_PAIRS) x = nt_load_pair(e.pre_b + r); else x = e.pre_b[r];
        if (x.v < best.v) best = x;
    }
    return best;
}

template <bool NT>
__device__ __forceinline__ Pair load_pair(const Pair *p) {
    if constexpr (NT) return nt_load_pair(p); else return *p;
}

// Whether two table entries lie in one 128-byte line.
__device__ __forceinline__ bool one_line(const Pair *a, const Pair *b) {
    return ((reinterpret_cast<uintptr_t>(a) ^ reinterpret_cast<uintptr_t>(b)) >> 7) == 0;
}

// A range across blocks bl < br of a table far beyond the caches (the
// lognormal 2^6 batch's bs=8 table is 2.4 GB), where every line read is a
// DRAM request, in three phases; in each every lane picks its addresses
// first and the phase's loads go out together, so lanes on different paths
// do not wait for each other's loads in turn.
// 1. Superset test: the two level-lev entries (lev = floor(log2(br-bl+1)),
//    at bl and a2 = br - 2^lev + 1) when they share a 128-byte line ("in
//    line"; always for a single window, a2 == bl); else the one
//    level-(lev+1) entry at bl when it fits ("wide"), read again as the two
//    level-lev entries when its argmin lies past block br, where it says
//    nothing about bl..br; else the two entries.
// 2. The test failed on one side, so the other side's whole blocks and the
//    failing side's edge pair may answer it: argmin before l -> blocks
//    bl+1..br (in line: the entries at bl+1 and a2 of the line just read,
//    or, when a2 == bl, the one window at bl+1, blocks bl+1..br+1), whose
//    argmin <= r makes it the answer for the whole blocks, with the suffix
//    pair of l loaded beside it; argmin in block br after r -> blocks
//    bl..br-1 (bl-1..br-1) and the prefix pair of r.
// 3. Otherwise the suffix pair of l, the blocks between (from the line when
//    a2 >= bl + 2) and the prefix pair of r.
// A CPU replay of the lognormal 2^6 batch counts 1.31 -> 1.28 DRAM lines per
// range against a superset test followed by the general edge path; the same
// GPU box ran the batch in 1.899 -> 1.858 ms, answers identical.  Earlier
// steps: retrying a failed test with the failing side cut off took it from
// 2.13 to 2.01 ms, the exact re-read of a wide window from 1.927 to 1.870.
// A second copy of rows 1..4, 64 bytes out of line phase, so that the two
// entries always share a line, cut DRAM requests by another 5% but pushed
// the randomly read footprint past the translation caches' reach, and was
// no faster.  On tables that stay in the caches the extra dependent reads
// cost more than they save (lognormal 2^14 at bs=1024), so the host uses
// this path only for tables over kTwoStageTableBytes.
template <bool NT_PAIRS>
__device__ __forceinline__ Pair answer_two_stage(const EdgeTables &e, const Blocks &k, int l, int r, int bl, int br) {
    const Pair *__restrict__ st = k.st;
    const int nb = k.nb, bs = k.bs;
    const int c = br - bl + 1;
    const int lev = 31 - __clz(c);
    const int a2 = br - (1 << lev) + 1;
    const Pair *row = st + (int64_t)lev * nb;
    const bool inl = a2 == bl || one_line(row + bl, row + a2);
    const bool wide = !inl && (int64_t)bl + ((int64_t)2 << lev) <= nb;
    // phase 1
    const Pair *p1 = wide ? row + nb + bl : row + bl;
    const Pair *p2 = wide ? p1 : row + a2;
    Pair x = *p1, y = *p2;
    Pair sup = (y.v < x.v) ? y : x;
    if (wide && (int64_t)sup.i >= (int64_t)(br + 1) * bs) {
        x = row[bl];
        y = row[a2];
        sup = (y.v < x.v) ? y : x;
    }
    if (sup.i >= l && sup.i <= r) return sup;
    // phase 2
    const bool left = sup.i < l;
    const Pair *q1, *q2;
    if (inl && a2 > bl) {
        q1 = left ? row + bl + 1 : row + bl;
        q2 = left ? row + a2 : row + a2 - 1;
    } else if (inl && left && br + 1 < nb) {
        q1 = q2 = row + bl + 1;
    } else if (inl && !left && bl > 0) {
        q1 = q2 = row + bl - 1;
    } else {
        const int a = left ? bl + 1 : bl;
        const int lev1 = 31 - __clz(c - 1);
        q1 = st + (int64_t)lev1 * nb + a;
        q2 = q1 + (c - 1 - (1 << lev1));
    }
    const Pair u = *q1, v = *q2, ed = load_pair<NT_PAIRS>(left ? e.suf_b + l : e.pre_b + r);
    const Pair s2 = (v.v < u.v) ? v : u;
    if (left && s2.i <= r) return (ed.v <= s2.v) ? ed : s2;   // the left part's indices are smaller: it wins ties
    if (!left && s2.i >= l) return (ed.v < s2.v) ? ed : s2;   // the right part's are larger: strictly smaller wins
    // phase 3
    const Pair other = load_pair<NT_PAIRS>(left ? e.pre_b + r : e.suf_b + l);
    Pair best = left ? ed : other;
    if (c > 2) {
        const Pair *m1, *m2;
        if (inl && a2 - bl >= 2) {
            m1 = row + bl + 1;
            m2 = row + a2 - 1;
        } else {
            const int lev2 = 31 - __clz(c - 2);
            m1 = st + (int64_t)lev2 * nb + bl + 1;
            m2 = m1 + (c - 2 - (1 << lev2));
        }
        const Pair g = *m1, h = *m2;
        const Pair mid = (h.v < g.v) ? h : g;
        if (mid.v < best.v) best = mid;
    }
    const Pair rp = left ? other : ed;
    if (rp.v < best.v) best = rp;
    return best;
}

// A range inside block bl: the superset test over its 32-element sub-blocks
// jl..jr, else the sub-block edge pairs and the sub-blocks between.
template <bool NT_PAIRS>
__device__ __forceinline__ Pair answer_in_block(const EdgeTables &e, const Blocks &k, int l, int r, int bl) {
    const int base = bl * k.bs;
    const int jl = (l - base) >> kSubShift, jr = (r - base) >> kSubShift;
    const Pair sup = table_min(e.st_s, (uint32_t)e.subs_total, (int64_t)bl * e.nsub + jl, jr - jl + 1);
    if (sup.i >= l && sup.i <= r) return sup;
    Pair best, right;
    if constexpr (NT_PAIRS) best = nt_load_pair(e.suf_s + l); else best = e.suf_s[l];
    if constexpr (NT_PAIRS) right = nt_load_pair(e.pre_s + r); else right = e.pre_s[r];
    if (jr - jl > 1) {
        const Pair mid = table_min(e.st_s, (uint32_t)e.subs_total, (int64_t)bl * e.nsub + jl + 1, jr - jl - 1);
        if (mid.v < best.v) best = mid;
    }
    if (right.v < best.v) best = right;
    return best;
}

// O(1) answer, from the edge tables, of a range not inside one 32-element
// sub-block: blocks differ -> suffix of the left block, whole blocks between
// (block table), prefix of the right block; same block -> the same three
// parts one level down (the range spans at least two sub-blocks of the
// block).  Parts are combined left to right with strict '<', so the leftmost
// minimum wins.  One lane per query: no reduction needed.  Before any part is
// read, the minimum of a window of whole blocks covering bl..br is the answer
// when its leftmost argmin lies in [l, r] (every element of [l, r] before it
// is strictly larger); otherwise answer_failed.
// LINES: the table is in global memory, where each read costs a 128-byte
// line.  NT_PAIRS: the per-element edge pairs are loaded non-temporally
// (measured: uniform lengths with the table in LDS 0.386 -> 0.342 ms; with
// the table in global memory it costs the lognormal 2^14 batch 10%, so that
// kernel keeps cached loads).
// TWO_STAGE: the phased path for tables beyond the caches
// (answer_two_stage), which the host picks by table size
// (kTwoStageTableBytes).
template <bool LINES = true, bool NT_PAIRS = false, bool TWO_STAGE = false>
__device__ __forceinline__ Pair answer_edges(const EdgeTables &e, const Blocks &k, int l, int r) {
    const int bl = block_of(l, k.div), br = block_of(r, k.div);
    if (bl != br) {
        if constexpr (TWO_STAGE) return answer_two_stage<NT_PAIRS>(e, k, l, r, bl, br);
        const Pair sup = superset_min<LINES>(k.st, k.nb, bl, br);
        if (sup.i >= l && sup.i <= r) return sup;
        return answer_failed<NT_PAIRS>(e, k, l, r, bl, br);
    }
    return answer_in_block<NT_PAIRS>(e, k, l, r, bl);
}

// Coarse blocks of 2^shift blocks, at most kMidMaxBlocks of them, and one
// row over them: entry i is the leftmost minimum of coarse blocks i and i+1
// (the last entry, of the last block alone).  64 KB at most: it sits in LDS
// beside the staging of two 16-wave workgroups per CU, so a range's superset
// test over coarse blocks costs no global read.
constexpr int kMidMaxBlocks = 8192;

// answer_edges with a superset test over the coarse blocks first, from the
// LDS row `mid` (mid_nb entries): a range whose blocks lie in coarse blocks
// cl..cr with cr - cl < 4 takes the entry at cl (one or two coarse blocks)
// or the entries at cl and cr-1 (three or four); when the window's leftmost
// minimum lies in [l, r] it is the answer.  Otherwise, and for a range
// inside one block, answer_edges as it is.
template <bool LINES, bool NT_PAIRS, bool TWO_STAGE>
__device__ __forceinline__ Pair answer_edges_mid(const EdgeTables &e, const Blocks &k, const Pair *mid, int mid_shift,
                                                 int l, int r) {
    const int bl = block_of(l, k.div), br = block_of(r, k.div);
    if (bl != br) {
        const int cl = bl >> mid_shift, cr = br >> mid_shift;
        if (cr - cl < 4) {
            const Pair x = mid[cl];
            const Pair y = mid[cr - cl < 2 ? cl : cr - 1];
            const Pair s = (y.v < x.v) ? y : x;
            if (s.i >= l && s.i <= r) return s;
        }
        if constexpr (TWO_STAGE) return answer_two_stage<NT_PAIRS>(e, k, l, r, bl, br);
        const Pair sup = superset_min<LINES>(k.st, k.nb, bl, br);
        if (sup.i >= l && sup.i <= r) return sup;
        return answer_failed<NT_PAIRS>(e, k, l, r, bl, br);
    }
    return answer_in_block<NT_PAIRS>(e, k, l, r, bl);
}

// Per-wave LDS staging of one tile's queries of a class.
struct TileStage {
    int l[kWave];
    int r[kWave];
    int lane[kWave];
};

// Per-wave LDS slots for one tile's scanned ranges: the lane group that
// answers a range writes its (value, index) at the range's own lane, which
// reads it back, so that every lane ends the tile with its answer in
// registers and the tile's results leave in one store per output.
struct TileResults {
    float v[kWave];
    int i[kWave];
};

// Answers the queries of one class in a tile (mask mk over the tile's lanes),
// G lanes per query; results go to out[qbase + lane].  G == 64: the whole
// wave per query, queries taken from the mask in lane order.  G < 64: the
// queries are staged in LDS and answered 64/G per round, every lane taking
// part in each round's DPP reduction.  Must be called by the whole wave.
// U_SHORT: float4 loads per lane in flight for short ranges with G < 64;
// PRED_SHORT: those loads issued U_SHORT at a time under their own bounds
// tests (scan_range's PRED).
template <int G, bool LONG, bool NT = true, bool TO_LDS = false, int U_SHORT = 1, bool PRED_SHORT = false>
__device__ __forceinline__ void answer_class(const Blocks &k, uint64_t mk, const TileMasks &m, int lane, int64_t qbase,
                                             float *__restrict__ out_v, int32_t *__restrict__ out_i,
                                             TileStage &stage, TileResults *res = nullptr) {
    constexpr int U = (G == kWave) ? (LONG ? kLongUnroll : 2) : (LONG ? 2 : U_SHORT);
    constexpr bool PRED = !LONG && G < kWave && PRED_SHORT;
    if constexpr (LONG) {
        // One lane per query first: a range whose enclosing blocks (across
        // blocks) or enclosing 32-element sub-blocks (inside one block) have
        // their leftmost minimum inside the range is answered here by two
        // table reads -- that minimum is then the range's, and no earlier
        // index of the range can hold it; only the others go through the
        // G-lane path.
        bool done = false;
        if ((mk >> lane) & 1) {
            const int lB = block_of(m.l, k.div), rB = block_of(m.r, k.div);
            Pair sup{INFINITY, -1};
            if (lB != rB) {
                sup = table_min(k.st, k.nb, lB, rB - lB + 1);
            } else if (k.sub) {
                const int base = lB * k.bs;
                const int jl = (m.l - base) >> kSubShift, jr = (m.r - base) >> kSubShift;
                sup = table_min(k.sub, (uint32_t)(k.nb * k.nsub), (int64_t)lB * k.nsub + jl, jr - jl + 1);
            }
            if (sup.i >= m.l && sup.i <= m.r) {
                write_result<NT>(out_v, out_i, qbase + lane, sup.v, sup.i);
                done = true;
            }
        }
        mk &= ~(uint64_t)__ballot(done);
    }
    if constexpr (G == kWave) {
        for (; mk; mk &= mk - 1) {
            const int src = __ffsll((unsigned long long)mk) - 1;
            const int l = __builtin_amdgcn_readlane(m.l, src);
            const int r = __builtin_amdgcn_readlane(m.r, src);
            float bv = INFINITY, mv = INFINITY;
            int bi = INT_MAX, mi = INT_MAX;
            answer_partial<kWave, U, LONG>(k, l, r, lane, bv, bi, mv, mi);
            wave_reduce(bv, bi);
            take_if_better(mv, mi, bv, bi);
            if (lane == 0) {
                if (bv == INFINITY) bi = l;
                if constexpr (TO_LDS) {
                    res->v[src] = bv;
                    res->i[src] = bi;
                } else {
                    write_result<NT>(out_v, out_i, qbase + src, bv, bi);
                }
            }
        }
    } else {
        constexpr int QPR = kWave / G;  // queries per round
        const int g = lane / G;
        const int gl = lane & (G - 1);
        if ((mk >> lane) & 1) {
            const uint32_t rank = mask_rank(mk);
            stage.l[rank] = m.l;
            stage.r[rank] = m.r;
            stage.lane[rank] = lane;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const int count = __popcll(mk);
        for (int base = 0; base < count; base += QPR) {
            const int s = base + g;
            const bool has = s < count;
            float bv = INFINITY, mv = INFINITY;
            int bi = INT_MAX, mi = INT_MAX;
            int ql = 0;
            if (has) {
                ql = stage.l[s];
                answer_partial<G, U, LONG, PRED>(k, ql, stage.r[s], gl, bv, bi, mv, mi);
            }
            group_reduce<G>(bv, bi);
            take_if_better(mv, mi, bv, bi);
            if (has && gl == 0) {
                if (bv == INFINITY) bi = ql;
                if constexpr (TO_LDS) {
                    res->v[stage.lane[s]] = bv;
                    res->i[stage.lane[s]] = bi;
                } else {
                    write_result<NT>(out_v, out_i, qbase + stage.lane[s], bv, bi);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// host plumbing
// ---------------------------------------------------------------------------
inline int device_cus() {
    static int g_num_cus = 0;
    if (g_num_cus > 0) return g_num_cus;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    g_num_cus = cus;
    return cus;
}

// floor(log2(count)) + 1: rows of a sparse table over count entries
inline int table_levels(int64_t count) {
    int levels = 0;
    while (((int64_t)1 << levels) <= count) ++levels;
    return levels;
}

// Device bytes an index over n elements in blocks of bs owns, per allocation
// (rmq_build allocates exactly these; the server's admission check uses
// the same figures):
//   st   -- the block sparse table, levels * nb pairs;
//   sub  -- without edge tables and bs > 32, the sparse table over the
//           32-element sub-block minima, levels_s * nb * nsub pairs;
//   edge -- with edge tables: 2 (bs <= 32) or 4 per-element pair arrays,
//           the sub-block sparse table (levels_s rows of nb * nsub pairs)
//           and two per-sub-block exclusive scans used by the build.
struct HandleSizes {
    int64_t nb, subs;
    int32_t levels, nsub, levels_s, mid_shift, mid_nb;
    int64_t st_bytes, sub_bytes, edge_bytes, mid_bytes;
    int64_t total() const { return st_bytes + sub_bytes + edge_bytes + mid_bytes; }
};

// rmq_build flags: RMQ_BUILD_EDGES, optionally with RMQ_BUILD_MID.
inline bool build_flags_ok(uint32_t flags) {
    if (flags & ~(uint32_t)(RMQ_BUILD_EDGES | RMQ_BUILD_MID)) return false;
    return !(flags & RMQ_BUILD_MID) || (flags & RMQ_BUILD_EDGES);
}

inline HandleSizes handle_sizes(int64_t n, int32_t bs, uint32_t flags) {
    HandleSizes z{};
    z.nb = (n + bs - 1) / bs;
    z.levels = table_levels(z.nb);
    z.nsub = (bs + kSub - 1) / kSub;
    z.subs = z.nb * z.nsub;
    z.levels_s = table_levels(z.nsub);
    z.st_bytes = (int64_t)sizeof(Pair) * z.nb * z.levels;
    if (flags & RMQ_BUILD_EDGES) {
        const int64_t per_elem = bs > kSub ? 4 : 2;
        z.edge_bytes = (int64_t)sizeof(Pair) * (per_elem * n + (int64_t)z.levels_s * z.subs + 2 * z.subs);
    } else if (bs > kSub) {
        z.sub_bytes = (int64_t)sizeof(Pair) * z.subs * z.levels_s;
    }
    if (flags & RMQ_BUILD_MID) {
        while (((z.nb - 1) >> z.mid_shift) + 1 > kMidMaxBlocks) ++z.mid_shift;
        z.mid_nb = (int32_t)(((z.nb - 1) >> z.mid_shift) + 1);
        z.mid_bytes = (int64_t)sizeof(Pair) * z.mid_nb;
    }
    return z;
}

inline int hip_rc(hipError_t e) {
    if (e == hipSuccess) return RMQ_OK;
    if (e == hipErrorOutOfMemory) return RMQ_ERR_OOM;
    (void)hipGetLastError();
    return RMQ_ERR_HIP;
}

#define RMQ_HIP(call)                         \
    do {                                      \
        const int rc_ = hip_rc((call));       \
        if (rc_ != RMQ_OK) return rc_;        \
    } while (0)

}  // namespace rmqx

struct rmq_handle {
    const float *A;
    int64_t n;
    int32_t bs;
    int64_t nb;
    int32_t levels;
    int device;
    rmqx::Pair *st;  // levels * nb pairs; level 0 = block minima
    int64_t st_bytes;
    uint32_t flags;
    rmqx::EdgeTables edges;  // all pointers null unless built with RMQ_BUILD_EDGES
    void *edge_mem;          // one allocation holding every edge array
    int64_t edge_bytes;
    rmqx::Pair *sub;         // sub-block minima (bs > 32): row 0 of edges.st_s, or sub_mem
    int32_t nsub;            // sub-blocks per block, ceil(bs / 32)
    void *sub_mem;           // own allocation when built without edge tables
    int64_t sub_bytes;
    rmqx::Pair *mid;         // RMQ_BUILD_MID: mid_nb coarse entries (blocks i and i+1 each), else null
    int32_t mid_shift;       // a coarse block is 2^mid_shift blocks
    int32_t mid_nb;          // coarse blocks, ceil(nb / 2^mid_shift) <= kMidMaxBlocks
    int64_t mid_bytes;
};

namespace rmqx {

inline Blocks blocks_of(const rmq_handle *h) {
    return Blocks{h->A, h->st, h->sub, block_div(h->bs), (int)h->n, h->bs, (int)h->nb, h->nsub};
}

inline int check_device(const rmq_handle *h) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return RMQ_ERR_HIP;
    return dev == h->device ? RMQ_OK : RMQ_ERR_DEVICE;
}

inline void fill_default_opts(const rmq_handle *h, rmq_query_opts_t *o) {
    (void)h;
    o->short_max = kDefaultShortMax;
    o->group_lanes = kDefaultGroupLanes;
    o->long_wgs = 0;
    o->long_lanes = kDefaultLongLanes;
}

inline int resolve_opts(const rmq_handle *h, const rmq_query_opts_t *in, rmq_query_opts_t *o) {
    fill_default_opts(h, o);
    if (in) {
        if (in->short_max < 0 || in->long_wgs < 0) return RMQ_ERR_INVALID_ARG;
        if (in->short_max > 0) o->short_max = in->short_max;
        auto lanes_ok = [](int g) { return g == 4 || g == 8 || g == 16 || g == 64; };
        if (in->group_lanes != 0) {
            if (!lanes_ok(in->group_lanes)) return RMQ_ERR_INVALID_ARG;
            o->group_lanes = in->group_lanes;
        }
        if (in->long_lanes != 0) {
            if (!lanes_ok(in->long_lanes)) return RMQ_ERR_INVALID_ARG;
            o->long_lanes = in->long_lanes;
        }
        if (in->long_wgs > 0) o->long_wgs = in->long_wgs;
    }
    return RMQ_OK;
}

int query_impl(const rmq_handle *h, const int32_t *d_Q, int64_t q, float *d_val, int32_t *d_idx,
               const rmq_query_opts_t *opts_in, hipStream_t s, rmq_query_times_t *times);

}  // namespace rmqx

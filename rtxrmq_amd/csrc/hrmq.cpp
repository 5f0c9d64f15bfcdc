// hrmq.cpp -- the host range-minimum structure behind the CLI's algs 1 and
// 101 (HRMQ, src/cpu_methods.h:80-159).  The reference links the authors'
// library hrmq/RMQRMM64.cpp (GPL-2.0); this file restates the published
// algorithm -- H. Ferrada and G. Navarro, "Improved range minimum queries",
// J. Discrete Algorithms 43 (2017) -- with a layout of its own:
//
//   * P, a balanced-parentheses sequence of 2(n+1) bits (1 = '(') of the
//     2d-min-heap of A: an opening parenthesis for a virtual root, then for
//     each A[k] in order, a ')' for every stacked j with A[j] >= A[k] (popped),
//     then '(' for k (pushed); finally ')' for everything left on the stack
//     and the root (hrmq/includes/RMQRMM64.h:8-17, hrmq/RMQRMM64.cpp:196-251).
//     Element k's '(' is the (k+2)-th 1 of P.
//   * excess E[p] = (1s - 0s) in P[0..p]; the answer for A[i..j] is the
//     element whose '(' follows the rightmost minimum of E over
//     [x-1, y-1], x = select1(i+2), y = select1(j+2) -- the position of
//     minimum depth (closest to the root) among the opening parentheses in
//     [x, y], rightmost on ties, read back as rank1 - 2
//     (hrmq/RMQRMM64.cpp:1369-1388,1463-1465).  Elements equal to a later
//     one are popped by it, so ties answer the RIGHTMOST minimum index.
//   * rank: a count of 1s per 512-bit superblock plus popcounts; select:
//     the superblock of every 1024th 1, then superblocks, words and bytes;
//     range minima of E: a complete binary min-tree over 256-bit blocks
//     (the range min-max tree), byte tables inside blocks.
//
// Values are compared as int32, as the reference does after reading the
// float array as int* (src/main.cu:117-125): for non-negative floats without
// -0.0 this is the float order.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <exception>
#include <new>
#include <thread>
#include <vector>

#include "rmq.h"

#define RMQ_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int64_t kSuperBits = 512;     // rank directory step (8 words)
constexpr int64_t kBlockBits = 256;     // leaves of the min-tree (4 words)
constexpr int64_t kSelectStep = 1024;   // 1s between select samples

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

// Per byte (bits taken low to high): excess change, minimum of the running
// excess after each bit (relative to before the byte), the rightmost offset
// reaching it, and the offset of the r-th 1.
struct ByteTables {
    int8_t sum[256];
    int8_t mn[256];
    uint8_t pos[256];
    uint8_t sel[256][8];
    ByteTables() {
        for (int v = 0; v < 256; ++v) {
            int e = 0, m = INT_MAX, mp = 0, c = 0;
            for (int b = 0; b < 8; ++b) {
                const bool one = (v >> b) & 1;
                e += one ? 1 : -1;
                if (e <= m) {
                    m = e;
                    mp = b;
                }
                if (one) sel[v][c++] = (uint8_t)b;
            }
            for (; c < 8; ++c) sel[v][c] = 0;
            sum[v] = (int8_t)e;
            mn[v] = (int8_t)m;
            pos[v] = (uint8_t)mp;
        }
    }
};
const ByteTables kT;

}  // namespace

struct rmq_hrmq {
    const int32_t *A = nullptr;
    int64_t n = 0;
    int64_t nbits = 0;
    std::vector<uint64_t> P;
    std::vector<uint32_t> R;      // 1s before each superblock (one extra entry: the total)
    std::vector<uint32_t> S;      // superblock holding the (t * kSelectStep + 1)-th 1
    std::vector<int32_t> tree;    // min-tree over block minima of E; leaves at [L, 2L)
    int64_t L = 1;

    bool bit(int64_t p) const { return (P[p >> 6] >> (p & 63)) & 1; }
    unsigned byte_at(int64_t p) const { return (unsigned)(P[p >> 6] >> (p & 63)) & 0xFFu; }  // p % 8 == 0

    // 1s in P[0..p]
    int64_t rank1(int64_t p) const {
        const int64_t w = p >> 6;
        int64_t r = R[p >> 9];
        for (int64_t k = (p >> 9) << 3; k < w; ++k) r += popc(P[k]);
        return r + popc(P[w] & (~0ULL >> (63 - (p & 63))));
    }

    // E[p - 1]: the excess before position p
    int64_t excess_before(int64_t p) const { return p == 0 ? 0 : 2 * rank1(p - 1) - p; }

    // position of the k-th 1 (k >= 1)
    int64_t select1(int64_t k) const {
        int64_t s = S[(k - 1) / kSelectStep];
        while ((int64_t)R[s + 1] < k) ++s;
        int64_t rem = k - R[s];
        int64_t w = s << 3;
        for (;; ++w) {
            const int c = popc(P[w]);
            if (rem <= c) break;
            rem -= c;
        }
        uint64_t x = P[w];
        int off = 0;
        for (;;) {
            const int c = popc(x & 0xFF);
            if (rem <= c) break;
            rem -= c;
            x >>= 8;
            off += 8;
        }
        return (w << 6) + off + kT.sel[x & 0xFF][rem - 1];
    }

    // Running excess over P[p0..p1] from e = E[p0-1]; (bv, bp) becomes the
    // rightmost minimum seen so far (ties go right).
    void scan(int64_t p0, int64_t p1, int64_t e, int64_t &bv, int64_t &bp) const {
        int64_t p = p0;
        for (; p <= p1 && (p & 7); ++p) {
            e += bit(p) ? 1 : -1;
            if (e <= bv) {
                bv = e;
                bp = p;
            }
        }
        for (; p + 7 <= p1; p += 8) {
            const unsigned b = byte_at(p);
            if (e + kT.mn[b] <= bv) {
                bv = e + kT.mn[b];
                bp = p + kT.pos[b];
            }
            e += kT.sum[b];
        }
        for (; p <= p1; ++p) {
            e += bit(p) ? 1 : -1;
            if (e <= bv) {
                bv = e;
                bp = p;
            }
        }
    }

    // Rightmost block in [lo, hi] whose minimum is the range's minimum, and
    // that minimum.
    int64_t tree_rightmost_min(int64_t lo, int64_t hi, int64_t *vmin) const {
        int64_t l = lo + L, r = hi + L + 1;
        int64_t left[64], right[64];
        int nl = 0, nr = 0;
        int32_t vm = INT32_MAX;
        while (l < r) {
            if (l & 1) {
                vm = std::min(vm, tree[l]);
                left[nl++] = l++;
            }
            if (r & 1) {
                --r;
                vm = std::min(vm, tree[r]);
                right[nr++] = r;
            }
            l >>= 1;
            r >>= 1;
        }
        int64_t node = -1;
        for (int k = 0; k < nr && node < 0; ++k)  // right[] runs right to left
            if (tree[right[k]] == vm) node = right[k];
        for (int k = nl - 1; k >= 0 && node < 0; --k)  // left[] runs left to right
            if (tree[left[k]] == vm) node = left[k];
        while (node < L) node = tree[2 * node + 1] == vm ? 2 * node + 1 : 2 * node;
        *vmin = vm;
        return node - L;
    }

    // Rightmost minimum of E over [a, b]: (value, position).
    void min_excess(int64_t a, int64_t b, int64_t &bv, int64_t &bp) const {
        bv = INT64_MAX;
        bp = -1;
        const int64_t ba = a / kBlockBits, bb = b / kBlockBits;
        if (ba == bb) {
            scan(a, b, excess_before(a), bv, bp);
            return;
        }
        scan(a, (ba + 1) * kBlockBits - 1, excess_before(a), bv, bp);
        if (bb - ba > 1) {
            int64_t vm = 0;
            const int64_t blk = tree_rightmost_min(ba + 1, bb - 1, &vm);
            if (vm <= bv) {  // the middle's rightmost minimum lies right of the left part's
                const int64_t s0 = blk * kBlockBits;
                bv = INT64_MAX;
                scan(s0, s0 + kBlockBits - 1, excess_before(s0), bv, bp);
            }
        }
        const int64_t s1 = bb * kBlockBits;
        scan(s1, b, excess_before(s1), bv, bp);
    }

    int32_t query(int32_t i, int32_t j) const {
        if (i == j) return i;
        const int64_t x = select1((int64_t)i + 2), y = select1((int64_t)j + 2);
        int64_t bv = 0, bp = 0;
        min_excess(x - 1, y - 1, bv, bp);
        // the answer's '(' is at bp + 1; rank1(bp) = (E[bp] + bp + 1) / 2
        return (int32_t)((bv + bp + 1) / 2 - 1);
    }

    int64_t bytes() const {
        return (int64_t)(P.size() * sizeof(uint64_t) + R.size() * sizeof(uint32_t) + S.size() * sizeof(uint32_t) +
                         tree.size() * sizeof(int32_t));
    }

    int build(const int32_t *a, int64_t len) {
        A = a;
        n = len;
        nbits = 2 * (n + 1);
        const int64_t supers = (nbits + kSuperBits - 1) / kSuperBits;
        P.assign((size_t)(supers * (kSuperBits / 64)), 0ULL);  // whole superblocks: rank reads stay in bounds
        // the 2d-min-heap's parentheses, with an explicit stack
        std::vector<int32_t> stack;
        stack.reserve(1024);
        int64_t pos = 0;
        auto open = [&]() {
            P[pos >> 6] |= 1ULL << (pos & 63);
            ++pos;
        };
        open();  // the virtual root
        for (int64_t k = 0; k < n; ++k) {
            const int32_t v = A[k];
            while (!stack.empty() && A[stack.back()] >= v) {
                stack.pop_back();
                ++pos;  // ')'
            }
            open();
            stack.push_back((int32_t)k);
        }
        pos += (int64_t)stack.size() + 1;  // close the stack and the root
        if (pos != nbits) return RMQ_ERR_INVALID_ARG;
        // rank directory and select samples
        R.assign((size_t)supers + 1, 0u);
        int64_t ones = 0;
        for (int64_t s = 0; s < supers; ++s) {
            R[s] = (uint32_t)ones;
            for (int64_t w = s * 8; w < s * 8 + 8; ++w) ones += popc(P[w]);
        }
        R[supers] = (uint32_t)ones;
        if (ones != n + 1) return RMQ_ERR_INVALID_ARG;
        S.assign((size_t)((ones + kSelectStep - 1) / kSelectStep), 0u);
        for (int64_t s = 0, t = 0; s < supers; ++s)
            while (t < (int64_t)S.size() && t * kSelectStep + 1 <= (int64_t)R[s + 1]) S[t++] = (uint32_t)s;
        // min-tree over the blocks' minima of E
        const int64_t blocks = (nbits + kBlockBits - 1) / kBlockBits;
        while (L < blocks) L <<= 1;
        tree.assign((size_t)(2 * L), INT32_MAX);
        int64_t e = 0;
        for (int64_t b = 0; b < blocks; ++b) {
            const int64_t p0 = b * kBlockBits, p1 = std::min(nbits, p0 + kBlockBits) - 1;
            int64_t bv = INT64_MAX, bp = -1;
            scan(p0, p1, e, bv, bp);
            tree[L + b] = (int32_t)bv;
            for (int64_t w = p0 >> 6; w <= p1 >> 6; ++w) e += 2 * popc(P[w]) - 64;  // whole words: P is zero-padded
        }
        for (int64_t v = L - 1; v >= 1; --v) tree[v] = std::min(tree[2 * v], tree[2 * v + 1]);
        return RMQ_OK;
    }
};

RMQ_API int rmq_hrmq_build(const int32_t *A, int64_t n, rmq_hrmq_t **out) {
    if (!out) return RMQ_ERR_INVALID_ARG;
    *out = nullptr;
    if (!A || n <= 0 || n > INT32_MAX - 1) return RMQ_ERR_INVALID_ARG;
    rmq_hrmq *h = new (std::nothrow) rmq_hrmq();
    if (!h) return RMQ_ERR_OOM;
    int rc;
    try {
        rc = h->build(A, n);
    } catch (const std::bad_alloc &) {
        rc = RMQ_ERR_OOM;
    }
    if (rc != RMQ_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return RMQ_OK;
}

RMQ_API int rmq_hrmq_query(const rmq_hrmq_t *h, const int32_t *Q, int64_t q, int32_t *idx, int32_t *val,
                           int32_t nthreads) {
    if (!h || q < 0 || (q > 0 && (!Q || (!idx && !val))) || nthreads < 0) return RMQ_ERR_INVALID_ARG;
    if (q == 0) return RMQ_OK;
    auto work = [&](int64_t lo, int64_t hi) {
        for (int64_t j = lo; j < hi; ++j) {
            const int32_t l = Q[2 * j], r = Q[2 * j + 1];
            int32_t k = -1, v = 0x7FC00000;  // invalid ranges: index -1, value NaN bits
            if (l >= 0 && l <= r && r < h->n) {
                k = h->query(l, r);
                v = h->A[k];
            }
            if (idx) idx[j] = k;
            if (val) val[j] = v;
        }
    };
    int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    if (q < 4096) nt = 1;
    if (nt == 1) {
        work(0, q);
        return RMQ_OK;
    }
    // static chunks, as OpenMP's default schedule in the reference's loop; a
    // chunk whose thread cannot be started runs on the calling thread
    std::vector<std::thread> pool;
    const int64_t chunk = (q + nt - 1) / nt;
    try {
        pool.reserve((size_t)nt);
    } catch (const std::exception &) {
    }
    for (int t = 1; t < nt; ++t) {
        const int64_t lo = t * chunk, hi = std::min<int64_t>(q, lo + chunk);
        if (lo >= hi) continue;
        try {
            pool.emplace_back(work, lo, hi);
        } catch (const std::exception &) {
            work(lo, hi);
        }
    }
    work(0, std::min<int64_t>(q, chunk));
    for (auto &th : pool) th.join();
    return RMQ_OK;
}

RMQ_API int rmq_hrmq_bytes(const rmq_hrmq_t *h, int64_t *bytes) {
    if (!h || !bytes) return RMQ_ERR_INVALID_ARG;
    *bytes = h->bytes();
    return RMQ_OK;
}

RMQ_API void rmq_hrmq_free(rmq_hrmq_t *h) { delete h; }

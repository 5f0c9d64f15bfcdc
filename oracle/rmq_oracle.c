/*
 * rmq_oracle.c -- CPU oracle for the batch range-minimum engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product path (rtxrmq_amd/) never links, loads or calls it.
 *
 * What each function restates, and what pins it (tests/test_oracle.py):
 *
 *   orc_cpu_min / orc_cpu_min_idx
 *       CPU BASE, algs 0/100: src/cpu_methods.h:7-16 (value) and :42-53
 *       (leftmost index).  Start at A[l], strict '<' scanning left to right,
 *       so the result is the leftmost minimum and its exact float bits.
 *       Pinned by: golden vectors from the real HRMQ (value column), closed
 *       forms (constant array -> l, strictly descending -> r, [5,5] -> l,
 *       [3,5,3] -> l), and numpy argmin (first occurrence) on random data.
 *
 *   orc_hrmq_idx
 *       HRMQ index semantics, alg 101: the index RMQRMM64::queryRMQ returns
 *       (hrmq/RMQRMM64.cpp:1463-1465) for the 2d-min-heap built with the
 *       '>=' pop rule (hrmq/RMQRMM64.cpp:211), over the float bits read as
 *       int32 (src/main.cu:118,123).  That is the RIGHTMOST minimum under
 *       int32 comparison.  Pinned by golden vectors written by the real HRMQ
 *       library compiled from the reference sources (tests/golden/).
 *
 *   orc_rmq_blocks
 *       The RTX_blocks decomposition, algs 5/105: block minima
 *       (src/device_tools.cuh:243-253, src/rtx.h:3-22), then per query either
 *       one in-block search (src/rtx_kernels.cu:116-135) or the whole blocks
 *       in between (:138-149) plus the left (:151-163) and right (:165-177)
 *       edge blocks, combined in index order so ties resolve leftmost.  The
 *       whole-block search uses a sparse table over block minima.
 *       Pinned by: exhaustive all-(l,r) comparison with orc_cpu_min_idx on
 *       small arrays for many block sizes (tests/test_oracle.py).
 *
 *   orc_fast_*
 *       An O(1)-per-query exact RMQ (per-block prefix/suffix argmins plus a
 *       sparse table over block argmins) so full-size runs (n=1e8, q=2^26)
 *       can be checked bit for bit on the host in seconds.  Not a reference
 *       restatement; pinned by exhaustive comparison with orc_cpu_min_idx on
 *       small arrays and by random comparison on medium arrays.
 *
 *   orc_gen_*
 *       The deterministic input generators (counter-based splitmix64).  The
 *       reference's generators are src/rand.cuh:7-43 (array) and :116-201
 *       (queries); these follow the same roles (array kinds, lr codes) with
 *       a stream that depends only on (seed, index), never on thread count.
 *       Pinned by distribution checks (range, mean, medians) in the tests.
 *
 * Preconditions (same as the reference's usable domain): inputs contain no
 * NaN; 0 <= l <= r < n; n <= 2^31-1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* CPU BASE (src/cpu_methods.h:7-16, :42-53)                                 */
/* ------------------------------------------------------------------------ */

ORC_API float orc_cpu_min(const float *A, int32_t l, int32_t r) {
    float m = A[l];
    for (int32_t i = l; i <= r; ++i)
        if (A[i] < m) m = A[i];
    return m;
}

ORC_API int32_t orc_cpu_min_idx(const float *A, int32_t l, int32_t r) {
    float m = A[l];
    int32_t mi = l;
    for (int32_t i = l; i <= r; ++i)
        if (A[i] < m) { m = A[i]; mi = i; }
    return mi;
}

/* Rightmost minimum under int32 comparison of the float bit patterns: the
 * answer HRMQ gives (hrmq/RMQRMM64.cpp:211 pops on A[top] >= A[i], so among
 * equal values the later one becomes the ancestor; queryRMQ at :1463-1465
 * returns that ancestor). */
ORC_API int32_t orc_hrmq_idx(const int32_t *Ai, int32_t l, int32_t r) {
    int32_t m = Ai[l];
    int32_t mi = l;
    for (int32_t i = l; i <= r; ++i)
        if (Ai[i] <= m) { m = Ai[i]; mi = i; }
    return mi;
}

/* Batched brute force with OpenMP over queries, as src/cpu_methods.h:28-31
 * and :65-68 do.  val / idx may be NULL. */
ORC_API void orc_rmq_brute(const float *A, int64_t n, const int32_t *Q, int64_t q,
                           float *val, int32_t *idx, int nt) {
    (void)n;
#ifdef _OPENMP
    if (nt > 0) omp_set_num_threads(nt);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t j = 0; j < q; ++j) {
        int32_t mi = orc_cpu_min_idx(A, Q[2 * j], Q[2 * j + 1]);
        if (idx) idx[j] = mi;
        if (val) val[j] = A[mi];
    }
}

ORC_API void orc_rmq_brute_hrmq(const int32_t *Ai, int64_t n, const int32_t *Q, int64_t q,
                                int32_t *idx, int nt) {
    (void)n;
#ifdef _OPENMP
    if (nt > 0) omp_set_num_threads(nt);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t j = 0; j < q; ++j)
        idx[j] = orc_hrmq_idx(Ai, Q[2 * j], Q[2 * j + 1]);
}

/* ------------------------------------------------------------------------ */
/* (value, index) pairs with the leftmost tie rule                           */
/* ------------------------------------------------------------------------ */

typedef struct { float v; int32_t i; } orc_pair;

/* 'b' replaces 'a' when strictly smaller, or equal with a smaller index. */
static inline orc_pair pick(orc_pair a, orc_pair b) {
    if (b.v < a.v) return b;
    if (!(a.v < b.v) && b.i < a.i) return b;
    return a;
}

static int32_t ilog2_i64(int64_t x) {
    int32_t k = -1;
    while (x) { x >>= 1; ++k; }
    return k;
}

/* Sparse table over 'base' (nb pairs): level k covers [i, i+2^k). */
typedef struct {
    int64_t nb;
    int32_t levels;
    orc_pair *t; /* levels * nb */
} orc_sparse;

static int orc_sparse_build(orc_sparse *s, const orc_pair *base, int64_t nb) {
    s->nb = nb;
    s->levels = nb > 0 ? ilog2_i64(nb) + 1 : 0;
    s->t = (orc_pair *)malloc(sizeof(orc_pair) * (size_t)(s->levels > 0 ? s->levels : 1) * (size_t)(nb > 0 ? nb : 1));
    if (!s->t) return -1;
    if (nb > 0) memcpy(s->t, base, sizeof(orc_pair) * (size_t)nb);
    for (int32_t k = 1; k < s->levels; ++k) {
        const orc_pair *prev = s->t + (size_t)(k - 1) * nb;
        orc_pair *cur = s->t + (size_t)k * nb;
        int64_t half = (int64_t)1 << (k - 1);
        for (int64_t i = 0; i + ((int64_t)1 << k) <= nb; ++i)
            cur[i] = pick(prev[i], prev[i + half]);
    }
    return 0;
}

/* Leftmost minimum over base[a..b], a <= b. */
static orc_pair orc_sparse_query(const orc_sparse *s, int64_t a, int64_t b) {
    int32_t k = ilog2_i64(b - a + 1);
    const orc_pair *lv = s->t + (size_t)k * s->nb;
    orc_pair x = lv[a];
    orc_pair y = lv[b - ((int64_t)1 << k) + 1];
    /* the two windows overlap; x holds the left one so it wins ties */
    return (y.v < x.v) ? y : x;
}

/* ------------------------------------------------------------------------ */
/* RTX_blocks decomposition (algs 5/105)                                     */
/* ------------------------------------------------------------------------ */

static orc_pair scan_pair(const float *A, int64_t a, int64_t b) {
    orc_pair p = {A[a], (int32_t)a};
    for (int64_t i = a + 1; i <= b; ++i)
        if (A[i] < p.v) { p.v = A[i]; p.i = (int32_t)i; }
    return p;
}

ORC_API int orc_rmq_blocks(const float *A, int64_t n, int32_t bs, const int32_t *Q, int64_t q,
                           float *val, int32_t *idx, int nt) {
    if (n <= 0 || bs <= 0) return -1;
    int64_t nb = (n + bs - 1) / bs; /* src/rtx.h:42 */
    orc_pair *blk = (orc_pair *)malloc(sizeof(orc_pair) * (size_t)nb);
    if (!blk) return -1;
    for (int64_t b = 0; b < nb; ++b) {
        int64_t lo = b * bs, hi = lo + bs - 1;
        if (hi > n - 1) hi = n - 1; /* last partial block, clamped */
        blk[b] = scan_pair(A, lo, hi);
    }
    orc_sparse st;
    if (orc_sparse_build(&st, blk, nb)) { free(blk); return -1; }
#ifdef _OPENMP
    if (nt > 0) omp_set_num_threads(nt);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t j = 0; j < q; ++j) {
        int64_t l = Q[2 * j], r = Q[2 * j + 1];
        int64_t lB = l / bs, rB = r / bs;
        orc_pair best;
        if (lB == rB) {
            best = scan_pair(A, l, r);
        } else {
            best = scan_pair(A, l, (lB + 1) * bs - 1);           /* left edge   */
            if (lB < rB - 1) {                                    /* whole blocks */
                orc_pair mid = orc_sparse_query(&st, lB + 1, rB - 1);
                if (mid.v < best.v) best = mid;
            }
            orc_pair right = scan_pair(A, rB * bs, r);            /* right edge  */
            if (right.v < best.v) best = right;
        }
        if (idx) idx[j] = best.i;
        if (val) val[j] = best.v;
    }
    free(st.t);
    free(blk);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Fast exact oracle: per-block prefix/suffix argmins + sparse table         */
/* ------------------------------------------------------------------------ */

typedef struct {
    const float *A;
    int64_t n;
    int32_t b;      /* block size */
    int32_t *pre;   /* pre[i] = leftmost argmin of A[blockstart(i) .. i] */
    int32_t *suf;   /* suf[i] = leftmost argmin of A[i .. blockend(i)]   */
    orc_sparse st;  /* over block argmins */
} orc_fast;

ORC_API void *orc_fast_build(const float *A, int64_t n, int32_t b, int nt) {
    if (n <= 0 || b <= 0) return NULL;
    orc_fast *f = (orc_fast *)calloc(1, sizeof(orc_fast));
    if (!f) return NULL;
    f->A = A; f->n = n; f->b = b;
    f->pre = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    f->suf = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int64_t nb = (n + b - 1) / b;
    orc_pair *blk = (orc_pair *)malloc(sizeof(orc_pair) * (size_t)nb);
    if (!f->pre || !f->suf || !blk) { free(f->pre); free(f->suf); free(blk); free(f); return NULL; }
#ifdef _OPENMP
    if (nt > 0) omp_set_num_threads(nt);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t B = 0; B < nb; ++B) {
        int64_t lo = B * b, hi = lo + b - 1;
        if (hi > n - 1) hi = n - 1;
        int32_t m = (int32_t)lo;
        for (int64_t i = lo; i <= hi; ++i) {
            if (A[i] < A[m]) m = (int32_t)i;
            f->pre[i] = m;
        }
        m = (int32_t)hi;
        for (int64_t i = hi; i >= lo; --i) {
            if (!(A[m] < A[i])) m = (int32_t)i; /* '<=' going left keeps leftmost */
            f->suf[i] = m;
        }
        blk[B].v = A[f->pre[hi]];
        blk[B].i = f->pre[hi];
    }
    if (orc_sparse_build(&f->st, blk, nb)) { free(f->pre); free(f->suf); free(blk); free(f); return NULL; }
    free(blk);
    return f;
}

ORC_API void orc_fast_query(const void *h, const int32_t *Q, int64_t q, float *val, int32_t *idx, int nt) {
    const orc_fast *f = (const orc_fast *)h;
    const float *A = f->A;
    const int64_t b = f->b;
#ifdef _OPENMP
    if (nt > 0) omp_set_num_threads(nt);
#pragma omp parallel for schedule(static, 4096)
#endif
    for (int64_t j = 0; j < q; ++j) {
        int64_t l = Q[2 * j], r = Q[2 * j + 1];
        int64_t lB = l / b, rB = r / b;
        orc_pair best;
        if (lB == rB) {
            best = scan_pair(A, l, r);
        } else {
            best.i = f->suf[l]; best.v = A[best.i];
            if (lB < rB - 1) {
                orc_pair mid = orc_sparse_query(&f->st, lB + 1, rB - 1);
                if (mid.v < best.v) best = mid;
            }
            orc_pair right; right.i = f->pre[r]; right.v = A[right.i];
            if (right.v < best.v) best = right;
        }
        if (idx) idx[j] = best.i;
        if (val) val[j] = best.v;
    }
}

ORC_API void orc_fast_free(void *h) {
    orc_fast *f = (orc_fast *)h;
    if (!f) return;
    free(f->pre); free(f->suf); free(f->st.t); free(f);
}

/* ------------------------------------------------------------------------ */
/* Deterministic generators                                                  */
/* ------------------------------------------------------------------------ */
/*
 * r(seed, stream, i) = mix64(key(seed, stream) + (i + 1) * PHI), with
 * key(seed, stream) = mix64(seed * PHI + stream) and mix64 the splitmix64
 * finaliser.  Streams: 0 array values, 1 left endpoints, 2 uniform lengths,
 * 3/4 the two Box-Muller uniforms.
 *
 * Array kinds (src/rand.cuh:15-21 draws U(0,1]; these draw exact grid values):
 *   0 uniform:  A[i] = k * 2^-24, k = top 24 bits of r      (U[0,1) exactly)
 *   1 ties16:   A[i] = (float)(top 4 bits of r)             ({0..15})
 *   2 desc:     bits(A[i]) = 0x3F800000 - 1 - i              (strictly descending)
 *   3 const:    A[i] = 1.0f
 * Query codes (lr, after src/rand.cuh:179-201):
 *   lr > 0   fixed length lr, l uniform in [0, n-lr]
 *   lr = -1  l uniform in [0,n), length uniform in [1, n-l]
 *   lr = -2  l uniform, length round(exp(N(ln 2^14, 1))), clamped to [1, n-l]
 *   lr = -3  l uniform, length round(exp(N(ln 2^6, 1))), clamped to [1, n-l]
 *   lr = -4  l uniform, length round(exp(N(ln max(1,n/2^8), 0.3))), clamped
 *   lr = -5  l uniform, length round(exp(N(ln max(1,n/2^15), 0.3))), clamped
 *   lr = -6  l uniform, r = n-1
 *   lr = -7  l uniform, length round(exp(N(14/log2(1e8) * ln n, 1))), clamped
 *   lr = -8  l uniform, length round(exp(N(6/log2(1e8) * ln n, 1))), clamped
 *            (the -2 / -3 medians 2^14, 2^6 at n=1e8, scaled to n)
 */
#define PHI 0x9E3779B97F4A7C15ULL

static inline uint64_t mix64(uint64_t z) {
    z ^= z >> 30; z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27; z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}

static inline uint64_t rng_key(uint64_t seed, uint64_t stream) { return mix64(seed * PHI + stream); }
static inline uint64_t rng_at(uint64_t key, uint64_t i) { return mix64(key + (i + 1) * PHI); }

/* uniform integer in [0, m) by 64x64->128 multiply-high */
static inline int64_t below(uint64_t r, uint64_t m) {
    return (int64_t)(((unsigned __int128)r * m) >> 64);
}

ORC_API int orc_gen_array(float *A, int64_t n, uint64_t seed, int32_t kind) {
    uint64_t key = rng_key(seed, 0);
    if (kind == 2 && n > 0x3F7FFFFFLL) return -1;
    for (int64_t i = 0; i < n; ++i) {
        switch (kind) {
        case 0: A[i] = (float)(rng_at(key, (uint64_t)i) >> 40) * 0x1p-24f; break;
        case 1: A[i] = (float)(rng_at(key, (uint64_t)i) >> 60); break;
        case 2: { uint32_t bits = 0x3F800000u - 1u - (uint32_t)i; memcpy(&A[i], &bits, 4); } break;
        case 3: A[i] = 1.0f; break;
        default: return -1;
        }
    }
    return 0;
}

static double lognormal_len(uint64_t k3, uint64_t k4, uint64_t i, double mu, double sigma) {
    double u1 = (double)((rng_at(k3, i) >> 11) + 1) * 0x1p-53; /* (0,1] */
    double u2 = (double)(rng_at(k4, i) >> 11) * 0x1p-53;       /* [0,1) */
    double rad = sqrt(-2.0 * log(u1));
    double ang = 6.283185307179586476925286766559 * u2;
    double z = rad * cos(ang);
    double s = sigma * z;
    double e = mu + s;
    return exp(e);
}

ORC_API int orc_gen_queries(int32_t *Q, int64_t q, int32_t n, int32_t lr, uint64_t seed, int nt) {
    if (n <= 0 || lr == 0 || lr > n || lr < -8) return -1;
    uint64_t k1 = rng_key(seed, 1), k2 = rng_key(seed, 2), k3 = rng_key(seed, 3), k4 = rng_key(seed, 4);
    double mu = 0.0, sigma = 1.0;
    if (lr == -2) mu = 14.0 * 0.69314718055994530942;
    if (lr == -3) mu = 6.0 * 0.69314718055994530942;
    if (lr == -4) { int64_t s = n / 256; mu = log((double)(s > 1 ? s : 1)); sigma = 0.3; }
    if (lr == -5) { int64_t s = n / 32768; mu = log((double)(s > 1 ? s : 1)); sigma = 0.3; }
    if (lr == -7) mu = 14.0 / log2(1e8) * log((double)n);
    if (lr == -8) mu = 6.0 / log2(1e8) * log((double)n);
#ifdef _OPENMP
    if (nt > 0) omp_set_num_threads(nt);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < q; ++i) {
        int64_t l, len;
        uint64_t ui = (uint64_t)i;
        if (lr > 0) {
            len = lr;
            l = below(rng_at(k1, ui), (uint64_t)(n - lr + 1));
        } else {
            l = below(rng_at(k1, ui), (uint64_t)n);
            int64_t room = (int64_t)n - l;
            if (lr == -1) {
                len = 1 + below(rng_at(k2, ui), (uint64_t)room);
            } else if (lr == -6) {
                len = room;
            } else {
                double x = lognormal_len(k3, k4, ui, mu, sigma);
                if (!(x < 4.0e18)) x = 4.0e18;
                len = llround(x);
                if (len < 1) len = 1;
                if (len > room) len = room;
            }
        }
        Q[2 * i] = (int32_t)l;
        Q[2 * i + 1] = (int32_t)(l + len - 1);
    }
    return 0;
}

// hrmq_shim.cpp -- C entry points over the reference's HRMQ (RMQRMM64).
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile together with the
// reference's own hrmq/RMQRMM64.cpp and hrmq/includes/Basic_rmq.cpp, where
// they lie under the reference tree, into oracle/_ref/libhrmq.so.  Used to
// write golden vectors (tests/golden/) and as the timed host baseline in
// bench.py's cpu_baseline leg.  Never part of the product.
//
// The query loop follows src/cpu_methods.h:104-110 (value) and :145-150
// (index): an OpenMP parallel-for over queries calling queryRMQ(l, r), with
// the array read as int32 bit patterns (src/main.cu:118,123).
#include <cstdint>
#include <cstdlib>
#include <new>
#include <omp.h>
#include "RMQRMM64.h"

// The reference allocates its bit vector without clearing it
// (hrmq/RMQRMM64.cpp:53, `P = new ulong[lenP]`) and sets only the first
// 2(n+1) bits, so the padding bits of the last word -- and any table word it
// never writes -- hold whatever the heap held before.  In a long-running
// process with recycled heap memory (or under MALLOC_PERTURB_) queries then
// read garbage and can fault.  These replacements, bound inside this library
// only (-Bsymbolic in oracle/Makefile), zero every allocation the library
// makes, so HRMQ's answers no longer depend on heap contents.
#define HRMQ_LOCAL __attribute__((visibility("hidden")))
HRMQ_LOCAL void *operator new(std::size_t sz) {
    void *p = std::calloc(1, sz ? sz : 1);
    if (!p) throw std::bad_alloc();
    return p;
}
HRMQ_LOCAL void *operator new[](std::size_t sz) { return operator new(sz); }
HRMQ_LOCAL void operator delete(void *p) noexcept { std::free(p); }
HRMQ_LOCAL void operator delete[](void *p) noexcept { std::free(p); }
HRMQ_LOCAL void operator delete(void *p, std::size_t) noexcept { std::free(p); }
HRMQ_LOCAL void operator delete[](void *p, std::size_t) noexcept { std::free(p); }

using namespace rmqrmm;

extern "C" {

__attribute__((visibility("default"))) void *hrmq_build(const int32_t *A, int64_t n) {
    // src/cpu_methods.h:92 -- single-threaded construction
    return new RMQRMM64(const_cast<int *>(reinterpret_cast<const int *>(A)), (unsigned long)n);
}

__attribute__((visibility("default"))) uint32_t hrmq_size_bytes(void *h) {
    return static_cast<RMQRMM64 *>(h)->getSize();
}

__attribute__((visibility("default"))) int64_t hrmq_query(void *h, int64_t l, int64_t r) {
    return (int64_t) static_cast<RMQRMM64 *>(h)->queryRMQ((unsigned long)l, (unsigned long)r);
}

// idx_out and/or val_out may be NULL.  val_out receives A[idx] (the int32
// bits, reinterpreted as float by the caller, as src/main.cu:120 does).
__attribute__((visibility("default"))) void hrmq_query_batch(void *h, const int32_t *A, const int32_t *Q,
                                                              int64_t q, int32_t *idx_out, int32_t *val_out,
                                                              int nt) {
    RMQRMM64 *rmq = static_cast<RMQRMM64 *>(h);
    if (nt > 0) omp_set_num_threads(nt);
#pragma omp parallel for
    for (int64_t j = 0; j < q; ++j) {
        int32_t idx = (int32_t)rmq->queryRMQ((unsigned long)Q[2 * j], (unsigned long)Q[2 * j + 1]);
        if (idx_out) idx_out[j] = idx;
        if (val_out) val_out[j] = A[idx];
    }
}

__attribute__((visibility("default"))) void hrmq_free(void *h) { delete static_cast<RMQRMM64 *>(h); }

}  // extern "C"

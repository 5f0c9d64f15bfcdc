"""Write the golden RMQ fixtures under tests/golden/ (run in the build container).

    make -C oracle && python tests/golden/make_golden.py

Each fixture (.npz, no pickles) holds an array A, queries Q (int32 {l, r}
pairs) and expected answers computed WITHOUT the repository's oracle:
  val, idx         leftmost minimum by numpy.argmin (first occurrence) on A[l:r+1]
                   -- the CPU BASE contract, src/cpu_methods.h:7-53;
  hrmq_idx,        what the reference's own HRMQ library (RMQRMM64, compiled from
  hrmq_val         hrmq/ by oracle/Makefile into oracle/_ref/libhrmq.so) returns:
                   queryRMQ(l, r) (hrmq/RMQRMM64.cpp:1463-1465) and A[idx] read as
                   int32 bits (src/cpu_methods.h:107-108, src/main.cu:118-120).
                   Present only where HRMQ's input contract holds (non-negative
                   floats, no -0.0: the bits are compared as int32).
The reference ships no golden vectors of its own; these pin the oracle.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402  (the real HRMQ wrapper lives there)


def leftmost(A: np.ndarray, Q: np.ndarray):
    idx = np.array([l + int(np.argmin(A[l:r + 1])) for l, r in Q], dtype=np.int32)
    return A[idx].copy(), idx


def all_pairs(n: int) -> np.ndarray:
    return np.array([(l, r) for l in range(n) for r in range(l, n)], dtype=np.int32)


def random_queries(rng, n: int, q: int) -> np.ndarray:
    l = rng.integers(0, n, q)
    kind = rng.integers(0, 4, q)
    ln = np.where(kind == 0, 1, np.where(kind == 1, rng.integers(1, 33, q),
                                         np.where(kind == 2, rng.integers(1, n + 1, q), n - l)))
    ln = np.clip(ln, 1, n - l)
    Q = np.stack([l, l + ln - 1], 1).astype(np.int32)
    extra = np.array([[0, n - 1], [0, 0], [n - 1, n - 1]], dtype=np.int32)
    return np.concatenate([Q, extra])


def cases():
    rng = np.random.default_rng(20241203)
    out = {}
    # U[0,1) on the 2^-24 grid, unique minima almost everywhere
    A = (rng.integers(0, 1 << 24, 1000) * 2.0 ** -24).astype(np.float32)
    out["uniform_n1000"] = (A, random_queries(rng, 1000, 4000), True)
    # heavy ties: integers 0..15
    A = rng.integers(0, 16, 5000).astype(np.float32)
    out["ties16_n5000"] = (A, random_queries(rng, 5000, 6000), True)
    # very heavy ties, every (l, r)
    A = rng.integers(0, 3, 60).astype(np.float32)
    out["ties3_n60_allpairs"] = (A, all_pairs(60), True)
    # strictly descending positive floats, every (l, r): min at r
    A = (np.uint32(0x3F800000) - 1 - np.arange(64, dtype=np.uint32)).view(np.float32)
    out["desc_n64_allpairs"] = (A, all_pairs(64), True)
    # strictly ascending: min at l
    A = np.arange(1, 65, dtype=np.float32)
    out["asc_n64_allpairs"] = (A, all_pairs(64), True)
    # constant array: leftmost = l, HRMQ = r
    A = np.full(50, 7.0, np.float32)
    out["const_n50_allpairs"] = (A, all_pairs(50), True)
    # closed forms: [5,5] -> leftmost 0 / rightmost 1, [3,5,3] -> leftmost 0 / rightmost 2
    out["closed_5_5"] = (np.array([5, 5], np.float32), all_pairs(2), True)
    out["closed_3_5_3"] = (np.array([3, 5, 3], np.float32), all_pairs(3), True)
    # signed zeros and infinities (outside HRMQ's int32 contract)
    A = np.array([0.0, -0.0, 1.0, -0.0, np.inf, np.inf, -np.inf, 2.0, -np.inf, 0.0], np.float32)
    out["special_n10_allpairs"] = (A, all_pairs(10), False)
    A = np.full(37, np.inf, np.float32)
    out["allinf_n37_allpairs"] = (A, all_pairs(37), False)
    return out


def main():
    if not O.HRMQ.available():
        raise SystemExit("oracle/_ref/libhrmq.so missing: run make -C oracle first")
    for name, (A, Q, hrmq_ok) in cases().items():
        val, idx = leftmost(A, Q)
        rec = {"A": A, "Q": Q, "val": val, "idx": idx}
        if hrmq_ok:
            h = O.HRMQ(A)
            hv, hi = h.query(Q, nt=1)
            h.close()
            rec["hrmq_idx"] = hi
            rec["hrmq_val"] = hv
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **rec)
        print(f"{name}: n={A.size} q={Q.shape[0]} hrmq={'yes' if hrmq_ok else 'n/a'} -> {os.path.basename(path)}")


if __name__ == "__main__":
    main()

"""Pins the CPU oracle (oracle/rmq_oracle.c) before anything is checked against it.

Anchors, none of which re-types the oracle's own loop:
  * golden vectors (tests/golden): leftmost answers from numpy.argmin and the
    reference HRMQ library's own outputs (compiled from hrmq/);
  * closed forms: constant array -> l (leftmost) / r (HRMQ); ascending -> l;
    strictly descending -> r; [5,5] and [3,5,3];
  * exhaustive small inputs: every (l, r) against numpy;
  * cross-checks between independently structured oracles (brute force, the
    RTX_blocks decomposition for many block sizes, the O(1) fast oracle), on
    tie-heavy data where a wrong tie rule or a dropped edge block shows up.
"""
import numpy as np
import pytest

from conftest import golden_cases, load_golden
from oracle import oracle as O


@pytest.mark.parametrize("name", golden_cases())
def test_brute_matches_golden(name):
    g = load_golden(name)
    val, idx = O.rmq_brute(g["A"], g["Q"])
    np.testing.assert_array_equal(idx, g["idx"])
    np.testing.assert_array_equal(val.view(np.int32), g["val"].view(np.int32))  # exact bits


@pytest.mark.parametrize("name", golden_cases())
def test_hrmq_semantics_match_reference_hrmq(name):
    g = load_golden(name)
    if "hrmq_idx" not in g:
        pytest.skip("outside HRMQ's int32 input contract")
    ri = O.rmq_brute_hrmq(g["A"], g["Q"])
    np.testing.assert_array_equal(ri, g["hrmq_idx"])
    # HRMQ's value output equals the CPU BASE value (the minimum is unique as a value)
    np.testing.assert_array_equal(g["hrmq_val"].view(np.int32), g["val"].view(np.int32))


@pytest.mark.parametrize("name", golden_cases())
@pytest.mark.parametrize("bs", [1, 2, 3, 4, 5, 7, 16, 33, 1 << 15])
def test_blocks_and_fast_match_golden(name, bs):
    g = load_golden(name)
    v, i = O.rmq_blocks(g["A"], g["Q"], bs)
    np.testing.assert_array_equal(i, g["idx"])
    np.testing.assert_array_equal(v.view(np.int32), g["val"].view(np.int32))
    f = O.FastOracle(g["A"], block=bs)
    fv, fi = f.query(g["Q"])
    f.close()
    np.testing.assert_array_equal(fi, g["idx"])
    np.testing.assert_array_equal(fv.view(np.int32), g["val"].view(np.int32))


def test_closed_forms():
    A = np.full(100, 3.0, np.float32)
    for l, r in [(0, 99), (5, 5), (17, 60)]:
        assert O.cpu_min_idx(A, l, r) == l
        assert O.hrmq_idx(A, l, r) == r
    A = (np.uint32(0x3F800000) - 1 - np.arange(100, dtype=np.uint32)).view(np.float32)
    for l, r in [(0, 99), (5, 5), (17, 60)]:
        assert O.cpu_min_idx(A, l, r) == r
        assert O.hrmq_idx(A, l, r) == r
        assert O.cpu_min(A, l, r) == A[r]
    assert O.cpu_min_idx(np.array([5, 5], np.float32), 0, 1) == 0
    assert O.hrmq_idx(np.array([5, 5], np.float32), 0, 1) == 1
    assert O.cpu_min_idx(np.array([3, 5, 3], np.float32), 0, 2) == 0
    assert O.hrmq_idx(np.array([3, 5, 3], np.float32), 0, 2) == 2


def test_signed_zero_keeps_leftmost_bits():
    A = np.array([1.0, 0.0, -0.0, 0.0], np.float32)
    v, i = O.rmq_brute(A, np.array([[0, 3], [2, 3]], np.int32))
    assert list(i) == [1, 2]
    assert v.view(np.uint32)[0] == 0x00000000 and v.view(np.uint32)[1] == 0x80000000


@pytest.mark.parametrize("n", [1, 2, 3, 5, 31, 64, 97])
@pytest.mark.parametrize("vals", [2, 16, 1 << 20])
def test_exhaustive_small_vs_numpy(n, vals):
    rng = np.random.default_rng(n * 1000 + vals)
    A = rng.integers(0, vals, n).astype(np.float32)
    Q = np.array([(l, r) for l in range(n) for r in range(l, n)], np.int32)
    expect = np.array([l + np.argmin(A[l:r + 1]) for l, r in Q], np.int32)
    _, bi = O.rmq_brute(A, Q)
    np.testing.assert_array_equal(bi, expect)
    for bs in (1, 3, 4, 8, n):
        _, ki = O.rmq_blocks(A, Q, bs)
        np.testing.assert_array_equal(ki, expect)
        f = O.FastOracle(A, block=bs)
        np.testing.assert_array_equal(f.query(Q)[1], expect)
        f.close()


def test_fast_oracle_random_medium():
    A = O.gen_array(1 << 18, 5, "ties16")
    for lr in (-1, -2, -3, 1, 300):
        Q = O.gen_queries(20000, A.size, lr, 9)
        bv, bi = O.rmq_brute(A, Q)
        f = O.FastOracle(A, block=1024)
        fv, fi = f.query(Q)
        f.close()
        np.testing.assert_array_equal(fi, bi)
        np.testing.assert_array_equal(fv, bv)


def test_reference_hrmq_random_ties():
    if not O.HRMQ.available():
        pytest.skip("reference HRMQ not built here")
    A = O.gen_array(50000, 3, "ties16")
    Q = O.gen_queries(30000, A.size, -3, 4)
    h = O.HRMQ(A)
    hv, hi = h.query(Q, nt=4)
    h.close()
    np.testing.assert_array_equal(hi, O.rmq_brute_hrmq(A, Q))
    np.testing.assert_array_equal(hv, O.rmq_brute(A, Q)[0])


class TestGenerators:
    def test_uniform_array_is_exact_grid(self):
        A = O.gen_array(1 << 20, 1321)
        k = A.astype(np.float64) * 2.0 ** 24
        assert np.all(k == np.floor(k)) and A.min() >= 0 and A.max() < 1
        assert abs(A.mean() - 0.5) < 2e-3

    def test_kinds(self):
        t = O.gen_array(1 << 16, 2, "ties16")
        assert set(np.unique(t)) <= set(range(16)) and len(np.unique(t)) == 16
        d = O.gen_array(1000, 2, "desc")
        assert np.all(np.diff(d) < 0) and d.min() > 0
        assert np.all(O.gen_array(10, 2, "const") == 1.0)

    @pytest.mark.parametrize("lr", [-1, -2, -3, -4, -5, -6, -7, -8, 1, 256])
    def test_queries_valid_and_deterministic(self, lr):
        n = 10 ** 6
        Q = O.gen_queries(100000, n, lr, 77, nt=4)
        l, r = Q[:, 0], Q[:, 1]
        assert np.all(l >= 0) and np.all(l <= r) and np.all(r < n)
        np.testing.assert_array_equal(Q, O.gen_queries(100000, n, lr, 77, nt=1))  # thread-count independent
        ln = r - l + 1
        if lr > 0:
            assert np.all(ln == lr)
        if lr == -6:
            assert np.all(r == n - 1)

    def test_length_distributions(self):
        n = 10 ** 8
        ln = lambda lr: (lambda Q: Q[:, 1] - Q[:, 0] + 1)(O.gen_queries(200000, n, lr, 5))
        assert abs(np.median(ln(-2)) / 2 ** 14 - 1) < 0.02   # lognormal median 2^14
        assert abs(np.median(ln(-3)) / 2 ** 6 - 1) < 0.05    # lognormal median 2^6
        u = ln(-1)
        assert abs(u.mean() / (n / 4) - 1) < 0.02           # E[len] = E[(n-l)/2] = n/4
        assert abs(np.std(np.log(ln(-2))) - 1.0) < 0.02

"""Coarse row (RMQ_BUILD_MID): a superset test over coarse blocks of 2^s
blocks, from a row the query kernel keeps in LDS (entry i: the leftmost
minimum of coarse blocks i and i+1), before the block table in global memory
(k_query_edges_mid).  Bar: bit-exact against
CPU BASE brute force (leftmost minimum, value bits), on arrays whose ties,
+-0 and +inf runs make the leftmost rule matter, with ranges that equal,
straddle and end inside coarse blocks, including the partial last one.
"""
import os
import sys

import numpy as np
import pytest
import torch

import rtxrmq_amd as R
from oracle import oracle as O

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (mid_shift_of: the coarse block size, checked against the library in test_abi)

pytestmark = pytest.mark.gpu

MID = R.BUILD_EDGES | R.BUILD_MID


def _dev(x, gpu):
    return torch.from_numpy(np.ascontiguousarray(x)).to(gpu)


def assert_exact(val, idx, ev, ei, what=""):
    bad = np.nonzero((idx != ei) | (val.view(np.int32) != ev.view(np.int32)))[0]
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5]}: got {idx[bad[:5]]} want {ei[bad[:5]]}"


def tie_array(n, seed):
    """Values in {-0, +0, 1, 2} with runs of +inf and a few negatives."""
    rng = np.random.default_rng(seed)
    pool = np.array([0.0, -0.0, 1.0, 2.0], np.float32)
    A = pool[rng.integers(0, pool.size, n)]
    A[(np.arange(n) // 3001) % 7 == 3] = np.inf
    A[rng.integers(0, n, 20)] = np.float32(-1.5)
    return A


def coarse_queries(n, bs, q, seed, max_coarse=6):
    """Lengths log-uniform up to max_coarse coarse blocks; ranges that are
    exactly 1..4 coarse blocks; ranges that straddle a coarse boundary by one
    element on each side; ranges ending at the array's end (the partial last
    coarse block) and starting at 0."""
    nb = -(-n // bs)
    coarse = bs << bench.mid_shift_of(nb)
    rng = np.random.default_rng(seed)
    l = rng.integers(0, n, q)
    ln = np.exp2(rng.uniform(1, np.log2(max_coarse * coarse), q)).astype(np.int64)
    r = np.minimum(l + ln - 1, n - 1)
    parts = [np.stack([l, r], 1)]
    starts = np.arange(0, n, coarse)
    for c in (1, 2, 3, 4):
        ll = starts[rng.integers(0, starts.size, 2000)]
        parts.append(np.stack([ll, np.minimum(ll + c * coarse - 1, n - 1)], 1))
        parts.append(np.stack([np.maximum(ll - 1, 0), np.minimum(ll + c * coarse, n - 1)], 1))
    k = 2000
    ln2 = np.exp2(rng.uniform(1, np.log2(3 * coarse), k)).astype(np.int64)
    parts.append(np.stack([np.maximum(n - ln2, 0), np.full(k, n - 1)], 1))
    parts.append(np.stack([np.zeros(k, np.int64), np.minimum(ln2 - 1, n - 1)], 1))
    return np.concatenate(parts).astype(np.int32), coarse


def test_kernel_names(gpu):
    """Coarse rows are used where the block table is read from global memory:
    a table that fits in LDS keeps the LDS-table kernel."""
    A = _dev(np.random.default_rng(3).random(1 << 20).astype(np.float32), gpu)
    opts = R.query_opts(short_max=32, group_lanes=8, long_lanes=16)
    with R.RMQIndex(A, 1 << 15, flags=MID) as ix:
        assert ix.kernels(opts) == ["k_query_edges_lds<8, 3>"]
    with R.RMQIndex(A, 64, flags=MID) as ix:
        assert ix.kernels(opts) == ["k_query_edges_mid<8, false>"]
    big = _dev(np.random.default_rng(4).random(1 << 24).astype(np.float32), gpu)
    with R.RMQIndex(big, 8, flags=MID) as ix:
        assert ix.kernels(opts) == ["k_query_edges_mid<8, true>"]
    with pytest.raises(R.RMQError):
        R.RMQIndex(A, 64, flags=R.BUILD_MID)  # coarse rows need edge tables


@pytest.mark.parametrize("group_lanes", [4, 16, 64])
@pytest.mark.parametrize("bs", [16, 64, 100, 1024])
@pytest.mark.parametrize("kind", ["ties", "uniform"])
def test_matches_brute_force(kind, bs, group_lanes, gpu):
    """n = 2,000,003 (a partial last block and coarse block): bs=16 and 64
    give coarse blocks of 256 elements (shift 4, 2), bs=100 of 400 (shift 2),
    bs=1024 of one block (shift 0: the coarse test covers one or two blocks
    with one entry)."""
    n = 2_000_003
    A = tie_array(n, bs) if kind == "ties" else O.gen_array(n, 7 + bs, "uniform")
    Q, coarse = coarse_queries(n, bs, 1 << 18, bs + group_lanes)
    ev, ei = O.rmq_brute(A, Q)
    opts = R.query_opts(short_max=1 if group_lanes != 64 else 32, group_lanes=group_lanes, long_lanes=4)
    A_dev, Q_dev = _dev(A, gpu), _dev(Q, gpu)
    with R.RMQIndex(A_dev, bs, flags=MID) as ix:
        assert ix.kernels(opts) == [f"k_query_edges_mid<{group_lanes}, false>"], (bs, ix.kernels(opts))
        v, i = ix.query(Q_dev, opts=opts)
        torch.cuda.synchronize()
        assert_exact(v.cpu().numpy(), i.cpu().numpy(), ev, ei, f"{kind} bs={bs} G={group_lanes} coarse={coarse}")
    # value-only and index-only outputs take the same path
    with R.RMQIndex(A_dev, bs, flags=MID) as ix:
        v, _ = ix.query(Q_dev, want_index=False, opts=opts)
        _, i = ix.query(Q_dev, want_value=False, opts=opts)
        torch.cuda.synchronize()
        assert_exact(v.cpu().numpy(), i.cpu().numpy(), ev, ei, f"{kind} bs={bs} split outputs")


@pytest.mark.parametrize("bs", [2, 8])
def test_large_table_path(bs, gpu):
    """Block tables over 64 MB take the phased path after a failed coarse
    test (k_query_edges_mid<G, true>): n = 2^24 + 5, coarse blocks of 2^11
    (bs=2) or 2^9 (bs=8) blocks, 4096 elements either way."""
    n = (1 << 24) + 5
    A = tie_array(n, 50 + bs)
    Q, coarse = coarse_queries(n, bs, 1 << 19, 60 + bs, max_coarse=4)
    ev, ei = O.rmq_brute(A, Q)
    opts = R.query_opts(short_max=1, group_lanes=8, long_lanes=16)
    with R.RMQIndex(_dev(A, gpu), bs, flags=MID) as ix:
        assert ix.kernels(opts) == ["k_query_edges_mid<8, true>"], ix.kernels(opts)
        v, i = ix.query(_dev(Q, gpu), opts=opts)
        torch.cuda.synchronize()
        assert_exact(v.cpu().numpy(), i.cpu().numpy(), ev, ei, f"bs={bs} coarse={coarse}")


def test_headline_size_lognormal_medium(gpu):
    """n=1e8, q=2^26 lognormal 2^14 lengths at bs=1024 (coarse blocks of 16
    blocks): every answer against the fast oracle, and identical to the same
    handle built without the coarse row."""
    n, q, bs = 100_000_000, 1 << 26, 1024
    A_dev = R.gen_array(n, 1321, "uniform", device=gpu)
    Q = R.gen_queries(q, n, -2, 1319)
    Q_dev = _dev(Q, gpu)
    opts = R.query_opts(short_max=8, group_lanes=4, long_lanes=64)
    with R.RMQIndex(A_dev, bs, flags=MID) as ix:
        assert ix.kernels(opts) == ["k_query_edges_mid<4, false>"]
        v, i = ix.query(Q_dev, opts=opts)
    with R.RMQIndex(A_dev, bs, flags=R.BUILD_EDGES) as ix:
        v0, i0 = ix.query(Q_dev, opts=opts)
    torch.cuda.synchronize()
    assert torch.equal(i, i0) and torch.equal(v.view(torch.int32), v0.view(torch.int32))
    v, i = v.cpu().numpy(), i.cpu().numpy()
    del Q_dev, v0, i0
    A = A_dev.cpu().numpy()
    f = O.FastOracle(A)
    fv, fi = f.query(Q)
    f.close()
    assert_exact(v, i, fv, fi, "n=1e8 lognormal 2^14 with coarse rows")

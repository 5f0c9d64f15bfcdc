"""Parity of the HIP engine (algs 5/105 through the C ABI) with the CPU oracle.

Bar: bit-exact.  Values are compared as int32 bit patterns (so -0.0 vs +0.0
counts), indices exactly, leftmost tie rule (CPU BASE, src/cpu_methods.h:7-53).
Small cases go against brute force and the golden fixtures; the headline
sizes (n=1e8, q=2^26; n=2^27 stress arrays) go against the O(1) fast oracle
plus size-independent properties (l <= idx <= r, A[idx] == value).
"""
import os

import numpy as np
import pytest
import torch

import rtxrmq_amd as R
from conftest import golden_cases, load_golden
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev(x, gpu):
    return torch.from_numpy(np.ascontiguousarray(x)).to(gpu)


def run_engine(A_np, Q_np, bs, gpu, opts=None, A_dev=None, flags=0):
    A = A_dev if A_dev is not None else _dev(A_np, gpu)
    Q = _dev(Q_np.reshape(-1, 2), gpu)
    with R.RMQIndex(A, bs, flags=flags) as ix:
        v, i = ix.query(Q, opts=opts)
        torch.cuda.synchronize()
        return v.cpu().numpy(), i.cpu().numpy()


def assert_exact(val, idx, ev, ei, what=""):
    bad = np.nonzero((idx != ei) | (val.view(np.int32) != ev.view(np.int32)))[0]
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5]}: got {idx[bad[:5]]} want {ei[bad[:5]]}"


# ---------------------------------------------------------------------------
FLAGS = [0, R.BUILD_EDGES]


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("name", golden_cases())
@pytest.mark.parametrize("bs", [1, 3, 4, 16, 33, 1 << 15])
def test_golden(name, bs, flags, gpu):
    g = load_golden(name)
    v, i = run_engine(g["A"], g["Q"], bs, gpu, flags=flags)
    assert_exact(v, i, g["val"], g["idx"], f"{name} bs={bs} flags={flags}")


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("n", [1, 2, 5, 63, 64, 65, 130, 257, 1000])
@pytest.mark.parametrize("bs", [1, 2, 3, 5, 8, 31, 32, 33, 64, 100, 1000])
def test_exhaustive_small(n, bs, flags, gpu):
    rng = np.random.default_rng(n * 7919 + bs)
    A = rng.integers(0, 4, n).astype(np.float32)
    Q = np.array([(l, r) for l in range(n) for r in range(l, n)], np.int32)
    ev, ei = O.rmq_brute(A, Q)
    v, i = run_engine(A, Q, bs, gpu, flags=flags)
    assert_exact(v, i, ev, ei, f"n={n} bs={bs} flags={flags}")


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("kind", ["ties4", "asc", "desc", "inf_runs"])
@pytest.mark.parametrize("bs", [256, 777, 1000, 4096])
def test_long_scans_through_sub_blocks(kind, bs, flags, gpu):
    """Ranges inside one block and edge parts that must be scanned, read
    through the 32-element sub-block minima (blocks that are and are not
    multiples of 32, heavy ties, monotone data, runs of +inf)."""
    n = 20011
    rng = np.random.default_rng(bs + len(kind))
    if kind == "ties4":
        A = rng.integers(0, 4, n).astype(np.float32)
    elif kind == "asc":
        A = np.arange(n, dtype=np.float32)
    elif kind == "desc":
        A = (np.uint32(0x3F800000) - 1 - np.arange(n, dtype=np.uint32)).view(np.float32)
    else:
        A = rng.integers(0, 8, n).astype(np.float32)
        A[(np.arange(n) // 300) % 3 == 1] = np.inf
    l = rng.integers(0, n, 60000)
    ln = np.exp2(rng.uniform(5, 14, l.size)).astype(np.int64)
    r = np.minimum(l + ln - 1, n - 1)
    Q = np.stack([l, r], 1).astype(np.int32)
    ev, ei = O.rmq_brute(A, Q)
    v, i = run_engine(A, Q, bs, gpu, flags=flags)
    assert_exact(v, i, ev, ei, f"{kind} bs={bs} flags={flags}")


@pytest.mark.parametrize("bs,flags", [(8, 0), (8, R.BUILD_EDGES), (37, R.BUILD_EDGES), (1000, 0),
                                      (1000, R.BUILD_EDGES), (1 << 15, 0), (1 << 17, R.BUILD_EDGES)])
def test_signed_zeros_infinities_and_denormals(bs, flags, gpu):
    """-0.0 and +0.0 compare equal, so the leftmost of them is the answer and
    its sign bit must come back exactly; -inf, +inf, denormals and negative
    values mixed in heavy ties, through every query kernel (global sparse
    table, table in LDS, tile scans) and every range class."""
    n = 300_007
    rng = np.random.default_rng(bs + flags)
    pool = np.array([0.0, -0.0, np.inf, 1e-45, 1.4e-44, 2.0], np.float32)  # the minimum is mostly a zero tie
    A = pool[rng.integers(0, pool.size, n)]
    run = (np.arange(n) // 5000) % 4 == 1  # stretches of positive values only: a denormal minimum
    A[run] = np.where(A[run] == 0, np.float32(1e-45), A[run])
    for x, k in ((-1e-45, 30), (-3.5, 20), (-np.inf, 10)):  # rare negatives
        A[rng.integers(0, n, k)] = np.float32(x)
    Q = np.concatenate([O.gen_queries(30_000, n, lr, bs) for lr in (-1, -2, -3, 1, 33)])
    ev, ei = O.rmq_brute(A, Q)
    for opts in (None, R.query_opts(short_max=1, group_lanes=4, long_lanes=4)):
        v, i = run_engine(A, Q, bs, gpu, opts=opts, flags=flags)
        assert_exact(v, i, ev, ei, f"bs={bs} flags={flags} opts={opts is not None}")
    bits = ev.view(np.uint32)
    assert (bits == 0).any() and (bits == 0x80000000).any() and (bits == 1).any()  # +0, -0, 1e-45 all answered


@pytest.mark.parametrize("bs", [2, 3, 8, 16])
def test_large_table_path_with_ties_zeros_and_inf_runs(bs, gpu):
    """The phased path of block tables beyond the caches (k_query_edges<G,
    true>: superset test, the other side's blocks with the failing side's
    edge pair, then both pairs and the blocks between) on an array whose
    table is large enough to take it: values in {-0, +0, 1, 2, 3} with runs
    of +inf, n not a multiple of bs, lengths log-uniform in [2, 2^12] and
    ranges ending at the array's end, against brute force."""
    n = (1 << 24) + 5
    rng = np.random.default_rng(900 + bs)
    pool = np.array([0.0, -0.0, 1.0, 2.0, 3.0], np.float32)
    A = pool[rng.integers(0, pool.size, n)]
    A[(np.arange(n) // 700) % 5 == 2] = np.inf
    q = 1 << 20
    l = rng.integers(0, n, q)
    ln = np.exp2(rng.uniform(1, 12, q)).astype(np.int64)
    r = np.minimum(l + ln - 1, n - 1)
    r[: q // 64] = n - 1  # ranges that end at the array's end
    l[: q // 64] = np.maximum(0, n - ln[: q // 64])
    Q = np.stack([l, r], 1).astype(np.int32)
    ev, ei = O.rmq_brute(A, Q)
    opts = R.query_opts(short_max=1, group_lanes=8, long_lanes=16)
    with R.RMQIndex(_dev(A, gpu), bs, flags=R.BUILD_EDGES) as ix:
        assert ix.kernels(opts) == ["k_query_edges<8, true>"], ix.kernels(opts)
        v, i = ix.query(_dev(Q, gpu), opts=opts)
        torch.cuda.synchronize()
        assert_exact(v.cpu().numpy(), i.cpu().numpy(), ev, ei, f"bs={bs}")


@pytest.fixture(scope="module")
def mixed_case():
    A = O.gen_array(200003, 11, "ties16")
    Q = np.concatenate([O.gen_queries(20000, A.size, lr, 3) for lr in (-1, -2, -3, 1, 5)])
    ev, ei = O.rmq_brute(A, Q)
    return A, Q, ev, ei


@pytest.mark.parametrize("group_lanes", [4, 8, 16, 64])
@pytest.mark.parametrize("long_lanes", [4, 16, 64])
@pytest.mark.parametrize("short_max", [1, 7, 64, 1024, 1 << 30])
def test_tuning_knobs_do_not_change_answers(group_lanes, long_lanes, short_max, mixed_case, gpu):
    A, Q, ev, ei = mixed_case
    opts = R.query_opts(short_max=short_max, group_lanes=group_lanes, long_lanes=long_lanes)
    for bs, flags in ((37, 0), (1000, 0), (37, R.BUILD_EDGES), (4096, R.BUILD_EDGES)):
        v, i = run_engine(A, Q, bs, gpu, opts=opts, flags=flags)
        assert_exact(v, i, ev, ei, f"bs={bs} flags={flags} G={group_lanes} GL={long_lanes} short_max={short_max}")


@pytest.mark.parametrize("short_max", [1, 64])
@pytest.mark.parametrize("group_lanes", [4, 8, 16, 64])
def test_short_ranges_at_every_alignment_and_the_array_end(group_lanes, short_max, gpu):
    """Every range of at most 40 elements of arrays whose length is and is not
    a multiple of 4: short ranges start and end at each position of a 16-byte
    chunk, and the last ones end in the array's partial final chunk.  With
    edge tables and short_max 1, only ranges inside one 32-element sub-block
    (one block when bs <= 32) are scanned; the rest take the O(1) path."""
    opts = R.query_opts(short_max=short_max, group_lanes=group_lanes, long_lanes=16)
    for n in (1, 3, 6, 61, 62, 63, 64, 1001, 1002, 1003, 4096):
        A = np.random.default_rng(n).integers(0, 5, n).astype(np.float32)
        A[::7] = np.inf
        Q = np.array([(l, r) for l in range(n) for r in range(l, min(n, l + 40))], np.int32)
        ev, ei = O.rmq_brute(A, Q)
        for bs, flags in ((8, 0), (8, R.BUILD_EDGES), (64, R.BUILD_EDGES), (100, R.BUILD_EDGES), (1000, 0)):
            v, i = run_engine(A, Q, bs, gpu, opts=opts, flags=flags)
            assert_exact(v, i, ev, ei, f"n={n} bs={bs} flags={flags} G={group_lanes} short_max={short_max}")


@pytest.mark.parametrize("long_lanes", [4, 8, 16, 64])
def test_long_in_block_scans_at_every_sub_block_alignment(long_lanes, gpu):
    """The long-range path without edge tables scans a range inside one block
    as a partial 32-element sub-block, the whole sub-blocks between (their
    minima or two sparse-table entries) and a partial sub-block, with every
    load issued before the first combine.  Block sizes that are not
    multiples of 4 (1001, 4099) start their sub-blocks at every offset of a
    16-byte chunk, so the partials' scalar heads and tails take every length;
    ranges start and end at every position of their sub-blocks, span 6 to
    128 sub-blocks, and some end at the array's end (bs=1001); ties and +inf runs make the
    leftmost rule matter.  Against brute force."""
    rng = np.random.default_rng(long_lanes)
    opts = R.query_opts(short_max=1, group_lanes=16, long_lanes=long_lanes)
    for bs, n in ((1001, 40_037), (4099, 81_983), (333, 20_011)):
        A = rng.integers(0, 6, n).astype(np.float32)
        A[(np.arange(n) // 97) % 5 == 1] = np.inf
        qs = []
        nb = -(-n // bs)
        for b in sorted(set(range(0, nb, 3)) | {nb - 1}):  # the partial last block too
            base = b * bs
            end = min(base + bs, n) - 1
            for _ in range(40):
                l = base + int(rng.integers(0, 64))
                r = end - int(rng.integers(0, 64))
                if r - l >= 5 * 32:
                    qs.append((l, r))
            for off in range(0, 36):  # every head offset against a fixed far end, and the reverse
                if base + off < end - 200:
                    qs.append((base + off, end - 37))
                    qs.append((base + 37, end - off))
        Q = np.array(qs, np.int32)
        ev, ei = O.rmq_brute(A, Q)
        v, i = run_engine(A, Q, bs, gpu, opts=opts)
        assert_exact(v, i, ev, ei, f"bs={bs} GL={long_lanes}")


def test_query_kernel_names(gpu):
    """rmq_query_kernels names what rmq_query launches: with edge tables the
    LDS-table kernel when the block sparse table fits in LDS (twice per CU),
    else the global-table one (retrying failed superset tests when the table
    exceeds 64 MB); without, the short and long tile kernels."""
    A = _dev(np.random.default_rng(3).random(1 << 20).astype(np.float32), gpu)
    opts = R.query_opts(short_max=32, group_lanes=8, long_lanes=16)
    with R.RMQIndex(A, 1 << 15, flags=R.BUILD_EDGES) as ix:      # 32 blocks: a few KB
        assert ix.kernels(opts) == ["k_query_edges_lds<8, 3>"]
    with R.RMQIndex(A, 64, flags=R.BUILD_EDGES) as ix:           # 16384 blocks x 15 levels: about 2 MB
        assert ix.kernels(opts) == ["k_query_edges<8, false>"]
    big = _dev(np.random.default_rng(4).random(1 << 24).astype(np.float32), gpu)
    with R.RMQIndex(big, 8, flags=R.BUILD_EDGES) as ix:          # 2^21 blocks x 22 levels: 369 MB
        assert ix.kernels(opts) == ["k_query_edges<8, true>"]
    with R.RMQIndex(A, 1 << 15) as ix:
        assert ix.kernels(opts) == ["k_query_tiles<8, false>", "k_query_tiles<16, true>"]
        assert ix.kernels() == ["k_query_tiles<8, false>", "k_query_tiles<4, true>"]  # the defaults
        o = ix.default_opts()
        assert (o.short_max, o.group_lanes, o.long_wgs, o.long_lanes) == (128, 8, 0, 4)


def test_autotuner_pick_and_cache(tmp_path, gpu):
    n = 1 << 22
    A_dev = R.gen_array(n, 5, device=gpu)
    Q = R.gen_queries(1 << 16, n, -2, 5)
    Q_dev = _dev(Q, gpu)
    cache = str(tmp_path / "tune.txt")
    r1 = R.tune(A_dev, Q_dev, bs_list=[256, 1024, 4096], cache_path=cache)
    assert r1["bs"] in (256, 1024, 4096) and r1["n_tried"] > 3 and not r1["from_cache"]
    assert r1["opts"].group_lanes in (4, 8, 16, 64) and r1["opts"].long_lanes in (4, 8, 16, 64)
    r2 = R.tune(A_dev, Q_dev, bs_list=[256, 1024, 4096], cache_path=cache)
    assert r2["from_cache"] and r2["bs"] == r1["bs"] and r2["key"] == r1["key"]
    assert r1["key"].startswith("t3/n2^22.0/") and r1["key"].endswith("/pos0.50/tie0/out=vi/bs=256,1024,4096"), \
        r1["key"]
    # another candidate list asks another question: a key of its own, timed afresh
    r3 = R.tune(A_dev, Q_dev, bs_list=[1024], cache_path=cache)
    assert not r3["from_cache"] and r3["bs"] == 1024 and r3["key"].endswith("/out=vi/bs=1024"), r3["key"]
    # a list longer than 24 characters is keyed by the FNV-1a hash of its text
    long_list = [1024, 2048, 4096, 8192, 16384]
    h = 2166136261
    for ch in ",".join(map(str, long_list)).encode():
        h = ((h ^ ch) * 16777619) & 0xFFFFFFFF
    r4 = R.tune(A_dev, Q_dev, bs_list=long_list, cache_path=cache)
    assert not r4["from_cache"] and r4["key"].endswith(f"/out=vi/bs#{h:08x}"), r4["key"]
    assert R.tune(A_dev, Q_dev, bs_list=[256, 1024, 4096], cache_path=cache)["bs"] == r1["bs"]
    # a restricted set of build flags is a question of its own too (rmq_tune_ex)
    r5 = R.tune(A_dev, Q_dev, bs_list=[16, 64], cache_path=cache, flags_allowed=R.BUILD_EDGES)
    assert not r5["from_cache"] and r5["key"].endswith("/bs=16,64/f1") and not r5["flags"] & R.BUILD_MID, r5
    r6 = R.tune(A_dev, Q_dev, bs_list=[16, 64], cache_path=cache, flags_allowed=0)
    assert r6["key"].endswith("/bs=16,64/f0") and r6["flags"] == 0, r6
    with pytest.raises(R.RMQError):
        R.tune(A_dev, Q_dev, bs_list=[16, 64], flags_allowed=R.BUILD_MID)  # coarse rows need edge tables
    with R.RMQIndex(A_dev, r1["bs"]) as ix:
        v, i = ix.query(Q_dev, opts=r1["opts"])
        torch.cuda.synchronize()
    ev, ei = O.rmq_brute(A_dev.cpu().numpy(), Q)
    assert_exact(v.cpu().numpy(), i.cpu().numpy(), ev, ei, "tuned config")


def test_tuner_key_separates_value_kinds_and_outputs(tmp_path, gpu):
    """Arrays that differ in tie density or trend, and tunings that time
    different outputs, get different cache keys: a pick made for one never
    serves the other."""
    n = 1 << 20
    Q_dev = _dev(R.gen_queries(1 << 14, n, -2, 9), gpu)
    cache = str(tmp_path / "tune.txt")
    keys = {}
    for kind in ("uniform", "ties16", "desc"):
        A_dev = R.gen_array(n, 9, kind, device=gpu)
        keys[kind] = R.tune(A_dev, Q_dev, bs_list=[1024], cache_path=cache)["key"]
    assert keys["uniform"].endswith("/pos0.50/tie0/out=vi/bs=1024"), keys
    assert "/tie2/" in keys["ties16"] and "/pos0.25/" in keys["ties16"], keys
    assert keys["desc"].endswith("/pos1.00/tie0/out=vi/bs=1024"), keys
    A_dev = R.gen_array(n, 9, "uniform", device=gpu)
    r_i = R.tune(A_dev, Q_dev, bs_list=[1024], cache_path=cache, want_value=False)
    r_v = R.tune(A_dev, Q_dev, bs_list=[1024], cache_path=cache, want_index=False)
    assert r_i["key"].endswith("/out=i/bs=1024") and r_v["key"].endswith("/out=v/bs=1024")
    assert not r_i["from_cache"] and not r_v["from_cache"]


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("bs", [33, 1024, 1 << 15, 1 << 18, 1 << 20])
def test_handle_bytes_is_what_the_build_allocates(bs, flags, gpu):
    """rmq_handle_bytes (the server's admission figure) equals the bytes the
    built handle reports, for plain and edge-table handles."""
    n = 3_000_017
    A_dev = R.gen_array(n, 3, device=gpu)
    with R.RMQIndex(A_dev, bs, flags=flags) as ix:
        assert R.handle_bytes(n, bs, flags) == ix.info()["bytes"]


@pytest.mark.parametrize("n", [1, 5, 7])
def test_autotuner_on_arrays_smaller_than_every_block_size(n, gpu):
    A = O.gen_array(n, 3, "ties16")
    A_dev = _dev(A, gpu)
    Q = np.array([(l, r) for l in range(n) for r in range(l, n)], np.int32)
    pick = R.tune(A_dev, _dev(Q, gpu))
    assert pick["bs"] == n and not pick["from_cache"]
    with R.RMQIndex(A_dev, pick["bs"], flags=pick["flags"]) as ix:
        v, i = ix.query(_dev(Q, gpu), opts=pick["opts"])
        torch.cuda.synchronize()
    ev, ei = O.rmq_brute(A, Q)
    assert_exact(v.cpu().numpy(), i.cpu().numpy(), ev, ei, f"n={n}")


def test_plain_small_n2e20(gpu):
    """n=2^20 U[0,1), q=2^20 uniform lengths, bs=2^15: engine vs fast oracle
    (all queries), vs CPU BASE brute force and GPU BASE (subsample), vs the
    reference HRMQ values when it is built."""
    n, q, bs = 1 << 20, 1 << 20, 1 << 15
    A_dev = R.gen_array(n, 1321, "uniform", device=gpu)
    A = A_dev.cpu().numpy()
    np.testing.assert_array_equal(A.view(np.int32), O.gen_array(n, 1321).view(np.int32))
    Q = R.gen_queries(q, n, -1, 1321)
    v, i = run_engine(A, Q, bs, gpu, A_dev=A_dev)
    f = O.FastOracle(A, nt=0)
    fv, fi = f.query(Q)
    f.close()
    assert_exact(v, i, fv, fi, "n=2^20 vs fast oracle")
    sub = np.random.default_rng(0).choice(q, 3000, replace=False)
    bv, bi = O.rmq_brute(A, Q[sub])
    assert_exact(v[sub], i[sub], bv, bi, "n=2^20 vs CPU BASE")
    gv, gi = R.gpu_base(A_dev, _dev(Q[sub], gpu))
    assert_exact(v[sub], i[sub], gv.cpu().numpy(), gi.cpu().numpy(), "n=2^20 vs GPU BASE")
    if O.HRMQ.available():
        h = O.HRMQ(A)
        hv, _ = h.query(Q)
        h.close()
        np.testing.assert_array_equal(v.view(np.int32), hv.view(np.int32))


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("lr", [-1, -2, -3])
def test_headline_full_size(lr, flags, gpu):
    """n=1e8, q=2^26, bs=2^15, each length distribution: every answer checked."""
    n, q, bs = 100_000_000, 1 << 26, 1 << 15
    A_dev = R.gen_array(n, 1321, "uniform", device=gpu)
    Q = R.gen_queries(q, n, lr, 1321 + lr)
    Q_dev = _dev(Q, gpu)
    with R.RMQIndex(A_dev, bs, flags=flags) as ix:
        v, i = ix.query(Q_dev)
        torch.cuda.synchronize()
        v, i = v.cpu().numpy(), i.cpu().numpy()
    A = A_dev.cpu().numpy()
    del Q_dev
    assert np.all(i >= Q[:, 0]) and np.all(i <= Q[:, 1])
    np.testing.assert_array_equal(A[i].view(np.int32), v.view(np.int32))
    f = O.FastOracle(A)
    fv, fi = f.query(Q)
    f.close()
    assert_exact(v, i, fv, fi, f"headline lr={lr}")


# The kernels the committed MI355X picks run for the three headline batches.
HEADLINE_KERNELS = {-1: ["k_query_edges_lds<16, 3>"], -2: ["k_query_edges_mid<8, false>"],
                    -3: ["k_query_edges_mid<16, true>"]}


@pytest.mark.parametrize("lr", [-1, -2, -3])
def test_committed_tuner_picks_full_size(lr, tmp_path, gpu):
    """The configurations the headline runs: n=1e8, each length distribution
    at the block size, edge tables and thresholds of the committed MI355X
    tuner cache (the LDS-table kernel for uniform lengths, the global-table
    kernel with coarse rows in LDS and its line-aware superset test for the
    lognormal ones), every answer checked against the O(1) oracle."""
    import shutil
    n, q = 100_000_000, 1 << 26
    A_dev = R.gen_array(n, 1321, "uniform", device=gpu)
    Q = R.gen_queries(q, n, lr, 1321 + lr)
    Q_dev = _dev(Q, gpu)
    cache = tmp_path / "picks.txt"  # a copy: a missing key is tuned here, not written to the committed file
    shutil.copy(os.path.join(os.path.dirname(R.__file__), "tune_cache_mi355x.txt"), cache)
    pick = R.tune(A_dev, Q_dev[: 1 << 22], cache_path=str(cache))
    assert pick["from_cache"], f"lr={lr}: key {pick['key']} missed the committed cache"
    with R.RMQIndex(A_dev, pick["bs"], flags=pick["flags"]) as ix:
        kern = ix.kernels(pick["opts"])
    assert kern == HEADLINE_KERNELS[lr], f"lr={lr}: {kern}"
    with R.RMQIndex(A_dev, pick["bs"], flags=pick["flags"]) as ix:
        v, i = ix.query(Q_dev, opts=pick["opts"])
        torch.cuda.synchronize()
        v, i = v.cpu().numpy(), i.cpu().numpy()
    del Q_dev
    A = A_dev.cpu().numpy()
    f = O.FastOracle(A)
    fv, fi = f.query(Q)
    f.close()
    assert_exact(v, i, fv, fi, f"lr={lr} pick={pick['key']} bs={pick['bs']} kernels={kern}")


# The kernels the committed picks restricted to bs = 2^15 run (edge tables for all three; coarse rows,
# here one block each, for the uniform and lognormal 2^14 batches).
LITERAL_BS_KERNELS = {-1: ["k_query_edges_mid<4, false>"], -2: ["k_query_edges_mid<16, false>"],
                      -3: ["k_query_edges<8, false>"]}


@pytest.mark.parametrize("lr", [-1, -2, -3])
def test_committed_picks_at_bs_2_15_full_size(lr, tmp_path, gpu):
    """The headline batches at the stated block size 2^15 with the structure
    the tuner restricted to it picks (edge tables and thresholds, from the
    committed cache), every answer checked against the O(1) oracle."""
    import shutil
    n, q = 100_000_000, 1 << 26
    A_dev = R.gen_array(n, 1321, "uniform", device=gpu)
    Q = R.gen_queries(q, n, lr, 1321 + lr)
    Q_dev = _dev(Q, gpu)
    cache = tmp_path / "picks.txt"
    shutil.copy(os.path.join(os.path.dirname(R.__file__), "tune_cache_mi355x.txt"), cache)
    pick = R.tune(A_dev, Q_dev[: 1 << 22], bs_list=[1 << 15], cache_path=str(cache))
    assert pick["from_cache"] and pick["bs"] == 1 << 15, f"lr={lr}: key {pick['key']} missed the committed cache"
    with R.RMQIndex(A_dev, pick["bs"], flags=pick["flags"]) as ix:
        kern = ix.kernels(pick["opts"])
        assert kern == LITERAL_BS_KERNELS[lr], f"lr={lr}: {kern}"
        v, i = ix.query(Q_dev, opts=pick["opts"])
        torch.cuda.synchronize()
        v, i = v.cpu().numpy(), i.cpu().numpy()
    del Q_dev
    A = A_dev.cpu().numpy()
    f = O.FastOracle(A)
    fv, fi = f.query(Q)
    f.close()
    assert_exact(v, i, fv, fi, f"lr={lr} bs=2^15 pick={pick['key']} kernels={kern}")


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("kind", ["ties16", "desc"])
@pytest.mark.parametrize("bs", [1 << 5, 1 << 10, 1 << 12, 1 << 15, 1 << 18])
def test_stress_n2_27(kind, bs, flags, gpu):
    """n=2^27 heavy-tie and strictly descending arrays, fixed lengths
    {1, 2^8, 2^16, n-l}, block sizes across the sweep (2^20 queries per length)."""
    n = 1 << 27
    A_dev = R.gen_array(n, 99, kind, device=gpu)
    A = A_dev.cpu().numpy()
    f = O.FastOracle(A)
    with R.RMQIndex(A_dev, bs, flags=flags) as ix:
        for lr in (1, 1 << 8, 1 << 16, -6):
            Q = R.gen_queries(1 << 20, n, lr, 7)
            v, i = ix.query(_dev(Q, gpu))
            torch.cuda.synchronize()
            fv, fi = f.query(Q)
            assert_exact(v.cpu().numpy(), i.cpu().numpy(), fv, fi, f"{kind} bs={bs} flags={flags} lr={lr}")
    f.close()


def test_value_only_and_index_only(gpu):
    A = O.gen_array(100000, 5, "ties16")
    Q = O.gen_queries(50000, A.size, -2, 5)
    ev, ei = O.rmq_brute(A, Q)
    with R.RMQIndex(_dev(A, gpu), 4096) as ix:
        v, none = ix.query(_dev(Q, gpu), want_value=True, want_index=False)
        none2, i = ix.query(_dev(Q, gpu), want_value=False, want_index=True)
        torch.cuda.synchronize()
    assert none is None and none2 is None
    np.testing.assert_array_equal(v.cpu().numpy().view(np.int32), ev.view(np.int32))
    np.testing.assert_array_equal(i.cpu().numpy(), ei)


def test_invalid_queries_are_flagged_not_faulted(gpu):
    A = O.gen_array(1000, 5)
    Q = np.array([[0, 999], [5, 4], [-1, 3], [10, 1000], [3, 3], [999, 999]], np.int32)
    v, i = run_engine(A, Q, 64, gpu)
    assert list(i[[1, 2, 3]]) == [-1, -1, -1] and np.all(np.isnan(v[[1, 2, 3]]))
    ev, ei = O.rmq_brute(A, Q[[0, 4, 5]])
    assert_exact(v[[0, 4, 5]], i[[0, 4, 5]], ev, ei)
    assert R.check_queries(_dev(Q, gpu), 1000) == 3


@pytest.mark.parametrize("outputs", ["vi", "v", "i"])
@pytest.mark.parametrize("bs,flags,family", [(1024, R.BUILD_EDGES, "k_query_edges_lds<"),
                                             (8, R.BUILD_EDGES, "k_query_edges<"), (4096, 0, "k_query_tiles<")])
def test_invalid_and_partial_tiles_every_kernel_and_output(bs, flags, family, outputs, gpu):
    """Each query kernel (LDS-table, global-table, tile pair) with each output
    kind: batches of 1, 63, 64, 65, 1000 and 4097 queries (full and partial
    64-query tiles, fewer tiles than waves), a tenth of them invalid (l > r,
    l < 0, r >= n), mixed with short and long ranges; valid answers equal
    brute force, invalid ones are (NaN, -1), and nothing past q is written."""
    n = 100_003
    A = O.gen_array(n, 17, "ties16")
    A_dev = _dev(A, gpu)
    rng = np.random.default_rng(bs + len(outputs))
    with R.RMQIndex(A_dev, bs, flags=flags) as ix:
        assert ix.kernels()[0].startswith(family), ix.kernels()
        for q in (1, 63, 64, 65, 1000, 4097):
            l = rng.integers(0, n, q)
            ln = np.where(rng.random(q) < 0.5, rng.integers(1, 40, q), rng.integers(1, 5000, q))
            Q = np.stack([l, np.minimum(n - 1, l + ln - 1)], 1).astype(np.int32)
            bad = rng.random(q) < 0.1
            kinds = rng.integers(0, 3, q)
            Q[bad & (kinds == 0)] = [[7, 3]]
            Q[bad & (kinds == 1)] = [[-2, 5]]
            Q[bad & (kinds == 2)] = [[n - 5, n + 2]]
            ev, ei = O.rmq_brute(A, np.where(bad[:, None], 0, Q))
            v = torch.full((q + 8,), 123.0, dtype=torch.float32, device=gpu)
            i = torch.full((q + 8,), 77, dtype=torch.int32, device=gpu)
            ix.query(_dev(Q, gpu), want_value="v" in outputs, want_index="i" in outputs,
                     out_val=v if "v" in outputs else None, out_idx=i if "i" in outputs else None)
            torch.cuda.synchronize()
            vh, ih = v.cpu().numpy(), i.cpu().numpy()
            if "i" in outputs:
                assert np.array_equal(ih[:q][~bad], ei[~bad]) and np.all(ih[:q][bad] == -1), f"q={q}"
                assert np.all(ih[q:] == 77), f"q={q}: index written past q"
            if "v" in outputs:
                assert np.array_equal(vh[:q][~bad].view(np.int32), ev[~bad].view(np.int32)), f"q={q}"
                assert np.all(np.isnan(vh[:q][bad])) and np.all(vh[q:] == 123.0), f"q={q}"


@pytest.mark.parametrize("flags", FLAGS)
def test_empty_and_single(flags, gpu):
    A = _dev(np.array([2.5], np.float32), gpu)
    with R.RMQIndex(A, 7, flags=flags) as ix:
        v, i = ix.query(torch.empty((0, 2), dtype=torch.int32, device=gpu))
        assert v.numel() == 0 and i.numel() == 0
        v, i = ix.query(_dev(np.array([[0, 0]], np.int32), gpu))
        assert i.item() == 0 and v.item() == 2.5
        info = ix.info()
        assert info["nb"] == 1 and info["levels"] == 1


def test_device_generator_matches_host(gpu):
    for kind in ("uniform", "ties16", "desc", "const"):
        d = R.gen_array(1 << 20, 1234, kind, device=gpu).cpu().numpy()
        np.testing.assert_array_equal(d.view(np.int32), R.gen_array_host(1 << 20, 1234, kind).view(np.int32))


@pytest.mark.parametrize("bs,flags", [(1 << 12, 0), (1 << 15, R.BUILD_EDGES), (64, R.BUILD_EDGES)])
def test_concurrent_streams_same_handle(bs, flags, gpu):
    """One read-only handle queried from three streams at once (tile kernels,
    LDS-table and global-table edge kernels)."""
    A = O.gen_array(1 << 20, 8)
    Qs = [O.gen_queries(1 << 16, A.size, lr, 8) for lr in (-1, -2, -3)]
    A_dev = _dev(A, gpu)
    streams = [torch.cuda.Stream() for _ in Qs]
    with R.RMQIndex(A_dev, bs, flags=flags) as ix:
        torch.cuda.synchronize()
        outs = []
        for s, Q in zip(streams, Qs):
            with torch.cuda.stream(s):
                Qd = _dev(Q, gpu)
                outs.append((Qd,) + ix.query(Qd, stream=s))
        torch.cuda.synchronize()
        for (Qd, v, i), Q in zip(outs, Qs):
            ev, ei = O.rmq_brute(A, Q)
            assert_exact(v.cpu().numpy(), i.cpu().numpy(), ev, ei)


@pytest.mark.parametrize("bs,flags", [(1 << 17, R.BUILD_EDGES), (64, R.BUILD_EDGES), (1 << 12, 0)])
def test_query_replays_from_a_hip_graph(bs, flags, gpu):
    """rmq_query allocates nothing and records no memset, so a caller can
    capture it in a HIP graph and replay it.  Each replay, with new queries
    copied into the captured input and the outputs cleared by the program's
    own eager work in between, matches the brute force -- for the three query
    paths: block table in LDS, in global memory, and no edge tables."""
    n, qn = 1 << 20, 1 << 15
    A = O.gen_array(n, 21, "ties16")
    A_dev = _dev(A, gpu)
    Qbuf = torch.empty((qn, 2), dtype=torch.int32, device=gpu)
    v = torch.empty(qn, dtype=torch.float32, device=gpu)
    i = torch.empty(qn, dtype=torch.int32, device=gpu)
    with R.RMQIndex(A_dev, bs, flags=flags) as ix:
        Qbuf.copy_(_dev(O.gen_queries(qn, n, -1, 1), gpu))
        ix.query(Qbuf, out_val=v, out_idx=i)  # loads the kernels before capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            ix.query(Qbuf, out_val=v, out_idx=i)
        for k, lr in enumerate((-1, -2, -3, 5, -1)):
            Q = O.gen_queries(qn, n, lr, 100 + k)
            Qbuf.copy_(_dev(Q, gpu))
            v.fill_(0.0)
            i.fill_(-7)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            ev, ei = O.rmq_brute(A, Q)
            assert_exact(v.cpu().numpy(), i.cpu().numpy(), ev, ei, f"replay {k} bs={bs} flags={flags}")
        kern = ix.kernels()
    assert kern[0].startswith({(1 << 17): "k_query_edges_lds", 64: "k_query_edges<", 1 << 12: "k_query_tiles"}[bs])


def test_rtx_rmq_mirror(gpu):
    n, q, bs = 1 << 16, 1 << 14, 1 << 10
    A_dev = R.gen_array(n, 3, device=gpu)
    Q = R.gen_queries(q, n, -2, 3)
    A = A_dev.cpu().numpy()
    ev, ei = O.rmq_brute(A, Q)
    r5 = R.rtx_rmq(R.ALG_GPU_RTX_BLOCKS, n, bs, q, A_dev, _dev(Q, gpu), reps=3)
    r105 = R.rtx_rmq(R.ALG_GPU_RTX_BLOCKS_IDX, n, bs, q, A_dev, _dev(Q, gpu), reps=3)
    np.testing.assert_array_equal(r5.output.view(np.int32), ev.view(np.int32))
    np.testing.assert_array_equal(r105.output, ei)
    assert r5.rmq_per_s > 0 and r105.reps == 3

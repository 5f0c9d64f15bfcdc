"""Batched mode: many arrays answered by one persistent launch (rmq_plan_*).

Parity bar as for the single-array path: bit-exact against the CPU oracle
(leftmost index, exact value bits), per array.  Graph replays are checked
against the eager launch, with the program's own eager work and
synchronisations between replays.
"""
import numpy as np
import pytest
import torch

import rtxrmq_amd as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev(x, gpu):
    return torch.from_numpy(np.ascontiguousarray(x)).to(gpu)


def make_batch(gpu, m, lo_log2, hi_log2, seed):
    rng = np.random.default_rng(seed)
    ns = np.exp2(rng.uniform(lo_log2, hi_log2, m)).astype(np.int64)
    arrays, queries = [], []
    for i, n in enumerate(ns):
        A = O.gen_array(int(n), seed * 1000 + i, "ties16" if i % 3 == 0 else "uniform")
        q = int(n)
        Q = np.concatenate([O.gen_queries(q // 3 + 1, int(n), lr, seed + i) for lr in (-1, -7, -8)])
        arrays.append(A)
        queries.append(Q)
    return arrays, queries


def expected(arrays, queries):
    return [O.rmq_brute(A, Q) for A, Q in zip(arrays, queries)]


def test_device_query_generator_matches_host(gpu):
    for lr in (-1, -6, 5):
        d = R.gen_queries_device(1 << 20, 10 ** 7, lr, 3, device=gpu).cpu().numpy()
        np.testing.assert_array_equal(d, R.gen_queries(1 << 20, 10 ** 7, lr, 3))
    for lr in (-2, -3, -7, -8):
        d = R.gen_queries_device(1 << 20, 10 ** 7, lr, 3, device=gpu).cpu().numpy()
        h = R.gen_queries(1 << 20, 10 ** 7, lr, 3)
        assert np.all(d[:, 0] >= 0) and np.all(d[:, 0] <= d[:, 1]) and np.all(d[:, 1] < 10 ** 7)
        assert np.mean(np.any(d != h, axis=1)) < 1e-4  # libm last-ulp differences only


@pytest.mark.parametrize("opts", [None, R.query_opts(short_max=16, group_lanes=4, long_lanes=8),
                                  R.query_opts(short_max=1000, group_lanes=16, long_lanes=64)])
@pytest.mark.parametrize("item_queries", [0, 97])
@pytest.mark.parametrize("flags", [0, R.BUILD_EDGES])
def test_plan_matches_oracle(opts, item_queries, flags, gpu):
    arrays, queries = make_batch(gpu, 40, 4, 14, 11)
    exp = expected(arrays, queries)
    bss = [1, 3, 64, 256, 1000]
    A_dev = [_dev(A, gpu) for A in arrays]
    Q_dev = [_dev(Q, gpu) for Q in queries]
    idx = [torch.full((Q.shape[0],), -7, dtype=torch.int32, device=gpu) for Q in queries]
    val = [torch.zeros(Q.shape[0], dtype=torch.float32, device=gpu) for Q in queries]
    indexes = [R.RMQIndex(A, bss[i % len(bss)], flags=flags if i % 2 else 0) for i, A in enumerate(A_dev)]
    with R.Plan(indexes, Q_dev, out_idx=idx, out_val=val, opts=opts, item_queries=item_queries) as plan:
        info = plan.info()
        assert info["arrays"] == 40 and info["queries"] == sum(Q.shape[0] for Q in queries)
        plan.launch()
        torch.cuda.synchronize()
    for k, ((ev, ei), i_t, v_t) in enumerate(zip(exp, idx, val)):
        np.testing.assert_array_equal(i_t.cpu().numpy(), ei, err_msg=f"array {k} index")
        np.testing.assert_array_equal(v_t.cpu().numpy().view(np.int32), ev.view(np.int32), err_msg=f"array {k} value")
    for ix in indexes:
        ix.close()


@pytest.mark.parametrize("flags", [0, R.BUILD_EDGES])
def test_plan_equals_per_array_loop_and_graph_replays(flags, gpu):
    arrays, queries = make_batch(gpu, 64, 6, 16, 23)
    A_dev = [_dev(A, gpu) for A in arrays]
    Q_dev = [_dev(Q, gpu) for Q in queries]
    indexes = [R.RMQIndex(A, 256, flags=flags) for A in A_dev]
    loop = []
    for ix, Q in zip(indexes, Q_dev):
        loop.append(ix.query(Q, want_value=False, want_index=True)[1])
    torch.cuda.synchronize()
    loop = [t.cpu().numpy() for t in loop]
    exp = expected(arrays, queries)
    for k, (l_, (ev, ei)) in enumerate(zip(loop, exp)):
        np.testing.assert_array_equal(l_, ei, err_msg=f"loop array {k}")
    idx = [torch.empty(Q.shape[0], dtype=torch.int32, device=gpu) for Q in queries]
    other = R.RMQIndex(A_dev[0], 7)
    with R.Plan(indexes, Q_dev, out_idx=idx) as plan:
        plan.launch()
        torch.cuda.synchronize()
        eager = [t.cpu().numpy() for t in idx]
        for k, e in enumerate(eager):
            np.testing.assert_array_equal(e, loop[k], err_msg=f"eager plan array {k}")
        for replay in range(3):
            for t in idx:
                t.fill_(-3)
            # unrelated eager work between replays
            other.query(Q_dev[0])
            torch.cuda.synchronize()
            plan.graph_launch()
            torch.cuda.synchronize()
            for k, t in enumerate(idx):
                np.testing.assert_array_equal(t.cpu().numpy(), eager[k], err_msg=f"replay {replay} array {k}")
    other.close()
    for ix in indexes:
        ix.close()


def test_empty_plan_and_empty_arrays(gpu):
    A = _dev(O.gen_array(100, 1), gpu)
    ix = R.RMQIndex(A, 8)
    Q0 = torch.empty((0, 2), dtype=torch.int32, device=gpu)
    Q1 = _dev(np.array([[0, 99], [5, 5]], np.int32), gpu)
    out0 = torch.empty(0, dtype=torch.int32, device=gpu)
    out1 = torch.empty(2, dtype=torch.int32, device=gpu)
    with R.Plan([ix, ix], [Q0, Q1], out_idx=[out0, out1]) as plan:
        assert plan.info()["items"] == 1
        plan.launch()
        plan.graph_launch()
        torch.cuda.synchronize()
    ev, ei = O.rmq_brute(A.cpu().numpy(), Q1.cpu().numpy())
    np.testing.assert_array_equal(out1.cpu().numpy(), ei)
    ix.close()


def test_output_buffers_are_checked(gpu):
    A = _dev(O.gen_array(1000, 2), gpu)
    Q = _dev(O.gen_queries(100, 1000, -1, 3), gpu)
    with R.RMQIndex(A, 64) as ix:
        for bad in (torch.empty(99, dtype=torch.int32, device=gpu), torch.empty(100, dtype=torch.int64, device=gpu),
                    torch.empty(100, dtype=torch.int32), torch.empty(200, dtype=torch.int32, device=gpu)[::2]):
            with pytest.raises(TypeError):
                ix.query(Q, want_value=False, out_idx=bad)
        with pytest.raises(TypeError):
            R.Plan([ix], [Q], out_idx=[torch.empty(50, dtype=torch.int32, device=gpu)])
        with pytest.raises(ValueError):
            R.Plan([ix], [Q])
        ok = torch.full((150,), -4, dtype=torch.int32, device=gpu)  # longer than q: only q entries written
        ix.query(Q, want_value=False, out_idx=ok)
        torch.cuda.synchronize()
    got = ok.cpu().numpy()
    np.testing.assert_array_equal(got[:100], O.rmq_brute(A.cpu().numpy(), Q.cpu().numpy())[1])
    assert np.all(got[100:] == -4)

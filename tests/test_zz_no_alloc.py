"""rmq_plan_launch_static, the allocation-free form of a plan launch, against
rmq_plan_create + rmq_plan_launch on the same handles and buffers: answers
bit for bit the same, on two calls in a row (the second with fewer arrays
and fewer queries, so stale descriptors or work items of the first call in
the shared static scratch would show), and against the O(1) oracle."""
import ctypes

import numpy as np
import pytest
import torch

import rtxrmq_amd as R
from rtxrmq_amd import _native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _workload(rng, m, max_log_n, gpu):
    """m arrays of log-uniform lengths, each with a mixed batch of queries
    (uniform, lognormal 2^14 and 2^6 scaled to n, and fixed lengths), built
    with and without edge tables at several block sizes."""
    arrays, indexes, Qs = [], [], []
    for k in range(m):
        n = int(np.exp2(rng.uniform(6, max_log_n)))
        A_np = O.gen_array(n, 100 + k, "ties16" if k % 3 == 0 else "uniform")
        q = int(rng.integers(1, 3 * n))
        parts = [O.gen_queries(q // 3 + 1, n, lr, 200 + k) for lr in (-1, -7, -8)]
        Q_np = np.concatenate(parts)[:q]
        A = torch.from_numpy(A_np).to(gpu)
        bs = int(rng.choice([8, 33, 256, 1024]))
        flags = R.BUILD_EDGES if k % 2 else 0
        arrays.append((A_np, Q_np))
        indexes.append(R.RMQIndex(A, bs, flags=flags))
        Qs.append(torch.from_numpy(Q_np).to(gpu))
    return arrays, indexes, Qs


def _outputs(Qs, gpu, fill):
    vals = [torch.full((Q.shape[0],), float(fill), dtype=torch.float32, device=gpu) for Q in Qs]
    idxs = [torch.full((Q.shape[0],), fill, dtype=torch.int32, device=gpu) for Q in Qs]
    return vals, idxs


def _launch_static(indexes, Qs, vals, idxs, opts=None, item_queries=0):
    m = len(indexes)
    P = ctypes.c_void_p * m
    hs = P(*[ix.handle for ix in indexes])
    qp = P(*[Q.data_ptr() for Q in Qs])
    qs = (ctypes.c_int64 * m)(*[Q.shape[0] for Q in Qs])
    vp = P(*[v.data_ptr() for v in vals])
    ip = P(*[i.data_ptr() for i in idxs])
    stream = torch.cuda.current_stream().cuda_stream or None
    return N.lib().rmq_plan_launch_static(hs, m, qp, qs, vp, ip, ctypes.byref(opts) if opts is not None else None,
                                          item_queries, stream)


@pytest.mark.parametrize("opts", [None, R.query_opts(short_max=8, group_lanes=4, long_lanes=16)])
def test_static_launch_matches_plan_twice(opts, gpu):
    rng = np.random.default_rng(2024)
    arrays, indexes, Qs = _workload(rng, 40, 17, gpu)
    calls = [(arrays, indexes, Qs), (arrays[3:10], indexes[3:10], [Q[: max(1, Q.shape[0] // 3)] for Q in Qs[3:10]])]
    for call, (arrs, ixs, qs) in enumerate(calls):
        print(f"static launch call {call}: {len(ixs)} arrays, {sum(Q.shape[0] for Q in qs)} queries", flush=True)
        want_v, want_i = _outputs(qs, gpu, -7)
        with R.Plan(ixs, qs, out_idx=want_i, out_val=want_v, opts=opts) as plan:
            plan.launch()
            torch.cuda.synchronize()
        got_v, got_i = _outputs(qs, gpu, -9)
        assert _launch_static(ixs, qs, got_v, got_i, opts=opts) == R.RMQ_OK
        torch.cuda.synchronize()
        for k, ((A_np, Q_np), Q) in enumerate(zip(arrs, qs)):
            gv, gi = got_v[k].cpu().numpy(), got_i[k].cpu().numpy()
            wv, wi = want_v[k].cpu().numpy(), want_i[k].cpu().numpy()
            assert np.array_equal(gi, wi) and np.array_equal(gv.view(np.int32), wv.view(np.int32)), \
                f"call {call} array {k}: static launch differs from the plan"
            f = O.FastOracle(A_np, nt=0)
            fv, fi = f.query(Q.cpu().numpy())
            f.close()
            assert np.array_equal(gi, fi) and np.array_equal(gv.view(np.int32), fv.view(np.int32)), \
                f"call {call} array {k}: static launch differs from the oracle"
    for ix in indexes:
        ix.close()


def test_static_launch_limits(gpu):
    """Sizes beyond the static scratch are refused before anything runs: more
    than 8192 arrays, or an explicit item size that needs more than 65536
    work items; with item_queries 0 the items are enlarged instead."""
    A = R.gen_array(1 << 20, 5, device=gpu)
    Q = R.gen_queries_device(1 << 20, 1 << 20, -3, 5, device=gpu)
    with R.RMQIndex(A, 64) as ix:
        v, i = _outputs([Q], gpu, 0)
        assert _launch_static([ix], [Q], v, i, item_queries=8) == N.RMQ_ERR_LIMIT
        assert _launch_static([ix], [Q], v, i, item_queries=0) == R.RMQ_OK
        torch.cuda.synchronize()
        wv, wi = ix.query(Q)
        torch.cuda.synchronize()
        assert torch.equal(i[0], wi) and torch.equal(v[0].view(torch.int32), wv.view(torch.int32))
    lib = R.lib()
    assert lib.rmq_plan_launch_static(None, 8193, None, None, None, None, None, 0, None) == N.RMQ_ERR_LIMIT

"""Serving loop: continuous batching over resident arrays (rmq_server_*).

Each answer must be the same alone and in the mix, and equal to the CPU
oracle's (tolerance 0: exact value bits and leftmost index).  Admission by
device memory is exercised with a small budget: requests wait for memory
instead of failing, a request that can never fit fails cleanly, and nothing
else is disturbed.
"""
import ctypes

import numpy as np
import pytest
import torch

import rtxrmq_amd as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def arrays(gpu):
    out = []
    for i, (n, kind) in enumerate([(1 << 16, "uniform"), (100_003, "ties16"), (1 << 20, "uniform"),
                                   (777, "ties16")]):
        A = O.gen_array(n, 40 + i, kind)
        out.append((A, torch.from_numpy(A).to(gpu)))
    return out


def request_payloads(arrays, count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        a = int(rng.integers(0, len(arrays)))
        n = arrays[a][0].size
        q = int(np.exp2(rng.uniform(4, 15)))
        lr = [-1, -7, -8][k % 3]
        reqs.append((a, O.gen_queries(q, n, lr, seed * 1000 + k)))
    return reqs


@pytest.mark.parametrize("flags", [0, R.BUILD_EDGES, "mixed"])
def test_alone_and_in_the_mix(arrays, flags, gpu):
    """Answers are the same alone and coalesced with others; with "mixed",
    arrays with and without edge tables share launches."""
    with R.Server(slots=3) as sv:
        ids = [sv.add_array(Ad, bs, (R.BUILD_EDGES if k % 2 else 0) if flags == "mixed" else flags)
               for k, ((_, Ad), bs) in enumerate(zip(arrays, (64, 1000, 256, 3)))]
        reqs = request_payloads(arrays, 48, 5)
        # request 0 alone
        a0, Q0 = reqs[0]
        alone = np.empty(Q0.shape[0], np.int32)
        alone_v = np.empty(Q0.shape[0], np.float32)
        tk = sv.submit(ids[a0], Q0, out_idx=alone, out_val=alone_v)
        assert sv.wait(tk)[0] == R.RMQ_OK
        # everything at once, request 0 included again
        outs, tickets = [], []
        for a, Q in reqs:
            oi = np.full(Q.shape[0], -5, np.int32)
            ov = np.zeros(Q.shape[0], np.float32)
            outs.append((oi, ov))
            tickets.append(sv.submit(ids[a], Q, out_idx=oi, out_val=ov))
        lats = []
        for tk in tickets:
            rc, lat = sv.wait(tk)
            assert rc == R.RMQ_OK
            lats.append(lat)
        st = sv.stats()
    np.testing.assert_array_equal(outs[0][0], alone)
    np.testing.assert_array_equal(outs[0][1].view(np.int32), alone_v.view(np.int32))
    for k, ((a, Q), (oi, ov)) in enumerate(zip(reqs, outs)):
        ev, ei = O.rmq_brute(arrays[a][0], Q)
        np.testing.assert_array_equal(oi, ei, err_msg=f"request {k}")
        np.testing.assert_array_equal(ov.view(np.int32), ev.view(np.int32), err_msg=f"request {k}")
    assert st["requests_done"] == 49 and st["failed"] == 0
    assert st["launches"] <= 49 and st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert min(lats) > 0


def test_memory_backpressure_and_oversized_request(arrays, gpu):
    budget = 1 << 20  # 1 MiB of request buffers in flight
    with R.Server(slots=2, mem_budget_bytes=budget) as sv:
        aid = sv.add_array(arrays[2][1], 256)
        n = arrays[2][0].size
        Qs = [O.gen_queries(40_000, n, -7, 90 + k) for k in range(12)]  # ~0.6 MiB each: one at a time
        outs = [np.empty(40_000, np.int32) for _ in Qs]
        tks = [sv.submit(aid, Q, out_idx=o) for Q, o in zip(Qs, outs)]
        big = O.gen_queries(200_000, n, -1, 7)  # 2.4 MiB: can never fit
        big_out = np.empty(200_000, np.int32)
        tk_big = sv.submit(aid, big, out_idx=big_out)
        for tk in tks:
            assert sv.wait(tk)[0] == R.RMQ_OK
        assert sv.wait(tk_big)[0] == R.RMQ_ERR_OOM
        st = sv.stats()
    assert st["peak_inflight_bytes"] <= budget and st["failed"] == 1 and st["requests_done"] == 12
    for Q, o in zip(Qs, outs):
        np.testing.assert_array_equal(o, O.rmq_brute(arrays[2][0], Q)[1])


def test_invalid_server_arguments(gpu):
    lib = R.lib()
    h = ctypes.c_void_p()
    cfg = R._native.ServerConfig(-1, 0, 0, 0, 0, 0, R.QueryOpts())
    assert lib.rmq_server_create(ctypes.byref(cfg), ctypes.byref(h)) == R.RMQ_ERR_INVALID_ARG
    with R.Server() as sv:
        with pytest.raises(R.RMQError):
            sv.submit(99, np.zeros((4, 2), np.int32), out_idx=np.empty(4, np.int32))
        A = torch.zeros(16, dtype=torch.float32, device=gpu)
        with pytest.raises(R.RMQError):
            sv.add_array(A, 4, 0x80)  # unknown build flag


def test_derived_budget_yields_to_resident_arrays(arrays, gpu):
    with R.Server(slots=1) as sv:
        b0 = sv.stats()["budget_bytes"]
        aid = sv.add_array(arrays[2][1], 64, R.BUILD_EDGES)  # ~40 MB of tables
        b1 = sv.stats()["budget_bytes"]
        Q = O.gen_queries(5000, arrays[2][0].size, -1, 3)
        out = np.empty(5000, np.int32)
        assert sv.wait(sv.submit(aid, Q, out_idx=out))[0] == R.RMQ_OK
    assert 0 < b1 < b0
    np.testing.assert_array_equal(out, O.rmq_brute(arrays[2][0], Q)[1])


@pytest.mark.parametrize("stage", [False, True])
def test_pinned_buffers_reused_with_new_contents(arrays, stage, gpu):
    """Pinned buffers are read and written by the kernel in place (unless
    staging is forced); refilling them between requests must be seen, and an
    interior slice must be written exactly, nothing around it."""
    A, Ad = arrays[2]
    n, q = A.size, 30000
    with R.Server(slots=2, stage_copies=stage) as sv:
        aid = sv.add_array(Ad, 256, R.BUILD_EDGES)
        Qp = torch.empty((q, 2), dtype=torch.int32, pin_memory=True)
        big_i = torch.empty(q + 6000, dtype=torch.int32, pin_memory=True)
        big_v = torch.empty(q + 6000, dtype=torch.float32, pin_memory=True)
        out_i, out_v = big_i[5000:5000 + q], big_v[5000:5000 + q]
        for k, lr in enumerate((-1, -7, -8, 5)):
            Q = O.gen_queries(q, n, lr, 500 + k)
            Qp.numpy()[:] = Q
            big_i.fill_(-9)
            big_v.fill_(-2.0)
            rc, _ = sv.wait(sv.submit(aid, Qp, out_idx=out_i, out_val=out_v))
            assert rc == R.RMQ_OK
            ev, ei = O.rmq_brute(A, Q)
            np.testing.assert_array_equal(out_i.numpy(), ei, err_msg=f"batch {k}")
            np.testing.assert_array_equal(out_v.numpy().view(np.int32), ev.view(np.int32), err_msg=f"batch {k}")
            assert np.all(big_i.numpy()[:5000] == -9) and np.all(big_i.numpy()[5000 + q:] == -9)
            assert np.all(big_v.numpy()[:5000] == -2.0) and np.all(big_v.numpy()[5000 + q:] == -2.0)
        # pinned queries with pageable outputs: staged
        Q = O.gen_queries(q, n, -1, 600)
        Qp.numpy()[:] = Q
        host_out = np.empty(q, np.int32)
        assert sv.wait(sv.submit(aid, Qp, out_idx=host_out))[0] == R.RMQ_OK
        np.testing.assert_array_equal(host_out, O.rmq_brute(A, Q)[1])
        st = sv.stats()
        with pytest.raises(TypeError):
            sv.submit(aid, Qp, out_idx=torch.empty(q, dtype=torch.int64, pin_memory=True))
        with pytest.raises(TypeError):
            sv.submit(aid, Qp, out_idx=np.empty(q - 1, np.int32))
    assert st["failed"] == 0 and st["peak_inflight_bytes"] > 0  # the staged request used device memory


def test_arrays_added_while_requests_run(arrays, gpu):
    """An array added while requests on others are in flight: the build runs
    on the server's own stream outside its lock, and the ids, the descriptor
    table and every answer stay right (arrays generated on the device just
    before being added: the add must see their finished values)."""
    A0, Ad0 = arrays[2]
    with R.Server(slots=2) as sv:
        a0 = sv.add_array(Ad0, 256, R.BUILD_EDGES)
        first = [O.gen_queries(200_000, A0.size, lr, 70 + k) for k, lr in enumerate((-1, -7, -8, -1))]
        first_out = [np.empty(Q.shape[0], np.int32) for Q in first]
        tks = [sv.submit(a0, Q, out_idx=o) for Q, o in zip(first, first_out)]
        added = []
        for k, (n, kind) in enumerate([(300_001, "ties16"), (1 << 19, "uniform")]):
            Ad = R.gen_array(n, 900 + k, kind, device=gpu)  # on torch's stream, not yet finished
            added.append((sv.add_array(Ad, (64, 1024)[k], R.BUILD_EDGES if k else 0), Ad))
        assert [aid for aid, _ in added] == [a0 + 1, a0 + 2]
        second, second_out = [], []
        for aid, Ad in added:
            Q = O.gen_queries(100_000, Ad.numel(), -8, aid)
            second.append(Q)
            second_out.append(np.empty(Q.shape[0], np.int32))
            tks.append(sv.submit(aid, Q, out_idx=second_out[-1]))
        assert all(sv.wait(tk)[0] == R.RMQ_OK for tk in tks)
        hosts = [Ad.cpu().numpy() for _, Ad in added]
    f = O.FastOracle(A0)
    for Q, o in zip(first, first_out):
        np.testing.assert_array_equal(o, f.query(Q)[1])
    f.close()
    for k, (A, Q, o) in enumerate(zip(hosts, second, second_out)):
        kind = ("ties16", "uniform")[k]
        np.testing.assert_array_equal(A.view(np.int32), O.gen_array(A.size, 900 + k, kind).view(np.int32))
        f = O.FastOracle(A)
        np.testing.assert_array_equal(o, f.query(Q)[1], err_msg=f"added array {k}")
        f.close()


def test_several_waiters_on_one_ticket_while_submits_rehash(arrays, gpu):
    """Three threads wait on one ticket while the main thread submits
    thousands of small requests (the ticket map rehashes many times under
    the waiters).  A waiter inside the wait when the request completes gets
    its status; one that arrives after the last waiter left finds the ticket
    forgotten (RMQ_ERR_INVALID_ARG), as a later wait does.  The request is
    large (2^23 staged queries) so the three normally all wait.  Every answer
    equals the oracle's."""
    import threading
    A, Ad = arrays[2]
    with R.Server(slots=2) as sv:
        aid = sv.add_array(Ad, 256, R.BUILD_EDGES)
        Qb = O.gen_queries(1 << 23, A.size, -1, 77)
        big_i = np.full(Qb.shape[0], -5, np.int32)
        tk = sv.submit(aid, Qb, out_idx=big_i)
        got = []
        waiters = [threading.Thread(target=lambda: got.append(sv.wait(tk)[0])) for _ in range(3)]
        for w in waiters:
            w.start()
        small = []
        for k in range(3000):
            Q = O.gen_queries(16, A.size, -3, 9000 + k)
            oi = np.full(16, -5, np.int32)
            small.append((Q, oi, sv.submit(aid, Q, out_idx=oi)))
        for w in waiters:
            w.join(timeout=60)
        assert not any(w.is_alive() for w in waiters)
        assert len(got) == 3 and R.RMQ_OK in got and set(got) <= {R.RMQ_OK, R.RMQ_ERR_INVALID_ARG}, got
        assert sv.wait(tk)[0] == R.RMQ_ERR_INVALID_ARG  # forgotten once its last waiter returned
        for Q, oi, t in small:
            assert sv.wait(t)[0] == R.RMQ_OK
    f = O.FastOracle(A)
    np.testing.assert_array_equal(big_i, f.query(Qb)[1])
    f.close()
    for Q, oi, _ in small:
        np.testing.assert_array_equal(oi, O.rmq_brute(A, Q)[1])

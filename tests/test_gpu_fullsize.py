"""Parity at the full sizes of the batched, served and stress configurations,
every answer against the O(1) CPU oracle (leftmost minimum, exact value
bits where values are returned; CPU BASE semantics, src/cpu_methods.h:42-53).

- batched: 4096 arrays with n_i log-uniform in [2^10, 2^22] (about 2e9
  elements), q_i = n_i queries of mixed lengths, int32 indices, the
  committed batched pick; the plan's eager launch and two HIP-graph replays.
  The arrays sit in one buffer after a lead gap of 2^30 elements, so about
  half of them start past element 2^31 of the allocation.
- served: 16 resident arrays with n log-uniform in [2^24, 2^27], requests of
  q log-uniform in [2^8, 2^22], submitted all at once (the mix) and a subset
  again one at a time (alone): the same answers both ways, and the oracle's.
- stress: n=2^27 tie-heavy and strictly descending arrays, fixed lengths
  1, 2^8, 2^16 and n-l, 2^26 queries each, at the committed tuner pick and
  at the literal bs=2^15.
"""
import os
import shutil

import numpy as np
import pytest
import torch

import rtxrmq_amd as R
from rtxrmq_amd import _native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CACHE = os.path.join(os.path.dirname(R.__file__), "tune_cache_mi355x.txt")
MIX = (-1, -7, -8)  # uniform, lognormal 2^14 and 2^6 scaled to n


def _cache_copy(tmp_path):
    dst = tmp_path / "picks.txt"
    shutil.copy(CACHE, dst)
    return str(dst)


def _fill_mixed(Q, n, seed):
    """Q (device, (q, 2) int32): thirds of uniform, lognormal 2^14-scaled and
    2^6-scaled lengths, as the batched and served workloads draw them."""
    parts = np.array_split(np.arange(Q.shape[0]), 3)
    for lr, part in zip(MIX, parts):
        if len(part):
            sub = Q[int(part[0]):int(part[-1]) + 1]
            N.check(N.lib().rmq_gen_queries(sub.data_ptr(), len(part), n, lr, seed, None), "rmq_gen_queries")


def test_batched_config_full_size(tmp_path, gpu):
    seed, m, gap = 1321, 4096, 1 << 30
    rng = np.random.default_rng(seed)
    ns = np.floor(np.exp2(rng.uniform(10, 22, m))).astype(np.int64)
    a_off = np.zeros(m + 1, np.int64)
    a_off[0] = gap
    a_off[1:] = gap + np.cumsum((ns + 63) // 64 * 64)
    q_off = np.zeros(m + 1, np.int64)
    q_off[1:] = np.cumsum(ns)
    assert int((a_off[:-1] > 1 << 31).sum()) > m // 4
    A_all = torch.empty(int(a_off[-1]), dtype=torch.float32, device=gpu)
    N.check(N.lib().rmq_gen_array(A_all.data_ptr(), A_all.numel(), seed, 0, None), "rmq_gen_array")
    Q_all = torch.empty((int(q_off[-1]), 2), dtype=torch.int32, device=gpu)
    A_views = [A_all[a_off[i]:a_off[i] + ns[i]] for i in range(m)]
    Q_views = [Q_all[q_off[i]:q_off[i + 1]] for i in range(m)]
    for i in range(m):
        _fill_mixed(Q_views[i], int(ns[i]), seed * 7919 + i)
    torch.cuda.synchronize()
    big = int(np.argmax(ns))
    pick = R.tune(A_views[big], Q_views[big], bs_list=[64, 128, 256, 512, 1024, 4096], cache_path=_cache_copy(tmp_path),
                  want_value=False, flags_allowed=R.BUILD_EDGES)
    assert pick["from_cache"], f"batched pick {pick['key']} missed the committed cache"
    print(f"batched full size: {m} arrays, {int(ns.sum())} elements, pick bs={pick['bs']} flags={pick['flags']}",
          flush=True)
    indexes = [R.RMQIndex(A, pick["bs"], flags=pick["flags"]) for A in A_views]
    out = torch.full((int(q_off[-1]),), -7, dtype=torch.int32, device=gpu)
    outs = [out[q_off[i]:q_off[i + 1]] for i in range(m)]
    with R.Plan(indexes, Q_views, out_idx=outs, opts=pick["opts"]) as plan:
        plan.graph_launch()
        torch.cuda.synchronize()
        first = out.clone()
        out.fill_(-9)
        torch.cuda.synchronize()
        plan.graph_launch()
        torch.cuda.synchronize()
        assert torch.equal(out, first), "second graph replay differs from the first"
        out.fill_(-5)
        plan.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, first), "eager launch differs from the graph replays"
    for ix in indexes:
        ix.close()
    got = first.cpu().numpy()
    Q_host = Q_all.cpu().numpy()
    del first, out, Q_all
    bad = 0
    for i in range(m):
        A_i = A_views[i].cpu().numpy()
        f = O.FastOracle(A_i)
        _, fi = f.query(Q_host[q_off[i]:q_off[i + 1]])
        f.close()
        bad += int((got[q_off[i]:q_off[i + 1]] != fi).sum())
    assert bad == 0, f"{bad} answers of {int(q_off[-1])} differ from the oracle"


def test_served_config_full_size(tmp_path, gpu):
    seed, m = 1321, 16
    rng = np.random.default_rng(seed + 5)
    ns = np.floor(np.exp2(rng.uniform(24, 27, m))).astype(np.int64)
    arrays = [R.gen_array(int(n), seed * 31 + i, "uniform", device=gpu) for i, n in enumerate(ns)]
    big = int(np.argmax(ns))
    sample = torch.cat([R.gen_queries_device(1 << 20, int(ns[big]), lr, 77, device=gpu) for lr in MIX])
    pick = R.tune(arrays[big], sample, bs_list=[16, 32, 64, 128, 256, 1024], cache_path=_cache_copy(tmp_path),
                  want_value=False, flags_allowed=R.BUILD_EDGES)
    assert pick["from_cache"], f"served pick {pick['key']} missed the committed cache"
    reqs = []  # (array, pinned queries)
    for k in range(96):
        a = k % m
        q = int(np.exp2(rng.uniform(8, 22)))
        Qd = torch.empty((q, 2), dtype=torch.int32, device=gpu)
        _fill_mixed(Qd, int(ns[a]), seed * 1009 + k)
        reqs.append((a, Qd.cpu().pin_memory()))
    print(f"served full size: {m} arrays, {int(ns.sum())} elements, {len(reqs)} requests, "
          f"{sum(Q.shape[0] for _, Q in reqs)} queries, pick bs={pick['bs']} flags={pick['flags']}", flush=True)
    with R.Server(slots=3, opts=pick["opts"]) as sv:
        ids = [sv.add_array(A, pick["bs"], pick["flags"]) for A in arrays]
        mixed = [torch.full((Q.shape[0],), -7, dtype=torch.int32).pin_memory() for _, Q in reqs]
        tickets = [sv.submit(ids[a], Q, out_idx=o) for (a, Q), o in zip(reqs, mixed)]
        assert all(sv.wait(t)[0] == R.RMQ_OK for t in tickets)
        for k in range(0, len(reqs), 6):  # alone: one request in flight at a time
            a, Q = reqs[k]
            alone = torch.full((Q.shape[0],), -9, dtype=torch.int32).pin_memory()
            assert sv.wait(sv.submit(ids[a], Q, out_idx=alone))[0] == R.RMQ_OK
            assert torch.equal(alone, mixed[k]), f"request {k}: alone and in the mix differ"
    for a in range(m):
        f = O.FastOracle(arrays[a].cpu().numpy())
        for k, (ak, Q) in enumerate(reqs):
            if ak == a:
                _, fi = f.query(Q.numpy())
                assert np.array_equal(mixed[k].numpy(), fi), f"request {k} (array {a}) differs from the oracle"
        f.close()


@pytest.mark.parametrize("kind", ["ties16", "desc"])
def test_stress_config_full_queries(kind, tmp_path, gpu):
    n, q, seed = 1 << 27, 1 << 26, 1321
    A_dev = R.gen_array(n, 99, kind, device=gpu)
    A = A_dev.cpu().numpy()
    f = O.FastOracle(A)
    cache = _cache_copy(tmp_path)
    v = torch.empty(q, dtype=torch.float32, device=gpu)
    i = torch.empty(q, dtype=torch.int32, device=gpu)
    for lr in (1, 1 << 8, 1 << 16, -6):
        Q = R.gen_queries_device(q, n, lr, seed, device=gpu)
        pick = R.tune(A_dev, R.gen_queries_device(1 << 24, n, lr, seed + 7919, device=gpu), cache_path=cache)
        assert pick["from_cache"], f"{kind} lr={lr}: pick {pick['key']} missed the committed cache"
        fv, fi = f.query(Q.cpu().numpy())
        for bs, flags, opts in ((pick["bs"], pick["flags"], pick["opts"]), (1 << 15, 0, None)):
            print(f"stress {kind} lr={lr} bs={bs} flags={flags}", flush=True)
            with R.RMQIndex(A_dev, bs, flags=flags) as ix:
                ix.query(Q, opts=opts, out_val=v, out_idx=i)
                torch.cuda.synchronize()
            gi, gv = i.cpu().numpy(), v.cpu().numpy()
            assert np.array_equal(gi, fi) and np.array_equal(gv.view(np.int32), fv.view(np.int32)), \
                f"{kind} lr={lr} bs={bs}: {int((gi != fi).sum())} index mismatches"
        del Q
    f.close()

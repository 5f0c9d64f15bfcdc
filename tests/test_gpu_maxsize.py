"""The largest array the int32 interface allows: n = 2^31 - 1.

Every index computation near the top of the int32 range (block starts, the
16-byte alignment of scans, thread ids of per-element build kernels, grid
sizes) is exercised by three structurally different indexes over the same
8.6 GB array -- bs=2^15 without edge tables, bs=1000 with them, bs=7 (a 29-
level table), bs=2^24 with edge tables (128 blocks: the block sparse table is
read from LDS) -- which must agree bit for bit.  Every answer is checked to lie
in [l, r] and to be the value stored there; ranges up to 4096 elements are
checked against the brute force on the device (algs 2/102), and long ranges
(including [0, n-1]) against a torch minimum and first match.
"""
import numpy as np
import pytest
import torch

import rtxrmq_amd as R

pytestmark = pytest.mark.gpu

N = 2 ** 31 - 1


def queries(rng):
    n = N
    parts = []
    l = rng.integers(0, n, 1024)
    ln = rng.integers(1, 4097, 1024)
    top = n - 1 - np.arange(256)  # ranges ending at or near the last element
    l[:256] = np.maximum(top - ln[:256] + 1, 0)
    parts.append(np.stack([l, np.minimum(l + ln - 1, n - 1)], 1))
    parts.append(R.gen_queries(1024, n, -2, 3))
    parts.append(R.gen_queries(1024, n, -1, 4))
    parts.append(R.gen_queries(1024, n, -6, 5))
    parts.append(np.array([[0, n - 1], [1, n - 1], [n - 1, n - 1], [n - 2, n - 1], [0, 0], [n - 5, n - 2]]))
    return np.concatenate(parts).astype(np.int32)


def test_maximum_array_size(gpu):
    A = R.gen_array(N, 17, "uniform", device=gpu)
    Q = queries(np.random.default_rng(5))
    Qd = torch.from_numpy(Q).to(gpu)
    results = []
    for bs, flags in ((1 << 15, 0), (1000, R.BUILD_EDGES), (7, 0), (1 << 24, R.BUILD_EDGES)):
        with R.RMQIndex(A, bs, flags=flags) as ix:
            if bs == 1 << 24:
                assert ix.kernels()[0].startswith("k_query_edges_lds<")
            v, i = ix.query(Qd)
            torch.cuda.synchronize()
            results.append((bs, flags, v.cpu().numpy(), i.cpu().numpy()))
    v0, i0 = results[0][2], results[0][3]
    for bs, flags, v, i in results[1:]:
        np.testing.assert_array_equal(i, i0, err_msg=f"bs={bs} flags={flags}")
        np.testing.assert_array_equal(v.view(np.int32), v0.view(np.int32), err_msg=f"bs={bs} flags={flags}")
    assert np.all(i0 >= Q[:, 0]) and np.all(i0 <= Q[:, 1])
    stored = A[torch.from_numpy(i0.astype(np.int64)).to(gpu)].cpu().numpy()
    np.testing.assert_array_equal(stored.view(np.int32), v0.view(np.int32))
    # short ranges: the device brute force
    short = (Q[:, 1] - Q[:, 0]) < 4096
    bv, bi = R.gpu_base(A, Qd[torch.from_numpy(short).to(gpu)])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(bi.cpu().numpy(), i0[short])
    np.testing.assert_array_equal(bv.cpu().numpy().view(np.int32), v0[short].view(np.int32))
    # long ranges: minimum and its first occurrence by torch
    long_ids = np.nonzero(~short)[0]
    pick = np.concatenate([long_ids[:4], np.nonzero((Q[:, 0] == 0) & (Q[:, 1] == N - 1))[0],
                           np.nonzero(Q[:, 1] == N - 1)[0][:3]])
    for k in pick:
        l, r = int(Q[k, 0]), int(Q[k, 1])
        seg = A[l:r + 1]
        m = seg.min()
        first = l + int(torch.nonzero(seg == m)[0, 0])
        assert first == i0[k] and np.float32(m.item()).view(np.int32) == v0[k].view(np.int32), (l, r)
    del A
    torch.cuda.empty_cache()

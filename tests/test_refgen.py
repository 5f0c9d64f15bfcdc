"""The reference-compatible query stream (rmq_gen_queries_ref) against an
independent restatement of the reference's generator (src/rand.cuh:116-201).

The restatement here uses numpy's MT19937 with the legacy (init_genrand)
seeding -- the same raw 32-bit stream as std::mt19937(seed) -- and spells
out the standard library's algorithms: uniform_int_distribution's Lemire
downscaling for a 32-bit engine (libstdc++ bits/uniform_int_dist.h),
generate_canonical<double> from two 32-bit draws, normal_distribution's
Marsaglia polar method with its cached second value, lognormal = exp(s*N+m).
"""
import math

import numpy as np
import pytest

import rtxrmq_amd as R


class MT:
    def __init__(self, seed):
        self.bg = np.random.MT19937()
        self.bg._legacy_seeding(seed & 0xFFFFFFFF)
        self.buf, self.k = [], 0

    def __call__(self):
        if self.k == len(self.buf):
            self.buf, self.k = [int(x) for x in self.bg.random_raw(4096)], 0
        v = self.buf[self.k]
        self.k += 1
        return v


def uniform_int(g, a, b):
    rng = b - a + 1
    prod = g() * rng
    low = prod & 0xFFFFFFFF
    if low < rng:
        thr = ((1 << 32) - rng) % rng
        while low < thr:
            prod = g() * rng
            low = prod & 0xFFFFFFFF
    return a + (prod >> 32)


def canonical(g):
    s = float(g())
    s += float(g()) * 4294967296.0
    r = s / 18446744073709551616.0
    return math.nextafter(1.0, 0.0) if r >= 1.0 else r


class Normal:
    def __init__(self):
        self.saved = None

    def __call__(self, g):
        if self.saved is not None:
            v, self.saved = self.saved, None
            return v
        while True:
            x = 2.0 * canonical(g) - 1.0
            y = 2.0 * canonical(g) - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2 * math.log(r2) / r2)
        self.saved = x * mult
        return y * mult


def ref_queries(q, n, lr, seed, nt):
    out = np.zeros((q, 2), np.int64)
    chunk = (q + nt - 1) // nt
    for tid in range(nt):
        g = MT((seed * tid) & 0xFFFFFFFF)
        begin, end = chunk * tid, chunk * tid + chunk
        if lr > 0:
            for i in range(begin, min(q, end)):
                l = uniform_int(g, 0, n - 1 - (lr - 1))
                out[i] = (l, l + lr - 1)
        elif lr == -1:
            for i in range(begin, min(q, end)):
                ln = uniform_int(g, 1, n)
                l = uniform_int(g, 0, n - 1 - (ln - 1))
                out[i] = (l, l + ln - 1)
        else:
            scale = {-2: int(math.pow(n, 0.6)), -3: int(math.pow(n, 0.3)), -4: max(1, n // 256),
                     -5: max(1, n // 32768)}[lr]
            nd, m, s = Normal(), math.log(scale), 0.3
            for i in range(begin, min(q, end)):
                while True:
                    x = math.exp(s * nd(g) + m)
                    ln = int(x) if x < 2147483648.0 else -(1 << 31)
                    if 0 < ln <= n:
                        break
                l = uniform_int(g, 0, n - 1 - (ln - 1))
                out[i] = (l, l + ln - 1)
    return out.astype(np.int32)


@pytest.mark.parametrize("lr", [-1, -2, -3, -4, -5, 1, 37])
@pytest.mark.parametrize("nt", [1, 3, 8])
@pytest.mark.parametrize("n,seed", [(1000, 7), (10 ** 6, 1321), (100_000_000, 123456789)])
def test_ref_stream_matches_restatement(lr, nt, n, seed):
    q = 1500
    got = R.gen_queries_ref(q, n, lr, seed, nt)
    np.testing.assert_array_equal(got, ref_queries(q, n, lr, seed, nt))
    assert np.all(got[:, 0] >= 0) and np.all(got[:, 0] <= got[:, 1]) and np.all(got[:, 1] < n)


def test_ref_stream_depends_on_nt_like_the_reference():
    a = R.gen_queries_ref(1000, 10 ** 6, -1, 5, 1)
    b = R.gen_queries_ref(1000, 10 ** 6, -1, 5, 4)
    assert not np.array_equal(a, b)          # thread 0 always seeds 0 (src/rand.cuh:140)
    np.testing.assert_array_equal(a[:250], b[:250])

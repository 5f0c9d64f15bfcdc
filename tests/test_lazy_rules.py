"""The rules that let a long range skip its edge blocks, checked by brute force.

The kernels answer a range [l, r] across blocks lB < rB from block minima
(each block's min and leftmost argmin) before touching any element:
  1. superset: the minimum of blocks lB..rB is the answer when its argmin lies
     in [l, r];
  2. exact edge: block lB's entry is the left part's answer when its argmin is
     >= l; block rB's entry is the right part's when its argmin is <= r;
  3. skipped edge: the left part is needed only when block lB's minimum is <=
     the best so far, the right part only when block rB's is strictly below.
The same rules run over 32-element sub-blocks inside one block.  Here they are
restated on plain Python lists and compared with a brute-force leftmost
minimum over every (l, r) of small arrays with ties and +inf, then bench.py's
byte model (which replays the same decisions on tensors) is compared with the
restatement's counts.
"""
import math
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

INF, IMAX = math.inf, 2 ** 31 - 1


def better(a, b):
    return a[0] < b[0] or (a[0] == b[0] and a[1] < b[1])


def lexmin(a, b):
    return a if better(a, b) else b


def brute(A, l, r):
    best = (INF, IMAX)
    for i in range(l, r + 1):
        best = lexmin((float(A[i]), i), best)
    return best


def block_mins(A, bs):
    """(min, leftmost argmin) per block; an all-+inf block gives its start."""
    return [brute(A, b, min(b + bs, len(A)) - 1) for b in range(0, len(A), bs)]


def superset_window(nb, lB, rB, lines):
    """Blocks [lB, hi] whose minimum the superset test reads: lB..rB from two
    level-lev sparse-table entries, or with lines (a global table, 256-byte
    aligned, 8-byte entries), when those two entries lie in different 128-byte
    lines and the window fits, the one level-(lev+1) entry at lB."""
    lB, rB = int(lB), int(rB)
    c = rB - lB + 1
    lev = c.bit_length() - 1
    a2 = rB - (1 << lev) + 1
    if lines and ((lev * nb + lB) * 8) // 128 != ((lev * nb + a2) * 8) // 128 and lB + (2 << lev) <= nb:
        return lB, lB + (2 << lev) - 1
    return lB, rB


def blocks_min(bm, lo, hi):
    best = (INF, IMAX)
    for b in range(lo, hi + 1):
        best = lexmin(bm[b], best)
    return best


def lazy_two_stage(A, bs, bm, l, r):
    """answer_two_stage (tables far beyond the caches), phase by phase, with
    the entries it reads counted as bench.py's model counts them; returns
    (answer, edge pairs read, superset hit, table entries read).  The table
    is 256-byte aligned with rows of nb 8-byte entries, so entry a of row lev
    lies in line (lev * nb + a) * 8 // 128; every entry read must be a valid
    window (a + 2^lev <= nb)."""
    nb = len(bm)

    def entry(lev, a):
        assert 0 <= a and a + (1 << lev) <= nb, (lev, a, nb)
        return blocks_min(bm, a, a + (1 << lev) - 1)

    def line(lev, a):
        return (lev * nb + a) * 8 // 128

    l, r = int(l), int(r)
    lB, rB = l // bs, r // bs
    c = rB - lB + 1
    lev = c.bit_length() - 1
    a2 = rB - (1 << lev) + 1
    inl = a2 == lB or line(lev, lB) == line(lev, a2)
    wide = not inl and lB + (2 << lev) <= nb
    if wide:
        sup, reads = entry(lev + 1, lB), 1
        if sup[1] >= (rB + 1) * bs:  # past block rB: the exact window after all
            sup, reads = blocks_min(bm, lB, rB), 3
    else:
        sup, reads = lexmin(entry(lev, lB), entry(lev, a2)), 2
    if l <= sup[1] <= r:
        return sup, 0, True, reads
    left = sup[1] < l
    reads += 2
    if inl and a2 > lB:
        s2 = blocks_min(bm, lB + 1, rB) if left else blocks_min(bm, lB, rB - 1)
        entry(lev, lB + 1 if left else lB), entry(lev, a2 if left else a2 - 1)
    elif inl and left and rB + 1 < nb:
        s2 = entry(lev, lB + 1)  # blocks lB+1 .. rB+1
    elif inl and not left and lB > 0:
        s2 = entry(lev, lB - 1)  # blocks lB-1 .. rB-1
    else:
        s2 = blocks_min(bm, lB + 1, rB) if left else blocks_min(bm, lB, rB - 1)
    suf = brute(A, l, lB * bs + bs - 1)
    pre = brute(A, rB * bs, r)
    if left and s2[1] <= r:
        return (suf if suf[0] <= s2[0] else s2), 1, False, reads
    if not left and s2[1] >= l:
        return (pre if pre[0] < s2[0] else s2), 1, False, reads
    best = suf
    if c > 2:
        reads += 2
        mid = blocks_min(bm, lB + 1, rB - 1)
        if mid[0] < best[0]:
            best = mid
    if pre[0] < best[0]:
        best = pre
    return best, 2, False, reads


def lazy(A, bs, bm, l, r, left_counts=True, lines=False, two_stage=False, mid_shift=None):
    """(answer, parts computed, superset hit, sparse-table entries read) for
    lB < rB (a window's minimum costs two entries, a widened window one),
    following rules
    1-3.  left_counts: a computed left part joins the best so far before the
    right part is decided (the edge-table path reads it as one pair); False
    for the scan path, whose left scan is spread over a lane group.
    two_stage: the large-table path (lazy_two_stage), always line-aware.
    mid_shift: rule 1 first over coarse blocks of 2^mid_shift blocks (a row in
    LDS whose entry i covers coarse blocks i and i+1; no table read) when the
    range's blocks lie in at most 4 of them: coarse blocks cl..cr, or cl and
    cl+1 when cr - cl < 2."""
    if mid_shift is not None:
        lB, rB = l // bs, r // bs
        cl, cr = lB >> mid_shift, rB >> mid_shift
        if cr - cl < 4:
            hi = cl + 1 if cr - cl < 2 else cr
            sup = blocks_min(bm, cl << mid_shift, min(len(bm), (hi + 1) << mid_shift) - 1)
            if l <= sup[1] <= r:
                return sup, 0, True, 0
    if two_stage:
        return lazy_two_stage(A, bs, bm, l, r)
    lB, rB = l // bs, r // bs
    lo, hi = superset_window(len(bm), lB, rB, lines)
    sup = blocks_min(bm, lo, hi)
    reads = 1 if hi > rB else 2
    if l <= sup[1] <= r:
        return sup, 0, True, reads
    best = (INF, IMAX)
    for b in range(lB + 1, rB):
        best = lexmin(bm[b], best)
    reads += (2 if rB - lB > 1 else 0) + 2  # whole blocks between; the two edge blocks' entries
    need = 0
    left = (INF, IMAX)
    el = bm[lB]
    if el[1] >= l:
        best = lexmin(el, best)
    elif el[0] <= best[0]:
        left = brute(A, l, lB * bs + bs - 1)
        need += 1
        if left_counts:
            best = lexmin(left, best)
    er = bm[rB]
    if er[1] <= r:
        best = lexmin(er, best)
    elif er[0] < best[0]:
        best = lexmin(brute(A, rB * bs, r), best)
        need += 1
    return lexmin(left, best), need, False, reads


def arrays(rng):
    out = []
    for n in (2, 7, 33, 64, 97):
        out += [rng.integers(0, 3, n).astype(np.float32), rng.integers(0, 16, n).astype(np.float32),
                np.full(n, 5.0, np.float32), np.arange(n, dtype=np.float32),
                (1000 - np.arange(n)).astype(np.float32), np.full(n, np.inf, np.float32)]
        x = rng.integers(0, 4, n).astype(np.float32)
        x[rng.random(n) < 0.5] = np.inf
        out.append(x)
    return out


@pytest.mark.parametrize("left_counts,lines,two_stage,mid_shift",
                         [(True, False, False, None), (False, False, False, None), (True, True, False, None),
                          (True, True, True, None), (True, True, False, 0), (True, True, False, 1),
                          (True, True, False, 2), (True, True, True, 1), (True, True, True, 2)])
@pytest.mark.parametrize("bs", [1, 2, 3, 8, 16, 33])
def test_rules_give_the_leftmost_minimum(bs, left_counts, lines, two_stage, mid_shift):
    """lines: the superset window grows to a level-(lev+1) window when the two
    level-lev entries lie in different 128-byte lines (any window of whole
    blocks that covers [l, r] is a valid superset); two_stage: a failed test
    is retried with the failing side cut off; mid_shift: a superset test over
    coarse blocks (the last one partial) comes first."""
    for A in arrays(np.random.default_rng(bs)):
        bm = block_mins(A, bs)
        for l in range(len(A)):
            for r in range(l, len(A)):
                if l // bs == r // bs:
                    continue
                got = lazy(A, bs, bm, l, r, left_counts, lines, two_stage, mid_shift)[0]
                exp = brute(A, l, r)
                assert got[1] == exp[1] and np.float32(got[0]).tobytes() == np.float32(exp[0]).tobytes(), \
                    (A.tolist(), bs, l, r, got, exp)


def test_superset_rule_inside_one_block_over_sub_blocks():
    """Sub-blocks of 32 restart at each block start (bs a multiple of 32 here)."""
    rng = np.random.default_rng(7)
    for A in arrays(rng) + [rng.integers(0, 8, 300).astype(np.float32)]:
        n = len(A)
        for bs in (64, 96):
            for l in range(n):
                for r in range(l, min(n, l + 100)):
                    if l // bs != r // bs:
                        continue
                    base = l // bs * bs
                    lo = base + (l - base) // 32 * 32
                    hi = min(base + (r - base) // 32 * 32 + 31, min(base + bs, n) - 1)
                    sup = brute(A, lo, hi)
                    if l <= sup[1] <= r:
                        assert sup == brute(A, l, r), (bs, l, r)


@pytest.mark.parametrize("edges,short_max,lines,two_stage,mid_shift",
                         [(False, 32, False, False, None), (True, 32, False, False, None),
                          (True, 1, False, False, None), (True, 1, True, False, None),
                          (True, 32, True, True, None), (True, 1, True, True, None),
                          (True, 1, True, False, 1), (True, 32, True, False, 3), (True, 1, True, True, 2)])
def test_byte_model_replays_the_rules(edges, short_max, lines, two_stage, mid_shift):
    """With edge tables a range inside one 32-element sub-block (one block
    when bs <= 32) is scanned whatever short_max is; every other range longer
    than short_max takes the O(1) path."""
    rng = np.random.default_rng(11 + edges + short_max + 5 * lines + 7 * two_stage + 13 * (mid_shift or 0))
    for trial in range(16):
        n = int(rng.integers(100, 2500))
        bs = int(rng.choice([8, 24, 64, 128, 256] if edges else [4, 16, 37, 64, 256]))
        A = rng.integers(0, int(rng.choice([3, 50, 10 ** 6])), n).astype(np.float32)
        if trial % 4 == 0 and not edges:
            A[rng.random(n) < 0.3] = np.inf
        Q = []
        for _ in range(300):
            l = int(rng.integers(0, n))
            r = int(rng.integers(l, min(n, l + int(rng.choice([40, 300, n])))))
            Q.append((l, r))
        Q = np.array(Q, np.int32)
        bm = block_mins(A, bs)
        parts = need = gathers = n_short = reads = 0
        for l, r in Q:
            lB, rB = l // bs, r // bs
            one_sub = lB == rB and (l - lB * bs) // 32 == (r - lB * bs) // 32
            if r - l + 1 <= short_max or (edges and one_sub):
                n_short += 1
                continue
            if lB == rB:
                if edges:
                    base = lB * bs
                    lo = base + (l - base) // 32 * 32
                    hi = min(base + (r - base) // 32 * 32 + 31, min(base + bs, n) - 1)
                    s = brute(A, lo, hi)
                    hit = l <= s[1] <= r
                    gathers += 0 if hit else 2
                    reads += 2 + (0 if hit or (r - base) // 32 - (l - base) // 32 <= 1 else 2)
                continue
            _, k, _, rd = lazy(A, bs, bm, l, r, left_counts=edges, lines=lines, two_stage=two_stage,
                               mid_shift=mid_shift)
            parts += 2
            need += k
            reads += rd
            gathers += k if edges else 0
        lvl = bench.BlockLevel(torch.from_numpy(A), bs)
        sub = bench.BlockLevel(torch.from_numpy(A), 32) if edges else None
        ab = bench.algorithmic_bytes(torch.from_numpy(Q), lvl, short_max, edges=edges, sub=sub,
                                     A=torch.from_numpy(A), lines=lines, two_stage=two_stage, mid_shift=mid_shift)
        assert ab["n_short"] == n_short
        assert abs(ab["edge_parts_computed"] - (need / parts if parts else 0.0)) < 1e-12
        assert ab["gathers"] == gathers
        assert ab["table_reads"] == reads
        for k in range(lvl.nb):  # block minima: leftmost argmin, all-+inf block -> its start
            assert int(lvl.I[0, k]) == bm[k][1]

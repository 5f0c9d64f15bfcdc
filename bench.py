"""Headline benchmark: batch RMQ at n=1e8, q=2^26, three length distributions.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--bs 32768]

A step is one query pass (alg 5 values + alg 105 indices in one launch set)
over each of the three query batches: uniform lengths (lr=-1), lognormal
medium (lr=-2) and lognormal small (lr=-3), 2^26 queries each, over n=1e8
float32 U[0,1) values.  Inputs are resident in HBM before timing starts.
Two runs: the literal block size bs=2^15 with default thresholds
(`fixed_bs_32768`), and the configuration the engine's autotuner picks per
batch (block size, short-range threshold, lanes per query) from a timed
2^20-query sample -- the way the reference's published runs tune --nb per
distribution (scripts/perf-fulltest.sh:45-58).  value = 3*q*K / (timed
seconds) of the tuned run; answers are bit-identical either way.

roofline: the dominant kernel of the run -- algorithmic bytes per query,
the block method read literally: B = 8 (query) + 4*S (elements scanned: the
range when it lies inside one block, else both edge-block parts) + T (two
8-byte table entries when whole blocks lie between) + 8 (value + index
out), summed over the queries that kernel processed and divided by its time
measured with HIP events on the stream it runs on.  achieved_model replays
the kernels' own decisions instead (skipped and exact edges, superset
tests, edge-table pairs): the bytes this design must move.  `traffic` is the
HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE
(profiles/traffic.json, written by --reduce-pmc), for exactly this workload
or null, and per batch measured_bytes_per_rmq / dram_read_lines_per_rmq
give the same counters per query, comparable across designs.  The query
kernels are bound by random reads (sparse-table entries, mostly from L2,
and edge pairs from DRAM) rather than bytes: `gather_ceiling` sets their
counts against rmq_probe_gather's random-read rates on the same GPU.

stress: n=2^27 tie-heavy and strictly descending arrays, fixed lengths
1, 2^8, 2^16 and n-l, the tuner's pick against the literal method swept
over bs = 2^10 .. 2^18, answers compared.

cpu_baseline: the reference's HRMQ (RMQRMM64, built from the reference's own
sources into oracle/_ref by oracle/Makefile) on this host's cores, threads
over queries as src/cpu_methods.h:104-110, on a bounded sample of each batch,
its answers checked against the GPU's; the engine's own HRMQ (rmq_hrmq_*) timed on the
same sample; CPU BASE (algs 0/100) and HRMQ at n=2^20, q=2^20 beside the GPU.
"""
from __future__ import annotations

import argparse
import faulthandler
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DISTS = (-1, -2, -3)
DIST_NAMES = {-1: "uniform", -2: "lognormal_medium", -3: "lognormal_small"}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


class BlockLevel:
    """Copy of an index's block level for the byte model: (min, leftmost
    argmin) per block and a sparse table over them, as torch tensors on the
    array's device, to replay which edge parts each query needs."""

    def __init__(self, A, bs: int):
        import torch
        n, dev = A.numel(), A.device
        nb = -(-n // bs)
        pad = torch.full((nb * bs,), float("inf"), dtype=torch.float32, device=dev)
        pad[:n] = A
        rows = pad.view(nb, bs)
        v = rows.min(dim=1).values
        pos = torch.arange(bs, device=dev).expand(nb, bs)
        j = torch.where(rows == v[:, None], pos, bs).min(dim=1).values  # leftmost; all +inf: 0
        i = torch.arange(nb, device=dev, dtype=torch.int64) * bs + j
        del pad, rows, pos
        self.bs, self.nb = bs, nb
        levels = max(1, nb.bit_length())
        self.V = torch.full((levels, nb), float("inf"), dtype=torch.float32, device=dev)
        self.I = torch.full((levels, nb), 2 ** 31 - 1, dtype=torch.int64, device=dev)
        self.V[0], self.I[0] = v, i
        for k in range(1, levels):
            h, cnt = 1 << (k - 1), nb - (1 << k) + 1
            xv, xi = self.V[k - 1, :cnt], self.I[k - 1, :cnt]
            yv, yi = self.V[k - 1, h:h + cnt], self.I[k - 1, h:h + cnt]
            t = yv < xv  # the left window wins ties
            self.V[k, :cnt] = torch.where(t, yv, xv)
            self.I[k, :cnt] = torch.where(t, yi, xi)

    def range_min(self, a, c):
        """Leftmost minimum of blocks [a, a+c) for c >= 1, elementwise."""
        import torch
        lev = (torch.frexp(c.double())[1] - 1).long()  # floor(log2 c), exact for c < 2^53
        b = a + c - (1 << lev)
        xv, xi, yv, yi = self.V[lev, a], self.I[lev, a], self.V[lev, b], self.I[lev, b]
        t = yv < xv
        return torch.where(t, yv, xv), torch.where(t, yi, xi)

    def superset(self, lB, rB, lines: bool, retry: bool = False):
        """(value, index, table entries read) of the superset test of blocks
        lB < rB: the minimum of lB..rB from two level-lev entries, or, with
        lines (a global table, 256-byte aligned), when those two lie in
        different 128-byte lines and the window fits, the one level-(lev+1)
        entry at lB (blocks lB .. lB + 2^(lev+1) - 1); with retry, when that
        entry's argmin lies past block rB, the two level-lev entries after it
        (three reads)."""
        import torch
        c = rB - lB + 1
        v, i = self.range_min(lB, c)
        reads = torch.full_like(lB, 2)
        if lines:
            lev = (torch.frexp(c.double())[1] - 1).long()
            base = lev * self.nb
            a2 = rB - torch.bitwise_left_shift(torch.ones_like(lev), lev) + 1
            split = ((base + lB) * 8) // 128 != ((base + a2) * 8) // 128
            one = split & (lB + torch.bitwise_left_shift(torch.full_like(lev, 2), lev) <= self.nb)
            lev1 = torch.where(one, lev + 1, lev)  # level lev+1 exists where one holds
            wv, wi = self.V[lev1, lB], self.I[lev1, lB]
            use = one & (wi < (rB + 1) * self.bs) if retry else one
            v, i = torch.where(use, wv, v), torch.where(use, wi, i)
            reads = torch.where(use, 1, torch.where(one, 3, 2))
        return v, i, reads


HIER_MIN = 4  # kHierMin: whole sub-blocks between the ends before a scan reads their minima
HIER_TABLE = 16  # kHierTable: more whole sub-blocks than this are read as two sparse-table entries


def scan_bytes(lo, hi, bs: int):
    """Bytes of the scan of [lo, hi] inside one block: element by element, or,
    with at least HIER_MIN whole 32-element sub-blocks between the ends (and
    bs > 32), the partial sub-blocks element by element and the whole ones as
    8-byte (min, argmin) pairs, or, beyond HIER_TABLE of them, as two entries
    (16 B) of the sparse table over sub-block minima."""
    import torch
    direct = 4 * (hi - lo + 1)
    if bs <= 32:
        return direct
    base = lo // bs * bs
    jl, jr = (lo - base) // 32, (hi - base) // 32
    whole = jr - jl - 1
    mid = torch.where(whole > HIER_TABLE, torch.full_like(whole, 16), 8 * whole)
    hier = 4 * (base + jl * 32 + 32 - lo) + mid + 4 * (hi - (base + jr * 32) + 1)
    return torch.where(whole >= HIER_MIN, hier, direct)


def left_part_min(A, sub, bs: int, l):
    """Minimum value of A[l .. end of l's block] (the edge pair suf_b[l]):
    the rest of l's 32-element sub-block, then the sub-blocks after it in the
    block.  A: the array tensor; sub: its 32-element level; bs <= 32 or a
    multiple of 32."""
    import torch
    n = A.numel()
    end = torch.clamp((l // bs + 1) * bs, max=n) - 1
    sub_end = end if bs <= 32 else torch.minimum(end, l // 32 * 32 + 31)  # at most 32 elements from l
    k = torch.arange(32, device=l.device)
    idx = l[:, None] + k[None, :]
    vals = torch.where(idx <= sub_end[:, None], A[torch.clamp(idx, max=n - 1)], float("inf"))
    v = vals.min(dim=1).values
    g0, g1 = l // 32 + 1, end // 32
    more = g1 >= g0
    if bs > 32 and bool(more.any()):
        mv, _ = sub.range_min(g0[more], g1[more] - g0[more] + 1)
        v[more] = torch.minimum(v[more], mv)
    return v


MID_MAX_BLOCKS = 8192  # coarse blocks of an RMQ_BUILD_MID handle at most (kMidMaxBlocks)


def mid_shift_of(nb: int) -> int:
    """log2 of the blocks per coarse block of an RMQ_BUILD_MID handle with nb
    blocks: the smallest that leaves at most MID_MAX_BLOCKS coarse blocks."""
    s = 0
    while ((nb - 1) >> s) + 1 > MID_MAX_BLOCKS:
        s += 1
    return s


def algorithmic_bytes(Q, lvl: BlockLevel, short_max: int, out_bytes: int = 8, edges: bool = False,
                      sub: BlockLevel | None = None, A=None, lines: bool = False, two_stage: bool = False,
                      mid_shift: int | None = None) -> dict:
    """Per pass: bytes in the long and the short class, their query counts,
    the fraction of cross-block edge parts computed, the random gathers of
    per-element edge pairs (with edge tables), the sparse-table entries read
    and the 128-byte lines of A the short ranges span.  Q: (q, 2) int32
    tensor on the level's device.

    Short class (<= short_max elements; with edge tables also every range
    inside one 32-element sub-block, one block when bs <= 32): B = 8 +
    4*len + out.  Ranges inside one block without edge tables: B = 8 + the
    scan's bytes (scan_bytes: 4 per element, or 8 per whole 32-element
    sub-block between the ends for long ones) + out.  Every other
    range first reads the minimum of the blocks it touches (of the 32-element
    sub-blocks it touches, inside one block; two table entries, 16 B): when
    its argmin lies in [l, r] that is the answer, B = 8 + 16 + out.
    Otherwise, inside one block (edge tables): + 16 (suffix and prefix pairs)
    + T.  Across blocks: + T + 16 (the two edge blocks' (min, argmin)) + each
    edge part the query needs, where an edge part is needed unless its
    block's leftmost argmin lies inside it (then that entry is exact) or its
    block minimum cannot win (left: > the best so far; right: >= it); a needed
    part costs its scan's bytes (scan_bytes), or 8 B (one edge pair) with
    edge tables.  two_stage (edge tables, large tables; answer_two_stage): a
    failed superset test is followed by the whole blocks of the side that did
    not fail (16 B) and the failing side's edge pair (8 B); when that does not
    answer, by the other edge pair (8 B) and the blocks between (16 B when
    whole blocks lie between).  With edge tables the left pair, once read, counts in the best so
    far before the right part is decided; a scanned left part (spread over a
    lane group) does not.  T = 16 when whole blocks (or sub-blocks) lie
    between.  sub: the 32-element level and A the array, with edge tables.
    lines: the superset test reads one level-up entry instead of two entries
    in different 128-byte lines (BlockLevel.superset), as the edge-table
    kernel with its table in global memory does; with two_stage as well, an
    argmin of that entry past block rB is followed by the two entries.
    mid_shift (edge tables with a coarse row, RMQ_BUILD_MID): a range across
    blocks whose blocks lie in coarse blocks (2^mid_shift blocks) cl..cr,
    cr - cl < 4, first takes from LDS the minimum of coarse blocks cl..cr, or
    of cl and cl+1 when cr - cl < 2 (the last coarse block alone at the end);
    when its argmin lies in [l, r] that is the answer, B = 8 + out, with no
    table read."""
    import torch
    bs = lvl.bs
    long_b = short_b = 0.0
    n_long = n_short = parts = needed = gathers = reads = short_lines = 0
    for c0 in range(0, Q.shape[0], 1 << 23):
        l = Q[c0:c0 + (1 << 23), 0].long()
        r = Q[c0:c0 + (1 << 23), 1].long()
        ln = r - l + 1
        is_long = ln > short_max
        if edges:  # inside one sub-block: scanned whatever short_max is
            base = l // bs * bs
            is_long &= ~((r // bs * bs == base) & ((l - base) // 32 == (r - base) // 32))
        short_b += float((8 + 4 * ln[~is_long] + out_bytes).sum())
        n_short += int((~is_long).sum())
        short_lines += int((r[~is_long] // 32 - l[~is_long] // 32 + 1).sum())
        l, r, ln = l[is_long], r[is_long], ln[is_long]
        n_long += int(l.numel())
        lB, rB = l // bs, r // bs
        same = lB == rB
        if edges:
            ls, rs = l[same], r[same]
            gl, gr = ls // 32, rs // 32  # the sub-block grid restarts at block starts: bs % 32 == 0
            sup_i = sub.range_min(gl, gr - gl + 1)[1]
            slow = ~((sup_i >= ls) & (sup_i <= rs))
            long_b += float((8 + 16 + slow.long() * (16 + 16 * (gr - gl > 1).long()) + out_bytes).sum())
            gathers += 2 * int(slow.sum())
            reads += 2 * int(ls.numel()) + 2 * int((slow & (gr - gl > 1)).sum())
        else:
            ls, rs = l[same], r[same]
            if sub is not None and bs % 32 == 0:
                # the one-lane test over the enclosing sub-blocks: two entries,
                # and the scan only when their leftmost minimum lies outside
                gl, gr = ls // 32, rs // 32
                sup_i = sub.range_min(gl, gr - gl + 1)[1]
                hit = (sup_i >= ls) & (sup_i <= rs)
                reads += 2 * int(ls.numel())
                long_b += float((8 + 16 + out_bytes) * int(hit.sum()))
                ls, rs = ls[~hit], rs[~hit]
                long_b += float(16 * int(ls.numel()))
            long_b += float((8 + scan_bytes(ls, rs, bs) + out_bytes).sum())
        x = ~same
        l, r, lB, rB = l[x], r[x], lB[x], rB[x]
        if edges and mid_shift is not None:
            cl, cr = lB >> mid_shift, rB >> mid_shift
            near = cr - cl < 4  # at most 4 coarse blocks: one or two entries of the LDS row
            a0 = cl << mid_shift
            hi = torch.where(cr - cl < 2, cl + 1, cr)
            end = torch.clamp((hi + 1) << mid_shift, max=lvl.nb)
            ci = lvl.range_min(a0[near], (end - a0)[near])[1]
            hit = torch.zeros_like(near)
            hit[near] = (ci >= l[near]) & (ci <= r[near])
            long_b += float(8 * int(hit.sum()) + out_bytes * int(hit.sum()))
            parts += 2 * int(hit.sum())
            x = ~hit
            l, r, lB, rB = l[x], r[x], lB[x], rB[x]
        _, sup_i, sr = lvl.superset(lB, rB, edges and lines, retry=edges and lines and two_stage)
        fast = (sup_i >= l) & (sup_i <= r)
        long_b += float((8 + 8 * sr[fast] + out_bytes).sum())
        parts += 2 * int(fast.sum())
        reads += int(sr.sum())
        x = ~fast
        l, r, lB, rB, sr, sup_i = l[x], r[x], lB[x], rB[x], sr[x], sup_i[x]
        if edges and two_stage:
            # answer_two_stage's phases 2 and 3 (phase 1 is the superset test
            # above): the blocks of the side that did not fail, from the line
            # just read when the two entries shared it, with the failing side's
            # edge pair; failing again, both edge pairs and the blocks between
            c = rB - lB + 1
            lev = (torch.frexp(c.double())[1] - 1).long()
            a2 = rB - torch.bitwise_left_shift(torch.ones_like(lev), lev) + 1
            base = lev * lvl.nb
            inl = (a2 == lB) | (((base + lB) * 8) // 128 == ((base + a2) * 8) // 128)
            left = sup_i < l
            q1 = torch.where(left, lB + 1, lB)
            q2 = torch.where(left, a2, a2 - 1)
            adj = inl & (a2 == lB) & torch.where(left, rB + 1 < lvl.nb, lB > 0)
            q1 = torch.where(adj, torch.where(left, lB + 1, lB - 1), q1)
            q2 = torch.where(adj, q1, q2)
            pair = inl & (a2 > lB)
            gv, gi = lvl.range_min(q1, c - 1)  # the general window over c-1 blocks
            # entries of the line (only lanes in pair | adj use them; the others' indexes are clamped)
            q1c, q2c = q1.clamp(0, lvl.nb - 1), q2.clamp(0, lvl.nb - 1)
            uv, ui = lvl.V[lev, q1c], lvl.I[lev, q1c]
            vv, vi = lvl.V[lev, q2c], lvl.I[lev, q2c]
            t = vv < uv
            s2i = torch.where(pair | adj, torch.where(t, vi, ui), gi)
            ok2 = torch.where(left, s2i <= r, s2i >= l)
            n2, n3 = int(ok2.sum()), int((~ok2).sum())
            mid = ~ok2 & (c > 2)
            reads += 2 * (n2 + n3) + 2 * int(mid.sum())
            gathers += n2 + 2 * n3
            needed += n2 + 2 * n3
            parts += 2 * (n2 + n3)
            long_b += float((8 + 8 * sr + 16 + 8 + 8 * (~ok2).long() + 16 * mid.long() + out_bytes).sum())
            continue
        s2b = torch.zeros_like(l)
        mv = torch.full_like(l, 0, dtype=torch.float32).fill_(float("inf"))
        mi = torch.full_like(l, 2 ** 31 - 1)
        mid = rB - lB > 1
        if bool(mid.any()):
            mv[mid], mi[mid] = lvl.range_min(lB[mid] + 1, rB[mid] - lB[mid] - 1)
        ev, ei = lvl.V[0, lB], lvl.I[0, lB]
        exact_l = ei >= l
        need_l = ~exact_l & (ev <= mv)
        cv = torch.where(exact_l & ((ev < mv) | ((ev == mv) & (ei < mi))), ev, mv)
        if edges and bool(need_l.any()):
            cv[need_l] = torch.minimum(cv[need_l], left_part_min(A, sub, bs, l[need_l]))
        fv, fi = lvl.V[0, rB], lvl.I[0, rB]
        need_r = (fi > r) & (fv < cv)
        parts += 2 * int(l.numel())
        reads += 2 * int(mid.sum()) + 2 * int(l.numel())
        n_need = int(need_l.sum() + need_r.sum())
        needed += n_need
        if edges:
            edge_b = 8 * need_l.long() + 8 * need_r.long()
            gathers += n_need
        else:
            edge_b = (need_l.long() * scan_bytes(l, (lB + 1) * bs - 1, bs)
                      + need_r.long() * scan_bytes(rB * bs, r, bs))
        long_b += float((8 + 8 * sr + s2b + 16 * mid.long() + 16 + edge_b + out_bytes).sum())
    return {"long": long_b, "short": short_b, "n_long": n_long, "n_short": n_short,
            "edge_parts_computed": needed / parts if parts else 0.0, "gathers": gathers, "table_reads": reads,
            "short_lines": short_lines}


XCDS = 8  # accelerator complex dies of an MI355X (GRBM_GUI_ACTIVE counts each)
TRAFFIC_PATH = os.path.join(ROOT, "profiles", "traffic.json")


def literal_bytes(Q, bs: int, out_bytes: int = 8, longer_than: int = 0) -> int:
    """Algorithmic bytes of a batch under the block method read literally,
    per RMQ: 8 (the query) + 4*S + T + out_bytes, where S is the range's
    length when it lies inside one block, else both edge-block parts
    scanned element by element, (bs - l mod bs) + (r mod bs + 1), and T is
    16 (two 8-byte table entries) when whole blocks lie between -- what the
    reference's three traces per query amount to (src/rtx_kernels.cu:85-180).
    Designs that skip edge parts move fewer bytes than this figure.
    longer_than: count only ranges of more elements than this."""
    import torch
    total = 0
    for c in range(0, Q.shape[0], 1 << 24):
        l = Q[c:c + (1 << 24), 0].to(torch.int64)
        r = Q[c:c + (1 << 24), 1].to(torch.int64)
        if longer_than:
            keep = r - l + 1 > longer_than
            l, r = l[keep], r[keep]
        lb, rb = l // bs, r // bs
        S = torch.where(lb == rb, r - l + 1, (bs - l % bs) + (r % bs + 1))
        T = torch.where(rb - lb > 1, 16, 0)
        total += int((8 + 4 * S + T + out_bytes).sum().item())
    return total


def load_traffic_record(kernel: str, workload: str):
    """The whole counter record of `kernel` for exactly this workload, or None."""
    if not os.path.exists(TRAFFIC_PATH):
        return None
    try:
        with open(TRAFFIC_PATH) as f:
            return json.load(f).get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None


def attach_measured(run: dict, workload: str, q: int) -> None:
    """Per batch, the HBM bytes and DRAM read lines per RMQ its kernels moved
    (calibrated FETCH_SIZE / WRITE_SIZE of this exact workload, when
    measured): a figure that stays comparable across designs, unlike either
    byte model.  Only for batches whose kernels no other batch launches:
    the counters of a kernel shared by several batches mix them."""
    users = {}
    for d in run["per_dist"].values():
        for sym in d["kernels"].values():
            users[sym] = users.get(sym, 0) + 1
    for name, d in run["per_dist"].items():
        if any(users[sym] > 1 for sym in d["kernels"].values()):
            continue
        recs = [load_traffic_record(sym, workload) for sym in d["kernels"].values()]
        if not recs or any(r is None for r in recs):
            continue
        fetch = sum(r["fetch_bytes_per_launch"] for r in recs)
        write = sum(r["write_bytes_per_launch"] for r in recs)
        d["measured_bytes_per_rmq"] = (fetch + write) / q
        d["dram_read_lines_per_rmq"] = fetch / 128.0 / q
        d["measured_GBps"] = (fetch + write) / (d["ms_per_pass"] / 1e3) / 1e9


def workload_key(n: int, q: int, bs_for: dict, opts_for: dict, edges_for: dict | None) -> str:
    """Names a timed region exactly: HBM traffic measured for one configuration
    is never reported for another."""
    parts = []
    for lr in DISTS:
        o = opts_for[lr]
        fl = int(edges_for[lr]) if edges_for else 0
        tag = ("+edges" if fl & 1 else "") + ("+mid" if fl & 2 else "")
        parts.append(f"{lr}:bs={bs_for[lr]}{tag}/s{o.short_max}/g{o.group_lanes}/l{o.long_lanes}")
    return f"n={n} q={q} value+index " + " ".join(parts)


def load_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, calibrated)
    measured for exactly this workload, else None."""
    if not os.path.exists(TRAFFIC_PATH):
        return None
    try:
        with open(TRAFFIC_PATH) as f:
            rec = json.load(f).get(workload, {}).get(kernel)
        return None if rec is None else rec["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def _counter_rows(d: str, counter: str) -> list[tuple[str, float]]:
    """(kernel name, value) per dispatch of one counter from a rocprofv3 --pmc output tree."""
    import csv
    import glob
    rows = []
    for path in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return [(k, v) for _, k, v in rows]


def reduce_pmc(d: str, copy_bytes: int) -> int:
    """Turns four rocprofv3 --pmc runs under `d` into profiles/traffic.json.

    d/fetch, d/write: FETCH_SIZE and WRITE_SIZE of a bench run whose JSON line
    is d/bench.json; d/copy_fetch, d/copy_write: the same counters for
    `bench.py --copy-probe copy_bytes` (each copy reads and writes copy_bytes).
    Each counter is scaled by (bytes the copy moves) / (what the counter read
    for it), then averaged over every dispatch of each query kernel."""
    with open(os.path.join(d, "bench.json")) as f:
        line = json.loads(f.read().strip().splitlines()[-1])
    cal = {}
    for c, sub in (("FETCH_SIZE", "copy_fetch"), ("WRITE_SIZE", "copy_write")):
        rows = _counter_rows(os.path.join(d, sub), c)
        kib = max(v for _, v in rows)  # the copy is the largest dispatch of the probe
        cal[c] = copy_bytes / (kib * 1024.0)
    per = {}
    for c, sub in (("FETCH_SIZE", "fetch"), ("WRITE_SIZE", "write")):
        for k, v in _counter_rows(os.path.join(d, sub), c):
            per.setdefault(k, {}).setdefault(c, []).append(v * 1024.0 * cal[c])
    # optional passes: L2 (d/l2), shader sequencer (d/sq), L1 (d/tcp)
    raw = {}
    for sub, names in (("l2", ("TCC_HIT_sum", "TCC_MISS_sum")),
                       ("sq", ("VALUBusy", "SQ_WAIT_INST_ANY", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES")),
                       ("tcp", ("TCP_TOTAL_CACHE_ACCESSES_sum", "TCP_TCC_READ_REQ_sum", "TA_BUSY_avr",
                                "GRBM_GUI_ACTIVE"))):
        if os.path.isdir(os.path.join(d, sub)):
            for c in names:
                for k, v in _counter_rows(os.path.join(d, sub), c):
                    raw.setdefault(k, {}).setdefault(c, []).append(v)

    def find(sym: str):
        # "k_query_tiles<4, false>" matches the demangled "rmqx::k_query_tiles<4, false>(...)"
        hits = [k for k in per if f"::{sym}(" in k or k.startswith(f"{sym}(")]
        return per[hits[0]] if len(hits) == 1 else None

    sections = [(line["roofline"]["workload"], line["per_kernel"])]
    sections += [(v["roofline"]["workload"], v["per_kernel"]) for k, v in line.items() if k.startswith("fixed_bs_")]
    # only the workloads this bench line ran, and of each only the kernels it launches now: figures
    # for older configurations and for kernels since renamed or replaced are dropped
    current = {}
    for wk, kernels in sections:
        current.setdefault(wk, set()).update(kernels)
    out = {}
    if os.path.exists(TRAFFIC_PATH):
        with open(TRAFFIC_PATH) as f:
            out = {w: {k: r for k, r in v.items() if k in current[w]}
                   for w, v in json.load(f).items() if w in current}
    seen = [sym for _, kernels in sections for sym in kernels]
    for workload, kernels in sections:
        for sym, info in kernels.items():
            if seen.count(sym) > 1:  # launched by both runs: its dispatches cannot be told apart
                log(f"traffic: {sym} runs in more than one section; not attributed")
                continue
            rec = find(sym)
            if not rec or "FETCH_SIZE" not in rec or "WRITE_SIZE" not in rec:
                continue
            fb, wb = float(np.mean(rec["FETCH_SIZE"])), float(np.mean(rec["WRITE_SIZE"]))
            hits = [k for k in raw if f"::{sym}(" in k or k.startswith(f"{sym}(")]
            cnt = {c: float(np.sum(v)) for c, v in raw[hits[0]].items()} if len(hits) == 1 else {}
            extra = {}
            if "TCC_HIT_sum" in cnt and cnt["TCC_HIT_sum"] + cnt.get("TCC_MISS_sum", 0) > 0:
                extra["l2_hit_pct"] = 100 * cnt["TCC_HIT_sum"] / (cnt["TCC_HIT_sum"] + cnt["TCC_MISS_sum"])
            if "VALUBusy" in cnt:
                extra["valu_busy_pct"] = float(np.mean(raw[hits[0]]["VALUBusy"]))
            if cnt.get("SQ_WAVE_CYCLES"):
                extra["wave_cycles_waiting_pct"] = 100 * cnt.get("SQ_WAIT_INST_ANY", 0) / cnt["SQ_WAVE_CYCLES"]
                if "SQ_WAIT_ANY" in cnt:  # s_waitcnt: waiting for memory
                    extra["wave_cycles_waitcnt_pct"] = 100 * cnt["SQ_WAIT_ANY"] / cnt["SQ_WAVE_CYCLES"]
            if cnt.get("TCP_TOTAL_CACHE_ACCESSES_sum"):
                extra["l1_hit_pct"] = 100 * (1 - cnt.get("TCP_TCC_READ_REQ_sum", 0) / cnt["TCP_TOTAL_CACHE_ACCESSES_sum"])
            if cnt.get("GRBM_GUI_ACTIVE") and "TA_BUSY_avr" in cnt:
                # GRBM_GUI_ACTIVE sums the GPU's XCDs; TA_BUSY_avr is one TA's busy cycles
                extra["ta_busy_pct"] = 100 * cnt["TA_BUSY_avr"] / (cnt["GRBM_GUI_ACTIVE"] / XCDS)
            out.setdefault(workload, {})[sym] = {
                **extra,
                "bytes_per_launch": fb + wb, "fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb,
                "dispatches": len(rec["FETCH_SIZE"]),
                "algorithmic_bytes_per_launch": info["algorithmic_bytes_per_launch"],
                "traffic_over_algorithmic": (fb + wb) / info["algorithmic_bytes_per_launch"],
                "model_bytes_per_launch": info.get("model_bytes_per_launch"),
                "traffic_over_model": ((fb + wb) / info["model_bytes_per_launch"]
                                       if info.get("model_bytes_per_launch") else None),
                "queries_per_launch": info.get("queries_per_launch"),
                "calibration": {"FETCH_SIZE": cal["FETCH_SIZE"], "WRITE_SIZE": cal["WRITE_SIZE"],
                                "copy_bytes": copy_bytes}}
            log(f"traffic: {sym}: {(fb + wb) / 1e6:.1f} MB/launch over {len(rec['FETCH_SIZE'])} dispatches, "
                f"{(fb + wb) / info['algorithmic_bytes_per_launch']:.2f}x the literal bytes, "
                f"{(fb + wb) / info['model_bytes_per_launch']:.2f}x the model's")
    with open(TRAFFIC_PATH, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    return 0


def host_info() -> dict:
    """CPU model, the cores this process may run on, and SMT, of the host."""
    model, siblings, cores = "unknown", None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "siblings" and siblings is None:
                    siblings = int(v)
                elif k == "cpu cores" and cores is None:
                    cores = int(v)
    except OSError:
        pass
    smt = None
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        if siblings and cores:
            smt = siblings > cores
    return {"cpu_model": model, "affinity_cores": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count(),
            "smt_active": smt, "threads_per_core": (siblings // cores) if siblings and cores else None}


def compare_hrmq(A_host: np.ndarray, val_g, idx_g, val_h, idx_h) -> dict:
    """GPU answers (leftmost index) against HRMQ's (rightmost index) on the
    same queries: values bit for bit; indices equal where the minimum is
    unique, and where they differ both must hold the minimum with HRMQ's
    further right (a tie)."""
    vb = val_g.view(np.int32) != val_h.view(np.int32)
    differ = idx_g != idx_h
    tie_ok = differ & (idx_h > idx_g) & (A_host.view(np.int32)[idx_h] == val_g.view(np.int32))
    return {"compared": int(val_g.size), "value_mismatches": int(vb.sum()),
            "index_ties": int(tie_ok.sum()), "index_mismatches": int((differ & ~tie_ok).sum())}


def cpu_baseline(A_host: np.ndarray, queries: dict, target_s: float, gpu_answers: dict,
                 plain_small: dict | None) -> dict | None:
    """The reference's HRMQ on this host's cores over a bounded sample of each
    batch, checked against the GPU's answers to the same queries; this
    build's own HRMQ (algs 1/101 of its CLI) on the same sample beside it;
    and CPU BASE (algs 0/100) at the small configuration."""
    try:
        from oracle import oracle as O
    except Exception as e:  # noqa: BLE001
        log(f"cpu_baseline unavailable: {e}")
        return None
    import rtxrmq_amd as R
    info = host_info()
    nt = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["affinity_cores"]
    if not O.HRMQ.available():
        log("cpu_baseline: oracle/_ref/libhrmq.so missing; skipping")
        return None
    log(f"cpu_baseline: {info['cpu_model']}, {info['affinity_cores']} cores in affinity, SMT {info['smt_active']}; "
        f"building reference HRMQ over n={A_host.size} (single thread)")
    t0 = time.perf_counter()
    h = O.HRMQ(A_host)
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    port = R.HostHRMQ(A_host)
    port_build_s = time.perf_counter() - t0
    log(f"cpu_baseline: reference built in {build_s:.2f} s, the engine's HRMQ in {port_build_s:.2f} s; "
        f"timing queries on {nt} threads")
    total_q, total_t, port_t, per = 0, 0.0, 0.0, {}
    for lr, Q in queries.items():
        pilot = Q[: 1 << 16]
        t0 = time.perf_counter()
        h.query(pilot, nt=nt, want_val=True, want_idx=True)
        rate = pilot.shape[0] / max(time.perf_counter() - t0, 1e-6)
        m = int(min(Q.shape[0], max(1 << 16, rate * target_s)))
        sample = Q[:m]
        t0 = time.perf_counter()
        hv, hi = h.query(sample, nt=nt, want_val=True, want_idx=True)
        dt = time.perf_counter() - t0
        t0 = time.perf_counter()
        pv, pi = port.query(sample, nthreads=nt)
        pdt = time.perf_counter() - t0
        gv, gi = gpu_answers[lr]
        cmp = compare_hrmq(A_host, gv[:m], gi[:m], hv, hi)
        port_bad = int(((pi != hi) | (pv.view(np.int32) != hv.view(np.int32))).sum())
        per[DIST_NAMES[lr]] = {"rmq_per_s": m / dt, "queries": m, "port_rmq_per_s": m / pdt,
                               "gpu_vs_hrmq": cmp, "port_vs_reference_mismatches": port_bad}
        log(f"cpu_baseline: {DIST_NAMES[lr]} reference {m / dt:.3e} RMQ/s, the engine's {m / pdt:.3e} RMQ/s over "
            f"{m} queries; GPU vs HRMQ: {cmp}; port vs reference mismatches {port_bad}")
        if cmp["value_mismatches"] or cmp["index_mismatches"] or port_bad:
            raise SystemExit(f"cpu_baseline: answers differ on {DIST_NAMES[lr]}")
        total_q += m
        total_t += dt
        port_t += pdt
    size, port_size = h.size_bytes(), port.bytes()
    h.close()
    port.close()
    small = None
    if plain_small:
        # CPU BASE (brute force from A[l], strict '<'), bounded sample; HRMQ on every query
        A1, Q1 = plain_small["A"], plain_small["Q"]
        pilot = Q1[:64]
        t0 = time.perf_counter()
        O.rmq_brute(A1, pilot, nt=nt)
        rate = pilot.shape[0] / max(time.perf_counter() - t0, 1e-6)
        m = int(min(Q1.shape[0], max(256, rate * target_s)))
        t0 = time.perf_counter()
        bv, bi = O.rmq_brute(A1, Q1[:m], nt=nt)
        base_dt = time.perf_counter() - t0
        g_v, g_i = plain_small["val"], plain_small["idx"]
        base_bad = int(((bi != g_i[:m]) | (bv.view(np.int32) != g_v[:m].view(np.int32))).sum())
        h1 = O.HRMQ(A1)
        t0 = time.perf_counter()
        hv, hi = h1.query(Q1, nt=nt, want_val=True, want_idx=True)
        h1_dt = time.perf_counter() - t0
        h1.close()
        cmp = compare_hrmq(A1, g_v, g_i, hv, hi)
        small = {"n": int(A1.size), "q": int(Q1.shape[0]), "lr": -1,
                 "gpu_rmq_per_s": plain_small["rmq_per_s"], "gpu_bs": plain_small["bs"],
                 "cpu_base_rmq_per_s": m / base_dt, "cpu_base_queries": m, "cpu_base_mismatches": base_bad,
                 "hrmq_rmq_per_s": Q1.shape[0] / h1_dt, "gpu_vs_hrmq": cmp}
        log(f"cpu_baseline: plain n=2^20 q=2^20: GPU {plain_small['rmq_per_s']:.3e} RMQ/s, CPU BASE {m / base_dt:.3e} RMQ/s "
            f"({m} queries, {base_bad} mismatches), HRMQ {Q1.shape[0] / h1_dt:.3e} RMQ/s, GPU vs HRMQ {cmp}")
        if base_bad or cmp["value_mismatches"] or cmp["index_mismatches"]:
            raise SystemExit("cpu_baseline: plain n=2^20 q=2^20 answers differ")
    return {
        "value": total_q / total_t,
        "unit": "RMQ/s",
        "cores": nt,
        "kind": "reference",
        "sample": (f"reference HRMQ (RMQRMM64, alg 1/101 value+index) on the first {total_q} queries of the same "
                   f"three batches (about {target_s:.0f}s each on {nt} threads); construction {build_s:.2f}s "
                   f"single-threaded, {size / 1e6:.1f} MB; answers checked against the GPU's"),
        "host": info,
        "port_rmq_per_s": total_q / port_t,
        "port_over_reference": total_t / port_t,
        "port_construction_s": port_build_s,
        "port_bytes": port_size,
        "per_dist": per,
        "plain_small": small,
    }


def timed_run(R, torch, index_for, opts_for, Q_dev, out_v, out_i, steps: int, warmup: int, label: str):
    """Warmup, then `steps` timed steps; a step is one query pass per batch."""
    def one_step():
        res = {}
        for lr in DISTS:
            _, _, t = index_for[lr].query_timed(Q_dev[lr], out_val=out_v, out_idx=out_i, opts=opts_for[lr])
            res[lr] = t
        return res

    for w in range(warmup):
        one_step()
        log(f"bench[{label}]: warmup {w + 1}/{warmup} done")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    steps_out = []
    for k in range(steps):
        steps_out.append(one_step())
        log(f"bench[{label}]: step {k + 1}/{steps}: " + ", ".join(
            f"{DIST_NAMES[lr]} {s['total_ms']:.2f} ms" for lr, s in steps_out[-1].items()))
    torch.cuda.synchronize()
    return steps_out, time.perf_counter() - t_start


def answers_digest(torch, index_for, opts_for, Q_dev, out_v, out_i) -> dict:
    """Per batch, (sum of the indexes, sum of the value bits as int32) of one
    more pass, computed on the device: equal digests of two configurations
    over the same queries say their answers agree."""
    out = {}
    for lr in DISTS:
        index_for[lr].query(Q_dev[lr], out_val=out_v, out_idx=out_i, opts=opts_for[lr])
        out[lr] = (int(out_i.to(torch.int64).sum()), int(out_v.view(torch.int32).to(torch.int64).sum()))
    return out


def summarise(A, Q_dev, q, steps_out, elapsed, levels, bs_for, opts_for, edges_for=None, kernels_for=None) -> dict:
    """RMQ/s, per-distribution figures and the roofline of the dominant kernel.
    levels: BlockLevel per block size (and 32, with edge tables); kernels_for:
    the kernel names each batch's handle launches (RMQIndex.kernels)."""
    per_dist = {}
    kernels = {}  # symbol -> [algorithmic bytes, ms, launches, edge-pair gathers, table reads, short-range lines]
    for lr in DISTS:
        e = bool(edges_for and edges_for[lr])
        if e and bs_for[lr] > 32 and bs_for[lr] % 32:
            raise ValueError("the byte model needs block sizes of at most 32 or multiples of 32 with edge tables")
        kname = kernels_for[lr][0] if kernels_for else ""
        lines = bool(e and not kname.startswith("k_query_edges_lds"))
        two = bool(e and kname.startswith(("k_query_edges<", "k_query_edges_mid<")) and kname.endswith(", true>"))
        mid = mid_shift_of(levels[bs_for[lr]].nb) if e and kname.startswith("k_query_edges_mid<") else None
        ab = algorithmic_bytes(Q_dev[lr], levels[bs_for[lr]], opts_for[lr].short_max, edges=e,
                               sub=levels.get(32), A=A, lines=lines, two_stage=two, mid_shift=mid)
        lb, sb, nl = ab["long"], ab["short"], ab["n_long"]
        lit = literal_bytes(Q_dev[lr], bs_for[lr])
        ms = [s[lr]["total_ms"] for s in steps_out]
        lms = [s[lr]["long_ms"] for s in steps_out]
        sms = [s[lr]["short_ms"] for s in steps_out]
        per_dist[DIST_NAMES[lr]] = {
            "rmq_per_s": q / (np.mean(ms) / 1e3),
            "ms_per_pass": float(np.mean(ms)),
            "bs": bs_for[lr],
            "opts": {"short_max": opts_for[lr].short_max, "group_lanes": opts_for[lr].group_lanes,
                     "long_lanes": opts_for[lr].long_lanes},
            "long_kernel_ms": float(np.mean(lms)),
            "short_kernel_ms": float(np.mean(sms)),
            "n_long": nl,
            "edge_parts_computed": ab["edge_parts_computed"],
            "algorithmic_GBps": (lb + sb) / (np.mean(ms) / 1e3) / 1e9,
            "frac_of_roofline": (lb + sb) / (np.mean(ms) / 1e3) / 1e9 / HBM_PEAK_GBS,
            "model_bytes_per_rmq": (lb + sb) / q,
            "literal_bytes_per_rmq": lit / q,
            "literal_frac_of_roofline": lit / (np.mean(ms) / 1e3) / 1e9 / HBM_PEAK_GBS,
        }
        if e:  # one pass answers every range (G lanes per scanned range)
            sym = kernels_for[lr][0] if kernels_for else f"k_query_edges<{opts_for[lr].group_lanes}>"
            per_dist[DIST_NAMES[lr]]["kernels"] = {"all": sym}
            launches = ((sym, lb + sb, lms, ab["gathers"], ab["table_reads"], ab["short_lines"], lit, q),)
        else:
            long_sym = f"k_query_tiles<{opts_for[lr].long_lanes}, true>"
            short_sym = f"k_query_tiles<{opts_for[lr].group_lanes}, false>"
            per_dist[DIST_NAMES[lr]]["kernels"] = {"long": long_sym, "short": short_sym}
            lit_long = literal_bytes(Q_dev[lr], bs_for[lr], longer_than=opts_for[lr].short_max)
            launches = ((long_sym, lb, lms, 0, ab["table_reads"], 0, lit_long, nl),
                        (short_sym, sb, sms, 0, 0, 0, lit - lit_long, q - nl))
        for sym, b, t, g, rd, ln, lit_b, nq in launches:
            acc = kernels.setdefault(sym, [0.0, 0.0, 0, 0, 0, 0, 0.0, 0])
            acc[0] += b * len(steps_out)
            acc[1] += float(np.sum(t))
            acc[2] += len(steps_out)
            acc[3] += g * len(steps_out)
            acc[4] += rd * len(steps_out)
            acc[5] += ln * len(steps_out)
            acc[6] += lit_b * len(steps_out)
            acc[7] += nq * len(steps_out)
    dom = max(kernels, key=lambda k: kernels[k][1])
    achieved = kernels[dom][6] / (kernels[dom][1] / 1e3) / 1e9
    achieved_model = kernels[dom][0] / (kernels[dom][1] / 1e3) / 1e9
    per_kernel = {k: {"ms": v[1], "launches": v[2], "avg_launch_ms": v[1] / v[2],
                      "queries_per_launch": v[7] / v[2],
                      "algorithmic_bytes_per_launch": v[6] / v[2],
                      "achieved_GBps": v[6] / max(v[1], 1e-9) * 1e3 / 1e9,
                      "model_bytes_per_launch": v[0] / v[2],
                      "model_GBps": v[0] / max(v[1], 1e-9) * 1e3 / 1e9} for k, v in kernels.items()}
    for k, v in kernels.items():
        if k.startswith("k_query_edges_lds"):
            per_kernel[k]["table_in_lds"] = True
        if k.startswith("k_query_edges_mid"):
            per_kernel[k]["coarse_rows_in_lds"] = True
        if v[3] or v[4]:
            per_kernel[k]["edge_gathers_per_launch"] = v[3] / v[2]
            per_kernel[k]["table_reads_per_launch"] = v[4] / v[2]
            per_kernel[k]["edge_gathers_per_s"] = v[3] / (v[1] / 1e3)
            per_kernel[k]["table_reads_per_s"] = v[4] / (v[1] / 1e3)
            per_kernel[k]["short_lines_per_launch"] = v[5] / v[2]
    return {"value": 3 * q * len(steps_out) / elapsed, "ms_per_step": elapsed / len(steps_out) * 1e3,
            "per_dist": per_dist, "dominant": dom, "achieved": achieved, "achieved_model": achieved_model,
            "per_kernel": per_kernel}


def run_batched(R, torch, dev, args) -> dict:
    """Config 4: m arrays with n_i log-uniform in [2^10, 2^22], q_i = n_i queries
    per array (1:1:1 uniform / lognormal 2^14-scaled / lognormal 2^6-scaled
    lengths), int32 leftmost indices; one persistent plan launch (HIP graph)
    against the per-array launch loop on the same handles and buffers."""
    import ctypes
    from rtxrmq_amd import _native as N

    rng = np.random.default_rng(args.seed)
    m = args.batch_arrays
    ns = np.floor(np.exp2(rng.uniform(10, 22, m))).astype(np.int64)
    a_off = np.zeros(m + 1, np.int64)
    a_off[1:] = np.cumsum((ns + 63) // 64 * 64)  # 256-byte aligned array starts
    q_off = np.zeros(m + 1, np.int64)
    q_off[1:] = np.cumsum(ns)
    total_q = int(q_off[-1])
    log(f"batched: {m} arrays, {int(ns.sum())} elements, {total_q} queries")
    A_all = torch.empty(int(a_off[-1]), dtype=torch.float32, device=dev)
    N.check(N.lib().rmq_gen_array(A_all.data_ptr(), A_all.numel(), args.seed, 0, None), "rmq_gen_array")
    Q_all = torch.empty((total_q, 2), dtype=torch.int32, device=dev)
    out_plan = torch.empty(total_q, dtype=torch.int32, device=dev)
    out_loop = torch.empty(total_q, dtype=torch.int32, device=dev)
    A_views, Q_views = [], []
    for i in range(m):
        n_i = int(ns[i])
        A_views.append(A_all[a_off[i]:a_off[i] + n_i])
        Qi = Q_all[q_off[i]:q_off[i + 1]]
        Q_views.append(Qi)
        parts = np.array_split(np.arange(n_i), 3)
        for lr, part in zip((-1, -7, -8), parts):
            if len(part):
                sub = Qi[int(part[0]):int(part[-1]) + 1]
                N.check(N.lib().rmq_gen_queries(sub.data_ptr(), len(part), n_i, lr, args.seed * 7919 + i, None),
                        "rmq_gen_queries")
    torch.cuda.synchronize()
    big = int(np.argmax(ns))
    # the plan's kernel reads no coarse rows: tuned among plain and edge-table handles
    pick = R.tune(A_views[big], Q_views[big], bs_list=[64, 128, 256, 512, 1024, 4096],
                  cache_path=args.tune_cache or None, refresh=args.retune, want_value=False,
                  flags_allowed=R.BUILD_EDGES)
    opts, bs, flags = pick["opts"], pick["bs"], pick["flags"]
    log(f"batched: tuned on the largest array (n={int(ns[big])}): bs={bs} edges={bool(flags)} "
        f"short_max={opts.short_max} group_lanes={opts.group_lanes} long_lanes={opts.long_lanes}")
    t0 = time.perf_counter()
    indexes = [R.RMQIndex(A, bs, flags=flags) for A in A_views]
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    outs_p = [out_plan[q_off[i]:q_off[i + 1]] for i in range(m)]
    outs_l = [out_loop[q_off[i]:q_off[i + 1]] for i in range(m)]
    t0 = time.perf_counter()
    plan = R.Plan(indexes, Q_views, out_idx=outs_p, opts=opts)
    plan_s = time.perf_counter() - t0
    P = ctypes.c_void_p * m
    hs = P(*[ix.handle for ix in indexes])
    qp = P(*[Q.data_ptr() for Q in Q_views])
    qs = (ctypes.c_int64 * m)(*[int(x) for x in ns])
    ip = P(*[t.data_ptr() for t in outs_l])

    def loop():
        N.check(N.lib().rmq_query_loop(hs, m, qp, qs, None, ip, ctypes.byref(opts), None), "rmq_query_loop")

    def timed(fn, steps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    t_graph = timed(plan.graph_launch, args.batch_steps)
    t_eager = timed(plan.launch, args.batch_steps)
    t_loop = timed(loop, args.batch_steps)
    identical = bool(torch.equal(out_plan, out_loop))
    log(f"batched: graph {total_q / t_graph:.4e} RMQ/s ({t_graph * 1e3:.1f} ms), eager plan "
        f"{total_q / t_eager:.4e}, per-array loop {total_q / t_loop:.4e} ({t_loop * 1e3:.1f} ms); "
        f"identical={identical}")
    info = plan.info()
    plan.close()
    for ix in indexes:
        ix.close()
    per_class = run_batched_classes(R, torch, args, ns, A_views, Q_views, q_off, out_loop, timed)
    return {"arrays": m, "elements": int(ns.sum()), "queries": total_q, "bs": bs, "edge_tables": bool(flags),
            "per_class": per_class,
            "opts": {"short_max": opts.short_max, "group_lanes": opts.group_lanes, "long_lanes": opts.long_lanes},
            "work_items": info["items"], "rmq_per_s_graph": total_q / t_graph, "rmq_per_s_plan": total_q / t_eager,
            "rmq_per_s_loop": total_q / t_loop, "speedup_vs_loop": t_loop / t_graph,
            "ms_graph": t_graph * 1e3, "ms_loop": t_loop * 1e3, "build_s": build_s, "plan_create_s": plan_s,
            "steps": args.batch_steps, "identical_to_loop": identical}


def run_batched_classes(R, torch, args, ns, A_views, Q_views, q_off, out_ref, timed) -> dict:
    """The batched workload with picks per size class: arrays of the same
    floor(log2 n) share the block size and edge tables the tuner picks on the
    class's largest array (index outputs, cached per key); the plan's
    thresholds and lanes are those of the class holding the most queries.
    Timed as one graph launch and as the per-array loop on the same handles;
    answers compared with the single-pick loop's (out_ref)."""
    import ctypes
    from rtxrmq_amd import _native as N
    m = len(A_views)
    total_q = int(q_off[-1])
    cls = np.floor(np.log2(ns)).astype(np.int64)
    picks = {}
    for c in sorted(set(cls.tolist())):
        members = np.nonzero(cls == c)[0]
        rep = int(members[np.argmax(ns[members])])
        picks[c] = R.tune(A_views[rep], Q_views[rep], bs_list=[16, 32, 64, 128, 256, 512, 1024, 4096],
                          cache_path=args.tune_cache or None, refresh=args.retune, want_value=False,
                          flags_allowed=R.BUILD_EDGES)
    q_per = {c: int(ns[cls == c].sum()) for c in picks}
    top = max(q_per, key=q_per.get)
    opts = picks[top]["opts"]
    indexes = [R.RMQIndex(A_views[i], picks[int(cls[i])]["bs"], flags=picks[int(cls[i])]["flags"]) for i in range(m)]
    torch.cuda.synchronize()
    out = torch.empty_like(out_ref)
    outs = [out[q_off[i]:q_off[i + 1]] for i in range(m)]
    plan = R.Plan(indexes, Q_views, out_idx=outs, opts=opts)
    P = ctypes.c_void_p * m
    hs = P(*[ix.handle for ix in indexes])
    qp = P(*[Q.data_ptr() for Q in Q_views])
    qs = (ctypes.c_int64 * m)(*[int(x) for x in ns])
    ip = P(*[t.data_ptr() for t in outs])

    def loop():
        N.check(N.lib().rmq_query_loop(hs, m, qp, qs, None, ip, ctypes.byref(opts), None), "rmq_query_loop")

    t_graph = timed(plan.graph_launch, args.batch_steps)
    identical_graph = bool(torch.equal(out, out_ref))
    t_loop = timed(loop, args.batch_steps)
    identical_loop = bool(torch.equal(out, out_ref))
    plan.close()
    for ix in indexes:
        ix.close()
    del out
    classes = {f"2^{c}": {"arrays": int((cls == c).sum()), "queries": q_per[c], "bs": picks[c]["bs"],
                          "edge_tables": bool(picks[c]["flags"]), "short_max": picks[c]["opts"].short_max,
                          "group_lanes": picks[c]["opts"].group_lanes, "long_lanes": picks[c]["opts"].long_lanes,
                          "source": "cache" if picks[c]["from_cache"] else "timed"} for c in picks}
    log(f"batched per class: graph {total_q / t_graph:.4e} RMQ/s ({t_graph * 1e3:.1f} ms), loop "
        f"{total_q / t_loop:.4e} ({t_loop * 1e3:.1f} ms); plan opts from class 2^{top}; "
        f"identical={identical_graph and identical_loop}; picks " +
        ", ".join(f"{k}:bs={v['bs']}{'+e' if v['edge_tables'] else ''}" for k, v in classes.items()))
    if not (identical_graph and identical_loop):
        raise SystemExit("batched per class: answers differ from the single-pick loop")
    return {"rmq_per_s_graph": total_q / t_graph, "rmq_per_s_loop": total_q / t_loop,
            "speedup_vs_loop": t_loop / t_graph, "ms_graph": t_graph * 1e3, "ms_loop": t_loop * 1e3,
            "plan_opts_from_class": f"2^{top}", "classes": classes, "identical_to_single_pick": True}


def run_served(R, torch, dev, args) -> dict:
    """Config 5: resident arrays, requests of q log-uniform in [2^8, 2^22]
    queries (1:1:1 uniform / lognormal 2^14-scaled / 2^6-scaled lengths, int32
    indices, host buffers in pinned memory) through the continuous-batching
    server: saturation throughput and a Poisson stream at a fixed offered
    rate (p50/p99 latency), against the plain per-request loop (upload,
    rmq_query, download, synchronise)."""

    rng = np.random.default_rng(args.seed + 5)
    m = args.serve_arrays
    ns = np.floor(np.exp2(rng.uniform(24, 27, m))).astype(np.int64)
    log(f"served: {m} resident arrays, {int(ns.sum())} elements")
    arrays = [R.gen_array(int(n), args.seed * 31 + i, "uniform", device=dev) for i, n in enumerate(ns)]
    big = int(np.argmax(ns))
    sample = torch.cat([R.gen_queries_device(1 << 20, int(ns[big]), lr, 77, device=dev) for lr in (-1, -7, -8)])
    # the server's kernel reads no coarse rows: tuned among plain and edge-table handles
    pick = R.tune(arrays[big], sample, bs_list=[16, 32, 64, 128, 256, 1024], cache_path=args.tune_cache or None,
                  refresh=args.retune, want_value=False, flags_allowed=R.BUILD_EDGES)
    bs, opts, flags = pick["bs"], pick["opts"], pick["flags"]
    log(f"served: bs={bs} edges={bool(flags)} short_max={opts.short_max} group_lanes={opts.group_lanes} "
        f"long_lanes={opts.long_lanes}")
    # payload pool: per array, serve_payloads query batches in pinned host memory
    payloads = []
    for i, n in enumerate(ns):
        for k in range(args.serve_payloads):
            q = int(np.exp2(rng.uniform(8, 22)))
            Q = torch.empty((q, 2), dtype=torch.int32, pin_memory=True)
            parts = np.array_split(np.arange(q), 3)
            Qn = Q.numpy()
            for lr, part in zip((-1, -7, -8), parts):
                if len(part):
                    Qn[part[0]:part[-1] + 1] = R.gen_queries(len(part), int(n), lr, args.seed * 1009 + i * 97 + k)
            payloads.append((i, Q))
    pq = np.array([Q.shape[0] for _, Q in payloads])
    log(f"served: {len(payloads)} payloads, mean q {pq.mean():.0f}")
    max_q = int(pq.max())

    # plain per-request loop on the same handles
    indexes = [R.RMQIndex(A, bs, flags=flags) for A in arrays]
    plain_out = {}
    out_pin = torch.empty(max_q, dtype=torch.int32, pin_memory=True)
    order = rng.integers(0, len(payloads), args.serve_plain_requests)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nq_plain = 0
    for p_id in order:
        a, Q = payloads[p_id]
        Qd = Q.to(dev, non_blocking=True)
        _, idx = indexes[a].query(Qd, want_value=False, want_index=True, opts=opts)
        out_pin[:idx.numel()].copy_(idx, non_blocking=True)
        torch.cuda.synchronize()
        nq_plain += Q.shape[0]
        if p_id not in plain_out:
            plain_out[int(p_id)] = out_pin[:idx.numel()].clone()
    plain_s = time.perf_counter() - t0
    for ix in indexes:
        ix.close()
    plain_rate = nq_plain / plain_s
    log(f"served: plain per-request loop {plain_rate:.4e} RMQ/s over {len(order)} requests")

    sv = R.Server(slots=args.serve_slots, opts=opts)
    ids = [sv.add_array(A, bs, flags) for A in arrays]
    arena = [torch.empty(0, dtype=torch.int32)]

    def run_stream(n_req, rate_per_s, tag):
        """Submits n_req random payloads (all at once when rate is None, else at
        Poisson times), each answered into its own slice of one pinned output
        arena; the answers are compared with the plain loop's after the timed
        window, so checking costs the server nothing."""
        choice = rng.integers(0, len(payloads), n_req)
        offs = np.concatenate([[0], np.cumsum([payloads[c][1].shape[0] for c in choice])])
        if arena[0].numel() < offs[-1]:
            arena[0] = torch.empty(int(offs[-1]), dtype=torch.int32, pin_memory=True)
        out = arena[0]
        tickets = []
        sv.stats(reset=True)
        gaps = rng.exponential(1.0 / rate_per_s, n_req) if rate_per_s else np.zeros(n_req)
        t_next = time.perf_counter()
        for k in range(n_req):
            t_next += gaps[k]
            while rate_per_s and time.perf_counter() < t_next:
                pass
            a, Q = payloads[int(choice[k])]
            tickets.append(sv.submit(ids[a], Q, out_idx=out[int(offs[k]):int(offs[k + 1])]))
        failed = sum(sv.wait(tk)[0] != R.RMQ_OK for tk in tickets)
        st = sv.stats()
        mismatches = failed
        for k, p_id in enumerate(choice):
            ref = plain_out.get(int(p_id))
            if ref is not None and not torch.equal(out[int(offs[k]):int(offs[k + 1])], ref):
                mismatches += 1
        rate = st["queries_done"] / (st["wall_ms"] / 1e3)
        log(f"served[{tag}]: {st['requests_done']} requests, {rate:.4e} RMQ/s, p50 {st['p50_ms']:.2f} ms, "
            f"p99 {st['p99_ms']:.2f} ms, {st['mean_requests_per_launch']:.1f} requests/launch, "
            f"launch: enqueue {st['mean_enqueue_ms']:.2f} ms, device {st['mean_device_ms']:.2f} ms, "
            f"wait {st['mean_launch_wait_ms']:.2f} ms; mismatches {mismatches}")
        return st, rate, mismatches

    run_stream(64, None, "warmup")
    sat, sat_rate, sat_bad = run_stream(args.serve_requests, None, "saturation")
    req_rate = sat["requests_done"] / (sat["wall_ms"] / 1e3)
    offered = args.serve_load * req_rate
    poi, poi_rate, poi_bad = run_stream(args.serve_requests, offered, f"poisson {offered:.0f} req/s")
    sv.close()
    staged = run_staged(R, args, arrays, bs, flags, opts, payloads, plain_out, rng)
    # pinned host blocks go back to the runtime now, while it is whole, rather
    # than from torch's allocator during library teardown at exit
    pq_count = len(payloads)
    del payloads, arena, out_pin, plain_out
    torch.cuda.synchronize()
    torch._C._host_emptyCache()
    return {"arrays": m, "elements": int(ns.sum()), "bs": bs, "edge_tables": bool(flags), "slots": args.serve_slots,
            "request_q": "log-uniform [2^8, 2^22]", "mean_q": float(pq.mean()), "payloads": pq_count,
            "plain_rmq_per_s": plain_rate, "plain_requests": len(order),
            "saturation": {"rmq_per_s": sat_rate, "requests": sat["requests_done"], "p50_ms": sat["p50_ms"],
                           "p99_ms": sat["p99_ms"], "requests_per_launch": sat["mean_requests_per_launch"]},
            "poisson": {"offered_requests_per_s": offered, "offered_fraction_of_saturation": args.serve_load,
                        "rmq_per_s": poi_rate, "requests": poi["requests_done"], "p50_ms": poi["p50_ms"],
                        "p99_ms": poi["p99_ms"], "mean_ms": poi["mean_ms"], "max_ms": poi["max_ms"],
                        "requests_per_launch": poi["mean_requests_per_launch"]},
            "staged_under_budget": staged,
            "speedup_vs_plain": sat_rate / plain_rate, "mismatches_vs_plain": sat_bad + poi_bad + staged["mismatches"],
            "budget_bytes": sat["budget_bytes"], "peak_inflight_bytes": max(sat["peak_inflight_bytes"],
                                                                            poi["peak_inflight_bytes"])}


def run_staged(R, args, arrays, bs, flags, opts, payloads, plain_out, rng) -> dict:
    """Requests in pageable host memory, staged through device memory, under
    a fixed budget (--serve-budget) far below what all of them need at once:
    the server admits them as launches in flight free their staging buffers.
    All are submitted at once; answers compared with the plain loop's."""
    budget = args.serve_budget
    with R.Server(slots=args.serve_slots, opts=opts, mem_budget_bytes=budget) as sv:
        ids = [sv.add_array(A, bs, flags) for A in arrays]
        choice = rng.integers(0, len(payloads), args.serve_staged_requests)
        Qs = [payloads[int(c)][1].numpy().copy() for c in choice]  # pageable copies
        outs = [np.empty(Q.shape[0], np.int32) for Q in Qs]
        need = sum(12 * Q.shape[0] for Q in Qs)
        sv.stats(reset=True)
        tickets = [sv.submit(ids[payloads[int(c)][0]], Q, out_idx=o) for c, Q, o in zip(choice, Qs, outs)]
        failed = sum(sv.wait(t)[0] != R.RMQ_OK for t in tickets)
        st = sv.stats()
    mismatches = failed
    for c, o in zip(choice, outs):
        ref = plain_out.get(int(c))
        if ref is not None and not np.array_equal(o, ref.numpy()):
            mismatches += 1
    rate = st["queries_done"] / (st["wall_ms"] / 1e3)
    log(f"served[staged, budget {budget / 2**20:.0f} MiB]: {st['requests_done']} requests needing "
        f"{need / 2**20:.0f} MiB at once, peak in flight {st['peak_inflight_bytes'] / 2**20:.1f} MiB, "
        f"{rate:.4e} RMQ/s, p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms, "
        f"{st['mean_requests_per_launch']:.1f} requests/launch; mismatches {mismatches}")
    if st["peak_inflight_bytes"] > budget:
        raise SystemExit("served: staged requests exceeded the memory budget")
    return {"budget_bytes": budget, "requests": st["requests_done"], "bytes_needed_at_once": need,
            "peak_inflight_bytes": st["peak_inflight_bytes"], "rmq_per_s": rate, "p50_ms": st["p50_ms"],
            "p99_ms": st["p99_ms"], "requests_per_launch": st["mean_requests_per_launch"], "mismatches": mismatches}


def run_stress(R, torch, dev, args) -> dict:
    """Config 3: n=2^27 arrays of heavy ties (values in {0..15}) and strictly
    descending values, fixed range lengths 1, 2^8, 2^16 and n-l (every range
    ends at n-1); per workload the autotuner's pick (from the tuner cache, or
    timed on a separately seeded stream of the same workload) and its RMQ/s,
    and the literal method (no edge tables, default lanes) swept over
    bs = 2^10 .. 2^18.  Every swept configuration's answers are compared in
    full with the tuned ones."""
    n, q = 1 << 27, args.stress_q
    sample = min(q, args.tune_sample)
    steps = max(1, args.stress_steps)
    out = {}
    v_t = torch.empty(q, dtype=torch.float32, device=dev)
    i_t = torch.empty(q, dtype=torch.int32, device=dev)
    v_f = torch.empty_like(v_t)
    i_f = torch.empty_like(i_t)

    def rate(ix, Q, opts, v, i):
        ix.query(Q, opts=opts, out_val=v, out_idx=i)
        ms = [ix.query_timed(Q, opts=opts, out_val=v, out_idx=i)[2]["total_ms"] for _ in range(steps)]
        return sum(ms) / len(ms)

    for kind in ("ties16", "desc"):
        A = R.gen_array(n, 99, kind, device=dev)
        for lr in (1, 1 << 8, 1 << 16, -6):
            Q = R.gen_queries_device(q, n, lr, args.seed)
            pick = R.tune(A, R.gen_queries_device(sample, n, lr, args.seed + 7919), cache_path=args.tune_cache or None,
                          refresh=args.retune)
            with R.RMQIndex(A, pick["bs"], flags=pick["flags"]) as ix:
                ms_t = rate(ix, Q, pick["opts"], v_t, i_t)
                kern = ix.kernels(pick["opts"])
            sweep, same = {}, True
            for lg in range(10, 19):  # the literal method (no edge tables, default lanes) at bs = 2^10 .. 2^18
                with R.RMQIndex(A, 1 << lg) as ix:
                    ms_f = rate(ix, Q, None, v_f, i_f)
                sweep[f"2^{lg}"] = q / (ms_f * 1e-3)
                same &= bool(torch.equal(i_t, i_f) and torch.equal(v_t.view(torch.int32), v_f.view(torch.int32)))
            o = pick["opts"]
            name = {1: "len 1", 256: "len 2^8", 65536: "len 2^16", -6: "len n-l"}[lr]
            out.setdefault(kind, {})[name] = {
                "tuned": {"bs": pick["bs"], "edges": bool(pick["flags"]), "short_max": o.short_max,
                          "source": "cache" if pick["from_cache"] else "timed",
                          "group_lanes": o.group_lanes, "long_lanes": o.long_lanes, "kernels": kern,
                          "ms_per_pass": ms_t, "rmq_per_s": q / (ms_t * 1e-3)},
                "bs_sweep_rmq_per_s": sweep,
                "answers_identical": same}
            best = max(sweep, key=sweep.get)
            log(f"stress {kind:6s} {name:9s} tuned bs={pick['bs']} edges={bool(pick['flags'])} "
                f"s{o.short_max}/g{o.group_lanes}/l{o.long_lanes} {q / (ms_t * 1e-3):.3e} RMQ/s; "
                f"sweep bs=2^15 {sweep['2^15']:.3e}, best bs={best} {sweep[best]:.3e}; identical={same}")
            if not same:
                raise SystemExit(f"stress {kind} {name}: tuned and swept answers differ")
            del Q
        del A
        torch.cuda.empty_cache()
    return {"n": n, "q": q, "tune_sample": sample, "steps": steps, "workloads": out}


def run_plain_small(R, torch, dev, args) -> dict:
    """The small configuration: n=2^20 U[0,1), 2^20 uniform-length queries,
    the literal bs=2^15, values and indices, timed on the GPU; its arrays
    and answers go to the host checks of the CPU baseline."""
    n1, q1, bs1 = 1 << 20, 1 << 20, 1 << 15
    A1 = R.gen_array(n1, args.seed, "uniform", device=dev)
    Q1 = R.gen_queries(q1, n1, -1, args.seed)
    Qd = torch.from_numpy(Q1).to(dev)
    with R.RMQIndex(A1, bs1) as ix:
        v, i = ix.query(Qd)
        ms = [ix.query_timed(Qd, out_val=v, out_idx=i)[2]["total_ms"] for _ in range(5)]
        torch.cuda.synchronize()
        out = {"A": A1.cpu().numpy(), "Q": Q1, "val": v.cpu().numpy(), "idx": i.cpu().numpy(),
               "rmq_per_s": q1 / (float(np.mean(ms)) / 1e3), "bs": bs1}
    log(f"plain small: n={n1} q={q1} bs={bs1}: {out['rmq_per_s']:.3e} RMQ/s")
    return out


def main() -> int:
    faulthandler.enable()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100_000_000)
    ap.add_argument("--q", type=int, default=1 << 26)
    ap.add_argument("--bs", type=int, default=1 << 15, help="block size of the fixed-bs run")
    ap.add_argument("--fixed-steps", type=int, default=2, help="timed steps of the fixed-bs run")
    ap.add_argument("--tune-sample", type=int, default=1 << 26,
                    help="queries per batch the autotuner times (the whole batch by default: picks made on "
                         "2^22 queries ran 6%% slower at full size)")
    ap.add_argument("--tune-cache", default=os.path.join(ROOT, "rtxrmq_amd", "tune_cache_mi355x.txt"),
                    help="autotuner cache (picks per n and length signature); '' for none")
    ap.add_argument("--retune", action="store_true", help="time the sweep even when the cache has a pick")
    ap.add_argument("--seed", type=int, default=1321)
    ap.add_argument("--cpu-seconds", type=float, default=4.0, help="wall seconds of HRMQ per batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-arrays", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--serve-arrays", type=int, default=16)
    ap.add_argument("--serve-payloads", type=int, default=8, help="query batches per resident array")
    ap.add_argument("--serve-requests", type=int, default=1500)
    ap.add_argument("--serve-plain-requests", type=int, default=300)
    ap.add_argument("--serve-slots", type=int, default=3)
    ap.add_argument("--serve-load", type=float, default=0.7, help="offered rate / saturation rate")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--serve-budget", type=int, default=64 << 20, help="device budget of the staged served run")
    ap.add_argument("--serve-staged-requests", type=int, default=400)
    ap.add_argument("--stress-q", type=int, default=1 << 24, help="queries per workload of the stress leg")
    ap.add_argument("--stress-steps", type=int, default=2, help="timed passes per configuration of the stress leg")
    ap.add_argument("--no-stress", action="store_true")
    ap.add_argument("--only", choices=["batched", "served", "stress"], default=None,
                    help="run one section alone and print its JSON object (profiling)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many bytes (counter calibration) and exit")
    ap.add_argument("--no-gather-probe", action="store_true", help="skip the random-gather ceiling probe")
    ap.add_argument("--no-fixed-tuned", action="store_true",
                    help="skip the run at --bs with edge tables and thresholds from the tuner restricted to it")
    ap.add_argument("--gather-sweep", default=None, metavar="MIB,MIB,...",
                    help="only time rmq_probe_gather over two tables of each size (MiB, powers of two) and exit")
    ap.add_argument("--reduce-pmc", default=None, metavar="DIR",
                    help="CPU only: turn rocprofv3 FETCH_SIZE/WRITE_SIZE runs under DIR into profiles/traffic.json")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="size the --copy-probe run used")
    args = ap.parse_args()

    if args.reduce_pmc:
        return reduce_pmc(args.reduce_pmc, args.copy_bytes)

    if args.copy_probe:
        import torch
        nel = args.copy_probe // 4
        src = torch.ones(nel, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 3 copies of {nel * 4} bytes")
        return 0

    if args.gather_sweep:
        # random 8-byte gathers against the footprint they are spread over:
        # the request rate of lines from the last-level cache and from DRAM
        import rtxrmq_amd as R
        for mib in (int(x) for x in args.gather_sweep.split(",")):
            r = R.probe_gather(table_bytes=mib << 20, count=1 << 26, reps=2, device="cuda")
            print(json.dumps({"tables_mib": [mib, mib], "footprint_bytes": 2 * (mib << 20),
                              "ms_per_launch": r["ms_per_launch"], "gathers_per_s": r["gathers_per_s"]}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); queries shard with no exchange "
              f"across GPUs, so the one-GPU rate is the unit of scale.", file=sys.stderr)
        return 2

    import torch

    import rtxrmq_amd as R

    dev = torch.device("cuda:0")
    if args.only:
        res = {"batched": run_batched, "served": run_served, "stress": run_stress}[args.only](R, torch, dev, args)
        print(json.dumps({args.only: res}), flush=True)
        return 0
    n, q = args.n, args.q
    log(f"bench: n={n} q={q} dists={DISTS} steps={args.steps} warmup={args.warmup}")

    A = R.gen_array(n, args.seed, "uniform", device=dev)
    Q_host = {lr: R.gen_queries(q, n, lr, args.seed + 17 * (-lr)) for lr in DISTS}
    Q_dev = {lr: torch.from_numpy(Q).to(dev) for lr, Q in Q_host.items()}
    out_v = torch.empty(q, dtype=torch.float32, device=dev)
    out_i = torch.empty(q, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    # --- the literal configuration: bs = 2^15, default thresholds -----------------
    t0 = time.perf_counter()
    fixed_index = R.RMQIndex(A, args.bs)
    torch.cuda.synchronize()
    fixed_build_ms = (time.perf_counter() - t0) * 1e3
    fixed_opts = fixed_index.default_opts()
    fixed_steps, fixed_elapsed = timed_run(R, torch, {lr: fixed_index for lr in DISTS},
                                           {lr: fixed_opts for lr in DISTS}, Q_dev, out_v, out_i,
                                           args.fixed_steps, 1, f"bs={args.bs}")
    levels = {args.bs: BlockLevel(A, args.bs)}
    if args.bs % 32 == 0:  # sub-block minima: the plain handle's in-block superset test
        levels[32] = BlockLevel(A, 32)
    fixed = summarise(A, Q_dev, q, fixed_steps, fixed_elapsed, levels, {lr: args.bs for lr in DISTS},
                      {lr: fixed_opts for lr in DISTS})
    fixed_digest = answers_digest(torch, {lr: fixed_index for lr in DISTS}, {lr: fixed_opts for lr in DISTS}, Q_dev,
                                  out_v, out_i)
    fixed_index.close()

    # --- autotuned configuration per distribution ----------------------------------
    picks = {}
    for lr in DISTS:
        # the tuner times a stream of the same distribution with another seed,
        # so the picks are not fitted to the very queries that are timed
        sample = R.gen_queries_device(min(q, args.tune_sample), n, lr, args.seed + 7919 + 17 * (-lr), device=dev)
        t0 = time.perf_counter()
        picks[lr] = R.tune(A, sample, cache_path=args.tune_cache or None, refresh=args.retune)
        p = picks[lr]
        log(f"bench: autotune {DIST_NAMES[lr]}: bs={p['bs']} edges={bool(p['flags'])} "
            f"mid={bool(p['flags'] & R.BUILD_MID)} short_max={p['opts'].short_max} "
            f"group_lanes={p['opts'].group_lanes} long_lanes={p['opts'].long_lanes} "
            f"({'cached' if p['from_cache'] else str(p['n_tried']) + ' configs'}, {time.perf_counter() - t0:.1f} s, "
            f"key {p['key']})")
    indexes, build_ms = {}, {}
    for lr in DISTS:
        key = (picks[lr]["bs"], picks[lr]["flags"])
        if key not in indexes:
            t0 = time.perf_counter()
            indexes[key] = R.RMQIndex(A, key[0], flags=key[1])
            torch.cuda.synchronize()
            build_ms[f"bs={key[0]}{' edges' if key[1] else ''}{' mid' if key[1] & R.BUILD_MID else ''}"] = \
                (time.perf_counter() - t0) * 1e3
    index_for = {lr: indexes[(picks[lr]["bs"], picks[lr]["flags"])] for lr in DISTS}
    opts_for = {lr: picks[lr]["opts"] for lr in DISTS}
    bs_for = {lr: picks[lr]["bs"] for lr in DISTS}
    steps_out, elapsed = timed_run(R, torch, index_for, opts_for, Q_dev, out_v, out_i, args.steps, args.warmup,
                                   "tuned")
    gpu_answers = {}  # the tuned answers, checked against HRMQ on the host after the timed region
    if not args.no_cpu_baseline:
        for lr in DISTS:
            index_for[lr].query(Q_dev[lr], out_val=out_v, out_idx=out_i, opts=opts_for[lr])
            torch.cuda.synchronize()
            gpu_answers[lr] = (out_v.cpu().numpy(), out_i.cpu().numpy())
    for b in (set(bs_for.values()) | ({32} if any(p["flags"] for p in picks.values()) else set())) - set(levels):
        levels[b] = BlockLevel(A, b)
    kernels_for = {lr: index_for[lr].kernels(opts_for[lr]) for lr in DISTS}
    tuned = summarise(A, Q_dev, q, steps_out, elapsed, levels, bs_for, opts_for,
                      {lr: picks[lr]["flags"] for lr in DISTS}, kernels_for)
    tuned_digest = answers_digest(torch, index_for, opts_for, Q_dev, out_v, out_i)

    # --- the literal block size with the tuner's structure --------------------------
    # rmq_tune restricted to bs = args.bs picks edge tables and thresholds per
    # distribution (its picks are cached under keys of their own)
    fixed_tuned = None
    if not args.no_fixed_tuned:
        fpicks = {}
        for lr in DISTS:
            sample = R.gen_queries_device(min(q, args.tune_sample), n, lr, args.seed + 7919 + 17 * (-lr), device=dev)
            fpicks[lr] = R.tune(A, sample, bs_list=[args.bs], cache_path=args.tune_cache or None, refresh=args.retune)
            del sample
        f_indexes = {}
        for lr in DISTS:
            if fpicks[lr]["flags"] not in f_indexes:
                f_indexes[fpicks[lr]["flags"]] = R.RMQIndex(A, args.bs, flags=fpicks[lr]["flags"])
        f_index_for = {lr: f_indexes[fpicks[lr]["flags"]] for lr in DISTS}
        f_opts = {lr: fpicks[lr]["opts"] for lr in DISTS}
        f_steps, f_elapsed = timed_run(R, torch, f_index_for, f_opts, Q_dev, out_v, out_i, args.fixed_steps, 1,
                                       f"bs={args.bs} tuned")
        if 32 not in levels:
            levels[32] = BlockLevel(A, 32)
        fixed_tuned = summarise(A, Q_dev, q, f_steps, f_elapsed, levels, {lr: args.bs for lr in DISTS}, f_opts,
                                {lr: fpicks[lr]["flags"] for lr in DISTS},
                                {lr: f_index_for[lr].kernels(f_opts[lr]) for lr in DISTS})
        f_digest = answers_digest(torch, f_index_for, f_opts, Q_dev, out_v, out_i)
        fixed_tuned["picks"] = {DIST_NAMES[lr]: {"edge_tables": bool(fpicks[lr]["flags"]),
                                                 "short_max": f_opts[lr].short_max,
                                                 "group_lanes": f_opts[lr].group_lanes,
                                                 "long_lanes": f_opts[lr].long_lanes,
                                                 "source": "cache" if fpicks[lr]["from_cache"] else "timed"}
                                for lr in DISTS}
        fixed_tuned["identical_to_tuned"] = f_digest == tuned_digest
        fixed_tuned["workload"] = workload_key(n, q, {lr: args.bs for lr in DISTS}, f_opts,
                                               {lr: fpicks[lr]["flags"] for lr in DISTS})
        fixed_tuned["steps"] = args.fixed_steps
        for ix in f_indexes.values():
            ix.close()
        log(f"bench: bs={args.bs} with the tuner's structure {fixed_tuned['value']:.4e} RMQ/s "
            f"(literal {fixed['value']:.4e}); answers identical to the tuned picks: {fixed_tuned['identical_to_tuned']}")
    log(f"bench: tuned {tuned['value']:.4e} RMQ/s; {tuned['dominant']} {tuned['achieved']:.1f} GB/s algorithmic "
        f"({tuned['achieved'] / HBM_PEAK_GBS:.3f} of peak); fixed bs={args.bs}: {fixed['value']:.4e} RMQ/s")
    for ix in indexes.values():
        ix.close()
    gather_ceiling = None
    dk = tuned["per_kernel"][tuned["dominant"]]
    if "table_reads_per_s" in dk and not args.no_gather_probe:
        l2 = R.probe_gather(table_bytes=1 << 21, device=dev)["gathers_per_s"]
        dram = R.probe_gather(table_bytes=1 << 30, device=dev)["gathers_per_s"]
        t_bound = ((0.0 if dk.get("table_in_lds") else dk["table_reads_per_launch"] / l2)
                   + (dk["edge_gathers_per_launch"] + dk["short_lines_per_launch"]) / dram)
        gather_ceiling = {
            "l2_table_reads_per_s": l2, "dram_gathers_per_s": dram,
            "kernel_table_reads_per_s": dk["table_reads_per_s"], "kernel_edge_gathers_per_s": dk["edge_gathers_per_s"],
            "kernel_short_lines_per_s": dk["short_lines_per_launch"] / (dk["avg_launch_ms"] / 1e3),
            "frac": t_bound / (dk["avg_launch_ms"] / 1e3),
            "probe": "rmq_probe_gather: 2^26 threads x 2 random 8 B reads from two 2 MiB tables (L2) and from two "
                     "1 GiB tables (DRAM); frac = (table reads / L2 rate + (edge-pair gathers + 128 B lines of "
                     "short ranges) / DRAM rate) / time"}
        log(f"bench: random-read ceilings: L2 tables {l2:.3e}/s, DRAM {dram:.3e}/s; {tuned['dominant']} "
            f"{dk['table_reads_per_s']:.3e} table reads/s + {dk['edge_gathers_per_s']:.3e} edge gathers/s "
            f"= {gather_ceiling['frac']:.3f} of those rates")

    edges_for = {lr: picks[lr]["flags"] for lr in DISTS}
    workload = workload_key(n, q, bs_for, opts_for, edges_for)
    fixed_workload = workload_key(n, q, {lr: args.bs for lr in DISTS}, {lr: fixed_opts for lr in DISTS}, None)
    del out_v, out_i
    batched = None if args.no_batched else run_batched(R, torch, dev, args)
    served = None if args.no_served else run_served(R, torch, dev, args)
    stress = None if args.no_stress else run_stress(R, torch, dev, args)
    cpu = None
    del levels
    if not args.no_cpu_baseline:
        plain_small = run_plain_small(R, torch, dev, args)
        cpu = cpu_baseline(A.cpu().numpy(), Q_host, args.cpu_seconds, gpu_answers, plain_small)

    attach_measured(tuned, workload, q)
    attach_measured(fixed, fixed_workload, q)
    traffic = load_traffic(tuned["dominant"], workload)
    dom_ms = tuned["per_kernel"][tuned["dominant"]]["avg_launch_ms"]
    fixed_traffic = load_traffic(fixed["dominant"], fixed_workload)
    fixed_ms = fixed["per_kernel"][fixed["dominant"]]["avg_launch_ms"]
    if gather_ceiling is not None:
        # the dominant kernel's measured read requests past the L2 (calibrated
        # counters of this workload) per second of its launches, against the
        # probe's rate of random requests from 1 GiB tables
        dom_dist = [d for d in tuned["per_dist"].values() if tuned["dominant"] in d["kernels"].values()]
        if len(dom_dist) == 1 and "dram_read_lines_per_rmq" in dom_dist[0]:
            req = dom_dist[0]["dram_read_lines_per_rmq"] * q / (dom_ms / 1e3)
            gather_ceiling["kernel_read_requests_per_s"] = req
            gather_ceiling["read_requests_over_probe"] = req / gather_ceiling["dram_gathers_per_s"]
    line = {
        "metric": "RMQ/s at n=1e8,q=2^26 per length dist, %HBM roofline; batched/served vs plain",
        "value": tuned["value"],
        "unit": "RMQ/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tuned["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: counter-based U[0,1) float32 array and host-generated queries, seed 1321",
        "config": {"workload": "rtx_blocks (alg 5 + alg 105, value+index), block size and thresholds autotuned "
                               "per distribution",
                   "n": n, "q": q, "dists": [DIST_NAMES[lr] for lr in DISTS],
                   "bs": {DIST_NAMES[lr]: bs_for[lr] for lr in DISTS},
                   "edge_tables": {DIST_NAMES[lr]: bool(picks[lr]["flags"]) for lr in DISTS},
                   "coarse_rows": {DIST_NAMES[lr]: bool(picks[lr]["flags"] & R.BUILD_MID) for lr in DISTS},
                   "construction_ms": {str(k): v for k, v in build_ms.items()},
                   "tune_sample": min(q, args.tune_sample),
                   "tune_source": {DIST_NAMES[lr]: ("cache" if picks[lr]["from_cache"] else "timed") for lr in DISTS}},
        "roofline": {"bound": "hbm", "achieved": tuned["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": tuned["achieved"] / HBM_PEAK_GBS,
                     "bytes_model": "literal: 8 + 4*S + T + 8 per RMQ (both edge parts scanned); "
                                    "achieved_model: the kernels' own decisions replayed",
                     "achieved_model": tuned["achieved_model"], "frac_model": tuned["achieved_model"] / HBM_PEAK_GBS,
                     "traffic": traffic, "kernel": tuned["dominant"],
                     # the HBM bytes the counters saw, over the same launch time: random
                     # 8-byte reads each move a whole 128-byte DRAM request
                     "traffic_GBps": traffic / (dom_ms / 1e3) / 1e9 if traffic else None,
                     "traffic_frac": traffic / (dom_ms / 1e3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "avg_launch_ms": tuned["per_kernel"][tuned["dominant"]]["avg_launch_ms"],
                     "algorithmic_bytes_per_launch":
                         tuned["per_kernel"][tuned["dominant"]]["algorithmic_bytes_per_launch"],
                     "workload": workload, "gather_ceiling": gather_ceiling},
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "stress": stress,
        "per_dist": tuned["per_dist"],
        "per_kernel": tuned["per_kernel"],
        f"fixed_bs_{args.bs}": {
            "value": fixed["value"], "ms_per_step": fixed["ms_per_step"], "steps": args.fixed_steps,
            "construction_ms": fixed_build_ms, "per_dist": fixed["per_dist"],
            "per_kernel": fixed["per_kernel"],
            "roofline": {"kernel": fixed["dominant"], "achieved": fixed["achieved"],
                         "frac": fixed["achieved"] / HBM_PEAK_GBS,
                         "achieved_model": fixed["achieved_model"],
                         "frac_model": fixed["achieved_model"] / HBM_PEAK_GBS,
                         "traffic": fixed_traffic,
                         "traffic_GBps": fixed_traffic / (fixed_ms / 1e3) / 1e9 if fixed_traffic else None,
                         "avg_launch_ms": fixed_ms,
                         "workload": fixed_workload},
            "identical_to_tuned": fixed_digest == tuned_digest},
    }
    if fixed_tuned is not None:
        ft_workload = fixed_tuned["workload"]
        attach_measured(fixed_tuned, ft_workload, q)
        ft_ms = fixed_tuned["per_kernel"][fixed_tuned["dominant"]]["avg_launch_ms"]
        ft_traffic = load_traffic(fixed_tuned["dominant"], ft_workload)
        line[f"fixed_bs_{args.bs}_tuned"] = {
            "value": fixed_tuned["value"], "ms_per_step": fixed_tuned["ms_per_step"], "steps": fixed_tuned["steps"],
            "picks": fixed_tuned["picks"], "identical_to_tuned": fixed_tuned["identical_to_tuned"],
            "per_dist": fixed_tuned["per_dist"], "per_kernel": fixed_tuned["per_kernel"],
            "roofline": {"kernel": fixed_tuned["dominant"], "achieved": fixed_tuned["achieved"],
                         "frac": fixed_tuned["achieved"] / HBM_PEAK_GBS,
                         "achieved_model": fixed_tuned["achieved_model"],
                         "frac_model": fixed_tuned["achieved_model"] / HBM_PEAK_GBS,
                         "traffic": ft_traffic,
                         "traffic_GBps": ft_traffic / (ft_ms / 1e3) / 1e9 if ft_traffic else None,
                         "avg_launch_ms": ft_ms, "workload": ft_workload}}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

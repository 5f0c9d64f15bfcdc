"""Headline benchmark: batch RMQ at n=1e8, q=2^26, three length distributions.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--bs 32768]

A step is one query pass (alg 5 values + alg 105 indices in one launch set)
over each of the three query batches: uniform lengths (lr=-1), lognormal
medium (lr=-2) and lognormal small (lr=-3), 2^26 queries each, against one
index built over n=1e8 float32 U[0,1) values with bs=2^15.  Inputs are
resident in HBM before timing starts.  value = 3*q*K / (timed seconds).

roofline: the dominant kernel (k_query_long, the edge-block wave scans) --
algorithmic bytes per query B = 8 (query) + 4*S (elements scanned) + T (two
8-byte table entries when whole blocks lie between) + 8 (value + index out),
summed over the queries that kernel processed, divided by its time measured
with HIP events on the stream it runs on.

cpu_baseline: the reference's HRMQ (RMQRMM64, built from the reference's own
sources into oracle/_ref by oracle/Makefile) on this host's cores, OpenMP
over queries as src/cpu_methods.h:104-110, on a bounded sample of each batch.
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


def algorithmic_bytes(Q: np.ndarray, bs: int, short_max: int, out_bytes: int = 8) -> tuple[float, float, int, int]:
    """(bytes in the long kernel, bytes in the short kernel, n_long, n_short) per pass."""
    l = Q[:, 0].astype(np.int64)
    r = Q[:, 1].astype(np.int64)
    ln = r - l + 1
    lB, rB = l // bs, r // bs
    same = lB == rB
    S = np.where(same, ln, (bs - l % bs) + (r % bs + 1))
    T = np.where(rB - lB > 1, 16, 0)
    B = 8 + 4 * S + T + out_bytes
    is_long = ln > short_max
    short_B = 8 + 4 * ln + out_bytes
    return float(B[is_long].sum()), float(short_B[~is_long].sum()), int(is_long.sum()), int((~is_long).sum())


def load_traffic(kernel: str, workload: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        rec = d.get(kernel, {})
        if rec.get("workload") == workload:
            return rec.get("bytes_per_launch")
    except (OSError, ValueError):
        return None
    return None


def cpu_baseline(A_host: np.ndarray, queries: dict, target_s: float) -> dict | None:
    """Reference HRMQ on the host cores over a bounded sample of each batch."""
    try:
        from oracle import oracle as O
    except Exception as e:  # noqa: BLE001
        log(f"cpu_baseline unavailable: {e}")
        return None
    nt = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    if not O.HRMQ.available():
        log("cpu_baseline: oracle/_ref/libhrmq.so missing; skipping")
        return None
    log(f"cpu_baseline: building reference HRMQ over n={A_host.size} (single thread)")
    t0 = time.perf_counter()
    h = O.HRMQ(A_host)
    build_s = time.perf_counter() - t0
    log(f"cpu_baseline: built in {build_s:.2f} s; timing queries on {nt} threads")
    total_q, total_t, per = 0, 0.0, {}
    for lr, Q in queries.items():
        pilot = Q[: 1 << 16]
        t0 = time.perf_counter()
        h.query(pilot, nt=nt, want_val=True, want_idx=True)
        rate = pilot.shape[0] / max(time.perf_counter() - t0, 1e-6)
        m = int(min(Q.shape[0], max(1 << 16, rate * target_s)))
        sample = Q[:m]
        t0 = time.perf_counter()
        h.query(sample, nt=nt, want_val=True, want_idx=True)
        dt = time.perf_counter() - t0
        per[DIST_NAMES[lr]] = {"rmq_per_s": m / dt, "queries": m}
        log(f"cpu_baseline: {DIST_NAMES[lr]} {m / dt:.3e} RMQ/s over {m} queries")
        total_q += m
        total_t += dt
    size = h.size_bytes()
    h.close()
    return {
        "value": total_q / total_t,
        "unit": "RMQ/s",
        "cores": nt,
        "kind": "reference",
        "sample": (f"reference HRMQ (RMQRMM64, alg 1/101 value+index) on the first {total_q} queries of the same "
                   f"three batches (about {target_s:.0f}s each on {nt} threads); construction {build_s:.2f}s "
                   f"single-threaded, {size / 1e6:.1f} MB"),
        "per_dist": per,
    }


def main() -> int:
    faulthandler.enable()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100_000_000)
    ap.add_argument("--q", type=int, default=1 << 26)
    ap.add_argument("--bs", type=int, default=1 << 15)
    ap.add_argument("--seed", type=int, default=1321)
    ap.add_argument("--cpu-seconds", type=float, default=1.5, help="wall seconds of HRMQ per batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); queries shard with no exchange "
              f"across GPUs, so the one-GPU rate is the unit of scale.", file=sys.stderr)
        return 2

    import torch

    import rtxrmq_amd as R

    dev = torch.device("cuda:0")
    n, q, bs = args.n, args.q, args.bs
    log(f"bench: n={n} q={q} bs={bs} dists={DISTS} steps={args.steps} warmup={args.warmup}")

    A = R.gen_array(n, args.seed, "uniform", device=dev)
    Q_host = {lr: R.gen_queries(q, n, lr, args.seed + 17 * (-lr)) for lr in DISTS}
    Q_dev = {lr: torch.from_numpy(Q).to(dev) for lr, Q in Q_host.items()}
    out_v = torch.empty(q, dtype=torch.float32, device=dev)
    out_i = torch.empty(q, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    index = R.RMQIndex(A, bs)
    torch.cuda.synchronize()
    construction_ms = (time.perf_counter() - t0) * 1e3
    opts = index.default_opts()
    log(f"bench: index built in {construction_ms:.1f} ms ({index.info()})")

    def one_step():
        res = {}
        for lr in DISTS:
            _, _, t = index.query_timed(Q_dev[lr], out_val=out_v, out_idx=out_i, opts=opts)
            res[lr] = t
        return res

    for w in range(args.warmup):
        one_step()
        log(f"bench: warmup {w + 1}/{args.warmup} done")

    torch.cuda.synchronize()
    t_start = time.perf_counter()
    steps = []
    for k in range(args.steps):
        steps.append(one_step())
        log(f"bench: step {k + 1}/{args.steps}: " + ", ".join(
            f"{DIST_NAMES[lr]} {s['total_ms']:.2f} ms" for lr, s in steps[-1].items()))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start

    total_queries = 3 * q * args.steps
    value = total_queries / elapsed
    per_dist = {}
    long_bytes = long_ms = short_bytes = short_ms = 0.0
    for lr in DISTS:
        lb, sb, nl, ns = algorithmic_bytes(Q_host[lr], bs, opts.short_max)
        ms = [s[lr]["total_ms"] for s in steps]
        lms = [s[lr]["long_ms"] for s in steps]
        sms = [s[lr]["short_ms"] for s in steps]
        per_dist[DIST_NAMES[lr]] = {
            "rmq_per_s": q / (np.mean(ms) / 1e3),
            "ms_per_pass": float(np.mean(ms)),
            "long_kernel_ms": float(np.mean(lms)),
            "short_kernel_ms": float(np.mean(sms)),
            "n_long": nl,
            "algorithmic_GBps": (lb + sb) / (np.mean(ms) / 1e3) / 1e9,
            "roofline_rmq_per_s": q / ((lb + sb) / (HBM_PEAK_GBS * 1e9)),
        }
        long_bytes += lb * args.steps
        long_ms += float(np.sum(lms))
        short_bytes += sb * args.steps
        short_ms += float(np.sum(sms))
    achieved = long_bytes / (long_ms / 1e3) / 1e9
    workload = f"n={n} q={q} bs={bs} lr=-1,-2,-3 value+index"
    traffic = load_traffic("k_query_long", workload)

    log(f"bench: long kernel {achieved:.1f} GB/s algorithmic ({achieved / HBM_PEAK_GBS:.3f} of peak)")
    cpu = None
    if not args.no_cpu_baseline:
        A_host = A.cpu().numpy()
        cpu = cpu_baseline(A_host, Q_host, args.cpu_seconds)

    line = {
        "metric": "RMQ/s at n=1e8, q=2^26 over the three length distributions (uniform, lognormal 2^14, "
                  "lognormal 2^6), bs=2^15",
        "value": value,
        "unit": "RMQ/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: counter-based U[0,1) float32 array and host-generated queries, seed 1321",
        "config": {"workload": "rtx_blocks plain (alg 5 + alg 105)", "n": n, "q": q, "bs": bs,
                   "dists": [DIST_NAMES[lr] for lr in DISTS], "outputs": "value+index",
                   "short_max": opts.short_max, "group_lanes": opts.group_lanes,
                   "construction_ms": construction_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_query_long"},
        "cpu_baseline": cpu,
        "per_dist": per_dist,
        "short_kernel": {"achieved_GBps": short_bytes / max(short_ms, 1e-9) * 1e3 / 1e9},
    }
    index.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

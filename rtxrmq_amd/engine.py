"""Host-side mirror of the reference's RTX_blocks operator over the C ABI.

The reference's interface for this path is
    template <typename T> T* rtx_rmq(int alg, int n, int bs, int q,
                                     float* darray, int2* dquery, CmdArgs args)
(src/rtx.h:24-177): build the block structure, run the query launch `reps`
times with a synchronisation after each (src/rtx.h:150-155), copy the result
to the host and report RMQ/s = q / (time per rep).  `rtx_rmq` below keeps
that contract (same alg ids, same argument meaning, host array out); the
`RMQIndex` class exposes the split build/query form of include/rmq.h.

torch is only the device-memory and stream plumbing; every computation runs
in librtxrmq.so.
"""
from __future__ import annotations

import atexit
import ctypes
import sys
import time
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

# algorithm ids, src/main.cu:17-32
ALG_CPU_BASE = 0
ALG_CPU_HRMQ = 1
ALG_GPU_BASE = 2
ALG_GPU_RTX_BLOCKS = 5
ALG_CPU_BASE_IDX = 100
ALG_CPU_HRMQ_IDX = 101
ALG_GPU_BASE_IDX = 102
ALG_GPU_RTX_BLOCKS_IDX = 105


def _torch():
    import torch
    return torch


def _stream_ptr(stream) -> int | None:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream or None


def _dev_f32(t, name: str):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous float32 CUDA tensor")
    return t


def _dev_queries(t, device=None):
    """Queries for a kernel: a contiguous int32 CUDA tensor of {l, r} pairs,
    on `device` (the array's) when one is given -- the kernels read it there."""
    torch = _torch()
    if (not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.int32
            or not t.is_contiguous() or t.numel() % 2):
        raise TypeError("queries must be a contiguous int32 CUDA tensor of {l, r} pairs")
    if device is not None and t.device != device:
        raise TypeError(f"queries are on {t.device}, the array on {device}")
    return t


def _dev_out(t, dtype: str, q: int, device, name: str):
    """A caller's output buffer: the kernels write q entries into it, so it
    must be a contiguous CUDA tensor of the right dtype, on the array's device,
    with at least q entries."""
    torch = _torch()
    if (not isinstance(t, torch.Tensor) or t.device != device or t.dtype != getattr(torch, dtype)
            or not t.is_contiguous() or t.numel() < q):
        raise TypeError(f"{name} must be a contiguous {dtype} tensor on {device} with at least {q} entries")
    return t


# Native objects still open.  They are closed by an exit hook registered at
# the first one's creation (so it runs before the exit hooks of modules
# imported earlier, torch's included): servers first (their threads, streams
# and pinned buffers), then plans, then indexes, all while the interpreter
# and the HIP runtime are still whole.  During interpreter finalization the
# finalizers do nothing: module globals (ctypes, this module's library
# handle) may already be gone by then, and the HIP runtime's own teardown
# follows.
_LIVE: "weakref.WeakSet" = weakref.WeakSet()
_EXIT_HOOK = [False]


def _track(obj) -> None:
    if not _EXIT_HOOK[0]:
        atexit.register(_close_all)
        _EXIT_HOOK[0] = True
    _LIVE.add(obj)


def _close_all() -> None:
    objs = list(_LIVE)
    for kind in (Server, Plan, RMQIndex, HostHRMQ):
        for o in objs:
            if isinstance(o, kind):
                try:
                    o.close()
                except Exception:  # noqa: BLE001 -- best effort at exit
                    pass


class _Native:
    """Closing shared by the native-object wrappers: close() releases the
    object, a `with` block closes it at its end, and the finalizer closes it
    unless the interpreter is finalizing.  The finalizer holds its own
    reference to sys.is_finalizing: module globals may already be None when
    an object is collected during teardown."""

    def close(self):  # pragma: no cover -- every subclass defines it
        raise NotImplementedError

    def __del__(self, _is_finalizing=sys.is_finalizing):
        if _is_finalizing():
            return
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def query_opts(short_max: int = 0, group_lanes: int = 0, long_wgs: int = 0, long_lanes: int = 0) -> N.QueryOpts:
    return N.QueryOpts(short_max, group_lanes, long_wgs, long_lanes)


class RMQIndex(_Native):
    """A built block index over a device array (rmq_build / rmq_query / rmq_free).

    The array tensor is kept referenced: the edge-block scans read it.
    """

    def __init__(self, A, bs: int, flags: int = 0, stream=None):
        self.A = _dev_f32(A, "A")
        self.n = int(A.numel())
        self.bs = int(bs)
        self.flags = int(flags)
        h = ctypes.c_void_p()
        N.check(N.lib().rmq_build(A.data_ptr(), self.n, self.bs, self.flags, _stream_ptr(stream), ctypes.byref(h)),
                "rmq_build")
        self._h = h
        _track(self)

    @property
    def handle(self) -> int:
        return self._h.value

    def info(self) -> dict:
        n, nb, byt = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        bs, lev = ctypes.c_int32(), ctypes.c_int32()
        N.check(N.lib().rmq_handle_info(self._h, ctypes.byref(n), ctypes.byref(bs), ctypes.byref(nb),
                                        ctypes.byref(lev), ctypes.byref(byt)), "rmq_handle_info")
        return {"n": n.value, "bs": bs.value, "nb": nb.value, "levels": lev.value, "bytes": byt.value}

    def kernels(self, opts: N.QueryOpts | None = None) -> list[str]:
        """The kernels a query on this handle launches, as a profiler names them."""
        buf = ctypes.create_string_buffer(256)
        N.check(N.lib().rmq_query_kernels(self._h, ctypes.byref(opts) if opts is not None else None, buf, 256),
                "rmq_query_kernels")
        return buf.value.decode().split(";")

    def default_opts(self) -> N.QueryOpts:
        o = N.QueryOpts()
        N.check(N.lib().rmq_query_opts_default(self._h, ctypes.byref(o)), "rmq_query_opts_default")
        return o

    def _outs(self, q: int, want_value: bool, want_index: bool, out_val, out_idx):
        torch = _torch()
        dev = self.A.device
        if want_value:
            out_val = (torch.empty(q, dtype=torch.float32, device=dev) if out_val is None
                       else _dev_out(out_val, "float32", q, dev, "out_val"))
        if want_index:
            out_idx = (torch.empty(q, dtype=torch.int32, device=dev) if out_idx is None
                       else _dev_out(out_idx, "int32", q, dev, "out_idx"))
        return (out_val if want_value else None), (out_idx if want_index else None)

    def query(self, Q, want_value: bool = True, want_index: bool = True, opts: N.QueryOpts | None = None,
              out_val=None, out_idx=None, stream=None):
        Q = _dev_queries(Q, self.A.device)
        q = Q.numel() // 2
        v, i = self._outs(q, want_value, want_index, out_val, out_idx)
        rc = N.lib().rmq_query_ex(self._h, Q.data_ptr(), q, v.data_ptr() if v is not None else None,
                                  i.data_ptr() if i is not None else None,
                                  ctypes.byref(opts) if opts is not None else None, _stream_ptr(stream))
        N.check(rc, "rmq_query")
        return v, i

    def query_timed(self, Q, want_value: bool = True, want_index: bool = True, opts: N.QueryOpts | None = None,
                    out_val=None, out_idx=None, stream=None):
        Q = _dev_queries(Q, self.A.device)
        q = Q.numel() // 2
        v, i = self._outs(q, want_value, want_index, out_val, out_idx)
        t = N.QueryTimes()
        rc = N.lib().rmq_query_timed(self._h, Q.data_ptr(), q, v.data_ptr() if v is not None else None,
                                     i.data_ptr() if i is not None else None,
                                     ctypes.byref(opts) if opts is not None else None, _stream_ptr(stream),
                                     ctypes.byref(t))
        N.check(rc, "rmq_query_timed")
        return v, i, {"short_ms": t.short_ms, "long_ms": t.long_ms, "total_ms": t.total_ms}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().rmq_free(self._h)
            self._h = ctypes.c_void_p()


class HostHRMQ(_Native):
    """The host HRMQ of algs 1/101 (rmq_hrmq_*): rightmost minimum index and
    its value bits, over values compared as int32 (a float array is read as
    its bit patterns, src/main.cu:117-125).  Runs on the CPU only."""

    def __init__(self, A: np.ndarray):
        A = np.ascontiguousarray(A)
        if A.dtype not in (np.dtype(np.float32), np.dtype(np.int32)) or A.ndim != 1:
            raise TypeError("HostHRMQ takes a 1-d float32 or int32 array")
        self.A = A.view(np.int32)
        h = ctypes.c_void_p()
        N.check(N.lib().rmq_hrmq_build(self.A.ctypes.data, self.A.size, ctypes.byref(h)), "rmq_hrmq_build")
        self._h = h
        _track(self)

    def query(self, Q: np.ndarray, nthreads: int = 0, want_value: bool = True, want_index: bool = True):
        """(value as the input's dtype or None, rightmost index or None)."""
        Q = np.ascontiguousarray(Q, dtype=np.int32).reshape(-1, 2)
        q = Q.shape[0]
        idx = np.empty(q, np.int32) if want_index else None
        val = np.empty(q, np.int32) if want_value else None
        N.check(N.lib().rmq_hrmq_query(self._h, Q.ctypes.data, q, idx.ctypes.data if idx is not None else None,
                                       val.ctypes.data if val is not None else None, nthreads), "rmq_hrmq_query")
        return (val.view(np.float32) if val is not None else None), idx

    def bytes(self) -> int:
        b = ctypes.c_int64()
        N.check(N.lib().rmq_hrmq_bytes(self._h, ctypes.byref(b)), "rmq_hrmq_bytes")
        return b.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().rmq_hrmq_free(self._h)
            self._h = ctypes.c_void_p()


def handle_bytes(n: int, bs: int, flags: int = 0) -> int:
    """Device bytes rmq_build(n, bs, flags) allocates (rmq_handle_bytes; no GPU needed)."""
    b = ctypes.c_int64()
    N.check(N.lib().rmq_handle_bytes(n, bs, flags, ctypes.byref(b)), "rmq_handle_bytes")
    return b.value


def gpu_base(A, Q, want_value: bool = True, want_index: bool = True, stream=None):
    """Brute force on the device (algs 2/102; the --check oracle)."""
    torch = _torch()
    A = _dev_f32(A, "A")
    Q = _dev_queries(Q, A.device)
    q = Q.numel() // 2
    v = torch.empty(q, dtype=torch.float32, device=A.device) if want_value else None
    i = torch.empty(q, dtype=torch.int32, device=A.device) if want_index else None
    N.check(N.lib().rmq_gpu_base(A.data_ptr(), A.numel(), Q.data_ptr(), q, v.data_ptr() if v is not None else None,
                                 i.data_ptr() if i is not None else None, _stream_ptr(stream)), "rmq_gpu_base")
    return v, i


def check_queries(Q, n: int, stream=None) -> int:
    """Number of queries with l < 0, l > r or r >= n (synchronising)."""
    Q = _dev_queries(Q)
    bad = ctypes.c_int64()
    rc = N.lib().rmq_query_check(Q.data_ptr(), Q.numel() // 2, n, _stream_ptr(stream), ctypes.byref(bad))
    if rc not in (N.RMQ_OK, N.RMQ_ERR_BAD_QUERY):
        N.check(rc, "rmq_query_check")
    return bad.value


def tune(A, Q_sample, bs_list=None, cache_path: str | None = None, refresh: bool = False, stream=None,
         want_value: bool = True, want_index: bool = True, flags_allowed: int = N.BUILD_EDGES | N.BUILD_MID) -> dict:
    """Autotuner pick (rmq_tune_ex) for this array and a sample of its queries,
    timing the outputs the caller will ask for (value and/or index), among
    the build flags in flags_allowed (N.BUILD_EDGES for handles that go into
    a Plan or Server, which read no coarse rows).

    With a cache file, a pick stored for the same key (tuner version, n,
    length percentiles, value signature, outputs) is returned without timing
    anything; refresh=True times the sweep anyway and appends the new pick
    (the last line for a key wins).
    Returns {"bs", "flags" (rmq_build flags), "opts" (QueryOpts), "best_ms", "n_tried", "from_cache", "key"}.
    """
    torch = _torch()
    A = _dev_f32(A, "A")
    Q_sample = _dev_queries(Q_sample, A.device)
    if not (want_value or want_index):
        raise ValueError("tune: time at least one output")
    q = Q_sample.numel() // 2
    v_scr = torch.empty(q, dtype=torch.float32, device=A.device) if want_value else None
    i_scr = torch.empty(q, dtype=torch.int32, device=A.device) if want_index else None
    arr = (ctypes.c_int32 * len(bs_list))(*bs_list) if bs_list else None
    res = N.TuneResult()
    use_cache = cache_path if (cache_path and not refresh) else None
    rc = N.lib().rmq_tune_ex(A.data_ptr(), A.numel(), Q_sample.data_ptr(), q, arr, len(bs_list) if bs_list else 0,
                             flags_allowed, v_scr.data_ptr() if v_scr is not None else None,
                             i_scr.data_ptr() if i_scr is not None else None,
                             use_cache.encode() if use_cache else None, _stream_ptr(stream), ctypes.byref(res))
    N.check(rc, "rmq_tune_ex")
    o = res.opts
    if cache_path and refresh:  # same line format rmq_tune writes
        with open(cache_path, "a") as f:
            f.write(f"{res.key.decode()} {res.bs} {o.short_max} {o.group_lanes} {o.long_lanes} {res.best_ms:.6f} "
                    f"{res.build_flags}\n")
    return {"bs": res.bs, "flags": res.build_flags,
            "opts": N.QueryOpts(o.short_max, o.group_lanes, o.long_wgs, o.long_lanes),
            "best_ms": res.best_ms, "n_tried": res.n_tried, "from_cache": bool(res.from_cache),
            "key": res.key.decode()}


def gen_array(n: int, seed: int, kind: str = "uniform", device="cuda", stream=None):
    """Device array from the counter-based generator (create_random_array_dev, src/rand.cuh:23-43)."""
    torch = _torch()
    A = torch.empty(n, dtype=torch.float32, device=device)
    N.check(N.lib().rmq_gen_array(A.data_ptr(), n, seed, N.GEN_KINDS[kind], _stream_ptr(stream)), "rmq_gen_array")
    return A


def gen_queries_ref(q: int, n: int, lr: int, seed: int, nt: int) -> np.ndarray:
    """The reference's query stream (random_queries_par_cpu, src/rand.cuh:116-201); depends on nt."""
    Q = np.empty((q, 2), np.int32)
    N.check(N.lib().rmq_gen_queries_ref(Q.ctypes.data, q, n, lr, seed, nt), "rmq_gen_queries_ref")
    return Q


def probe_gather(table_bytes: int = 1 << 30, count: int = 1 << 26, reps: int = 5, device="cuda",
                 stream=None) -> dict:
    """Random-request ceiling of this GPU (rmq_probe_gather): count threads
    each gather two 8-byte entries at hashed positions of two tables of
    table_bytes each.  Returns the launch time and random gathers per second."""
    torch = _torch()
    entries = table_bytes // 8
    if entries < 2 or entries & (entries - 1):
        raise ValueError("table_bytes must be 8 times a power of two")
    t0 = torch.zeros(entries, dtype=torch.int64, device=device)
    t1 = torch.ones(entries, dtype=torch.int64, device=device)
    out = torch.empty(count, dtype=torch.int64, device=device)
    ms = ctypes.c_float()
    N.check(N.lib().rmq_probe_gather(t0.data_ptr(), t1.data_ptr(), entries, count, out.data_ptr(), reps,
                                     _stream_ptr(stream), ctypes.byref(ms)), "rmq_probe_gather")
    del t0, t1, out
    return {"ms_per_launch": ms.value, "gathers_per_s": 2 * count / (ms.value / 1e3), "table_bytes": table_bytes,
            "count": count, "reps": reps}


def gen_queries_device(q: int, n: int, lr: int, seed: int, device="cuda", stream=None):
    """Device queries from the same counter-based stream (rmq_gen_queries); (q, 2) int32 tensor."""
    torch = _torch()
    Q = torch.empty((q, 2), dtype=torch.int32, device=device)
    N.check(N.lib().rmq_gen_queries(Q.data_ptr(), q, n, lr, seed, _stream_ptr(stream)), "rmq_gen_queries")
    return Q


class Plan(_Native):
    """Many arrays (or requests) answered by one persistent launch (rmq_plan_*).

    indexes[i] answers Qs[i] into out_idx[i] and/or out_val[i].  The tensors
    are referenced for the plan's lifetime (the device work list holds their
    addresses); their contents may change between launches.
    """

    def __init__(self, indexes, Qs, out_idx=None, out_val=None, opts: N.QueryOpts | None = None,
                 item_queries: int = 0, stream=None):
        m = len(indexes)
        if len(Qs) != m or (out_idx is not None and len(out_idx) != m) or (out_val is not None and len(out_val) != m):
            raise ValueError("indexes, Qs and outputs must have the same length")
        Qs = [_dev_queries(Q, ix.A.device) for ix, Q in zip(indexes, Qs)]
        for k, (ix, Q) in enumerate(zip(indexes, Qs)):
            q = Q.numel() // 2
            i_t = out_idx[k] if out_idx is not None else None
            v_t = out_val[k] if out_val is not None else None
            if i_t is not None:
                _dev_out(i_t, "int32", q, ix.A.device, f"out_idx[{k}]")
            if v_t is not None:
                _dev_out(v_t, "float32", q, ix.A.device, f"out_val[{k}]")
            if q and i_t is None and v_t is None:
                raise ValueError(f"batch {k} has queries but no output")
        self._keep = (list(indexes), Qs, out_idx, out_val)
        P = ctypes.c_void_p * m
        hs = P(*[ix.handle for ix in indexes])
        qp = P(*[Q.data_ptr() for Q in Qs])
        qs = (ctypes.c_int64 * m)(*[Q.numel() // 2 for Q in Qs])
        ip = P(*[t.data_ptr() if t is not None else None for t in out_idx]) if out_idx is not None else None
        vp = P(*[t.data_ptr() if t is not None else None for t in out_val]) if out_val is not None else None
        h = ctypes.c_void_p()
        rc = N.lib().rmq_plan_create(hs, m, qp, qs, vp, ip, ctypes.byref(opts) if opts is not None else None,
                                     item_queries, _stream_ptr(stream), ctypes.byref(h))
        N.check(rc, "rmq_plan_create")
        self._h = h
        _track(self)

    def launch(self, stream=None):
        N.check(N.lib().rmq_plan_launch(self._h, _stream_ptr(stream)), "rmq_plan_launch")

    def graph_launch(self, stream=None):
        N.check(N.lib().rmq_plan_graph_launch(self._h, _stream_ptr(stream)), "rmq_plan_graph_launch")

    def info(self) -> dict:
        a, it, q = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        N.check(N.lib().rmq_plan_info(self._h, ctypes.byref(a), ctypes.byref(it), ctypes.byref(q)), "rmq_plan_info")
        return {"arrays": a.value, "items": it.value, "queries": q.value}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().rmq_plan_free(self._h)
            self._h = ctypes.c_void_p()


class Server(_Native):
    """Continuous-batching request loop over resident arrays (rmq_server_*).

    add_array keeps the array tensor referenced; submit keeps the request's
    host buffers referenced until wait returns.  Host buffers may be numpy
    arrays (staged through device memory) or pinned torch tensors, which the
    kernel reads and writes directly unless stage_copies is set.
    """

    def __init__(self, slots: int = 0, mem_budget_bytes: int = 0, max_batch_queries: int = 0,
                 item_queries: int = 0, opts: N.QueryOpts | None = None, stage_copies: bool = False):
        cfg = N.ServerConfig(slots, 0, mem_budget_bytes, max_batch_queries, item_queries,
                             N.SERVER_STAGE_COPIES if stage_copies else 0,
                             opts if opts is not None else N.QueryOpts())
        h = ctypes.c_void_p()
        N.check(N.lib().rmq_server_create(ctypes.byref(cfg), ctypes.byref(h)), "rmq_server_create")
        self._h = h
        _track(self)
        self._arrays = []
        self._live = {}

    def add_array(self, A, bs: int, flags: int = 0) -> int:
        """Builds A's tables on the server's own stream, which is not ordered
        after the caller's: the current stream (which wrote A, as torch and
        gen_array do) is synchronised first.  An array written on another
        stream must be finished before this call."""
        A = _dev_f32(A, "A")
        _torch().cuda.current_stream(A.device).synchronize()
        aid = ctypes.c_int32()
        N.check(N.lib().rmq_server_add_array_ex(self._h, A.data_ptr(), A.numel(), bs, flags, ctypes.byref(aid)),
                "rmq_server_add_array")
        self._arrays.append(A)
        return aid.value

    @staticmethod
    def _host_ptr(x, kind: str, min_count: int):
        """(object to keep alive, address) of a contiguous host buffer of at
        least min_count elements of `kind` ("int32" or "float32")."""
        if x is None:
            return None, None
        if isinstance(x, np.ndarray):
            if not x.flags["C_CONTIGUOUS"] or x.dtype != np.dtype(kind) or x.size < min_count:
                raise TypeError(f"host buffer must be a contiguous {kind} array of at least {min_count} entries")
            return x, x.ctypes.data
        torch = _torch()
        if (x.device.type != "cpu" or not x.is_contiguous() or x.dtype != getattr(torch, kind)
                or x.numel() < min_count):
            raise TypeError(f"host buffer must be a contiguous CPU {kind} tensor of at least {min_count} entries")
        return x, x.data_ptr()

    def submit(self, array_id: int, Q, out_idx=None, out_val=None) -> int:
        """Q: (q, 2) int32 host queries; out_idx / out_val: host outputs of q entries."""
        q = (Q.size if isinstance(Q, np.ndarray) else Q.numel()) // 2
        keep_q, pq = self._host_ptr(Q, "int32", 2 * q)
        keep_i, pi = self._host_ptr(out_idx, "int32", q)
        keep_v, pv = self._host_ptr(out_val, "float32", q)
        tk = ctypes.c_uint64()
        N.check(N.lib().rmq_server_submit(self._h, array_id, pq, q, pi, pv, ctypes.byref(tk)), "rmq_server_submit")
        self._live[tk.value] = (keep_q, keep_i, keep_v)
        return tk.value

    def wait(self, ticket: int) -> tuple[int, float]:
        """(status, latency_ms); status is RMQ_OK or a negative code."""
        lat = ctypes.c_double()
        rc = N.lib().rmq_server_wait(self._h, ticket, ctypes.byref(lat))
        self._live.pop(ticket, None)
        return rc, lat.value

    def stats(self, reset: bool = False) -> dict:
        """Counters since the last reset; latency percentiles over the latest
        2^20 requests of them at most (rmq_server_stats)."""
        st = N.ServerStats()
        N.check(N.lib().rmq_server_stats(self._h, ctypes.byref(st), 1 if reset else 0), "rmq_server_stats")
        return {name: getattr(st, name) for name, _ in N.ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().rmq_server_destroy(self._h)
            self._h = ctypes.c_void_p()
            self._live.clear()


def gen_array_host(n: int, seed: int, kind: str = "uniform") -> np.ndarray:
    A = np.empty(n, np.float32)
    N.check(N.lib().rmq_gen_array_host(A.ctypes.data, n, seed, N.GEN_KINDS[kind]), "rmq_gen_array_host")
    return A


def gen_queries(q: int, n: int, lr: int, seed: int, nthreads: int = 0) -> np.ndarray:
    """Host queries (random_queries_par_cpu, src/rand.cuh:179-201); shape (q, 2) int32."""
    Q = np.empty((q, 2), np.int32)
    N.check(N.lib().rmq_gen_queries_host(Q.ctypes.data, q, n, lr, seed, nthreads), "rmq_gen_queries_host")
    return Q


@dataclass
class RunResult:
    """What the reference prints and writes per run (src/rtx.h:160-162, src/tools.h:299-336)."""
    output: np.ndarray
    time_ms: float                 # total over reps
    reps: int
    q: int
    construction_ms: float
    index_bytes: int
    extra: dict = field(default_factory=dict)

    @property
    def rmq_per_s(self) -> float:
        return self.q / (self.time_ms / self.reps / 1000.0)

    @property
    def ns_per_rmq(self) -> float:
        return self.time_ms / self.reps * 1e6 / self.q


def rtx_rmq(alg: int, n: int, bs: int, q: int, darray, dquery, reps: int = 10, opts: N.QueryOpts | None = None,
            stream=None) -> RunResult:
    """Same contract as the reference's rtx_rmq<T> (src/rtx.h:24-177) for algs 5 and 105.

    Builds the index, runs `reps` query passes each followed by a device
    synchronisation (src/rtx.h:151-154), and returns the host output: float32
    minima for alg 5, leftmost int32 indices for alg 105.
    """
    torch = _torch()
    if alg not in (ALG_GPU_RTX_BLOCKS, ALG_GPU_RTX_BLOCKS_IDX):
        raise ValueError(f"rtx_rmq handles algs 5 and 105, not {alg}")
    if darray.numel() != n or dquery.numel() != 2 * q:
        raise ValueError("n / q do not match the device buffers")
    want_value = alg < 100
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    index = RMQIndex(darray, bs, stream=stream)
    torch.cuda.synchronize()
    construction_ms = (time.perf_counter() - t0) * 1e3
    try:
        out_v, out_i = index._outs(q, want_value, not want_value, None, None)
        t0 = time.perf_counter()
        for _ in range(reps):
            index.query(dquery, want_value, not want_value, opts=opts, out_val=out_v, out_idx=out_i, stream=stream)
            torch.cuda.synchronize()
        time_ms = (time.perf_counter() - t0) * 1e3
        out = (out_v if want_value else out_i).cpu().numpy()
        return RunResult(out, time_ms, reps, q, construction_ms, index.info()["bytes"])
    finally:
        index.close()

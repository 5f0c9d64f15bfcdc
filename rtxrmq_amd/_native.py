"""ctypes binding of librtxrmq.so (the C ABI declared in include/rmq.h).

The library is built in-tree (python -m rtxrmq_amd.build).  There is no
fallback: when the library is missing every entry point raises, so nothing
ever silently runs on the host instead of the GPU.
"""
from __future__ import annotations

import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "librtxrmq.so")
HEADER = os.path.join(ROOT, "include", "rmq.h")

RMQ_OK = 0
RMQ_ERR_INVALID_ARG = -1
RMQ_ERR_OOM = -2
RMQ_ERR_HIP = -3
RMQ_ERR_DEVICE = -4
RMQ_ERR_BAD_QUERY = -5
RMQ_ERR_LIMIT = -6

GEN_KINDS = {"uniform": 0, "ties16": 1, "desc": 2, "const": 3}
BUILD_EDGES = 1  # RMQ_BUILD_EDGES
BUILD_MID = 2  # RMQ_BUILD_MID (with BUILD_EDGES)
SERVER_STAGE_COPIES = 1  # RMQ_SERVER_STAGE_COPIES

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64


class RMQError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = strerror(code) if _lib_or_none() is not None else f"code {code}"
        super().__init__(f"{what}: {msg} ({code})" if what else f"{msg} ({code})")


class QueryOpts(ctypes.Structure):
    """rmq_query_opts_t"""
    _fields_ = [("short_max", _i32), ("group_lanes", _i32), ("long_wgs", _i32), ("long_lanes", _i32)]


class QueryTimes(ctypes.Structure):
    """rmq_query_times_t"""
    _fields_ = [("short_ms", ctypes.c_float), ("long_ms", ctypes.c_float), ("total_ms", ctypes.c_float)]


class TuneResult(ctypes.Structure):
    """rmq_tune_result_t"""
    _fields_ = [("bs", _i32), ("build_flags", _u32), ("opts", QueryOpts), ("best_ms", ctypes.c_float), ("n_tried", _i32),
                ("from_cache", _i32), ("key", ctypes.c_char * 96)]


class ServerConfig(ctypes.Structure):
    """rmq_server_config_t"""
    _fields_ = [("slots", _i32), ("max_arrays", _i32), ("mem_budget_bytes", _i64), ("max_batch_queries", _i64),
                ("item_queries", _i32), ("flags", _i32), ("opts", QueryOpts)]


class ServerStats(ctypes.Structure):
    """rmq_server_stats_t"""
    _fields_ = [("requests_done", _i64), ("queries_done", _i64), ("launches", _i64), ("failed", _i64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("mean_requests_per_launch", ctypes.c_double), ("budget_bytes", _i64),
                ("peak_inflight_bytes", _i64), ("mean_enqueue_ms", ctypes.c_double),
                ("mean_device_ms", ctypes.c_double), ("mean_launch_wait_ms", ctypes.c_double)]


_LIB = None


def _declare(lib):
    sig = {
        "rmq_build": (ctypes.c_int, [_vp, _i64, _i32, _u32, _vp, ctypes.POINTER(_vp)]),
        "rmq_query": (ctypes.c_int, [_vp, _vp, _i64, _vp, _vp, _vp]),
        "rmq_query_ex": (ctypes.c_int, [_vp, _vp, _i64, _vp, _vp, ctypes.POINTER(QueryOpts), _vp]),
        "rmq_query_timed": (ctypes.c_int, [_vp, _vp, _i64, _vp, _vp, ctypes.POINTER(QueryOpts), _vp,
                                           ctypes.POINTER(QueryTimes)]),
        "rmq_query_check": (ctypes.c_int, [_vp, _i64, _i64, _vp, ctypes.POINTER(_i64)]),
        "rmq_free": (None, [_vp]),
        "rmq_handle_info": (ctypes.c_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i32), ctypes.POINTER(_i64),
                                           ctypes.POINTER(_i32), ctypes.POINTER(_i64)]),
        "rmq_handle_bytes": (ctypes.c_int, [_i64, _i32, _u32, ctypes.POINTER(_i64)]),
        "rmq_query_opts_default": (ctypes.c_int, [_vp, ctypes.POINTER(QueryOpts)]),
        "rmq_query_kernels": (ctypes.c_int, [_vp, ctypes.POINTER(QueryOpts), ctypes.c_char_p, _i64]),
        "rmq_gpu_base": (ctypes.c_int, [_vp, _i64, _vp, _i64, _vp, _vp, _vp]),
        "rmq_hrmq_build": (ctypes.c_int, [_vp, _i64, ctypes.POINTER(_vp)]),
        "rmq_hrmq_query": (ctypes.c_int, [_vp, _vp, _i64, _vp, _vp, _i32]),
        "rmq_hrmq_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(_i64)]),
        "rmq_hrmq_free": (None, [_vp]),
        "rmq_tune": (ctypes.c_int, [_vp, _i64, _vp, _i64, _vp, _i32, _vp, _vp, ctypes.c_char_p, _vp,
                                    ctypes.POINTER(TuneResult)]),
        "rmq_tune_ex": (ctypes.c_int, [_vp, _i64, _vp, _i64, _vp, _i32, _u32, _vp, _vp, ctypes.c_char_p, _vp,
                                       ctypes.POINTER(TuneResult)]),
        "rmq_gen_array": (ctypes.c_int, [_vp, _i64, _u64, _i32, _vp]),
        "rmq_gen_array_host": (ctypes.c_int, [_vp, _i64, _u64, _i32]),
        "rmq_gen_queries_host": (ctypes.c_int, [_vp, _i64, _i32, _i32, _u64, _i32]),
        "rmq_gen_queries": (ctypes.c_int, [_vp, _i64, _i32, _i32, _u64, _vp]),
        "rmq_gen_queries_ref": (ctypes.c_int, [_vp, _i64, _i32, _i32, _i32, _i32]),
        "rmq_plan_create": (ctypes.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, ctypes.POINTER(QueryOpts), _i32, _vp,
                                           ctypes.POINTER(_vp)]),
        "rmq_plan_launch": (ctypes.c_int, [_vp, _vp]),
        "rmq_plan_graph_launch": (ctypes.c_int, [_vp, _vp]),
        "rmq_plan_info": (ctypes.c_int, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i64)]),
        "rmq_plan_free": (None, [_vp]),
        "rmq_query_loop": (ctypes.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, ctypes.POINTER(QueryOpts), _vp]),
        "rmq_plan_launch_static": (ctypes.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, ctypes.POINTER(QueryOpts), _i32,
                                                  _vp]),
        "rmq_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]),
        "rmq_server_add_array": (ctypes.c_int, [_vp, _vp, _i64, _i32, ctypes.POINTER(_i32)]),
        "rmq_server_add_array_ex": (ctypes.c_int, [_vp, _vp, _i64, _i32, _u32, ctypes.POINTER(_i32)]),
        "rmq_server_submit": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _vp, ctypes.POINTER(_u64)]),
        "rmq_server_wait": (ctypes.c_int, [_vp, _u64, ctypes.POINTER(ctypes.c_double)]),
        "rmq_server_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ServerStats), _i32]),
        "rmq_server_destroy": (None, [_vp]),
        "rmq_probe_gather": (ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _i32, _vp, ctypes.POINTER(ctypes.c_float)]),
        "rmq_strerror": (ctypes.c_char_p, [ctypes.c_int]),
        "rmq_version": (ctypes.c_char_p, []),
    }
    optional = {}
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for name, (res, args) in optional.items():
        if hasattr(lib, name):
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args


def _lib_or_none():
    return _LIB


def lib():
    """The loaded librtxrmq.so; raises when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library missing: {LIB_PATH} (build with: python -m rtxrmq_amd.build)")
        handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        _declare(handle)
        _LIB = handle
    return _LIB


def strerror(code: int) -> str:
    return lib().rmq_strerror(code).decode()


def check(rc: int, what: str = "") -> None:
    if rc != RMQ_OK:
        raise RMQError(rc, what)


def header_functions() -> list[str]:
    """Names of every function include/rmq.h declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rmq_[a-z0-9_]+)\s*\(", text)))

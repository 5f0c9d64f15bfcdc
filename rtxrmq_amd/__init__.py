"""rtxrmq_amd -- MI355X-native batch range-minimum queries (the RTX_blocks path).

The kernels and the C ABI live in librtxrmq.so (rtxrmq_amd/csrc, include/rmq.h);
this package is the host-side mirror of the reference's operator interface.
"""
from ._native import (BUILD_EDGES, BUILD_MID, GEN_KINDS, RMQ_ERR_BAD_QUERY, RMQ_ERR_DEVICE, RMQ_ERR_HIP, RMQ_ERR_INVALID_ARG,
                      RMQ_ERR_OOM, RMQ_OK, QueryOpts, QueryTimes, RMQError, check, header_functions, lib,
                      strerror)
from .engine import (ALG_CPU_BASE, ALG_CPU_BASE_IDX, ALG_CPU_HRMQ, ALG_CPU_HRMQ_IDX, ALG_GPU_BASE,
                     ALG_GPU_BASE_IDX, ALG_GPU_RTX_BLOCKS, ALG_GPU_RTX_BLOCKS_IDX, RMQIndex, RunResult,
                     Plan, Server, HostHRMQ, check_queries, gen_array, handle_bytes, gen_array_host, gen_queries, gen_queries_device, gen_queries_ref, gpu_base,
                     probe_gather, query_opts, rtx_rmq, tune)

__all__ = [
    "BUILD_EDGES", "BUILD_MID", "GEN_KINDS", "RMQ_OK", "RMQ_ERR_INVALID_ARG", "RMQ_ERR_OOM", "RMQ_ERR_HIP", "RMQ_ERR_DEVICE",
    "RMQ_ERR_BAD_QUERY", "QueryOpts", "QueryTimes", "RMQError", "check", "header_functions", "lib", "strerror",
    "ALG_CPU_BASE", "ALG_CPU_BASE_IDX", "ALG_CPU_HRMQ", "ALG_CPU_HRMQ_IDX", "ALG_GPU_BASE", "ALG_GPU_BASE_IDX",
    "ALG_GPU_RTX_BLOCKS", "ALG_GPU_RTX_BLOCKS_IDX", "RMQIndex", "RunResult", "check_queries", "gen_array",
    "gen_array_host", "handle_bytes", "HostHRMQ", "gen_queries", "gen_queries_device", "gen_queries_ref", "gpu_base", "probe_gather", "query_opts", "rtx_rmq", "tune", "Plan", "Server",
]

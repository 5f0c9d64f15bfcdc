"""ctypes wrapper over the CPU oracle (oracle/liboracle.so) and the reference
HRMQ built from its own sources (oracle/_ref/libhrmq.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed host
baseline, never as the thing measured.  See rmq_oracle.c for what each
function restates (reference file:line) and what pins it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
HRMQ_PATH = os.path.join(HERE, "_ref", "libhrmq.so")
REF_ROOT = "/root/reference"

_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_u64 = ctypes.c_uint64
_vp = ctypes.c_void_p

ARRAY_KINDS = {"uniform": 0, "ties16": 1, "desc": 2, "const": 3}


def build(with_ref: bool = True) -> None:
    """Compile liboracle.so (and _ref/libhrmq.so when the reference is present)."""
    target = ["all"] if with_ref else ["oracle"]
    subprocess.run(["make", "-s", "-C", HERE, *target], check=True)


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_vp)


class _Lib:
    _lib = None

    @classmethod
    def get(cls):
        if cls._lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"oracle not built: {LIB_PATH} missing (run make -C oracle)")
            lib = ctypes.CDLL(LIB_PATH)
            lib.orc_cpu_min.argtypes = [_vp, _i32, _i32]
            lib.orc_cpu_min.restype = ctypes.c_float
            lib.orc_cpu_min_idx.argtypes = [_vp, _i32, _i32]
            lib.orc_cpu_min_idx.restype = _i32
            lib.orc_hrmq_idx.argtypes = [_vp, _i32, _i32]
            lib.orc_hrmq_idx.restype = _i32
            lib.orc_rmq_brute.argtypes = [_vp, _i64, _vp, _i64, _vp, _vp, ctypes.c_int]
            lib.orc_rmq_brute.restype = None
            lib.orc_rmq_brute_hrmq.argtypes = [_vp, _i64, _vp, _i64, _vp, ctypes.c_int]
            lib.orc_rmq_brute_hrmq.restype = None
            lib.orc_rmq_blocks.argtypes = [_vp, _i64, _i32, _vp, _i64, _vp, _vp, ctypes.c_int]
            lib.orc_rmq_blocks.restype = ctypes.c_int
            lib.orc_fast_build.argtypes = [_vp, _i64, _i32, ctypes.c_int]
            lib.orc_fast_build.restype = _vp
            lib.orc_fast_query.argtypes = [_vp, _vp, _i64, _vp, _vp, ctypes.c_int]
            lib.orc_fast_query.restype = None
            lib.orc_fast_free.argtypes = [_vp]
            lib.orc_fast_free.restype = None
            lib.orc_gen_array.argtypes = [_vp, _i64, _u64, _i32]
            lib.orc_gen_array.restype = ctypes.c_int
            lib.orc_gen_queries.argtypes = [_vp, _i64, _i32, _i32, _u64, ctypes.c_int]
            lib.orc_gen_queries.restype = ctypes.c_int
            cls._lib = lib
        return cls._lib


def _check_inputs(A: np.ndarray, Q: np.ndarray):
    A = np.ascontiguousarray(A, dtype=np.float32)
    Q = np.ascontiguousarray(Q, dtype=np.int32).reshape(-1, 2)
    return A, Q


def cpu_min(A: np.ndarray, l: int, r: int) -> np.float32:
    A = np.ascontiguousarray(A, dtype=np.float32)
    return np.float32(_Lib.get().orc_cpu_min(_ptr(A), l, r))


def cpu_min_idx(A: np.ndarray, l: int, r: int) -> int:
    A = np.ascontiguousarray(A, dtype=np.float32)
    return int(_Lib.get().orc_cpu_min_idx(_ptr(A), l, r))


def hrmq_idx(A: np.ndarray, l: int, r: int) -> int:
    Ai = np.ascontiguousarray(A, dtype=np.float32).view(np.int32)
    return int(_Lib.get().orc_hrmq_idx(_ptr(Ai), l, r))


def rmq_brute(A: np.ndarray, Q: np.ndarray, nt: int = 0):
    """CPU BASE (algs 0/100): values and leftmost indices."""
    A, Q = _check_inputs(A, Q)
    q = Q.shape[0]
    val = np.empty(q, np.float32)
    idx = np.empty(q, np.int32)
    _Lib.get().orc_rmq_brute(_ptr(A), A.size, _ptr(Q), q, _ptr(val), _ptr(idx), nt)
    return val, idx


def rmq_brute_hrmq(A: np.ndarray, Q: np.ndarray, nt: int = 0) -> np.ndarray:
    """HRMQ index semantics (alg 101): rightmost minimum over int32 bits."""
    A, Q = _check_inputs(A, Q)
    Ai = A.view(np.int32)
    idx = np.empty(Q.shape[0], np.int32)
    _Lib.get().orc_rmq_brute_hrmq(_ptr(Ai), A.size, _ptr(Q), Q.shape[0], _ptr(idx), nt)
    return idx


def rmq_blocks(A: np.ndarray, Q: np.ndarray, bs: int, nt: int = 0):
    """RTX_blocks decomposition (algs 5/105) restated on the CPU."""
    A, Q = _check_inputs(A, Q)
    q = Q.shape[0]
    val = np.empty(q, np.float32)
    idx = np.empty(q, np.int32)
    rc = _Lib.get().orc_rmq_blocks(_ptr(A), A.size, bs, _ptr(Q), q, _ptr(val), _ptr(idx), nt)
    if rc != 0:
        raise ValueError("orc_rmq_blocks failed")
    return val, idx


class FastOracle:
    """O(1)-per-query exact oracle for full-size checks (values + leftmost idx)."""

    def __init__(self, A: np.ndarray, block: int = 1024, nt: int = 0):
        self.A = np.ascontiguousarray(A, dtype=np.float32)
        self.h = _Lib.get().orc_fast_build(_ptr(self.A), self.A.size, block, nt)
        if not self.h:
            raise MemoryError("orc_fast_build failed")

    def query(self, Q: np.ndarray, nt: int = 0):
        Q = np.ascontiguousarray(Q, dtype=np.int32).reshape(-1, 2)
        q = Q.shape[0]
        val = np.empty(q, np.float32)
        idx = np.empty(q, np.int32)
        _Lib.get().orc_fast_query(self.h, _ptr(Q), q, _ptr(val), _ptr(idx), nt)
        return val, idx

    def close(self):
        if self.h:
            _Lib.get().orc_fast_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def gen_array(n: int, seed: int, kind: str = "uniform") -> np.ndarray:
    A = np.empty(n, np.float32)
    if _Lib.get().orc_gen_array(_ptr(A), n, seed, ARRAY_KINDS[kind]) != 0:
        raise ValueError(f"bad array kind/size: {kind}, n={n}")
    return A


def gen_queries(q: int, n: int, lr: int, seed: int, nt: int = 0) -> np.ndarray:
    Q = np.empty((q, 2), np.int32)
    if _Lib.get().orc_gen_queries(_ptr(Q), q, n, lr, seed, nt) != 0:
        raise ValueError(f"bad query spec: n={n}, lr={lr}")
    return Q


# ---------------------------------------------------------------------------
# reference HRMQ (compiled from the reference's sources into oracle/_ref)
# ---------------------------------------------------------------------------

class HRMQ:
    """The reference's RMQRMM64 (alg 1/101) via oracle/_ref/libhrmq.so."""

    _lib = None

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(HRMQ_PATH)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(HRMQ_PATH):
                raise RuntimeError(f"reference HRMQ not built: {HRMQ_PATH} missing")
            lib = ctypes.CDLL(HRMQ_PATH)
            lib.hrmq_build.argtypes = [_vp, _i64]
            lib.hrmq_build.restype = _vp
            lib.hrmq_size_bytes.argtypes = [_vp]
            lib.hrmq_size_bytes.restype = ctypes.c_uint32
            lib.hrmq_query.argtypes = [_vp, _i64, _i64]
            lib.hrmq_query.restype = _i64
            lib.hrmq_query_batch.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, ctypes.c_int]
            lib.hrmq_query_batch.restype = None
            lib.hrmq_free.argtypes = [_vp]
            lib.hrmq_free.restype = None
            cls._lib = lib
        return cls._lib

    def __init__(self, A: np.ndarray):
        self.Ai = np.ascontiguousarray(A, dtype=np.float32).view(np.int32)
        self.h = self.lib().hrmq_build(_ptr(self.Ai), self.Ai.size)
        if not self.h:
            raise MemoryError("hrmq_build failed")

    def size_bytes(self) -> int:
        return int(self.lib().hrmq_size_bytes(self.h))

    def query(self, Q: np.ndarray, nt: int = 0, want_val: bool = True, want_idx: bool = True):
        Q = np.ascontiguousarray(Q, dtype=np.int32).reshape(-1, 2)
        q = Q.shape[0]
        idx = np.empty(q, np.int32) if want_idx else None
        val = np.empty(q, np.int32) if want_val else None
        self.lib().hrmq_query_batch(self.h, _ptr(self.Ai), _ptr(Q), q, _ptr(idx), _ptr(val), nt)
        return (None if val is None else val.view(np.float32)), idx

    def close(self):
        if self.h:
            self.lib().hrmq_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

"""C-ABI checks that need no GPU: the library loads, exports every function
include/rmq.h declares, validates arguments before touching the device, and
its host-side generators agree with the oracle's."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import rtxrmq_amd as R
from rtxrmq_amd import _native as N
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_loads_and_exports_header():
    lib = R.lib()
    names = R.header_functions()
    assert {"rmq_build", "rmq_query", "rmq_free"} <= set(names)
    for name in names:
        assert hasattr(lib, name), f"{name} declared in include/rmq.h but not exported"
    dyn = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in dyn.splitlines() if line.strip()}
    assert set(names) <= exported


def test_handle_bytes_without_a_gpu():
    """rmq_handle_bytes is host arithmetic: validated arguments, and for
    bs <= 32 without edge tables only the block sparse table (nb pairs times
    floor(log2 nb) + 1 levels, 8 bytes each)."""
    lib = R.lib()
    b = ctypes.c_int64()
    assert lib.rmq_handle_bytes(0, 4, 0, ctypes.byref(b)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_handle_bytes(10, 0, 0, ctypes.byref(b)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_handle_bytes(10, 4, 7, ctypes.byref(b)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_handle_bytes(1 << 31, 4, 0, ctypes.byref(b)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_handle_bytes(10, 4, 0, None) == R.RMQ_ERR_INVALID_ARG
    for n, bs in ((1, 1), (100, 8), (1000, 32), (12345, 7)):
        nb = -(-n // bs)
        assert R.handle_bytes(n, bs) == 8 * nb * nb.bit_length()
    # edge tables add per-element pairs: 2 per element up to bs = 32, 4 above
    assert R.handle_bytes(1 << 20, 32, R.BUILD_EDGES) - R.handle_bytes(1 << 20, 32) >= 16 << 20
    assert R.handle_bytes(1 << 20, 1024, R.BUILD_EDGES) - R.handle_bytes(1 << 20, 1024, 0) >= 24 << 20


def test_coarse_rows_flag_and_size_without_a_gpu():
    """RMQ_BUILD_MID is accepted only with RMQ_BUILD_EDGES, and adds one row
    of 8-byte entries over coarse blocks of 2^s blocks, s the smallest that
    leaves at most 8192 of them: the same rule bench.py's byte model uses to
    replay the coarse test (bench.mid_shift_of), checked here against the
    library's own accounting for block counts around each power of two."""
    import sys
    sys.path.insert(0, ROOT)
    import bench
    lib = R.lib()
    b = ctypes.c_int64()
    assert lib.rmq_handle_bytes(1000, 4, R.BUILD_MID, ctypes.byref(b)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_handle_bytes(1000, 4, R.BUILD_EDGES | R.BUILD_MID | 4, ctypes.byref(b)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_handle_bytes(1000, 4, R.BUILD_EDGES | R.BUILD_MID, ctypes.byref(b)) == R.RMQ_OK
    for nb in (1, 2, 8191, 8192, 8193, 16384, 16385, 97657, 1 << 20, (1 << 20) + 1, 12_500_000, (1 << 30) + 3):
        bs = 2
        n = min(nb * bs, 2 ** 31 - 1)
        nb_real = -(-n // bs)
        s = bench.mid_shift_of(nb_real)
        coarse = ((nb_real - 1) >> s) + 1
        assert coarse <= 8192 and (s == 0 or ((nb_real - 1) >> (s - 1)) + 1 > 8192)
        extra = R.handle_bytes(n, bs, R.BUILD_EDGES | R.BUILD_MID) - R.handle_bytes(n, bs, R.BUILD_EDGES)
        assert extra == 8 * coarse, (nb_real, s, extra)


def test_strerror_and_version():
    assert R.strerror(R.RMQ_OK) == "ok"
    for code in (R.RMQ_ERR_INVALID_ARG, R.RMQ_ERR_OOM, R.RMQ_ERR_HIP, R.RMQ_ERR_DEVICE, R.RMQ_ERR_BAD_QUERY):
        assert R.strerror(code) not in ("ok", "unknown error")
    assert b"gfx950" in R.lib().rmq_version()


def test_argument_validation_before_device_use():
    lib = R.lib()
    h = ctypes.c_void_p()
    fake = ctypes.c_void_p(0x1000)
    ms = ctypes.c_float()
    for args in ((None, fake, 1 << 10, 64, fake, 1), (fake, fake, 1000, 64, fake, 1), (fake, fake, 1, 64, fake, 1),
                 (fake, fake, 1 << 10, 0, fake, 1), (fake, fake, 1 << 10, 64, None, 1), (fake, fake, 1 << 10, 64, fake, 0)):
        assert lib.rmq_probe_gather(*args, None, ctypes.byref(ms)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_build(None, 10, 4, 0, None, ctypes.byref(h)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_build(fake, 0, 4, 0, None, ctypes.byref(h)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_build(fake, 10, 0, 0, None, ctypes.byref(h)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_build(fake, 1 << 31, 4, 0, None, ctypes.byref(h)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_build(fake, 10, 4, 7, None, ctypes.byref(h)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_build(ctypes.c_void_p(0x1004), 10, 4, 0, None, ctypes.byref(h)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_build(fake, 10, 4, 0, None, None) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_query(None, fake, 4, fake, fake, None) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_gpu_base(None, 10, fake, 1, fake, None, None) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_gpu_base(fake, 10, fake, 1, None, None, None) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_gen_array(None, 10, 1, 0, None) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_gen_array(fake, 10, 1, 9, None) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_query_check(fake, 1, 10, None, None) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_handle_info(None, None, None, None, None, None) == R.RMQ_ERR_INVALID_ARG
    plan = ctypes.c_void_p()
    assert lib.rmq_plan_create(None, -1, None, None, None, None, None, 0, None, ctypes.byref(plan)) == \
        R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_plan_create(None, 2, None, None, None, None, None, 0, None, ctypes.byref(plan)) == \
        R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_plan_launch(None, None) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_gen_queries(None, 5, 10, -1, 1, None) == R.RMQ_ERR_INVALID_ARG
    cfg = N.ServerConfig(-1, 0, 0, 0, 0, 0, R.QueryOpts())
    srv = ctypes.c_void_p()
    assert lib.rmq_server_create(ctypes.byref(cfg), ctypes.byref(srv)) == R.RMQ_ERR_INVALID_ARG
    cfg = N.ServerConfig(0, 0, 0, 0, 0, 2, R.QueryOpts())  # unknown server flag
    assert lib.rmq_server_create(ctypes.byref(cfg), ctypes.byref(srv)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_server_submit(None, 0, fake, 4, fake, None, None) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_server_wait(None, 1, None) == R.RMQ_ERR_INVALID_ARG
    lib.rmq_free(None)  # no-op
    lib.rmq_plan_free(None)  # no-op
    lib.rmq_server_destroy(None)  # no-op


def test_query_generator_rejects_bad_specs():
    Q = np.empty((4, 2), np.int32)
    lib = R.lib()
    for n, lr in [(0, -1), (10, 0), (10, 11), (10, -9)]:
        assert lib.rmq_gen_queries_host(Q.ctypes.data, 4, n, lr, 1, 1) == R.RMQ_ERR_INVALID_ARG


@pytest.mark.parametrize("lr", [-1, -2, -3, -4, -5, -6, -7, -8, 1, 1000])
def test_product_query_generator_matches_oracle(lr):
    a = R.gen_queries(200000, 10 ** 8, lr, 1321, nthreads=4)
    b = O.gen_queries(200000, 10 ** 8, lr, 1321, nt=3)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("kind", ["uniform", "ties16", "desc", "const"])
def test_product_array_generator_matches_oracle(kind):
    a = R.gen_array_host(100003, 42, kind)
    b = O.gen_array(100003, 42, kind)
    np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))


def test_product_path_has_no_cpu_fallback():
    """The product package never imports the oracle."""
    import pathlib
    import re
    pkg = pathlib.Path(N.PKG)
    pat = re.compile(r"(^\s*(from|import)\s+oracle)|liboracle|libhrmq|_ref/", re.M)
    for f in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.cpp")):
        assert not pat.search(f.read_text()), f"{f} reaches into oracle/"


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_compiles_and_links_for_c_and_cpp(lang, tmp_path):
    """include/rmq.h is the boundary a C or C++ host includes in place of the
    reference's rtx_rmq<T> (src/rtx.h:24-177): it must compile on its own as
    C11 and C++17, and a program linked against librtxrmq.so must run (calls
    that need no device)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc" if lang == "c" else "g++")
    if cc is None:
        pytest.skip("no host compiler")
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('''#include "rmq.h"
#include <stdio.h>
#include <string.h>
int main(void) {
    rmq_query_opts_t o = {0, 0, 0, 0};
    rmq_server_config_t c;
    memset(&c, 0, sizeof c);
    (void)o;
    if (rmq_build(0, 10, 4, 0, 0, 0) != RMQ_ERR_INVALID_ARG) return 2;
    printf("%s|%s\\n", rmq_version(), rmq_strerror(RMQ_ERR_OOM));
    return (int)sizeof(c) > 0 ? 0 : 3;
}
''')
    lib_dir = os.path.join(N.PKG, "lib")
    exe = tmp_path / "t"
    std = "-std=c11" if lang == "c" else "-std=c++17"
    r = subprocess.run([cc, std, "-Wall", "-Werror", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe),
                        f"-L{lib_dir}", "-lrtxrmq", f"-Wl,-rpath,{lib_dir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "gfx950" in r.stdout

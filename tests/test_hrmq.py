"""The host HRMQ of algs 1/101 (rmq_hrmq_*, CPU only): rightmost minimum
index and its value bits, over values compared as int32, as the reference's
rmq_rmm_par / rmq_rmm_par_idx return them (src/cpu_methods.h:80-159,
hrmq/RMQRMM64.cpp:1463-1465, src/main.cu:117-125).

Pinned by: the golden fixtures written by the reference's own HRMQ library;
a brute-force rightmost minimum over int32 bits (orc_rmq_brute_hrmq);
closed forms (constant -> r, strictly descending -> r, ascending -> l); and,
where the reference tree is present, the real HRMQ built from its sources on
the same random inputs."""
import numpy as np
import pytest

import rtxrmq_amd as R
from conftest import golden_cases, load_golden
from oracle import oracle as O


def _hrmq(A, Q, nthreads=0):
    with R.HostHRMQ(A) as h:
        return h.query(Q, nthreads=nthreads)


@pytest.mark.parametrize("name", [c for c in golden_cases() if "hrmq_idx" in load_golden(c)])
def test_golden_vectors_of_the_reference_library(name):
    g = load_golden(name)
    v, i = _hrmq(g["A"], g["Q"])
    np.testing.assert_array_equal(i, g["hrmq_idx"])
    np.testing.assert_array_equal(v.view(np.int32), g["hrmq_val"].view(np.int32))


def test_closed_forms():
    n = 300
    Q = np.array([(l, r) for l in range(0, n, 7) for r in range(l, n, 5)], np.int32)
    _, i = _hrmq(np.ones(n, np.float32), Q)
    np.testing.assert_array_equal(i, Q[:, 1])                      # constant: rightmost = r
    _, i = _hrmq((np.uint32(0x3F800000) - 1 - np.arange(n, dtype=np.uint32)).view(np.float32), Q)
    np.testing.assert_array_equal(i, Q[:, 1])                      # strictly descending: r
    _, i = _hrmq(np.arange(n, dtype=np.float32), Q)
    np.testing.assert_array_equal(i, Q[:, 0])                      # strictly ascending: l
    _, i = _hrmq(np.array([5, 5], np.float32), np.array([[0, 1]], np.int32))
    assert i[0] == 1
    _, i = _hrmq(np.array([3, 5, 3], np.float32), np.array([[0, 2]], np.int32))
    assert i[0] == 2


@pytest.mark.parametrize("kind", ["ties2", "ties16", "uniform", "normal", "desc_runs", "zigzag"])
@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 255, 256, 257, 511, 512, 513, 4097, 70001])
def test_rightmost_minimum_vs_brute_force(kind, n):
    """Sizes around the 64-bit word, the 256-bit block and the 512-bit rank
    step (n elements give 2n+2 parentheses); negative floats ('normal') are
    ordered by their int32 bits, as in the reference."""
    rng = np.random.default_rng(n * 31 + len(kind))
    if kind == "ties2":
        A = rng.integers(0, 2, n).astype(np.float32)
    elif kind == "ties16":
        A = rng.integers(0, 16, n).astype(np.float32)
    elif kind == "uniform":
        A = (rng.integers(0, 1 << 24, n) * 2.0 ** -24).astype(np.float32)
    elif kind == "normal":
        A = rng.standard_normal(n).astype(np.float32)
    elif kind == "desc_runs":
        A = (np.arange(n)[::-1] // 7).astype(np.float32)
    else:
        A = (np.arange(n) % 5).astype(np.float32)
    if n <= 65:
        Q = np.array([(l, r) for l in range(n) for r in range(l, n)], np.int32)
    else:
        l = rng.integers(0, n, 20000)
        ln = np.where(rng.random(20000) < 0.5, rng.integers(1, 600, 20000), rng.integers(1, n + 1, 20000))
        Q = np.stack([l, np.minimum(n - 1, l + ln - 1)], 1).astype(np.int32)
    v, i = _hrmq(A, Q, nthreads=3)
    want = O.rmq_brute_hrmq(A, Q)
    np.testing.assert_array_equal(i, want)
    np.testing.assert_array_equal(v.view(np.int32), A.view(np.int32)[want])


def test_thread_count_does_not_change_answers():
    A = O.gen_array(200_000, 3, "ties16")
    Q = O.gen_queries(100_000, A.size, -2, 3)
    with R.HostHRMQ(A) as h:
        v1, i1 = h.query(Q, nthreads=1)
        v8, i8 = h.query(Q, nthreads=8)
    np.testing.assert_array_equal(i1, i8)
    np.testing.assert_array_equal(v1.view(np.int32), v8.view(np.int32))


def test_invalid_queries_and_arguments():
    A = np.arange(10, dtype=np.float32)
    Q = np.array([[0, 9], [5, 4], [-1, 3], [2, 10]], np.int32)
    v, i = _hrmq(A, Q)
    assert list(i) == [0, -1, -1, -1] and np.isnan(v[1:]).all()
    lib = R.lib()
    import ctypes
    h = ctypes.c_void_p()
    assert lib.rmq_hrmq_build(None, 10, ctypes.byref(h)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_hrmq_build(A.ctypes.data, 0, ctypes.byref(h)) == R.RMQ_ERR_INVALID_ARG
    assert lib.rmq_hrmq_query(None, Q.ctypes.data, 4, None, None, 0) == R.RMQ_ERR_INVALID_ARG
    for bad in (np.arange(10, dtype=np.float64), np.arange(10, dtype=np.int16), np.zeros((2, 5), np.float32)):
        with pytest.raises(TypeError):
            R.HostHRMQ(bad)


@pytest.mark.skipif(not O.HRMQ.available(), reason="reference HRMQ not built (oracle/_ref)")
@pytest.mark.parametrize("kind,lr", [("uniform", -1), ("ties16", -2), ("uniform", -3), ("desc", -8)])
def test_same_answers_as_the_reference_library(kind, lr):
    n = 1 << 20
    A = O.gen_array(n, 11, kind)
    Q = O.gen_queries(1 << 18, n, lr, 11)
    v, i = _hrmq(A, Q)
    ref = O.HRMQ(A)
    rv, ri = ref.query(Q)
    ref.close()
    np.testing.assert_array_equal(i, ri)
    np.testing.assert_array_equal(v.view(np.int32), rv.view(np.int32))


def test_objects_alive_at_interpreter_exit(tmp_path):
    """Objects still referenced when the interpreter exits (one closed by its
    `with` block, one never closed, both in reference cycles so that only the
    last collection frees them) leave no finalizer error behind: the exit hook
    closes the open one, and finalizers that run after module globals are
    cleared do nothing."""
    import subprocess
    import sys

    from conftest import ROOT
    script = tmp_path / "exit_alive.py"
    script.write_text(
        "import sys\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "import numpy as np\n"
        "import rtxrmq_amd as R\n"
        "A = np.arange(100, dtype=np.float32)\n"
        "with R.HostHRMQ(A) as closed:\n"
        "    closed.query(np.array([[0, 99]], np.int32))\n"
        "closed.cycle = closed\n"          # collected only by the last collection, after modules are cleared
        "still_open = R.HostHRMQ(A)\n"
        "still_open.cycle = still_open\n"
        "print('done', flush=True)\n")
    p = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout.strip() == "done"
    assert "Exception ignored" not in p.stderr and "Traceback" not in p.stderr, p.stderr


def test_finalizer_after_module_globals_cleared():
    """Interpreter teardown sets a module's globals to None before the last
    objects are collected; a finalizer that runs then must not raise (the
    failure was "'NoneType' object is not callable" from RMQIndex.__del__)."""
    from rtxrmq_amd import engine as E
    h = R.HostHRMQ(np.arange(10, dtype=np.float32))
    h.close()
    saved = {k: v for k, v in vars(E).items() if not k.startswith("__")}
    try:
        for k in saved:
            setattr(E, k, None)
        h.__del__()
    finally:
        for k, v in saved.items():
            setattr(E, k, v)

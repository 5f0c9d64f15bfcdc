"""The ./rtxrmq command line: same arguments, alg ids, options and CSV rows as
the reference driver (src/main.cu:51-171, src/tools.h:28-189,268-336)."""
import os
import re
import subprocess

import numpy as np
import pytest

from rtxrmq_amd import _native as N

CLI = os.path.join(N.PKG, "bin", "rtxrmq")


ANSI = re.compile(r"\x1b\[[0-9;]*m")


def run(*args, cwd=None):
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=300, cwd=cwd)
    r.stdout = ANSI.sub("", r.stdout)
    r.stderr = ANSI.sub("", r.stderr)
    return r


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")


def test_usage_and_argument_errors():
    r = run()
    assert r.returncode == 1 and "usage: ./rtxrmq <n> <q> <lr> <alg>" in r.stderr
    assert run(1000, 10, 0, 5).returncode == 1          # lr == 0 rejected (src/tools.h:87-90)
    r = run(1000, 10, 2000, 5)                          # lr > n rejected (src/tools.h:91-94)
    assert r.returncode == 1 and "lr=2000 > n=1000" in r.stderr
    r = run(1000, 10, -1, 3)                            # RT-core algs are outside this engine
    assert r.returncode == 2 and "outside this engine" in r.stderr
    assert run(1000, 10, -1, 5, "--array", "nope").returncode == 1


@pytest.mark.parametrize("alg", [0, 100])
def test_cpu_base_writes_reference_csv_row(tmp_path, alg):
    csv = tmp_path / "t.csv"
    r = run(3000, 2000, -3, alg, "--nt", 2, "--reps", 2, "--seed", 9, f"--save-time={csv}")
    assert r.returncode == 0, r.stderr
    assert "Benchmark Finished" in r.stdout
    row = csv.read_text().strip().split(",")
    # device,alg,reps,n,bs,q,lr,t,q/s,ns/q,construction,outbuffer,tempbuffer
    assert len(row) == 13 and row[0].startswith("CPU ") and row[1] == "[CPU] BASE"
    assert row[2:7] == ["2", "3000", "0", "2000", "-3"]  # bs zeroed for non-block algs (src/tools.h:163-173)
    assert float(row[8]) > 0 and float(row[9]) > 0


@pytest.mark.parametrize("alg", [1, 101])
def test_hrmq_algs_write_reference_csv_row(tmp_path, alg):
    """Algs 1/101 (host HRMQ, src/cpu_methods.h:80-159): construction and the
    structure's size in the row's construction and outbuffer columns."""
    csv = tmp_path / "h.csv"
    r = run(200_000, 50_000, -2, alg, "--nt", 4, "--reps", 2, "--seed", 9, f"--save-time={csv}")
    assert r.returncode == 0, r.stderr
    assert "Creating MinMaxTree" in r.stdout and "Benchmark Finished" in r.stdout
    row = csv.read_text().strip().split(",")
    assert len(row) == 13 and row[0].startswith("CPU ") and row[1] == "[CPU] HRMQ"
    assert row[2:7] == ["2", "200000", "0", "50000", "-2"]
    assert float(row[8]) > 0 and float(row[10]) > 0 and float(row[11]) > 0


def test_perf_fulltest_hrmq_lines_parse(tmp_path):
    """The reference sweep's HRMQ invocations (scripts/perf-fulltest.sh:33-42
    through scripts/perf-benchmark.sh:47-51): ./rtxrmq n q lr 1 --nb 1 --reps R
    --nt N --dev D --save-time=FILE --seed S, at small sizes."""
    csv = tmp_path / "p.csv"
    for lr in (-1, -2, -3):
        r = run(1 << 16, 1 << 14, lr, 1, "--nb", 1, "--reps", 2, "--nt", 2, "--dev", 0, f"--save-time={csv}",
                "--seed", 1321)
        assert r.returncode == 0, r.stderr
    assert len(csv.read_text().strip().splitlines()) == 3


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [1, 101])
@pytest.mark.parametrize("kind", ["uniform", "ties16"])
def test_hrmq_algs_pass_check(alg, kind, gpu):
    """--check against GPU BASE: values bit for bit (alg 1); for alg 101 the
    rightmost index must hold the minimum value inside [l, r]."""
    r = run(1 << 20, 1 << 16, -2, alg, "--nt", 8, "--reps", 1, "--seed", 3, "--array", kind, "--check")
    assert r.returncode == 0 and "pass (0 mismatches)" in r.stdout, r.stdout[-1500:]


def test_reference_query_stream_option(tmp_path):
    csv = tmp_path / "r.csv"
    r = run(5000, 3000, -2, 100, "--qgen", "ref", "--nt", 4, "--reps", 1, "--seed", 11, f"--save-time={csv}")
    assert r.returncode == 0, r.stderr
    assert run(5000, 10, -1, 0, "--qgen", "bogus").returncode == 1


def test_nb_overrides_bs_and_bs_clamps_to_n(tmp_path):
    # only argument handling runs here (CPU alg); bs/nb semantics are printed
    r = run(1000, 10, 5, 0, "--nb", 7, "--nt", 1, "--reps", 1, "--seed", 1)
    assert r.returncode == 0 and "bs   = 0" in r.stdout  # CPU alg: zeroed like the reference


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [5, 105, 2, 102])
@pytest.mark.parametrize("lr", [-1, -3, 77])
def test_gpu_algs_pass_check(alg, lr, tmp_path, gpu):
    csv = tmp_path / "g.csv"
    r = run(1 << 20, 1 << 16, lr, alg, "--bs", 1000, "--reps", 3, "--seed", 1321, "--nt", 8, "--check",
            f"--save-time={csv}")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "pass (0 mismatches)" in r.stdout
    row = csv.read_text().strip().split(",")
    assert len(row) == 13 and row[1] == ("[GPU] RTX_blocks" if alg % 100 == 5 else "[GPU] BASE")
    if alg % 100 == 5:
        assert row[4] == "1000" and float(row[11]) > 0  # bs and the index size (MB)


@pytest.mark.gpu
def test_gpu_tie_arrays_and_nb(gpu):
    for kind in ("ties16", "desc", "const"):
        r = run(300_007, 50_000, -2, 105, "--nb", 97, "--reps", 1, "--seed", 4, "--array", kind, "--check")
        assert r.returncode == 0 and "pass (0 mismatches)" in r.stdout, r.stdout[-1500:]


def test_engine_options_parse_for_cpu_algs():
    r = run(3000, 100, -1, 0, "--edges", "--tune", "--nt", 1, "--reps", 1, "--seed", 1)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [5, 105])
def test_gpu_edges_and_autotune(alg, tmp_path, gpu):
    r = run(1 << 20, 1 << 16, -2, alg, "--edges", "--bs", 1000, "--reps", 2, "--seed", 5, "--check")
    assert r.returncode == 0 and "pass (0 mismatches)" in r.stdout, r.stdout[-1500:]
    assert "edge tables" in r.stdout
    csv, cache = tmp_path / "t.csv", tmp_path / "picks.txt"
    args = (1 << 20, 1 << 16, -1, alg, f"--tune={cache}", "--reps", 2, "--seed", 5, "--check", f"--save-time={csv}")
    r = run(*args)
    assert r.returncode == 0 and "pass (0 mismatches)" in r.stdout, r.stdout[-1500:]
    m = re.search(r"\[bs=(\d+) edges=(\d)", r.stdout)
    assert m and "; timed]" in r.stdout
    assert csv.read_text().strip().split(",")[4] == m.group(1)  # the CSV row carries the picked bs
    r = run(*args)
    assert r.returncode == 0 and "; cached]" in r.stdout and "pass (0 mismatches)" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [5, 105])
def test_gpu_coarse_rows(alg, gpu):
    """--edges --mid builds coarse rows (the block table of 2^14 blocks is
    read from global memory, so the query pass tries them first); answers
    pass the on-device brute-force check."""
    r = run(1 << 20, 1 << 16, -2, alg, "--edges", "--mid", "--bs", 64, "--reps", 2, "--seed", 9, "--check")
    assert r.returncode == 0 and "pass (0 mismatches)" in r.stdout, r.stdout[-1500:]
    assert "edge tables, coarse rows" in r.stdout


@pytest.mark.gpu
def test_gpu_autotune_keeps_a_given_block_size(tmp_path, gpu):
    """--tune with --nb (the reference's sweep parameter) tunes edge tables and
    thresholds at that block size only; the row and the pick carry it."""
    n, nb = 1 << 20, 64
    csv = tmp_path / "t.csv"
    r = run(n, 1 << 16, -3, 105, "--nb", nb, "--tune", "--reps", 2, "--seed", 6, "--check", f"--save-time={csv}")
    assert r.returncode == 0 and "pass (0 mismatches)" in r.stdout, r.stdout[-1500:]
    m = re.search(r"\[bs=(\d+) edges=(\d)", r.stdout)
    assert m and int(m.group(1)) == n // nb, r.stdout[-1500:]
    assert csv.read_text().strip().split(",")[4] == str(n // nb)


@pytest.mark.gpu
def test_gpu_save_power_writes_reference_columns(tmp_path, gpu):
    """--save-power samples the amdgpu hwmon sensor every 100 ms during the
    query loop into the reference's six columns (common/nvmlPower.cpp:37-84):
    timestep, power, accumulated energy, average power, dt, accumulated time;
    and prints RMQ/J as the reference's plots compute it (q * reps / energy,
    scripts/plot_rmq_per_joule.py:61-76)."""
    pfile = tmp_path / "p.dat"
    r = run(1 << 20, 1 << 16, -1, 2, "--reps", 4, "--seed", 3, f"--save-power={pfile}")
    assert r.returncode == 0, r.stderr[-1500:]
    if "no readable amdgpu power sensor" in r.stderr:
        pytest.skip("no readable power sensor on this machine")
    rows = [ln.split() for ln in pfile.read_text().splitlines() if ln.strip() and not ln.startswith("#")]
    assert len(rows) >= 2 and all(len(x) == 6 for x in rows)
    vals = np.array(rows, dtype=np.float64)
    assert np.all(vals[:, 1] > 0) and np.all(np.diff(vals[:, 2]) >= 0) and np.all(np.diff(vals[:, 5]) > 0)
    assert "power:" in r.stdout
    m = re.search(r"energy efficiency: ([0-9.e+]+) RMQ/J", r.stdout)  # q * reps / J
    assert m and float(m.group(1)) > 0


def test_save_power_hrmq_alg(tmp_path):
    """--save-power for alg 1 (HRMQ): the package energy of a fake powercap
    tree sampled into the reference's six columns, and RMQ/J printed, as the
    reference reads RAPL around rmq_rmm_par (src/cpu_methods.h:100-102)."""
    import threading
    import time

    root = tmp_path / "powercap"
    pkg = root / "intel-rapl:0"
    pkg.mkdir(parents=True)
    (pkg / "name").write_text("package-0\n")
    (pkg / "energy_uj").write_text("0\n")
    (pkg / "max_energy_range_uj").write_text("1000000000\n")
    stop = threading.Event()

    def tick():
        e = 0
        while not stop.is_set():
            time.sleep(0.05)
            e += 500_000
            tmp = pkg / "energy_uj.tmp"
            tmp.write_text(f"{e}\n")
            os.replace(tmp, pkg / "energy_uj")

    th = threading.Thread(target=tick)
    th.start()
    pfile = tmp_path / "hrmq.dat"
    try:
        env = dict(os.environ, RTXRMQ_POWERCAP=str(root))
        r = subprocess.run([CLI, "1000000", "1000000", "-1", "1", "--nt", "1", "--reps", "12", "--seed", "3",
                            f"--save-power={pfile}"], capture_output=True, text=True, timeout=300, env=env)
    finally:
        stop.set()
        th.join()
    out = ANSI.sub("", r.stdout)
    assert r.returncode == 0, r.stderr[-1500:]
    rows = [ln.split() for ln in pfile.read_text().splitlines() if ln.strip() and not ln.startswith("#")]
    assert len(rows) >= 1 and all(len(x) == 6 for x in rows)
    assert "RMQ/J" in out


def test_save_power_cpu_alg_reads_powercap_packages(tmp_path):
    """CPU algs sample the CPU packages' energy every 500 ms, as the reference
    samples RAPL around its CPU methods (common/nvmlPower.cpp:216-274), into
    the same six columns.  A fake powercap tree: a package zone whose counter
    a thread advances by 0.5 J every 50 ms and wraps at 30 J (after 3 s), and a non-package
    zone that must be ignored."""
    import threading
    import time

    root = tmp_path / "powercap"
    pkg, core = root / "intel-rapl:0", root / "intel-rapl:0:0"
    for z, name in ((pkg, "package-0"), (core, "core")):
        z.mkdir(parents=True)
        (z / "name").write_text(name + "\n")
        (z / "energy_uj").write_text("0\n")
        (z / "max_energy_range_uj").write_text("30000000\n")
    stop = threading.Event()

    def tick():
        e = 0
        while not stop.is_set():
            time.sleep(0.05)
            e = (e + 500_000) % 30_000_000
            tmp = pkg / "energy_uj.tmp"
            tmp.write_text(f"{e}\n")
            os.replace(tmp, pkg / "energy_uj")
            (core / "energy_uj").write_text(f"{e * 7}\n")

    th = threading.Thread(target=tick)
    th.start()
    pfile = tmp_path / "cpu.dat"
    try:
        env = dict(os.environ, RTXRMQ_POWERCAP=str(root))
        r = subprocess.run([CLI, "200000", "50000", "-2", "0", "--nt", "2", "--reps", "5", "--seed", "3",
                            f"--save-power={pfile}"], capture_output=True, text=True, timeout=300, env=env)
    finally:
        stop.set()
        th.join()
    out = ANSI.sub("", r.stdout)
    assert r.returncode == 0, r.stderr[-1500:]
    rows = [ln.split() for ln in pfile.read_text().splitlines() if ln.strip() and not ln.startswith("#")]
    assert len(rows) >= 4 and all(len(x) == 6 for x in rows)
    vals = np.array(rows, dtype=np.float64)
    assert vals[-1, 5] > 3.0  # the counter wrapped at least once meanwhile
    assert np.all(np.diff(vals[:, 2]) >= 0) and np.all(vals[:, 1] >= 0)
    # about 10 J/s from the one package zone (the core zone ignored), wraps included
    rate = vals[-1, 2] / vals[-1, 5]
    assert 5.0 < rate < 15.0, (rate, vals[-3:])
    assert "CPU power:" in out and "(1 package)" in out
    m = re.search(r"energy efficiency: ([0-9.e+]+) RMQ/J", out)
    assert m and float(m.group(1)) > 0

"""The command-line counterparts of the reference's benchmarks: clo_rng_bench
(option checks on the CPU; every output format against the oracle's streams
on the GPU), clo_sort_bench and clo_scan_bench (report lines, file output,
the sortedness / serial-scan checks)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cl-ops_amd", "bin", "clo_rng_bench")


def _run(args, cwd=None, timeout=60):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, timeout=timeout)


def test_cli_built():
    assert os.access(CLI, os.X_OK), "build the library first (make -C cl-ops_amd)"
    r = _run(["--help"])
    assert r.returncode == 0 and b"file-tsv, file-dh, stdout-bin, stdout-uint" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["-o", "foo"], b"Error: Unknown output 'foo'."),
    (["--bits=33"], b"Error: Number of bits must be between 1 and 32."),
    (["-b", "0"], b"Error: Number of bits must be between 1 and 32."),
    (["-n", "0"], b"Error: Continuous generation can only be performed to stdout."),
    (["-n0", "-ofile-dh"], b"Error: Continuous generation can only be performed to stdout."),
    (["--frobnicate"], b"Error: Unknown option --frobnicate"),
    (["-g", "x"], b"Error: Cannot parse integer value 'x' for -g"),
])
def test_cli_argument_errors(args, msg):
    """clo_rng_bench.c:146-158: the checks and their messages; exit status is
    the error code (CLO_ERROR_ARGS = 2)."""
    r = _run(args)
    assert r.returncode == 2 and msg in r.stderr


def _stream(oracle, gen, gws, runs, seed, hash_id=None, bits=32, maxint=0):
    """Run-major numbers: run r holds draw r of work-items 0..gws-1."""
    if hash_id is None:
        st = oracle.rng_init_host_mt(gen, gws, seed)
    else:
        st = oracle.rng_init_dev_gid(gen, gws, seed, hash_id)
    return oracle.rng_fill(gen, st, runs, bits, maxint).reshape(runs, gws)


@pytest.mark.gpu
def test_cli_file_tsv(oracle, tmp_path):
    r = _run(["-r", "xorshift128", "-o", "file-tsv", "-g", "1000", "-n", "3", "-s", "42"],
             cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert b"Random number generator (seed): xorshift128 (42)" in r.stdout
    want = _stream(oracle, "xorshift128", 1000, 3, 42)
    text = "".join("".join("%d\t" % v for v in row) + "\n" for row in want)
    assert (tmp_path / "out_xorshift128_host_mt.tsv").read_text() == text


@pytest.mark.gpu
def test_cli_file_dh_gid_hash(oracle, tmp_path):
    r = _run(["--rng=lcg", "--output=file-dh", "--globalsize=777", "--runs=5",
              "--gid-hash=KNUTH(x)", "--bits=24"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    want = _stream(oracle, "lcg", 777, 5, 0, hash_id=1, bits=24)
    text = "type: d\ncount: %d\nnumbit: 24\n" % (777 * 5) + "".join(
        "%d\n" % v for v in want.ravel())
    assert (tmp_path / "out_lcg_gid_KNUTH(x).dh.txt").read_text() == text


@pytest.mark.gpu
@pytest.mark.parametrize("gen", ["mwc64x", "parkmiller", "tauslcg", "xorshift64"])
def test_cli_stdout_uint_and_bin(oracle, gen):
    r = _run(["-r", gen, "-o", "stdout-uint", "-g", "2048", "-n", "4", "-s", "7", "-b", "17"])
    assert r.returncode == 0, r.stderr
    want = _stream(oracle, gen, 2048, 4, 7, bits=17)
    assert r.stdout == "".join("%d\n" % v for v in want.ravel()).encode()
    r = _run(["-r", gen, "-o", "stdout-bin", "-g", "2048", "-n", "4", "-s", "7", "-m", "1000"])
    assert r.returncode == 0, r.stderr
    want = _stream(oracle, gen, 2048, 4, 7, maxint=1000)
    assert np.array_equal(np.frombuffer(r.stdout, dtype=np.uint32), want.ravel())


@pytest.mark.gpu
def test_cli_continuous_stdout(oracle):
    """runs = 0: generation continues until the reader closes the stream."""
    gws, nbytes = 4096, 4096 * 4 * 37
    p = subprocess.Popen([CLI, "-r", "xorshift128", "-o", "stdout-bin", "-g", str(gws), "-n",
                          "0", "-s", "3"], stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    try:
        data = b""
        while len(data) < nbytes:
            chunk = p.stdout.read(nbytes - len(data))
            assert chunk, p.stderr.read()
            data += chunk
    finally:
        p.stdout.close()
        rc = p.wait(timeout=60)
    assert rc == 0
    want = _stream(oracle, "xorshift128", gws, 37, 3)
    assert np.array_equal(np.frombuffer(data, dtype=np.uint32), want.ravel())


# ---- clo_sort_bench / clo_scan_bench -----------------------------------------

SORT_CLI = os.path.join(ROOT, "cl-ops_amd", "bin", "clo_sort_bench")
SCAN_CLI = os.path.join(ROOT, "cl-ops_amd", "bin", "clo_scan_bench")


def _run_cli(cli, args, cwd=None, timeout=120):
    return subprocess.run([cli] + args, cwd=cwd, capture_output=True, timeout=timeout)


def test_bench_clis_built_and_check_arguments():
    for cli in (SORT_CLI, SCAN_CLI):
        r = _run_cli(cli, ["--help"])
        assert r.returncode == 0 and r.stdout.startswith(b"Usage: ")
    r = _run_cli(SORT_CLI, ["-t", "foo"])
    assert r.returncode == 6 and b"Error: Unknown type 'foo'" in r.stderr  # CLO_ERROR_UNKNOWN_TYPE
    r = _run_cli(SCAN_CLI, ["--bogus"])
    assert r.returncode == 2


def _rates(stdout, pattern):
    import re
    return re.findall(pattern, stdout.decode())


@pytest.mark.gpu
@pytest.mark.parametrize("algo,typ,ok", [
    ("satradix", "ulong", True), ("abitonic", "uint", True), ("gselect", "short", True),
    ("sbitonic", "float", True), ("abitonic", "double", True),
    ("satradix", "int", False),  # unsigned key order, as the reference: a signed check fails
])
def test_sort_bench_cli(tmp_path, algo, typ, ok):
    maxpo2 = 12 if algo != "gselect" else 10
    r = _run_cli(SORT_CLI, ["-a", algo, "-t", typ, "-n", str(maxpo2), "-r", "2", "-o", "b.tsv"],
                 cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert b"Type of elements to sort: %s" % typ.encode() in r.stdout
    lines = _rates(r.stdout, r"- 2\^(\d+): ([0-9.]+) Mkeys/s (.*)")
    assert [int(n) for n, _, _ in lines] == list(range(4, maxpo2 + 1))
    assert all(float(v) > 0 for _, v, _ in lines)
    if ok:
        assert all(msg.strip() == "" for _, _, msg in lines)
    else:  # 2^4 random ints include negatives: they sort after the positives
        assert any("(sort did not work)" in msg for _, _, msg in lines)
    rows = (tmp_path / "b.tsv").read_text().splitlines()
    assert len(rows) == maxpo2 and all(len(row.split("\t")) == 3 for row in rows)
    assert int(rows[maxpo2 - 1].split("\t")[1]) > 0 and rows[0] == "0\t0\t0"


@pytest.mark.gpu
def test_scan_bench_cli(tmp_path):
    r = _run_cli(SCAN_CLI, ["-n", "12", "-r", "2", "-o", "s.tsv"], cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    lines = _rates(r.stdout, r"- +(\d+) : ([0-9.]+) MValues/s (.*)")
    assert [int(n) for n, _, _ in lines] == [4 << k for k in range(12)]
    assert all(msg.strip() == "" for _, _, msg in lines)
    rows = (tmp_path / "s.tsv").read_text().splitlines()
    assert len(rows) == 12 and all(len(row.split("\t")) == 3 for row in rows)
    # a uchar sum overflows past a few elements; --no-check skips the check
    r = _run_cli(SCAN_CLI, ["-t", "uchar", "-y", "uchar", "-i", "64", "-n", "3"], cwd=tmp_path)
    assert r.returncode == 0 and all("[Overflow]" in m for _, _, m in
                                     _rates(r.stdout, r"- +(\d+) : ([0-9.]+) MValues/s (.*)"))
    r = _run_cli(SCAN_CLI, ["-u", "-n", "3", "-t", "int", "-y", "long"], cwd=tmp_path)
    assert r.returncode == 0 and all("[Unverified]" in m for _, _, m in
                                     _rates(r.stdout, r"- +(\d+) : ([0-9.]+) MValues/s (.*)"))


# ---- a C99 caller of the public headers ----------------------------------------

C_CALLER = os.path.join(ROOT, "cl-ops_amd", "bin", "c_caller")


def test_c_caller_built():
    """Built by gcc -std=c99 -pedantic -Werror against include/ alone."""
    assert os.access(C_CALLER, os.X_OK)


@pytest.mark.gpu
def test_c_caller_runs():
    r = subprocess.run([C_CALLER], capture_output=True, timeout=120)
    assert r.returncode == 0 and r.stdout == b"c_caller: ok\n", r.stderr

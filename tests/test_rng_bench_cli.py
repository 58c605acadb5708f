"""clo_rng_bench, the command-line counterpart of the reference's RNG bench:
option checks (CPU) and every output format against the oracle's streams
(GPU)."""
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

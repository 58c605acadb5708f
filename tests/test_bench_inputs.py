"""The benchmark programs' input generator (cl-ops_amd/cli/grand.hpp), which
restates GLib's GRand and the reference's clo_bench_rand
(src/benchmarks/clo_bench.c:67-142), against GLib's own outputs recorded in
tests/golden/glib_grand.json (tests/golden/make_glib_grand.py calls the
system libglib).  A small driver (tests/golden/grand_driver.cpp) is built
with g++ and prints the header's draws."""
import json
import os
import shutil
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("grand") / "grand_driver")
    subprocess.run([gxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
                    "-I" + os.path.join(ROOT, "cl-ops_amd", "cli"),
                    "-I" + os.path.join(ROOT, "include"),
                    os.path.join(GOLDEN, "grand_driver.cpp"), "-o", exe], check=True)

    def run(*args):
        r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, check=True,
                           timeout=30)
        return r.stdout.decode().split()
    return run


@pytest.fixture(scope="module")
def fixture():
    with open(os.path.join(GOLDEN, "glib_grand.json")) as f:
        return json.load(f)["seeds"]


def test_words_and_doubles_match_glib(driver, fixture):
    for seed, rec in fixture.items():
        words = [int(x) for x in driver(seed, "ints", 1000)]
        assert words[:32] == rec["first32"] and words[999] == rec["word1000"], seed
        got = [float.fromhex(x) for x in driver(seed, "doubles", 8)]
        assert got == [float.fromhex(x) for x in rec["double_first8"]], seed


def test_int_range_matches_glib(driver, fixture):
    for seed, rec in fixture.items():
        for key, want in rec["int_range"].items():
            b, e = key.split(",")
            assert [int(x) for x in driver(seed, "range", b, e, 8)] == want, (seed, key)


def test_double_range_matches_glib(driver, fixture):
    for seed, rec in fixture.items():
        for key, want in rec["double_range"].items():
            b, e = (float.fromhex(x) for x in key.split(","))
            got = driver(seed, "drange", repr(b), repr(e), 8)
            assert [float.fromhex(x) for x in got] == [float.fromhex(x) for x in want], \
                (seed, key)


def _bench(driver, seed, t, n, fmt):
    return [struct.unpack("<" + fmt, bytes.fromhex(x))[0] for x in driver(seed, "bench", t, n)]


def test_bench_rand_types(driver, fixture):
    """clo_bench_rand per type from the GLib draws of the fixture: integer
    types take g_rand_int_range over the type's range; uint and ulong scale
    g_rand_double by the type's maximum and truncate; float and double take
    g_rand_double_range over (MIN, MAX) of the type."""
    for seed, rec in fixture.items():
        ir, dr = rec["int_range"], rec["double_range"]
        assert _bench(driver, seed, 0, 8, "b") == ir["-128,127"]
        assert _bench(driver, seed, 1, 8, "B") == ir["0,255"]
        assert _bench(driver, seed, 2, 8, "h") == ir["-32768,32767"]
        assert _bench(driver, seed, 3, 8, "H") == ir["0,65535"]
        assert _bench(driver, seed, 4, 8, "i") == ir["-2147483648,2147483647"]
        d = [float.fromhex(x) for x in rec["double_first8"]]
        assert _bench(driver, seed, 5, 4, "I") == [int(x * 4294967295.0) for x in d[:4]]
        assert _bench(driver, seed, 7, 4, "Q") == [int(x * 2.0 ** 64) for x in d[:4]]
        fkey = "0x1.0000000000000p-126,0x1.fffffe0000000p+127"
        want_f = [struct.unpack("<f", struct.pack("<f", float.fromhex(x)))[0]
                  for x in dr[fkey]]
        assert _bench(driver, seed, 9, 8, "f") == want_f
        dkey = "0x1.0000000000000p-1022,0x1.fffffffffffffp+1023"
        assert _bench(driver, seed, 10, 8, "d") == [float.fromhex(x) for x in dr[dkey]]
        hkey = "-0x1.0000000000000p+10,0x1.0000000000000p+10"
        assert _bench(driver, seed, 8, 8, "H") == [int(float.fromhex(x)) & 0xFFFF
                                                   for x in dr[hkey]]


def test_bench_rand_long_sign_and_magnitude(driver, fixture):
    """long: |value| = trunc(g_rand_double * 2^63) from the first two words,
    sign from bit 15 of the third (the build draws the double first)."""
    for seed, rec in fixture.items():
        w = rec["first32"]
        v = _bench(driver, seed, 6, 1, "q")[0]
        r = (w[0] * 2.0 ** -32 + w[1]) * 2.0 ** -32
        neg = bool(w[2] & (1 << 15))
        assert v == (-int(r * 2.0 ** 63) if neg else int(r * 2.0 ** 63)), seed

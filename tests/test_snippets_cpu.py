"""Sorts with user key / compare snippets, CPU side: the oracle's generic
bitonic network and gselect rank pinned by independent results, and the
run-time kernel text compiled for gfx950 with hipRTC for every element type
and representative snippets."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RTC = os.path.join(ROOT, "cl-ops_amd", "csrc", "snippet_kernels.rtc")


# ---- oracle pins ------------------------------------------------------------

@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint32, np.int64])
def test_network_matches_c_restatement(oracle, dtype):
    """Identity key, ascending and descending compares: the generic network
    equals the C restatement of sbitonic (a separate implementation)."""
    rs = np.random.default_rng(3)
    for n in (2, 8, 64, 1024, 4096):
        info = np.iinfo(dtype)
        x = rs.integers(info.min, info.max, n, endpoint=True).astype(dtype)
        x[: n // 4] = x[n // 2]  # duplicates
        assert np.array_equal(oracle.bitonic_network(x), oracle.sbitonic(x))
        assert np.array_equal(oracle.bitonic_network(x, compare=np.less),
                              oracle.sbitonic(x, desc=True))


def test_network_tie_order_hand_derived(oracle):
    """All keys equal: COMPARE is false, so a comparator swaps exactly when its
    direction is descending.  Worked by hand from sbitonic.cl for n = 4
    (stage 1 swaps pair (2,3); stage 2 swaps nothing) and n = 8 (stage 1
    swaps (2,3) and (6,7); stage 2 swaps (4,6), (5,7), then (4,5), (6,7);
    stage 3 swaps nothing)."""
    zero = lambda x: np.zeros(x.shape, dtype=np.uint8)  # noqa: E731
    assert oracle.bitonic_network(np.arange(4), key=zero).tolist() == [0, 1, 3, 2]
    assert oracle.bitonic_network(np.arange(8), key=zero).tolist() == [0, 1, 3, 2, 6, 7, 5, 4]


def test_network_brute_force_sorts(oracle):
    """Every 0/1 input of length 16 (the 0-1 principle) comes out sorted, with
    the payload carried along: the network is a sorting network with this
    comparator rule."""
    n = 16
    for bits in range(1 << n):
        x = ((np.uint32(bits) >> np.arange(n, dtype=np.uint32)) & 1).astype(np.uint32)
        y = oracle.bitonic_network((x << 8) | np.arange(n, dtype=np.uint32),
                                   key=lambda v: v >> 8)
        assert np.all(np.diff((y >> 8).astype(np.int64)) >= 0)
        assert sorted(y & 0xFF) == list(range(n))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 100, 1000, 3000])
def test_padded_network(oracle, n):
    """Other sizes: sorted by key, a permutation, and for distinct keys the
    plain sort; at a power of two the padded network is the reference's."""
    rs = np.random.default_rng(n)
    x = rs.permutation(n * 7)[:n].astype(np.uint64) << np.uint64(20)
    x |= rs.integers(0, 1 << 20, n).astype(np.uint64)
    hi = lambda v: (v >> np.uint64(40)).astype(np.uint32)  # noqa: E731
    y = oracle.bitonic_network(x, key=hi, padded=True)
    assert np.all(np.diff(hi(y).astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(y), np.sort(x))
    assert np.array_equal(oracle.bitonic_network(x, padded=True), np.sort(x))
    if n & (n - 1) == 0:
        assert np.array_equal(y, oracle.bitonic_network(x, key=hi))


def test_gselect_rank_is_stable_sort(oracle):
    """For a strict order whose ties are key equality, the rank formula is
    the stable sort: Python's sorted() (stable) gives the same order."""
    rs = np.random.default_rng(9)
    x = rs.integers(0, 1 << 62, 2000, dtype=np.uint64)
    v = np.arange(x.size, dtype=np.uint32)
    key = lambda a: (a & np.uint64(0xFF)).astype(np.uint8)  # noqa: E731
    for compare, rev in ((np.greater, False), (np.less, True)):
        got, gv = oracle.gselect_rank(x, v, key=key, compare=compare)
        # sorted() keeps equal keys in input order, also with reverse=True
        order = sorted(range(x.size), key=lambda i: int(x[i]) & 0xFF, reverse=rev)
        assert np.array_equal(gv, np.array(order, dtype=np.uint32))
        assert np.array_equal(got, x[order])


def test_gselect_rank_matches_c_restatement(oracle):
    rs = np.random.default_rng(10)
    k = rs.integers(-50, 50, 3000).astype(np.int32)
    v = rs.integers(0, 1 << 32, k.size, dtype=np.uint64).astype(np.uint32)
    for desc in (False, True):
        got = oracle.gselect_rank(k, v, compare=np.less if desc else np.greater)
        want = oracle.gselect(k, v, desc=desc)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_gselect_rank_rejects_non_strict_compare(oracle):
    with pytest.raises(ValueError):
        oracle.gselect_rank(np.arange(10, dtype=np.uint32), compare=np.greater_equal)


# ---- the run-time kernels compile for gfx950 ---------------------------------

_TYPES = {"char": 1, "uchar": 1, "short": 2, "ushort": 2, "int": 4, "uint": 4, "long": 8,
          "ulong": 8, "half": 2, "float": 4, "double": 8}
_BITS = {1: "uchar", 2: "ushort", 4: "uint", 8: "ulong"}


def _compile(src, opts=()):
    rtc = ctypes.CDLL("libhiprtc.so")
    prog = ctypes.c_void_p()
    assert rtc.hiprtcCreateProgram(ctypes.byref(prog), src.encode(), b"s.hip", 0, None,
                                   None) == 0
    args = [b"-O3", b"--offload-arch=gfx950", b"-Xclang", b"-target-feature", b"-Xclang",
            b"-packed-fp32-ops"] + [o.encode() for o in opts]
    arr = (ctypes.c_char_p * len(args))(*args)
    rc = rtc.hiprtcCompileProgram(prog, len(args), arr)
    sz = ctypes.c_size_t()
    rtc.hiprtcGetProgramLogSize(prog, ctypes.byref(sz))
    log = ctypes.create_string_buffer(sz.value + 1)
    rtc.hiprtcGetProgramLog(prog, log)
    code = ctypes.c_size_t()
    rtc.hiprtcGetCodeSize(prog, ctypes.byref(code))
    rtc.hiprtcDestroyProgram(ctypes.byref(prog))
    return rc, log.value.decode(errors="replace"), code.value


def _source(elem, key, compare="((a) > (b))", get_key="(x)"):
    # The prelude the library writes before the kernel text.
    return ("#define CLO_SORT_ELEM_TYPE %s\n#define CLO_SORT_KEY_TYPE %s\n"
            "#define CLO_SORT_COMPARE(a, b) %s\n#define CLO_SORT_KEY_GET(x) %s\n"
            "#define CLO_SNIP_KEY_BITS %s\n" % (elem, key, compare, get_key, _BITS[_TYPES[key]])
            + open(RTC).read())


@pytest.mark.parametrize("elem", sorted(_TYPES))
def test_kernels_compile_every_type(elem):
    rc, log, size = _compile(_source(elem, elem))
    assert rc == 0 and size > 0, log


@pytest.mark.parametrize("elem,key,compare,get_key,opts", [
    ("ulong", "uint", "((a) > (b))", "((uint)((x) >> 32))", ()),
    ("float", "uint", "((a) > (b))", "(as_uint(x) ^ ((uint)(as_int(x) >> 31) | 0x80000000u))",
     ()),
    ("double", "double", "((a) < (b))", "(x)", ()),
    ("ulong", "ushort", "((a) > (b))", "((ushort)((x) >> SHIFT))", ("-DSHIFT=48",)),
    ("int", "int", "(abs(a) > abs(b))", "(x)", ()),
])
def test_kernels_compile_snippets(elem, key, compare, get_key, opts):
    rc, log, size = _compile(_source(elem, key, compare, get_key), opts)
    assert rc == 0 and size > 0, log


def test_bad_snippet_reports_log():
    rc, log, _ = _compile(_source("uint", "uint", compare="((a).x > (b).x)"))
    assert rc != 0 and "error" in log


def test_embedded_text_is_the_file():
    """The library carries snippet_kernels.rtc verbatim (generated include)."""
    inc = os.path.join(ROOT, "cl-ops_amd", "build", "snippet_kernels.inc")
    if not os.path.exists(inc):
        pytest.skip("library not built")
    body = "".join(
        bytes(m, "utf-8").decode("unicode_escape") + "\n"
        for m in re.findall(r'^"(.*)\\n"$', open(inc).read(), re.M))
    assert body == "".join(line + "\n" for line in open(RTC).read().splitlines())

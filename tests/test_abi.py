"""C-ABI checks that need no GPU: libcl_ops.so loads, exports every function
and data symbol the public headers declare, and its host-only helpers behave
as CL_Ops' src/cl_ops/common/clo_common.c:54-223."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def _declared_functions():
    names = set()
    for fn in sorted(os.listdir(INC)):
        if not fn.endswith(".h") or fn == "clo_rng_device.h":
            continue
        text = open(os.path.join(INC, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        # drop struct bodies (function-pointer fields are not exports)
        text = re.sub(r"typedef struct [^{;]*\{.*?\}[^;]*;", "", text, flags=re.S)
        text = re.sub(r"struct [a-z_]+ \{.*?\};", "", text, flags=re.S)
        for m in re.finditer(r"\b((?:clo|ccl)_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
        for m in re.finditer(r"extern const [A-Za-z_ ]+?\b((?:clo|ccl)_[a-z0-9_]+)(\[\])?;", text):
            names.add(m.group(1))
    return names


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "cl-ops_amd", "lib",
                                                                    "libcl_ops.so")],
                         check=True, capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_loads(clops):
    assert b"gfx950" in clops.lib().clo_util_build_info()


def test_every_declared_symbol_is_exported(clops):
    declared = _declared_functions()
    assert len(declared) > 60
    missing = sorted(declared - _exported())
    assert not missing, missing


def test_vtable_data_symbols(clops):
    exp = _exported()
    for sym in ("clo_sort_satradix_def", "clo_sort_sbitonic_def", "clo_sort_abitonic_def",
                "clo_sort_gselect_def", "clo_scan_blelloch_def", "clo_rng_infos"):
        assert sym in exp


def test_rng_infos_table(clops):
    class Info(ctypes.Structure):
        _fields_ = [("name", ctypes.c_char_p), ("src", ctypes.c_char_p),
                    ("seed_size", ctypes.c_size_t)]

    # the reference's table, in its order (src/cl_ops/rng/clo_rng.c:60-68)
    tab = (Info * 7).in_dll(clops.lib(), "clo_rng_infos")
    got = [(tab[i].name, tab[i].seed_size) for i in range(7)]
    assert got == [(b"lcg", 8), (b"xorshift64", 8), (b"xorshift128", 16), (b"mwc64x", 8),
                   (b"parkmiller", 4), (b"tauslcg", 16), (None, 0)]
    assert b"namespace clo_rng_xorshift128" in tab[2].src
    assert b"namespace clo_rng_tauslcg" in tab[5].src


def test_type_table(clops):
    L = clops.lib()
    names = ["char", "uchar", "short", "ushort", "int", "uint", "long", "ulong", "half",
             "float", "double"]
    sizes = [1, 1, 2, 2, 4, 4, 8, 8, 2, 4, 8]
    for i, (nm, sz) in enumerate(zip(names, sizes)):
        assert L.clo_type_get_name(i) == nm.encode()
        assert L.clo_type_sizeof(i) == sz
        assert clops._call(L.clo_type_by_name, nm.encode()) == i
    with pytest.raises(clops.ClopsError) as e:
        clops._call(L.clo_type_by_name, b"quad")
    assert e.value.code == clops.CLO_ERROR_UNKNOWN_TYPE
    assert e.value.domain == L.clo_error_quark()


def _ref_nlpo2(x):
    # clo_common.c:141-151 on 32-bit unsigned arithmetic
    if (x & (x - 1)) & 0xFFFFFFFF == 0:
        return x
    for s in (1, 2, 4, 8, 16):
        x |= x >> s
    return (x + 1) & 0xFFFFFFFF


def test_bit_utilities(clops):
    L = clops.lib()
    vals = [0, 1, 2, 3, 5, 255, 256, 257, 1000, (1 << 31) - 1, 1 << 31, (1 << 31) + 1,
            0xFFFFFFFF]
    for x in vals:
        assert L.clo_nlpo2(x) == _ref_nlpo2(x)
        assert L.clo_ones32(x) == bin(x).count("1")
    for x in [1, 2, 8, 16, 256, 1 << 20, 12, 40]:
        assert L.clo_tzc(x) == (x & -x).bit_length() - 1
    for x in [0, 1, 10, 1000, 65535]:
        assert L.clo_sum(x) == (x * (x + 1) // 2) & 0xFFFFFFFF


def test_no_context_without_device(clops):
    # With no GPU, creating a context fails with an error, not a crash.
    import numpy as np  # noqa: F401
    L = clops.lib()
    try:
        ctx = clops._call(L.ccl_context_new_from_device_index, 4096)
    except clops.ClopsError as e:
        assert e.domain == L.ccl_error_quark()
    else:
        raise AssertionError("device 4096 should not exist: %r" % ctx)


def test_log_level_setter(clops):
    L = clops.lib()
    L.clo_set_log_level.argtypes = [ctypes.c_int]
    L.clo_get_log_level.restype = ctypes.c_int
    L.clo_set_log_level(2)
    assert L.clo_get_log_level() == 2
    L.clo_set_log_level(0)
    assert L.clo_get_log_level() == 0
    L.clo_set_log_level(-1)  # back to the CLO_LOG environment variable
    assert L.clo_get_log_level() == int(os.environ.get("CLO_LOG", "0") or 0)

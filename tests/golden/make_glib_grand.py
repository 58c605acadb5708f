"""Writes glib_grand.json: words of GLib's GRand for a few seeds.

The reference seeds HOST_MT generators with GRand (src/cl_ops/rng/clo_rng.c:
196-203: g_rand_new_with_seed((guint32)main_seed), then g_rand_int per 32-bit
seed word) and draws its benchmark inputs with g_rand_double
(src/benchmarks/clo_bench.c:105-117).  This script calls the system's
libglib-2.0 through ctypes (the GLib runtime library installed on the
machine; nothing from the reference is run) and records, each from a fresh
generator per seed:
  - the first 32 g_rand_int words and word #1000;
  - the first 8 g_rand_double values;
  - 8 g_rand_int_range values for each range the benchmarks use (and two
    small ones that exercise the rejection of high words);
  - 8 g_rand_double_range values for each range the benchmarks use.
Usage: python tests/golden/make_glib_grand.py  (rewrites the fixture).
"""
import ctypes
import json
import os

SEEDS = [0, 1, 1234, 5489, 0xDEADBEEF]
# clo_bench.c:67-142 ranges (char, uchar, short, ushort, int) and two small
# ones whose rejection threshold is below 2^32 - 1.
INT_RANGES = [(-128, 127), (0, 255), (-32768, 32767), (0, 65535),
              (-2147483648, 2147483647), (0, 3), (5, 1000)]
# float (FLT_MIN, FLT_MAX), double (DBL_MIN, DBL_MAX), half (-1024, 1024).
DOUBLE_RANGES = [(float.fromhex("0x1p-126"), float.fromhex("0x1.fffffep+127")),
                 (float.fromhex("0x1p-1022"), float.fromhex("0x1.fffffffffffffp+1023")),
                 (-1024.0, 1024.0)]


def main():
    g = ctypes.CDLL("libglib-2.0.so.0")
    g.g_rand_new_with_seed.restype = ctypes.c_void_p
    g.g_rand_new_with_seed.argtypes = [ctypes.c_uint32]
    g.g_rand_int.restype = ctypes.c_uint32
    g.g_rand_int.argtypes = [ctypes.c_void_p]
    g.g_rand_double.restype = ctypes.c_double
    g.g_rand_double.argtypes = [ctypes.c_void_p]
    g.g_rand_free.argtypes = [ctypes.c_void_p]
    g.g_rand_int_range.restype = ctypes.c_int32
    g.g_rand_int_range.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    g.g_rand_double_range.restype = ctypes.c_double
    g.g_rand_double_range.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    ver = "%d.%d.%d" % tuple(ctypes.c_uint.in_dll(g, n).value
                             for n in ("glib_major_version", "glib_minor_version",
                                       "glib_micro_version"))
    out = {"glib_version": ver, "seeds": {}}
    for s in SEEDS:
        r = g.g_rand_new_with_seed(s)
        words = [g.g_rand_int(r) for _ in range(1000)]
        g.g_rand_free(r)
        r = g.g_rand_new_with_seed(s)
        doubles = [g.g_rand_double(r) for _ in range(8)]
        g.g_rand_free(r)
        ranges = {}
        for b, e in INT_RANGES:
            r = g.g_rand_new_with_seed(s)
            ranges["%d,%d" % (b, e)] = [g.g_rand_int_range(r, b, e) for _ in range(8)]
            g.g_rand_free(r)
        dranges = {}
        for b, e in DOUBLE_RANGES:
            r = g.g_rand_new_with_seed(s)
            dranges["%s,%s" % (b.hex(), e.hex())] = [g.g_rand_double_range(r, b, e).hex()
                                                     for _ in range(8)]
            g.g_rand_free(r)
        out["seeds"][str(s)] = {"first32": words[:32], "word1000": words[999],
                                "double_first8": [d.hex() for d in doubles],
                                "int_range": ranges, "double_range": dranges}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "glib_grand.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, "GLib", ver)


if __name__ == "__main__":
    main()

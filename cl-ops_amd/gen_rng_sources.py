"""Build step: extract the generator sections of include/clo_rng_device.h into
C string literals (rng_sources.inc) so that clo_rng_get_source() returns the
same text the library's own kernels are compiled from."""
import re
import sys

NAMES = {"lcg": "kSrcLcg", "xorshift64": "kSrcXorshift64", "xorshift128": "kSrcXorshift128",
         "mwc64x": "kSrcMwc64x", "parkmiller": "kSrcParkmiller", "tauslcg": "kSrcTauslcg",
         "api": "kSrcApi"}


def main(header, out):
    text = open(header).read()
    sections = dict(re.findall(r"/\* @@section (\w+) \*/\n(.*?)/\* @@end \*/", text, re.S))
    missing = set(NAMES) - set(sections)
    if missing:
        raise SystemExit("missing sections: %s" % sorted(missing))
    with open(out, "w") as f:
        f.write("// Generated from include/clo_rng_device.h; do not edit.\n")
        for key, var in NAMES.items():
            body = sections[key]
            lit = "".join('"%s\\n"\n' % line.replace("\\", "\\\\").replace('"', '\\"')
                          for line in body.splitlines())
            f.write("static const char %s[] =\n%s;\n" % (var, lit))


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])

"""Build step: write a text file as one C string literal so that the library
carries the exact text it compiles at run time (the snippet sort kernels)."""
import sys


def main(src, var, out):
    text = open(src).read()
    lit = "".join('"%s\\n"\n' % line.replace("\\", "\\\\").replace('"', '\\"')
                  for line in text.splitlines())
    with open(out, "w") as f:
        f.write("// Generated from %s; do not edit.\n" % src)
        f.write("static const char %s[] =\n%s;\n" % (var, lit))


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2], sys.argv[3])

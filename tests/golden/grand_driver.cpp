// Prints draws of the benchmark programs' GRand restatement
// (cl-ops_amd/cli/grand.hpp) for tests/test_bench_inputs.py, which compares
// them with GLib's own outputs in glib_grand.json.
//   grand_driver SEED ints N | doubles N | range BEGIN END N |
//                drange BEGIN END N | bench TYPE N
// Doubles print as C99 hex floats; bench prints each element's bytes in hex.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "grand.hpp"

int main(int argc, char** argv) {
	if (argc < 4) return 2;
	cli::GRand g(static_cast<uint32_t>(strtoul(argv[1], nullptr, 0)));
	const char* cmd = argv[2];
	if (!strcmp(cmd, "ints")) {
		for (long i = strtol(argv[3], nullptr, 0); i > 0; --i) printf("%u\n", g.next_u32());
	} else if (!strcmp(cmd, "doubles")) {
		for (long i = strtol(argv[3], nullptr, 0); i > 0; --i) printf("%a\n", g.next_double());
	} else if (!strcmp(cmd, "range") && argc >= 6) {
		const int32_t b = static_cast<int32_t>(strtol(argv[3], nullptr, 0));
		const int32_t e = static_cast<int32_t>(strtol(argv[4], nullptr, 0));
		for (long i = strtol(argv[5], nullptr, 0); i > 0; --i) printf("%d\n", g.int_range(b, e));
	} else if (!strcmp(cmd, "drange") && argc >= 6) {
		const double b = strtod(argv[3], nullptr), e = strtod(argv[4], nullptr);
		for (long i = strtol(argv[5], nullptr, 0); i > 0; --i) printf("%a\n", g.double_range(b, e));
	} else if (!strcmp(cmd, "bench") && argc >= 5) {
		// Element sizes by CloType (char .. double), so this program needs only
		// the headers, not the library.
		static const int kSize[] = {1, 1, 2, 2, 4, 4, 8, 8, 2, 4, 8};
		const long ti = strtol(argv[3], nullptr, 0);
		if (ti < 0 || ti > 10) return 3;
		const CloType t = static_cast<CloType>(ti);
		unsigned char buf[8];
		for (long i = strtol(argv[4], nullptr, 0); i > 0; --i) {
			if (!cli::bench_rand(g, t, buf)) return 3;
			for (int k = 0; k < kSize[ti]; ++k) printf("%02x", buf[k]);
			printf("\n");
		}
	} else {
		return 2;
	}
	return 0;
}

// clo_rng_bench: streams of a library generator written in the formats
// statistical test suites read (CL_Ops' src/benchmarks/clo_rng_bench.c:
// options :36-81, checks :146-158, files and formats :204-274, the per-run
// loop :302-324).  Same options, file names, formats and error exits.
//
// Run r writes the gws numbers of draw r (work-item order), as the reference
// does; draws come from clo_rng_fill in batches of runs instead of one launch
// and one blocking read per run.  runs = 0 generates until the output closes
// (stdout only).

#include <cerrno>
#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cl_ops.h"

namespace {

struct Args {
	std::string rng = "lcg";
	std::string output = "file-tsv";
	size_t gws = 262144;
	size_t lws = 256;
	unsigned runs = 10;
	int dev = -1;
	uint32_t seed = CLO_DEFAULT_SEED;
	std::string hash;
	bool have_hash = false;
	unsigned bits = 32;
	unsigned maxint = 0;
};

void usage(FILE* f) {
	fprintf(f,
		"Usage: clo_rng_bench [OPTION...] - Test RNGs\n"
		"  -r, --rng=RNG             Random number generator: lcg, xorshift64, xorshift128,\n"
		"                            mwc64x, parkmiller, tauslcg (default is lcg)\n"
		"  -o, --output=OUTPUT       Output: file-tsv, file-dh, stdout-bin, stdout-uint\n"
		"                            (default: file-tsv)\n"
		"  -g, --globalsize=SIZE     Global work size (default is 262144)\n"
		"  -l, --localsize=SIZE      Local work size (default is 256)\n"
		"  -n, --runs=SIZE           Random numbers per workitem (default is 10, 0 means\n"
		"                            continuous generation)\n"
		"  -d, --device=INDEX        Device index\n"
		"  -s, --rng-seed=SEED       Seed for random number generator (default is 0)\n"
		"  -h, --gid-hash=HASH       Use workitem GID-based seeds instead of MT derived seeds\n"
		"                            from host; the hash to apply (KNUTH(x), XS1(x) or code\n"
		"                            modifying x, e.g. x = x << 2)\n"
		"  -b, --bits                Number of bits in unsigned integers to produce\n"
		"                            (default 32)\n"
		"  -m, --max                 Maximum integer to produce, overrides --bits\n");
}

bool parse_uint(const char* s, unsigned long long* v) {
	char* end = nullptr;
	errno = 0;
	*v = strtoull(s, &end, 0);
	return errno == 0 && end && *end == '\0' && end != s;
}

// GOption-style parsing: -x VALUE, -xVALUE, --long VALUE, --long=VALUE.
int parse(int argc, char** argv, Args* a) {
	struct Opt {
		char s;
		const char* l;
	} const opts[] = {{'r', "rng"}, {'o', "output"}, {'g', "globalsize"}, {'l', "localsize"},
		{'n', "runs"}, {'d', "device"}, {'s', "rng-seed"}, {'h', "gid-hash"}, {'b', "bits"},
		{'m', "max"}};
	for (int i = 1; i < argc; ++i) {
		const char* arg = argv[i];
		if (!strcmp(arg, "--help") || !strcmp(arg, "-?")) {
			usage(stdout);
			exit(0);
		}
		char key = 0;
		const char* val = nullptr;
		if (arg[0] == '-' && arg[1] == '-') {
			const char* name = arg + 2;
			const char* eq = strchr(name, '=');
			const size_t len = eq ? static_cast<size_t>(eq - name) : strlen(name);
			for (const Opt& o : opts)
				if (strlen(o.l) == len && !strncmp(o.l, name, len)) key = o.s;
			if (eq) val = eq + 1;
		} else if (arg[0] == '-' && arg[1]) {
			for (const Opt& o : opts)
				if (o.s == arg[1]) key = o.s;
			if (arg[2]) val = arg + 2;
		}
		if (!key) {
			fprintf(stderr, "Error: Unknown option %s\n", arg);
			return CLO_ERROR_ARGS;
		}
		if (!val) {
			if (i + 1 >= argc) {
				fprintf(stderr, "Error: Missing argument for %s\n", arg);
				return CLO_ERROR_ARGS;
			}
			val = argv[++i];
		}
		unsigned long long v = 0;
		const bool numeric = strchr("glndsbm", key) != nullptr;
		if (numeric && !parse_uint(val, &v) && !(key == 'd' && !strcmp(val, "-1"))) {
			fprintf(stderr, "Error: Cannot parse integer value '%s' for %s\n", val, arg);
			return CLO_ERROR_ARGS;
		}
		switch (key) {
		case 'r': a->rng = val; break;
		case 'o': a->output = val; break;
		case 'g': a->gws = static_cast<size_t>(v); break;
		case 'l': a->lws = static_cast<size_t>(v); break;
		case 'n': a->runs = static_cast<unsigned>(v); break;
		case 'd': a->dev = strcmp(val, "-1") ? static_cast<int>(v) : -1; break;
		case 's': a->seed = static_cast<uint32_t>(v); break;
		case 'h':
			a->hash = val;
			a->have_hash = true;
			break;
		case 'b': a->bits = static_cast<unsigned>(v); break;
		case 'm': a->maxint = static_cast<unsigned>(v); break;
		}
	}
	return 0;
}

int fail(GError* err) {
	const int status = err ? err->code : CLO_ERROR_ARGS;
	fprintf(stderr, "Error: %s\n", err && err->message ? err->message : "unknown");
	clo_error_free(err);
	return status;
}

// Appends the decimal digits of v.
inline char* put_uint(char* p, uint32_t v) {
	char tmp[10];
	int k = 0;
	do {
		tmp[k++] = static_cast<char>('0' + v % 10u);
		v /= 10u;
	} while (v);
	while (k) *p++ = tmp[--k];
	return p;
}

}  // namespace

int main(int argc, char** argv) {
	Args a;
	int st = parse(argc, argv, &a);
	if (st) return st;
	GError* err = nullptr;
	const bool to_stdout = a.output.compare(0, 6, "stdout") == 0;
	// clo_rng_bench.c:146-158
	if (a.output != "file-tsv" && a.output != "file-dh" && a.output != "stdout-bin" &&
		a.output != "stdout-uint") {
		fprintf(stderr, "Error: Unknown output '%s'.\n", a.output.c_str());
		return CLO_ERROR_ARGS;
	}
	if (a.bits > 32 || a.bits < 1) {
		fprintf(stderr, "Error: Number of bits must be between 1 and 32.\n");
		return CLO_ERROR_ARGS;
	}
	if (a.runs == 0 && !to_stdout) {
		fprintf(stderr, "Error: Continuous generation can only be performed to stdout.\n");
		return CLO_ERROR_ARGS;
	}
	if (a.gws == 0) {
		fprintf(stderr, "Error: Global work size must be positive.\n");
		return CLO_ERROR_ARGS;
	}
	// A closed stdout ends continuous generation with EPIPE, not a signal.
	signal(SIGPIPE, SIG_IGN);

	CCLContext* ctx = ccl_context_new_from_device_index(a.dev < 0 ? 0 : a.dev, &err);
	if (!ctx) return fail(err);
	CCLQueue* q = ccl_queue_new(ctx, nullptr, 0, &err);
	if (!q) return fail(err);
	const CloRngSeedType seed_type = a.have_hash ? CLO_RNG_SEED_DEV_GID : CLO_RNG_SEED_HOST_MT;
	CloRng* rng = clo_rng_new(a.rng.c_str(), seed_type, nullptr, a.gws, a.seed,
		a.have_hash ? a.hash.c_str() : nullptr, ctx, q, &err);
	if (!rng) return fail(err);

	// clo_rng_bench.c:204-274: where the numbers go and how they are separated.
	FILE* out = stdout;
	std::string filename;
	bool raw = false;
	const char* sep_field = "\n";
	const char* sep_line = "";
	if (a.output == "stdout-bin") {
		raw = true;
	} else if (a.output == "file-dh" || a.output == "file-tsv") {
		const bool dh = a.output == "file-dh";
		filename = std::string("out_") + a.rng + "_" + (a.have_hash ? "gid_" : "host_") +
			(a.have_hash ? a.hash : "mt") + (dh ? ".dh.txt" : ".tsv");
		if (!dh) {
			sep_field = "\t";
			sep_line = "\n";
		}
		out = fopen(filename.c_str(), "w");
		if (!out) {
			fprintf(stderr, "Error: Unable to create output file '%s'.\n", filename.c_str());
			return CLO_ERROR_OPENFILE;
		}
		if (dh) {
			fprintf(out, "type: d\n");
			fprintf(out, "count: %d\n", static_cast<int>(a.gws * a.runs));
			fprintf(out, "numbit: %d\n", a.bits);
		}
	}
	if (!to_stdout) {
		printf("\n   =========================== Selected options ============================\n\n");
		printf("     Random number generator (seed): %s (%u)\n", a.rng.c_str(), a.seed);
		printf("     Seeds in workitems: %s %s\n",
			a.have_hash ? "GID-based, hash:" : "Host-based,",
			a.have_hash ? a.hash.c_str() : "Mersenne Twister");
		printf("     Global/local worksizes: %d/%d\n", static_cast<int>(a.gws),
			static_cast<int>(a.lws));
		printf("     Number of runs: %d\n", a.runs);
		printf("     Number of bits / Maximum integer: %d / %u\n", a.bits,
			static_cast<unsigned>((1ul << a.bits) - 1));
		printf("     Compiler Options: %s\n", a.maxint ? " -D CLO_RNG_BENCHMARK_MAXINT" : "");
		printf("\n   =========================== Execution status ============================\n\n");
		printf("     Producing random numbers...\n");
		fflush(stdout);
	}
	const auto t0 = std::chrono::steady_clock::now();

	// Batches of runs: up to 64 MiB of numbers per fill.
	const size_t batch_max = a.gws >= (16u << 20) ? 1 : (16u << 20) / a.gws;
	const size_t batch_runs = a.runs ? (a.runs < batch_max ? a.runs : batch_max) : batch_max;
	CCLBuffer* dev = ccl_buffer_new(ctx, CL_MEM_READ_WRITE, batch_runs * a.gws * 4, nullptr,
		&err);
	if (!dev) return fail(err);
	std::vector<uint32_t> host(batch_runs * a.gws);
	std::vector<char> text;
	if (!raw) text.resize(a.gws * 12 + 16);
	const size_t lf = strlen(sep_field), ll = strlen(sep_line);
	bool io_ok = true;
	int io_errno = 0;
	for (unsigned long long done = 0; io_ok && (a.runs == 0 || done < a.runs);) {
		const size_t nb = a.runs ? ((a.runs - done) < batch_runs ? (a.runs - done) : batch_runs)
			: batch_runs;
		if (!clo_rng_fill(rng, q, dev, static_cast<cl_uint>(nb), a.bits, a.maxint, &err) ||
			!ccl_buffer_enqueue_read(dev, q, CL_TRUE, 0, nb * a.gws * 4, host.data(), nullptr,
				&err))
			return fail(err);
		for (size_t r = 0; r < nb && io_ok; ++r) {
			const uint32_t* row = host.data() + r * a.gws;
			if (raw) {
				io_ok = fwrite(row, 4, a.gws, out) == a.gws;
				if (!io_ok) io_errno = errno;
				continue;
			}
			// Rows are formatted in pieces of at most a.gws numbers.
			char* p = text.data();
			for (size_t i = 0; i < a.gws; ++i) {
				p = put_uint(p, row[i]);
				memcpy(p, sep_field, lf);
				p += lf;
			}
			memcpy(p, sep_line, ll);
			p += ll;
			const size_t len = static_cast<size_t>(p - text.data());
			io_ok = fwrite(text.data(), 1, len, out) == len;
			if (!io_ok) io_errno = errno;
		}
		done += nb;
	}
	if (io_ok && fflush(out) != 0) {
		io_ok = false;
		io_errno = errno;
	}
	const double secs =
		std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	if (!to_stdout) {
		printf("     Finished, ellapsed time: %lfs\n", secs);
		printf("     Done...\n");
	}
	if (out != stdout) fclose(out);
	ccl_buffer_destroy(dev);
	clo_rng_destroy(rng);
	ccl_queue_destroy(q);
	ccl_context_destroy(ctx);
	// A reader closing the stream is how continuous generation ends.
	if (!io_ok && !(a.runs == 0 && io_errno == EPIPE)) {
		fprintf(stderr, "Error: Unable to write the output.\n");
		return CLO_ERROR_STREAM_WRITE;
	}
	return CLO_SUCCESS;
}

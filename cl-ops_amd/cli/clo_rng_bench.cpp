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
#include "cli_common.hpp"

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

// GOption's forms (cli_common.hpp); -d takes -1 for the default device.
int parse(int argc, char** argv, Args* a) {
	unsigned long long gws = a->gws, lws = a->lws, runs = a->runs, seed = a->seed, bits = a->bits,
		maxint = a->maxint;
	long long dev = a->dev;
	std::string usage_text;
	{
		char* buf = nullptr;
		size_t len = 0;
		FILE* f = open_memstream(&buf, &len);
		usage(f);
		fclose(f);
		usage_text.assign(buf, len);
		free(buf);
	}
	const std::vector<cli::Option> opts = {{'r', "rng", cli::kString, &a->rng, nullptr},
		{'o', "output", cli::kString, &a->output, nullptr},
		{'g', "globalsize", cli::kUInt, &gws, nullptr},
		{'l', "localsize", cli::kUInt, &lws, nullptr},
		{'n', "runs", cli::kUInt, &runs, nullptr},
		{'d', "device", cli::kInt, &dev, nullptr},
		{'s', "rng-seed", cli::kUInt, &seed, nullptr},
		{'h', "gid-hash", cli::kString, &a->hash, &a->have_hash},
		{'b', "bits", cli::kUInt, &bits, nullptr},
		{'m', "max", cli::kUInt, &maxint, nullptr}};
	const int st = cli::parse(argc, argv, opts, usage_text.c_str());
	a->gws = static_cast<size_t>(gws);
	a->lws = static_cast<size_t>(lws);
	a->runs = static_cast<unsigned>(runs);
	a->dev = static_cast<int>(dev);
	a->seed = static_cast<uint32_t>(seed);
	a->bits = static_cast<unsigned>(bits);
	a->maxint = static_cast<unsigned>(maxint);
	return st;
}

int fail(GError* err) { return cli::fail(err); }

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

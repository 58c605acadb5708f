// clo_sort_bench: the sort benchmark of CL_Ops (src/benchmarks/clo_sort_bench.c:
// options :36-78, loop :178-247, file output :249-259).  For N = 4..maxpo2,
// `runs` sorts of 2^N random elements through clo_sort_with_host_data; the
// time of a run is the kernel time recorded on the profiling exec queue;
// each size prints its Mkeys/s and whether the result came out sorted.
//
// Inputs are the reference's: one GRand seeded with the bench seed, each
// element drawn with clo_bench_rand (grand.hpp), so a given seed sorts the
// same arrays as the reference benchmark.

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cl_ops.h"
#include "cli_common.hpp"
#include "grand.hpp"

namespace {

const char kUsage[] =
	"Usage: clo_sort_bench [OPTION...] - Test sorting algorithms\n"
	"  -a, --algorithm=ALGORITHM   Sorting algorithm to use (default is sbitonic)\n"
	"  -g, --alg-opts=STRING       Algorithm options\n"
	"  -r, --runs=RUNS             Number of runs (default is 1)\n"
	"  -l, --localsize=SIZE        Maximum local work size (default is auto-select)\n"
	"  -d, --device=INDEX          Device index\n"
	"  -s, --rng-seed=SEED         Seed for random number generator (default is 0)\n"
	"  -t, --type=TYPE             Type of elements to sort (default uint)\n"
	"  -n, --maxpo2=N              Log2 of the maximum number of elements to sort, e.g.\n"
	"                              2^N (default N=24)\n"
	"  -o, --out=FILENAME          File where to output sorting benchmarks (default is no\n"
	"                              file output)\n"
	"  -c, --compiler=STRING       Compiler options\n";

// Elements i and i+1 in order for the element type (clo_bench.c:31-70).
bool in_order(CloType t, const unsigned char* a, const unsigned char* b) {
	switch (t) {
#define CMP(T)                                                                                     \
	{                                                                                              \
		T x, y;                                                                                    \
		memcpy(&x, a, sizeof(T));                                                                  \
		memcpy(&y, b, sizeof(T));                                                                  \
		return !(x > y);                                                                           \
	}
	case CLO_CHAR: CMP(int8_t)
	case CLO_UCHAR: CMP(uint8_t)
	case CLO_SHORT: CMP(int16_t)
	case CLO_USHORT: CMP(uint16_t)
	case CLO_INT: CMP(int32_t)
	case CLO_UINT: CMP(uint32_t)
	case CLO_LONG: CMP(int64_t)
	case CLO_ULONG: CMP(uint64_t)
	case CLO_FLOAT: CMP(float)
	case CLO_DOUBLE: CMP(double)
#undef CMP
	default: return true;
	}
}

}  // namespace

int main(int argc, char** argv) {
	std::string algorithm = "sbitonic", alg_opts, type = "uint", out, compiler;
	unsigned long long runs = 1, lws = 0, seed = CLO_DEFAULT_SEED, maxpo2 = 24;
	long long dev = -1;
	bool have_compiler = false;
	const std::vector<cli::Option> opts = {{'a', "algorithm", cli::kString, &algorithm, nullptr},
		{'g', "alg-opts", cli::kString, &alg_opts, nullptr},
		{'r', "runs", cli::kUInt, &runs, nullptr}, {'l', "localsize", cli::kUInt, &lws, nullptr},
		{'d', "device", cli::kInt, &dev, nullptr}, {'s', "rng-seed", cli::kUInt, &seed, nullptr},
		{'t', "type", cli::kString, &type, nullptr},
		{'n', "maxpo2", cli::kUInt, &maxpo2, nullptr},
		{'o', "out", cli::kString, &out, nullptr},
		{'c', "compiler", cli::kString, &compiler, &have_compiler}};
	int st = cli::parse(argc, argv, opts, kUsage);
	if (st) return st;
	GError* err = nullptr;
	CloType et = clo_type_by_name(type.c_str(), &err);
	if (err) return cli::fail(err);
	if (et == CLO_HALF) {
		// The reference checks half results as unsigned shorts on the host
		// (clo_bench.c:52) while the device compares half floats.
		fprintf(stderr, "Error: Type 'half' is not supported by this benchmark.\n");
		return CLO_ERROR_ARGS;
	}
	if (maxpo2 < 4 || maxpo2 > 31 || runs == 0) {
		fprintf(stderr, "Error: maxpo2 must be in 4..31 and runs positive.\n");
		return CLO_ERROR_ARGS;
	}
	const size_t bytes = clo_type_sizeof(et);
	seed = static_cast<uint32_t>(seed);

	CCLContext* ctx = ccl_context_new_from_device_index(dev < 0 ? 0 : static_cast<int>(dev),
		&err);
	if (!ctx) return cli::fail(err);
	CloSort* sorter = clo_sort_new(algorithm.c_str(), alg_opts.c_str(), ctx, &et, nullptr, nullptr,
		nullptr, have_compiler ? compiler.c_str() : nullptr, &err);
	if (!sorter) return cli::fail(err);
	CCLQueue* cq_exec = ccl_queue_new(ctx, nullptr, CL_QUEUE_PROFILING_ENABLE, &err);
	if (!cq_exec) return cli::fail(err);
	CCLQueue* cq_comm = ccl_queue_new(ctx, nullptr, 0, &err);
	if (!cq_comm) return cli::fail(err);

	printf("\n   =========================== Selected options ============================\n\n");
	printf("     Random number generator seed: %u\n", static_cast<unsigned>(seed));
	printf("     Maximum local worksize (0 is auto-select): %d\n", static_cast<int>(lws));
	printf("     Type of elements to sort: %s\n", clo_type_get_name(et));
	printf("     Number of runs: %d\n", static_cast<int>(runs));
	printf("     Compiler Options: %s\n", have_compiler ? compiler.c_str() : "(null)");

	const size_t max_elems = static_cast<size_t>(1) << maxpo2;
	std::vector<unsigned char> host(bytes * max_elems);
	cli::GRand rng(static_cast<uint32_t>(seed));
	std::vector<std::vector<unsigned long>> bench(maxpo2, std::vector<unsigned long>(runs, 0));

	for (unsigned N = 4; N <= maxpo2; ++N) {
		const size_t num = static_cast<size_t>(1) << N;
		bool sorted_ok = true;
		for (unsigned long long r = 0; r < runs; ++r) {
			for (size_t i = 0; i < num; ++i) cli::bench_rand(rng, et, host.data() + bytes * i);
			ccl_prof_reset(cq_exec);
			if (!clo_sort_with_host_data(sorter, cq_exec, cq_comm, host.data(), host.data(), num,
					static_cast<size_t>(lws), &err))
				return cli::fail(err);
			double ms = 0;
			ccl_prof_get(cq_exec, "", &ms, &err);
			if (err) return cli::fail(err);
			bench[N - 1][r] = static_cast<unsigned long>(ms * 1e6);  // ns
			sorted_ok = true;
			for (size_t i = 0; i + 1 < num; ++i)
				if (!in_order(et, host.data() + bytes * i, host.data() + bytes * (i + 1))) {
					sorted_ok = false;
					break;
				}
		}
		unsigned long long total = 0;
		for (unsigned long long r = 0; r < runs; ++r) total += bench[N - 1][r];
		printf("       - 2^%d: %lf Mkeys/s %s\n", N,
			(1e-6 * static_cast<double>(num) * static_cast<double>(runs)) /
				(static_cast<double>(total) * 1e-9),
			sorted_ok ? "" : "(sort did not work)");
		fflush(stdout);
	}
	if (!out.empty()) {
		FILE* f = fopen(out.c_str(), "w");
		if (!f) {
			fprintf(stderr, "Error: Unable to create output file '%s'.\n", out.c_str());
			return CLO_ERROR_OPENFILE;
		}
		for (unsigned i = 0; i < maxpo2; ++i) {
			fprintf(f, "%u", i);
			for (unsigned long long j = 0; j < runs; ++j) fprintf(f, "\t%lu", bench[i][j]);
			fprintf(f, "\n");
		}
		fclose(f);
	}
	clo_sort_destroy(sorter);
	ccl_queue_destroy(cq_exec);
	ccl_queue_destroy(cq_comm);
	ccl_context_destroy(ctx);
	return CLO_SUCCESS;
}

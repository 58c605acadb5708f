// clo_scan_bench: the scan benchmark of CL_Ops (src/benchmarks/clo_scan_bench.c:
// options :43-78, loop :203-291, file output :293-303).  Starting from
// init-elems and doubling num-doub times, `runs` exclusive scans of random
// values in [0, 127] through clo_scan_with_host_data; a run's time is the
// kernel time recorded on the profiling exec queue; each size prints its
// MValues/s and the outcome of a serial host check ("[Overflow]" when the
// sum no longer fits the sum type, "[Scan did not work]" on a mismatch,
// "[Unverified]" with --no-check).
//
// Values are the reference's: one GRand seeded with the bench seed, value =
// (ulong)(g_rand_double * 128) per element (grand.hpp), so a given seed
// scans the same arrays as the reference benchmark.

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cl_ops.h"
#include "cli_common.hpp"
#include "grand.hpp"

namespace {

const char kUsage[] =
	"Usage: clo_scan_bench [OPTION...] - Test CL_Ops scan implementations\n"
	"  -r, --runs=RUNS            Number of runs (default is 1)\n"
	"  -l, --localsize=SIZE       Maximum local work size (default is 0, auto)\n"
	"  -d, --device=INDEX         Device index\n"
	"  -s, --rng-seed=SEED        Seed for random number generator (default is 0)\n"
	"  -t, --type=TYPE            Type of elements to scan (default uint)\n"
	"  -y, --type-sum=TYPE        Type of elements in scan result (default ulong)\n"
	"  -i, --init-elems=SIZE      The starting number of elements to scan (default is 4)\n"
	"  -n, --num-doub=SIZE        Number of times the number of elements is doubled\n"
	"                             (default is 24)\n"
	"  -c, --compiler=STRING      Compiler options\n"
	"  -a, --algorithm=ALGORITHM  Scan algorithm to use (default is blelloch)\n"
	"  -p, --alg-opts=STRING      Algorithm options\n"
	"  -u, --no-check             Don't check scan with serial version\n"
	"  -o, --out=FILENAME         File where to output scan benchmarks (default is no\n"
	"                             file output)\n";

}  // namespace

int main(int argc, char** argv) {
	std::string type = "uint", type_sum = "ulong", compiler, algorithm = "blelloch", alg_opts, out;
	unsigned long long runs = 1, lws = 0, seed = CLO_DEFAULT_SEED, init_elems = 4, num_doub = 24;
	long long dev = -1;
	bool no_check = false, have_compiler = false;
	const std::vector<cli::Option> opts = {{'r', "runs", cli::kUInt, &runs, nullptr},
		{'l', "localsize", cli::kUInt, &lws, nullptr}, {'d', "device", cli::kInt, &dev, nullptr},
		{'s', "rng-seed", cli::kUInt, &seed, nullptr}, {'t', "type", cli::kString, &type, nullptr},
		{'y', "type-sum", cli::kString, &type_sum, nullptr},
		{'i', "init-elems", cli::kUInt, &init_elems, nullptr},
		{'n', "num-doub", cli::kUInt, &num_doub, nullptr},
		{'c', "compiler", cli::kString, &compiler, &have_compiler},
		{'a', "algorithm", cli::kString, &algorithm, nullptr},
		{'p', "alg-opts", cli::kString, &alg_opts, nullptr},
		{'u', "no-check", cli::kFlag, &no_check, nullptr},
		{'o', "out", cli::kString, &out, nullptr}};
	int st = cli::parse(argc, argv, opts, kUsage);
	if (st) return st;
	GError* err = nullptr;
	const CloType et = clo_type_by_name(type.c_str(), &err);
	if (err) return cli::fail(err);
	const CloType stp = clo_type_by_name(type_sum.c_str(), &err);
	if (err) return cli::fail(err);
	if (init_elems == 0 || runs == 0 || (init_elems << num_doub) > (1ull << 31)) {
		fprintf(stderr, "Error: init-elems * 2^num-doub must be in 1..2^31 and runs positive.\n");
		return CLO_ERROR_ARGS;
	}
	const size_t bytes = clo_type_sizeof(et), bytes_sum = clo_type_sizeof(stp);
	seed = static_cast<uint32_t>(seed);

	CCLContext* ctx = ccl_context_new_from_device_index(dev < 0 ? 0 : static_cast<int>(dev),
		&err);
	if (!ctx) return cli::fail(err);
	CloScan* scanner = clo_scan_new(algorithm.c_str(), alg_opts.c_str(), ctx, et, stp,
		have_compiler ? compiler.c_str() : nullptr, &err);
	if (!scanner) return cli::fail(err);
	CCLQueue* cq_exec = ccl_queue_new(ctx, nullptr, CL_QUEUE_PROFILING_ENABLE, &err);
	if (!cq_exec) return cli::fail(err);
	CCLQueue* cq_comm = ccl_queue_new(ctx, nullptr, 0, &err);
	if (!cq_comm) return cli::fail(err);

	printf("\n   =========================== Selected options ============================\n\n");
	printf("     Random number generator seed: %u\n", static_cast<unsigned>(seed));
	printf("     Maximum local worksize (0 = auto): %d\n", static_cast<int>(lws));
	printf("     Type of elements to scan: %s\n", clo_type_get_name(et));
	printf("     Type of elements in scan result: %s\n", clo_type_get_name(stp));
	printf("     Starting number of elements: %d\n", static_cast<int>(init_elems));
	printf("     Number of times number of elements will be doubled: %d\n",
		static_cast<int>(num_doub));
	printf("     Number of runs: %d\n", static_cast<int>(runs));
	printf("     Compiler Options: %s\n", have_compiler ? compiler.c_str() : "(null)");

	const size_t max_elems = static_cast<size_t>(init_elems) << num_doub;
	std::vector<unsigned char> host(bytes * max_elems), scanned(bytes_sum * max_elems);
	cli::GRand rng(static_cast<uint32_t>(seed));
	std::vector<std::vector<double>> bench(num_doub, std::vector<double>(runs, 0.0));
	const uint64_t max_sum = bytes_sum >= 8 ? ~0ull : ((1ull << (8 * bytes_sum)) - 1);
	size_t num = init_elems;
	for (unsigned N = 1; N <= num_doub; ++N, num *= 2) {
		const char* scan_ok = "";
		for (unsigned long long r = 0; r < runs; ++r) {
			// clo_scan_bench.c:219-224: 0..127, so any element type holds it.
			for (size_t i = 0; i < num; ++i) {
				const uint64_t v = static_cast<uint64_t>(rng.next_double() * 128);
				memcpy(host.data() + bytes * i, &v, bytes);
			}
			ccl_prof_reset(cq_exec);
			if (!clo_scan_with_host_data(scanner, cq_exec, cq_comm, host.data(), scanned.data(),
					num, static_cast<size_t>(lws), &err))
				return cli::fail(err);
			double ms = 0;
			ccl_prof_get(cq_exec, "", &ms, &err);
			if (err) return cli::fail(err);
			bench[N - 1][r] = ms * 1e6;  // ns
			if (no_check) {
				scan_ok = "[Unverified]";
				continue;
			}
			// clo_scan_bench.c:243-268: serial exclusive scan and overflow check.
			scan_ok = "";
			uint64_t value_host = 0;
			for (size_t i = 0; i < num; ++i) {
				if (i) {
					uint64_t prev = 0;
					memcpy(&prev, host.data() + bytes * (i - 1), bytes);
					value_host += prev;
				}
				if (value_host > max_sum) {
					scan_ok = "[Overflow]";
					break;
				}
				uint64_t value_dev = 0;
				memcpy(&value_dev, scanned.data() + bytes_sum * i, bytes_sum);
				if (value_dev != value_host) {
					scan_ok = "[Scan did not work]";
					break;
				}
			}
		}
		double total = 0;
		for (unsigned long long r = 0; r < runs; ++r) total += bench[N - 1][r];
		printf("       - %10d : %f MValues/s %s\n", static_cast<int>(num),
			(1e-6 * static_cast<double>(num) * static_cast<double>(runs)) / (total * 1e-9),
			scan_ok);
		fflush(stdout);
	}
	if (!out.empty()) {
		FILE* f = fopen(out.c_str(), "w");
		if (!f) {
			fprintf(stderr, "Error: Unable to create output file '%s'.\n", out.c_str());
			return CLO_ERROR_OPENFILE;
		}
		for (unsigned i = 0; i < num_doub; ++i) {
			fprintf(f, "%u", i);
			for (unsigned long long j = 0; j < runs; ++j) fprintf(f, "\t%lf", bench[i][j]);
			fprintf(f, "\n");
		}
		fclose(f);
	}
	clo_scan_destroy(scanner);
	ccl_queue_destroy(cq_exec);
	ccl_queue_destroy(cq_comm);
	ccl_context_destroy(ctx);
	return CLO_SUCCESS;
}

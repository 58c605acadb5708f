/* A C99 caller of libcl_ops.so, written as a CL_Ops user writes one
 * (compare src/benchmarks/clo_sort_bench.c): a stable key-value sort, a
 * scan and an RNG fill through the public headers only, each checked on the
 * host.  Prints "c_caller: ok" and exits with 0 on success, or prints the
 * library error and exits with its code. */

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cl_ops.h>

static cl_ulong next_u64(cl_ulong* s) { /* splitmix64 */
	cl_ulong z = (*s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

static int fail(GError* err, const char* what) {
	int code = err ? err->code : -1;
	fprintf(stderr, "c_caller: %s: %s\n", what, err && err->message ? err->message : "failed");
	clo_error_free(err);
	return code ? code : 1;
}

int main(void) {
	GError* err = NULL;
	const size_t n = 1u << 20;
	size_t i;
	cl_ulong seed = 42;
	CloType kt = CLO_ULONG;
	CCLContext* ctx;
	CCLQueue* q;
	CloSort* sorter;
	CloScan* scanner;
	CloRng *rng_a, *rng_b;
	CCLBuffer *out_a, *out_b;
	cl_ulong* keys = malloc(n * sizeof(cl_ulong));
	cl_uint* vals = malloc(n * sizeof(cl_uint));
	unsigned char* seen = calloc(n, 1);
	cl_uint* values = malloc(n * sizeof(cl_uint));
	cl_ulong* sums = malloc(n * sizeof(cl_ulong));
	cl_uint *draws_a = malloc(4096 * 4 * sizeof(cl_uint)), *draws_b = malloc(4096 * 4 * 4);
	if (!keys || !vals || !seen || !values || !sums || !draws_a || !draws_b) return 1;

	ctx = ccl_context_new_from_device_index(0, &err);
	if (!ctx) return fail(err, "context");
	q = ccl_queue_new(ctx, NULL, CL_QUEUE_PROFILING_ENABLE, &err);
	if (!q) return fail(err, "queue");

	/* Key-value sort: 12-bit keys (many ties), payload = input position. */
	for (i = 0; i < n; ++i) {
		keys[i] = next_u64(&seed) & 0xFFF;
		vals[i] = (cl_uint) i;
	}
	sorter = clo_sort_new("satradix", "radix=256", ctx, &kt, NULL, NULL, NULL, NULL, &err);
	if (!sorter) return fail(err, "clo_sort_new");
	if (!clo_sort_kv_with_host_data(sorter, q, NULL, keys, vals, n, 0, &err))
		return fail(err, "clo_sort_kv_with_host_data");
	for (i = 0; i < n; ++i) {
		if (vals[i] >= n || seen[vals[i]]) return fail(NULL, "payloads are not a permutation");
		seen[vals[i]] = 1;
		if (i && (keys[i - 1] > keys[i] || (keys[i - 1] == keys[i] && vals[i - 1] > vals[i])))
			return fail(NULL, "not a stable sort");
	}

	/* Exclusive scan, uint -> ulong. */
	for (i = 0; i < n; ++i) values[i] = (cl_uint) (next_u64(&seed) >> 40);
	scanner = clo_scan_new("blelloch", NULL, ctx, CLO_UINT, CLO_ULONG, NULL, &err);
	if (!scanner) return fail(err, "clo_scan_new");
	if (!clo_scan_with_host_data(scanner, q, NULL, values, sums, n, 0, &err))
		return fail(err, "clo_scan_with_host_data");
	{
		cl_ulong run = 0;
		for (i = 0; i < n; ++i) {
			if (sums[i] != run) return fail(NULL, "scan mismatch");
			run += values[i];
		}
	}

	/* Two generators from one host seed give one stream. */
	rng_a = clo_rng_new("xorshift128", CLO_RNG_SEED_HOST_MT, NULL, 4096, 7, NULL, ctx, q, &err);
	rng_b = rng_a ? clo_rng_new("xorshift128", CLO_RNG_SEED_HOST_MT, NULL, 4096, 7, NULL, ctx, q,
		&err) : NULL;
	if (!rng_a || !rng_b) return fail(err, "clo_rng_new");
	out_a = ccl_buffer_new(ctx, CL_MEM_READ_WRITE, 4096 * 4 * 4, NULL, &err);
	out_b = out_a ? ccl_buffer_new(ctx, CL_MEM_READ_WRITE, 4096 * 4 * 4, NULL, &err) : NULL;
	if (!out_b) return fail(err, "ccl_buffer_new");
	if (!clo_rng_fill(rng_a, q, out_a, 4, 32, 0, &err) ||
		!clo_rng_fill(rng_b, q, out_b, 4, 32, 0, &err) ||
		!ccl_buffer_enqueue_read(out_a, q, CL_TRUE, 0, 4096 * 16, draws_a, NULL, &err) ||
		!ccl_buffer_enqueue_read(out_b, q, CL_TRUE, 0, 4096 * 16, draws_b, NULL, &err))
		return fail(err, "clo_rng_fill");
	if (memcmp(draws_a, draws_b, 4096 * 16) != 0) return fail(NULL, "rng streams differ");

	clo_rng_destroy(rng_a);
	clo_rng_destroy(rng_b);
	ccl_buffer_destroy(out_a);
	ccl_buffer_destroy(out_b);
	clo_scan_destroy(scanner);
	clo_sort_destroy(sorter);
	ccl_queue_destroy(q);
	ccl_context_destroy(ctx);
	free(keys);
	free(vals);
	free(seen);
	free(values);
	free(sums);
	free(draws_a);
	free(draws_b);
	printf("c_caller: ok\n");
	return 0;
}

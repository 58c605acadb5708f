/* Sanitizer driver for the oracle (test infrastructure, built by
 * `make -C oracle asan` and `make -C oracle tsan`): runs every threaded
 * oracle function with 1 and with 4 host threads on small seeded inputs and
 * requires identical results, plus the single-threaded functions once, so
 * AddressSanitizer / UndefinedBehaviorSanitizer see every routine and
 * ThreadSanitizer sees the worker threads of orc_parallel_for. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int fails = 0;

static void check(int ok, const char* what) {
	if (!ok) {
		fprintf(stderr, "MISMATCH: %s\n", what);
		++fails;
	}
}

static void satradix_case(int width, int bits, size_t n) {
	uint64_t* raw = malloc(n * 8 + 8);
	orc_generate(0, 0x5A + width + bits, 0, 0, n, raw);
	void* k1 = malloc(n * 8 + 8);
	void* k4 = malloc(n * 8 + 8);
	void* kr = malloc(n * 8 + 8);
	uint32_t* v1 = malloc(n * 4 + 4);
	uint32_t* v4 = malloc(n * 4 + 4);
	uint32_t* vr = malloc(n * 4 + 4);
	for (size_t i = 0; i < n; ++i) {
		switch (width) {
		case 1: ((uint8_t*)k1)[i] = (uint8_t)raw[i]; break;
		case 2: ((uint16_t*)k1)[i] = (uint16_t)raw[i]; break;
		case 4: ((uint32_t*)k1)[i] = (uint32_t)(raw[i] % 5000); break;
		default: ((uint64_t*)k1)[i] = raw[i]; break;
		}
		v1[i] = (uint32_t)i;
	}
	memcpy(k4, k1, n * (size_t)width);
	memcpy(kr, k1, n * (size_t)width);
	memcpy(v4, v1, n * 4);
	memcpy(vr, v1, n * 4);
	orc_satradix(k1, v1, n, width, bits, 256, 1);
	orc_satradix(k4, v4, n, width, bits, 256, 4);
	orc_stable_sort(kr, vr, n, width);
	check(memcmp(k1, k4, n * (size_t)width) == 0 && memcmp(v1, v4, n * 4) == 0,
		"satradix 1 vs 4 threads");
	check(memcmp(k1, kr, n * (size_t)width) == 0 && memcmp(v1, vr, n * 4) == 0,
		"satradix vs stable sort");
	free(raw);
	free(k1);
	free(k4);
	free(kr);
	free(v1);
	free(v4);
	free(vr);
}

static void scan_case(size_t n, int sum_width) {
	uint32_t* in = malloc(n * 4 + 4);
	orc_generate(3, 0x5C + n, 0xFF, 0, n, in);
	void* o1 = malloc(n * 8 + 8);
	void* o4 = malloc(n * 8 + 8);
	void* oc = malloc(n * 8 + 8);
	orc_scan_blelloch(in, o1, sum_width, n, 64, 1);
	orc_scan_blelloch(in, o4, sum_width, n, 64, 4);
	orc_scan(in, 4, 0, oc, sum_width, n);
	check(memcmp(o1, o4, n * (size_t)sum_width) == 0, "blelloch 1 vs 4 threads");
	check(memcmp(o1, oc, n * (size_t)sum_width) == 0, "blelloch vs scan");
	free(in);
	free(o1);
	free(o4);
	free(oc);
}

static void rng_case(int gen) {
	const size_t G = 3001;
	const uint32_t draws = 5;
	const size_t ss = orc_rng_state_size(gen);
	unsigned char* s1 = malloc(G * ss);
	unsigned char* s4 = malloc(G * ss);
	uint32_t* o1 = malloc(G * draws * 4);
	uint32_t* o4 = malloc(G * draws * 4);
	orc_rng_init_host_mt(gen, G, 1234, s1);
	orc_rng_init_dev_gid(gen, G, 99, 2, s4);
	orc_rng_init_host_mt(gen, G, 1234, s4);
	orc_rng_fill(gen, s1, G, draws, 32, 0, o1, 1);
	orc_rng_fill(gen, s4, G, draws, 32, 0, o4, 4);
	check(memcmp(o1, o4, G * draws * 4) == 0 && memcmp(s1, s4, G * ss) == 0,
		"rng fill 1 vs 4 threads");
	free(s1);
	free(s4);
	free(o1);
	free(o4);
}

int main(void) {
	const int widths[4] = {1, 2, 4, 8};
	for (int w = 0; w < 4; ++w) {
		satradix_case(widths[w], 4, 3001);
		satradix_case(widths[w], 8, 20000);
	}
	scan_case(1, 4);
	scan_case(100003, 4);
	scan_case(100003, 8);
	for (int g = 0; g < 6; ++g) rng_case(g);
	/* single-threaded routines */
	{
		const size_t n = 1024;
		int32_t k[1024];
		for (size_t i = 0; i < n; ++i) k[i] = (int32_t)(orc_mix64(i) & 0xFFFF) - 30000;
		orc_sbitonic(k, n, 4, 1, 0);
		int sorted = 1;
		for (size_t i = 1; i < n; ++i) sorted &= k[i - 1] <= k[i];
		check(sorted, "sbitonic");
		int32_t ko[1024];
		uint32_t vi[1024], vo[1024];
		for (size_t i = 0; i < n; ++i) vi[i] = (uint32_t)i;
		orc_gselect(k, vi, ko, vo, n, 4, 1, 1);
		check(ko[0] >= ko[n - 1], "gselect descending");
		double cdf[4096];
		orc_zipf_cdf(cdf, 4096, 1.1);
		uint64_t z[512];
		orc_generate_zipf(cdf, 4096, 7, 512, z);
		uint32_t mt[700];
		orc_mt19937(5489, 700, mt);
		unsigned char st[16];
		orc_rng_from_seed(5, 0x0123456789ABCDEFull, st);
		(void)orc_rng_next(5, st);
		(void)orc_swapped_position(3, 12345);
	}
	printf("oracle sanitizer driver: %d mismatches\n", fails);
	return fails ? 1 : 0;
}

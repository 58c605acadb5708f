/*
 * oracle.h - CPU restatement of CL_Ops' hot-path algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the
 * benchmark's CPU-baseline leg may load this library, and only as the checker
 * or the timed CPU baseline, never as the product path.
 *
 * Each function cites the reference file:line it restates.  Nothing here
 * shares code with the HIP library.
 */
#ifndef CLO_ORACLE_H_
#define CLO_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded inputs (the build's own generator, see clo_util.h) ----------- */
uint64_t orc_mix64(uint64_t z);
/* kind: 0 U64, 1 U32_HI, 2 U64_MASK, 3 U32_MASK, 4 IOTA_U32, 6 U32_MOD,
 * 7 U64_MOD, 8/9 NEARLY_SORTED_U32/U64, 10/11 REVERSED_U32/U64 (same
 * numbering and formulas as CloGenKind). */
uint64_t orc_swapped_position(uint64_t seed, uint64_t c);
void orc_generate(int kind, uint64_t seed, uint64_t param, uint64_t start, size_t n, void* out);
void orc_zipf_cdf(double* cdf, size_t ranks, double s);
void orc_generate_zipf(const double* cdf, size_t ranks, uint64_t seed, size_t n, uint64_t* out);

/* ---- sorts ---------------------------------------------------------------- */
/* satradix restated with the reference's work decomposition
 * (src/cl_ops/sort/clo_sort_satradix.c:117-335, .cl:34-258): per pass of
 * `bits`-bit digits, each block of `lws` elements is split-sorted bit by bit,
 * digit offsets/counters are taken per block, the digit-major counters are
 * exclusive-scanned, and every element is scattered to
 * counters_sum[d*numWG+wg] + lid - offsets[wg*R+d].  Blocks are spread over
 * `threads` host threads.  `width` = key bytes (1, 2, 4, 8); vals (uint32) may
 * be NULL.  Keys compare as unsigned.  Any n (the last block is partial). */
void orc_satradix(void* keys, uint32_t* vals, size_t n, int width, int bits, size_t lws,
	int threads);

/* Stable LSD counting sort by unsigned key (8-bit digits), the plain closed
 * form of the same result. */
void orc_stable_sort(void* keys, uint32_t* vals, size_t n, int width);

/* Bitonic network of src/cl_ops/sort/clo_sort_sbitonic.cl:38-69 driven as in
 * clo_sort_sbitonic.c:102-118, on n = power of two elements.  `sign` = keys
 * are signed; `desc` = compare snippet ((a) < (b)). */
void orc_sbitonic(void* keys, size_t n, int width, int sign, int desc);

/* gselect rank sort (src/cl_ops/sort/clo_sort_gselect.cl:38-57): out-of-place
 * stable sort by rank.  vals may be NULL. */
void orc_gselect(const void* keys, const uint32_t* vals, void* keys_out, uint32_t* vals_out,
	size_t n, int width, int sign, int desc);

/* ---- scan ----------------------------------------------------------------- */
/* Exclusive scan: out[i] = sum_{j<i} (SUM)in[j] mod 2^(8*sum_width), the
 * element converted as C converts `in_width`-byte (signed if in_sign) values.
 * (src/cl_ops/scan/clo_scan_blelloch.cl:49-211, checked the same way as
 * src/benchmarks/clo_scan_bench.c:246-271). */
void orc_scan(const void* in, int in_width, int in_sign, void* out, int sum_width, size_t n);

/* The reference's three-kernel Blelloch decomposition
 * (src/cl_ops/scan/clo_scan_blelloch.c:78-214): per work-group serial blocks of
 * 2*lws with a tree up/down sweep, a scan of the work-group sums, and the
 * add-back, with work-groups spread over `threads` host threads.  uint32 in,
 * uint32 or uint64 sums. */
void orc_scan_blelloch(const uint32_t* in, void* out, int sum_width, size_t n, size_t lws,
	int threads);

/* ---- RNG ------------------------------------------------------------------ */
/* gen: 0 lcg, 1 xorshift64, 2 xorshift128, 3 mwc64x, 4 parkmiller, 5 tauslcg
 * (src/cl_ops/rng/clo_rng.c:60-68).  States are stored as in device memory:
 * lcg/xorshift64 uint64, xorshift128 4 x uint32 (x,y,z,w), mwc64x 2 x uint32
 * (x,c), parkmiller int32, tauslcg 4 x uint32 (x,y,z,w). */
size_t orc_rng_state_size(int gen);
/* clo_ulong2statetype of each generator. */
void orc_rng_from_seed(int gen, uint64_t seed, void* state);
/* The Tausworthe step of tauslcg (src/cl_ops/rng/clo_rng_tauslcg.cl:52-55):
 * ((z & m) << s3) ^ (((z << s1) ^ z) >> s2). */
uint32_t orc_taus(uint32_t z, int s1, int s2, int s3, uint32_t m);
/* One clo_rng_next (src/cl_ops/rng/clo_rng_<gen>.cl). */
uint32_t orc_rng_next(int gen, void* state);
/* DEV_GID seeds (src/cl_ops/rng/clo_rng_init.cl:47-60): state[gid] =
 * conv(HASH(gid + main_seed)); hash 0 none, 1 KNUTH(x), 2 XS1(x). */
void orc_rng_init_dev_gid(int gen, size_t count, uint64_t main_seed, int hash, void* states);
/* HOST_MT seeds (src/cl_ops/rng/clo_rng.c:168-234). */
void orc_rng_init_host_mt(int gen, size_t count, uint64_t main_seed, void* states);
/* MT19937 words (init_genrand seeding). */
void orc_mt19937(uint32_t seed, size_t n, uint32_t* out);
/* The reference bench kernel run `draws` times (src/benchmarks/
 * clo_rng_bench.cl:23-37 + clo_rng_bench.c:302-324):
 * out[r*G + gid] = maxint ? next % maxint : next >> (32 - bits).  Work-items
 * are spread over `threads` host threads. */
void orc_rng_fill(int gen, void* states, size_t G, uint32_t draws, uint32_t bits, uint32_t maxint,
	uint32_t* out, int threads);

#ifdef __cplusplus
}
#endif

#endif

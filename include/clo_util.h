/*
 * clo_util.h - seeded input generators and a copy kernel (additions with no
 * counterpart in CL_Ops).  Benchmarks and tests generate their inputs on the
 * device with these, from the counter-based generator
 *   U(seed, i) = mix64(seed + (i + 1) * 0x9E3779B97F4A7C15)
 * where mix64 is the splitmix64 finalizer (Steele, Lea & Flood 2014), so a
 * CPU restatement produces the same words.
 */
#ifndef CLO_UTIL_H_
#define CLO_UTIL_H_

#include "clo_common.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
	CLO_GEN_U64 = 0,       /* U(seed, i) */
	CLO_GEN_U32_HI = 1,    /* U(seed, i) >> 32 */
	CLO_GEN_U64_MASK = 2,  /* U(seed, i) & param */
	CLO_GEN_U32_MASK = 3,  /* (U(seed, i) >> 32) & param */
	CLO_GEN_IOTA_U32 = 4,  /* (uint32) i */
	CLO_GEN_ZIPF_U64 = 5,  /* mix64(rank), rank from the CDF table, see below */
	CLO_GEN_U32_MOD = 6,   /* (U(seed, i) >> 32) % param */
	CLO_GEN_U64_MOD = 7,   /* U(seed, i) % param */
	/* "Presorted with 1% random swaps": position i holds s(i) * param
	 * (truncated to the width), s a product of disjoint transpositions:
	 * blocks of 100 positions are paired, block b with b ^ m for
	 * m = 1 + U(seed ^ 0x2545f4914f6cdd1d, 0) % 1023, and the pair {lo, hi}
	 * exchanges positions 100 lo + u % 100 and 100 hi + (u >> 32) % 100,
	 * u = U(seed ^ 0x5bd1e995, lo): 2 positions in every 200 move. */
	CLO_GEN_NEARLY_SORTED_U32 = 8,
	CLO_GEN_NEARLY_SORTED_U64 = 9,
	/* Descending: all-ones minus i * param. */
	CLO_GEN_REVERSED_U32 = 10,
	CLO_GEN_REVERSED_U64 = 11
} CloGenKind;

/* Writes n words of the kind's width (4 or 8 bytes) to out at element
 * offset 0, element i from counter i + start. */
CCLEvent* clo_util_generate(CCLQueue* cq, CCLBuffer* out, size_t n,
	CloGenKind kind, cl_ulong seed, cl_ulong param, cl_ulong start,
	GError** err);

/* Zipf(s) keys over `ranks` ranks: rank r (0-based) has probability
 * proportional to (r+1)^-s.  cdf is a device buffer of `ranks` doubles,
 * cdf[r] = P(rank <= r) (built on the host by clo_util_zipf_cdf).  Element
 * i = mix64(rank) where rank is the smallest r with cdf[r] > u,
 * u = (U(seed, i) >> 11) * 2^-53. */
/* As clo_util_generate_zipf with 32-bit elements: (uint32)(mix64(rank) >> 32). */
CCLEvent* clo_util_generate_zipf_u32(CCLQueue* cq, CCLBuffer* out, size_t n,
	CCLBuffer* cdf, size_t ranks, cl_ulong seed, GError** err);
CCLEvent* clo_util_generate_zipf(CCLQueue* cq, CCLBuffer* out, size_t n,
	CCLBuffer* cdf, size_t ranks, cl_ulong seed, GError** err);
void clo_util_zipf_cdf(double* cdf, size_t ranks, double s);

/* Plain device copy kernel of `bytes` bytes (dst and src 16-byte aligned).
 * Used to calibrate memory-traffic counters against a known byte count. */
CCLEvent* clo_util_copy(CCLQueue* cq, CCLBuffer* dst, CCLBuffer* src,
	size_t bytes, GError** err);

/* Bandwidth probes (benchmark utilities).  clo_util_stream_copy: `bytes`
 * (a multiple of 16 KiB) copied by blocks of one 16 KiB tile each, four
 * 16-byte loads in flight per thread, tiles XCD-contiguous: the streaming
 * ceiling.  clo_util_scatter_runs: the write pattern of one uniform 8-bit
 * radix pass over n (u64 key, u32 value) pairs (n a multiple of 8192) and
 * nothing else: tile t of 8192 pairs read contiguously and written as runs
 * of `run` (4, 8, 16, 32, 64 or 128) pairs, run d to d * (n * run / 8192) +
 * t * run; 32 is what uniform keys give an 8192-key tile with 256 digits, 4
 * what they would give with 2048 (11-bit digits). */
CCLEvent* clo_util_stream_copy(CCLQueue* cq, CCLBuffer* dst, CCLBuffer* src,
	size_t bytes, GError** err);
CCLEvent* clo_util_scatter_runs(CCLQueue* cq, CCLBuffer* kdst, CCLBuffer* vdst,
	CCLBuffer* ksrc, CCLBuffer* vsrc, size_t n, cl_uint run, GError** err);

/* Order-independent 64-bit multiset hash of n elements of `width` bytes
 * (sum of mix64(x) mod 2^64), computed on the device; blocks. */
cl_ulong clo_util_multiset_hash(CCLQueue* cq, CCLBuffer* buf, size_t n,
	size_t width, GError** err);

/* Checks on the device that n elements of `width` bytes (unsigned) are
 * non-decreasing; if vals is non-NULL also that vals (uint32) increase
 * within runs of equal keys (stability when vals held the input index).
 * Returns the number of violations; blocks. */
size_t clo_util_check_sorted(CCLQueue* cq, CCLBuffer* keys, CCLBuffer* vals,
	size_t n, size_t width, GError** err);

/* Checks a key-value sort whose input values were the indices 0..n-1:
 * vals_out must be a permutation of 0..n-1 and keys_out[i] must equal
 * keys_in[vals_out[i]] (unsigned keys of `width` bytes, n < 2^32).  Returns
 * the number of violations (out-of-range or repeated indices, mismatched
 * keys); blocks.  Together with clo_util_check_sorted on (keys_out,
 * vals_out) a zero result means the output is exactly the stable sort of
 * the input pairs. */
size_t clo_util_check_pairs(CCLQueue* cq, CCLBuffer* keys_in, CCLBuffer* keys_out,
	CCLBuffer* vals_out, size_t n, size_t width, GError** err);

/* Library build description (compiler, target, kernel variants). */
const char* clo_util_build_info(void);

/* Number of 32-bit words that differ between two device buffers of `words`
 * words (blocking). */
size_t clo_util_count_diff(CCLQueue* cq, CCLBuffer* a, CCLBuffer* b, size_t words,
	GError** err);

#ifdef __cplusplus
}
#endif
#endif

/*
 * oracle.c - CPU restatement of CL_Ops' hot-path algorithms.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by the tests, the smoke check and the
 * benchmark's CPU-baseline leg, never by the product path.  Every function
 * cites the reference file:line it restates.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * Parallel-for over [0, count) split into `threads` contiguous chunks.
 */
typedef void (*orc_range_fn)(void* ctx, size_t lo, size_t hi);

struct orc_job {
	orc_range_fn fn;
	void* ctx;
	size_t lo, hi;
};

static void* orc_job_run(void* p) {
	struct orc_job* j = (struct orc_job*)p;
	j->fn(j->ctx, j->lo, j->hi);
	return NULL;
}

static void orc_parallel_for(size_t count, int threads, orc_range_fn fn, void* ctx) {
	if (threads <= 1 || count < 2) {
		fn(ctx, 0, count);
		return;
	}
	if ((size_t)threads > count) threads = (int)count;
	pthread_t tid[256];
	struct orc_job jobs[256];
	if (threads > 256) threads = 256;
	size_t chunk = (count + threads - 1) / threads;
	int started = 0;
	for (int t = 0; t < threads; ++t) {
		size_t lo = (size_t)t * chunk, hi = lo + chunk;
		if (lo >= count) break;
		if (hi > count) hi = count;
		jobs[t].fn = fn;
		jobs[t].ctx = ctx;
		jobs[t].lo = lo;
		jobs[t].hi = hi;
		if (pthread_create(&tid[t], NULL, orc_job_run, &jobs[t]) != 0) {
			fn(ctx, lo, hi);
			jobs[t].fn = NULL;
		}
		++started;
	}
	for (int t = 0; t < started; ++t)
		if (jobs[t].fn) pthread_join(tid[t], NULL);
}

/* ---------------------------------------------------------------------------
 * Seeded inputs.  U(seed, i) = mix64(seed + (i+1) * 0x9E3779B97F4A7C15), mix64
 * the splitmix64 finalizer (Steele, Lea & Flood 2014).
 */
uint64_t orc_mix64(uint64_t z) {
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

static uint64_t orc_U(uint64_t seed, uint64_t i) {
	return orc_mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

/* "Presorted with 1% random swaps": position c of the ascending sequence
 * holds the value of position orc_swapped_position(seed, c).  Positions form
 * blocks of 100; block b is paired with block b ^ m, m = 1 + U(seed ^ S2, 0)
 * % 1023 (an involution inside each aligned group of 1024 blocks), and each
 * pair {lo, hi} exchanges one position of each: x = 100 lo + u % 100 with
 * y = 100 hi + (u >> 32) % 100, u = U(seed ^ S1, lo).  Every position is in
 * at most one exchange, so 2 positions in 200 move. */
uint64_t orc_swapped_position(uint64_t seed, uint64_t c) {
	const uint64_t m = 1 + orc_U(seed ^ 0x2545f4914f6cdd1dull, 0) % 1023;
	const uint64_t b = c / 100, b2 = b ^ m;
	const uint64_t lo = b < b2 ? b : b2, hi = b < b2 ? b2 : b;
	const uint64_t u = orc_U(seed ^ 0x5bd1e995ull, lo);
	const uint64_t x = lo * 100 + u % 100, y = hi * 100 + (u >> 32) % 100;
	return c == x ? y : (c == y ? x : c);
}

void orc_generate(int kind, uint64_t seed, uint64_t param, uint64_t start, size_t n, void* out) {
	for (size_t i = 0; i < n; ++i) {
		uint64_t c = i + start;
		switch (kind) {
		case 0: ((uint64_t*)out)[i] = orc_U(seed, c); break;
		case 1: ((uint32_t*)out)[i] = (uint32_t)(orc_U(seed, c) >> 32); break;
		case 2: ((uint64_t*)out)[i] = orc_U(seed, c) & param; break;
		case 3: ((uint32_t*)out)[i] = (uint32_t)(orc_U(seed, c) >> 32) & (uint32_t)param; break;
		case 4: ((uint32_t*)out)[i] = (uint32_t)c; break;
		case 6: ((uint32_t*)out)[i] = (uint32_t)(orc_U(seed, c) >> 32) % (uint32_t)param; break;
		case 7: ((uint64_t*)out)[i] = orc_U(seed, c) % param; break;
		case 8: ((uint32_t*)out)[i] = (uint32_t)(orc_swapped_position(seed, c) * param); break;
		case 9: ((uint64_t*)out)[i] = orc_swapped_position(seed, c) * param; break;
		case 10: ((uint32_t*)out)[i] = 0xFFFFFFFFu - (uint32_t)(c * param); break;
		case 11: ((uint64_t*)out)[i] = ~0ull - c * param; break;
		default: break;
		}
	}
}

void orc_zipf_cdf(double* cdf, size_t ranks, double s) {
	/* P(rank <= r) for weights (r+1)^-s, sums in long double. */
	long double tot = 0.0L, acc = 0.0L;
	for (size_t r = 0; r < ranks; ++r) tot += powl((long double)(r + 1), -s);
	for (size_t r = 0; r < ranks; ++r) {
		acc += powl((long double)(r + 1), -s);
		cdf[r] = (double)(acc / tot);
	}
	if (ranks) cdf[ranks - 1] = 1.0;
}

void orc_generate_zipf(const double* cdf, size_t ranks, uint64_t seed, size_t n, uint64_t* out) {
	for (size_t i = 0; i < n; ++i) {
		double u = (double)(orc_U(seed, i) >> 11) * 0x1.0p-53;
		size_t lo = 0, hi = ranks - 1;
		while (lo < hi) {
			size_t mid = (lo + hi) >> 1;
			if (cdf[mid] > u) hi = mid;
			else lo = mid + 1;
		}
		out[i] = orc_mix64(lo);
	}
}

/* ---------------------------------------------------------------------------
 * satradix with the reference's decomposition, one instance per key width.
 */
#define ORC_SATRADIX(T, SUF)                                                              \
	struct orc_sr_##SUF {                                                                 \
		const T* kin;                                                                     \
		const uint32_t* vin;                                                              \
		T* kout;                                                                          \
		uint32_t* vout;                                                                   \
		T* ktile;                                                                         \
		uint32_t* vtile;                                                                  \
		uint32_t* offsets;                                                                \
		uint32_t* counters;                                                               \
		size_t n, lws, nwg;                                                               \
		unsigned start, bits;                                                             \
	};                                                                                    \
	/* satradix_localsort (.cl:34-123) + satradix_histogram (.cl:125-222). */            \
	static void orc_sr_local_##SUF(void* p, size_t lo, size_t hi) {                       \
		struct orc_sr_##SUF* c = (struct orc_sr_##SUF*)p;                                 \
		const size_t R = (size_t)1 << c->bits;                                            \
		T* tk = (T*)malloc(c->lws * sizeof(T));                                           \
		uint32_t* tv = c->vin ? (uint32_t*)malloc(c->lws * 4) : NULL;                     \
		for (size_t wg = lo; wg < hi; ++wg) {                                             \
			const size_t b0 = wg * c->lws;                                                \
			const size_t cnt = (c->n - b0) < c->lws ? (c->n - b0) : c->lws;               \
			T* k = c->ktile + b0;                                                         \
			uint32_t* v = c->vtile ? c->vtile + b0 : NULL;                                \
			memcpy(k, c->kin + b0, cnt * sizeof(T));                                      \
			if (v) memcpy(v, c->vin + b0, cnt * 4);                                       \
			/* one stable split (zeros first) per digit bit, .cl:52-118 */                \
			for (unsigned b = c->start; b < c->start + c->bits && b < 8 * sizeof(T); ++b) { \
				size_t zeros = 0;                                                         \
				for (size_t i = 0; i < cnt; ++i) zeros += !((k[i] >> b) & 1);            \
				size_t z = 0, o = zeros;                                                  \
				for (size_t i = 0; i < cnt; ++i) {                                        \
					size_t pos = ((k[i] >> b) & 1) ? o++ : z++;                           \
					tk[pos] = k[i];                                                       \
					if (v) tv[pos] = v[i];                                                \
				}                                                                         \
				memcpy(k, tk, cnt * sizeof(T));                                           \
				if (v) memcpy(v, tv, cnt * 4);                                            \
			}                                                                             \
			/* per-block digit starts and counts, .cl:217-220 */                          \
			for (size_t d = 0; d < R; ++d) {                                              \
				c->offsets[wg * R + d] = 0;                                               \
				c->counters[d * c->nwg + wg] = 0;                                         \
			}                                                                             \
			for (size_t i = cnt; i-- > 0;) {                                              \
				size_t d = (size_t)((k[i] >> c->start) & (R - 1));                        \
				c->offsets[wg * R + d] = (uint32_t)i;                                     \
				c->counters[d * c->nwg + wg]++;                                           \
			}                                                                             \
		}                                                                                 \
		free(tk);                                                                         \
		free(tv);                                                                         \
	}                                                                                     \
	/* satradix_scatter (.cl:224-258). */                                                 \
	static void orc_sr_scatter_##SUF(void* p, size_t lo, size_t hi) {                     \
		struct orc_sr_##SUF* c = (struct orc_sr_##SUF*)p;                                 \
		const size_t R = (size_t)1 << c->bits;                                            \
		for (size_t wg = lo; wg < hi; ++wg) {                                             \
			const size_t b0 = wg * c->lws;                                                \
			const size_t cnt = (c->n - b0) < c->lws ? (c->n - b0) : c->lws;               \
			for (size_t lid = 0; lid < cnt; ++lid) {                                      \
				T key = c->ktile[b0 + lid];                                               \
				size_t d = (size_t)((key >> c->start) & (R - 1));                         \
				size_t o = c->counters[d * c->nwg + wg] + lid - c->offsets[wg * R + d];   \
				c->kout[o] = key;                                                         \
				if (c->vtile) c->vout[o] = c->vtile[b0 + lid];                            \
			}                                                                             \
		}                                                                                 \
	}                                                                                     \
	static void orc_satradix_##SUF(T* keys, uint32_t* vals, size_t n, int bits, size_t lws,  \
		int threads) {                                                                    \
		const size_t R = (size_t)1 << bits;                                               \
		const size_t nwg = (n + lws - 1) / lws;                                           \
		const unsigned passes = (unsigned)((sizeof(T) * 8 + bits - 1) / bits);            \
		struct orc_sr_##SUF c;                                                            \
		T* kaux = (T*)malloc(n * sizeof(T));                                              \
		uint32_t* vaux = vals ? (uint32_t*)malloc(n * 4) : NULL;                          \
		c.ktile = (T*)malloc(n * sizeof(T));                                              \
		c.vtile = vals ? (uint32_t*)malloc(n * 4) : NULL;                                 \
		c.offsets = (uint32_t*)malloc(nwg * R * 4);                                       \
		c.counters = (uint32_t*)malloc(nwg * R * 4);                                      \
		c.n = n;                                                                          \
		c.lws = lws;                                                                      \
		c.nwg = nwg;                                                                      \
		c.bits = (unsigned)bits;                                                          \
		T* src = keys;                                                                    \
		uint32_t* vsrc = vals;                                                            \
		T* dst = kaux;                                                                    \
		uint32_t* vdst = vaux;                                                            \
		for (unsigned p = 0; p < passes; ++p) {                                           \
			c.start = p * (unsigned)bits;                                                 \
			c.kin = src;                                                                  \
			c.vin = vsrc;                                                                 \
			c.kout = dst;                                                                 \
			c.vout = vdst;                                                                \
			orc_parallel_for(nwg, threads, orc_sr_local_##SUF, &c);                       \
			/* exclusive scan of the digit-major counters (.c:298-300) */                 \
			uint32_t run = 0;                                                             \
			for (size_t i = 0; i < nwg * R; ++i) {                                        \
				uint32_t t = c.counters[i];                                               \
				c.counters[i] = run;                                                      \
				run += t;                                                                 \
			}                                                                             \
			orc_parallel_for(nwg, threads, orc_sr_scatter_##SUF, &c);                     \
			T* tk = src;                                                                  \
			src = dst;                                                                    \
			dst = tk;                                                                     \
			uint32_t* tv = vsrc;                                                          \
			vsrc = vdst;                                                                  \
			vdst = tv;                                                                    \
		}                                                                                 \
		if (src != keys) {                                                                \
			memcpy(keys, src, n * sizeof(T));                                             \
			if (vals) memcpy(vals, vsrc, n * 4);                                          \
		}                                                                                 \
		free(kaux);                                                                       \
		free(vaux);                                                                       \
		free(c.ktile);                                                                    \
		free(c.vtile);                                                                    \
		free(c.offsets);                                                                  \
		free(c.counters);                                                                 \
	}

ORC_SATRADIX(uint8_t, u8)
ORC_SATRADIX(uint16_t, u16)
ORC_SATRADIX(uint32_t, u32)
ORC_SATRADIX(uint64_t, u64)

void orc_satradix(void* keys, uint32_t* vals, size_t n, int width, int bits, size_t lws,
	int threads) {
	if (n == 0 || bits < 1 || bits > 16 || lws == 0) return;
	switch (width) {
	case 1: orc_satradix_u8((uint8_t*)keys, vals, n, bits, lws, threads); break;
	case 2: orc_satradix_u16((uint16_t*)keys, vals, n, bits, lws, threads); break;
	case 4: orc_satradix_u32((uint32_t*)keys, vals, n, bits, lws, threads); break;
	case 8: orc_satradix_u64((uint64_t*)keys, vals, n, bits, lws, threads); break;
	default: break;
	}
}

/* ---------------------------------------------------------------------------
 * Plain stable LSD counting sort (8-bit digits).
 */
static uint64_t orc_load(const void* p, size_t i, int w) {
	switch (w) {
	case 1: return ((const uint8_t*)p)[i];
	case 2: return ((const uint16_t*)p)[i];
	case 4: return ((const uint32_t*)p)[i];
	default: return ((const uint64_t*)p)[i];
	}
}

static void orc_store(void* p, size_t i, int w, uint64_t v) {
	switch (w) {
	case 1: ((uint8_t*)p)[i] = (uint8_t)v; break;
	case 2: ((uint16_t*)p)[i] = (uint16_t)v; break;
	case 4: ((uint32_t*)p)[i] = (uint32_t)v; break;
	default: ((uint64_t*)p)[i] = v; break;
	}
}

void orc_stable_sort(void* keys, uint32_t* vals, size_t n, int width) {
	if (n < 2) return;
	uint64_t* k = (uint64_t*)malloc(n * 8);
	uint64_t* k2 = (uint64_t*)malloc(n * 8);
	uint32_t* v2 = vals ? (uint32_t*)malloc(n * 4) : NULL;
	for (size_t i = 0; i < n; ++i) k[i] = orc_load(keys, i, width);
	for (int p = 0; p < width; ++p) {
		size_t cnt[257] = {0};
		const int sh = 8 * p;
		for (size_t i = 0; i < n; ++i) cnt[((k[i] >> sh) & 255) + 1]++;
		for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
		for (size_t i = 0; i < n; ++i) {
			size_t o = cnt[(k[i] >> sh) & 255]++;
			k2[o] = k[i];
			if (vals) v2[o] = vals[i];
		}
		uint64_t* t = k;
		k = k2;
		k2 = t;
		if (vals) memcpy(vals, v2, n * 4);
	}
	for (size_t i = 0; i < n; ++i) orc_store(keys, i, width, k[i]);
	free(k);
	free(k2);
	free(v2);
}

/* ---------------------------------------------------------------------------
 * Bitonic network of sbitonic.cl:38-69, loop of sbitonic.c:102-118.
 */
static int orc_key_gt(uint64_t a, uint64_t b, int width, int sign) {
	if (!sign) return a > b;
	const int sh = 64 - 8 * width;
	int64_t sa = (int64_t)(a << sh) >> sh, sb = (int64_t)(b << sh) >> sh;
	return sa > sb;
}

void orc_sbitonic(void* keys, size_t n, int width, int sign, int desc) {
	if (n < 2) return;
	unsigned stages = 0;
	while (((size_t)1 << stages) < n) ++stages;
	const size_t gws = n / 2;
	for (unsigned stage = 1; stage <= stages; ++stage) {
		for (unsigned step = stage; step > 0; --step) {
			const size_t stride = (size_t)1 << (step - 1);
			for (size_t gid = 0; gid < gws; ++gid) {
				size_t i1 = gid + (gid / stride) * stride, i2 = i1 + stride;
				uint64_t d1 = orc_load(keys, i1, width), d2 = orc_load(keys, i2, width);
				int dsc = (int)((gid >> (stage - 1)) & 1);
				/* CLO_SORT_COMPARE(k1, k2): ((a) > (b)) or ((a) < (b)) */
				int cmp = desc ? orc_key_gt(d2, d1, width, sign) : orc_key_gt(d1, d2, width, sign);
				if (cmp ^ dsc) {
					orc_store(keys, i1, width, d2);
					orc_store(keys, i2, width, d1);
				}
			}
		}
	}
}

/* gselect.cl:38-57 */
void orc_gselect(const void* keys, const uint32_t* vals, void* keys_out, uint32_t* vals_out,
	size_t n, int width, int sign, int desc) {
	for (size_t g = 0; g < n; ++g) {
		uint64_t kg = orc_load(keys, g, width);
		size_t pos = 0;
		for (size_t i = 0; i < n; ++i) {
			uint64_t ki = orc_load(keys, i, width);
			int cmp = desc ? orc_key_gt(ki, kg, width, sign) : orc_key_gt(kg, ki, width, sign);
			if (cmp || (ki == kg && i < g)) ++pos;
		}
		orc_store(keys_out, pos, width, kg);
		if (vals) vals_out[pos] = vals[g];
	}
}

/* ---------------------------------------------------------------------------
 * Scans.
 */
static uint64_t orc_convert(uint64_t raw, int w, int sign) {
	if (!sign || w == 8) return raw;
	const int sh = 64 - 8 * w;
	return (uint64_t)((int64_t)(raw << sh) >> sh); /* C sign extension */
}

void orc_scan(const void* in, int in_width, int in_sign, void* out, int sum_width, size_t n) {
	uint64_t acc = 0;
	for (size_t i = 0; i < n; ++i) {
		orc_store(out, i, sum_width, acc);
		acc += orc_convert(orc_load(in, i, in_width), in_width, in_sign);
	}
}

struct orc_bl {
	const uint32_t* in;
	void* out;
	uint64_t* wgsum;
	int sw;
	size_t n, lws, bpw, nwg;
};

/* workgroupScan (clo_scan_blelloch.cl:49-126): serial blocks of 2*lws per
 * work-group, tree up-sweep and down-sweep per block, carried block sums. */
static void orc_bl_wgscan(void* p, size_t lo, size_t hi) {
	struct orc_bl* c = (struct orc_bl*)p;
	const size_t bs = 2 * c->lws;
	uint64_t* aux = (uint64_t*)malloc(bs * 8);
	const uint64_t mask = c->sw == 8 ? ~0ull : 0xFFFFFFFFull;
	for (size_t wg = lo; wg < hi; ++wg) {
		uint64_t in_sum = 0;
		for (size_t b = 0; b < c->bpw; ++b) {
			size_t blk = wg * c->bpw + b;
			if (blk * bs >= c->n) break;
			size_t g0 = blk * bs;
			for (size_t i = 0; i < bs; ++i) aux[i] = g0 + i < c->n ? c->in[g0 + i] : 0;
			size_t offset = 1;
			for (size_t d = bs >> 1; d > 0; d >>= 1) {
				for (size_t lid = 0; lid < d; ++lid) {
					size_t ai = offset * (2 * lid + 1) - 1, bi = offset * (2 * lid + 2) - 1;
					aux[bi] = (aux[bi] + aux[ai]) & mask;
				}
				offset *= 2;
			}
			uint64_t prev = in_sum;
			in_sum = (in_sum + aux[bs - 1]) & mask;
			aux[bs - 1] = 0;
			for (size_t d = 1; d < bs; d *= 2) {
				offset >>= 1;
				for (size_t lid = 0; lid < d; ++lid) {
					size_t ai = offset * (2 * lid + 1) - 1, bi = offset * (2 * lid + 2) - 1;
					uint64_t t = aux[ai];
					aux[ai] = aux[bi];
					aux[bi] = (aux[bi] + t) & mask;
				}
			}
			for (size_t i = 0; i < bs && g0 + i < c->n; ++i)
				orc_store(c->out, g0 + i, c->sw, (aux[i] + prev) & mask);
		}
		c->wgsum[wg] = in_sum;
	}
	free(aux);
}

/* addWorkgroupSums (clo_scan_blelloch.cl:193-211). */
static void orc_bl_add(void* p, size_t lo, size_t hi) {
	struct orc_bl* c = (struct orc_bl*)p;
	const size_t per = c->bpw * 2 * c->lws;
	for (size_t wg = lo; wg < hi; ++wg) {
		const uint64_t add = c->wgsum[wg];
		for (size_t i = wg * per; i < (wg + 1) * per && i < c->n; ++i)
			orc_store(c->out, i, c->sw, orc_load(c->out, i, c->sw) + add);
	}
}

void orc_scan_blelloch(const uint32_t* in, void* out, int sum_width, size_t n, size_t lws,
	int threads) {
	if (n == 0) return;
	struct orc_bl c;
	c.in = in;
	c.out = out;
	c.sw = sum_width;
	c.n = n;
	c.lws = lws;
	/* clo_scan_blelloch.c:130-140: at most lws^2 work-items, lws per group. */
	size_t gws = ((n / 2 + lws - 1) / lws) * lws;
	if (gws > lws * lws) gws = lws * lws;
	if (gws == 0) gws = lws;
	c.nwg = gws / lws;
	const size_t bs = 2 * lws;
	c.bpw = ((n + bs - 1) / bs + c.nwg - 1) / c.nwg;
	c.wgsum = (uint64_t*)calloc(c.nwg, 8);
	orc_parallel_for(c.nwg, threads, orc_bl_wgscan, &c);
	/* workgroupSumsScan (clo_scan_blelloch.cl:134-182): exclusive scan. */
	uint64_t run = 0;
	for (size_t w = 0; w < c.nwg; ++w) {
		uint64_t t = c.wgsum[w];
		c.wgsum[w] = run;
		run += t;
	}
	orc_parallel_for(c.nwg, threads, orc_bl_add, &c);
	free(c.wgsum);
}

/* ---------------------------------------------------------------------------
 * RNG (src/cl_ops/rng/clo_rng_*.cl).
 */
size_t orc_rng_state_size(int gen) {
	static const size_t sz[6] = {8, 8, 16, 8, 4, 16}; /* clo_rng.c:60-68 */
	return (gen >= 0 && gen < 6) ? sz[gen] : 0;
}

void orc_rng_from_seed(int gen, uint64_t seed, void* state) {
	switch (gen) {
	case 0: /* clo_rng_lcg.cl:32 */
	case 1: /* clo_rng_xorshift64.cl:29 */
		*(uint64_t*)state = seed;
		break;
	case 2: { /* clo_rng_xorshift128.cl:30 */
		uint32_t* s = (uint32_t*)state;
		s[0] = (uint32_t)(seed & 0xFFFFFFFFull);
		s[1] = (uint32_t)((seed >> 16) & 0xFFFFFFFFull);
		s[2] = (uint32_t)((seed >> 32) & 0xFFFFFFFFull);
		s[3] = (uint32_t)((seed >> 46) & 0xFFFFFFFFull);
		break;
	}
	case 3: { /* clo_rng_mwc64x.cl:32, as_uint2: little-endian halves */
		uint32_t* s = (uint32_t*)state;
		s[0] = (uint32_t)seed;
		s[1] = (uint32_t)(seed >> 32);
		break;
	}
	case 4: /* clo_rng_parkmiller.cl:31 */
		*(int32_t*)state = (int32_t)(uint32_t)(seed & 0xFFFFFFFFull);
		break;
	case 5: { /* clo_rng_tauslcg.cl:35, as_uint4((ulong2)(seed, seed)) */
		uint32_t* s = (uint32_t*)state;
		s[0] = s[2] = (uint32_t)seed;
		s[1] = s[3] = (uint32_t)(seed >> 32);
		break;
	}
	default: break;
	}
}

/* One Tausworthe step (clo_rng_tauslcg.cl:52-55). */
uint32_t orc_taus(uint32_t z, int s1, int s2, int s3, uint32_t m) {
	uint32_t b = ((z << s1) ^ z) >> s2;
	return ((z & m) << s3) ^ b;
}

uint32_t orc_rng_next(int gen, void* state) {
	switch (gen) {
	case 0: { /* clo_rng_lcg.cl:43-59 */
		uint64_t s = *(uint64_t*)state;
		s = (s * 0x5DEECE66Dull + 0xBull) & ((1ull << 48) - 1);
		*(uint64_t*)state = s;
		return (uint32_t)(s >> 16);
	}
	case 1: { /* clo_rng_xorshift64.cl:40-63 */
		uint64_t s = *(uint64_t*)state;
		s ^= (s << 21);
		s ^= (s >> 35);
		s ^= (s << 4);
		*(uint64_t*)state = s;
		return (uint32_t)s;
	}
	case 2: { /* clo_rng_xorshift128.cl:41-59 */
		uint32_t* s = (uint32_t*)state;
		uint32_t t = s[0] ^ (s[0] << 11);
		s[0] = s[1];
		s[1] = s[2];
		s[2] = s[3];
		s[3] = s[3] ^ (s[3] >> 19) ^ (t ^ (t >> 8));
		return s[3];
	}
	case 3: { /* clo_rng_mwc64x.cl:43-63 */
		uint32_t* s = (uint32_t*)state;
		const uint32_t A = 4294883355u;
		uint32_t x = s[0], c = s[1];
		uint32_t res = x ^ c;
		uint32_t hi = (uint32_t)(((uint64_t)x * A) >> 32);
		x = x * A + c;
		c = hi + (x < c);
		s[0] = x;
		s[1] = c;
		return res;
	}
	case 4: { /* clo_rng_parkmiller.cl:43-59: C remainder truncates toward 0 */
		int32_t s = *(int32_t*)state;
		s = (int32_t)(((int64_t)s * 16807) % 2147483647);
		*(int32_t*)state = s;
		return ((uint32_t)s) << 1;
	}
	case 5: { /* clo_rng_tauslcg.cl:52-105 */
		uint32_t* s = (uint32_t*)state;
		uint32_t x = s[0];
		s[0] = orc_taus(s[1], 13, 19, 12, 4294967294u);
		s[1] = orc_taus(s[2], 2, 25, 4, 4294967288u);
		s[2] = orc_taus(s[3], 3, 11, 17, 4294967294u);
		s[3] = 1664525u * x + 1013904223u;
		return s[0];
	}
	default: return 0;
	}
}

static uint64_t orc_hash(uint64_t x, int hash) {
	if (hash == 1) { /* KNUTH(x), clo_rng_init.cl:27 */
		x = (x * 2654435761ull) % 0x100000000ull;
	} else if (hash == 2) { /* XS1(x), clo_rng_init.cl:30-33 */
		x = ((x >> 16) ^ x) * 0x45d9f3bull;
		x = ((x >> 16) ^ x) * 0x45d9f3bull;
		x = ((x >> 16) ^ x);
	}
	return x;
}

void orc_rng_init_dev_gid(int gen, size_t count, uint64_t main_seed, int hash, void* states) {
	const size_t ss = orc_rng_state_size(gen);
	for (size_t gid = 0; gid < count; ++gid)
		orc_rng_from_seed(gen, orc_hash(gid + main_seed, hash), (char*)states + gid * ss);
}

/* MT19937, Matsumoto & Nishimura (1998), init_genrand seeding. */
void orc_mt19937(uint32_t seed, size_t n, uint32_t* out) {
	uint32_t mt[624];
	int mti;
	mt[0] = seed;
	for (mti = 1; mti < 624; ++mti)
		mt[mti] = 1812433253u * (mt[mti - 1] ^ (mt[mti - 1] >> 30)) + (uint32_t)mti;
	for (size_t i = 0; i < n; ++i) {
		if (mti >= 624) {
			int kk;
			uint32_t y;
			for (kk = 0; kk < 624 - 397; ++kk) {
				y = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
				mt[kk] = mt[kk + 397] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
			}
			for (; kk < 623; ++kk) {
				y = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
				mt[kk] = mt[kk + (397 - 624)] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
			}
			y = (mt[623] & 0x80000000u) | (mt[0] & 0x7fffffffu);
			mt[623] = mt[396] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
			mti = 0;
		}
		uint32_t y = mt[mti++];
		y ^= (y >> 11);
		y ^= (y << 7) & 0x9d2c5680u;
		y ^= (y << 15) & 0xefc60000u;
		y ^= (y >> 18);
		out[i] = y;
	}
}

void orc_rng_init_host_mt(int gen, size_t count, uint64_t main_seed, void* states) {
	/* clo_rng.c:196-203: g_rand_new_with_seed((guint32) main_seed); word i of
	 * the seeds buffer = i-th g_rand_int. */
	orc_mt19937((uint32_t)main_seed, count * orc_rng_state_size(gen) / 4, (uint32_t*)states);
}

struct orc_fill {
	int gen;
	char* states;
	size_t G, ss;
	uint32_t draws, bits, maxint;
	uint32_t* out;
};

static void orc_fill_range(void* p, size_t lo, size_t hi) {
	struct orc_fill* c = (struct orc_fill*)p;
	for (size_t gid = lo; gid < hi; ++gid) {
		void* st = c->states + gid * c->ss;
		for (uint32_t r = 0; r < c->draws; ++r) {
			uint32_t v = orc_rng_next(c->gen, st);
			c->out[(size_t)r * c->G + gid] = c->maxint ? v % c->maxint : v >> (32 - c->bits);
		}
	}
}

void orc_rng_fill(int gen, void* states, size_t G, uint32_t draws, uint32_t bits, uint32_t maxint,
	uint32_t* out, int threads) {
	struct orc_fill c;
	c.gen = gen;
	c.states = (char*)states;
	c.G = G;
	c.ss = orc_rng_state_size(gen);
	c.draws = draws;
	c.bits = bits;
	c.maxint = maxint;
	c.out = out;
	orc_parallel_for(G, threads, orc_fill_range, &c);
}

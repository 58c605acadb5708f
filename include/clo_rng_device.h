/*
 * clo_rng_device.h - HIP device functions of the generators, for kernels
 * written by callers (the counterpart of the OpenCL source CL_Ops hands out
 * through clo_rng_get_source, src/cl_ops/rng/clo_rng.c:438-446).
 *
 * Each generator lives in its own namespace with the reference's names:
 *   clo_statetype                     per-work-item state
 *   clo_ulong2statetype(seed)         seed -> state (the seed conversion)
 *   clo_rng_step(state&)              one draw, state in registers
 *   clo_rng_next(states, index)       one draw from states[index] (global)
 * and the workitem API of clo_rng_api.cl/clo_rng_workitem.cl:
 *   clo_rng_next_int{,2,4,8}(states, n)  draws at gid + k*global_size, % n
 * Use one generator per kernel: `using namespace clo_rng_xorshift128;`.
 *
 * clo_rng_get_source() returns the text of one generator section followed by
 * the API section.  The lines between "@@section" markers are extracted
 * verbatim at build time.
 */
#ifndef CLO_RNG_DEVICE_H_
#define CLO_RNG_DEVICE_H_

#include <hip/hip_runtime.h>

/* @@section lcg */
/* Linear congruential generator of java.util.Random
 * (src/cl_ops/rng/clo_rng_lcg.cl:32-59): 48-bit state, output = top 32. */
namespace clo_rng_lcg {
typedef unsigned long long clo_statetype;
__host__ __device__ inline clo_statetype clo_ulong2statetype(unsigned long long seed) {
	return seed;
}
__host__ __device__ inline unsigned int clo_rng_step(clo_statetype& s) {
	s = (s * 0x5DEECE66DULL + 0xBULL) & ((1ULL << 48) - 1);
	return (unsigned int)(s >> 16);
}
__device__ inline unsigned int clo_rng_next(clo_statetype* states, unsigned int index) {
	clo_statetype s = states[index];
	unsigned int r = clo_rng_step(s);
	states[index] = s;
	return r;
}
}  // namespace clo_rng_lcg
/* @@end */

/* @@section xorshift64 */
/* Marsaglia xorshift (21, 35, 4) on 64 bits, output = low 32 bits
 * (src/cl_ops/rng/clo_rng_xorshift64.cl:29-63). */
namespace clo_rng_xorshift64 {
typedef unsigned long long clo_statetype;
__host__ __device__ inline clo_statetype clo_ulong2statetype(unsigned long long seed) {
	return seed;
}
__host__ __device__ inline unsigned int clo_rng_step(clo_statetype& s) {
	s ^= (s << 21);
	s ^= (s >> 35);
	s ^= (s << 4);
	return (unsigned int)s;
}
__device__ inline unsigned int clo_rng_next(clo_statetype* states, unsigned int index) {
	clo_statetype s = states[index];
	unsigned int r = clo_rng_step(s);
	states[index] = s;
	return r;
}
}  // namespace clo_rng_xorshift64
/* @@end */

/* @@section xorshift128 */
/* Marsaglia xor128 (src/cl_ops/rng/clo_rng_xorshift128.cl:30-59).  The state
 * is (x, y, z, w) in memory order; the seed conversion takes 32-bit windows of
 * the 64-bit seed at bit offsets 0, 16, 32 and 46. */
namespace clo_rng_xorshift128 {
typedef uint4 clo_statetype;
__host__ __device__ inline clo_statetype clo_ulong2statetype(unsigned long long seed) {
	clo_statetype s;
	s.x = (unsigned int)(seed & 0xFFFFFFFFULL);
	s.y = (unsigned int)((seed >> 16) & 0xFFFFFFFFULL);
	s.z = (unsigned int)((seed >> 32) & 0xFFFFFFFFULL);
	s.w = (unsigned int)((seed >> 46) & 0xFFFFFFFFULL);
	return s;
}
__host__ __device__ inline unsigned int clo_rng_step(clo_statetype& s) {
	unsigned int t = s.x ^ (s.x << 11);
	s.x = s.y;
	s.y = s.z;
	s.z = s.w;
	s.w = s.w ^ (s.w >> 19) ^ (t ^ (t >> 8));
	return s.w;
}
__device__ inline unsigned int clo_rng_next(clo_statetype* states, unsigned int index) {
	clo_statetype s = states[index];
	unsigned int r = clo_rng_step(s);
	states[index] = s;
	return r;
}
}  // namespace clo_rng_xorshift128
/* @@end */

/* @@section mwc64x */
/* Multiply-with-carry, A = 4294883355 (src/cl_ops/rng/clo_rng_mwc64x.cl:32-63).
 * State (x, c) = (low, high) 32-bit halves of the seed; output x ^ c. */
namespace clo_rng_mwc64x {
typedef uint2 clo_statetype;
__host__ __device__ inline clo_statetype clo_ulong2statetype(unsigned long long seed) {
	clo_statetype s;
	s.x = (unsigned int)(seed & 0xFFFFFFFFULL);
	s.y = (unsigned int)(seed >> 32);
	return s;
}
__host__ __device__ inline unsigned int clo_rng_step(clo_statetype& s) {
	const unsigned int A = 4294883355U;
	unsigned int x = s.x, c = s.y;
	unsigned int res = x ^ c;
	unsigned long long p = (unsigned long long)x * A;
	unsigned int hi = (unsigned int)(p >> 32);
	x = x * A + c;
	c = hi + (x < c);
	s.x = x;
	s.y = c;
	return res;
}
__device__ inline unsigned int clo_rng_next(clo_statetype* states, unsigned int index) {
	clo_statetype s = states[index];
	unsigned int r = clo_rng_step(s);
	states[index] = s;
	return r;
}
}  // namespace clo_rng_mwc64x
/* @@end */

/* @@section parkmiller */
/* Park-Miller minimal standard, a = 16807, m = 2^31 - 1, on a signed 32-bit
 * state with C's truncating remainder (src/cl_ops/rng/clo_rng_parkmiller.cl:
 * 31-59); output = state << 1.  The remainder uses m = 2^31 - 1:
 * p mod m = (p & m) + (p >> 31), reduced once, applied to |p|. */
namespace clo_rng_parkmiller {
typedef int clo_statetype;
__host__ __device__ inline clo_statetype clo_ulong2statetype(unsigned long long seed) {
	return (int)(unsigned int)(seed & 0xFFFFFFFFULL);
}
__host__ __device__ inline unsigned int clo_rng_step(clo_statetype& s) {
	long long p = (long long)s * 16807LL;
	unsigned long long a = (unsigned long long)(p < 0 ? -p : p);
	unsigned long long r = (a & 0x7FFFFFFFULL) + (a >> 31);
	if (r >= 0x7FFFFFFFULL) r -= 0x7FFFFFFFULL;
	s = p < 0 ? -(int)r : (int)r;
	return ((unsigned int)s) << 1;
}
__device__ inline unsigned int clo_rng_next(clo_statetype* states, unsigned int index) {
	clo_statetype s = states[index];
	unsigned int r = clo_rng_step(s);
	states[index] = s;
	return r;
}
}  // namespace clo_rng_parkmiller
/* @@end */

/* @@section tauslcg */
/* Hybrid Tausworthe/LCG generator as CL_Ops runs it
 * (src/cl_ops/rng/clo_rng_tauslcg.cl:31-105): state (x, y, z, w) = the 64-bit
 * seed's (low, high, low, high) 32-bit halves.  One step moves every word one
 * slot down while advancing it: x <- T1(y), y <- T2(z), z <- T3(w),
 * w <- 1664525 * x + 1013904223 (the old x), and returns the new x, where
 * T(v) = ((v & m) << s3) ^ (((v << s1) ^ v) >> s2) with (s1, s2, s3, m) =
 * (13, 19, 12, ~1), (2, 25, 4, ~7), (3, 11, 17, ~1). */
namespace clo_rng_tauslcg {
typedef uint4 clo_statetype;
__host__ __device__ inline clo_statetype clo_ulong2statetype(unsigned long long seed) {
	clo_statetype s;
	s.x = (unsigned int)(seed & 0xFFFFFFFFULL);
	s.y = (unsigned int)(seed >> 32);
	s.z = s.x;
	s.w = s.y;
	return s;
}
__host__ __device__ inline unsigned int clo_taus(unsigned int v, int s1, int s2, int s3,
	unsigned int m) {
	return ((v & m) << s3) ^ (((v << s1) ^ v) >> s2);
}
__host__ __device__ inline unsigned int clo_rng_step(clo_statetype& s) {
	const unsigned int x = s.x;
	s.x = clo_taus(s.y, 13, 19, 12, 4294967294U);
	s.y = clo_taus(s.z, 2, 25, 4, 4294967288U);
	s.z = clo_taus(s.w, 3, 11, 17, 4294967294U);
	s.w = 1664525U * x + 1013904223U;
	return s.x;
}
__device__ inline unsigned int clo_rng_next(clo_statetype* states, unsigned int index) {
	clo_statetype s = states[index];
	unsigned int r = clo_rng_step(s);
	states[index] = s;
	return r;
}
}  // namespace clo_rng_tauslcg
/* @@end */

/* @@section api */
/* Workitem API (src/cl_ops/rng/clo_rng_api.cl:33-105 and
 * clo_rng_workitem.cl:24-32) over whichever generator namespace is in use:
 * draw k of a work-item comes from state gid + k * global_size. */
#ifndef CLO_RNG_API_DEFINED
#define CLO_RNG_API_DEFINED
#define CLO_RNG_GLOBAL_SIZE() (gridDim.x * blockDim.x)
#define CLO_RNG_GID1() (blockIdx.x * blockDim.x + threadIdx.x)
#define CLO_RNG_DEFINE_API(NS)                                                        \
	namespace NS {                                                                    \
	__device__ inline unsigned int clo_rng_next_int(clo_statetype* st, unsigned int n) { \
		return clo_rng_next(st, CLO_RNG_GID1()) % n;                                  \
	}                                                                                 \
	__device__ inline uint2 clo_rng_next_int2(clo_statetype* st, unsigned int n) {    \
		unsigned int g = CLO_RNG_GID1(), G = CLO_RNG_GLOBAL_SIZE();                    \
		return make_uint2(clo_rng_next(st, g) % n, clo_rng_next(st, G + g) % n);       \
	}                                                                                 \
	__device__ inline uint4 clo_rng_next_int4(clo_statetype* st, unsigned int n) {    \
		unsigned int g = CLO_RNG_GID1(), G = CLO_RNG_GLOBAL_SIZE();                    \
		return make_uint4(clo_rng_next(st, g) % n, clo_rng_next(st, G + g) % n,        \
			clo_rng_next(st, 2 * G + g) % n, clo_rng_next(st, 3 * G + g) % n);         \
	}                                                                                 \
	__device__ inline void clo_rng_next_int8(clo_statetype* st, unsigned int n,       \
		unsigned int out[8]) {                                                        \
		unsigned int g = CLO_RNG_GID1(), G = CLO_RNG_GLOBAL_SIZE();                    \
		for (unsigned int k = 0; k < 8; ++k) out[k] = clo_rng_next(st, k * G + g) % n; \
	}                                                                                 \
	}
#endif
/* @@end */

CLO_RNG_DEFINE_API(clo_rng_lcg)
CLO_RNG_DEFINE_API(clo_rng_xorshift64)
CLO_RNG_DEFINE_API(clo_rng_xorshift128)
CLO_RNG_DEFINE_API(clo_rng_mwc64x)
CLO_RNG_DEFINE_API(clo_rng_parkmiller)
CLO_RNG_DEFINE_API(clo_rng_tauslcg)

#endif /* CLO_RNG_DEVICE_H_ */

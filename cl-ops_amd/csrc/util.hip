// Seeded device input generators, a calibration copy kernel and device-side
// result checks (additions with no counterpart in CL_Ops).

#include <cmath>
#include <cstring>

#include "clo_internal.hpp"
#include "dev_common.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

constexpr u64 kGolden = 0x9E3779B97F4A7C15ull;
constexpr u64 kPerturb = 0x5bd1e995ull;

// splitmix64 finalizer (Steele, Lea & Flood, OOPSLA 2014).
__host__ __device__ __forceinline__ u64 mix64(u64 z) {
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

__device__ __forceinline__ u64 U(u64 seed, u64 i) { return mix64(seed + (i + 1) * kGolden); }

// Position whose ascending value position c holds in the "presorted with 1%
// random swaps" sequence: blocks of 100 positions are paired (block b with
// b ^ m, m = 1 + U(seed ^ kPairing, 0) % 1023) and each pair exchanges one
// position of each block, chosen by u = U(seed ^ kPerturb, lower block).
constexpr u64 kPairing = 0x2545f4914f6cdd1dull;
__device__ __forceinline__ u64 swapped_position(u64 seed, u64 c) {
	const u64 m = 1 + U(seed ^ kPairing, 0) % 1023;
	const u64 b = c / 100, b2 = b ^ m;
	const u64 lo = b < b2 ? b : b2, hi = b < b2 ? b2 : b;
	const u64 u = U(seed ^ kPerturb, lo);
	const u64 x = lo * 100 + u % 100, y = hi * 100 + (u >> 32) % 100;
	return c == x ? y : (c == y ? x : c);
}

template <int KIND>
__global__ __launch_bounds__(256) void util_generate(void* out, u64 n, u64 seed, u64 param,
	u64 start) {
	const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
	for (u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		const u64 c = i + start;
		switch (KIND) {
		case CLO_GEN_U64:
			static_cast<u64*>(out)[i] = U(seed, c);
			break;
		case CLO_GEN_U32_HI:
			static_cast<u32*>(out)[i] = static_cast<u32>(U(seed, c) >> 32);
			break;
		case CLO_GEN_U64_MASK:
			static_cast<u64*>(out)[i] = U(seed, c) & param;
			break;
		case CLO_GEN_U32_MASK:
			static_cast<u32*>(out)[i] = static_cast<u32>(U(seed, c) >> 32) & static_cast<u32>(param);
			break;
		case CLO_GEN_IOTA_U32:
			static_cast<u32*>(out)[i] = static_cast<u32>(c);
			break;
		case CLO_GEN_U32_MOD:
			static_cast<u32*>(out)[i] = static_cast<u32>(U(seed, c) >> 32) % static_cast<u32>(param);
			break;
		case CLO_GEN_U64_MOD:
			static_cast<u64*>(out)[i] = U(seed, c) % param;
			break;
		case CLO_GEN_NEARLY_SORTED_U32:
			static_cast<u32*>(out)[i] = static_cast<u32>(swapped_position(seed, c) * param);
			break;
		case CLO_GEN_NEARLY_SORTED_U64:
			static_cast<u64*>(out)[i] = swapped_position(seed, c) * param;
			break;
		case CLO_GEN_REVERSED_U32:
			static_cast<u32*>(out)[i] = 0xFFFFFFFFu - static_cast<u32>(c * param);
			break;
		case CLO_GEN_REVERSED_U64:
			static_cast<u64*>(out)[i] = ~0ull - c * param;
			break;
		}
	}
}

template <typename T>
__global__ __launch_bounds__(256) void util_generate_zipf(T* out, u64 n, const double* cdf,
	u64 ranks, u64 seed) {
	const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
	for (u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		const double u = static_cast<double>(U(seed, i) >> 11) * 0x1.0p-53;
		// Smallest r with cdf[r] > u.
		u64 lo = 0, hi = ranks - 1;
		while (lo < hi) {
			u64 mid = (lo + hi) >> 1;
			if (cdf[mid] > u) hi = mid;
			else lo = mid + 1;
		}
		const u64 key = mix64(lo);
		out[i] = sizeof(T) == 8 ? static_cast<T>(key) : static_cast<T>(key >> 32);
	}
}

__global__ __launch_bounds__(256) void util_copy(uint4* dst, const uint4* src, u64 n16,
	unsigned char* dtail, const unsigned char* stail, u32 ntail) {
	const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
	u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
	for (; i < n16; i += stride) dst[i] = src[i];
	if (blockIdx.x == 0 && threadIdx.x < ntail) dtail[threadIdx.x] = stail[threadIdx.x];
}

// Streaming copy for the bandwidth ceiling: each block moves one 16 KiB
// tile, four 16-byte loads in flight per thread before its stores, tiles
// XCD-contiguous.
constexpr u32 kStreamPer = 4;
constexpr u32 kStreamTile = 256 * kStreamPer;  // uint4 per tile

__global__ __launch_bounds__(256) void util_stream_copy(uint4* __restrict__ dst,
	const uint4* __restrict__ src, u32 tiles) {
	const u32 t = xcd_tile(tiles);
	if (t >= tiles) return;
	const u64 base = static_cast<u64>(t) * kStreamTile + threadIdx.x;
	uint4 v[kStreamPer];
#pragma unroll
	for (u32 j = 0; j < kStreamPer; ++j) v[j] = src[base + j * 256];
#pragma unroll
	for (u32 j = 0; j < kStreamPer; ++j) dst[base + j * 256] = v[j];
}

// The write pattern of one uniform 8-bit radix pass over (u64 key, u32
// value) pairs with nothing else: tile t of 8192 pairs is read contiguously
// and written as 8192 / run runs of `run` pairs, run d of tile t to
// d * (n / runs_per_tile) + t * run (every bucket receives `run` pairs per
// tile, as uniform keys give), consecutive lanes writing consecutive pairs.
constexpr u32 kProbeTile = 8192;

__global__ __launch_bounds__(256) void util_scatter_runs(u64* __restrict__ kdst,
	u32* __restrict__ vdst, const u64* __restrict__ ksrc, const u32* __restrict__ vsrc, u32 tiles,
	u32 run) {
	const u32 t = xcd_tile(tiles);
	if (t >= tiles) return;
	const u64 n = static_cast<u64>(tiles) * kProbeTile;
	const u32 runs = kProbeTile / run;
	const u64 bucket = n / runs;
	constexpr u32 kPer = kProbeTile / 256;
	u64 k[kPer];
	u32 v[kPer];
	const u64 base = static_cast<u64>(t) * kProbeTile;
#pragma unroll
	for (u32 j = 0; j < kPer; ++j) {
		k[j] = ksrc[base + j * 256 + threadIdx.x];
		v[j] = vsrc[base + j * 256 + threadIdx.x];
	}
#pragma unroll
	for (u32 j = 0; j < kPer; ++j) {
		const u32 i = j * 256 + threadIdx.x;
		const u64 o = (i / run) * bucket + static_cast<u64>(t) * run + (i % run);
		kdst[o] = k[j];
		vdst[o] = v[j];
	}
}

template <typename T>
__global__ __launch_bounds__(256) void util_hash(const T* x, u64 n, u64* acc) {
	const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
	u64 s = 0;
	for (u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
		s += mix64(static_cast<u64>(x[i]));
	s = wave_sum(s);
	if (lane_id() == 0) atomicAdd(reinterpret_cast<unsigned long long*>(acc),
		static_cast<unsigned long long>(s));
}

template <typename T>
__global__ __launch_bounds__(256) void util_check_sorted(const T* k, const u32* v, u64 n,
	u64* bad) {
	const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
	u64 c = 0;
	for (u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i + 1 < n; i += stride) {
		T a = k[i], b = k[i + 1];
		if (a > b) ++c;
		else if (v && a == b && !(v[i] < v[i + 1])) ++c;
	}
	c = wave_sum(c);
	if (lane_id() == 0 && c) atomicAdd(reinterpret_cast<unsigned long long*>(bad),
		static_cast<unsigned long long>(c));
}

// Pairing and permutation: vout[i] must be an index < n that no other i
// names (a bitmap of n bits records the indices seen), and kout[i] must be
// the input key at that index.
template <typename T>
__global__ __launch_bounds__(256) void util_check_pairs(const T* kin, const T* kout,
	const u32* vout, u64 n, u32* seen, u64* bad) {
	const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
	u64 c = 0;
	for (u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		const u32 v = vout[i];
		if (v >= n) {
			++c;
			continue;
		}
		const u32 bit = 1u << (v & 31u);
		if (atomicOr(&seen[v >> 5], bit) & bit) ++c;
		if (kout[i] != kin[v]) ++c;
	}
	c = wave_sum(c);
	if (lane_id() == 0 && c) atomicAdd(reinterpret_cast<unsigned long long*>(bad),
		static_cast<unsigned long long>(c));
}

__global__ __launch_bounds__(256) void util_count_diff(const u32* a, const u32* b, u64 n,
	u64* bad) {
	const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
	u64 c = 0;
	for (u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
		c += a[i] != b[i] ? 1u : 0u;
	c = wave_sum(c);
	if (lane_id() == 0 && c) atomicAdd(reinterpret_cast<unsigned long long*>(bad),
		static_cast<unsigned long long>(c));
}

u32 grid_for(u64 n) {
	u64 g = (n + 255) / 256;
	if (g > 8192) g = 8192;
	if (g == 0) g = 1;
	return static_cast<u32>(g);
}

}  // namespace

extern "C" {

CCLEvent* clo_util_generate(CCLQueue* cq, CCLBuffer* out, size_t n, CloGenKind kind,
	cl_ulong seed, cl_ulong param, cl_ulong start, GError** err) {
	if (!cq || !out) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL queue or buffer");
		return nullptr;
	}
	size_t w = (kind == CLO_GEN_U64 || kind == CLO_GEN_U64_MASK || kind == CLO_GEN_ZIPF_U64 ||
		kind == CLO_GEN_U64_MOD || kind == CLO_GEN_NEARLY_SORTED_U64 ||
		kind == CLO_GEN_REVERSED_U64) ? 8 : 4;
	if (kind == CLO_GEN_ZIPF_U64 || static_cast<int>(kind) < 0 || kind > CLO_GEN_REVERSED_U64) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "generator kind %d not valid here", kind);
		return nullptr;
	}
	if ((kind == CLO_GEN_U32_MOD || kind == CLO_GEN_U64_MOD) && param == 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "modulus must be non-zero");
		return nullptr;
	}
	if (out->size < n * w) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "buffer too small");
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	Launch L(cq, "util_generate");
	u32 g = grid_for(n);
	switch (kind) {
	case CLO_GEN_U64: util_generate<CLO_GEN_U64><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	case CLO_GEN_U32_HI: util_generate<CLO_GEN_U32_HI><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	case CLO_GEN_U64_MASK: util_generate<CLO_GEN_U64_MASK><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	case CLO_GEN_U32_MASK: util_generate<CLO_GEN_U32_MASK><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	case CLO_GEN_IOTA_U32: util_generate<CLO_GEN_IOTA_U32><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	case CLO_GEN_U32_MOD: util_generate<CLO_GEN_U32_MOD><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	case CLO_GEN_U64_MOD: util_generate<CLO_GEN_U64_MOD><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	case CLO_GEN_NEARLY_SORTED_U32: util_generate<CLO_GEN_NEARLY_SORTED_U32><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	case CLO_GEN_NEARLY_SORTED_U64: util_generate<CLO_GEN_NEARLY_SORTED_U64><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	case CLO_GEN_REVERSED_U32: util_generate<CLO_GEN_REVERSED_U32><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	case CLO_GEN_REVERSED_U64: util_generate<CLO_GEN_REVERSED_U64><<<g, 256, 0, cq->stream>>>(out->ptr, n, seed, param, start); break;
	default: break;
	}
	if (!L.done(err)) return nullptr;
	return record_event(cq, "util_generate", err);
}

void clo_util_zipf_cdf(double* cdf, size_t ranks, double s) {
	// Normalised partial sums of (r+1)^-s, accumulated in long double.
	long double tot = 0.0L;
	for (size_t r = 0; r < ranks; ++r) tot += powl(static_cast<long double>(r + 1), -s);
	long double acc = 0.0L;
	for (size_t r = 0; r < ranks; ++r) {
		acc += powl(static_cast<long double>(r + 1), -s);
		cdf[r] = static_cast<double>(acc / tot);
	}
	if (ranks) cdf[ranks - 1] = 1.0;
}

}  // extern "C"

namespace {

template <typename T>
CCLEvent* generate_zipf(CCLQueue* cq, CCLBuffer* out, size_t n, CCLBuffer* cdf, size_t ranks,
	cl_ulong seed, GError** err) {
	if (err && *err) return nullptr;
	if (!cq || !out || !cdf || ranks == 0 || out->size < n * sizeof(T) ||
		cdf->size < ranks * 8) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "invalid zipf generator arguments");
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	Launch L(cq, "util_generate_zipf");
	util_generate_zipf<T><<<grid_for(n), 256, 0, cq->stream>>>(static_cast<T*>(out->ptr), n,
		static_cast<const double*>(cdf->ptr), ranks, seed);
	if (!L.done(err)) return nullptr;
	return record_event(cq, "util_generate_zipf", err);
}

}  // namespace

extern "C" {

CCLEvent* clo_util_generate_zipf(CCLQueue* cq, CCLBuffer* out, size_t n, CCLBuffer* cdf,
	size_t ranks, cl_ulong seed, GError** err) {
	return generate_zipf<u64>(cq, out, n, cdf, ranks, seed, err);
}

CCLEvent* clo_util_generate_zipf_u32(CCLQueue* cq, CCLBuffer* out, size_t n, CCLBuffer* cdf,
	size_t ranks, cl_ulong seed, GError** err) {
	return generate_zipf<u32>(cq, out, n, cdf, ranks, seed, err);
}

CCLEvent* clo_util_copy(CCLQueue* cq, CCLBuffer* dst, CCLBuffer* src, size_t bytes,
	GError** err) {
	if (!cq || !dst || !src || dst->size < bytes || src->size < bytes ||
		(reinterpret_cast<uintptr_t>(dst->ptr) | reinterpret_cast<uintptr_t>(src->ptr)) % 16) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "invalid copy arguments");
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	const u64 n16 = bytes / 16;
	const u32 ntail = static_cast<u32>(bytes % 16);
	Launch L(cq, "util_copy");
	u64 g = (n16 + 255) / 256;
	if (g > 4096) g = 4096;
	if (g == 0) g = 1;
	util_copy<<<static_cast<u32>(g), 256, 0, cq->stream>>>(static_cast<uint4*>(dst->ptr),
		static_cast<const uint4*>(src->ptr), n16,
		static_cast<unsigned char*>(dst->ptr) + n16 * 16,
		static_cast<const unsigned char*>(src->ptr) + n16 * 16, ntail);
	if (!L.done(err)) return nullptr;
	return record_event(cq, "util_copy", err);
}

CCLEvent* clo_util_stream_copy(CCLQueue* cq, CCLBuffer* dst, CCLBuffer* src, size_t bytes,
	GError** err) {
	const size_t tile_bytes = static_cast<size_t>(kStreamTile) * 16;
	if (!cq || !dst || !src || dst->size < bytes || src->size < bytes || bytes % tile_bytes ||
		bytes / tile_bytes > 0xFFFFFFF0u ||
		(reinterpret_cast<uintptr_t>(dst->ptr) | reinterpret_cast<uintptr_t>(src->ptr)) % 16) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "stream copy takes 16 KiB multiples, aligned");
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	const u32 tiles = static_cast<u32>(bytes / tile_bytes);
	if (tiles) {
		Launch L(cq, "util_stream_copy");
		util_stream_copy<<<xcd_grid(tiles), 256, 0, cq->stream>>>(static_cast<uint4*>(dst->ptr),
			static_cast<const uint4*>(src->ptr), tiles);
		if (!L.done(err)) return nullptr;
	}
	return record_event(cq, "util_stream_copy", err);
}

CCLEvent* clo_util_scatter_runs(CCLQueue* cq, CCLBuffer* kdst, CCLBuffer* vdst, CCLBuffer* ksrc,
	CCLBuffer* vsrc, size_t n, cl_uint run, GError** err) {
	if (!cq || !kdst || !vdst || !ksrc || !vsrc || n % kProbeTile || n / kProbeTile > 0xFFFFFFF0u ||
		(run != 4 && run != 8 && run != 16 && run != 32 && run != 64 && run != 128) || kdst->size < n * 8 || ksrc->size < n * 8 ||
		vdst->size < n * 4 || vsrc->size < n * 4) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"scatter probe takes n a multiple of 8192 pairs and runs of 4 to 128 pairs (powers of two)");
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	const u32 tiles = static_cast<u32>(n / kProbeTile);
	if (tiles) {
		Launch L(cq, "util_scatter_runs");
		util_scatter_runs<<<xcd_grid(tiles), 256, 0, cq->stream>>>(static_cast<u64*>(kdst->ptr),
			static_cast<u32*>(vdst->ptr), static_cast<const u64*>(ksrc->ptr),
			static_cast<const u32*>(vsrc->ptr), tiles, run);
		if (!L.done(err)) return nullptr;
	}
	return record_event(cq, "util_scatter_runs", err);
}

static bool scratch_u64(CCLQueue* cq, u64** p, GError** err) {
	if (hip_failed(hipMallocAsync(reinterpret_cast<void**>(p), 8, cq->stream), err,
			"hipMallocAsync"))
		return false;
	return !hip_failed(hipMemsetAsync(*p, 0, 8, cq->stream), err, "hipMemsetAsync");
}

static bool read_u64(CCLQueue* cq, u64* p, u64* out, GError** err) {
	bool ok = !hip_failed(hipMemcpyAsync(out, p, 8, hipMemcpyDeviceToHost, cq->stream), err,
		"hipMemcpyAsync");
	ok = ok && !hip_failed(hipStreamSynchronize(cq->stream), err, "hipStreamSynchronize");
	(void)hipFreeAsync(p, cq->stream);
	return ok;
}

cl_ulong clo_util_multiset_hash(CCLQueue* cq, CCLBuffer* buf, size_t n, size_t width,
	GError** err) {
	if (!cq || !buf || buf->size < n * width) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "invalid hash arguments");
		return 0;
	}
	if (!use_device(cq->ctx, err)) return 0;
	u64* acc = nullptr;
	if (!scratch_u64(cq, &acc, err)) return 0;
	u32 g = grid_for(n);
	if (g > 2048) g = 2048;
	switch (width) {
	case 1: util_hash<u8><<<g, 256, 0, cq->stream>>>(static_cast<const u8*>(buf->ptr), n, acc); break;
	case 2: util_hash<u16><<<g, 256, 0, cq->stream>>>(static_cast<const u16*>(buf->ptr), n, acc); break;
	case 4: util_hash<u32><<<g, 256, 0, cq->stream>>>(static_cast<const u32*>(buf->ptr), n, acc); break;
	case 8: util_hash<u64><<<g, 256, 0, cq->stream>>>(static_cast<const u64*>(buf->ptr), n, acc); break;
	default:
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "width must be 1, 2, 4 or 8");
		(void)hipFreeAsync(acc, cq->stream);
		return 0;
	}
	u64 r = 0;
	if (!read_u64(cq, acc, &r, err)) return 0;
	return r;
}

size_t clo_util_check_sorted(CCLQueue* cq, CCLBuffer* keys, CCLBuffer* vals, size_t n,
	size_t width, GError** err) {
	if (!cq || !keys || keys->size < n * width || (vals && vals->size < n * 4)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "invalid check arguments");
		return static_cast<size_t>(-1);
	}
	if (!use_device(cq->ctx, err)) return static_cast<size_t>(-1);
	u64* bad = nullptr;
	if (!scratch_u64(cq, &bad, err)) return static_cast<size_t>(-1);
	u32 g = grid_for(n);
	if (g > 2048) g = 2048;
	const u32* v = vals ? static_cast<const u32*>(vals->ptr) : nullptr;
	switch (width) {
	case 1: util_check_sorted<u8><<<g, 256, 0, cq->stream>>>(static_cast<const u8*>(keys->ptr), v, n, bad); break;
	case 2: util_check_sorted<u16><<<g, 256, 0, cq->stream>>>(static_cast<const u16*>(keys->ptr), v, n, bad); break;
	case 4: util_check_sorted<u32><<<g, 256, 0, cq->stream>>>(static_cast<const u32*>(keys->ptr), v, n, bad); break;
	case 8: util_check_sorted<u64><<<g, 256, 0, cq->stream>>>(static_cast<const u64*>(keys->ptr), v, n, bad); break;
	default:
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "width must be 1, 2, 4 or 8");
		(void)hipFreeAsync(bad, cq->stream);
		return static_cast<size_t>(-1);
	}
	u64 r = 0;
	if (!read_u64(cq, bad, &r, err)) return static_cast<size_t>(-1);
	return static_cast<size_t>(r);
}

size_t clo_util_check_pairs(CCLQueue* cq, CCLBuffer* keys_in, CCLBuffer* keys_out,
	CCLBuffer* vals_out, size_t n, size_t width, GError** err) {
	if (!cq || !keys_in || !keys_out || !vals_out || keys_in->size < n * width ||
		keys_out->size < n * width || vals_out->size < n * 4 || n > 0xFFFFFFFFull) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "invalid pair-check arguments");
		return static_cast<size_t>(-1);
	}
	if (!use_device(cq->ctx, err)) return static_cast<size_t>(-1);
	u64* bad = nullptr;
	if (!scratch_u64(cq, &bad, err)) return static_cast<size_t>(-1);
	u32* seen = nullptr;
	const size_t words = (n + 31) / 32;
	if (hip_failed(hipMallocAsync(reinterpret_cast<void**>(&seen), words * 4 + 4, cq->stream),
			err, "hipMallocAsync") ||
		hip_failed(hipMemsetAsync(seen, 0, words * 4 + 4, cq->stream), err, "hipMemsetAsync")) {
		(void)hipFreeAsync(bad, cq->stream);
		return static_cast<size_t>(-1);
	}
	u32 g = grid_for(n);
	if (g > 2048) g = 2048;
	const u32* v = static_cast<const u32*>(vals_out->ptr);
	switch (width) {
	case 1: util_check_pairs<u8><<<g, 256, 0, cq->stream>>>(static_cast<const u8*>(keys_in->ptr), static_cast<const u8*>(keys_out->ptr), v, n, seen, bad); break;
	case 2: util_check_pairs<u16><<<g, 256, 0, cq->stream>>>(static_cast<const u16*>(keys_in->ptr), static_cast<const u16*>(keys_out->ptr), v, n, seen, bad); break;
	case 4: util_check_pairs<u32><<<g, 256, 0, cq->stream>>>(static_cast<const u32*>(keys_in->ptr), static_cast<const u32*>(keys_out->ptr), v, n, seen, bad); break;
	case 8: util_check_pairs<u64><<<g, 256, 0, cq->stream>>>(static_cast<const u64*>(keys_in->ptr), static_cast<const u64*>(keys_out->ptr), v, n, seen, bad); break;
	default:
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "width must be 1, 2, 4 or 8");
		(void)hipFreeAsync(seen, cq->stream);
		(void)hipFreeAsync(bad, cq->stream);
		return static_cast<size_t>(-1);
	}
	(void)hipFreeAsync(seen, cq->stream);
	u64 r = 0;
	if (!read_u64(cq, bad, &r, err)) return static_cast<size_t>(-1);
	return static_cast<size_t>(r);
}

size_t clo_util_count_diff(CCLQueue* cq, CCLBuffer* a, CCLBuffer* b, size_t words,
	GError** err) {
	if (!cq || !a || !b || a->size < words * 4 || b->size < words * 4) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "invalid compare arguments");
		return static_cast<size_t>(-1);
	}
	if (!use_device(cq->ctx, err)) return static_cast<size_t>(-1);
	u64* bad = nullptr;
	if (!scratch_u64(cq, &bad, err)) return static_cast<size_t>(-1);
	u32 g = grid_for(words);
	if (g > 2048) g = 2048;
	util_count_diff<<<g, 256, 0, cq->stream>>>(static_cast<const u32*>(a->ptr),
		static_cast<const u32*>(b->ptr), words, bad);
	u64 r = 0;
	if (!read_u64(cq, bad, &r, err)) return static_cast<size_t>(-1);
	return static_cast<size_t>(r);
}

const char* clo_util_build_info(void) {
	return "cl-ops MI355X: gfx950, clang " __clang_version__ ", wave64, satradix reduce-then-scan";
}

}  // extern "C"

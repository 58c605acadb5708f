// Device helpers shared by the gfx950 kernels: wave64 lane arithmetic,
// block-wide scans, and the look-back status words.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace clo {
namespace dev {

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;

constexpr int kWave = 64;

__device__ __forceinline__ u32 lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ u32 wave_id() { return threadIdx.x / kWave; }

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ u32 count_below(u64 mask) {
	return __builtin_amdgcn_mbcnt_hi(static_cast<u32>(mask >> 32),
		__builtin_amdgcn_mbcnt_lo(static_cast<u32>(mask), 0u));
}

__device__ __forceinline__ u64 ballot(bool p) { return __ballot(p); }

// Length bin of a segment for the segmented sort's work-class table: four
// bins per octave, from the leading one and the two bits after it (len 0 and
// 1 share bin 0), capped at 63.  The same function on host and device.
__host__ __device__ inline u32 seg_len_bin(u32 len) {
	if (len < 2) return 0;
	const u32 lg = 31u - static_cast<u32>(__builtin_clz(len));
	const u32 sub = lg >= 2 ? (len >> (lg - 2)) & 3u : (len << (2 - lg)) & 3u;
	const u32 b = 4 * lg + sub;
	return b < 63 ? b : 63;
}

// ---- XCD-aware tile mapping -------------------------------------------------
// MI355X dispatches consecutive blocks round-robin over its 8 XCDs, each with
// its own L2.  Kernels whose tiles are independent map block b to tile
// (b % 8) * per + b / 8 so that each XCD works through a contiguous range of
// tiles: neighbouring tiles then share an L2 (their partial cache lines of
// the same output runs merge there instead of being written back from 8
// caches).  Any mapping is correct; this one is the fast one.  Launch
// xcd_grid(tiles) blocks; blocks whose tile is >= tiles return.
constexpr u32 kXcds = 8;

__device__ __forceinline__ u32 xcd_tile(u32 tiles) {
	const u32 per = (tiles + kXcds - 1) / kXcds;
	return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

inline u32 xcd_grid(u32 tiles) { return kXcds * ((tiles + kXcds - 1) / kXcds); }

// xcd_tile with each XCD's range walked from its end: a kernel that follows
// one over the same tiles starts on the data the first read last, which the
// memory-side cache may still hold.
__device__ __forceinline__ u32 xcd_tile_rev(u32 tiles) {
	const u32 per = (tiles + kXcds - 1) / kXcds;
	return (blockIdx.x % kXcds) * per + (per - 1 - blockIdx.x / kXcds);
}

// Inclusive wave-wide prefix sum.
template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T x) {
	const u32 lane = lane_id();
#pragma unroll
	for (int off = 1; off < kWave; off <<= 1) {
		T y = __shfl_up(x, off, kWave);
		if (lane >= static_cast<u32>(off)) x += y;
	}
	return x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
	for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
	return x;
}

// Exclusive block-wide prefix sum for a block of NT threads (NT multiple of
// 64).  `s_wsum` holds NT/64 elements of shared scratch.  Ends with a
// barrier, so s_wsum may be reused afterwards.
template <int NT, typename T>
__device__ __forceinline__ T block_exclusive_sum(T v, T* s_wsum, T* total) {
	constexpr int W = NT / kWave;
	T inc = wave_inclusive_sum(v);
	if (lane_id() == kWave - 1) s_wsum[wave_id()] = inc;
	__syncthreads();
	T pre = 0, tot = 0;
	const u32 w = wave_id();
#pragma unroll
	for (int i = 0; i < W; ++i) {
		T s = s_wsum[i];
		if (static_cast<u32>(i) < w) pre += s;
		tot += s;
	}
	if (total) *total = tot;
	__syncthreads();
	return pre + inc - v;
}

// Lanes of the wave holding the same `bits`-bit digit as this lane, among
// the lanes where `valid` is set: AND over the digit bits of
// (ballot of bit b) matched against this lane's bit b.  The mismatches are
// gathered as OR of (ballot ^ sext(bit)) on 32-bit halves, one
// v_bitop3_b32 (0xF6: a | (b ^ c)) per half and bit on gfx950.
// MAXBITS bounds the unrolled loop; bits <= MAXBITS.
template <u32 MAXBITS>
__device__ __forceinline__ u64 match_peers(u32 dig, u32 bits, bool valid) {
	const u64 pv = ballot(valid);
	u32 mlo = 0, mhi = 0;
#pragma unroll
	for (u32 bt = 0; bt < MAXBITS; ++bt) {
		if (bt >= bits) break;
		const u32 sb = static_cast<u32>(__builtin_amdgcn_sbfe(static_cast<int>(dig),
			static_cast<int>(bt), 1));
		const u64 bb = ballot(static_cast<int>(sb) < 0);
		mlo = __builtin_amdgcn_bitop3_b32(mlo, static_cast<u32>(bb), sb, 0xF6);
		mhi = __builtin_amdgcn_bitop3_b32(mhi, static_cast<u32>(bb >> 32), sb, 0xF6);
	}
	return pv & ~((static_cast<u64>(mhi) << 32) | mlo);
}

// Lanes holding the same digit when every lane of the wave holds an element.
template <u32 MAXBITS>
__device__ __forceinline__ u64 match_peers_full(u32 dig, u32 bits) {
	u32 mlo = 0, mhi = 0;
#pragma unroll
	for (u32 bt = 0; bt < MAXBITS; ++bt) {
		if (bt >= bits) break;
		const u32 sb = static_cast<u32>(__builtin_amdgcn_sbfe(static_cast<int>(dig),
			static_cast<int>(bt), 1));
		const u64 bb = ballot(static_cast<int>(sb) < 0);
		mlo = __builtin_amdgcn_bitop3_b32(mlo, static_cast<u32>(bb), sb, 0xF6);
		mhi = __builtin_amdgcn_bitop3_b32(mhi, static_cast<u32>(bb >> 32), sb, 0xF6);
	}
	return ~((static_cast<u64>(mhi) << 32) | mlo);
}

// ---- look-back status words -----------------------------------------------
// 64-bit word: high half = (epoch << 2) | flag, low half = 32-bit value.
// The epoch (30 bits) is unique per kernel launch, so words left over from
// earlier launches read as "not ready" and the array never needs clearing.
constexpr u32 kFlagAggregate = 1;
constexpr u32 kFlagPrefix = 2;

__device__ __forceinline__ void status_store(u64* p, u32 epoch, u32 flag, u32 value) {
	u64 w = (static_cast<u64>((epoch << 2) | flag) << 32) | value;
	__hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ u64 status_load(const u64* p) {
	return __hip_atomic_load(const_cast<u64*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ u32 status_flag(u64 w, u32 epoch) {
	u32 hi = static_cast<u32>(w >> 32);
	return (hi >> 2) == epoch ? (hi & 3u) : 0u;
}

}  // namespace dev
}  // namespace clo

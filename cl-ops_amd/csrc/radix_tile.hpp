// One stable digit pass over one tile of consecutive keys: wave64-ballot
// ranking, LDS staging in digit order and a scatter of contiguous runs to
// per-digit output offsets (the satradix downsweep).
#pragma once

#include "dev_common.hpp"

namespace clo {
namespace dev {

constexpr int kTileRadix = 256;

// LDS of one tile pass.  8-byte keys pass through LDS as two 32-bit halves
// (high halves, then low halves) in a buffer of 4 bytes per key, so a tile
// of 8192 keys needs 36 KB instead of 68 KB.
template <typename K, bool KV, int NT, int ITEMS>
struct TileSmem {
	static constexpr u32 kTile = NT * ITEMS;
	static constexpr bool kSplit = sizeof(K) == 8;
	static constexpr u32 kStageBytes =
		kTile * (kSplit ? 4u : (sizeof(K) > 4 ? sizeof(K) : (KV ? 4u : sizeof(K))));
	unsigned char stage[kStageBytes] __attribute__((aligned(16)));
	u32 whist[NT / kWave][kTileRadix];
	u32 tstart[kTileRadix];
	u32 gofs[kTileRadix];
	u32 wsum[NT / kWave];
};

// Where a tile pass reads and writes.  Keys [base, base + count) of src
// (count <= NT * ITEMS) go to dst, the element of digit d at rank i within
// the tile's keys of that digit to goff(d) + i, where thread d < 256 holds
// goff(d) in `goff`.  vsrc/vdst: u32 payloads (KV).  ndig != nullptr also
// writes the digit at bit `nshift` of every moved key at its output position.
template <typename K>
struct TilePass {
	const K* src;
	const u32* vsrc;
	K* dst;
	u32* vdst;
	u8* ndig;
	u32 base;
	u32 count;
	u32 shift;
	u32 nshift;
};

// Stable ranking of the tile, element order wave-striped: wave w owns
// elements [w*64*ITEMS, (w+1)*64*ITEMS), item j of lane l is element
// w*64*ITEMS+j*64+l.  Processing items j = 0.. in order with per-wave running
// digit counts gives each element the number of equal-digit elements before
// it (stability).  FIXED_BITS = 8 compiles the digit-match loop for 8-bit
// digits; 0 takes the width from `bits`.  FULL: count == NT * ITEMS, no
// per-element bounds checks.  Contains barriers: call from uniform control
// flow with sm.whist zeroed by the caller (before a barrier).
template <typename K, bool KV, int NT, int ITEMS, int FIXED_BITS, bool FULL>
__device__ __forceinline__ void tile_pass(TileSmem<K, KV, NT, ITEMS>& sm, const TilePass<K>& tp,
	u32 bits, u32 goff) {
	using S = TileSmem<K, KV, NT, ITEMS>;
	constexpr u32 kTile = S::kTile;
	constexpr int NW = NT / kWave;
	const u32 tid = threadIdx.x;
	const u32 lane = lane_id();
	const u32 w = wave_id();
	const u32 mask = (1u << bits) - 1u;
	const u32 shift = tp.shift;
	const u32 count = FULL ? kTile : tp.count;
	const K* __restrict__ src = tp.src + tp.base;
	K* __restrict__ dst = tp.dst;

	// Load (coalesced: 64 consecutive elements per wave instruction).  A wave
	// issuing its loads (and later its scatter stores) runs at raised priority
	// so its memory requests go out ahead of other waves' ranking work
	// (measured: 2% on the 2^28 u64 KV downsweep).
	const u32 wbase = w * kWave * ITEMS;
	K k[ITEMS];
	u32 v[ITEMS];
	__builtin_amdgcn_s_setprio(3);
#pragma unroll
	for (int j = 0; j < ITEMS; ++j) {
		const u32 idx = wbase + j * kWave + lane;
		k[j] = (FULL || idx < count) ? src[idx] : K(0);
	}
	if (KV) {
		const u32* __restrict__ vsrc = tp.vsrc + tp.base;
#pragma unroll
		for (int j = 0; j < ITEMS; ++j) {
			const u32 idx = wbase + j * kWave + lane;
			v[j] = (FULL || idx < count) ? vsrc[idx] : 0u;
		}
	}
	__builtin_amdgcn_s_setprio(0);
	__syncthreads();

	// Rank inside the wave.  The lanes holding one digit (found with one ballot
	// per digit bit) form a peer group; its lowest lane adds the group's size
	// to the wave's running digit count with one LDS atomic that returns the
	// previous count.  The atomics of all items issue back to back (a wave's
	// LDS operations complete in order, so item j sees the counts of items
	// before it); each lane then takes its group's base from the leader.
	// Items go in groups of kGroup (8 atomics in flight, 8 extra registers).
	constexpr int kGroup = ITEMS < 8 ? ITEMS : 8;
	static_assert(ITEMS % kGroup == 0, "whole groups");
	constexpr u32 kMatch = FIXED_BITS ? FIXED_BITS : 8;
	u32 rank[ITEMS];  // (leader lane << 16) | peers below, then the rank
#pragma unroll
	for (int j0 = 0; j0 < ITEMS; j0 += kGroup) {
		u32 got[kGroup];
#pragma unroll
		for (int jj = 0; jj < kGroup; ++jj) {
			const int j = j0 + jj;
			const bool valid = FULL || (wbase + j * kWave + lane) < count;
			const u32 dig = static_cast<u32>(k[j] >> shift) & mask;
			const u64 peers = FULL ? match_peers_full<kMatch>(dig, bits)
				: match_peers<kMatch>(dig, bits, valid);
			const u32 below = count_below(peers);
			// An invalid lane has no peers; its leader (lane 63) is never used.
			const u32 leader = static_cast<u32>(__builtin_ctzll(peers | (1ull << 63)));
			got[jj] = 0;
			if (valid && below == 0)
				got[jj] = atomicAdd(&sm.whist[w][dig], static_cast<u32>(__popcll(peers)));
			rank[j] = (leader << 16) | below;
		}
#pragma unroll
		for (int jj = 0; jj < kGroup; ++jj) {
			const int j = j0 + jj;
			rank[j] = static_cast<u32>(__shfl(static_cast<int>(got[jj]),
				static_cast<int>(rank[j] >> 16), kWave)) + (rank[j] & 0xFFFFu);
		}
	}
	__syncthreads();

	// Per-digit tile totals and per-wave exclusive offsets.
	u32 total = 0;
	if (tid < kTileRadix) {
#pragma unroll
		for (int ww = 0; ww < NW; ++ww) {
			const u32 c = sm.whist[ww][tid];
			sm.whist[ww][tid] = total;
			total += c;
		}
	}
	const u32 tstart = block_exclusive_sum<NT>(total, sm.wsum, static_cast<u32*>(nullptr));
	if (tid < kTileRadix) {
		sm.tstart[tid] = tstart;
		sm.gofs[tid] = goff - tstart;
	}
	__syncthreads();

	// Stage the keys in digit order in LDS; thread tid then writes elements
	// tid, tid + NT, ... of the staged tile to their run in the output.
	u32 oidx[ITEMS];
	const u32 nshift = tp.nshift;
	u8* __restrict__ ndig = tp.ndig;
	if constexpr (S::kSplit) {
		u32* s32 = reinterpret_cast<u32*>(sm.stage);
#pragma unroll
		for (int j = 0; j < ITEMS; ++j) {
			const bool valid = FULL || (wbase + j * kWave + lane) < count;
			const u32 dig = static_cast<u32>(k[j] >> shift) & mask;
			if (valid) {
				rank[j] += sm.tstart[dig] + sm.whist[w][dig];
				s32[rank[j]] = static_cast<u32>(static_cast<u64>(k[j]) >> 32);
			}
		}
		__syncthreads();
		u32 hi[ITEMS];
#pragma unroll
		for (int r = 0; r < ITEMS; ++r) {
			const u32 i = r * NT + tid;
			hi[r] = (FULL || i < count) ? s32[i] : 0u;
		}
		__syncthreads();
#pragma unroll
		for (int j = 0; j < ITEMS; ++j) {
			const bool valid = FULL || (wbase + j * kWave + lane) < count;
			if (valid) s32[rank[j]] = static_cast<u32>(k[j]);
		}
		__syncthreads();
		__builtin_amdgcn_s_setprio(2);
#pragma unroll
		for (int r = 0; r < ITEMS; ++r) {
			const u32 i = r * NT + tid;
			if (FULL || i < count) {
				const K key = static_cast<K>((static_cast<u64>(hi[r]) << 32) | s32[i]);
				const u32 dig = static_cast<u32>(key >> shift) & mask;
				const u32 o = sm.gofs[dig] + i;
				dst[o] = key;
				if (ndig) ndig[o] = static_cast<u8>(static_cast<u32>(key >> nshift) & mask);
				oidx[r] = o;
			}
		}
	} else {
		K* s_keys = reinterpret_cast<K*>(sm.stage);
#pragma unroll
		for (int j = 0; j < ITEMS; ++j) {
			const bool valid = FULL || (wbase + j * kWave + lane) < count;
			const u32 dig = static_cast<u32>(k[j] >> shift) & mask;
			if (valid) {
				rank[j] += sm.tstart[dig] + sm.whist[w][dig];
				s_keys[rank[j]] = k[j];
			}
		}
		__syncthreads();
		__builtin_amdgcn_s_setprio(2);
		if (ndig) {
#pragma unroll
			for (int r = 0; r < ITEMS; ++r) {
				const u32 i = r * NT + tid;
				if (FULL || i < count) {
					const K key = s_keys[i];
					const u32 dig = static_cast<u32>(key >> shift) & mask;
					const u32 o = sm.gofs[dig] + i;
					dst[o] = key;
					ndig[o] = static_cast<u8>(static_cast<u32>(key >> nshift) & mask);
					oidx[r] = o;
				}
			}
		} else {
#pragma unroll
			for (int r = 0; r < ITEMS; ++r) {
				const u32 i = r * NT + tid;
				if (FULL || i < count) {
					const K key = s_keys[i];
					const u32 dig = static_cast<u32>(key >> shift) & mask;
					const u32 o = sm.gofs[dig] + i;
					dst[o] = key;
					oidx[r] = o;
				}
			}
		}
	}
	if (KV) {
		u32* s_vals = reinterpret_cast<u32*>(sm.stage);
		u32* __restrict__ vdst = tp.vdst;
		__syncthreads();
#pragma unroll
		for (int j = 0; j < ITEMS; ++j) {
			const bool valid = FULL || (wbase + j * kWave + lane) < count;
			if (valid) s_vals[rank[j]] = v[j];
		}
		__syncthreads();
#pragma unroll
		for (int r = 0; r < ITEMS; ++r) {
			const u32 i = r * NT + tid;
			if (FULL || i < count) vdst[oidx[r]] = s_vals[i];
		}
	}
	__builtin_amdgcn_s_setprio(0);
}

}  // namespace dev
}  // namespace clo

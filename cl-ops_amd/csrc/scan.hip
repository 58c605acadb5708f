// Scanner factory (CL_Ops' src/cl_ops/scan/clo_scan_abstract.c:74-567) and
// the "blelloch" exclusive scan for gfx950.
//
// Result contract (clo_scan_blelloch.cl:49-211): out[i] = sum_{j<i} (SUM)in[j]
// with the element converted to the sum type as C does (sign extension for
// signed elements) and the sum taken modulo 2^(8*sizeof(SUM)).
//
// Three kernels, the reference's three-phase structure (workgroupScan,
// workgroupSumsScan, addWorkgroupSums: clo_scan_blelloch.c:78-214) reordered
// as reduce-then-scan so that the data is written once:
//   blelloch_tile_reduce  per tile of 4096 elements: its sum (segmented: its
//                         (segment-start flag, sum after the last start))
//   blelloch_tiles_scan   one block: exclusive scan of the tile sums
//   blelloch_tile_scan    per tile: scan in LDS, plus the tile's carry-in
// The data is read twice and written once; no block ever waits for another.
// (A single-pass decoupled look-back scan was measured on MI355X at 20-30% of
// HBM peak: its cross-tile look-back chain, not the data, set the time.)
// Segmented scans take segment starts from a sorted offsets array: the sum
// restarts at every start.

#include <atomic>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "clo_internal.hpp"
#include "dev_common.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;  // default items per thread (4096-element tiles)
constexpr int kTilesBlock = 1024;  // threads of the tile-sum scan

__host__ __device__ constexpr u32 pad(u32 i) { return i + (i >> 5); }

// (flag, value) pair of the segmented scan: flag = a segment starts inside,
// value = sum since the last segment start (or of everything without one).
template <typename Acc>
struct SegPair {
	u32 f;
	Acc v;
};

template <typename Acc>
__device__ __forceinline__ SegPair<Acc> seg_combine(SegPair<Acc> a, SegPair<Acc> b) {
	return {a.f | b.f, b.f ? b.v : a.v + b.v};
}

// Exclusive block scan of (flag, value) pairs over NT threads; *total = all
// threads combined.  Ends with a barrier.
template <int NT, typename Acc>
__device__ SegPair<Acc> block_exclusive_segscan(SegPair<Acc> mine, SegPair<Acc>* s_w,
	SegPair<Acc>* total) {
	constexpr int W = NT / kWave;
	const u32 lane = lane_id();
	SegPair<Acc> inc = mine;
#pragma unroll
	for (int off = 1; off < kWave; off <<= 1) {
		SegPair<Acc> o;
		o.f = __shfl_up(inc.f, off, kWave);
		o.v = __shfl_up(inc.v, off, kWave);
		if (lane >= static_cast<u32>(off)) inc = seg_combine(o, inc);
	}
	if (lane == kWave - 1) s_w[wave_id()] = inc;
	__syncthreads();
	SegPair<Acc> pre = {0u, Acc(0)}, tot = {0u, Acc(0)};
#pragma unroll
	for (int i = 0; i < W; ++i) {
		if (static_cast<u32>(i) < wave_id()) pre = seg_combine(pre, s_w[i]);
		tot = seg_combine(tot, s_w[i]);
	}
	*total = tot;
	__syncthreads();
	SegPair<Acc> ex;
	ex.f = __shfl_up(inc.f, 1, kWave);
	ex.v = __shfl_up(inc.v, 1, kWave);
	if (lane == 0) ex = {0u, Acc(0)};
	return seg_combine(pre, ex);
}

// First index j with off[j] >= x (off non-decreasing).
__device__ __forceinline__ u32 first_at_or_after(const u32* off, u32 nseg, u64 x) {
	u32 lo = 0, hi = nseg;
	while (lo < hi) {
		const u32 mid = (lo + hi) >> 1;
		if (off[mid] < x) lo = mid + 1;
		else hi = mid;
	}
	return lo;
}

// Bitmap (in LDS) of the segment starts inside the tile at `base`.  Contains
// barriers.
template <int IT>
__device__ void mark_heads(u32* s_head, u32* s_j0, const u32* __restrict__ off, u32 nseg,
	u64 base) {
	constexpr u32 kScanTile = kScanBlock * IT;
	const u32 tid = threadIdx.x;
	for (u32 i = tid; i < kScanTile / 32; i += kScanBlock) s_head[i] = 0;
	if (tid == 0) *s_j0 = first_at_or_after(off, nseg, base);
	__syncthreads();
	for (u32 j = *s_j0 + tid; j < nseg; j += kScanBlock) {
		const u64 o = off[j];
		if (o >= base + kScanTile) break;
		if (o < base) continue;  // offsets that are not monotone
		const u32 p = static_cast<u32>(o - base);
		atomicOr(&s_head[p >> 5], 1u << (p & 31));
	}
	__syncthreads();
}

// Stages the tile (striped, coalesced loads) in LDS converted to the sum type.
template <typename In, typename Acc, int IT>
__device__ __forceinline__ void stage_tile(Acc* s_t, const In* __restrict__ in, u64 base, u64 n) {
	const u32 tid = threadIdx.x;
#pragma unroll
	for (int j = 0; j < IT; ++j) {
		const u32 i = j * kScanBlock + tid;
		const u64 g = base + i;
		s_t[pad(i)] = g < n ? static_cast<Acc>(in[g]) : Acc(0);
	}
	__syncthreads();
}

// Thread's blocked items: (any segment start, sum after the last start).
template <typename Acc, bool SEG, int IT>
__device__ __forceinline__ SegPair<Acc> thread_items(const Acc* s_t, const u32* s_head,
	u64* heads_out) {
	static_assert(IT <= 64, "one head bit per item");
	const u32 tid = threadIdx.x;
	u64 heads = 0;
	Acc run = 0;
	u32 any = 0;
#pragma unroll
	for (int j = 0; j < IT; ++j) {
		const u32 p = tid * IT + j;
		const Acc x = s_t[pad(p)];
		if (SEG && ((s_head[p >> 5] >> (p & 31)) & 1u)) {
			heads |= 1ull << j;
			run = 0;
			any = 1;
		}
		run += x;
	}
	*heads_out = heads;
	return {any, run};
}

template <typename In, typename Acc, bool SEG, int IT>
__global__ __launch_bounds__(kScanBlock) void blelloch_tile_reduce(const In* __restrict__ in, u64 n,
	u32 tiles, const u32* __restrict__ off, u32 nseg, Acc* __restrict__ tsum,
	u32* __restrict__ tflag) {
	constexpr u32 kScanTile = kScanBlock * IT;
	__shared__ Acc s_t[SEG ? pad(kScanTile) + 1 : 1];
	__shared__ u32 s_head[SEG ? kScanTile / 32 : 1];
	__shared__ Acc s_wsum[kScanBlock / kWave];
	__shared__ SegPair<Acc> s_pw[kScanBlock / kWave];
	__shared__ u32 s_j0;
	const u32 tile = xcd_tile(tiles);
	if (tile >= tiles) return;
	const u64 base = static_cast<u64>(tile) * kScanTile;
	const u32 tid = threadIdx.x;
	if (!SEG) {
		// Order does not matter for a plain sum: striped loads, block reduce.
		Acc s = 0;
#pragma unroll
		for (int j = 0; j < IT; ++j) {
			const u64 g = base + j * kScanBlock + tid;
			if (g < n) s += static_cast<Acc>(in[g]);
		}
		Acc tot;
		block_exclusive_sum<kScanBlock>(s, s_wsum, &tot);
		if (tid == 0) tsum[tile] = tot;
		return;
	}
	mark_heads<IT>(s_head, &s_j0, off, nseg, base);
	stage_tile<In, Acc, IT>(s_t, in, base, n);
	u64 heads;
	const SegPair<Acc> mine = thread_items<Acc, true, IT>(s_t, s_head, &heads);
	SegPair<Acc> tot;
	block_exclusive_segscan<kScanBlock, Acc>(mine, s_pw, &tot);
	if (tid == 0) {
		tsum[tile] = tot.v;
		tflag[tile] = tot.f;
	}
}

// One block: tsum[t] <- carry into tile t (exclusive, segmented when SEG).
// Rounds of kTilesBlock * kTilesItems tile sums are staged in LDS with
// coalesced loads; each thread then combines kTilesItems consecutive ones.
// `chain` (plain scans, may be NULL) carries a running total between calls
// on consecutive chunks of one array: the scan starts from *chain and leaves
// *chain + the chunk's total there.
constexpr u32 kTilesItems = 8;
constexpr u32 kTilesRound = kTilesBlock * kTilesItems;

template <typename Acc, bool SEG>
__global__ __launch_bounds__(kTilesBlock) void blelloch_tiles_scan(Acc* __restrict__ tsum,
	u32* __restrict__ tflag, u32 tiles, Acc* __restrict__ chain) {
	__shared__ Acc s_v[pad(kTilesRound)];
	__shared__ u8 s_f[SEG ? kTilesRound : 1];
	__shared__ Acc s_wsum[kTilesBlock / kWave];
	__shared__ SegPair<Acc> s_pw[kTilesBlock / kWave];
	const u32 tid = threadIdx.x;
	SegPair<Acc> carry = {0u, (!SEG && chain) ? *chain : Acc(0)};
	// Values of the next round are loaded while the current one is scanned.
	Acc nv[kTilesItems];
	u8 nf[kTilesItems];
#pragma unroll
	for (u32 j = 0; j < kTilesItems; ++j) {
		const u32 i = j * kTilesBlock + tid;
		nv[j] = i < tiles ? tsum[i] : Acc(0);
		nf[j] = (SEG && i < tiles) ? static_cast<u8>(tflag[i] != 0) : u8(0);
	}
	for (u32 base = 0; base < tiles; base += kTilesRound) {
#pragma unroll
		for (u32 j = 0; j < kTilesItems; ++j) {
			const u32 i = j * kTilesBlock + tid;
			s_v[pad(i)] = nv[j];
			if (SEG) s_f[i] = nf[j];
		}
		__syncthreads();
		const u32 nb = base + kTilesRound;
#pragma unroll
		for (u32 j = 0; j < kTilesItems; ++j) {
			const u32 i = nb + j * kTilesBlock + tid;
			nv[j] = i < tiles ? tsum[i] : Acc(0);
			nf[j] = (SEG && i < tiles) ? static_cast<u8>(tflag[i] != 0) : u8(0);
		}
		SegPair<Acc> x[kTilesItems];
		SegPair<Acc> mine = {0u, Acc(0)};
#pragma unroll
		for (u32 k = 0; k < kTilesItems; ++k) {
			const u32 i = tid * kTilesItems + k;
			x[k] = {SEG ? static_cast<u32>(s_f[i]) : 0u, s_v[pad(i)]};
			mine = seg_combine(mine, x[k]);
		}
		SegPair<Acc> run, tot;
		if (SEG) {
			run = block_exclusive_segscan<kTilesBlock, Acc>(mine, s_pw, &tot);
		} else {
			Acc total;
			run = {0u, block_exclusive_sum<kTilesBlock>(mine.v, s_wsum, &total)};
			tot = {0u, total};
		}
		run = seg_combine(carry, run);
#pragma unroll
		for (u32 k = 0; k < kTilesItems; ++k) {
			s_v[pad(tid * kTilesItems + k)] = run.v;
			run = seg_combine(run, x[k]);
		}
		__syncthreads();
#pragma unroll
		for (u32 j = 0; j < kTilesItems; ++j) {
			const u32 i = j * kTilesBlock + tid;
			if (base + i < tiles) tsum[base + i] = s_v[pad(i)];
		}
		carry = seg_combine(carry, tot);
		__syncthreads();
	}
	// Every thread read *chain before the first barrier above.
	if (!SEG && chain && tid == 0) *chain = carry.v;
}

template <typename In, typename Acc, int OUTW, bool SEG, int IT>
__global__ __launch_bounds__(kScanBlock) void blelloch_tile_scan(const In* __restrict__ in,
	void* __restrict__ out, u64 n, u32 tiles, const Acc* __restrict__ carry_in,
	const u32* __restrict__ off, u32 nseg) {
	constexpr u32 kScanTile = kScanBlock * IT;
	__shared__ Acc s_t[pad(kScanTile) + 1];
	__shared__ u32 s_head[SEG ? kScanTile / 32 : 1];
	__shared__ Acc s_wsum[kScanBlock / kWave];
	__shared__ SegPair<Acc> s_pw[kScanBlock / kWave];
	__shared__ u32 s_j0;
	const u32 tile = xcd_tile_rev(tiles);
	if (tile >= tiles) return;
	const u64 base = static_cast<u64>(tile) * kScanTile;
	const u32 tid = threadIdx.x;
	const Acc carry = carry_in[tile];
	if (SEG) mark_heads<IT>(s_head, &s_j0, off, nseg, base);
	stage_tile<In, Acc, IT>(s_t, in, base, n);
	u64 heads;
	const SegPair<Acc> mine = thread_items<Acc, SEG, IT>(s_t, s_head, &heads);
	Acc acc;
	if (SEG) {
		SegPair<Acc> tot;
		const SegPair<Acc> ex = block_exclusive_segscan<kScanBlock, Acc>(mine, s_pw, &tot);
		acc = ex.f ? ex.v : carry + ex.v;
	} else {
		Acc tot;
		acc = carry + block_exclusive_sum<kScanBlock>(mine.v, s_wsum, &tot);
	}
#pragma unroll
	for (int j = 0; j < IT; ++j) {
		if (SEG && ((heads >> j) & 1ull)) acc = 0;
		const u32 p = pad(tid * IT + j);
		const Acc x = s_t[p];
		s_t[p] = acc;
		acc += x;
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < IT; ++j) {
		const u32 i = j * kScanBlock + tid;
		const u64 g = base + i;
		if (g < n) {
			const Acc r = s_t[pad(i)];
			if (OUTW == 1) static_cast<u8*>(out)[g] = static_cast<u8>(r);
			else if (OUTW == 2) static_cast<u16*>(out)[g] = static_cast<u16>(r);
			else if (OUTW == 4) static_cast<u32*>(out)[g] = static_cast<u32>(r);
			else static_cast<u64*>(out)[g] = static_cast<u64>(r);
		}
	}
}

// ---- vectorised plain scans (unsigned 32/64-bit elements, 16-byte aligned
// buffers).  The reduce reads 16-byte vectors in striped order; the scan has
// every thread own IT consecutive elements (loaded and stored as 16-byte
// vectors) and needs LDS only for the block-level scan of thread sums.

template <typename T, int IT>
__device__ __forceinline__ void load_vec(const T* __restrict__ p, T (&x)[IT]) {
	static_assert((IT * sizeof(T)) % 16 == 0, "whole 16-byte vectors");
	const uint4* __restrict__ q = reinterpret_cast<const uint4*>(p);
#pragma unroll
	for (int k = 0; k < static_cast<int>(IT * sizeof(T) / 16); ++k) {
		const uint4 t = q[k];
		if constexpr (sizeof(T) == 4) {
			x[4 * k] = t.x;
			x[4 * k + 1] = t.y;
			x[4 * k + 2] = t.z;
			x[4 * k + 3] = t.w;
		} else {
			x[2 * k] = static_cast<T>(t.x) | (static_cast<T>(t.y) << 32);
			x[2 * k + 1] = static_cast<T>(t.z) | (static_cast<T>(t.w) << 32);
		}
	}
}

template <typename T, int IT>
__device__ __forceinline__ void store_vec(T* __restrict__ p, const T (&x)[IT]) {
	static_assert((IT * sizeof(T)) % 16 == 0, "whole 16-byte vectors");
	uint4* __restrict__ q = reinterpret_cast<uint4*>(p);
#pragma unroll
	for (int k = 0; k < static_cast<int>(IT * sizeof(T) / 16); ++k) {
		uint4 t;
		if constexpr (sizeof(T) == 4) {
			t = make_uint4(x[4 * k], x[4 * k + 1], x[4 * k + 2], x[4 * k + 3]);
		} else {
			t = make_uint4(static_cast<u32>(x[2 * k]), static_cast<u32>(x[2 * k] >> 32),
				static_cast<u32>(x[2 * k + 1]), static_cast<u32>(x[2 * k + 1] >> 32));
		}
		q[k] = t;
	}
}

// Thread tid's IT consecutive elements of the tile at `base` (16-byte loads
// when the whole run is inside the array).
template <typename In, int IT>
__device__ __forceinline__ void load_run(const In* __restrict__ in, u64 first, u64 n, In (&x)[IT]) {
	if (first + IT <= n) {
		load_vec<In, IT>(in + first, x);
	} else {
#pragma unroll
		for (int j = 0; j < IT; ++j) x[j] = first + j < n ? in[first + j] : In(0);
	}
}

// Segment-start bits of thread tid's IT consecutive elements (bit j: element
// tid * IT + j of the tile), from mark_heads' bitmap.
template <int IT>
__device__ __forceinline__ u32 run_heads(const u32* s_head) {
	static_assert(IT <= 32 && 32 % IT == 0, "a thread's bits inside one bitmap word");
	const u32 p0 = threadIdx.x * IT;
	const u32 w = s_head[p0 >> 5] >> (p0 & 31);
	return IT == 32 ? w : (w & ((1u << IT) - 1u));
}

// Segmented scans take the blocked layout (thread tid owns elements
// [tid * IT, (tid + 1) * IT) of the tile), so that a thread's start bits are
// consecutive bits of the tile's bitmap and its elements stay in registers:
// no LDS staging of the data, as in the plain vector kernels.
template <typename In, typename Acc, int IT>
__device__ __forceinline__ SegPair<Acc> run_pair(const In (&x)[IT], u32 heads) {
	Acc run = 0;
#pragma unroll
	for (int j = 0; j < IT; ++j) {
		if ((heads >> j) & 1u) run = 0;
		run += static_cast<Acc>(x[j]);
	}
	return {heads != 0u ? 1u : 0u, run};
}

template <typename In, typename Acc, bool SEG, int IT>
__global__ __launch_bounds__(kScanBlock) void blelloch_tile_reduce_vec(const In* __restrict__ in,
	u64 n, u32 tiles, Acc* __restrict__ tsum, const u32* __restrict__ off, u32 nseg,
	u32* __restrict__ tflag) {
	constexpr u32 kScanTile = kScanBlock * IT;
	constexpr int PER = 16 / sizeof(In);
	__shared__ Acc s_wsum[kScanBlock / kWave];
	__shared__ u32 s_head[SEG ? kScanTile / 32 : 1];
	__shared__ SegPair<Acc> s_pw[SEG ? kScanBlock / kWave : 1];
	__shared__ u32 s_j0;
	const u32 tile = xcd_tile(tiles);
	if (tile >= tiles) return;
	const u64 base = static_cast<u64>(tile) * kScanTile;
	const u32 tid = threadIdx.x;
	if constexpr (SEG) {
		mark_heads<IT>(s_head, &s_j0, off, nseg, base);
		In x[IT];
		load_run<In, IT>(in, base + static_cast<u64>(tid) * IT, n, x);
		SegPair<Acc> tot;
		block_exclusive_segscan<kScanBlock, Acc>(run_pair<In, Acc, IT>(x, run_heads<IT>(s_head)),
			s_pw, &tot);
		if (tid == 0) {
			tsum[tile] = tot.v;
			tflag[tile] = tot.f;
		}
		return;
	}
	Acc s = 0;
	if (base + kScanTile <= n) {
		In x[IT];
#pragma unroll
		for (int j = 0; j < IT / PER; ++j) {
			In v[PER];
			load_vec<In, PER>(in + base + static_cast<u64>(j * kScanBlock + tid) * PER, v);
#pragma unroll
			for (int e = 0; e < PER; ++e) x[j * PER + e] = v[e];
		}
#pragma unroll
		for (int j = 0; j < IT; ++j) s += static_cast<Acc>(x[j]);
	} else {
		for (u64 g = base + tid; g < n; g += kScanBlock) s += static_cast<Acc>(in[g]);
	}
	Acc tot;
	block_exclusive_sum<kScanBlock>(s, s_wsum, &tot);
	if (tid == 0) tsum[tile] = tot;
}

template <typename In, typename Acc, bool SEG, int IT>
__global__ __launch_bounds__(kScanBlock) void blelloch_tile_scan_vec(const In* __restrict__ in,
	Acc* __restrict__ out, u64 n, u32 tiles, const Acc* __restrict__ carry_in,
	const u32* __restrict__ off, u32 nseg) {
	constexpr u32 kScanTile = kScanBlock * IT;
	__shared__ Acc s_wsum[kScanBlock / kWave];
	__shared__ u32 s_head[SEG ? kScanTile / 32 : 1];
	__shared__ SegPair<Acc> s_pw[SEG ? kScanBlock / kWave : 1];
	__shared__ u32 s_j0;
	const u32 tile = xcd_tile_rev(tiles);
	if (tile >= tiles) return;
	const u32 tid = threadIdx.x;
	const u64 first = static_cast<u64>(tile) * kScanTile + static_cast<u64>(tid) * IT;
	const bool full = first + IT <= n;
	if (SEG) mark_heads<IT>(s_head, &s_j0, off, nseg, static_cast<u64>(tile) * kScanTile);
	In x[IT];
	load_run<In, IT>(in, first, n, x);
	Acc run;
	u32 heads = 0;
	if constexpr (SEG) {
		heads = run_heads<IT>(s_head);
		SegPair<Acc> tot;
		const SegPair<Acc> ex = block_exclusive_segscan<kScanBlock, Acc>(
			run_pair<In, Acc, IT>(x, heads), s_pw, &tot);
		run = ex.f ? ex.v : carry_in[tile] + ex.v;
	} else {
		Acc sum = 0;
#pragma unroll
		for (int j = 0; j < IT; ++j) sum += static_cast<Acc>(x[j]);
		Acc tot;
		run = carry_in[tile] + block_exclusive_sum<kScanBlock>(sum, s_wsum, &tot);
	}
	Acc y[IT];
#pragma unroll
	for (int j = 0; j < IT; ++j) {
		if (SEG && ((heads >> j) & 1u)) run = 0;
		y[j] = run;
		run += static_cast<Acc>(x[j]);
	}
	if (full) {
		store_vec<Acc, IT>(out + first, y);
	} else {
#pragma unroll
		for (int j = 0; j < IT; ++j)
			if (first + j < n) out[first + j] = y[j];
	}
}

// ---- single pass: decoupled look-back --------------------------------------
// Plain unsigned 32/64-bit scans on 16-byte aligned buffers of at most
// kOnePassMaxTiles tiles.  On MI355X a device-coherent status read takes about
// a microsecond, and a tile's inclusive prefix only appears after its own
// look-back, so the frontier of known prefixes advances about one poll width
// per microsecond: for large scans that is slower than reading the input
// twice (2^26 u32: 0.32 ms single pass vs 0.17 ms reduce-then-scan), while up
// to 2^21 elements the single launch beats three (2^20: 10.6 vs 14.2 us).
// Tile t = block
// t (linear order: each XCD dispatches its blocks in index order, so the
// lowest unfinished tile is always resident and every wait below ends).  A
// tile loads its elements, publishes its aggregate (A) in its status words,
// then wave 0 walks back over the statuses 64 tiles at a time, adding
// aggregates until it meets an inclusive prefix (P); the tile publishes its
// own P and writes its scanned elements.  The input is read once.
//
// Status words (one per 32 bits of the sum) pack value, epoch and flag into
// one 64-bit relaxed device-scope atomic, so no fences are needed.  The
// epoch comes from a device word (the first word of the status buffer; the
// tiles' words start at the third).  Every tile reads the epoch and the chunk
// base before it publishes its aggregate; the last tile advances the epoch
// once every tile has published, so words of earlier launches read as "not
// ready" and the status array never needs clearing (also across graph
// replays).
//
// Fallback: a tile whose look-back has found a predecessor unpublished for
// `polls` rounds sums that predecessor's input itself (a tile's aggregate is
// the sum of its elements, whoever computes it), so every wait is bounded
// and the result is exact even if a predecessor is stalled, for instance on
// a preempted GPU.  Correct launches stay far below kLookbackPolls; a test
// hook lowers it to force the fallback.  A prefix found through a fallback,
// or through a prefix so found, is published "late" (kFlagPrefixLate): an
// earlier tile may not have started.  An ordinary prefix means every earlier
// tile has published, so the last tile advances the epoch at once unless its
// walk was late; then it first waits until every tile has published (each
// one will: no wait is unbounded).
constexpr u32 kLookbackPolls = 1u << 12;
constexpr u32 kFlagPrefixLate = 3;
constexpr u32 kOnePassMaxTiles = 512;
std::atomic<u32> g_lookback_polls{kLookbackPolls};

template <typename Acc>
__device__ __forceinline__ void publish(u64* st, u32 tile, u32 epoch, u32 flag, Acc v) {
	if constexpr (sizeof(Acc) == 8) {
		status_store(st + 2 * static_cast<u64>(tile), epoch, flag, static_cast<u32>(v));
		status_store(st + 2 * static_cast<u64>(tile) + 1, epoch, flag, static_cast<u32>(v >> 32));
	} else {
		status_store(st + tile, epoch, flag, static_cast<u32>(v));
	}
}

// Flag (0 = not ready) and value of tile `i`'s status.
template <typename Acc>
__device__ __forceinline__ u32 peek(const u64* st, u64 i, u32 epoch, Acc* v) {
	if constexpr (sizeof(Acc) == 8) {
		const u64 lo = status_load(st + 2 * i), hi = status_load(st + 2 * i + 1);
		const u32 f = status_flag(lo, epoch);
		if (f != status_flag(hi, epoch)) return 0u;  // halves of different updates
		*v = static_cast<Acc>(static_cast<u32>(lo)) | (static_cast<Acc>(static_cast<u32>(hi)) << 32);
		return f;
	} else {
		const u64 w = status_load(st + i);
		*v = static_cast<Acc>(static_cast<u32>(w));
		return status_flag(w, epoch);
	}
}

// Exclusive prefix of tile t > 0, by the whole block: each round every
// thread reads kLookPer statuses, so one round covers kLookSpan predecessors
// (the distance back to the nearest inclusive prefix is about the number of
// tiles in flight, so wide rounds keep the walk to one or two device-coherent
// reads).  Offset d of a round is tile j - d; thread tid takes offsets
// tid + kScanBlock * k (coalesced).  After `polls` rounds that found a
// predecessor unpublished, the block sums the input of each unpublished tile
// up to the nearest prefix itself (`base`, the chunk base, belongs to tile 0).
// `late` is set (uniformly) if the walk used a fallback or ended at a late
// prefix.
constexpr u32 kLookPer = 4;
constexpr u32 kLookSpan = kScanBlock * kLookPer;

template <typename Acc>
struct LookShared {
	u32 wmin[kScanBlock / kWave];
	u32 wx[kScanBlock / kWave];
	Acc wsum[kScanBlock / kWave];
	u32 miss[kLookSpan / 32];  // offsets of this round found unpublished
};

// Sum of tile `idx`'s elements, by the whole block (a barrier inside).
template <typename In, typename Acc, int IT>
__device__ Acc tile_sum(const In* in, u64 n, u64 idx, LookShared<Acc>& sh) {
	constexpr int W = kScanBlock / kWave;
	const u64 first = idx * (kScanBlock * IT) + static_cast<u64>(threadIdx.x) * IT;
	In x[IT];
	if (first + IT <= n) {
		load_vec<In, IT>(in + first, x);
	} else {
#pragma unroll
		for (int j = 0; j < IT; ++j) x[j] = first + j < n ? in[first + j] : In(0);
	}
	Acc s = 0;
#pragma unroll
	for (int j = 0; j < IT; ++j) s += static_cast<Acc>(x[j]);
	s = wave_sum(s);
	if (lane_id() == 0) sh.wsum[wave_id()] = s;
	__syncthreads();
	Acc tot = 0;
#pragma unroll
	for (int i = 0; i < W; ++i) tot += sh.wsum[i];
	__syncthreads();
	return tot;
}

template <typename In, typename Acc, int IT>
__device__ Acc look_back(const u64* st, u32 t, u32 epoch, const In* in, u64 n, u32 polls,
	Acc base, LookShared<Acc>& sh, bool* late) {
	constexpr int W = kScanBlock / kWave;
	const u32 tid = threadIdx.x, lane = lane_id(), w = wave_id();
	Acc excl = 0;
	int64_t j = static_cast<int64_t>(t) - 1;
	u32 spins = 0;
	bool was_late = false;
	while (j >= 0) {
		u32 f[kLookPer];
		Acc v[kLookPer];
		u32 my_min = kLookSpan;  // smallest offset holding an inclusive prefix
#pragma unroll
		for (u32 k = 0; k < kLookPer; ++k) {
			const u32 d = tid + kScanBlock * k;
			const int64_t idx = j - static_cast<int64_t>(d);
			v[k] = 0;
			f[k] = kFlagPrefix;  // before tile 0: never reached (tile 0 publishes P)
			if (idx >= 0) f[k] = peek<Acc>(st, static_cast<u64>(idx), epoch, &v[k]);
			if (f[k] >= kFlagPrefix && d < my_min) my_min = d;
		}
		// block minimum of the prefix offsets
		u32 m = my_min;
#pragma unroll
		for (int off = kWave / 2; off > 0; off >>= 1) {
			const u32 o = static_cast<u32>(__shfl_xor(static_cast<int>(m), off, kWave));
			m = o < m ? o : m;
		}
		if (lane == 0) sh.wmin[w] = m;
		if (tid < kLookSpan / 32) sh.miss[tid] = 0u;
		__syncthreads();
		u32 first_p = sh.wmin[0];
#pragma unroll
		for (int i = 1; i < W; ++i) first_p = sh.wmin[i] < first_p ? sh.wmin[i] : first_p;
		// any tile up to the prefix not published yet, and the sum of those
		// that are
		bool missing = false, lp = false;
		Acc part = 0;
#pragma unroll
		for (u32 k = 0; k < kLookPer; ++k) {
			const u32 d = tid + kScanBlock * k;
			if (d <= first_p) {
				if (f[k] == 0u) {
					missing = true;
					if (spins >= polls) atomicOr(&sh.miss[d >> 5], 1u << (d & 31u));
				} else {
					part += v[k];
					lp |= f[k] == kFlagPrefixLate;
				}
			}
		}
		const u64 mb = ballot(missing), lb = ballot(lp);
		const Acc wsum = wave_sum(part);
		if (lane == 0) {
			sh.wx[w] = (mb != 0 ? 1u : 0u) | (lb != 0 ? 2u : 0u);
			sh.wsum[w] = wsum;
		}
		__syncthreads();
		bool any_missing = false, any_late = false;
		Acc sum = 0;
#pragma unroll
		for (int i = 0; i < W; ++i) {
			any_missing |= (sh.wx[i] & 1u) != 0;
			any_late |= (sh.wx[i] & 2u) != 0;
			sum += sh.wsum[i];
		}
		__syncthreads();  // the shared slots are reused below and by the next round
		if (any_missing) {
			if (spins < polls) {
				++spins;
				__builtin_amdgcn_s_sleep(1);
				continue;
			}
			// Fallback: sum the unpublished tiles' inputs (the offsets
			// flagged in sh.miss; uniform across the block).
			was_late = true;
			for (u32 wi = 0; wi < kLookSpan / 32; ++wi) {
				u32 bits = sh.miss[wi];
				while (bits) {
					const u32 d = wi * 32 + static_cast<u32>(__builtin_ctz(bits));
					bits &= bits - 1u;
					const u64 idx = static_cast<u64>(j - static_cast<int64_t>(d));
					sum += tile_sum<In, Acc, IT>(in, n, idx, sh);
					if (idx == 0) sum += base;
				}
			}
		}
		excl += sum;
		was_late |= any_late;
		if (first_p < kLookSpan) break;
		j -= kLookSpan;
	}
	*late = was_late;
	return excl;
}

// Waits, by the whole block, until every tile of this launch has published
// a status word (and so has read the epoch and the chunk base).
template <typename Acc>
__device__ void wait_all_published(const u64* st, u32 tiles, u32 epoch) {
	constexpr u32 kWords = sizeof(Acc) / 4;
	for (;;) {
		int unpublished = 0;
		for (u32 i = threadIdx.x; i < tiles; i += kScanBlock)
			unpublished |= status_flag(status_load(st + static_cast<u64>(i) * kWords), epoch) == 0u;
		if (!__syncthreads_or(unpublished)) return;
		__builtin_amdgcn_s_sleep(1);
	}
}

template <typename In, typename Acc, int IT>
__global__ __launch_bounds__(kScanBlock) void blelloch_onepass(const In* __restrict__ in,
	Acc* __restrict__ out, u64 n, u32 tiles, u64* __restrict__ status, u32* __restrict__ epoch_p,
	Acc* __restrict__ chain, u32 polls) {
	constexpr u32 kScanTile = kScanBlock * IT;
	__shared__ Acc s_wsum[kScanBlock / kWave];
	__shared__ Acc s_prefix;
	__shared__ Acc s_base;
	__shared__ u32 s_epoch;
	__shared__ LookShared<Acc> s_look;
	const u32 t = blockIdx.x;
	const u32 tid = threadIdx.x;
	if (tid == 0) {  // read before this tile publishes (barriers in between)
		s_epoch = *epoch_p & 0x3FFFFFFFu;
		s_base = chain ? *chain : Acc(0);
	}
	const u64 first = static_cast<u64>(t) * kScanTile + static_cast<u64>(tid) * IT;
	const bool full = first + IT <= n;
	In x[IT];
	if (full) {
		load_vec<In, IT>(in + first, x);
	} else {
#pragma unroll
		for (int j = 0; j < IT; ++j) x[j] = first + j < n ? in[first + j] : In(0);
	}
	Acc sum = 0;
#pragma unroll
	for (int j = 0; j < IT; ++j) sum += static_cast<Acc>(x[j]);
	Acc agg;
	Acc run = block_exclusive_sum<kScanBlock>(sum, s_wsum, &agg);  // barriers: s_epoch, s_base
	const u32 epoch = s_epoch;
	const Acc base = s_base;
	bool late = false;
	if (t == 0) {
		if (tid == 0) {
			publish<Acc>(status, 0, epoch, kFlagPrefix, base + agg);
			s_prefix = base;
		}
	} else {
		if (tid == 0) publish<Acc>(status, t, epoch, kFlagAggregate, agg);
		const Acc excl = look_back<In, Acc, IT>(status, t, epoch, in, n, polls, base, s_look,
			&late);
		if (tid == 0) {
			publish<Acc>(status, t, epoch, late ? kFlagPrefixLate : kFlagPrefix, excl + agg);
			s_prefix = excl;
		}
	}
	if (t == tiles - 1 && late) wait_all_published<Acc>(status, tiles, epoch);
	__syncthreads();
	const Acc prefix = s_prefix;
	if (tid == 0 && t == tiles - 1) {
		if (chain) *chain = prefix + agg;
		*epoch_p = (epoch + 1u) & 0x3FFFFFFFu;
	}
	run += prefix;
	Acc y[IT];
#pragma unroll
	for (int j = 0; j < IT; ++j) {
		y[j] = run;
		run += static_cast<Acc>(x[j]);
	}
	if (full) {
		store_vec<Acc, IT>(out + first, y);
	} else {
#pragma unroll
		for (int j = 0; j < IT; ++j)
			if (first + j < n) out[first + j] = y[j];
	}
}

struct BlellochData {
	Scratch tiles;   // per-tile sums (then carries) and segment flags
	Scratch status;  // single pass: status words, then the epoch word (zeroed once)
	BlellochData() { status.zero_on_alloc = true; }
};

struct ScanLaunch {
	const void* in;
	void* out;
	u64 n;
	u32 tiles;
	void* tsum;
	u32* tflag;
	const u32* off;  // segment starts (segmented scan) or nullptr
	u32 nseg;
	u32 items;  // items per thread: 16, or 8 / 32 for unsigned plain scans
	void* chain;  // running total across chunks (plain scans) or nullptr
	u64* status;  // single pass: epoch word, an unused word, then the tiles' status words
	CCLQueue* q;
};

template <typename In, typename Acc, int OUTW, bool SEG, int IT>
bool launch_scan3(const ScanLaunch& a, GError** err) {
	hipStream_t st = a.q->stream;
	const u32 grid = xcd_grid(a.tiles);
	const In* in = static_cast<const In*>(a.in);
	Acc* tsum = static_cast<Acc*>(a.tsum);
	if constexpr ((std::is_same<In, u32>::value || std::is_same<In, u64>::value) &&
		sizeof(Acc) == OUTW) {
		if (!SEG && a.status && a.tiles <= kOnePassMaxTiles &&
			((reinterpret_cast<uintptr_t>(a.in) | reinterpret_cast<uintptr_t>(a.out)) & 15) == 0) {
			Launch L(a.q, CLO_SCAN_BLELLOCH_KNAME_ONEPASS);
			// status[0]: epoch; status[1]: unused; tiles' words from status[2]
			blelloch_onepass<In, Acc, IT><<<a.tiles, kScanBlock, 0, st>>>(in,
				static_cast<Acc*>(a.out), a.n, a.tiles, a.status + 2,
				reinterpret_cast<u32*>(a.status), static_cast<Acc*>(a.chain),
				g_lookback_polls.load(std::memory_order_relaxed));
			return L.done(err);
		}
		if (((reinterpret_cast<uintptr_t>(a.in) | reinterpret_cast<uintptr_t>(a.out)) & 15) == 0) {
			{
				Launch L(a.q, CLO_SCAN_BLELLOCH_KNAME_REDUCE);
				blelloch_tile_reduce_vec<In, Acc, SEG, IT><<<grid, kScanBlock, 0, st>>>(in, a.n,
					a.tiles, tsum, a.off, a.nseg, a.tflag);
				if (!L.done(err)) return false;
			}
			{
				Launch L(a.q, CLO_SCAN_BLELLOCH_KNAME_TILES);
				blelloch_tiles_scan<Acc, SEG><<<1, kTilesBlock, 0, st>>>(tsum, a.tflag, a.tiles,
					static_cast<Acc*>(a.chain));
				if (!L.done(err)) return false;
			}
			Launch L(a.q, CLO_SCAN_BLELLOCH_KNAME_SCAN);
			blelloch_tile_scan_vec<In, Acc, SEG, IT><<<grid, kScanBlock, 0, st>>>(in,
				static_cast<Acc*>(a.out), a.n, a.tiles, tsum, a.off, a.nseg);
			return L.done(err);
		}
	}
	{
		Launch L(a.q, CLO_SCAN_BLELLOCH_KNAME_REDUCE);
		blelloch_tile_reduce<In, Acc, SEG, IT><<<grid, kScanBlock, 0, st>>>(in, a.n, a.tiles, a.off,
			a.nseg, tsum, a.tflag);
		if (!L.done(err)) return false;
	}
	{
		Launch L(a.q, CLO_SCAN_BLELLOCH_KNAME_TILES);
		blelloch_tiles_scan<Acc, SEG><<<1, kTilesBlock, 0, st>>>(tsum, a.tflag, a.tiles,
			static_cast<Acc*>(a.chain));
		if (!L.done(err)) return false;
	}
	{
		Launch L(a.q, CLO_SCAN_BLELLOCH_KNAME_SCAN);
		blelloch_tile_scan<In, Acc, OUTW, SEG, IT><<<grid, kScanBlock, 0, st>>>(in, a.out, a.n,
			a.tiles, tsum, a.off, a.nseg);
		if (!L.done(err)) return false;
	}
	return true;
}

template <typename In, typename Acc, int OUTW>
bool launch_scan(const ScanLaunch& a, GError** err) {
	if (a.off) return launch_scan3<In, Acc, OUTW, true, kScanItems>(a, err);
	if constexpr ((std::is_same<In, u32>::value || std::is_same<In, u64>::value) && OUTW >= 4) {
		if (a.items == 8) return launch_scan3<In, Acc, OUTW, false, 8>(a, err);
		if (a.items == 32) return launch_scan3<In, Acc, OUTW, false, 32>(a, err);
	}
	if (a.items != kScanItems) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"items per thread %u needs an unsigned 32/64-bit plain scan", a.items);
		return false;
	}
	return launch_scan3<In, Acc, OUTW, false, kScanItems>(a, err);
}

template <typename In>
bool launch_for_in(size_t sum_size, const ScanLaunch& a, GError** err) {
	switch (sum_size) {
	case 1: return launch_scan<In, u32, 1>(a, err);
	case 2: return launch_scan<In, u32, 2>(a, err);
	case 4: return launch_scan<In, u32, 4>(a, err);
	case 8: return launch_scan<In, u64, 8>(a, err);
	}
	set_error(err, kQuarkClo, CLO_ERROR_ARGS, "unsupported sum size %zu", sum_size);
	return false;
}

const char* blelloch_init(CloScan* scanner, const char* options, GError** err) {
	// clo_scan_blelloch.c:43-45: any option is an error.
	if (options && strlen(options) > 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "Invalid options for blelloch scan.");
		return nullptr;
	}
	scanner->data = new BlellochData();
	return "blelloch (gfx950: tile reduce + tile-sum scan + tile scan)";
}

void blelloch_finalize(CloScan* scanner) {
	delete static_cast<BlellochData*>(scanner->data);
	scanner->data = nullptr;
}

// Workspace carved from a caller-provided device region (the no-allocation
// scan): the single-pass status words first (kWsStatusBytes, zeroed once when
// the region is set up), then the tile sums, then the segment flags.
constexpr size_t kWsStatusBytes = ((2 + 2 * static_cast<size_t>(kOnePassMaxTiles)) * 8 + 255) /
	256 * 256;

// Launches the scan of numel elements at `in` into `out` on cq_exec.  `ws`
// (device memory) replaces the scanner's own scratch when set; the caller
// has checked that it is large enough.
bool scan_ptrs(CloScan* scanner, CCLQueue* cq_exec, const void* in, void* out, size_t numel,
	const u32* off, u32 nseg, GError** err, u32 items, void* chain,
	unsigned char* ws = nullptr) {
	BlellochData* bd = static_cast<BlellochData*>(scanner->data);
	if (items != 8 && items != 16 && items != 32) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "items per thread must be 8, 16 or 32");
		return false;
	}
	const size_t tiles = div_ceil(numel, static_cast<size_t>(kScanBlock) * items);
	if (tiles > 0x7FFFFFFFull) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "numel too large");
		return false;
	}
	const size_t ss = clo_type_sizeof(scanner->sum_type);
	const size_t acc = ss == 8 ? 8 : 4;
	const size_t sum_bytes = (tiles * acc + 255) / 256 * 256;
	void* tiles_ptr = ws ? static_cast<void*>(ws + kWsStatusBytes) : nullptr;
	if (!ws) {
		if (!bd->tiles.ensure(sum_bytes + tiles * 4, err)) return false;
		tiles_ptr = bd->tiles.ptr;
	}
	// Single pass (plain scans): the epoch word, an unused word and (sum
	// bytes / 4) status words per tile.  Growing reallocates zeroed memory:
	// epoch 0, no status.
	u64* status = nullptr;
	// In-place scans (in == out, equal widths) take reduce-then-scan: the
	// single pass may re-read a slow predecessor tile's input after that tile
	// has overwritten it with its prefixes.
	if (!off && tiles <= kOnePassMaxTiles && in != out) {
		if (ws) {
			status = reinterpret_cast<u64*>(ws);
		} else {
			if (!bd->status.ensure((2 + tiles * (acc / 4)) * sizeof(u64), err)) return false;
			status = static_cast<u64*>(bd->status.ptr);
		}
	}
	ScanLaunch a;
	a.in = in;
	a.out = out;
	a.n = numel;
	a.tiles = static_cast<u32>(tiles);
	a.tsum = tiles_ptr;
	a.tflag = reinterpret_cast<u32*>(static_cast<char*>(tiles_ptr) + sum_bytes);
	a.off = off;
	a.nseg = nseg;
	a.items = items;
	a.chain = chain;
	a.status = status;
	a.q = cq_exec;
	CLO_LOG(1, "blelloch: numel=%zu elem=%zu B sum=%zu B tiles=%zu of 256x%u %s%s", numel,
		clo_type_sizeof(scanner->elem_type), ss, tiles, items,
		off ? "segmented (reduce-then-scan)"
			: status ? "single pass (decoupled look-back) if plain unsigned, else reduce-then-scan"
					 : "reduce-then-scan",
		ws ? ", caller workspace" : "");
	bool ok = false;
	switch (scanner->elem_type) {
	case CLO_CHAR: ok = launch_for_in<int8_t>(ss, a, err); break;
	case CLO_UCHAR: ok = launch_for_in<u8>(ss, a, err); break;
	case CLO_SHORT: ok = launch_for_in<int16_t>(ss, a, err); break;
	case CLO_USHORT: ok = launch_for_in<u16>(ss, a, err); break;
	case CLO_INT: ok = launch_for_in<int32_t>(ss, a, err); break;
	case CLO_UINT: ok = launch_for_in<u32>(ss, a, err); break;
	case CLO_LONG: ok = launch_for_in<int64_t>(ss, a, err); break;
	case CLO_ULONG: ok = launch_for_in<u64>(ss, a, err); break;
	default:
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "unsupported scan types");
		break;
	}
	return ok;
}

CCLEvent* scan_common(CloScan* scanner, CCLQueue* cq_exec, CCLBuffer* data_in,
	CCLBuffer* data_out, size_t numel, const u32* off, u32 nseg, GError** err,
	u32 items = kScanItems) {
	if (numel > 0 &&
		!scan_ptrs(scanner, cq_exec, data_in->ptr, data_out->ptr, numel, off, nseg, err, items,
			nullptr))
		return nullptr;
	return record_event(cq_exec, "blelloch_done", err);
}

CCLEvent* blelloch_scan_with_device_data(CloScan* scanner, CCLQueue* cq_exec, CCLQueue* cq_comm,
	CCLBuffer* data_in, CCLBuffer* data_out, size_t numel, size_t lws_max, GError** err) {
	(void)cq_comm;
	(void)lws_max;
	return scan_common(scanner, cq_exec, data_in, data_out, numel, nullptr, 0, err);
}

cl_uint blelloch_get_num_kernels(CloScan* scanner, GError** err) {
	(void)scanner;
	(void)err;
	return CLO_SCAN_BLELLOCH_NUM_KERNELS;
}

const char* blelloch_get_kernel_name(CloScan* scanner, cl_uint i, GError** err) {
	(void)scanner;
	if (i >= CLO_SCAN_BLELLOCH_NUM_KERNELS) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "kernel index %u out of range", i);
		return nullptr;
	}
	static const char* names[] = CLO_SCAN_BLELLOCH_KERNELNAMES;
	return names[i];
}

size_t blelloch_get_localmem_usage(CloScan* scanner, cl_uint i, size_t lws_max, size_t numel,
	GError** err) {
	(void)lws_max;
	(void)numel;
	if (i >= CLO_SCAN_BLELLOCH_NUM_KERNELS) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "kernel index %u out of range", i);
		return 0;
	}
	const size_t acc = clo_type_sizeof(scanner->sum_type) > 4 ? 8 : 4;
	constexpr size_t kScanTile = static_cast<size_t>(kScanBlock) * kScanItems;
	if (i == 1) return pad(kTilesRound) * acc + kTilesRound + (kTilesBlock / kWave) * (acc + 8);
	if (i == 3) return (kScanBlock / kWave + 1) * acc;
	return (pad(kScanTile) + 1 + 2 * (kScanBlock / kWave)) * acc + kScanTile / 8 + 4;
}

}  // namespace

bool clo_blelloch_scan_chunk(CloScan* scanner, CCLQueue* cq, const void* in, void* out,
	size_t numel, void* chain, GError** err) {
	if (numel == 0) return true;
	return scan_ptrs(scanner, cq, in, out, numel, nullptr, 0, err, kScanItems, chain);
}

CCLEvent* clo_blelloch_scan_items(CloScan* scanner, CCLQueue* cq, CCLBuffer* in, CCLBuffer* out,
	size_t numel, uint32_t items, GError** err) {
	if (err && *err) return nullptr;
	if (!scanner || !cq || !in || !out || strcmp(scanner->impl.name, "blelloch") != 0 ||
		in->size < numel * clo_type_sizeof(scanner->elem_type) ||
		out->size < numel * clo_type_sizeof(scanner->sum_type)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad blelloch scan arguments");
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	return scan_common(scanner, cq, in, out, numel, nullptr, 0, err, items);
}

extern "C" {

// Host-pass only: the vtable holds host function addresses.
#if !defined(__HIP_DEVICE_COMPILE__)
const CloScanImplDef clo_scan_blelloch_def = {"blelloch", blelloch_init, blelloch_finalize,
	blelloch_scan_with_device_data, blelloch_get_num_kernels, blelloch_get_kernel_name,
	blelloch_get_localmem_usage};
#endif


CloScan* clo_scan_new(const char* type, const char* options, CCLContext* ctx, CloType elem_type,
	CloType sum_type, const char* compiler_opts, GError** err) {
	(void)compiler_opts;
	if (err && *err) return nullptr;
	if (!ctx) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "clo_scan_new: NULL context");
		return nullptr;
	}
	const CloScanImplDef* impl = nullptr;
	if (type && strcmp(type, clo_scan_blelloch_def.name) == 0) impl = &clo_scan_blelloch_def;
	if (!impl) {
		set_error(err, kQuarkClo, CLO_ERROR_IMPL_NOT_FOUND,
			"The requested scan implementation, '%s', was not found.", type ? type : "(null)");
		return nullptr;
	}
	if (!type_is_integer(elem_type) || !type_is_integer(sum_type)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"Scan types must be integer types (char..ulong); floating-point sums are not "
			"reproducible bit for bit.");
		return nullptr;
	}
	CloScan* s = new clo_scan();
	s->impl = *impl;
	s->ctx = ctx;
	ccl_context_ref(ctx);
	s->elem_type = elem_type;
	s->sum_type = sum_type;
	s->data = nullptr;
	GError* e = nullptr;
	const char* src = s->impl.init(s, options, &e);
	if (!src) {
		if (e) {
			if (err && !*err) *err = e;
			else clo_error_free(e);
		}
		clo_scan_destroy(s);
		return nullptr;
	}
	return s;
}

void clo_scan_destroy(CloScan* scan) {
	if (!scan) return;
	if (scan->impl.finalize && scan->data) scan->impl.finalize(scan);
	if (scan->host_cache) host_cache_free(scan->host_cache);
	ccl_context_destroy(scan->ctx);
	delete scan;
}

// Input and output either disjoint or the same address with equal element
// and sum widths (an in-place scan); anything else overlaps destructively.
static bool scan_placement_ok(CloScan* sc, CCLBuffer* in, CCLBuffer* out, size_t numel,
	GError** err) {
	const size_t ib = numel * clo_type_sizeof(sc->elem_type);
	const size_t ob = numel * clo_type_sizeof(sc->sum_type);
	const uintptr_t x = reinterpret_cast<uintptr_t>(in->ptr);
	const uintptr_t y = reinterpret_cast<uintptr_t>(out->ptr);
	const bool overlap = ib && ob && x < y + ob && y < x + ib;
	if (overlap && !(x == y && ib == ob)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"scan output overlaps the input (only an in-place scan with equal element and sum "
			"widths may share it)");
		return false;
	}
	return true;
}

CCLEvent* clo_scan_with_device_data(CloScan* scanner, CCLQueue* cq_exec, CCLQueue* cq_comm,
	CCLBuffer* data_in, CCLBuffer* data_out, size_t numel, size_t lws_max, GError** err) {
	ApiRange api_range("clo_scan_with_device_data");
	if (err && *err) return nullptr;
	if (!scanner || !cq_exec || !data_in || !data_out) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL scanner, queue or buffer");
		return nullptr;
	}
	if (data_in->size < numel * clo_type_sizeof(scanner->elem_type) ||
		data_out->size < numel * clo_type_sizeof(scanner->sum_type)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "buffers smaller than numel elements");
		return nullptr;
	}
	if (!scan_placement_ok(scanner, data_in, data_out, numel, err)) return nullptr;
	if (!use_device(cq_exec->ctx, err)) return nullptr;
	return scanner->impl.scan_with_device_data(scanner, cq_exec, cq_comm, data_in, data_out,
		numel, lws_max, err);
}

CCLEvent* clo_scan_segmented_with_device_data(CloScan* scanner, CCLQueue* cq_exec,
	CCLQueue* cq_comm, CCLBuffer* data_in, CCLBuffer* data_out, CCLBuffer* offsets, size_t nseg,
	size_t numel, GError** err) {
	ApiRange api_range("clo_scan_segmented_with_device_data");
	(void)cq_comm;
	if (err && *err) return nullptr;
	if (!scanner || !cq_exec || !data_in || !data_out || !offsets) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL scanner, queue or buffer");
		return nullptr;
	}
	if (data_in->size < numel * clo_type_sizeof(scanner->elem_type) ||
		data_out->size < numel * clo_type_sizeof(scanner->sum_type) || offsets->size < nseg * 4 ||
		nseg > 0xFFFFFFFFull || numel > 0xFFFFFFFFull) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"buffers smaller than numel elements / nseg offsets, or sizes beyond 2^32-1");
		return nullptr;
	}
	if (!scan_placement_ok(scanner, data_in, data_out, numel, err)) return nullptr;
	if (!use_device(cq_exec->ctx, err)) return nullptr;
	if (strcmp(scanner->impl.name, "blelloch") != 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "segmented scan needs a blelloch scanner");
		return nullptr;
	}
	return scan_common(scanner, cq_exec, data_in, data_out, numel,
		static_cast<const u32*>(offsets->ptr), static_cast<u32>(nseg), err);
}

cl_bool clo_scan_with_host_data(CloScan* scanner, CCLQueue* cq_exec, CCLQueue* cq_comm,
	void* data_in, void* data_out, size_t numel, size_t lws_max, GError** err) {
	ApiRange api_range("clo_scan_with_host_data");
	(void)lws_max;
	if (err && *err) return CL_FALSE;
	if (!scanner || !data_in || !data_out) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL scanner or host pointer");
		return CL_FALSE;
	}
	CCLQueue* own = nullptr;
	if (!cq_exec) {
		own = ccl_queue_new(scanner->ctx, nullptr, 0, err);
		if (!own) return CL_FALSE;
		cq_exec = own;
	}
	if (!cq_comm) cq_comm = cq_exec;
	const bool ok = scan_host_path(scanner, cq_exec, cq_comm, data_in, data_out, numel, err);
	if (own) ccl_queue_destroy(own);
	return ok ? CL_TRUE : CL_FALSE;
}

CCLContext* clo_scan_get_context(CloScan* scanner) { return scanner ? scanner->ctx : nullptr; }
CCLProgram* clo_scan_get_program(CloScan* scanner) {
	(void)scanner;
	return nullptr;
}
CloType clo_scan_get_elem_type(CloScan* scanner) { return scanner->elem_type; }
size_t clo_scan_get_element_size(CloScan* scanner) { return clo_type_sizeof(scanner->elem_type); }
CloType clo_scan_get_sum_type(CloScan* scanner) { return scanner->sum_type; }
size_t clo_scan_get_sum_size(CloScan* scanner) { return clo_type_sizeof(scanner->sum_type); }
void* clo_scan_get_data(CloScan* scanner) { return scanner ? scanner->data : nullptr; }
void clo_scan_set_data(CloScan* scanner, void* data) {
	if (scanner) scanner->data = data;
}
cl_uint clo_scan_get_num_kernels(CloScan* scanner, GError** err) {
	return scanner ? scanner->impl.get_num_kernels(scanner, err) : 0;
}
const char* clo_scan_get_kernel_name(CloScan* scanner, cl_uint i, GError** err) {
	return scanner ? scanner->impl.get_kernel_name(scanner, i, err) : nullptr;
}
size_t clo_scan_get_localmem_usage(CloScan* scanner, cl_uint i, size_t lws_max, size_t numel,
	GError** err) {
	return scanner ? scanner->impl.get_localmem_usage(scanner, i, lws_max, numel, err) : 0;
}

void clo_scan_debug_lookback_polls(int polls) {
	g_lookback_polls.store(polls < 0 ? kLookbackPolls : static_cast<u32>(polls),
		std::memory_order_relaxed);
}

}  // extern "C"

// ---- no-allocation scan --------------------------------------------------
// Workspace of clo_scan_with_device_data_noalloc: a static array that every
// GPU the library runs on holds until the process ends, each at its own
// address.  Host code never passes the name `scratch` to a kernel (on the
// host it is only a stand-in for the symbol): it uses the device address
// hipGetSymbolAddress returns for the current device, kept per device.
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mu;
static unsigned char* g_noalloc_ptr[kNoAllocMaxDevices];

// Clears the single-pass status words (epoch 0, no status published): a
// kernel rather than a memset, so a first call made under graph capture
// records no memset node.
__global__ __launch_bounds__(256) void noalloc_zero(u64* p, u32 words) {
	for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x)
		p[i] = 0;
}

// Largest numel the workspace holds for a sum width: status words, then
// (acc + 4) bytes per tile of kScanBlock * kScanItems elements (the tile sums
// rounded up to 256 bytes).
static size_t noalloc_max_numel(size_t sum_size) {
	const size_t acc = sum_size == 8 ? 8 : 4;
	const size_t avail = kNoAllocScratchBytes - kWsStatusBytes - 256;
	return avail / (acc + 4) * (static_cast<size_t>(kScanBlock) * kScanItems);
}

extern "C" {

size_t clo_scan_noalloc_max_numel(CloScan* scanner) {
	return scanner ? noalloc_max_numel(clo_type_sizeof(scanner->sum_type)) : 0;
}

cl_bool clo_scan_with_device_data_noalloc(CloScan* scanner, CCLQueue* cq_exec,
	CCLBuffer* data_in, CCLBuffer* data_out, size_t numel, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!scanner || !cq_exec || !data_in || !data_out) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL scanner, queue or buffer");
		return CL_FALSE;
	}
	if (strcmp(scanner->impl.name, "blelloch") != 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "the no-allocation scan needs a blelloch scanner");
		return CL_FALSE;
	}
	const size_t maxn = noalloc_max_numel(clo_type_sizeof(scanner->sum_type));
	if (numel > maxn) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"numel %zu exceeds the no-allocation scan's limit of %zu elements for %zu-byte "
			"sums (its workspace is %zu bytes)", numel, maxn,
			clo_type_sizeof(scanner->sum_type), kNoAllocScratchBytes);
		return CL_FALSE;
	}
	if (data_in->size < numel * clo_type_sizeof(scanner->elem_type) ||
		data_out->size < numel * clo_type_sizeof(scanner->sum_type)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "buffers smaller than numel elements");
		return CL_FALSE;
	}
	if (!scan_placement_ok(scanner, data_in, data_out, numel, err)) return CL_FALSE;
	if (!use_device(cq_exec->ctx, err)) return CL_FALSE;
	const int dev = cq_exec->ctx->dev;
	if (dev < 0 || dev >= kNoAllocMaxDevices) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "device index %d out of range", dev);
		return CL_FALSE;
	}
	unsigned char* ws = nullptr;
	{
		std::lock_guard<std::mutex> lock(g_noalloc_mu);
		ws = g_noalloc_ptr[dev];
		if (!ws) {
			// Set-up, once per device with that device current: the array's
			// address on this GPU, and its status words cleared ahead of the
			// first scan on the same queue.
			void* p = nullptr;
			if (hip_failed(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), err,
					"hipGetSymbolAddress"))
				return CL_FALSE;
			ws = static_cast<unsigned char*>(p);
			noalloc_zero<<<4, 256, 0, cq_exec->stream>>>(reinterpret_cast<u64*>(ws),
				static_cast<u32>(kWsStatusBytes / 8));
			if (hip_failed(hipGetLastError(), err, "noalloc_zero")) return CL_FALSE;
			g_noalloc_ptr[dev] = ws;
		}
	}
	if (numel == 0) return CL_TRUE;
	g_launch_quiet = true;
	const bool ok = scan_ptrs(scanner, cq_exec, data_in->ptr, data_out->ptr, numel, nullptr, 0,
		err, kScanItems, nullptr, ws);
	g_launch_quiet = false;
	return ok ? CL_TRUE : CL_FALSE;
}

}  // extern "C"

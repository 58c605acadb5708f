// satradix: stable LSD radix sort for gfx950.
//
// Result contract (CL_Ops' src/cl_ops/sort/clo_sort_satradix.c:117-335 and
// clo_sort_satradix.cl:34-258): a stable sort of the elements by the unsigned
// value of their bits, ascending, whatever the compare snippet (the reference
// kernels never use CLO_SORT_COMPARE).  Key-value pairs keep the payloads of
// equal keys in input order.
//
// Kernel plan per call (all on cq_exec, no host synchronisation, no
// inter-block waiting):
//   satradix_upsweep    first pass, per tile (4096 to 16384 keys): 16-bit
//                       digit counts from the keys, and their bitwise AND/OR
//   satradix_upsweep_digits  later passes, one wave per tile: counts of the
//                       digit bytes the previous pass wrote (1 byte per key)
//   satradix_plan       one block, only for sorts of at most 256 tiles: which
//                       digits vary between keys (the other passes are
//                       skipped) and the buffers of each pass
//   satradix_reduce     per 256 tiles and quarter of the digits: sums of the
//                       tiles' digit counts (first pass: also the group's
//                       AND/OR)
//   satradix_scan       per 256 tiles and quarter of the digits: tile counts
//                       -> global output offsets (each block derives its
//                       bases from all group sums; in the first pass it
//                       derives the plan from the group AND/OR); sorts of at
//                       most 256 tiles skip reduce and scan: each downsweep
//                       block sums the tile counts itself
//   satradix_downsweep  per tile: wave64 ballot ranking, LDS-staged scatter
//                       to the offsets (contiguous runs per digit), and the
//                       next executed pass's digit byte of every key
//   satradix_copy       only when an in-place sort executed an odd number of
//                       passes (the result sits in the scratch buffer)
// The reduce-then-scan structure reads the keys once per pass, plus one byte
// per key for the upsweep.  A single-pass decoupled look-back variant was measured on
// MI355X: the per-block ticket atomic serialises (~11 ns per block) and the
// look-back walks over ~30 predecessors per tile in 256-digit status rows,
// which made the pass 2.3x slower than the scatter alone.

#include <cstring>
#include <string>

#include "clo_internal.hpp"
#include "dev_common.hpp"
#include "radix_tile.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

constexpr int kRadixMax = 256;
constexpr int kMaxPasses = 64;
constexpr u32 kScanTiles = 256;  // tiles per group of the counter reduce/scan

enum BufSel : u32 { kIn = 0, kAux = 1, kOut = 2 };

// One digit pass.  `dsrc` = 1: the pass's upsweep counts the digit bytes the
// previous executed pass's downsweep wrote (1 byte per key) instead of reading
// the keys again; `nact` = 1: this pass's downsweep writes those bytes for the
// next executed pass, whose digit starts at bit `nshift`.
struct PassDesc {
	u32 active;
	u32 shift;
	u32 src;
	u32 dst;
	u32 dsrc;
	u32 nact;
	u32 nshift;
	u32 pad;
};

struct SortPlan {
	PassDesc pass[kMaxPasses];
	u32 copy_active;
	u32 copy_src;
	u32 copy_dst;
	u32 executed;
};

template <typename K>
struct Bufs {
	K* k[3];
	u32* v[3];
};

// Digit counts of one tile of NT*ITEMS keys for the first pass, which reads
// the input and also records the tile's bitwise AND and OR of the keys, from
// which the plan finds the digits that are equal in every key.
template <typename K, int NT, int ITEMS>
__global__ __launch_bounds__(NT) void satradix_upsweep(const K* __restrict__ src, u32 n,
	u32 tiles, u32 bits, u16* __restrict__ counts, K* __restrict__ andor) {
	constexpr u32 kTile = NT * ITEMS;
	constexpr int NW = NT / kWave;
	static_assert(NT >= kRadixMax, "one thread per digit bin");
	// Four copies of each wave's histogram, lane l adding to copy l % 4:
	// lanes holding the same digit collide on one counter a quarter as often.
	// The counting is bound by these LDS atomics; on 2^28 u64 keys the
	// upsweep takes 0.124 ms (one copy: 0.129) on uniform keys and 0.254 ms
	// (one copy: 0.66, two: 0.44) on Zipf keys, whose top key is 11.5%.
	__shared__ __attribute__((aligned(16))) u32 s_h[NW][4 * kRadixMax];
	__shared__ K s_red[2 * NW];
	const u32 tile = xcd_tile(tiles);
	if (tile >= tiles) return;
	const u32 tid = threadIdx.x;
	const u32 lane = lane_id();
	const u32 w = wave_id();
	for (u32 i = lane; i < kRadixMax; i += kWave)
		reinterpret_cast<uint4*>(s_h[w])[i] = make_uint4(0u, 0u, 0u, 0u);
	__syncthreads();
	const u32 copy = lane & 3u;
	const u32 mask = (1u << bits) - 1u;
	const u32 wbase = tile * kTile + w * kWave * ITEMS;
	K k[ITEMS];
#pragma unroll
	for (int j = 0; j < ITEMS; ++j) {
		const u32 idx = wbase + j * kWave + lane;
		k[j] = idx < n ? src[idx] : K(0);
	}
	K ka = ~K(0), ko = K(0);
#pragma unroll
	for (int j = 0; j < ITEMS; ++j) {
		if (wbase + j * kWave + lane < n) {
			atomicAdd(&s_h[w][(static_cast<u32>(k[j]) & mask) * 4 + copy], 1u);
			ka &= k[j];
			ko |= k[j];
		}
	}
#pragma unroll
	for (int off = kWave / 2; off > 0; off >>= 1) {
		ka &= __shfl_xor(ka, off, kWave);
		ko |= __shfl_xor(ko, off, kWave);
	}
	if (lane == 0) {
		s_red[w] = ka;
		s_red[NW + w] = ko;
	}
	__syncthreads();
	if (tid < kRadixMax) {
		u32 c = 0;
#pragma unroll
		for (int ww = 0; ww < NW; ++ww) {
			const uint4 h = reinterpret_cast<const uint4*>(s_h[ww])[tid];
			c += h.x + h.y + h.z + h.w;
		}
		counts[static_cast<size_t>(tile) * kRadixMax + tid] = static_cast<u16>(c);
	}
	if (tid == 0) {
		K a = ~K(0), o = K(0);
#pragma unroll
		for (int ww = 0; ww < NW; ++ww) {
			a &= s_red[ww];
			o |= s_red[NW + ww];
		}
		andor[2 * static_cast<size_t>(tile)] = a;
		andor[2 * static_cast<size_t>(tile) + 1] = o;
	}
}

// Digit counts of one tile for the passes after the first: one wave per tile
// of TILE keys.  A pass whose previous executed pass wrote digit bytes
// (dsrc) reads those, 1 byte per key (TILE / 256 words per lane, all loads in
// flight); otherwise (the first executed pass is not pass 0) it reads the
// keys.  The wave keeps kDigitCopies copies of its histogram, lane l adding to
// copy l % kDigitCopies.  Measured alone on 2^28 digit bytes in 8192-digit
// tiles: 0.059 ms per pass, against 0.070 ms for four 64-lane waves per tile
// with four copies each, and 0.043 ms for reading the bytes alone; 2 and 16
// copies, copy-major layouts and two waves per tile were slower.  Inside the
// 2^28-pair sort (after a downsweep) it takes 0.083 ms against about 0.085.
constexpr int kDigitCopies = 8;

template <typename K, int TILE>
__global__ __launch_bounds__(kWave) void satradix_upsweep_digits(Bufs<K> b,
	const SortPlan* __restrict__ plan, u32 pass, u32 n, u32 tiles, u32 bits,
	u16* __restrict__ counts, const u8* __restrict__ digits) {
	constexpr u32 C = kDigitCopies;
	constexpr int WPL = TILE / 4 / kWave;  // digit words per lane
	static_assert(TILE % (4 * kWave) == 0, "whole digit words per lane");
	__shared__ __attribute__((aligned(16))) u32 s_h[C * kRadixMax];
	const PassDesc d = plan->pass[pass];
	if (!d.active) return;
	const u32 tile = xcd_tile(tiles);
	if (tile >= tiles) return;
	const u32 lane = threadIdx.x;
	const u32 base = tile * static_cast<u32>(TILE);
	const u32 copy = lane % C;
	if (d.dsrc && base + TILE <= n) {
		// TILE is a multiple of 4 and the digit array is 16-byte aligned.
		const u32* __restrict__ dw = reinterpret_cast<const u32*>(digits + base);
		u32 x[WPL];
#pragma unroll
		for (int j = 0; j < WPL; ++j) x[j] = dw[j * kWave + lane];
#pragma unroll
		for (u32 i = lane; i < C * kRadixMax / 4; i += kWave)
			reinterpret_cast<uint4*>(s_h)[i] = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
		for (int j = 0; j < WPL; ++j) {
#pragma unroll
			for (int q = 0; q < 4; ++q)
				atomicAdd(&s_h[((x[j] >> (8 * q)) & 0xFFu) * C + copy], 1u);
		}
	} else {
#pragma unroll
		for (u32 i = lane; i < C * kRadixMax / 4; i += kWave)
			reinterpret_cast<uint4*>(s_h)[i] = make_uint4(0u, 0u, 0u, 0u);
		const u32 end = (n - base) < static_cast<u32>(TILE) ? n : base + TILE;
		if (d.dsrc) {
			for (u32 i = base + lane; i < end; i += kWave) atomicAdd(&s_h[digits[i] * C + copy], 1u);
		} else {
			const K* __restrict__ src = b.k[d.src];
			const u32 mask = (1u << bits) - 1u;
			for (u32 i = base + lane; i < end; i += kWave)
				atomicAdd(&s_h[(static_cast<u32>(src[i] >> d.shift) & mask) * C + copy], 1u);
		}
	}
	__syncthreads();
#pragma unroll
	for (u32 dg = lane; dg < kRadixMax; dg += kWave) {
		u32 c = 0;
#pragma unroll
		for (u32 k = 0; k < C; k += 4) {
			const uint4 h = reinterpret_cast<const uint4*>(s_h)[(dg * C + k) / 4];
			c += h.x + h.y + h.z + h.w;
		}
		counts[static_cast<size_t>(tile) * kRadixMax + dg] = static_cast<u16>(c);
	}
}

// The pass plan from the bits that differ between keys, written by one full
// wave: lane p describes pass p (kMaxPasses == kWave).  Executed pass k (k-th
// active one) reads the buffer pass k-1 wrote and writes kAux or kIn/kOut in
// turn, so that an out-of-place sort ends in kOut; the digit-byte hand-over
// links each executed pass with the next one.
static_assert(kMaxPasses == kWave, "one lane per pass");

__device__ __forceinline__ u32 plan_dst(u32 k, u32 executed, u32 in_place) {
	if (in_place) return (k % 2 == 0) ? kAux : kIn;
	return ((executed - 1 - k) % 2 == 0) ? kOut : kAux;  // the last one lands in kOut
}

template <typename K>
__device__ void plan_write(SortPlan* __restrict__ plan, K diff, u32 passes, u32 bits,
	u32 in_place) {
	const u32 p = lane_id();
	const u32 mask = (1u << bits) - 1u;
	const bool act = p < passes && (static_cast<u32>(diff >> (p * bits)) & mask) != 0u;
	const u64 am = ballot(act);
	const u32 executed = static_cast<u32>(__popcll(am));
	const u32 k = count_below(am);  // executed passes before pass p
	PassDesc d = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
	if (act) {
		d.active = 1;
		d.shift = p * bits;
		d.src = k == 0 ? kIn : plan_dst(k - 1, executed, in_place);
		d.dst = plan_dst(k, executed, in_place);
		d.dsrc = k > 0 ? 1u : 0u;
		const u64 later = p + 1 < 64 ? (am >> (p + 1)) : 0ull;
		if (later) {
			d.nact = 1;
			d.nshift = (p + 1 + static_cast<u32>(__builtin_ctzll(later))) * bits;
		}
	}
	plan->pass[p] = d;
	if (p == 0) {
		plan->executed = executed;
		plan->copy_active = 0;
		plan->copy_src = 0;
		plan->copy_dst = 0;
		if (in_place && (executed % 2 == 1)) {
			plan->copy_active = 1;
			plan->copy_src = kAux;
			plan->copy_dst = kIn;
		} else if (!in_place && executed == 0) {
			plan->copy_active = 1;
			plan->copy_src = kIn;
			plan->copy_dst = kOut;
		}
	}
}

// Which digit passes run (a digit equal in every key needs no pass), and the
// source and destination buffer of each executed pass.  One block of
// kPlanThreads; each thread keeps kPlanUnroll tile loads in flight.  Launched
// only for sorts of at most kScanTiles tiles: larger ones derive the plan in
// the first pass's counter scan from per-group AND/OR.
constexpr int kPlanThreads = 1024;
constexpr u32 kPlanUnroll = 8;

template <typename K>
__global__ __launch_bounds__(kPlanThreads) void satradix_plan(SortPlan* __restrict__ plan,
	const K* __restrict__ andor, u32 tiles, u32 passes, u32 bits, u32 in_place) {
	constexpr int kPW = kPlanThreads / kWave;
	__shared__ K s_red[2 * kPW];
	const u32 tid = threadIdx.x;
	K a = ~K(0), o = K(0);
	for (u32 t0 = 0; t0 < tiles; t0 += kPlanThreads * kPlanUnroll) {
		K xa[kPlanUnroll], xo[kPlanUnroll];
#pragma unroll
		for (u32 j = 0; j < kPlanUnroll; ++j) {
			const u32 t = t0 + j * kPlanThreads + tid;
			xa[j] = t < tiles ? andor[2 * static_cast<size_t>(t)] : ~K(0);
			xo[j] = t < tiles ? andor[2 * static_cast<size_t>(t) + 1] : K(0);
		}
#pragma unroll
		for (u32 j = 0; j < kPlanUnroll; ++j) {
			a &= xa[j];
			o |= xo[j];
		}
	}
#pragma unroll
	for (int off = kWave / 2; off > 0; off >>= 1) {
		a &= __shfl_xor(a, off, kWave);
		o |= __shfl_xor(o, off, kWave);
	}
	if (lane_id() == 0) {
		s_red[wave_id()] = a;
		s_red[kPW + wave_id()] = o;
	}
	__syncthreads();
	if (wave_id() != 0) return;
	a = ~K(0);
	o = K(0);
	for (int ww = 0; ww < kPW; ++ww) {
		a &= s_red[ww];
		o |= s_red[kPW + ww];
	}
	plan_write<K>(plan, a ^ o, passes, bits, in_place);
}

// Counter scan.  Tiles form groups of kScanTiles, handled by blocks of
// kGroupThreads = 4 x 256 threads.
constexpr int kGroupThreads = 4 * kRadixMax;
// Sorts of at most kScanTiles tiles (one group) launch no counter scan: each
// downsweep block sums the tile counts itself, up to 256 KB of reads per
// block, which measured cheaper than a one-block scan launch (2^14..2^20
// u32 keys: 0.078-0.099 -> 0.059-0.094 ms per sort).

// Blocks of the grouped counter kernels split the digits DSPLIT ways
// (grid.y): a block handles kDigits digits of one group of kScanTiles tiles,
// thread (q, d) the kPer consecutive tiles of part q.  Splitting gives the
// 2^28-key sort's 128 groups 512 blocks instead of 128, which fill the 256
// CUs (measured: reduce + scan 0.062 -> 0.032 ms per pass; 2 and 4 tie).
template <u32 DSPLIT>
struct GroupShape {
	static constexpr u32 kDigits = kRadixMax / DSPLIT;
	static constexpr u32 kParts = kGroupThreads / kDigits;
	static constexpr u32 kPer = kScanTiles / kParts;
};
constexpr u32 kDigitSplit = 4;

// Loads this thread's PER count rows (digit d) from tile t0 on.
template <u32 PER>
__device__ __forceinline__ void load_part(const u16* __restrict__ counts, u32 t0, u32 tiles,
	u32 d, u32 (&c)[PER]) {
#pragma unroll
	for (u32 i = 0; i < PER; ++i)
		c[i] = (t0 + i < tiles) ? counts[static_cast<size_t>(t0 + i) * kRadixMax + d] : 0u;
}

// partial[g][d] = sum of counts[t][d] over the tiles of group g.  In the first
// pass (first = 1, before any plan exists) the blocks of the first digit
// quarter also reduce their group's per-tile AND/OR into gandor[2g], [2g+1],
// from which the counter scan derives the plan.
template <typename K, u32 DSPLIT>
__global__ __launch_bounds__(kGroupThreads) void satradix_reduce(const SortPlan* __restrict__ plan,
	u32 pass, const u16* __restrict__ counts, u32 tiles, u32* __restrict__ partial,
	const K* __restrict__ andor, K* __restrict__ gandor, u32 first) {
	using S = GroupShape<DSPLIT>;
	__shared__ u32 s_q[S::kParts][S::kDigits];
	__shared__ K s_red[2 * (kGroupThreads / kWave)];
	if (!first && !plan->pass[pass].active) return;
	const u32 dl = threadIdx.x % S::kDigits, q = threadIdx.x / S::kDigits;
	const u32 d = blockIdx.y * S::kDigits + dl;
	u32 c[S::kPer];
	load_part(counts, blockIdx.x * kScanTiles + q * S::kPer, tiles, d, c);
	if (first && blockIdx.y == 0) {
		// kScanTiles <= kGroupThreads: one tile's AND/OR per thread.
		const u32 t = blockIdx.x * kScanTiles + threadIdx.x;
		K a = ~K(0), o = K(0);
		if (threadIdx.x < kScanTiles && t < tiles) {
			a = andor[2 * static_cast<size_t>(t)];
			o = andor[2 * static_cast<size_t>(t) + 1];
		}
#pragma unroll
		for (int off = kWave / 2; off > 0; off >>= 1) {
			a &= __shfl_xor(a, off, kWave);
			o |= __shfl_xor(o, off, kWave);
		}
		if (lane_id() == 0) {
			s_red[wave_id()] = a;
			s_red[kGroupThreads / kWave + wave_id()] = o;
		}
	}
	u32 s = 0;
#pragma unroll
	for (u32 i = 0; i < S::kPer; ++i) s += c[i];
	s_q[q][dl] = s;
	__syncthreads();
	if (q == 0) {
		u32 tot = 0;
#pragma unroll
		for (u32 qq = 0; qq < S::kParts; ++qq) tot += s_q[qq][dl];
		partial[blockIdx.x * kRadixMax + d] = tot;
	}
	if (first && blockIdx.y == 0 && threadIdx.x == 0) {
		K a = ~K(0), o = K(0);
#pragma unroll
		for (u32 w = 0; w < kGroupThreads / kWave; ++w) {
			a &= s_red[w];
			o |= s_red[kGroupThreads / kWave + w];
		}
		gandor[2 * blockIdx.x] = a;
		gandor[2 * blockIdx.x + 1] = o;
	}
}

// Turns the tile counts into global output offsets:
// offs[t][d] = (#keys with a smaller digit) + (#keys with digit d in tiles < t).
// Each block first derives, from the group sums of all groups, its digits'
// bases: (#keys with a smaller digit) + (#keys with digit d in earlier
// groups); no separate launch scans the group sums.  In the first pass
// (first = 1) every block derives from the group AND/OR whether the pass
// runs, and block (0, 0) writes the plan of the whole sort, which every later
// kernel reads.
template <typename K, u32 DSPLIT>
__global__ __launch_bounds__(kGroupThreads) void satradix_scan(SortPlan* __restrict__ plan,
	u32 pass, const u16* __restrict__ counts, u32* __restrict__ offs, u32 tiles,
	const u32* __restrict__ partial, const K* __restrict__ gandor, u32 first, u32 passes,
	u32 bits, u32 in_place) {
	using S = GroupShape<DSPLIT>;
	constexpr u32 kGQ = kGroupThreads / kRadixMax;  // threads per digit over the group sums
	__shared__ u32 s_q[S::kParts][S::kDigits];
	__shared__ u32 s_tot[kGQ][kRadixMax];
	__shared__ u32 s_pre[kGQ][kRadixMax];
	__shared__ u32 s_base[kRadixMax];
	__shared__ u32 s_wsum[kGroupThreads / kWave];
	const u32 tid = threadIdx.x;
	if (first) {
		// AND/OR over the groups, by wave 0: at most 2^32 / 4096 / 256 groups.
		if (wave_id() == 0) {
			K a = ~K(0), o = K(0);
			for (u32 g = lane_id(); g < gridDim.x; g += kWave) {
				a &= gandor[2 * g];
				o |= gandor[2 * g + 1];
			}
#pragma unroll
			for (int off = kWave / 2; off > 0; off >>= 1) {
				a &= __shfl_xor(a, off, kWave);
				o |= __shfl_xor(o, off, kWave);
			}
			const K diff = a ^ o;
			if (blockIdx.x == 0 && blockIdx.y == 0) plan_write<K>(plan, diff, passes, bits, in_place);
			if (lane_id() == 0) s_wsum[0] = (static_cast<u32>(diff) & ((1u << bits) - 1u)) ? 1u : 0u;
		}
		__syncthreads();
		const u32 active = s_wsum[0];
		__syncthreads();
		if (!active) return;
	} else if (!plan->pass[pass].active) {
		return;
	}
	const u32 dl = tid % S::kDigits, q = tid / S::kDigits;
	const u32 d = blockIdx.y * S::kDigits + dl;
	const u32 t0 = blockIdx.x * kScanTiles + q * S::kPer;
	u32 c[S::kPer];
	load_part(counts, t0, tiles, d, c);
	{
		// Digit totals over all groups, and the sums over groups before this
		// one; thread (gq, dd) takes every kGQ-th group, 8 loads in flight.
		const u32 g = blockIdx.x, groups = gridDim.x;
		const u32 dd = tid % kRadixMax, gq = tid / kRadixMax;
		u32 tot = 0, pre = 0, g2 = gq;
		for (; g2 + 7 * kGQ < groups; g2 += 8 * kGQ) {
			u32 v[8];
#pragma unroll
			for (u32 j = 0; j < 8; ++j) v[j] = partial[(g2 + j * kGQ) * kRadixMax + dd];
#pragma unroll
			for (u32 j = 0; j < 8; ++j) {
				tot += v[j];
				pre += (g2 + j * kGQ < g) ? v[j] : 0u;
			}
		}
		for (; g2 < groups; g2 += kGQ) {
			const u32 v = partial[g2 * kRadixMax + dd];
			tot += v;
			pre += (g2 < g) ? v : 0u;
		}
		s_tot[gq][dd] = tot;
		s_pre[gq][dd] = pre;
	}
	__syncthreads();
	{
		u32 tsum = 0, psum = 0;
		if (tid < kRadixMax) {
#pragma unroll
			for (u32 k = 0; k < kGQ; ++k) {
				tsum += s_tot[k][tid];
				psum += s_pre[k][tid];
			}
		}
		const u32 ex = block_exclusive_sum<kGroupThreads>(tsum, s_wsum,
			static_cast<u32*>(nullptr));
		if (tid < kRadixMax) s_base[tid] = ex + psum;
	}
	u32 s = 0;
#pragma unroll
	for (u32 i = 0; i < S::kPer; ++i) s += c[i];
	s_q[q][dl] = s;
	__syncthreads();
	u32 run = s_base[d];
	for (u32 qq = 0; qq < q; ++qq) run += s_q[qq][dl];
#pragma unroll
	for (u32 i = 0; i < S::kPer; ++i) {
		if (t0 + i < tiles) offs[static_cast<size_t>(t0 + i) * kRadixMax + d] = run;
		run += c[i];
	}
}

// One digit pass over one tile: stable ranking with wave ballots, then a
// scatter staged through LDS so that global writes are contiguous runs.
//
// Element order inside a tile is wave-striped: wave w owns elements
// [w*64*ITEMS, (w+1)*64*ITEMS), item j of lane l is element w*64*ITEMS+j*64+l.
// Processing items j = 0.. in order with per-wave running digit counts gives
// each element the number of equal-digit elements before it (stability).
// FIXED_BITS = 8 compiles the digit-match loop for the default 8-bit digits;
// 0 takes the digit width from `bits` (radix option).
// fused_scan (sorts of at most kScanTiles tiles): every block derives its own
// tile's offsets from the upsweep's tile counts (`counts`), in place of a
// separate counter-scan launch; otherwise `offs` holds the counter scan's
// offsets.
//
// Staging: 8-byte keys pass through LDS as two 32-bit halves (high halves,
// then low halves) in a buffer of 4 bytes per key, so the block needs 36 KB
// of LDS instead of 68 KB.  Full tiles (every tile but a ragged last one) run
// a copy of the body without per-element bounds checks.  Measured on the 2^28
// u64-key/u32-value sort: 1.326 -> 1.298 ms per downsweep.
// The per-tile work lives in radix_tile.hpp (tile_pass), shared with the MSD
// partition levels.  Measured on the 2^28 u64-key/u32-value sort: split
// staging of 8-byte keys and the check-free full-tile body take the
// downsweep from 1.326 to 1.298 ms.
template <typename K, bool KV, int NT, int ITEMS, int FIXED_BITS>
__global__ __launch_bounds__(NT) void satradix_downsweep(Bufs<K> b,
	const SortPlan* __restrict__ plan, u32 pass, u32 n, u32 tiles, u32 bits_arg,
	const u32* __restrict__ offs, const u16* __restrict__ counts, u8* __restrict__ ndig,
	u32 fused_scan) {
	const u32 bits = FIXED_BITS ? static_cast<u32>(FIXED_BITS) : bits_arg;
	constexpr u32 kTile = NT * ITEMS;
	static_assert(NT >= kRadixMax, "one thread per digit bin");
	__shared__ TileSmem<K, KV, NT, ITEMS> sm;

	const PassDesc d = plan->pass[pass];
	if (!d.active) return;  // uniform: the digit is equal in every key
	const u32 tile = xcd_tile(tiles);
	if (tile >= tiles) return;
	const u32 tid = threadIdx.x;
	const u32 lane = lane_id();
	const u32 w = wave_id();
	for (u32 i = lane; i < kRadixMax; i += kWave) sm.whist[w][i] = 0;
	u32 goff = 0;
	if (fused_scan) {
		// Offset of digit d = (#keys with a smaller digit) + (#keys with digit
		// d in earlier tiles), from the counts of all tiles (<= kScanTiles
		// rows of 512 B, read row by row, 16 rows in flight).
		u32 pre = 0, tot = 0;
		if (tid < kRadixMax) {
			u32 t2 = 0;
			for (; t2 + 16 <= tiles; t2 += 16) {
				u32 c[16];
#pragma unroll
				for (u32 j = 0; j < 16; ++j) c[j] = counts[static_cast<size_t>(t2 + j) * kRadixMax + tid];
#pragma unroll
				for (u32 j = 0; j < 16; ++j) {
					tot += c[j];
					pre += (t2 + j < tile) ? c[j] : 0u;
				}
			}
			for (; t2 < tiles; ++t2) {
				const u32 c = counts[static_cast<size_t>(t2) * kRadixMax + tid];
				tot += c;
				pre += (t2 < tile) ? c : 0u;
			}
		}
		const u32 base = block_exclusive_sum<NT>(tid < kRadixMax ? tot : 0u, sm.wsum,
			static_cast<u32*>(nullptr));
		goff = tid < kRadixMax ? base + pre : 0u;
	} else if (tid < kRadixMax) {
		goff = offs[static_cast<size_t>(tile) * kRadixMax + tid];
	}
	TilePass<K> tp;
	tp.src = b.k[d.src];
	tp.vsrc = b.v[d.src];
	tp.dst = b.k[d.dst];
	tp.vdst = b.v[d.dst];
	tp.ndig = d.nact ? ndig : nullptr;
	tp.base = tile * kTile;
	tp.count = (n - tp.base) < kTile ? (n - tp.base) : kTile;
	tp.shift = d.shift;
	tp.nshift = d.nshift;
	if (tp.count == kTile)
		tile_pass<K, KV, NT, ITEMS, FIXED_BITS, true>(sm, tp, bits, goff);
	else
		tile_pass<K, KV, NT, ITEMS, FIXED_BITS, false>(sm, tp, bits, goff);
}

// ---------------------------------------------------------------------------
template <typename K, bool KV>
__global__ __launch_bounds__(256) void satradix_copy(Bufs<K> b, const SortPlan* __restrict__ plan,
	u32 n) {
	if (!plan->copy_active) return;
	const K* __restrict__ ks = b.k[plan->copy_src];
	K* __restrict__ kd = b.k[plan->copy_dst];
	const u32* __restrict__ vs = b.v[plan->copy_src];
	u32* __restrict__ vd = b.v[plan->copy_dst];
	const u32 stride = gridDim.x * blockDim.x;
	for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		kd[i] = ks[i];
		if (KV) vd[i] = vs[i];
	}
}

// ---------------------------------------------------------------------------
// Host side.

// Tile shapes (threads per block x items per thread):
//   0: 256 x 32 (8192 keys)   1: 512 x 16 (8192)   2: 256 x 16 (4096)
//   3: 512 x 32 (16384)
// The "wg" and "ipt" options pick one; otherwise auto_shape() does.
struct TileShape {
	u32 nt, ipt;
};
constexpr TileShape kShapes[] = {{256, 32}, {512, 16}, {256, 16}, {512, 32}};
constexpr int kNumShapes = 4;

struct SatradixData {
	u32 bits = 8;
	int shape = -1;  // index into kShapes; -1 = automatic
	std::string scan_type = "blelloch";
	Scratch aux_keys;
	Scratch aux_vals;
	Scratch counts;  // tiles x 256 digit counts (16-bit: a tile holds at most 16384 keys)
	Scratch offs;    // tiles x 256 output offsets (the counter scan's result)
	Scratch meta;    // plan + per-tile AND/OR + partial sums + per-group AND/OR
	Scratch digits;  // next pass's digit byte per key
	CCLQueue* last_queue = nullptr;
};

inline SatradixData* sdata(CloSort* s) { return static_cast<SatradixData*>(s->data); }

template <typename K, bool KV, int NT, int ITEMS>
bool launch_all(SatradixData* sd, CCLQueue* q, K* kin, u32* vin, K* kout, u32* vout, size_t n,
	GError** err) {
	const u32 keybits = sizeof(K) * 8;
	const u32 bits = sd->bits;
	const u32 passes = (keybits + bits - 1) / bits;
	const bool in_place = (kout == nullptr || kout == kin);
	const u32 tile = NT * ITEMS;
	const u32 tiles = static_cast<u32>(div_ceil(n, tile));
	const u32 nblk = static_cast<u32>(div_ceil(tiles, kScanTiles));
	const size_t plan_bytes = (sizeof(SortPlan) + 255) / 256 * 256;
	const size_t andor_bytes = (2 * static_cast<size_t>(tiles) * sizeof(K) + 255) / 256 * 256;
	CLO_LOG(1, "satradix: numel=%zu key=%zu B%s radix=%u (bits_in_digit=%u) tile=%dx%d tiles=%u "
		"passes<=%u%s", n, sizeof(K), KV ? " +u32 values" : "", 1u << bits, bits, NT, ITEMS, tiles,
		passes, in_place ? " in place" : "");
	if (!sd->aux_keys.ensure(n * sizeof(K), err)) return false;
	if (KV && !sd->aux_vals.ensure(n * sizeof(u32), err)) return false;
	const bool fused = nblk == 1;
	const size_t gandor_bytes = (2 * static_cast<size_t>(nblk) * sizeof(K) + 255) / 256 * 256;
	if (!sd->counts.ensure(static_cast<size_t>(tiles) * kRadixMax * sizeof(u16), err)) return false;
	if (!fused && !sd->offs.ensure(static_cast<size_t>(tiles) * kRadixMax * sizeof(u32), err))
		return false;
	if (!sd->meta.ensure(plan_bytes + andor_bytes + static_cast<size_t>(nblk) * kRadixMax * 4 +
			gandor_bytes, err))
		return false;
	if (passes > 1 && !sd->digits.ensure((n + 15) / 16 * 16, err)) return false;
	u8* digits = passes > 1 ? static_cast<u8*>(sd->digits.ptr) : nullptr;
	char* mb = static_cast<char*>(sd->meta.ptr);
	SortPlan* plan = reinterpret_cast<SortPlan*>(mb);
	K* andor = reinterpret_cast<K*>(mb + plan_bytes);
	u32* partial = reinterpret_cast<u32*>(mb + plan_bytes + andor_bytes);
	K* gandor = reinterpret_cast<K*>(mb + plan_bytes + andor_bytes +
		static_cast<size_t>(nblk) * kRadixMax * 4);
	u16* counts = static_cast<u16*>(sd->counts.ptr);
	u32* offs = fused ? nullptr : static_cast<u32*>(sd->offs.ptr);

	Bufs<K> b;
	b.k[kIn] = kin;
	b.k[kAux] = static_cast<K*>(sd->aux_keys.ptr);
	b.k[kOut] = in_place ? kin : kout;
	b.v[kIn] = vin;
	b.v[kAux] = KV ? static_cast<u32*>(sd->aux_vals.ptr) : nullptr;
	b.v[kOut] = KV ? (in_place ? vin : vout) : nullptr;
	const u32 nn = static_cast<u32>(n);
	hipStream_t st = q->stream;

	const u32 grid = xcd_grid(tiles);
	{
		Launch L(q, CLO_SORT_SATRADIX_KNAME_UPSWEEP);
		satradix_upsweep<K, NT, ITEMS><<<grid, NT, 0, st>>>(kin, nn, tiles, bits, counts, andor);
		if (!L.done(err)) return false;
	}
	if (fused) {
		// One group of tiles: no counter scan to derive the plan in.
		Launch L(q, CLO_SORT_SATRADIX_KNAME_PLAN);
		satradix_plan<K><<<1, kPlanThreads, 0, st>>>(plan, andor, tiles, passes, bits, in_place ? 1u : 0u);
		if (!L.done(err)) return false;
	}
	for (u32 p = 0; p < passes; ++p) {
		if (p > 0) {
			Launch L(q, CLO_SORT_SATRADIX_KNAME_UPSWEEP);
			satradix_upsweep_digits<K, NT * ITEMS><<<grid, kWave, 0, st>>>(b, plan, p, nn, tiles,
				bits, counts, digits);
			if (!L.done(err)) return false;
		}
		const u32 first = p == 0 ? 1u : 0u;
		if (!fused) {
			{
				Launch L(q, CLO_SORT_SATRADIX_KNAME_REDUCE);
				satradix_reduce<K, kDigitSplit><<<dim3(nblk, kDigitSplit), kGroupThreads, 0, st>>>(plan,
					p, counts, tiles, partial, andor, gandor, first);
				if (!L.done(err)) return false;
			}
			Launch L(q, CLO_SORT_SATRADIX_KNAME_SCAN);
			satradix_scan<K, kDigitSplit><<<dim3(nblk, kDigitSplit), kGroupThreads, 0, st>>>(plan, p,
				counts, offs, tiles, partial, gandor, first, passes, bits, in_place ? 1u : 0u);
			if (!L.done(err)) return false;
		}
		{
			Launch L(q, CLO_SORT_SATRADIX_KNAME_DOWNSWEEP);
			if constexpr (NT == 256 && ITEMS == 32) {
				if (bits != 8) {
					satradix_downsweep<K, KV, NT, ITEMS, 0><<<grid, NT, 0, st>>>(b, plan, p, nn,
						tiles, bits, offs, counts, digits, fused ? 1u : 0u);
					if (!L.done(err)) return false;
					continue;
				}
			}
			satradix_downsweep<K, KV, NT, ITEMS, 8><<<grid, NT, 0, st>>>(b, plan, p, nn, tiles,
				bits, offs, counts, digits, fused ? 1u : 0u);
			if (!L.done(err)) return false;
		}
	}
	{
		Launch L(q, CLO_SORT_SATRADIX_KNAME_COPY);
		u32 grid = static_cast<u32>(div_ceil(n, 256 * 8));
		if (grid > 4096) grid = 4096;
		satradix_copy<K, KV><<<grid, 256, 0, st>>>(b, plan, nn);
		if (!L.done(err)) return false;
	}
	sd->last_queue = q;
	return true;
}

// Default tile shape, measured on MI355X.  At 2^28 keys 512x16 is 7% faster
// for u64 keys alone and 12% for u32 keys with values, 256x32 wins or ties
// elsewhere, and 16384-key tiles run one block per CU and lose 20-40%.  Up to
// 2^20 keys the 4096-key tile (256x16) is 10-25% faster: twice the blocks,
// and still one group of tiles, whose offsets the downsweep sums itself.
inline int auto_shape(size_t ksize, bool kv, size_t n) {
	if (n <= (static_cast<size_t>(1) << 20)) return 2;
	if ((ksize == 8 && !kv) || (ksize == 4 && kv)) return 1;
	return 0;
}

template <typename K, bool KV>
bool launch_shape(SatradixData* sd, CCLQueue* q, void* kin, u32* vin, void* kout, u32* vout,
	size_t n, GError** err) {
	int sh = sd->shape >= 0 ? sd->shape : auto_shape(sizeof(K), KV, n);
	if (sd->bits != 8) sh = 0;  // other digit widths: the generic-width kernels
	K* ki = static_cast<K*>(kin);
	K* ko = static_cast<K*>(kout);
	switch (sh) {
	case 1: return launch_all<K, KV, 512, 16>(sd, q, ki, vin, ko, vout, n, err);
	case 2: return launch_all<K, KV, 256, 16>(sd, q, ki, vin, ko, vout, n, err);
	case 3: return launch_all<K, KV, 512, 32>(sd, q, ki, vin, ko, vout, n, err);
	default: return launch_all<K, KV, 256, 32>(sd, q, ki, vin, ko, vout, n, err);
	}
}

template <bool KV>
bool dispatch(SatradixData* sd, CCLQueue* q, size_t ksize, void* kin, u32* vin, void* kout,
	u32* vout, size_t n, GError** err) {
	switch (ksize) {
	case 1: return launch_shape<u8, KV>(sd, q, kin, vin, kout, vout, n, err);
	case 2: return launch_shape<u16, KV>(sd, q, kin, vin, kout, vout, n, err);
	case 4: return launch_shape<u32, KV>(sd, q, kin, vin, kout, vout, n, err);
	case 8: return launch_shape<u64, KV>(sd, q, kin, vin, kout, vout, n, err);
	}
	set_error(err, kQuarkClo, CLO_ERROR_ARGS, "unsupported element size %zu", ksize);
	return false;
}

CCLEvent* sort_common(CloSort* sorter, CCLQueue* cq_exec, CCLQueue* cq_comm, CCLBuffer* kin,
	CCLBuffer* vin, CCLBuffer* kout, CCLBuffer* vout, size_t numel, GError** err) {
	(void)cq_comm;  // no transfers: the data never leaves the exec queue
	SatradixData* sd = sdata(sorter);
	const size_t ks = clo_type_sizeof(sorter->elem_type);
	if (kout && kout->ptr == kin->ptr) kout = nullptr;
	if (vout && vin && vout->ptr == vin->ptr) vout = nullptr;
	if (numel <= 1) {
		// Nothing to order; an out-of-place call still receives the data.
		if (kout && numel == 1) {
			if (hip_failed(hipMemcpyAsync(kout->ptr, kin->ptr, ks, hipMemcpyDeviceToDevice,
					cq_exec->stream), err, "hipMemcpyAsync"))
				return nullptr;
			if (vout && hip_failed(hipMemcpyAsync(vout->ptr, vin->ptr, 4,
					hipMemcpyDeviceToDevice, cq_exec->stream), err, "hipMemcpyAsync"))
				return nullptr;
		}
		return record_event(cq_exec, "satradix_done", err);
	}
	bool ok = vin ? dispatch<true>(sd, cq_exec, ks, kin->ptr, static_cast<u32*>(vin->ptr),
		kout ? kout->ptr : nullptr, vout ? static_cast<u32*>(vout->ptr) : nullptr, numel, err)
		: dispatch<false>(sd, cq_exec, ks, kin->ptr, nullptr, kout ? kout->ptr : nullptr,
		nullptr, numel, err);
	if (!ok) return nullptr;
	return record_event(cq_exec, "satradix_done", err);
}

// ---- vtable functions (clo_sort_satradix.c:342-687) -----------------------

const char* satradix_init(CloSort* sorter, const char* options, GError** err) {
	std::vector<Opt> opts;
	if (!parse_options(options, opts, "satradix sort", err)) return nullptr;
	SatradixData* sd = new SatradixData();
	sorter->data = sd;
	u32 radix = 256;
	long wg = 0, ipt = 0;
	for (auto& o : opts) {
		if (o.key == "wg") {
			wg = atol(o.value.c_str());
			if (wg != 256 && wg != 512) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS, "wg must be 256 or 512.");
				return nullptr;
			}
		} else if (o.key == "ipt") {
			ipt = atol(o.value.c_str());
			if (ipt != 16 && ipt != 32) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS, "ipt must be 16 or 32.");
				return nullptr;
			}
		} else if (o.key == "radix") {
			long r = atol(o.value.c_str());
			if (r <= 0 || clo_ones32(static_cast<unsigned>(r)) != 1) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS, "Radix must be a power of 2.");
				return nullptr;
			}
			if (r < 2 || r > kRadixMax) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS,
					"Radix must be between 2 and %d.", kRadixMax);
				return nullptr;
			}
			radix = static_cast<u32>(r);
		} else if (o.key == "segtile") {
			const long t = atol(o.value.c_str());
			const size_t ks = clo_type_sizeof(sorter->elem_type);
			if (!(t == 8 || t == 16 || (t == 32 && ks == 4))) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS,
					"segtile must be 8, 16 or 32 (8 or 16 for 64-bit keys).");
				return nullptr;
			}
			sorter->seg_tile = static_cast<u32>(t);
		} else if (o.key == "seggrid") {
			const long g = atol(o.value.c_str());
			if (g < 0 || g > 65536) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS, "seggrid must be in 0..65536.");
				return nullptr;
			}
			sorter->seg_grid = static_cast<u32>(g);
		} else if (o.key == "segmsd") {
			const long m = atol(o.value.c_str());
			if (o.value.empty() || (m != 0 && m != 1)) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS, "segmsd must be 0 or 1.");
				return nullptr;
			}
			sorter->seg_msd = static_cast<int32_t>(m);
		} else if (o.key == "segclass") {
			// One digit per length bin from bin 0: the least work class
			// (0 chunk, 1 medium, 2 large); bins past the string keep 0.
			const std::string& v = o.value;
			bool ok = v.size() <= 64;
			u64 cls[2] = {0, 0};
			for (size_t i = 0; ok && i < v.size(); ++i) {
				if (v[i] < '0' || v[i] > '2') {
					ok = false;
					break;
				}
				cls[i / 32] |= static_cast<u64>(v[i] - '0') << (2 * (i % 32));
			}
			if (!ok) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS,
					"segclass takes up to 64 digits 0, 1 or 2 (one per length bin).");
				return nullptr;
			}
			sorter->seg_cls[0] = cls[0];
			sorter->seg_cls[1] = cls[1];
		} else if (o.key.size() >= 4 && strncasecmp(o.key.c_str(), "scan", 4) == 0) {
			if (o.key.size() == 4) sd->scan_type = o.value;
		} else {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS,
				"Invalid option key '%s' for satradix sort.", o.key.c_str());
			return nullptr;
		}
	}
	if (sd->scan_type != "blelloch") {
		set_error(err, kQuarkClo, CLO_ERROR_IMPL_NOT_FOUND,
			"The requested scan implementation, '%s', was not found.", sd->scan_type.c_str());
		return nullptr;
	}
	sd->bits = clo_tzc(static_cast<int>(radix));
	if (wg || ipt) {
		if (!wg) wg = 256;
		if (!ipt) ipt = wg == 512 ? 16 : 32;
		sd->shape = -1;
		for (int i = 0; i < kNumShapes; ++i)
			if (kShapes[i].nt == static_cast<u32>(wg) && kShapes[i].ipt == static_cast<u32>(ipt))
				sd->shape = i;
		if (sd->shape < 0) {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS,
				"Unsupported tile shape wg=%ld ipt=%ld.", wg, ipt);
			return nullptr;
		}
		if (sd->bits != 8 && sd->shape != 0) {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS,
				"Digit widths other than 8 bits run the 256x32 tile shape only.");
			return nullptr;
		}
	}
	return "satradix (gfx950: upsweep + counter scan + downsweep per digit pass)";
}

void satradix_finalize(CloSort* sorter) {
	SatradixData* sd = sdata(sorter);
	if (sd) {
		(void)hipSetDevice(sorter->ctx->dev);
		delete sd;
	}
	sorter->data = nullptr;
}

CCLEvent* satradix_sort_with_device_data(CloSort* sorter, CCLQueue* cq_exec, CCLQueue* cq_comm,
	CCLBuffer* data_in, CCLBuffer* data_out, size_t numel, size_t lws_max, GError** err) {
	(void)lws_max;  // tile shapes are fixed per kernel variant
	return sort_common(sorter, cq_exec, cq_comm, data_in, nullptr, data_out, nullptr, numel,
		err);
}

cl_uint satradix_get_num_kernels(CloSort* sorter, GError** err) {
	(void)sorter;
	(void)err;
	return CLO_SORT_SATRADIX_NUM_KERNELS;
}

const char* satradix_get_kernel_name(CloSort* sorter, cl_uint i, GError** err) {
	(void)sorter;
	static const char* names[] = CLO_SORT_SATRADIX_KERNELNAMES;
	if (i >= CLO_SORT_SATRADIX_NUM_KERNELS) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "kernel index %u out of range", i);
		return nullptr;
	}
	return names[i];
}

size_t satradix_get_localmem_usage(CloSort* sorter, cl_uint i, size_t lws_max, size_t numel,
	GError** err) {
	(void)lws_max;
	(void)numel;
	const size_t ks = clo_type_sizeof(sorter->elem_type);
	const SatradixData* sd = sdata(sorter);
	const int sh = sd && sd->shape >= 0 ? sd->shape : auto_shape(ks, true, numel);
	const size_t nt = kShapes[sh].nt, ipt = kShapes[sh].ipt;
	switch (i) {
	case CLO_SORT_SATRADIX_KIDX_UPSWEEP:
		return (nt / kWave) * 4 * kRadixMax * 4 + 2 * (nt / kWave) * ks;
	case CLO_SORT_SATRADIX_KIDX_PLAN:
		return 2 * (kPlanThreads / kWave) * ks;
	case CLO_SORT_SATRADIX_KIDX_REDUCE:
		return 4 * kRadixMax * 4;
	case CLO_SORT_SATRADIX_KIDX_SCAN:
		return (3 * kGroupThreads + kRadixMax + kGroupThreads / kWave) * 4;
	case CLO_SORT_SATRADIX_KIDX_DOWNSWEEP:
		return nt * ipt * (ks == 8 ? 4 : (ks > 4 ? ks : 4)) + (nt / kWave) * kRadixMax * 4 +
			2 * kRadixMax * 4 + (nt / kWave) * 4;
	case CLO_SORT_SATRADIX_KIDX_COPY:
		return 0;
	}
	set_error(err, kQuarkClo, CLO_ERROR_ARGS, "kernel index %u out of range", i);
	return 0;
}

}  // namespace

CCLEvent* clo_satradix_sort_kv(CloSort* sorter, CCLQueue* cq_exec, CCLQueue* cq_comm,
	CCLBuffer* keys_in, CCLBuffer* vals_in, CCLBuffer* keys_out, CCLBuffer* vals_out,
	size_t numel, GError** err) {
	return sort_common(sorter, cq_exec, cq_comm, keys_in, vals_in, keys_out, vals_out, numel,
		err);
}

bool clo_satradix_sort_pairs(CloSort* sorter, CCLQueue* q, size_t ksize, void* keys,
	uint32_t* vals, void* keys_out, uint32_t* vals_out, size_t n, GError** err) {
	if (n <= 1) {
		if (n == 1 && keys_out &&
			(hip_failed(hipMemcpyAsync(keys_out, keys, ksize, hipMemcpyDeviceToDevice,
				q->stream), err, "hipMemcpyAsync") ||
			 hip_failed(hipMemcpyAsync(vals_out, vals, 4, hipMemcpyDeviceToDevice, q->stream),
				err, "hipMemcpyAsync")))
			return false;
		return true;
	}
	return dispatch<true>(sdata(sorter), q, ksize, keys, vals, keys_out, vals_out, n, err);
}

cl_uint clo_satradix_last_passes(CloSort* sorter, GError** err) {
	SatradixData* sd = sdata(sorter);
	if (!sd || !sd->meta.ptr || !sd->last_queue) return 0;
	u32 executed = 0;
	SortPlan* plan = static_cast<SortPlan*>(sd->meta.ptr);
	if (hip_failed(hipMemcpyAsync(&executed, &plan->executed, sizeof(u32),
			hipMemcpyDeviceToHost, sd->last_queue->stream), err, "hipMemcpyAsync") ||
		hip_failed(hipStreamSynchronize(sd->last_queue->stream), err, "hipStreamSynchronize"))
		return 0;
	return executed;
}

// Host-pass only: the vtable holds host function addresses.
#if !defined(__HIP_DEVICE_COMPILE__)
extern "C" const CloSortImplDef clo_sort_satradix_def = {"satradix", CL_TRUE, satradix_init,
	satradix_finalize, satradix_sort_with_device_data, satradix_get_num_kernels,
	satradix_get_kernel_name, satradix_get_localmem_usage};
#endif


// sbitonic and abitonic (CL_Ops' src/cl_ops/sort/clo_sort_sbitonic.{c,cl} and
// clo_sort_abitonic.{c,cl}) for gfx950.
//
// Both run one bitonic sorting network over N = nlpo2(n) positions; they
// differ only in how the network's steps are grouped into kernel launches.
// The network is the all-ascending form: stage s (blocks of B = 2^s) starts
// with a "flip" step pairing i with i ^ (B-1) inside the block, followed by
// half-cleaner steps pairing i with i ^ 2^(t-1), t = s-1..1; every comparator
// puts the element that sorts first at the lower index.  Positions >= n hold
// virtual elements that sort after everything: a comparator whose upper
// position is >= n does nothing, so real elements never leave [0, n) and the
// caller's buffer needs no padding.  For keys-only sorts the output equals
// the reference's (a sorted permutation is unique up to equal, identical keys).
//
// Order and signedness are folded into an XOR mask applied on load and store:
// signed keys flip the sign bit, descending order complements the key, and the
// kernels then sort unsigned words ascending.

#include <cstdint>
#include <cstring>

#include "clo_internal.hpp"
#include "dev_common.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

constexpr int kBitBlock = 256;

// Tiles of the register-tail kernel: 2^11 .. 2^15 keys (2^14 for 8-byte
// keys: 136-139 KB of LDS with the padding), LDS sized to the tile so that
// small tiles keep many blocks resident per CU.
constexpr u32 kTileLogMin = 11;
constexpr u32 tile_log_cap(size_t key_bytes) { return key_bytes >= 8 ? 14u : 15u; }
// Default tile per key width, measured on MI355X (abitonic 2^24 keys, tile
// sort and local passes in registers): 4-byte keys 2^14 keys with 32 per
// thread (0.750 ms; 2^14 x 16: 0.785, 2^13 x 32: 0.841, 2^15 x 32: 1.246),
// 8-byte keys 2^12 with 16 per thread (1.850 ms; 2^13 x 16: 1.908, 2^13 x 32:
// 2.082, 2^14 x 16: 2.167; 2^15 x 32 for 4-byte keys: 0.758), 2-byte keys
// 2^15 with 32 (0.566 ms; 2^14 x 32: 0.606, 2^13 x 16: 0.707), 1-byte keys
// 2^15 with 32 (0.515 ms; 2^14 x 32: 0.556, 2^13 x 16: 0.670).
template <typename T>
struct TileMax {
	static constexpr u32 kLog = sizeof(T) >= 8 ? 12 : sizeof(T) == 4 ? 14 : 15;
};
// The plain LDS kernels (one synchronised step at a time) take tiles of up
// to 2^11 keys.
constexpr u32 kPlainTileLog = 11;
constexpr u32 kPlainTileMax = 1u << kPlainTileLog;

__device__ __forceinline__ u64 insert_zero_bits(u64 v, u32 pos, u32 q) {
	const u64 low = v & ((1ull << pos) - 1ull);
	return ((v >> pos) << (pos + q)) | low;
}

// Q consecutive steps of stage s starting at step t (t == s: the first is the
// flip step), each thread holding the 2^Q positions one comparator group
// touches.  Q = 1 is one network step (the sbitonic kernel).
// Indices are 32-bit: sizes are capped at 2^31 keys.  FULL (n a power of
// two): no position is virtual, so no bounds checks.
template <typename T, int Q, bool FULL>
__global__ __launch_bounds__(kBitBlock) void bitonic_group(T* __restrict__ data, u32 n, u32 npo2,
	u32 s, u32 t, T xmask) {
	constexpr int V = 1 << Q;
	const u32 tid = blockIdx.x * kBitBlock + threadIdx.x;
	if (tid >= (npo2 >> Q)) return;
	const u32 pos = t - Q;  // lowest free bit
	const u32 x = ((tid >> pos) << t) | (tid & ((1u << pos) - 1u));
	u32 masks[Q];
#pragma unroll
	for (int r = 0; r < Q; ++r) masks[r] = 1u << (t - 1 - r);
	if (t == s) masks[0] = (1u << s) - 1u;
	u32 idx[V];
	T val[V];
#pragma unroll
	for (int j = 0; j < V; ++j) {
		u32 e = x;
#pragma unroll
		for (int r = 0; r < Q; ++r)
			if (j & (1 << r)) e ^= masks[r];
		idx[j] = e;
		val[j] = (FULL || e < n) ? static_cast<T>(data[e] ^ xmask) : T(0);
	}
#pragma unroll
	for (int r = 0; r < Q; ++r) {
#pragma unroll
		for (int j = 0; j < V; ++j) {
			if (j & (1 << r)) continue;
			const int k = j | (1 << r);
			const bool jlo = idx[j] < idx[k];
			if (!FULL && (jlo ? idx[k] : idx[j]) >= n) continue;
			const T a = val[j], b = val[k];
			const T mn = a < b ? a : b, mx = a < b ? b : a;
			val[j] = jlo ? mn : mx;
			val[k] = jlo ? mx : mn;
		}
	}
#pragma unroll
	for (int j = 0; j < V; ++j)
		if (FULL || idx[j] < n) data[idx[j]] = static_cast<T>(val[j] ^ xmask);
}

// One comparator of the step (s, t) inside an LDS tile; tile position p of the
// pair index i.  Returns lo/hi tile positions.
__device__ __forceinline__ void pair_pos(u32 i, u32 s, u32 t, u32* lo, u32* hi) {
	if (t == s) {
		const u32 half = 1u << (s - 1);
		const u32 blk = i >> (s - 1);
		const u32 off = i & (half - 1u);
		*lo = (blk << s) + off;
		*hi = (blk << s) + (2u * half - 1u - off);
	} else {
		const u32 h = 1u << (t - 1);
		*lo = ((i >> (t - 1)) << t) + (i & (h - 1u));
		*hi = *lo + h;
	}
}

// Steps t = min(s, t_top_limit)..1 of stages s_first..s_last on an LDS
// tile whose first `lim` positions are real (comparators reaching past them
// do nothing: those positions are virtual +inf).
template <typename T>
__device__ __forceinline__ void tile_steps(T* s_t, u32 lim, u32 tile_n, u32 s_first, u32 s_last,
	u32 t_top_limit) {
	for (u32 s = s_first; s <= s_last; ++s) {
		const u32 t0 = s < t_top_limit ? s : t_top_limit;
		for (u32 t = t0; t >= 1; --t) {
			for (u32 i = threadIdx.x; i < tile_n / 2; i += kBitBlock) {
				u32 lo, hi;
				pair_pos(i, s, t, &lo, &hi);
				if (hi < lim) {
					const T a = s_t[lo], b = s_t[hi];
					if (b < a) {
						s_t[lo] = b;
						s_t[hi] = a;
					}
				}
			}
			__syncthreads();
		}
	}
}

// Stages s_first..s_last inside each tile of 2^k positions (k <= 11), one
// synchronised LDS step at a time: the tile sort (stages 1..k) or the local
// half-cleaners of one stage s > k.
template <typename T>
__global__ __launch_bounds__(kBitBlock) void bitonic_plain(T* __restrict__ data, u64 n, u32 k,
	u32 s_first, u32 s_last, T xmask) {
	__shared__ T s_t[kPlainTileMax];
	const u32 tile_n = 1u << k;
	const u64 base = static_cast<u64>(blockIdx.x) << k;
	if (base >= n) return;
	const u32 lim = n - base < tile_n ? static_cast<u32>(n - base) : tile_n;
	T* const tile = data + base;
	for (u32 i = threadIdx.x; i < lim; i += kBitBlock) s_t[i] = static_cast<T>(tile[i] ^ xmask);
	__syncthreads();
	tile_steps(s_t, lim, tile_n, s_first, s_last, k);
	for (u32 i = threadIdx.x; i < lim; i += kBitBlock) tile[i] = static_cast<T>(s_t[i] ^ xmask);
}

// ---- tile kernels with register tails --------------------------------------
// Each thread owns R (16 or 32) consecutive tile positions in registers.
// Steps whose comparator distance is below R (t <= log2 R, and the flip steps
// of stages s <= log2 R) run in registers; the other steps go through LDS,
// addressed with a 1-in-32 padding so that a thread's consecutive keys sit in
// distinct banks.

__host__ __device__ constexpr u32 bpad(u32 i) { return i + (i >> 5); }

template <typename T>
__device__ __forceinline__ void cmpswap(T& a, T& b) {
	const T mn = a < b ? a : b;
	const T mx = a < b ? b : a;
	a = mn;
	b = mx;
}

// Half-cleaner of distance H on the thread's keys.  Positions past the
// array hold the largest key (all ones after the order mask): every
// comparator of the all-ascending network puts its minimum at the lower
// position, so those stay above every real key and no comparator needs a
// bounds check.  (Per-comparator checks made the compiler index the key array
// dynamically, which put it in scratch memory.)
template <typename T, int R, int H>
__device__ __forceinline__ void reg_half(T (&x)[R]) {
#pragma unroll
	for (int j = 0; j < R; ++j)
		if (!(j & H)) cmpswap(x[j], x[j | H]);
}

// Flip step of blocks of B positions: j pairs with j ^ (B - 1).
template <typename T, int R, int B>
__device__ __forceinline__ void reg_flip(T (&x)[R]) {
#pragma unroll
	for (int j = 0; j < R; ++j)
		if (!(j & (B / 2))) cmpswap(x[j], x[j ^ (B - 1)]);
}

// Steps t = min(t_hi, TT)..1 of stage s (the step t == s is the flip).
template <typename T, int R, int TT>
__device__ __forceinline__ void reg_tail(T (&x)[R], u32 s, u32 t_hi) {
	if constexpr (TT >= 1) {
		if (t_hi >= TT) {
			if (s == TT) reg_flip<T, R, (1 << TT)>(x);
			else reg_half<T, R, (1 << (TT - 1))>(x);
		}
		reg_tail<T, R, TT - 1>(x, s, t_hi);
	}
}

// Q consecutive steps t, t-1, .., t-Q+1 of stage s inside the LDS tile (each
// thread takes whole comparator groups of 2^Q positions), the same grouping as
// bitonic_group uses for global steps.
template <typename T, int Q, int NT>
__device__ __forceinline__ void lds_group(T* s_t, u64 base, u64 n, u32 tile_n, u32 s, u32 t) {
	constexpr int V = 1 << Q;
	const u32 pos = t - Q;
	u32 masks[Q];
#pragma unroll
	for (int r = 0; r < Q; ++r) masks[r] = 1u << (t - 1 - r);
	if (t == s) masks[0] = (1u << s) - 1u;
	for (u32 g = threadIdx.x; g < (tile_n >> Q); g += NT) {
		const u32 x = static_cast<u32>(insert_zero_bits(g, pos, Q));
		u32 idx[V];
		T val[V];
#pragma unroll
		for (int j = 0; j < V; ++j) {
			u32 e = x;
#pragma unroll
			for (int r = 0; r < Q; ++r)
				if (j & (1 << r)) e ^= masks[r];
			idx[j] = e;
			val[j] = s_t[bpad(e)];
		}
#pragma unroll
		for (int r = 0; r < Q; ++r) {
#pragma unroll
			for (int j = 0; j < V; ++j) {
				if (j & (1 << r)) continue;
				const int k = j | (1 << r);
				const bool jlo = idx[j] < idx[k];
				const u32 hi = jlo ? idx[k] : idx[j];
				if (base + hi >= n) continue;
				const T a = val[j], b = val[k];
				const T mn = a < b ? a : b, mx = a < b ? b : a;
				val[j] = jlo ? mn : mx;
				val[k] = jlo ? mx : mn;
			}
		}
#pragma unroll
		for (int j = 0; j < V; ++j) s_t[bpad(idx[j])] = val[j];
	}
}

// Stages s_first..s_last of the network restricted to one tile of 2^k
// positions (k >= 5): steps t = min(s, k)..1 of each stage.  Serves both the
// tile sort (stages 1..k) and the local half-cleaners of a stage s > k.
// R keys per thread; LDS for 2^LOGT keys (k <= LOGT); 2^LOGT / R threads.
template <typename T, int R, int LOGT>
__global__ __launch_bounds__((1 << LOGT) / R) void bitonic_tile_reg(
	T* __restrict__ data, u64 n, u32 k, u32 s_first, u32 s_last, T xmask) {
	constexpr u32 NT = (1u << LOGT) / R;
	constexpr u32 LR = R == 32 ? 5 : 4;  // log2 R
	static_assert(R == 16 || R == 32, "16 or 32 keys per thread");
	__shared__ T s_t[bpad(1u << LOGT)];
	const u32 tile_n = 1u << k;
	const u64 base = static_cast<u64>(blockIdx.x) << k;
	if (base >= n) return;
	const u32 tid = threadIdx.x;
	const bool owner = tid * R < tile_n;
	const u32 p0 = tid * R;
	const u64 left = n - base;  // > 0
	const u32 valid = !owner ? 0u : (left <= p0 ? 0u : (left - p0 >= R ? R
		: static_cast<u32>(left - p0)));
	// The tile moves between HBM and LDS in lane-contiguous 16-byte words
	// (whole, aligned tiles) or lane-contiguous keys, so every wave access is
	// coalesced; each thread then takes its R consecutive keys from LDS.
	T* const tile = data + base;
	const u32 cnt = left < tile_n ? static_cast<u32>(left) : tile_n;
	constexpr u32 E = 16 / sizeof(T);  // keys per 16-byte word
	const bool vec = cnt == tile_n && (reinterpret_cast<uintptr_t>(tile) & 15) == 0;
	if (vec) {
		for (u32 w = tid; w < tile_n / E; w += NT) {
			const uint4 v = reinterpret_cast<const uint4*>(tile)[w];
			T e[E];
			__builtin_memcpy(e, &v, 16);
#pragma unroll
			for (u32 i = 0; i < E; ++i) s_t[bpad(w * E + i)] = static_cast<T>(e[i] ^ xmask);
		}
	} else {
		for (u32 i = tid; i < cnt; i += NT) s_t[bpad(i)] = static_cast<T>(tile[i] ^ xmask);
	}
	__syncthreads();
	T x[R];
	if (owner) {
#pragma unroll
		for (int j = 0; j < R; ++j)
			x[j] = static_cast<u32>(j) < valid ? s_t[bpad(p0 + j)] : static_cast<T>(~T(0));
	}
	for (u32 s = s_first; s <= s_last; ++s) {
		const u32 t_top = s < k ? s : k;
		if (t_top > LR) {
			__syncthreads();
			if (owner) {
#pragma unroll
				for (int j = 0; j < R; ++j) s_t[bpad(p0 + j)] = x[j];
			}
			__syncthreads();
			for (u32 t = t_top; t > LR;) {
				const u32 m = t - LR;
				if (m >= 3) {
					lds_group<T, 3, NT>(s_t, base, n, tile_n, s, t);
					t -= 3;
				} else if (m == 2) {
					lds_group<T, 2, NT>(s_t, base, n, tile_n, s, t);
					t -= 2;
				} else {
					lds_group<T, 1, NT>(s_t, base, n, tile_n, s, t);
					t -= 1;
				}
				__syncthreads();
			}
			if (owner) {
#pragma unroll
				for (int j = 0; j < R; ++j) x[j] = s_t[bpad(p0 + j)];
			}
		}
		if (owner) reg_tail<T, R, LR>(x, s, t_top < LR ? t_top : LR);
	}
	__syncthreads();  // every thread has read its keys of the last LDS step
	if (owner) {
#pragma unroll
		for (int j = 0; j < R; ++j) s_t[bpad(p0 + j)] = x[j];
	}
	__syncthreads();
	if (vec) {
		for (u32 w = tid; w < tile_n / E; w += NT) {
			T e[E];
#pragma unroll
			for (u32 i = 0; i < E; ++i) e[i] = static_cast<T>(s_t[bpad(w * E + i)] ^ xmask);
			uint4 v;
			__builtin_memcpy(&v, e, 16);
			reinterpret_cast<uint4*>(tile)[w] = v;
		}
	} else {
		for (u32 i = tid; i < cnt; i += NT) tile[i] = static_cast<T>(s_t[bpad(i)] ^ xmask);
	}
}

// ---- local half-cleaner passes in registers ------------------------------
// A local pass of stage s > k (steps t = k..1 on every tile of 2^LOGT keys,
// all half-cleaners: the stage's flip step is a global one) runs every step
// in registers.  A thread holding R = 2^LR keys at stride 2^A (positions
// hi * 2^(A+LR) + m * 2^A + lo, m = 0..R-1) owns every comparator of the
// LR steps with distances 2^A .. 2^(A+LR-1); the steps go in chunks of LR
// from the top, each chunk in its own layout, with one LDS round trip
// between chunks (2 for 2^14 keys at 32 per thread, against 5 for the
// register-tail kernel's grouped LDS steps).

// Padded LDS index of the thread's key m at stride 2^A: the position's three
// parts (hi << (A + LR), m << A, lo < 2^A) have disjoint bits, so bpad splits
// over their sum and the part of m is a compile-time offset.
template <int A, int LR>
__device__ __forceinline__ u32 layout_base(u32 tid) {
	return bpad((tid >> A) << (A + LR)) + bpad(tid & ((1u << A) - 1u));
}
template <int A>
__device__ __forceinline__ constexpr u32 layout_off(int m) {
	return bpad(static_cast<u32>(m) << A);
}

// Half-cleaners of distances 2^EHI .. 2^ELO (relative to the thread's stride).
template <typename T, int R, int EHI, int ELO>
__device__ __forceinline__ void reg_halves(T (&x)[R]) {
	if constexpr (EHI >= ELO) {
		reg_half<T, R, (1 << EHI)>(x);
		reg_halves<T, R, EHI - 1, ELO>(x);
	}
}

// Chunk C (steps THI = LOGT - C*LR down to A + 1) and the chunks after it.
template <typename T, int R, int LR, int LOGT, int C>
__device__ __forceinline__ void half_chunks(T (&x)[R], T* s_t) {
	constexpr int THI = LOGT - C * LR;
	if constexpr (THI >= 1) {
		constexpr int A = THI - LR > 0 ? THI - LR : 0;
		T* const b = s_t + layout_base<A, LR>(threadIdx.x);
#pragma unroll
		for (int m = 0; m < R; ++m) x[m] = b[layout_off<A>(m)];
		reg_halves<T, R, THI - 1 - A, 0>(x);
#pragma unroll
		for (int m = 0; m < R; ++m) b[layout_off<A>(m)] = x[m];
		__syncthreads();
		half_chunks<T, R, LR, LOGT, C + 1>(x, s_t);
	}
}

// ---- tile sort in registers ----------------------------------------------
// Stages 1..LOGT on one tile, in the direction form of the network: at stage
// S every comparator puts the smaller key first where bit S of its lower
// position is clear and last where it is set, so each block of 2^S keys
// comes out ascending or descending by that bit and two neighbours form the
// bitonic input of stage S + 1; bit LOGT of a tile position is always clear,
// so the last stage sorts the tile ascending, as the global stages expect.
// Any sorting network gives the same keys, so the output matches the flip
// form of the network that the global steps use.  Steps go in chunks of up
// to LR in the layouts of bitonic_tile_half, with one LDS round trip between
// chunks (22 for 2^14 keys at 32 per thread, against 27 for the grouped LDS
// steps of bitonic_tile_reg), every index and direction known at compile
// time except a thread-uniform one, which is folded into the keys (x ^ ~0
// reverses the order) while they sit in registers.

// Half-cleaner of distance H register slots whose direction is bit SB of the
// slot index (SB < 0: ascending).
template <typename T, int R, int H, int SB>
__device__ __forceinline__ void reg_half_dir(T (&x)[R]) {
#pragma unroll
	for (int j = 0; j < R; ++j)
		if (!(j & H)) {
			if (SB >= 0 && ((j >> (SB < 0 ? 0 : SB)) & 1)) cmpswap(x[j | H], x[j]);
			else cmpswap(x[j], x[j | H]);
		}
}

template <typename T, int R, int EHI, int SB>
__device__ __forceinline__ void reg_halves_dir(T (&x)[R]) {
	if constexpr (EHI >= 0) {
		reg_half_dir<T, R, (1 << EHI), SB>(x);
		reg_halves_dir<T, R, EHI - 1, SB>(x);
	}
}

// Stage S, steps TT..A+1 at stride 2^A (A = max(0, TT - LR)), then the rest of
// the stage and the stages after it.  The keys are in LDS on entry and on exit.
template <typename T, int R, int LR, int LOGT, int S, int TT>
__device__ __forceinline__ void sort_chunks(T (&x)[R], T* s_t) {
	if constexpr (S <= LOGT) {
		if constexpr (TT < 1) {
			sort_chunks<T, R, LR, LOGT, S + 1, S + 1>(x, s_t);
		} else {
			constexpr int A = TT - LR > 0 ? TT - LR : 0;
			constexpr int SB = S - A < LR ? S - A : -1;
			const u32 tid = threadIdx.x;
			T* const b = s_t + layout_base<A, LR>(tid);
			T dm = T(0);
			if constexpr (SB < 0) dm = (((tid >> A) >> (S - A - LR)) & 1u) ? static_cast<T>(~T(0)) : T(0);
#pragma unroll
			for (int m = 0; m < R; ++m) x[m] = static_cast<T>(b[layout_off<A>(m)] ^ dm);
			reg_halves_dir<T, R, TT - 1 - A, SB>(x);
#pragma unroll
			for (int m = 0; m < R; ++m) b[layout_off<A>(m)] = static_cast<T>(x[m] ^ dm);
			__syncthreads();
			sort_chunks<T, R, LR, LOGT, S, A>(x, s_t);
		}
	}
}

// Stages S..LR in the thread-contiguous layout, without leaving registers.
template <typename T, int R, int LR, int S>
__device__ __forceinline__ void reg_stages(T (&x)[R], T dm_last) {
	if constexpr (S <= LR) {
		if constexpr (S < LR) {
			reg_halves_dir<T, R, S - 1, S>(x);
		} else {
#pragma unroll
			for (int m = 0; m < R; ++m) x[m] ^= dm_last;
			reg_halves_dir<T, R, S - 1, -1>(x);
#pragma unroll
			for (int m = 0; m < R; ++m) x[m] ^= dm_last;
		}
		reg_stages<T, R, LR, S + 1>(x, dm_last);
	}
}

// Waves per SIMD (the second launch bound) of the blocks per CU that the
// kernel's LDS allows, at most 4 (128 VGPRs): without the bound the compiler
// takes up to 256 VGPRs for 32 keys per thread and one 2^14-key block of
// 4-byte keys fills a CU alone.
template <typename T, int R, int LOGT>
constexpr int tile_sort_waves_per_simd() {
	int b = static_cast<int>(163840u / (bpad(1u << LOGT) * sizeof(T)));
	const int waves = (1 << LOGT) / R / 64;
	while (b > 1 && b * waves > 16) --b;
	const int w = (b < 1 ? 1 : b) * waves / 4;
	return w < 1 ? 1 : w;
}

template <typename T, int R, int LOGT>
__global__ __launch_bounds__((1 << LOGT) / R, (tile_sort_waves_per_simd<T, R, LOGT>())) void bitonic_tile_sort(T* __restrict__ data, u64 n,
	T xmask) {
	constexpr u32 NT = (1u << LOGT) / R;
	constexpr int LR = R == 32 ? 5 : 4;
	static_assert(R == 16 || R == 32, "16 or 32 keys per thread");
	static_assert(LOGT >= 2 * LR, "at least two chunks");
	constexpr u32 tile_n = 1u << LOGT;
	__shared__ T s_t[bpad(tile_n)];
	const u64 base = static_cast<u64>(blockIdx.x) << LOGT;
	if (base >= n) return;
	const u32 tid = threadIdx.x;
	T* const tile = data + base;
	const u64 left = n - base;
	const u32 cnt = left < tile_n ? static_cast<u32>(left) : tile_n;
	constexpr u32 E = 16 / sizeof(T);
	const bool vec = cnt == tile_n && (reinterpret_cast<uintptr_t>(tile) & 15) == 0;
	if (vec) {
		for (u32 w = tid; w < tile_n / E; w += NT) {
			const uint4 v = reinterpret_cast<const uint4*>(tile)[w];
			T e[E];
			__builtin_memcpy(e, &v, 16);
#pragma unroll
			for (u32 i = 0; i < E; ++i) s_t[bpad(w * E + i)] = static_cast<T>(e[i] ^ xmask);
		}
	} else {
		for (u32 i = tid; i < tile_n; i += NT)
			s_t[bpad(i)] = i < cnt ? static_cast<T>(tile[i] ^ xmask) : static_cast<T>(~T(0));
	}
	__syncthreads();
	T x[R];
#pragma unroll
	for (int m = 0; m < R; ++m) x[m] = s_t[bpad(tid * R + m)];
	// Stage LR's direction is bit LR of the position, bit 0 of the thread.
	reg_stages<T, R, LR, 1>(x, (tid & 1u) ? static_cast<T>(~T(0)) : T(0));
#pragma unroll
	for (int m = 0; m < R; ++m) s_t[bpad(tid * R + m)] = x[m];
	__syncthreads();
	sort_chunks<T, R, LR, LOGT, LR + 1, LR + 1>(x, s_t);
	if (vec) {
		for (u32 w = tid; w < tile_n / E; w += NT) {
			T e[E];
#pragma unroll
			for (u32 i = 0; i < E; ++i) e[i] = static_cast<T>(s_t[bpad(w * E + i)] ^ xmask);
			uint4 v;
			__builtin_memcpy(&v, e, 16);
			reinterpret_cast<uint4*>(tile)[w] = v;
		}
	} else {
		for (u32 i = tid; i < cnt; i += NT) tile[i] = static_cast<T>(s_t[bpad(i)] ^ xmask);
	}
}

// One local pass on tiles of exactly 2^LOGT keys (the last tile may be
// ragged: positions past the array hold the largest key, as in
// bitonic_tile_reg, and are never stored).
template <typename T, int R, int LOGT>
__global__ __launch_bounds__((1 << LOGT) / R) void bitonic_tile_half(T* __restrict__ data, u64 n,
	T xmask) {
	constexpr u32 NT = (1u << LOGT) / R;
	constexpr int LR = R == 32 ? 5 : 4;
	static_assert(R == 16 || R == 32, "16 or 32 keys per thread");
	static_assert(LOGT >= 2 * LR, "at least two chunks");
	constexpr u32 tile_n = 1u << LOGT;
	__shared__ T s_t[bpad(tile_n)];
	const u64 base = static_cast<u64>(blockIdx.x) << LOGT;
	if (base >= n) return;
	const u32 tid = threadIdx.x;
	T* const tile = data + base;
	const u64 left = n - base;
	const u32 cnt = left < tile_n ? static_cast<u32>(left) : tile_n;
	constexpr u32 E = 16 / sizeof(T);  // keys per 16-byte word
	const bool vec = cnt == tile_n && (reinterpret_cast<uintptr_t>(tile) & 15) == 0;
	if (vec) {
		for (u32 w = tid; w < tile_n / E; w += NT) {
			const uint4 v = reinterpret_cast<const uint4*>(tile)[w];
			T e[E];
			__builtin_memcpy(e, &v, 16);
#pragma unroll
			for (u32 i = 0; i < E; ++i) s_t[bpad(w * E + i)] = static_cast<T>(e[i] ^ xmask);
		}
	} else {
		for (u32 i = tid; i < tile_n; i += NT)
			s_t[bpad(i)] = i < cnt ? static_cast<T>(tile[i] ^ xmask) : static_cast<T>(~T(0));
	}
	__syncthreads();
	T x[R];
	half_chunks<T, R, LR, LOGT, 0>(x, s_t);
	if (vec) {
		for (u32 w = tid; w < tile_n / E; w += NT) {
			T e[E];
#pragma unroll
			for (u32 i = 0; i < E; ++i) e[i] = static_cast<T>(s_t[bpad(w * E + i)] ^ xmask);
			uint4 v;
			__builtin_memcpy(&v, e, 16);
			reinterpret_cast<uint4*>(tile)[w] = v;
		}
	} else {
		for (u32 i = tid; i < cnt; i += NT) tile[i] = static_cast<T>(s_t[bpad(i)] ^ xmask);
	}
}

// Keys per thread held in registers by the tile kernels: 0 runs the plain
// LDS kernel (tiles of 2^11), 16 or 32 the register kernels; the default is
// 16 for 8-byte keys and 32 otherwise (see TileMax).
constexpr int default_regkeys(size_t key_bytes) { return key_bytes <= 4 ? 32 : 16; }
// Global steps fused per launch when the maxps option is not given: 5 (2^5
// keys per thread) for keys of up to 4 bytes, 4 for 8-byte keys (measured on
// 2^24 keys: u32 1.34 -> 1.29 ms with 5; 6 steps, and 5 for u64, were slower).
// The option keeps the reference's range, 1..4.
constexpr u32 kAutoMaxSteps = 5;

struct BitonicData {
	bool advanced;
	u32 minps = 1, maxps = 4, maxsfs = 0xFFFFFFFFu;
	bool maxps_set = false;  // without the option, see kAutoMaxSteps
	int regkeys = -1;  // keys per thread of the tile kernels; -1: the default
	u32 last_sweeps = 0;
};

// log2 of the LDS tile: the maxsfs option (at most 2^15 keys, 2^14 for 8-byte keys) or the
// default of the key width.
inline int regkeys_of(const BitonicData* bd, size_t key_bytes) {
	return bd->regkeys >= 0 ? bd->regkeys : default_regkeys(key_bytes);
}
inline u32 tile_log(const BitonicData* bd, u32 dflt, size_t key_bytes) {
	const u32 cap = regkeys_of(bd, key_bytes) == 0 ? kPlainTileLog : tile_log_cap(key_bytes);
	if (bd->maxsfs == 0xFFFFFFFFu) return dflt < cap ? dflt : cap;
	return bd->maxsfs < cap ? bd->maxsfs : cap;
}

template <typename T>
T make_xmask(const CloSort* s) {
	T m = 0;
	if (type_is_signed(s->elem_type)) m ^= static_cast<T>(T(1) << (sizeof(T) * 8 - 1));
	if (s->descending) m ^= static_cast<T>(~T(0));
	return m;
}

template <typename T, bool FULL>
void group_kernel(u32 grid, hipStream_t st, T* d, u32 n, u32 npo2, u32 s, u32 t, u32 Qn, T xm) {
	switch (Qn) {
	case 1: bitonic_group<T, 1, FULL><<<grid, kBitBlock, 0, st>>>(d, n, npo2, s, t, xm); break;
	case 2: bitonic_group<T, 2, FULL><<<grid, kBitBlock, 0, st>>>(d, n, npo2, s, t, xm); break;
	case 3: bitonic_group<T, 3, FULL><<<grid, kBitBlock, 0, st>>>(d, n, npo2, s, t, xm); break;
	case 4: bitonic_group<T, 4, FULL><<<grid, kBitBlock, 0, st>>>(d, n, npo2, s, t, xm); break;
	case 5: bitonic_group<T, 5, FULL><<<grid, kBitBlock, 0, st>>>(d, n, npo2, s, t, xm); break;
	}
}

template <typename T>
bool launch_group(CCLQueue* q, T* d, u64 n, u64 npo2, u32 s, u32 t, u32 Qn, T xm, const char* nm,
	GError** err) {
	const u64 threads = npo2 >> Qn;
	const u32 grid = static_cast<u32>((threads + kBitBlock - 1) / kBitBlock);
	Launch L(q, nm);
	if (n == npo2)
		group_kernel<T, true>(grid, q->stream, d, static_cast<u32>(n), static_cast<u32>(npo2), s,
			t, Qn, xm);
	else
		group_kernel<T, false>(grid, q->stream, d, static_cast<u32>(n), static_cast<u32>(npo2),
			s, t, Qn, xm);
	return L.done(err);
}

// Stages s_first..s_last inside tiles of 2^k keys: at the default tiles
// (2^12..2^15 keys, 16 or 32 per thread) the tile sort and the local passes
// keep every step in registers (bitonic_tile_sort, bitonic_tile_half); other
// tiles of k >= 5 take the register-tail kernel, smaller ones the plain LDS
// kernels.
template <typename T>
bool launch_tile(CCLQueue* q, const BitonicData* bd, T* d, u64 n, u32 k, u32 s_first,
	u32 s_last, T xm, const char* nm, GError** err) {
	const u32 grid = static_cast<u32>((n + (1ull << k) - 1) >> k);
	const int r = regkeys_of(bd, sizeof(T));
	const u32 lt = k < kTileLogMin ? kTileLogMin : k;  // LDS tile of the instance
	Launch L(q, nm);
	if (s_first == 1 && s_last == k && k == lt && r != 0) {
		// The tile sort: every step in registers.
#define CLO_TILE_SORT(R_, LT_)                                                                   \
	bitonic_tile_sort<T, R_, LT_><<<grid, (1u << LT_) / R_, 0, q->stream>>>(d, n, xm)
		bool done = true;
		if (r == 16 && lt == 12) CLO_TILE_SORT(16, 12);
		else if (r == 16 && lt == 13) CLO_TILE_SORT(16, 13);
		else if (r == 16 && lt == 14) CLO_TILE_SORT(16, 14);
		else if (r == 32 && lt == 13) CLO_TILE_SORT(32, 13);
		else if (r == 32 && lt == 14) CLO_TILE_SORT(32, 14);
		else if constexpr (sizeof(T) < 8) {
			if (r == 32 && lt == 15) CLO_TILE_SORT(32, 15);
			else done = false;
		} else done = false;
#undef CLO_TILE_SORT
		if (done) return L.done(err);
	}
	if (s_first > k && s_first == s_last && k == lt && r != 0) {
		// A local pass (half-cleaners only): every step in registers.
#define CLO_TILE_HALF(R_, LT_)                                                                   \
	bitonic_tile_half<T, R_, LT_><<<grid, (1u << LT_) / R_, 0, q->stream>>>(d, n, xm)
		bool done = true;
		if (r == 16 && lt == 12) CLO_TILE_HALF(16, 12);
		else if (r == 16 && lt == 13) CLO_TILE_HALF(16, 13);
		else if (r == 16 && lt == 14) CLO_TILE_HALF(16, 14);
		else if (r == 32 && lt == 13) CLO_TILE_HALF(32, 13);
		else if (r == 32 && lt == 14) CLO_TILE_HALF(32, 14);
		else if constexpr (sizeof(T) < 8) {
			if (r == 32 && lt == 15) CLO_TILE_HALF(32, 15);
			else done = false;
		} else done = false;
#undef CLO_TILE_HALF
		if (done) return L.done(err);
	}
#define CLO_TILE_REG(R_, LT_)                                                                    \
	bitonic_tile_reg<T, R_, LT_><<<grid, (1u << LT_) / R_, 0, q->stream>>>(d, n, k, s_first,   \
		s_last, xm)
	if (k >= 5 && r != 0) {
		if (r == 16 && lt == 11) CLO_TILE_REG(16, 11);
		else if (r == 16 && lt == 12) CLO_TILE_REG(16, 12);
		else if (r == 16 && lt == 13) CLO_TILE_REG(16, 13);
		else if (r == 16 && lt == 14) CLO_TILE_REG(16, 14);
		else if (lt == 11) CLO_TILE_REG(32, 11);
		else if (lt == 12) CLO_TILE_REG(32, 12);
		else if (lt == 13) CLO_TILE_REG(32, 13);
		else if (lt == 14) CLO_TILE_REG(32, 14);
		else if constexpr (sizeof(T) < 8) CLO_TILE_REG(32, 15);  // 2^15 keys: 1024 x 32
	}
#undef CLO_TILE_REG
	else
		bitonic_plain<T><<<grid, kBitBlock, 0, q->stream>>>(d, n, k, s_first, s_last, xm);
	return L.done(err);
}

template <typename T>
bool run_network(CloSort* sorter, BitonicData* bd, CCLQueue* q, T* d, u64 n, GError** err) {
	const T xm = make_xmask<T>(sorter);
	u64 npo2 = 1;
	u32 S = 0;
	while (npo2 < n) {
		npo2 <<= 1;
		++S;
	}
	u32 sweeps = 0;
	if (!bd->advanced) {
		// sbitonic: one launch per (stage, step), clo_sort_sbitonic.c:102-118.
		for (u32 s = 1; s <= S; ++s)
			for (u32 t = s; t >= 1; --t) {
				if (!launch_group<T>(q, d, n, npo2, s, t, 1, xm, CLO_SORT_SBITONIC_KNAME, err))
					return false;
				++sweeps;
			}
		bd->last_sweeps = sweeps;
		CLO_LOG(1, "sbitonic: numel=%zu (network %u = 2^%u) launches=%u (one per stage and step)",
			static_cast<size_t>(n), static_cast<unsigned>(npo2), S, sweeps);
		return true;
	}
	// abitonic strategy: stages 1..k in one LDS tile kernel, then per stage the
	// global steps in register groups of up to maxps steps (at least minps when
	// that many remain) and the last k steps in one LDS kernel.
	u32 k = tile_log(bd, TileMax<T>::kLog, sizeof(T));
	if (k > S) k = S;
	if (k >= 1) {
		if (!launch_tile<T>(q, bd, d, n, k, 1, k, xm, CLO_SORT_ABITONIC_KNAME_TILESORT, err))
			return false;
		++sweeps;
	}
	static const char* gname[kAutoMaxSteps + 1] = {"", CLO_SORT_ABITONIC_KNAME_ANY,
		CLO_SORT_ABITONIC_KNAME_PRIV_2S4V, CLO_SORT_ABITONIC_KNAME_PRIV_3S8V,
		CLO_SORT_ABITONIC_KNAME_PRIV_4S16V, CLO_SORT_ABITONIC_KNAME_PRIV_5S32V};
	const u32 qmax = bd->maxps_set ? bd->maxps : (sizeof(T) <= 4 ? kAutoMaxSteps : 4u);
	for (u32 s = k + 1; s <= S; ++s) {
		u32 t = s;
		const u32 t_stop = k;  // steps t > t_stop are global
		while (t > t_stop) {
			u32 rem = t - t_stop;
			u32 Qn = rem < qmax ? rem : qmax;
			if (Qn < bd->minps && rem >= bd->minps) Qn = bd->minps;
			if (Qn > t) Qn = t;
			if (!launch_group<T>(q, d, n, npo2, s, t, Qn, xm, gname[Qn], err)) return false;
			++sweeps;
			t -= Qn;
		}
		if (k >= 1) {
			if (!launch_tile<T>(q, bd, d, n, k, s, s, xm, CLO_SORT_ABITONIC_KNAME_LOCAL, err))
				return false;
			++sweeps;
		}
	}
	bd->last_sweeps = sweeps;
	CLO_LOG(1, "abitonic: numel=%zu (network 2^%u) LDS tile 2^%u, register groups of <= %u steps, "
		"launches=%u", static_cast<size_t>(n), S, k, qmax, sweeps);
	return true;
}

CCLEvent* bitonic_sort(CloSort* sorter, CCLQueue* cq_exec, CCLQueue* cq_comm, CCLBuffer* in,
	CCLBuffer* out, size_t numel, size_t lws_max, GError** err) {
	(void)lws_max;
	(void)cq_comm;
	BitonicData* bd = static_cast<BitonicData*>(sorter->data);
	const size_t es = clo_type_sizeof(sorter->elem_type);
	void* d = in->ptr;
	if (out && out->ptr != in->ptr) {
		// Sort a copy in data_out; data_in stays unchanged.
		if (numel && hip_failed(hipMemcpyAsync(out->ptr, in->ptr, numel * es,
				hipMemcpyDeviceToDevice, cq_exec->stream), err, "hipMemcpyAsync"))
			return nullptr;
		d = out->ptr;
	}
	bool ok = true;
	bd->last_sweeps = 0;
	if (numel > 1) {
		switch (es) {
		case 1: ok = run_network<u8>(sorter, bd, cq_exec, static_cast<u8*>(d), numel, err); break;
		case 2: ok = run_network<u16>(sorter, bd, cq_exec, static_cast<u16*>(d), numel, err); break;
		case 4: ok = run_network<u32>(sorter, bd, cq_exec, static_cast<u32*>(d), numel, err); break;
		case 8: ok = run_network<u64>(sorter, bd, cq_exec, static_cast<u64*>(d), numel, err); break;
		default: ok = false;
		}
	}
	if (!ok) return nullptr;
	return record_event(cq_exec, bd->advanced ? "abitonic_done" : "sbitonic_done", err);
}

const char* sbitonic_init(CloSort* sorter, const char* options, GError** err) {
	(void)options;  // clo_sort_sbitonic.c ignores options
	(void)err;
	BitonicData* bd = new BitonicData();
	bd->advanced = false;
	sorter->data = bd;
	return "sbitonic (gfx950: one launch per network step)";
}

const char* abitonic_init(CloSort* sorter, const char* options, GError** err) {
	// clo_sort_abitonic.c:458-567.
	BitonicData* bd = new BitonicData();
	bd->advanced = true;
	sorter->data = bd;
	std::vector<Opt> opts;
	if (!parse_options(options, opts, "abitonic sort", err)) return nullptr;
	for (auto& o : opts) {
		const int v = atoi(o.value.c_str());
		if (o.key == "minps") {
			if (v < 1 || v > 4) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS, "Option 'minps' must be between 1 and 4.");
				return nullptr;
			}
			bd->minps = static_cast<u32>(v);
		} else if (o.key == "maxps") {
			if (v < 1 || v > 4) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS, "Option 'maxps' must be between 1 and 4.");
				return nullptr;
			}
			bd->maxps = static_cast<u32>(v);
			bd->maxps_set = true;
		} else if (o.key == "maxsfs") {
			bd->maxsfs = static_cast<u32>(v);
		} else if (o.key == "regkeys") {
			if (v != 0 && v != 16 && v != 32) {
				set_error(err, kQuarkClo, CLO_ERROR_ARGS, "Option 'regkeys' must be 0, 16 or 32.");
				return nullptr;
			}
			bd->regkeys = static_cast<int>(v);
		} else {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS, "Invalid option key '%s' for abitonic sort.",
				o.key.c_str());
			return nullptr;
		}
	}
	if (bd->maxps < bd->minps) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"'minps' (%u) must be less or equal than 'maxps' (%u).", bd->minps, bd->maxps);
		return nullptr;
	}
	return "abitonic (gfx950: LDS tiles + register-fused global steps)";
}

void bitonic_finalize(CloSort* sorter) {
	delete static_cast<BitonicData*>(sorter->data);
	sorter->data = nullptr;
}

cl_uint sbitonic_num_kernels(CloSort*, GError**) { return 1; }
const char* sbitonic_kernel_name(CloSort*, cl_uint i, GError** err) {
	if (i != 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "kernel index %u out of range", i);
		return nullptr;
	}
	return CLO_SORT_SBITONIC_KNAME;
}
size_t sbitonic_localmem(CloSort*, cl_uint, size_t, size_t, GError**) { return 0; }

cl_uint abitonic_num_kernels(CloSort*, GError**) { return CLO_SORT_ABITONIC_NUM_KERNELS; }
const char* abitonic_kernel_name(CloSort*, cl_uint i, GError** err) {
	static const char* names[] = CLO_SORT_ABITONIC_KERNELNAMES;
	if (i >= CLO_SORT_ABITONIC_NUM_KERNELS) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "kernel index %u out of range", i);
		return nullptr;
	}
	return names[i];
}
size_t abitonic_localmem(CloSort* sorter, cl_uint i, size_t, size_t, GError** err) {
	if (i >= CLO_SORT_ABITONIC_NUM_KERNELS) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "kernel index %u out of range", i);
		return 0;
	}
	const size_t es = clo_type_sizeof(sorter->elem_type);
	if (i < 4 || i > 5) return 0;  // register-only kernels
	const BitonicData* bd = static_cast<const BitonicData*>(sorter->data);
	const u32 k = tile_log(bd, es >= 8 ? TileMax<u64>::kLog : es == 4 ? TileMax<u32>::kLog
		: TileMax<u16>::kLog, es);
	if (k < 5 || regkeys_of(bd, es) == 0) return kPlainTileMax * es;
	return bpad(1u << (k < kTileLogMin ? kTileLogMin : k)) * es;
}

}  // namespace

extern "C" {

// Host-pass only: the vtable holds host function addresses.
#if !defined(__HIP_DEVICE_COMPILE__)
const CloSortImplDef clo_sort_sbitonic_def = {"sbitonic", CL_TRUE, sbitonic_init,
	bitonic_finalize, bitonic_sort, sbitonic_num_kernels, sbitonic_kernel_name, sbitonic_localmem};
#endif


// Host-pass only: the vtable holds host function addresses.
#if !defined(__HIP_DEVICE_COMPILE__)
const CloSortImplDef clo_sort_abitonic_def = {"abitonic", CL_TRUE, abitonic_init,
	bitonic_finalize, bitonic_sort, abitonic_num_kernels, abitonic_kernel_name, abitonic_localmem};
#endif


cl_uint clo_sort_abitonic_get_last_sweeps(CloSort* sorter) {
	if (!sorter || !sorter->data) return 0;
	if (strcmp(sorter->impl.name, "abitonic") != 0 && strcmp(sorter->impl.name, "sbitonic") != 0)
		return 0;
	return static_cast<BitonicData*>(sorter->data)->last_sweeps;
}

}  // extern "C"

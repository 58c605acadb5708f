// Batched segmented sort (an addition with no counterpart in CL_Ops): every
// segment [offsets[j], offsets[j+1]) of one array is stable-sorted by key in
// one persistent launch.
//
// Work items, taken largest first from one atomic queue:
//   large  (len > CAP)      one block, LSD digit passes through global memory
//                           (chunks of CAP staged in LDS, coalesced runs)
//   medium (W < len <= CAP) one block, all digit passes inside LDS
//   chunk  (len <= W)       every small segment starting in the element window
//                           [c*W, (c+1)*W) is packed into one LDS tile and
//                           sorted by (segment, key): key passes, then passes
//                           over the local segment index
// with CAP = 256 * IPT elements and W = CAP / 2 (so a chunk never exceeds
// CAP).  Digit passes whose digit is equal over the whole item are skipped.
//
// The work class of each segment can be raised per length: a table of 64
// length bins (four per octave) gives the least class of every bin, and a
// segment takes the greater of that and the least class that holds it.  An
// all-zero table is the rule above.  A small segment raised out of the chunk
// class stays out of its window's chunk (the chunk gathers only its own
// segments), so every segment is sorted by exactly one item.

#include <cstring>

#include "clo_internal.hpp"
#include "radix_block.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

// Large segments are listed in index order, and again by length class c:
// lengths in (CAP * 2^c, CAP * 2^(c+1)], the last class everything longer;
// class 0 also takes the segments of at most CAP keys that a work-class table
// sends to the large path.  The bucketing kernel takes the classes longest
// first, so the work left when its queue runs dry is a few short segments,
// not a long one that started last (32768 segments of log-uniform lengths up
// to 65536, 32-bit keys: 1.54 -> 1.50 ms).  The digit-pass kernel keeps index
// order: its passes over a large segment go through global memory, and all
// the longest segments at once hold more of it than the caches do (key-value
// sorts of the same workload 4.27 -> 5.27 ms longest first).
constexpr u32 kSegLenClasses = 8;

struct SegMeta {
	u32 n_medium;
	u32 n_large;
	u32 n_chunks;
	u32 head;      // work queue of the digit-pass kernel
	u32 head_msd;  // work queue of the keys-only bucketing kernel
	u32 n_fb;      // large keys-only segments left to the digit passes
	u32 n_cls[kSegLenClasses];  // large segments per length class
};

// Where each length class's list starts in the large list: class c holds at
// most numel / (CAP * 2^c) segments (each is longer than CAP * 2^c) and
// class 0 at most nseg, so the lists take at most nseg + 2 * numel / CAP + 7
// entries.  Returns the total.
struct SegClassBase {
	u32 b[kSegLenClasses];
};
inline size_t seg_class_layout(size_t nseg, size_t numel, size_t cap, SegClassBase* base) {
	size_t t = 0;
	for (u32 c = 0; c < kSegLenClasses; ++c) {
		base->b[c] = static_cast<u32>(t);
		const size_t m = numel / (cap << c) + 1;
		t += c == 0 ? nseg : (m < nseg ? m : nseg);
	}
	return t;
}

__device__ __forceinline__ u32 seg_len_class(u32 len, u32 CAP) {
	const u32 q = (len - 1) / CAP;  // len >= 1
	if (q == 0) return 0;
	const u32 c = 31u - static_cast<u32>(__builtin_clz(q));
	return c < kSegLenClasses - 1 ? c : kSegLenClasses - 1;
}

__global__ void seg_reset(SegMeta* m) {
	const u32 t = threadIdx.x;
	if (t == 0) {
		m->n_medium = 0;
		m->n_large = 0;
		m->n_chunks = 0;
		m->head = 0;
		m->head_msd = 0;
		m->n_fb = 0;
	}
	if (t < kSegLenClasses) m->n_cls[t] = 0;
}

// Segment j = [off[j], off[j+1]) (the last one ends at numel), clamped into
// [0, numel]: malformed offsets (decreasing, or past numel) give a wrong
// order but never an access outside the arrays.
struct SegRange {
	u32 start, len;
};
__device__ __forceinline__ SegRange seg_range(const u32* off, u32 j, u32 nseg, u32 numel) {
	u32 end = j + 1 < nseg ? off[j + 1] : numel;
	if (end > numel) end = numel;
	u32 start = off[j];
	if (start > end) start = end;
	return {start, end - start};
}

// Work class (0 chunk, 1 medium, 2 large): the table's entry for the length
// bin (two bits per bin, bins 0..31 in lo and 32..63 in hi), raised to the
// least class that can hold the segment.
__device__ __forceinline__ u32 seg_class(u64 lo, u64 hi, u32 len, u32 W, u32 CAP) {
	const u32 b = seg_len_bin(len);
	u32 c = static_cast<u32>(((b < 32 ? lo : hi) >> (2 * (b & 31))) & 3u);
	if (c > 2) c = 2;
	if (len > W && c < 1) c = 1;
	if (len > CAP) c = 2;
	if (len == 0) c = 0;
	return c;
}

// Appends the medium and large segments' indices and the windows of W
// elements in which small segments start (chunk items), one atomic per wave
// and list; wrange[window] receives the range [j0, j1) of the segments
// starting in the window.  Small segments starting in one window are
// consecutive (a longer segment pushes its successor's start past the
// window), so the first of them records j0 and the last j1; windows without
// a small segment never become work items.
__global__ __launch_bounds__(256) void seg_classify(const u32* __restrict__ off, u32 nseg,
	u32 numel, u32 W, u32 CAP, u64 cls_lo, u64 cls_hi, u32* __restrict__ medium,
	u32* __restrict__ large_ix, u32* __restrict__ large, SegClassBase lbase,
	u32* __restrict__ chunks,
	uint2* __restrict__ wrange, SegMeta* m) {
	const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
	SegRange sr = {0u, 0u};
	if (j < nseg) sr = seg_range(off, j, nseg, numel);
	const u32 cls = j < nseg ? seg_class(cls_lo, cls_hi, sr.len, W, CAP) : 0u;
	const bool is_med = j < nseg && cls == 1;
	const bool is_lrg = j < nseg && cls == 2;
	// Windows are built over every segment of at most W elements, whatever
	// its class; the chunk leaves out those of another class.
	const bool is_sml = j < nseg && sr.len <= W;
	bool head = false;
	if (is_sml) {
		const u32 win = sr.start / W;
		bool first = j == 0, last = j + 1 == nseg;
		if (!first) {
			const SegRange pr = seg_range(off, j - 1, nseg, numel);
			first = pr.len > W || pr.start / W != win;
		}
		if (!last) {
			const SegRange nx = seg_range(off, j + 1, nseg, numel);
			last = nx.len > W || nx.start / W != win;
		}
		if (first) wrange[win].x = j;
		if (last) wrange[win].y = j + 1;
		head = first;
	}
	// One atomic per list and block (the counters share one cache line that
	// every block updates): per-wave counts, a block total and base per list,
	// then each wave's share.  Lists: medium, large (index order), chunks,
	// and large by length class.
	constexpr u32 kLists = 3 + kSegLenClasses;
	constexpr u32 kWaves = 256 / kWave;
	__shared__ u32 s_cnt[kLists][kWaves];
	const u32 w = wave_id();
	const u32 lc = is_lrg ? seg_len_class(sr.len, CAP) : 0u;
	u64 bal[kLists];
	bal[0] = ballot(is_med);
	bal[1] = ballot(is_lrg);
	bal[2] = ballot(head);
#pragma unroll
	for (u32 c = 0; c < kSegLenClasses; ++c) bal[3 + c] = ballot(is_lrg && lc == c);
	if (lane_id() == 0) {
#pragma unroll
		for (u32 l = 0; l < kLists; ++l) s_cnt[l][w] = static_cast<u32>(__popcll(bal[l]));
	}
	__syncthreads();
	if (threadIdx.x < kLists) {
		const u32 l = threadIdx.x;
		u32 tot = 0;
#pragma unroll
		for (u32 ww = 0; ww < kWaves; ++ww) tot += s_cnt[l][ww];
		u32* ctr = l == 0 ? &m->n_medium : l == 1 ? &m->n_large : l == 2 ? &m->n_chunks
			: &m->n_cls[l - 3];
		u32 run = tot ? atomicAdd(ctr, tot) : 0u;
#pragma unroll
		for (u32 ww = 0; ww < kWaves; ++ww) {
			const u32 c = s_cnt[l][ww];
			s_cnt[l][ww] = run;
			run += c;
		}
	}
	__syncthreads();
	if (is_med) medium[s_cnt[0][w] + count_below(bal[0])] = j;
	if (is_lrg) large_ix[s_cnt[1][w] + count_below(bal[1])] = j;
	if (head) chunks[s_cnt[2][w] + count_below(bal[2])] = sr.start / W;
	if (is_lrg) {
		u32 pos = 0;
#pragma unroll
		for (u32 c = 0; c < kSegLenClasses; ++c)
			if (lc == c) pos = lbase.b[c] + s_cnt[3 + c][w] + count_below(bal[3 + c]);
		large[pos] = j;
	}
}

template <typename K>
struct SegArgs {
	const K* kin;
	const u32* vin;
	K* kout;   // final destination (== kin for in-place)
	u32* vout;
	const u32* kin32;  // composite sorts (K = u64): the 32-bit keys in and out
	u32* kout32;
	K* kscr;   // numel-sized scratch for large segments
	u32* vscr;
	const u32* off;
	const u32* medium;
	const u32* large_ix;  // large segments in index order
	const u32* large;     // and by length class (SegClassBase)
	SegClassBase lbase;
	const u32* chunks;    // windows holding small segment starts
	const uint2* wrange;  // per window: segments [x, y) starting in it
	uint2* fb;            // keys alone: (start, len) of large segments for the digit passes
	u32 fb_mode;          // digit-pass kernel after the bucketing kernel (keys alone)
	SegMeta* meta;
	u64 cls_lo, cls_hi;   // per-length-bin least work class (seg_class)
	u32 nseg;
	u32 numel;
};

// The large segments in work-queue order: the length classes longest first.
// The class sizes are read once per block (the counters share a cache line
// with the work-queue heads that every block updates).
struct LargeOrder {
	u32 n[kSegLenClasses];
};
__device__ __forceinline__ LargeOrder large_order(const SegMeta* m) {
	LargeOrder o;
#pragma unroll
	for (u32 c = 0; c < kSegLenClasses; ++c)
		o.n[c] = __builtin_amdgcn_readfirstlane(m->n_cls[c]);
	return o;
}
// Large segment number t (below meta->n_large) of the work queue.
template <typename K>
__device__ __forceinline__ u32 large_seg(const SegArgs<K>& a, const LargeOrder& o, u32 t) {
#pragma unroll
	for (int c = kSegLenClasses - 1; c > 0; --c) {
		if (t < o.n[c]) return a.large[a.lbase.b[c] + t];
		t -= o.n[c];
	}
	return a.large[a.lbase.b[0] + t];
}

template <typename K, bool KV, int IPT>
struct SegSmem {
	static constexpr u32 kCap = kRB * IPT;
	static constexpr u32 kW = kCap / 2;
	K k[kCap];
	u32 v[KV ? kCap : 1];
	u16 sid[kCap];     // chunk items: each key's segment in the window,
	u32 lpos[kW + 1];  // each segment's start in the tile
	u32 segj[kW];      // and in the array
	u32 hist[sizeof(K)][256];
	u32 base[256];
	u32 gofs[256];
	u32 red[2 * kRBWaves];
	u32 misc[4];
	RankSmem rs;
};

template <typename K>
__device__ __forceinline__ u32 digit_of(K k, u32 shift) {
	return static_cast<u32>(k >> shift) & 0xFFu;
}

// Copies with the loads of up to kBatch rounds issued before their stores (a
// plain load-store loop would wait on each load in turn), kBatch registers
// per thread.
constexpr int kBatch = 8;

// Global -> LDS copy of n <= IPT * kRB elements.
template <typename T, int IPT>
__device__ __forceinline__ void g2s(T* dst, const T* __restrict__ src, u32 n) {
	constexpr int B = IPT < kBatch ? IPT : kBatch;
#pragma unroll
	for (int r0 = 0; r0 < IPT; r0 += B) {
		T t[B];
#pragma unroll
		for (int r = 0; r < B; ++r) {
			const u32 i = (r0 + r) * kRB + threadIdx.x;
			t[r] = i < n ? src[i] : T(0);
		}
#pragma unroll
		for (int r = 0; r < B; ++r) {
			const u32 i = (r0 + r) * kRB + threadIdx.x;
			if (i < n) dst[i] = t[r];
		}
	}
}

// LDS -> global copy of n <= IPT * kRB elements.
template <typename T, int IPT>
__device__ __forceinline__ void s2g(T* __restrict__ dst, const T* src, u32 n) {
#pragma unroll
	for (int r = 0; r < IPT; ++r) {
		const u32 i = r * kRB + threadIdx.x;
		if (i < n) dst[i] = src[i];
	}
}

// Global -> global copy of any length.
template <typename T, int IPT>
__device__ __forceinline__ void g2g(T* __restrict__ dst, const T* __restrict__ src, u32 n) {
	constexpr u32 R = kBatch * kRB;
	for (u32 c0 = 0; c0 < n; c0 += R) {
		T t[kBatch];
#pragma unroll
		for (int r = 0; r < kBatch; ++r) {
			const u32 i = c0 + r * kRB + threadIdx.x;
			t[r] = i < n ? src[i] : T(0);
		}
#pragma unroll
		for (int r = 0; r < kBatch; ++r) {
			const u32 i = c0 + r * kRB + threadIdx.x;
			if (i < n) dst[i] = t[r];
		}
	}
}

// Key-value sorts of 32-bit keys, out of place, on the keys-alone bucketing
// kernel (COMP): it sorts 64-bit composites, key << 32 | position in the
// segment.  Composites are unique and ordered as the stable order of the
// pairs, so the unstable bucketing gives the stable result; the output takes
// each composite's key and the value at its position.
template <typename K, bool COMP>
__device__ __forceinline__ K seg_key(const SegArgs<K>& a, u32 g, u32 pos) {
	if constexpr (COMP) return (static_cast<K>(a.kin32[g]) << 32) | pos;
	else return a.kin[g];
}

// Output element g (of the segment starting at seg0) from its sorted key.
template <typename K, bool COMP>
__device__ __forceinline__ void seg_put(const SegArgs<K>& a, u32 g, u32 seg0, K c) {
	if constexpr (COMP) {
		a.kout32[g] = static_cast<u32>(c >> 32);
		a.vout[g] = a.vin[seg0 + static_cast<u32>(c)];
	} else {
		a.kout[g] = c;
	}
}

// Segment [start, start + len) to the output unchanged (keys, and in COMP
// the values).
template <typename K, int IPT, bool COMP>
__device__ __forceinline__ void seg_copy(const SegArgs<K>& a, u32 start, u32 len) {
	if constexpr (COMP) {
		g2g<u32, IPT>(a.kout32 + start, a.kin32 + start, len);
		g2g<u32, IPT>(a.vout + start, a.vin + start, len);
	} else {
		if (a.kout != a.kin) g2g<K, IPT>(a.kout + start, a.kin + start, len);
	}
}

// The n <= IPT * kRB keys of segment elements [g0, g0 + n) (positions
// pos0.. in their segment) into LDS.
template <typename K, int IPT, bool COMP>
__device__ __forceinline__ void g2s_in(const SegArgs<K>& a, K* dst, u32 g0, u32 pos0, u32 n) {
	if constexpr (COMP) {
		constexpr int B = IPT < kBatch ? IPT : kBatch;
#pragma unroll
		for (int r0 = 0; r0 < IPT; r0 += B) {
			K t[B];
#pragma unroll
			for (int r = 0; r < B; ++r) {
				const u32 i = (r0 + r) * kRB + threadIdx.x;
				t[r] = i < n ? seg_key<K, COMP>(a, g0 + i, pos0 + i) : K(0);
			}
#pragma unroll
			for (int r = 0; r < B; ++r) {
				const u32 i = (r0 + r) * kRB + threadIdx.x;
				if (i < n) dst[i] = t[r];
			}
		}
	} else {
		g2s<K, IPT>(dst, a.kin + g0, n);
	}
}

// Sorted keys src[0, n) to output elements [g0, g0 + n) of the segment
// starting at seg0.
template <typename K, int IPT, bool COMP>
__device__ __forceinline__ void s2g_out(const SegArgs<K>& a, u32 g0, u32 seg0, const K* src, u32 n) {
	if constexpr (COMP) {
#pragma unroll
		for (int r = 0; r < IPT; ++r) {
			const u32 i = r * kRB + threadIdx.x;
			if (i < n) seg_put<K, COMP>(a, g0 + i, seg0, src[i]);
		}
	} else {
		s2g<K, IPT>(a.kout + g0, src, n);
	}
}

// Block-wide AND and OR of the keys in s.k[0, n).
template <typename K, bool KV, int IPT>
__device__ void and_or_lds(SegSmem<K, KV, IPT>& s, u32 n, K* a_out, K* o_out) {
	K a = ~K(0), o = K(0);
	for (u32 i = threadIdx.x; i < n; i += kRB) {
		a &= s.k[i];
		o |= s.k[i];
	}
#pragma unroll
	for (int off = kWave / 2; off > 0; off >>= 1) {
		a &= __shfl_xor(a, off, kWave);
		o |= __shfl_xor(o, off, kWave);
	}
	K* red = reinterpret_cast<K*>(s.hist);  // scratch: 2 * waves values
	if (lane_id() == 0) {
		red[wave_id()] = a;
		red[kRBWaves + wave_id()] = o;
	}
	__syncthreads();
	a = ~K(0);
	o = K(0);
#pragma unroll
	for (int w = 0; w < kRBWaves; ++w) {
		a &= red[w];
		o |= red[kRBWaves + w];
	}
	__syncthreads();
	*a_out = a;
	*o_out = o;
}

// One stable LDS pass over s.k/s.v/s.sid[0, n): digit from the key (shift)
// or, when by_sid, from the local segment index.
template <typename K, bool KV, int IPT>
__device__ void lds_pass(SegSmem<K, KV, IPT>& s, u32 n, u32 shift, bool by_sid, bool carry_sid) {
	const u32 lane = lane_id();
	const u32 w = wave_id();
	K k[IPT];
	u32 v[IPT];
	u32 sd[IPT];
	u32 dig[IPT];
	bool valid[IPT];
	u32 pos[IPT];
#pragma unroll
	for (int j = 0; j < IPT; ++j) {
		const u32 e = w * kWave * IPT + j * kWave + lane;
		valid[j] = e < n;
		k[j] = valid[j] ? s.k[e] : K(0);
		if (KV) v[j] = valid[j] ? s.v[e] : 0u;
		sd[j] = (carry_sid && valid[j]) ? s.sid[e] : 0u;
		dig[j] = by_sid ? ((sd[j] >> shift) & 0xFFu) : digit_of(k[j], shift);
	}
	block_rank<IPT>(dig, valid, pos, 8, s.rs);
#pragma unroll
	for (int j = 0; j < IPT; ++j) {
		if (!valid[j]) continue;
		s.k[pos[j]] = k[j];
		if (KV) s.v[pos[j]] = v[j];
		if (carry_sid) s.sid[pos[j]] = static_cast<u16>(sd[j]);
	}
	__syncthreads();
}

template <typename K, bool KV, int IPT>
__device__ void sort_key_passes(SegSmem<K, KV, IPT>& s, u32 n, bool carry_sid) {
	K a, o;
	and_or_lds<K, KV, IPT>(s, n, &a, &o);
	const K diff = a ^ o;
	for (u32 p = 0; p < sizeof(K); ++p) {
		if (((diff >> (8 * p)) & 0xFF) == 0) continue;  // uniform: every key shares the digit
		lds_pass<K, KV, IPT>(s, n, 8 * p, false, carry_sid);
	}
}


// ---- keys-only tiles: unstable MSD bucketing ------------------------------
// For keys alone the order of equal keys cannot be seen (they are equal bit
// patterns), so a tile may be sorted by any method.  One bucketing level
// over s.k[base, base + m) takes the range's minimum and maximum, puts key x
// in bucket (x - min) >> shift, with shift such that the key range spans
// kCap buckets (one key per bucket on average for keys spread over their
// range), scatters the keys to their buckets with LDS atomic cursors and
// moves every key of a bucket of 2..kMsdMaxBucket keys to its rank in the
// bucket (msd_place).  Larger buckets go on a stack in LDS and get a level
// of their own; each level narrows the key
// range by a factor kCap, so every range ends sorted, and a range of equal
// keys ends at once.  A level costs a few VALU operations per key, where a
// stable digit pass costs about 50 (wave ballots over 8 digit bits).
constexpr u32 kMsdMaxBucket = 32;
// Large segments of msd_chunks_min<K>() to kMsdChunks tiles are partitioned
// by top byte in one read (sort_large_msd).  The lower bound is measured:
// 2^27 32-bit keys in 4096-key segments 1.09 ms from 5 tiles, 1.19 from 2;
// 64-bit keys 1.49 ms from 2 tiles, 1.55 from 5.
template <typename K>
constexpr u32 msd_chunks_min() {
	return sizeof(K) == 4 ? 5u : 2u;
}
constexpr u32 kMsdChunks = 32;

template <typename K>
__device__ __forceinline__ u32 top_bit(K x) {
	return sizeof(K) == 8 ? 63u - static_cast<u32>(__builtin_clzll(static_cast<u64>(x)))
		: 31u - static_cast<u32>(__builtin_clz(static_cast<u32>(x)));
}

// LDS of the keys-only bucketing kernel.  k holds the tile; mcnt the bucket
// counts, then cursors; mstack the ranges waiting for a level of their own.
// Large segments use hist/base/gofs (bucket sizes and starts, window starts,
// chunk cursors) and co/rsrc/rdst (chunk-major partition), chunk items sid/lpos/segj (gather and write-out) and
// smin/sshift (each small segment's minimum key and bucket shift).
template <typename K, int IPT>
struct MsdSmem {
	static constexpr u32 kCap = kRB * IPT;
	static constexpr u32 kW = kCap / 2;
	K k[kCap];
	u32 mcnt[kCap];
	uint2 mstack[kCap / 32];
	union {
		struct {  // large segments
			u32 hist[4][256];
			u32 base[256];
			u32 gofs[256];
			u16 co[kMsdChunks][256];  // each chunk's bucket starts
			u32 rsrc[kMsdChunks];     // a window's runs: start in the scratch array
			u32 rdst[kMsdChunks];     // and in the window
		};
		struct {  // chunk items
			u16 sid[kCap];
			u32 lpos[kW + 1];
			u32 segj[kW];
			K smin[kW];
			u8 sshift[kW];
		};
	};
	K red[2 * kRBWaves];
	u32 misc[4];
	u32 wsum[kRBWaves];
};

// After a level's scatter over s.k[base, base + m) (s.mcnt[b] = end of
// bucket b, which starts at s.mcnt[b - 1], or at base for b = 0;
// bucket_of(x, i) = bucket of key x in slot i): the key in every slot of a
// bucket of 2..kMsdMaxBucket keys counts the keys of its bucket that precede
// it (smaller, or equal in an earlier slot) and moves to that rank; a larger
// bucket goes on the stack, pushed by its first slot.  Work per key, not per
// bucket: neighbouring lanes take neighbouring slots, mostly of one bucket,
// and loop as long as the largest bucket among them, four independent LDS
// loads a step, whatever the number of buckets.  (One thread per bucket,
// sorting it by a network or by insertion, measured slower: 32768 segments
// of log-uniform lengths up to 65536, 32-bit keys, 1.83 ms against 1.61.)  Contains barriers: call from
// uniform control flow.
template <typename K, int IPT, typename BucketOf>
__device__ __forceinline__ void msd_place(MsdSmem<K, IPT>& s, u32 base, u32 m,
	BucketOf bucket_of) {
	constexpr u32 kNone = ~0u;
	K x[IPT];
	u32 np[IPT];
#pragma unroll
	for (int j = 0; j < IPT; ++j) {
		const u32 i = base + j * kRB + threadIdx.x;
		np[j] = kNone;
		if (i >= base + m) continue;
		x[j] = s.k[i];
		const u32 b = bucket_of(x[j], i);
		const u32 e = s.mcnt[b];
		const u32 b0 = b == 0 ? base : s.mcnt[b - 1];
		const u32 mb = e - b0;
		if (mb > kMsdMaxBucket) {
			if (i == b0) {
				const u32 slot = atomicAdd(&s.misc[2], 1u);
				s.mstack[slot] = make_uint2(b0, mb);
			}
		} else if (mb > 1) {
			u32 r = b0;
			// Slots t >= e are not counted; their loads stay inside the LDS
			// block (s.k is followed by s.mcnt, and t < e + 3 <= kCap + 3).
			for (u32 t = b0; t < e; t += 4) {
				K y[4];
#pragma unroll
				for (int u = 0; u < 4; ++u) y[u] = s.k[t + u];
#pragma unroll
				for (int u = 0; u < 4; ++u)
					r += (t + u < e && (y[u] < x[j] || (y[u] == x[j] && t + u < i))) ? 1u : 0u;
			}
			if (r != i) np[j] = r;
		}
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < IPT; ++j)
		if (np[j] != kNone) s.k[np[j]] = x[j];
	__syncthreads();
}

// One level over s.k[base, base + m), m <= kCap.  Contains barriers: call from
// uniform control flow.  Pushes the buckets of more than kMsdMaxBucket keys
// onto s.mstack (count in s.misc[2]).  KNOWN: every key lies in [klo, khi]
// (klo < khi), which stands in for the range's minimum and maximum and saves
// their block reduction.
template <typename K, int IPT, bool KNOWN = false>
__device__ __forceinline__ void msd_level(MsdSmem<K, IPT>& s, u32 base, u32 m, K klo = K(0),
	K khi = K(0)) {
	using S = MsdSmem<K, IPT>;
	constexpr u32 NB = S::kCap;  // buckets (a power of two)
	constexpr u32 WB = static_cast<u32>(__builtin_ctz(NB));  // bucket bits
	constexpr u32 BPT = NB / kRB;  // buckets per thread in the scan
	static_assert((NB & (NB - 1)) == 0 && BPT % 4 == 0, "power-of-two tile, 16-byte rows");
	const u32 tid = threadIdx.x;
	const u32 lane = lane_id();
	const u32 w = wave_id();
	K k[IPT];
	bool valid[IPT];
	K lo = ~K(0), hi = K(0);
#pragma unroll
	for (int j = 0; j < IPT; ++j) {
		const u32 e = j * kRB + tid;
		valid[j] = e < m;
		k[j] = valid[j] ? s.k[base + e] : K(0);
		if (!KNOWN && valid[j]) {
			lo = k[j] < lo ? k[j] : lo;
			hi = k[j] > hi ? k[j] : hi;
		}
	}
	if constexpr (KNOWN) {
		lo = klo;
		hi = khi;
#pragma unroll
		for (u32 q = 0; q < BPT / 4; ++q)
			reinterpret_cast<uint4*>(s.mcnt)[tid * (BPT / 4) + q] = make_uint4(0u, 0u, 0u, 0u);
		__syncthreads();
	} else {
		// Block min/max; every thread zeroes its BPT counters meanwhile.
#pragma unroll
		for (int off = kWave / 2; off > 0; off >>= 1) {
			const K l2 = __shfl_xor(lo, off, kWave);
			const K h2 = __shfl_xor(hi, off, kWave);
			lo = l2 < lo ? l2 : lo;
			hi = h2 > hi ? h2 : hi;
		}
		K* red = s.red;
		if (lane == 0) {
			red[w] = lo;
			red[kRBWaves + w] = hi;
		}
#pragma unroll
		for (u32 q = 0; q < BPT / 4; ++q)
			reinterpret_cast<uint4*>(s.mcnt)[tid * (BPT / 4) + q] = make_uint4(0u, 0u, 0u, 0u);
		__syncthreads();
		lo = ~K(0);
		hi = K(0);
#pragma unroll
		for (int ww = 0; ww < kRBWaves; ++ww) {
			lo = red[ww] < lo ? red[ww] : lo;
			hi = red[kRBWaves + ww] > hi ? red[kRBWaves + ww] : hi;
		}
		__syncthreads();  // red is read before the next level writes it
		if (hi == lo) return;  // all keys equal (uniform over the block)
	}
	// Bucket of key x: (x - lo) >> shift, below NB for every key of the range
	// and non-decreasing in x.
	const u32 hb = top_bit(static_cast<K>(hi - lo));
	const u32 shift = hb + 1 >= WB ? hb + 1 - WB : 0u;
	const bool exact = shift == 0;  // one key value per bucket
#pragma unroll
	for (int j = 0; j < IPT; ++j)
		if (valid[j]) atomicAdd(&s.mcnt[static_cast<u32>((k[j] - lo) >> shift)], 1u);
	__syncthreads();
	// Thread tid scans buckets [tid * BPT, (tid + 1) * BPT).
	u32 c[BPT];
#pragma unroll
	for (u32 q = 0; q < BPT / 4; ++q) {
		const uint4 x = reinterpret_cast<const uint4*>(s.mcnt)[tid * (BPT / 4) + q];
		c[4 * q] = x.x;
		c[4 * q + 1] = x.y;
		c[4 * q + 2] = x.z;
		c[4 * q + 3] = x.w;
	}
	u32 tot = 0;
#pragma unroll
	for (u32 i = 0; i < BPT; ++i) tot += c[i];
	u32 run = base + block_exclusive_sum<kRB>(tot, s.wsum, static_cast<u32*>(nullptr));
#pragma unroll
	for (u32 i = 0; i < BPT; ++i) {
		const u32 x = c[i];
		c[i] = run;
		run += x;
	}
#pragma unroll
	for (u32 q = 0; q < BPT / 4; ++q)
		reinterpret_cast<uint4*>(s.mcnt)[tid * (BPT / 4) + q] =
			make_uint4(c[4 * q], c[4 * q + 1], c[4 * q + 2], c[4 * q + 3]);
	__syncthreads();
	// Scatter (every key is in registers): the bucket's cursor, then +1.
#pragma unroll
	for (int j = 0; j < IPT; ++j) {
		if (valid[j]) {
			const u32 p = atomicAdd(&s.mcnt[static_cast<u32>((k[j] - lo) >> shift)], 1u);
			s.k[p] = k[j];
		}
	}
	__syncthreads();
	if (exact) return;
	// Cursors are now bucket ends.
	msd_place<K, IPT>(s, base, m,
		[lo, shift](K x, u32) { return static_cast<u32>((x - lo) >> shift); });
}

// Sorts s.k[0, n), n <= kCap, keys alone.  The stack holds disjoint ranges
// of more than kMsdMaxBucket keys, so it never exceeds kCap / 33 entries.
// KNOWN: every key lies in [klo, khi], klo < khi (msd_level).
template <typename K, int IPT, bool KNOWN = false>
__device__ __forceinline__ void msd_tile_sort(MsdSmem<K, IPT>& s, u32 n, K klo = K(0),
	K khi = K(0)) {
	if (threadIdx.x == 0) s.misc[2] = 0;
	__syncthreads();
	msd_level<K, IPT, KNOWN>(s, 0, n, klo, khi);
	for (;;) {
		__syncthreads();
		const u32 sz = s.misc[2];
		if (sz == 0) break;
		const uint2 r = s.mstack[sz - 1];
		__syncthreads();
		if (threadIdx.x == 0) s.misc[2] = sz - 1;
		msd_level<K, IPT>(s, r.x, r.y);
	}
}

// Medium segment: whole segment in LDS.
template <typename K, bool KV, int IPT>
__device__ void sort_medium(const SegArgs<K>& a, SegSmem<K, KV, IPT>& s, u32 j) {
	const SegRange sr = seg_range(a.off, j, a.nseg, a.numel);
	const u32 start = sr.start;
	const u32 len = sr.len;
	g2s<K, IPT>(s.k, a.kin + start, len);
	if (KV) g2s<u32, IPT>(s.v, a.vin + start, len);
	__syncthreads();
	sort_key_passes<K, KV, IPT>(s, len, false);
	s2g<K, IPT>(a.kout + start, s.k, len);
	if (KV) s2g<u32, IPT>(a.vout + start, s.v, len);
	__syncthreads();
}

// Chunk of small segments starting in [c*W, (c+1)*W).
template <typename K, bool KV, int IPT>
__device__ void sort_chunk(const SegArgs<K>& a, SegSmem<K, KV, IPT>& s, u32 c) {
	constexpr u32 W = SegSmem<K, KV, IPT>::kW;
	constexpr u32 CAP = SegSmem<K, KV, IPT>::kCap;
	if (threadIdx.x == 0) {
		const uint2 r = a.wrange[c];
		s.misc[0] = r.x;
		s.misc[1] = r.y > r.x ? r.y - r.x : 0u;  // offsets that are not monotone: none
	}
	__syncthreads();
	const u32 j0 = s.misc[0];
	const u32 mall = s.misc[1];  // segments starting in the window (empty ones included)
	// Compact the non-empty small segments (distinct starts inside a window of
	// W elements, so at most W of them): segj[r] = segment start, lpos[r] =
	// its start in the LDS tile.
	u32 carry = 0, m = 0;
	for (u32 r0 = 0; r0 < mall; r0 += kRB) {
		const u32 r = r0 + threadIdx.x;
		u32 len = 0, start = 0;
		const u32 j = j0 + r;
		if (r < mall) {
			const SegRange sr = seg_range(a.off, j, a.nseg, a.numel);
			start = sr.start;
			len = (sr.len > W || seg_class(a.cls_lo, a.cls_hi, sr.len, W, CAP) != 0) ? 0u
				: sr.len;
		}
		u32 tot, cnt;
		const u32 ex = block_exclusive_sum<kRB>(len, s.rs.wsum, &tot);
		const u32 exc = block_exclusive_sum<kRB>(len ? 1u : 0u, s.rs.wsum, &cnt);
		if (len && m + exc < W) {
			s.segj[m + exc] = start;  // the segment's start in the array
			s.lpos[m + exc] = carry + ex;
		}
		carry += tot;
		m += cnt;
	}
	// Well-formed offsets give m <= W and carry <= CAP (distinct starts in a
	// window of W elements, each segment at most W long).  Malformed ones can
	// exceed both: the chunk is then left out, never written past LDS.
	if (threadIdx.x == 0) s.lpos[m < W ? m : W] = carry;
	__syncthreads();
	const u32 n = (m > W || carry > CAP) ? 0u : carry;
	if (n == 0) return;
	// Gather: element i belongs to the included segment r with
	// lpos[r] <= i < lpos[r+1] (fixed-length binary search, m <= W < 2^13),
	// then every thread's global loads are issued together.
	constexpr int B = IPT < kBatch ? IPT : kBatch;
#pragma unroll
	for (int r0 = 0; r0 < IPT; r0 += B) {
		u32 gidx[B];
		u32 gsid[B];
#pragma unroll
		for (int r = 0; r < B; ++r) {
			const u32 i = (r0 + r) * kRB + threadIdx.x;
			u32 l = 0;
#pragma unroll
			for (int b = 12; b >= 0; --b) {
				const u32 c = l + (1u << b);
				if (c < m && s.lpos[c] <= i) l = c;
			}
			gsid[r] = l;
			gidx[r] = s.segj[l] + (i - s.lpos[l]);
		}
		K tk[B];
		u32 tv[B];
#pragma unroll
		for (int r = 0; r < B; ++r) {
			const bool in = (r0 + r) * kRB + threadIdx.x < n;
			tk[r] = in ? a.kin[gidx[r]] : K(0);
			if (KV) tv[r] = in ? a.vin[gidx[r]] : 0u;
		}
#pragma unroll
		for (int r = 0; r < B; ++r) {
			const u32 i = (r0 + r) * kRB + threadIdx.x;
			if (i < n) {
				s.k[i] = tk[r];
				if (KV) s.v[i] = tv[r];
				s.sid[i] = static_cast<u16>(gsid[r]);
			}
		}
	}
	__syncthreads();
	const bool multi = m > 1;
	sort_key_passes<K, KV, IPT>(s, n, multi);
	if (multi) {
		// Stable regroup by segment: LSD passes over the local index.
		lds_pass<K, KV, IPT>(s, n, 0, true, true);
		if (m > 256) lds_pass<K, KV, IPT>(s, n, 8, true, true);
	}
#pragma unroll
	for (int rr = 0; rr < IPT; ++rr) {
		const u32 i = rr * kRB + threadIdx.x;
		if (i < n) {
			const u32 r = multi ? s.sid[i] : 0u;
			const u32 g = s.segj[r] + (i - s.lpos[r]);
			a.kout[g] = s.k[i];
			if (KV) a.vout[g] = s.v[i];
		}
	}
	__syncthreads();
}

// One chunk of a large segment in one digit pass: stable rank in the block,
// staged in digit order in LDS, then written to the pass's destination run
// of each digit (s.base: the segment's next free position per digit, advanced
// by the chunk's digit counts).  FULL: cnt == CAP, no per-element checks.
template <typename K, bool KV, int IPT, bool FULL>
__device__ __forceinline__ void large_chunk(SegSmem<K, KV, IPT>& s, const K (&k)[IPT],
	const u32 (&v)[IPT], u32 cnt, u32 shift, K* __restrict__ dst, u32* __restrict__ vdst) {
	const u32 tid = threadIdx.x;
	const u32 lane = lane_id();
	const u32 w = wave_id();
	u32 dig[IPT];
	bool valid[IPT];
	u32 pos[IPT];
#pragma unroll
	for (int jj = 0; jj < IPT; ++jj) {
		const u32 e = w * kWave * IPT + jj * kWave + lane;
		valid[jj] = FULL || e < cnt;
		dig[jj] = digit_of(k[jj], shift);
	}
	block_rank<IPT, FULL>(dig, valid, pos, 8, s.rs);
	// Output position of staged element i of digit d: gofs[d] + i.
	s.gofs[tid] = s.base[tid] - s.rs.tstart[tid];
#pragma unroll
	for (int jj = 0; jj < IPT; ++jj)
		if (FULL || valid[jj]) s.k[pos[jj]] = k[jj];
	__syncthreads();
	u32 gidx[IPT];
#pragma unroll
	for (int r = 0; r < IPT; ++r) {
		const u32 i = r * kRB + tid;
		if (FULL || i < cnt) {
			const K key = s.k[i];
			gidx[r] = s.gofs[digit_of(key, shift)] + i;
			dst[gidx[r]] = key;
		}
	}
	if (KV) {
		__syncthreads();
#pragma unroll
		for (int jj = 0; jj < IPT; ++jj)
			if (FULL || valid[jj]) s.v[pos[jj]] = v[jj];
		__syncthreads();
#pragma unroll
		for (int r = 0; r < IPT; ++r) {
			const u32 i = r * kRB + tid;
			if (FULL || i < cnt) vdst[gidx[r]] = s.v[i];
		}
	}
	__syncthreads();
	s.base[tid] += s.rs.total[tid];
	__syncthreads();
}


// Large segment, keys alone: one unstable partition of the segment into 256
// buckets, then windows of whole buckets, each under CAP keys, sorted in LDS
// by msd_tile_sort and written to the output.  The buckets are the top byte
// of the key or, when a top-byte bucket holds more than CAP/2 keys, the
// segment's key range (minimum to maximum) over 256 buckets.  A window holds
// the buckets whose start lies in one range of span = CAP - (largest bucket)
// positions, so it holds fewer than CAP keys.
//
// A segment of msd_chunks_min<K>() to kMsdChunks chunks (of CAP keys) is read
// once: each chunk is grouped by top-byte bucket in LDS and written back to
// its own place in the scratch array (whole cache lines), its bucket starts
// kept in s.co, and a window gathers one run from every chunk.  Otherwise
// (other lengths, or range buckets) a pass counts the buckets, each chunk's
// buckets are then stored to the bucket's place in the segment, and windows
// read one contiguous range.  Storing the runs of about CAP/256 keys that
// every chunk has per bucket costs more than reading them back once the
// segments in flight outgrow the caches (65536-key segments, 32-bit keys:
// 1.17 ms stored, 1.02 ms read); for a few chunks the runs still merge in
// the cache and the gather's run search is the larger cost (4096-key
// segments: 1.07 ms stored, 1.24 ms read).
//
// Returns false, having written nothing to the output, when a range bucket
// is over CAP/2 keys (keys concentrated in a small part of their range): the
// segment is then left to the stable digit passes.
template <typename K, int IPT, bool COMP>
__device__ __forceinline__ bool sort_large_msd(const SegArgs<K>& a, MsdSmem<K, IPT>& s,
	u32 start, u32 len) {
	using S = MsdSmem<K, IPT>;
	constexpr u32 CAP = S::kCap;
	constexpr u32 HALF = CAP / 2;
	constexpr u32 TS = sizeof(K) * 8 - 8;  // top-byte shift
	const u32 tid = threadIdx.x;
	const u32 lane = lane_id();
	const u32 w = wave_id();
	u32* hcnt = s.hist[0];  // bucket sizes, then starts
	u32* wbk = s.hist[1];   // first bucket of each window
	u32* ccnt = s.hist[2];  // per chunk: bucket counts, then cursors
	u32* wbeg = s.hist[3];  // window starts (at most one per bucket)
	// The segment's keys through base pointers.
	const K* __restrict__ ksrc = nullptr;
	const u32* __restrict__ ksrc32 = nullptr;
	if constexpr (COMP) ksrc32 = a.kin32 + start;
	else ksrc = a.kin + start;
	auto src = [&](u32 e) -> K {
		if constexpr (COMP) return (static_cast<K>(ksrc32[e]) << 32) | e;
		else return ksrc[e];
	};
	K* __restrict__ scr = a.kscr + start;
	K* __restrict__ out = a.kout + start;  // keys alone
	const u32 nch = len / CAP + (len % CAP != 0 ? 1u : 0u);
	const bool cm = nch >= msd_chunks_min<K>() && nch <= kMsdChunks;  // chunk-major (uniform)
	hcnt[tid] = 0;
	ccnt[tid] = 0;
	if (tid == 0) s.misc[2] = 0;
	__syncthreads();
	K lo = ~K(0), hi = K(0);
	if (cm) {
		// Pass A, chunk by chunk (the next chunk's loads in flight).
		K kn[IPT];
#pragma unroll
		for (int jj = 0; jj < IPT; ++jj) {
			const u32 e = jj * kRB + tid;
			kn[jj] = e < len ? src(e) : K(0);
		}
		for (u32 c = 0; c < nch; ++c) {
			const u32 c0 = c * CAP;
			const u32 cnt = (len - c0) < CAP ? (len - c0) : CAP;
			K k[IPT];
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj) k[jj] = kn[jj];
			if (c + 1 < nch) {
				const u32 c1 = c0 + CAP;
				const u32 cnt1 = (len - c1) < CAP ? (len - c1) : CAP;
#pragma unroll
				for (int jj = 0; jj < IPT; ++jj) {
					const u32 e = jj * kRB + tid;
					kn[jj] = e < cnt1 ? src(c1 + e) : K(0);
				}
			}
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj) {
				if (jj * kRB + tid < cnt) {
					lo = k[jj] < lo ? k[jj] : lo;
					hi = k[jj] > hi ? k[jj] : hi;
					atomicAdd(&ccnt[static_cast<u32>(k[jj] >> TS)], 1u);
				}
			}
			__syncthreads();
			const u32 cc = ccnt[tid];
			const u32 cs = block_exclusive_sum<kRB>(cc, s.wsum, static_cast<u32*>(nullptr));
			ccnt[tid] = cs;
			hcnt[tid] += cc;
			s.co[c][tid] = static_cast<u16>(cs);
			__syncthreads();
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj)
				if (jj * kRB + tid < cnt)
					s.k[atomicAdd(&ccnt[static_cast<u32>(k[jj] >> TS)], 1u)] = k[jj];
			__syncthreads();
#pragma unroll
			for (int r = 0; r < IPT; ++r) {
				const u32 i = r * kRB + tid;
				if (i < cnt) scr[c0 + i] = s.k[i];
			}
			ccnt[tid] = 0;
			__syncthreads();
		}
	} else {
		// Pass 1: minimum, maximum and the top-byte bucket sizes.
		for (u32 c0 = 0; c0 < len; c0 += kBatch * kRB) {
			K t[kBatch];
#pragma unroll
			for (int r = 0; r < kBatch; ++r) {
				const u32 i = c0 + r * kRB + tid;
				t[r] = i < len ? src(i) : src(0);
			}
#pragma unroll
			for (int r = 0; r < kBatch; ++r) {
				lo = t[r] < lo ? t[r] : lo;
				hi = t[r] > hi ? t[r] : hi;
				if (c0 + r * kRB + tid < len) atomicAdd(&hcnt[static_cast<u32>(t[r] >> TS)], 1u);
			}
		}
	}
#pragma unroll
	for (int off = kWave / 2; off > 0; off >>= 1) {
		const K l2 = __shfl_xor(lo, off, kWave);
		const K h2 = __shfl_xor(hi, off, kWave);
		lo = l2 < lo ? l2 : lo;
		hi = h2 > hi ? h2 : hi;
	}
	K* red = s.red;
	if (lane == 0) {
		red[w] = lo;
		red[kRBWaves + w] = hi;
	}
	__syncthreads();
	if (hcnt[tid] > HALF) atomicOr(&s.misc[2], 1u);
	lo = ~K(0);
	hi = K(0);
#pragma unroll
	for (int ww = 0; ww < kRBWaves; ++ww) {
		lo = red[ww] < lo ? red[ww] : lo;
		hi = red[kRBWaves + ww] > hi ? red[kRBWaves + ww] : hi;
	}
	if (hi == lo) {  // uniform over the block: every key equal
		seg_copy<K, IPT, COMP>(a, start, len);
		return true;
	}
	__syncthreads();
	// Bucket of key x: (x - bbase) >> shift, below 256 and non-decreasing in x.
	K bbase = K(0);
	u32 shift = TS;
	const bool top = s.misc[2] == 0;  // uniform: read after the barrier
	if (!top) {
		// Pass 2: the segment's key range over 256 buckets.
		const u32 hb = top_bit(static_cast<K>(hi - lo));
		bbase = lo;
		shift = hb + 1 >= 8 ? hb + 1 - 8 : 0u;
		__syncthreads();
		hcnt[tid] = 0;
		if (tid == 0) s.misc[2] = 0;
		__syncthreads();
		for (u32 c0 = 0; c0 < len; c0 += kBatch * kRB) {
			K t[kBatch];
#pragma unroll
			for (int r = 0; r < kBatch; ++r) {
				const u32 i = c0 + r * kRB + tid;
				t[r] = i < len ? src(i) : K(0);
			}
#pragma unroll
			for (int r = 0; r < kBatch; ++r)
				if (c0 + r * kRB + tid < len)
					atomicAdd(&hcnt[static_cast<u32>((t[r] - bbase) >> shift)], 1u);
		}
		__syncthreads();
		if (hcnt[tid] > HALF) atomicOr(&s.misc[2], 1u);
	}
	const u32 c = hcnt[tid];
	u32 cmax = c;
#pragma unroll
	for (int off = kWave / 2; off > 0; off >>= 1) {
		const u32 o = __shfl_xor(cmax, off, kWave);
		cmax = o > cmax ? o : cmax;
	}
	if (tid == 0) s.misc[1] = 0;
	const u32 b0 = block_exclusive_sum<kRB>(c, s.wsum, static_cast<u32*>(nullptr));
	if (s.misc[2] != 0) return false;  // read after the barriers above: uniform
	if (lane == 0) atomicMax(&s.misc[1], cmax);
	s.base[tid] = b0;  // next free position of the bucket in the scratch array
	hcnt[tid] = b0;    // bucket starts
	__syncthreads();
	{
		// Bucket tid opens a window when its start lies in another range of
		// span = CAP - (largest bucket) positions than its predecessor's: a
		// window's buckets start within span positions of each other and the
		// last is at most CAP - span keys, so the window holds fewer than CAP
		// keys (span >= CAP/2, as no bucket exceeds CAP/2).
		const u32 span = CAP - s.misc[1];
		const u32 prev = tid == 0 ? 0u : hcnt[tid - 1];
		const bool op = tid == 0 || (b0 / span) != (prev / span);
		u32 nw = 0;
		const u32 idx = block_exclusive_sum<kRB>(op ? 1u : 0u, s.wsum, &nw);
		if (op) {
			wbeg[idx] = b0;
			wbk[idx] = tid;
		}
		if (tid == 0) s.misc[0] = nw;
	}
	const bool gather = cm && top;  // windows gather their runs from the chunks
	if (!gather) {
		// Pass 3: partition, chunk by chunk (the next chunk's loads in
		// flight), each chunk staged in LDS by bucket so that every bucket's
		// keys go out as one run.  (Every key taking its slot from the
		// bucket's LDS cursor and being stored straight to the scratch array
		// measured slower: 32768 segments of log-uniform lengths 2.30 -> 2.40
		// ms, 65536-key segments 1.54 -> 2.45 ms.)
		ccnt[tid] = 0;
		K kn[IPT];
#pragma unroll
		for (int jj = 0; jj < IPT; ++jj) {
			const u32 e = jj * kRB + tid;
			kn[jj] = e < len ? src(e) : K(0);
		}
		__syncthreads();
		for (u32 c0 = 0; c0 < len; c0 += CAP) {
			const u32 cnt = (len - c0) < CAP ? (len - c0) : CAP;
			K k[IPT];
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj) k[jj] = kn[jj];
			const u32 c1 = c0 + CAP;
			if (c1 < len) {
				const u32 cnt1 = (len - c1) < CAP ? (len - c1) : CAP;
#pragma unroll
				for (int jj = 0; jj < IPT; ++jj) {
					const u32 e = jj * kRB + tid;
					kn[jj] = e < cnt1 ? src(c1 + e) : K(0);
				}
			}
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj)
				if (jj * kRB + tid < cnt)
					atomicAdd(&ccnt[static_cast<u32>((k[jj] - bbase) >> shift)], 1u);
			__syncthreads();
			const u32 cc = ccnt[tid];
			const u32 cs = block_exclusive_sum<kRB>(cc, s.wsum, static_cast<u32*>(nullptr));
			ccnt[tid] = cs;                 // chunk cursor of bucket tid
			s.gofs[tid] = s.base[tid] - cs;  // staged element i of bucket d goes to gofs[d] + i
			__syncthreads();
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj)
				if (jj * kRB + tid < cnt)
					s.k[atomicAdd(&ccnt[static_cast<u32>((k[jj] - bbase) >> shift)], 1u)] = k[jj];
			__syncthreads();
#pragma unroll
			for (int r = 0; r < IPT; ++r) {
				const u32 i = r * kRB + tid;
				if (i < cnt) {
					const K key = s.k[i];
					scr[s.gofs[static_cast<u32>((key - bbase) >> shift)] + i] = key;
				}
			}
			__syncthreads();
			s.base[tid] += cc;
			ccnt[tid] = 0;
			__syncthreads();
		}
	} else {
		__syncthreads();
	}
	// Windows: load (the next window's keys in flight during the sort), sort
	// in LDS, write to the output.  Gathered: window wi is the run of its
	// buckets [wbk[wi], wbk[wi + 1]) in every chunk, run c starting at
	// s.rsrc[c] in the scratch array and at s.rdst[c] in the window.
	const u32 nw = s.misc[0];
	auto runs = [&](u32 wi) {  // wave 0 builds window wi's run table
		if (w != 0) return;
		const u32 bl = wbk[wi];
		const u32 bh = wi + 1 < nw ? wbk[wi + 1] : 256u;
		u32 ln = 0, rs = 0;
		if (lane < nch) {
			const u32 cntc = (len - lane * CAP) < CAP ? (len - lane * CAP) : CAP;
			rs = s.co[lane][bl];
			ln = (bh < 256u ? static_cast<u32>(s.co[lane][bh]) : cntc) - rs;
		}
		u32 inc = ln;
#pragma unroll
		for (u32 d = 1; d < kMsdChunks; d <<= 1) {
			const u32 t = __shfl_up(inc, d, kWave);
			if (lane >= d) inc += t;
		}
		if (lane < nch) {
			s.rsrc[lane] = lane * CAP + rs;
			s.rdst[lane] = inc - ln;
		}
	};
	K kw[IPT];
	auto load = [&](u32 wi) {  // window wi's keys into kw
		const u32 p = wbeg[wi];
		const u32 n = (wi + 1 < nw ? wbeg[wi + 1] : len) - p;
		if (gather) {
			// Run of element e: the last c < nch with rdst[c] <= e, by a
			// binary search taken one step at a time for all IPT elements
			// (IPT independent LDS loads a step).
			u32 l[IPT];
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj) l[jj] = 0;
#pragma unroll
			for (u32 b = kMsdChunks / 2; b > 0; b >>= 1) {
				if (b >= nch) continue;  // uniform: steps of at least nch find nothing
				u32 v[IPT];
#pragma unroll
				for (int jj = 0; jj < IPT; ++jj) {
					const u32 c = l[jj] + b;
					v[jj] = s.rdst[c < nch ? c : 0u];
				}
#pragma unroll
				for (int jj = 0; jj < IPT; ++jj)
					if (l[jj] + b < nch && v[jj] <= jj * kRB + tid) l[jj] += b;
			}
			u32 src_i[IPT];
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj)
				src_i[jj] = s.rsrc[l[jj]] + (jj * kRB + tid - s.rdst[l[jj]]);
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj)
				kw[jj] = jj * kRB + tid < n ? scr[src_i[jj]] : K(0);
		} else {
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj) {
				const u32 e = jj * kRB + tid;
				kw[jj] = e < n ? scr[p + e] : K(0);
			}
		}
	};
	if (gather) runs(0);
	__syncthreads();
	load(0);
	for (u32 wi = 0; wi < nw; ++wi) {
		const u32 p0 = wbeg[wi];
		const u32 n = (wi + 1 < nw ? wbeg[wi + 1] : len) - p0;
#pragma unroll
		for (int jj = 0; jj < IPT; ++jj) {
			const u32 e = jj * kRB + tid;
			if (e < n) s.k[e] = kw[jj];
		}
		if (gather && wi + 1 < nw) runs(wi + 1);
		__syncthreads();
		if (wi + 1 < nw) load(wi + 1);
		if (n > 1) {
			if (top) {
				// The window's top-byte buckets [bl, bh) bound its keys.
				const u32 bl = wbk[wi];
				const u32 bh = wi + 1 < nw ? wbk[wi + 1] : 256u;
				const K wlo = static_cast<K>(bl) << TS;
				const K whi = bh >= 256u ? ~K(0) : (static_cast<K>(bh) << TS) - 1;
				msd_tile_sort<K, IPT, true>(s, n, wlo, whi);
			} else {
				msd_tile_sort<K, IPT>(s, n);
			}
		}
		// Keys alone store through the segment's own restrict pointer (the
		// generic store measured 4-6% slower: 1.37 -> 1.42 ms for 32768
		// segments of log-uniform lengths, 32-bit keys).
		if constexpr (COMP) s2g_out<K, IPT, COMP>(a, start + p0, start, s.k, n);
		else s2g<K, IPT>(out + p0, s.k, n);
		__syncthreads();
	}
	return true;
}

// Medium segment, keys alone: whole segment in LDS, msd_tile_sort.
template <typename K, int IPT, bool COMP>
__device__ __forceinline__ void sort_medium_msd(const SegArgs<K>& a, MsdSmem<K, IPT>& s,
	u32 start, u32 len) {
	g2s_in<K, IPT, COMP>(a, s.k, start, 0, len);
	__syncthreads();
	if (len > 1) msd_tile_sort<K, IPT>(s, len);
	s2g_out<K, IPT, COMP>(a, start, start, s.k, len);
	__syncthreads();
}

// Chunk of small segments, keys alone: the segments starting in
// [c*W, (c+1)*W) gathered into one tile as in sort_chunk, then one segmented
// bucketing level: segment r (at tile positions [lpos[r], lpos[r+1]), n_r
// keys) owns buckets lpos[r] + ((x - min_r) >> shift_r), at most n_r of them,
// so every key stays inside its segment's range and one pass of counts,
// cursors and small-bucket sorts orders every segment at once; buckets over
// kMsdMaxBucket keys get levels of their own as in msd_tile_sort.
template <typename K, int IPT, bool COMP>
__device__ __forceinline__ void sort_chunk_msd(const SegArgs<K>& a, MsdSmem<K, IPT>& s, u32 c) {
	using S = MsdSmem<K, IPT>;
	constexpr u32 W = S::kW;
	constexpr u32 CAP = S::kCap;
	const u32 tid = threadIdx.x;
	if (tid == 0) {
		const uint2 r = a.wrange[c];
		s.misc[0] = r.x;
		s.misc[1] = r.y > r.x ? r.y - r.x : 0u;  // offsets that are not monotone: none
	}
	__syncthreads();
	const u32 j0 = s.misc[0];
	const u32 mall = s.misc[1];
	// The window's non-empty small segments: segj[r] = start in the array,
	// lpos[r] = start in the tile (at most W of them, CAP keys in all).
	u32 carry = 0, m = 0;
	for (u32 r0 = 0; r0 < mall; r0 += kRB) {
		const u32 r = r0 + tid;
		u32 len = 0, start = 0;
		const u32 j = j0 + r;
		if (r < mall) {
			const SegRange sr = seg_range(a.off, j, a.nseg, a.numel);
			start = sr.start;
			len = (sr.len > W || seg_class(a.cls_lo, a.cls_hi, sr.len, W, CAP) != 0) ? 0u
				: sr.len;
		}
		u32 tot, cnt;
		const u32 ex = block_exclusive_sum<kRB>(len, s.wsum, &tot);
		const u32 exc = block_exclusive_sum<kRB>(len ? 1u : 0u, s.wsum, &cnt);
		if (len && m + exc < W) {
			s.segj[m + exc] = start;
			s.lpos[m + exc] = carry + ex;
		}
		carry += tot;
		m += cnt;
	}
	if (tid == 0) s.lpos[m < W ? m : W] = carry;
	__syncthreads();
	const u32 n = (m > W || carry > CAP) ? 0u : carry;
	if (n == 0) return;
	// Gather (segment of element i by binary search over lpos) and each
	// segment's minimum and maximum (the maximum kept in mcnt, unused yet).
	K* smax = reinterpret_cast<K*>(s.mcnt);  // W keys fit in CAP words
	for (u32 r = tid; r < m; r += kRB) {
		s.smin[r] = ~K(0);
		smax[r] = K(0);
	}
	__syncthreads();
	constexpr int B = IPT < kBatch ? IPT : kBatch;
#pragma unroll
	for (int r0 = 0; r0 < IPT; r0 += B) {
		u32 gidx[B];
		u32 gsid[B];
#pragma unroll
		for (int r = 0; r < B; ++r) {
			const u32 i = (r0 + r) * kRB + tid;
			u32 l = 0;
#pragma unroll
			for (int b = 12; b >= 0; --b) {
				if ((1u << b) >= m) continue;  // uniform: no segment that far
				const u32 cc = l + (1u << b);
				if (cc < m && s.lpos[cc] <= i) l = cc;
			}
			gsid[r] = l;
			gidx[r] = s.segj[l] + (i - s.lpos[l]);
		}
		K tk[B];
#pragma unroll
		for (int r = 0; r < B; ++r) {
			const bool in = (r0 + r) * kRB + tid < n;
			tk[r] = in ? seg_key<K, COMP>(a, gidx[r], gidx[r] - s.segj[gsid[r]]) : K(0);
		}
#pragma unroll
		for (int r = 0; r < B; ++r) {
			const u32 i = (r0 + r) * kRB + tid;
			const bool in = i < n;
			if (in) {
				s.k[i] = tk[r];
				s.sid[i] = static_cast<u16>(gsid[r]);
			}
			// A wave whose elements all lie in one segment (every wave but
			// the few at segment boundaries, for segments of 64 keys and
			// more) reduces them first and takes one atomic per bound.
			const u32 g0 = __builtin_amdgcn_readfirstlane(gsid[r]);
			if (ballot(in && gsid[r] != g0) == 0 && ballot(in) == ~0ull) {
				K lo = tk[r], hi = tk[r];
#pragma unroll
				for (int off = kWave / 2; off > 0; off >>= 1) {
					const K l2 = __shfl_xor(lo, off, kWave);
					const K h2 = __shfl_xor(hi, off, kWave);
					lo = l2 < lo ? l2 : lo;
					hi = h2 > hi ? h2 : hi;
				}
				if (lane_id() == 0) {
					atomicMin(&s.smin[g0], lo);
					atomicMax(&smax[g0], hi);
				}
			} else if (in) {
				atomicMin(&s.smin[gsid[r]], tk[r]);
				atomicMax(&smax[gsid[r]], tk[r]);
			}
		}
	}
	__syncthreads();
	// Bucket shift of segment r: (max - min) >> shift below the largest power
	// of two not above its length.
	for (u32 r = tid; r < m; r += kRB) {
		const u32 nr = s.lpos[r + 1] - s.lpos[r];
		const K range = smax[r] - s.smin[r];
		const u32 lb = 31u - static_cast<u32>(__builtin_clz(nr));  // nr >= 1
		u32 sh = 0;
		if (range != 0) {
			const u32 hb = top_bit(range);
			sh = hb + 1 > lb ? hb + 1 - lb : 0u;
		}
		s.sshift[r] = static_cast<u8>(sh);
	}
	__syncthreads();
	for (u32 i = tid; i < n; i += kRB) s.mcnt[i] = 0;
	if (tid == 0) s.misc[2] = 0;
	__syncthreads();
	// Counts, cursors, scatter: every key stays in its segment's range.
	K k[IPT];
	u32 bk[IPT];
#pragma unroll
	for (int j = 0; j < IPT; ++j) {
		const u32 i = j * kRB + tid;
		bk[j] = 0;
		if (i < n) {
			k[j] = s.k[i];
			const u32 r = s.sid[i];
			bk[j] = s.lpos[r] + static_cast<u32>((k[j] - s.smin[r]) >> s.sshift[r]);
			atomicAdd(&s.mcnt[bk[j]], 1u);
		}
	}
	__syncthreads();
	{
		constexpr u32 BPT = CAP / kRB;  // buckets per thread (= IPT)
		u32 cnt[BPT];
		u32 tot = 0;
#pragma unroll
		for (u32 q = 0; q < BPT; ++q) {
			const u32 b = tid * BPT + q;
			cnt[q] = b < n ? s.mcnt[b] : 0u;
			tot += cnt[q];
		}
		u32 run = block_exclusive_sum<kRB>(tot, s.wsum, static_cast<u32*>(nullptr));
#pragma unroll
		for (u32 q = 0; q < BPT; ++q) {
			const u32 b = tid * BPT + q;
			if (b < n) s.mcnt[b] = run;
			run += cnt[q];
		}
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < IPT; ++j)
		if (j * kRB + tid < n) s.k[atomicAdd(&s.mcnt[bk[j]], 1u)] = k[j];
	__syncthreads();
	// Every key to its rank in its bucket (slot i stays in the segment of
	// slot i); buckets over kMsdMaxBucket keys on the stack.
	msd_place<K, IPT>(s, 0u, n, [&s](K x, u32 i) {
		const u32 r = s.sid[i];
		return s.lpos[r] + static_cast<u32>((x - s.smin[r]) >> s.sshift[r]);
	});
	for (;;) {
		const u32 sz = s.misc[2];
		if (sz == 0) break;
		const uint2 r = s.mstack[sz - 1];
		__syncthreads();
		if (tid == 0) s.misc[2] = sz - 1;
		msd_level<K, IPT>(s, r.x, r.y);
		__syncthreads();
	}
	// Write out: element i of segment r goes to segj[r] + (i - lpos[r]).
#pragma unroll
	for (int j = 0; j < IPT; ++j) {
		const u32 i = j * kRB + tid;
		if (i < n) {
			const u32 r = s.sid[i];
			seg_put<K, COMP>(a, s.segj[r] + (i - s.lpos[r]), s.segj[r], s.k[i]);
		}
	}
	__syncthreads();
}

// Large segment: LSD passes through global memory.
template <typename K, bool KV, int IPT>
__device__ void sort_large(const SegArgs<K>& a, SegSmem<K, KV, IPT>& s, u32 start, u32 len) {
	constexpr u32 CAP = SegSmem<K, KV, IPT>::kCap;
	constexpr u32 P = sizeof(K);
	const u32 tid = threadIdx.x;
	for (u32 i = tid; i < P * 256; i += kRB) (&s.hist[0][0])[i] = 0;
	__syncthreads();
	// One read: every digit histogram plus AND/OR of the keys.
	K ka = ~K(0), ko = K(0);
	for (u32 c0 = 0; c0 < len; c0 += kBatch * kRB) {
		K t[kBatch];
#pragma unroll
		for (int r = 0; r < kBatch; ++r) {
			const u32 i = c0 + r * kRB + tid;
			t[r] = i < len ? a.kin[start + i] : K(0);
		}
#pragma unroll
		for (int r = 0; r < kBatch; ++r) {
			if (c0 + r * kRB + tid >= len) continue;
			ka &= t[r];
			ko |= t[r];
#pragma unroll
			for (u32 p = 0; p < P; ++p) atomicAdd(&s.hist[p][digit_of(t[r], 8 * p)], 1u);
		}
	}
#pragma unroll
	for (int off = kWave / 2; off > 0; off >>= 1) {
		ka &= __shfl_xor(ka, off, kWave);
		ko |= __shfl_xor(ko, off, kWave);
	}
	K* red = reinterpret_cast<K*>(s.k);
	if (lane_id() == 0) {
		red[wave_id()] = ka;
		red[kRBWaves + wave_id()] = ko;
	}
	__syncthreads();
	K diff = 0;
	{
		K aa = ~K(0), oo = K(0);
#pragma unroll
		for (int w = 0; w < kRBWaves; ++w) {
			aa &= red[w];
			oo |= red[kRBWaves + w];
		}
		diff = aa ^ oo;
	}
	u32 executed = 0;
#pragma unroll
	for (u32 p = 0; p < P; ++p) executed += ((diff >> (8 * p)) & 0xFF) ? 1u : 0u;
	__syncthreads();
	const bool in_place = a.kout == a.kin;
	const K* src = a.kin + start;
	const u32* vsrc = KV ? a.vin + start : nullptr;
	u32 k_ex = 0;
	for (u32 p = 0; p < P; ++p) {
		if (((diff >> (8 * p)) & 0xFF) == 0) continue;
		// destination of executed pass k_ex (the last one lands in kout when possible)
		bool to_out;
		if (in_place) to_out = (k_ex % 2) == 1;
		else to_out = ((executed - 1 - k_ex) % 2) == 0;
		K* dst = (to_out ? a.kout : a.kscr) + start;
		u32* vdst = KV ? (to_out ? a.vout : a.vscr) + start : nullptr;
		const u32 shift = 8 * p;
		{
			const u32 c = s.hist[p][tid];
			const u32 ex = block_exclusive_sum<kRB>(c, s.rs.wsum, static_cast<u32*>(nullptr));
			s.base[tid] = ex;
		}
		__syncthreads();
		// Chunks are software-pipelined: the loads of chunk c+1 are issued
		// before chunk c is ranked and scattered, so their latency overlaps
		// that work instead of stalling the block at the top of every chunk.
		const u32 lane = lane_id();
		const u32 w = wave_id();
		K kn[IPT];
		u32 vn[IPT];
#pragma unroll
		for (int jj = 0; jj < IPT; ++jj) {
			const u32 e = w * kWave * IPT + jj * kWave + lane;
			kn[jj] = e < len ? src[e] : K(0);
			if (KV) vn[jj] = e < len ? vsrc[e] : 0u;
		}
		for (u32 c0 = 0; c0 < len; c0 += CAP) {
			const u32 cnt = (len - c0) < CAP ? (len - c0) : CAP;
			K k[IPT];
			u32 v[IPT];
#pragma unroll
			for (int jj = 0; jj < IPT; ++jj) {
				k[jj] = kn[jj];
				if (KV) v[jj] = vn[jj];
			}
			const u32 c1 = c0 + CAP;
			if (c1 < len) {
				const u32 cnt1 = (len - c1) < CAP ? (len - c1) : CAP;
#pragma unroll
				for (int jj = 0; jj < IPT; ++jj) {
					const u32 e = w * kWave * IPT + jj * kWave + lane;
					kn[jj] = e < cnt1 ? src[c1 + e] : K(0);
					if (KV) vn[jj] = e < cnt1 ? vsrc[c1 + e] : 0u;
				}
			}
			// The check-free body for full chunks only for keys alone: with
			// values the second copy of the body made the kernel spill.
			if constexpr (!KV) {
				if (cnt == CAP) {
					large_chunk<K, KV, IPT, true>(s, k, v, cnt, shift, dst, vdst);
					continue;
				}
			}
			large_chunk<K, KV, IPT, false>(s, k, v, cnt, shift, dst, vdst);
		}
		src = dst;
		vsrc = vdst;
		++k_ex;
	}
	// Result not yet in kout: in place with an odd pass count, or out of place
	// with nothing to do.
	const K* fin = src;
	if (fin != a.kout + start) {
		g2g<K, IPT>(a.kout + start, fin, len);
		if (KV) g2g<u32, IPT>(a.vout + start, vsrc, len);
	}
	__syncthreads();
}

// Blocks per CU the register allocation aims for: five for the 8-item
// 32-bit keys-only tile, three for the 16-item one, fewer where the shape's
// registers and LDS cannot reach that.
template <typename K, bool KV, int IPT>
constexpr int seg_min_blocks() {
	return (sizeof(K) == 4 && !KV && IPT <= 8) ? 5
		: (sizeof(K) == 4 && !KV && IPT <= 16) ? 3
		: (IPT >= 32 || (KV && sizeof(K) == 8)) ? 1 : 2;
}

// FB (key-value sorts): the fallback list of a composite sort on the
// bucketing kernel, copied to the output and sorted there in place; a
// separate instantiation, so that the others keep their register budgets.
template <typename K, bool KV, int IPT, bool FB = false>
__global__ __launch_bounds__(kRB, (seg_min_blocks<K, KV, IPT>())) void seg_sort_persistent(
	SegArgs<K> a) {
	__shared__ SegSmem<K, KV, IPT> s;
	// fb_mode (keys alone, after seg_msd_persistent): every segment is done
	// except those in the fallback list, which sit in the output array and
	// are sorted there in place.
	const bool fb = (!KV || FB) && a.fb_mode;
	const u32 nl = fb ? a.meta->n_fb : a.meta->n_large;
	const u32 nm = fb ? 0u : a.meta->n_medium;
	const u32 total = nl + nm + (fb ? 0u : a.meta->n_chunks);
	for (;;) {
		if (threadIdx.x == 0) s.misc[3] = atomicAdd(&a.meta->head, 1u);
		__syncthreads();
		// Uniform over the block: scalar loads and addresses downstream.
		const u32 item = __builtin_amdgcn_readfirstlane(s.misc[3]);
		__syncthreads();
		if (item >= total) break;
		if (item < nl) {
			if constexpr (KV && FB) {
				const uint2 r = a.fb[item];
				SegArgs<K> b = a;
				b.kin = a.kout;
				b.vin = a.vout;
				sort_large<K, KV, IPT>(b, s, __builtin_amdgcn_readfirstlane(r.x),
					__builtin_amdgcn_readfirstlane(r.y));
			} else if constexpr (KV) {
				const SegRange sr = seg_range(a.off, a.large_ix[item], a.nseg, a.numel);
				sort_large<K, KV, IPT>(a, s, __builtin_amdgcn_readfirstlane(sr.start),
					__builtin_amdgcn_readfirstlane(sr.len));
			} else {
				SegArgs<K> b = a;
				u32 start, len;
				if (a.fb_mode) {
					const uint2 r = a.fb[item];
					start = r.x;
					len = r.y;
					b.kin = a.kout;
				} else {
					const SegRange sr = seg_range(a.off, a.large_ix[item], a.nseg, a.numel);
					start = sr.start;
					len = sr.len;
				}
				sort_large<K, KV, IPT>(b, s, __builtin_amdgcn_readfirstlane(start),
					__builtin_amdgcn_readfirstlane(len));
			}
		} else if (item < nl + nm) {
			sort_medium<K, KV, IPT>(a, s, a.medium[item - nl]);
		} else {
			sort_chunk<K, KV, IPT>(a, s, a.chunks[item - nl - nm]);
		}
	}
}

// Medium or chunk items per work ticket of the bucketing kernel.
constexpr u32 kSmallBatch = 8;

// Resident blocks per CU the bucketing kernel is compiled for (its register
// budget): four 256x8 tiles of 32-bit keys fit the LDS of a CU, three of
// 64-bit keys.
template <typename K, int IPT>
constexpr int msd_min_blocks() {
	return IPT <= 8 ? (sizeof(K) == 4 ? 4 : 3) : IPT <= 16 ? 2 : 1;
}

// Keys alone: every segment, sorted by unstable MSD bucketing (the order of
// equal keys cannot be seen): large segments, longest first, then medium
// ones, then the chunk items.  A large segment whose keys crowd into part of
// their range is copied to the output unsorted and listed for
// seg_sort_persistent (fb_mode), which runs the stable digit passes on it in
// place.
template <typename K, int IPT, bool COMP>
__global__ __launch_bounds__(kRB, (msd_min_blocks<K, IPT>())) void seg_msd_persistent(
	SegArgs<K> a) {
	__shared__ MsdSmem<K, IPT> s;
	const LargeOrder lord = large_order(a.meta);
	const u32 nl = a.meta->n_large;
	const u32 nlm = nl + a.meta->n_medium;
	const u32 total = nlm + a.meta->n_chunks;
	// Tickets: one per large segment, then one per kSmallBatch medium or chunk
	// items (one atomic per item on the one queue counter bounds a queue of
	// one-tile segments).
	const u32 tickets = nl + (total - nl + kSmallBatch - 1) / kSmallBatch;
	if (threadIdx.x == 0) s.misc[3] = atomicAdd(&a.meta->head_msd, 1u);
	__syncthreads();
	u32 t = __builtin_amdgcn_readfirstlane(s.misc[3]);
	__syncthreads();
	while (t < tickets) {
		// The next ticket is taken now, its round trip under this item's work.
		u32 ticket = 0;
		if (threadIdx.x == 0) ticket = atomicAdd(&a.meta->head_msd, 1u);
		if (t < nl) {
			const SegRange sr = seg_range(a.off, large_seg(a, lord, t), a.nseg, a.numel);
			const u32 start = __builtin_amdgcn_readfirstlane(sr.start);
			const u32 len = __builtin_amdgcn_readfirstlane(sr.len);
			if (!sort_large_msd<K, IPT, COMP>(a, s, start, len)) {
				__syncthreads();
				seg_copy<K, IPT, COMP>(a, start, len);
				if (threadIdx.x == 0) a.fb[atomicAdd(&a.meta->n_fb, 1u)] = make_uint2(start, len);
			}
		} else {
			const u32 i0 = nl + (t - nl) * kSmallBatch;
			const u32 i1 = total - i0 < kSmallBatch ? total : i0 + kSmallBatch;
			// Medium items: the next one's keys are in flight during the
			// current one's sort.
			K kw[IPT];
			u32 nstart = 0, nlen = 0;
			auto fetch = [&](u32 item) {
				const SegRange sr = seg_range(a.off, a.medium[item - nl], a.nseg, a.numel);
				nstart = __builtin_amdgcn_readfirstlane(sr.start);
				nlen = __builtin_amdgcn_readfirstlane(sr.len);
#pragma unroll
				for (int jj = 0; jj < IPT; ++jj) {
					const u32 e = jj * kRB + threadIdx.x;
					kw[jj] = e < nlen ? seg_key<K, COMP>(a, nstart + e, e) : K(0);
				}
			};
			if (i0 < nlm) fetch(i0);
			for (u32 item = i0; item < i1; ++item) {
				if (item >= nlm) {
					sort_chunk_msd<K, IPT, COMP>(a, s, a.chunks[item - nlm]);
					continue;
				}
				const u32 start = nstart, len = nlen;
#pragma unroll
				for (int jj = 0; jj < IPT; ++jj) {
					const u32 e = jj * kRB + threadIdx.x;
					if (e < len) s.k[e] = kw[jj];
				}
				if (item + 1 < i1 && item + 1 < nlm) fetch(item + 1);
				__syncthreads();
				if (len > 1) msd_tile_sort<K, IPT>(s, len);
				s2g_out<K, IPT, COMP>(a, start, start, s.k, len);
				__syncthreads();
			}
		}
		if (threadIdx.x == 0) s.misc[3] = ticket;
		__syncthreads();
		t = __builtin_amdgcn_readfirstlane(s.misc[3]);
		__syncthreads();
	}
}

struct SegData {
	Scratch lists;  // medium + large index lists
	Scratch meta;
	Scratch kscr;
	Scratch vscr;
};

template <typename K, bool KV, int IPT>
bool seg_launch(CCLQueue* q, SegData* sd, const SegArgs<K>& a0, u32 grid, bool msd, GError** err) {
	constexpr u32 CAP = kRB * IPT;
	constexpr u32 W = CAP / 2;
	SegArgs<K> a = a0;
	hipStream_t st = q->stream;
	{
		Launch L(q, "segsort_reset");
		seg_reset<<<1, 64, 0, st>>>(a.meta);
		if (!L.done(err)) return false;
	}
	{
		Launch L(q, "segsort_classify");
		const u32 g = static_cast<u32>(div_ceil(a.nseg, 256));
		seg_classify<<<g ? g : 1, 256, 0, st>>>(a.off, a.nseg, a.numel, W, CAP, a.cls_lo,
			a.cls_hi, const_cast<u32*>(a.medium), const_cast<u32*>(a.large_ix),
			const_cast<u32*>(a.large), a.lbase,
			const_cast<u32*>(a.chunks),
			const_cast<uint2*>(a.wrange), a.meta);
		if (!L.done(err)) return false;
	}
	if constexpr (!KV) {
		if (msd) {
			// Keys alone: large and medium segments by MSD bucketing first.
			Launch L(q, "segsort_msd");
			seg_msd_persistent<K, IPT, false><<<grid, kRB, 0, st>>>(a);
			if (!L.done(err)) return false;
			a.fb_mode = 1;
		}
	}
	{
		Launch L(q, "segsort_persistent");
		seg_sort_persistent<K, KV, IPT><<<grid, kRB, 0, st>>>(a);
		if (!L.done(err)) return false;
	}
	return true;
}

// Key-value sort of 32-bit keys out of place: 64-bit composites on the
// bucketing kernel (tile IPT), then the key-value digit passes (tile 8) on
// the segments it left, sorted in place in the output.  a64 and a32 share
// the lists, the metadata and the scratch array.
template <int IPT>
bool seg_launch_comp(CCLQueue* q, const SegArgs<u64>& a64, const SegArgs<u32>& a32, u32 grid,
	GError** err) {
	constexpr u32 CAP = kRB * IPT;
	constexpr u32 W = CAP / 2;
	hipStream_t st = q->stream;
	{
		Launch L(q, "segsort_reset");
		seg_reset<<<1, 64, 0, st>>>(a64.meta);
		if (!L.done(err)) return false;
	}
	{
		Launch L(q, "segsort_classify");
		const u32 g = static_cast<u32>(div_ceil(a64.nseg, 256));
		seg_classify<<<g ? g : 1, 256, 0, st>>>(a64.off, a64.nseg, a64.numel, W, CAP, a64.cls_lo,
			a64.cls_hi, const_cast<u32*>(a64.medium), const_cast<u32*>(a64.large_ix),
			const_cast<u32*>(a64.large), a64.lbase, const_cast<u32*>(a64.chunks),
			const_cast<uint2*>(a64.wrange), a64.meta);
		if (!L.done(err)) return false;
	}
	{
		Launch L(q, "segsort_msd");
		seg_msd_persistent<u64, IPT, true><<<grid, kRB, 0, st>>>(a64);
		if (!L.done(err)) return false;
	}
	SegArgs<u32> b = a32;
	b.fb_mode = 1;
	Launch L(q, "segsort_persistent");
	seg_sort_persistent<u32, true, 8, true><<<grid, kRB, 0, st>>>(b);
	return L.done(err);
}

}  // namespace

struct clo_segsort_cache {
	SegData d;
};

// Entry point used by clo_sort_segmented_with_device_data (satradix sorter).
CCLEvent* clo_satradix_sort_segmented(CloSort* sorter, CCLQueue* cq, CCLBuffer* kin,
	CCLBuffer* vin, CCLBuffer* kout, CCLBuffer* vout, CCLBuffer* offsets, size_t nseg,
	size_t numel, u32 tile_items, u32 grid, void** cache, GError** err) {
	(void)sorter;
	if (!*cache) *cache = new clo_segsort_cache();
	SegData* sd = &static_cast<clo_segsort_cache*>(*cache)->d;
	const size_t ks = clo_type_sizeof(sorter->elem_type);
	if (numel == 0 || nseg == 0) return record_event(cq, "segsort_done", err);
	if (tile_items != 8 && tile_items != 16 && tile_items != 32) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "segmented tile items must be 8, 16 or 32");
		return nullptr;
	}
	const bool kv = vin != nullptr;
	const bool inplace = kout == nullptr || kout->ptr == kin->ptr;
	// Key-value sorts of 32-bit keys take the bucketing kernel on composites
	// (key << 32 | position), tile 8 unless a segtile option asks for 16.  The
	// values are gathered from their input positions when the output is
	// written; in place they are first copied to the value scratch array, so
	// that the gather reads values no output has overwritten (a segment's keys
	// are all read before its output is written).
	const bool comp = kv && ks == 4 && sorter->seg_msd != 0 &&
		(sorter->seg_tile == 0 || sorter->seg_tile == 8 || sorter->seg_tile == 16);
	if (comp) tile_items = sorter->seg_tile ? sorter->seg_tile : 8;
	// medium and large lists (nseg entries each), large lists by length
	// class, chunk list (nseg entries), then [j0, j1) per window of the
	// smallest tile's W = 1024 elements and the fallback list (nseg entries).
	const size_t nwin = numel / 1024 + 2;
	SegClassBase lbase;
	const size_t nlarge = seg_class_layout(nseg, numel, 256 * static_cast<size_t>(tile_items), &lbase);
	const size_t nlists = (3 * nseg + nlarge + 3) / 4 * 4 + 4;
	if (!sd->lists.ensure(nlists * sizeof(u32) + (nwin + nseg) * sizeof(uint2), err) ||
		!sd->meta.ensure(sizeof(SegMeta), err) ||
		!sd->kscr.ensure(numel * (comp ? 8 : ks), err) || (vin && !sd->vscr.ensure(numel * 4, err)))
		return nullptr;
	if (grid == 0) {
		int dev = cq->ctx->dev, cus = 0;
		if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
			cus <= 0)
			cus = 256;
		grid = static_cast<u32>(cus) * ((vin && ks == 8) ? kSegBlocksPerCuKv64 : kSegBlocksPerCuDefault);
	}
	u32* lists = static_cast<u32*>(sd->lists.ptr);
	CLO_LOG(1, "segmented sort: nseg=%zu numel=%zu key=%zu B%s tile=256x%u (CAP %u) grid=%u "
		"classes=%s%s", nseg, numel, ks, kv ? " +u32 values" : "", tile_items, 256 * tile_items,
		grid, (sorter->seg_cls[0] | sorter->seg_cls[1]) ? "per length bin" : "by length",
		comp ? " composite keys" : "");
	// MSD bucketing for keys alone (32768 segments of log-uniform lengths up to
	// 65536, tile 8: 32-bit keys 1.6 ms against 2.6 on the digit passes, 64-bit
	// keys 2.2 against 7.5).
	const bool msd = !kv && sorter->seg_msd != 0;
#define SEG_ARGS(K)                                                                          \
	SegArgs<K> a;                                                                            \
	a.kin = static_cast<const K*>(kin->ptr);                                                 \
	a.vin = kv ? static_cast<const u32*>(vin->ptr) : nullptr;                                \
	a.kout = inplace ? static_cast<K*>(kin->ptr) : static_cast<K*>(kout->ptr);                \
	a.vout = kv ? (inplace ? static_cast<u32*>(vin->ptr) : static_cast<u32*>(vout->ptr))     \
		: nullptr;                                                                           \
	a.kin32 = nullptr;                                                                       \
	a.kout32 = nullptr;                                                                      \
	a.kscr = static_cast<K*>(sd->kscr.ptr);                                                  \
	a.vscr = kv ? static_cast<u32*>(sd->vscr.ptr) : nullptr;                                  \
	a.off = static_cast<const u32*>(offsets->ptr);                                           \
	a.medium = lists;                                                                        \
	a.large_ix = lists + nseg;                                                               \
	a.large = lists + 2 * nseg;                                                              \
	a.lbase = lbase;                                                                         \
	a.chunks = lists + 2 * nseg + nlarge;                                                    \
	a.wrange = reinterpret_cast<const uint2*>(lists + nlists);                               \
	a.fb = reinterpret_cast<uint2*>(lists + nlists) + nwin;                                  \
	a.fb_mode = 0;                                                                           \
	a.meta = static_cast<SegMeta*>(sd->meta.ptr);                                            \
	a.cls_lo = sorter->seg_cls[0];                                                           \
	a.cls_hi = sorter->seg_cls[1];                                                           \
	a.nseg = static_cast<u32>(nseg);                                                         \
	a.numel = static_cast<u32>(numel);                                                       \

	bool ok = false;
	if (comp) {
		SEG_ARGS(u32)
		SegArgs<u64> c = {};
		c.vin = a.vin;
		c.vout = a.vout;
		if (inplace) {
			if (hipMemcpyAsync(a.vscr, a.vin, numel * sizeof(u32), hipMemcpyDeviceToDevice,
					cq->stream) != hipSuccess) {
				set_error(err, kQuarkClo, CLO_ERROR_LIBRARY, "segmented sort: value copy failed");
				return nullptr;
			}
			c.vin = a.vscr;
		}
		c.kin32 = a.kin;
		c.kout32 = a.kout;
		c.kscr = static_cast<u64*>(sd->kscr.ptr);
		c.off = a.off;
		c.medium = a.medium;
		c.large_ix = a.large_ix;
		c.large = a.large;
		c.lbase = a.lbase;
		c.chunks = a.chunks;
		c.wrange = a.wrange;
		c.fb = a.fb;
		c.fb_mode = 0;
		c.meta = a.meta;
		c.cls_lo = a.cls_lo;
		c.cls_hi = a.cls_hi;
		c.nseg = a.nseg;
		c.numel = a.numel;
		ok = tile_items == 16 ? seg_launch_comp<16>(cq, c, a, grid, err)
			: seg_launch_comp<8>(cq, c, a, grid, err);
	} else if (ks == 4) {
		SEG_ARGS(u32)
		switch (tile_items) {
		case 8: ok = kv ? seg_launch<u32, true, 8>(cq, sd, a, grid, msd, err) : seg_launch<u32, false, 8>(cq, sd, a, grid, msd, err); break;
		case 16: ok = kv ? seg_launch<u32, true, 16>(cq, sd, a, grid, msd, err) : seg_launch<u32, false, 16>(cq, sd, a, grid, msd, err); break;
		case 32: ok = kv ? seg_launch<u32, true, 32>(cq, sd, a, grid, msd, err) : seg_launch<u32, false, 32>(cq, sd, a, grid, msd, err); break;
		default:
			set_error(err, kQuarkClo, CLO_ERROR_ARGS, "segmented tile items must be 8, 16 or 32");
			return nullptr;
		}
	} else if (ks == 8) {
		SEG_ARGS(u64)
		switch (tile_items) {
		case 8: ok = kv ? seg_launch<u64, true, 8>(cq, sd, a, grid, msd, err) : seg_launch<u64, false, 8>(cq, sd, a, grid, msd, err); break;
		case 16: ok = kv ? seg_launch<u64, true, 16>(cq, sd, a, grid, msd, err) : seg_launch<u64, false, 16>(cq, sd, a, grid, msd, err); break;
		default:
			set_error(err, kQuarkClo, CLO_ERROR_ARGS, "segmented tile items for 64-bit keys must be 8 or 16");
			return nullptr;
		}
	} else {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "segmented sort takes 32- or 64-bit keys");
		return nullptr;
	}
#undef SEG_ARGS
	if (!ok) return nullptr;
	return record_event(cq, "segsort_done", err);
}

void clo_satradix_segsort_cache_free(void* cache) {
	delete static_cast<clo_segsort_cache*>(cache);
}

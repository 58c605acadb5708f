// satradix: stable LSD radix sort for gfx950.
//
// Result contract (CL_Ops' src/cl_ops/sort/clo_sort_satradix.c:117-335 and
// clo_sort_satradix.cl:34-258): a stable sort of the elements by the unsigned
// value of their bits, ascending, whatever the compare snippet (the reference
// kernels never use CLO_SORT_COMPARE).  Key-value pairs keep the payloads of
// equal keys in input order.
//
// Kernel plan per call (all on cq_exec, no host synchronisation):
//   satradix_histogram  one read of the keys -> digit counts of every pass
//   satradix_offsets    one block: exclusive digit starts per pass, which
//                       passes are trivial (one digit holds every key), and
//                       the source/destination buffer of every executed pass
//   satradix_scatter    per digit pass: tile load, wave64 ballot ranking,
//                       decoupled look-back of per-digit tile prefixes,
//                       LDS-staged scatter (one read + one write of the data)
//   satradix_copy       only when an in-place sort executed an odd number of
//                       passes (the result sits in the scratch buffer)

#include <cstring>
#include <string>

#include "clo_internal.hpp"
#include "dev_common.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

constexpr int kRadixMax = 256;
constexpr int kBlock = 256;  // == kRadixMax: one thread per digit bin
constexpr int kWaves = kBlock / kWave;
constexpr int kMaxPasses = 64;
constexpr int kHistBlock = 256;
constexpr int kHistLds = 2048;  // max passes * bins over all digit widths

enum BufSel : u32 { kIn = 0, kAux = 1, kOut = 2 };

struct PassDesc {
	u32 active;
	u32 shift;
	u32 src;
	u32 dst;
};

struct SortPlan {
	PassDesc pass[kMaxPasses];
	u32 copy_active;
	u32 copy_src;
	u32 copy_dst;
	u32 executed;
};

// Device-side metadata block of one sorter.
struct Meta {
	SortPlan plan;
	u32 tickets[kMaxPasses];
	u32 hist[kMaxPasses * kRadixMax];  // zero between calls (offsets clears it)
	u32 base[kMaxPasses * kRadixMax];
};

template <typename K>
struct Bufs {
	K* k[3];
	u32* v[3];
};

// ---------------------------------------------------------------------------
// Histogram of every digit position, one read of the keys.
template <typename K>
__global__ __launch_bounds__(kHistBlock) void satradix_histogram(const K* __restrict__ keys,
	u32 n, u32* __restrict__ hist, u32 passes, u32 bits) {
	__shared__ u32 s_h[kHistLds];
	const u32 nbins = 1u << bits;
	const u32 mask = nbins - 1u;
	const u32 total = passes * nbins;
	for (u32 i = threadIdx.x; i < total; i += kHistBlock) s_h[i] = 0;
	__syncthreads();
	const u32 stride = gridDim.x * kHistBlock;
	for (u32 i = blockIdx.x * kHistBlock + threadIdx.x; i < n; i += stride) {
		K k = keys[i];
		for (u32 p = 0; p < passes; ++p) {
			u32 d = static_cast<u32>(k >> (p * bits)) & mask;
			atomicAdd(&s_h[p * nbins + d], 1u);
		}
	}
	__syncthreads();
	for (u32 i = threadIdx.x; i < total; i += kHistBlock) {
		u32 c = s_h[i];
		if (c) atomicAdd(&hist[(i / nbins) * kRadixMax + (i % nbins)], c);
	}
}

// ---------------------------------------------------------------------------
// Digit starts per pass + the pass plan.  One block of kBlock threads.
__global__ __launch_bounds__(kBlock) void satradix_offsets(Meta* __restrict__ meta, u32 n,
	u32 passes, u32 bits, u32 in_place) {
	__shared__ u32 s_wsum[kWaves];
	__shared__ u32 s_trivial[kMaxPasses];
	const u32 t = threadIdx.x;
	const u32 nbins = 1u << bits;
	for (u32 p = 0; p < passes; ++p) {
		u32 c = 0;
		if (t < nbins) {
			c = meta->hist[p * kRadixMax + t];
			meta->hist[p * kRadixMax + t] = 0;  // ready for the next call
		}
		u32 excl = block_exclusive_sum<kBlock>(c, s_wsum, static_cast<u32*>(nullptr));
		if (t < nbins) meta->base[p * kRadixMax + t] = excl;
		int triv = __syncthreads_or(c == n ? 1 : 0);
		if (t == 0) s_trivial[p] = triv;
		if (t == 0) meta->tickets[p] = 0;
	}
	__syncthreads();
	if (t != 0) return;
	u32 executed = 0;
	for (u32 p = 0; p < passes; ++p) executed += s_trivial[p] ? 0u : 1u;
	SortPlan& pl = meta->plan;
	u32 k = 0;
	u32 cur = kIn;
	for (u32 p = 0; p < kMaxPasses; ++p) {
		PassDesc d = {0u, 0u, 0u, 0u};
		if (p < passes && !s_trivial[p]) {
			u32 dst;
			if (in_place) {
				dst = (k % 2 == 0) ? kAux : kIn;
			} else {
				// The last executed pass must land in kOut.
				dst = ((executed - 1 - k) % 2 == 0) ? kOut : kAux;
			}
			d.active = 1;
			d.shift = p * bits;
			d.src = cur;
			d.dst = dst;
			cur = dst;
			++k;
		}
		pl.pass[p] = d;
	}
	pl.executed = executed;
	pl.copy_active = 0;
	pl.copy_src = 0;
	pl.copy_dst = 0;
	if (in_place && (executed % 2 == 1)) {
		pl.copy_active = 1;
		pl.copy_src = kAux;
		pl.copy_dst = kIn;
	} else if (!in_place && executed == 0) {
		pl.copy_active = 1;
		pl.copy_src = kIn;
		pl.copy_dst = kOut;
	}
}

// ---------------------------------------------------------------------------
// One digit pass: stable ranking of a tile of kBlock*ITEMS elements with wave
// ballots, decoupled look-back for the tile's per-digit global offsets, and a
// scatter staged through LDS so that global writes are contiguous runs.
//
// Element order inside a tile is wave-striped: wave w owns elements
// [w*64*ITEMS, (w+1)*64*ITEMS), item j of lane l is element w*64*ITEMS+j*64+l.
// Processing items j = 0.. in order with per-wave running digit counts gives
// each element the number of equal-digit elements before it (stability).
template <typename K, bool KV, int ITEMS>
__global__ __launch_bounds__(kBlock) void satradix_scatter(Bufs<K> b,
	const Meta* __restrict__ meta, u32* __restrict__ tickets, u64* __restrict__ status,
	u32 pass, u32 n, u32 bits, u32 epoch) {
	constexpr u32 kTile = kBlock * ITEMS;
	constexpr u32 kStageBytes = kTile * (sizeof(K) > 4 ? sizeof(K) : (KV ? 4 : sizeof(K)));
	__shared__ __attribute__((aligned(16))) unsigned char s_stage[kStageBytes];
	__shared__ u32 s_whist[kWaves][kRadixMax];
	__shared__ u32 s_tstart[kRadixMax];
	__shared__ u32 s_gofs[kRadixMax];
	__shared__ u32 s_wsum[kWaves];
	__shared__ u32 s_ticket;

	const PassDesc d = meta->plan.pass[pass];
	if (!d.active) return;  // uniform: trivial digit, nothing moves
	const u32 tid = threadIdx.x;
	if (tid == 0) s_ticket = atomicAdd(&tickets[pass], 1u);
	const u32 lane = lane_id();
	const u32 w = wave_id();
	for (u32 i = lane; i < kRadixMax; i += kWave) s_whist[w][i] = 0;
	__syncthreads();
	const u32 tile = s_ticket;
	const u32 tile_base = tile * kTile;
	const u32 nbins = 1u << bits;
	const u32 mask = nbins - 1u;
	const u32 shift = d.shift;
	const K* __restrict__ src = b.k[d.src];
	K* __restrict__ dst = b.k[d.dst];

	// Load (coalesced: 64 consecutive elements per wave instruction).
	const u32 wbase = tile_base + w * kWave * ITEMS;
	K k[ITEMS];
	u32 v[ITEMS];
#pragma unroll
	for (int j = 0; j < ITEMS; ++j) {
		u32 idx = wbase + j * kWave + lane;
		k[j] = idx < n ? src[idx] : K(0);
	}
	if (KV) {
		const u32* __restrict__ vsrc = b.v[d.src];
#pragma unroll
		for (int j = 0; j < ITEMS; ++j) {
			u32 idx = wbase + j * kWave + lane;
			v[j] = idx < n ? vsrc[idx] : 0u;
		}
	}

	// Rank inside the wave.
	u32 rank[ITEMS];
#pragma unroll
	for (int j = 0; j < ITEMS; ++j) {
		const bool valid = (wbase + j * kWave + lane) < n;
		const u32 dig = static_cast<u32>(k[j] >> shift) & mask;
		u64 peers = ballot(valid);
		for (u32 bt = 0; bt < bits; ++bt) {
			const bool bit = (dig >> bt) & 1u;
			const u64 bb = ballot(bit);
			peers &= bit ? bb : ~bb;
		}
		const u32 below = count_below(peers);
		u32 base = 0;
		if (valid) {
			base = s_whist[w][dig];
			if (below == 0) s_whist[w][dig] = base + static_cast<u32>(__popcll(peers));
		}
		rank[j] = base + below;
	}
	__syncthreads();

	// Per-digit tile totals, per-wave exclusive offsets, aggregate publish.
	u32 total = 0;
	{
		u32 run = 0;
#pragma unroll
		for (int ww = 0; ww < kWaves; ++ww) {
			u32 c = s_whist[ww][tid];
			s_whist[ww][tid] = run;
			run += c;
		}
		total = run;
		if (tid < nbins)
			status_store(&status[static_cast<size_t>(tile) * kRadixMax + tid], epoch,
				tile == 0 ? kFlagPrefix : kFlagAggregate, total);
	}
	const u32 tstart = block_exclusive_sum<kBlock>(total, s_wsum, static_cast<u32*>(nullptr));
	s_tstart[tid] = tstart;

	// Decoupled look-back: exclusive prefix of this digit over earlier tiles.
	u32 excl = 0;
	if (tile > 0 && tid < nbins) {
		u32 pred = tile - 1;
		for (;;) {
			const u64 word = status_load(&status[static_cast<size_t>(pred) * kRadixMax + tid]);
			const u32 f = status_flag(word, epoch);
			if (f == 0) {
				__builtin_amdgcn_s_sleep(1);
				continue;
			}
			excl += static_cast<u32>(word);
			if (f == kFlagPrefix) break;
			--pred;
		}
		status_store(&status[static_cast<size_t>(tile) * kRadixMax + tid], epoch, kFlagPrefix,
			excl + total);
	}
	if (tid < nbins) s_gofs[tid] = meta->base[pass * kRadixMax + tid] + excl - tstart;
	__syncthreads();

	// Stage keys in digit order in LDS.
	K* s_keys = reinterpret_cast<K*>(s_stage);
#pragma unroll
	for (int j = 0; j < ITEMS; ++j) {
		const bool valid = (wbase + j * kWave + lane) < n;
		const u32 dig = static_cast<u32>(k[j] >> shift) & mask;
		if (valid) {
			rank[j] += s_tstart[dig] + s_whist[w][dig];
			s_keys[rank[j]] = k[j];
		}
	}
	__syncthreads();

	const u32 count = (n - tile_base) < kTile ? (n - tile_base) : kTile;
	u32 oidx[ITEMS];
#pragma unroll
	for (int r = 0; r < ITEMS; ++r) {
		const u32 i = r * kBlock + tid;
		if (i < count) {
			const K key = s_keys[i];
			const u32 dig = static_cast<u32>(key >> shift) & mask;
			const u32 o = s_gofs[dig] + i;
			dst[o] = key;
			oidx[r] = o;
		}
	}
	if (KV) {
		u32* s_vals = reinterpret_cast<u32*>(s_stage);
		u32* __restrict__ vdst = b.v[d.dst];
		__syncthreads();
#pragma unroll
		for (int j = 0; j < ITEMS; ++j) {
			const bool valid = (wbase + j * kWave + lane) < n;
			if (valid) s_vals[rank[j]] = v[j];
		}
		__syncthreads();
#pragma unroll
		for (int r = 0; r < ITEMS; ++r) {
			const u32 i = r * kBlock + tid;
			if (i < count) vdst[oidx[r]] = s_vals[i];
		}
	}
}

// ---------------------------------------------------------------------------
template <typename K, bool KV>
__global__ __launch_bounds__(256) void satradix_copy(Bufs<K> b, const Meta* __restrict__ meta,
	u32 n) {
	const SortPlan& pl = meta->plan;
	if (!pl.copy_active) return;
	const K* __restrict__ ks = b.k[pl.copy_src];
	K* __restrict__ kd = b.k[pl.copy_dst];
	const u32* __restrict__ vs = b.v[pl.copy_src];
	u32* __restrict__ vd = b.v[pl.copy_dst];
	const u32 stride = gridDim.x * blockDim.x;
	for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		kd[i] = ks[i];
		if (KV) vd[i] = vs[i];
	}
}

// ---------------------------------------------------------------------------
// Host side.

constexpr int kItemsDefault = 16;

struct SatradixData {
	u32 bits = 8;
	std::string scan_type = "blelloch";
	Scratch aux_keys;
	Scratch aux_vals;
	Scratch status;
	Scratch meta;
	u32 epoch = 0;
	CCLQueue* last_queue = nullptr;
	SatradixData() {
		status.zero_on_alloc = true;
		meta.zero_on_alloc = true;
	}
};

inline SatradixData* sdata(CloSort* s) { return static_cast<SatradixData*>(s->data); }

template <typename K, bool KV, int ITEMS>
bool launch_all(SatradixData* sd, CCLQueue* q, K* kin, u32* vin, K* kout, u32* vout, size_t n,
	GError** err) {
	const u32 keybits = sizeof(K) * 8;
	const u32 bits = sd->bits;
	const u32 passes = (keybits + bits - 1) / bits;
	const bool in_place = (kout == nullptr || kout == kin);
	const u32 tile = kBlock * ITEMS;
	const size_t tiles = div_ceil(n, tile);
	if (!sd->aux_keys.ensure(n * sizeof(K), err)) return false;
	if (KV && !sd->aux_vals.ensure(n * sizeof(u32), err)) return false;
	if (!sd->status.ensure(tiles * kRadixMax * sizeof(u64), err)) return false;
	if (!sd->meta.ensure(sizeof(Meta), err)) return false;
	Meta* meta = static_cast<Meta*>(sd->meta.ptr);
	u64* status = static_cast<u64*>(sd->status.ptr);

	Bufs<K> b;
	b.k[kIn] = kin;
	b.k[kAux] = static_cast<K*>(sd->aux_keys.ptr);
	b.k[kOut] = in_place ? kin : kout;
	b.v[kIn] = vin;
	b.v[kAux] = KV ? static_cast<u32*>(sd->aux_vals.ptr) : nullptr;
	b.v[kOut] = KV ? (in_place ? vin : vout) : nullptr;
	const u32 nn = static_cast<u32>(n);
	hipStream_t st = q->stream;

	{
		Launch L(q, CLO_SORT_SATRADIX_KNAME_HISTOGRAM);
		u32 grid = static_cast<u32>(div_ceil(n, kHistBlock * 16));
		if (grid > 2048) grid = 2048;
		if (grid == 0) grid = 1;
		satradix_histogram<K><<<grid, kHistBlock, 0, st>>>(kin, nn, meta->hist, passes, bits);
		if (!L.done(err)) return false;
	}
	{
		Launch L(q, CLO_SORT_SATRADIX_KNAME_OFFSETS);
		satradix_offsets<<<1, kBlock, 0, st>>>(meta, nn, passes, bits, in_place ? 1u : 0u);
		if (!L.done(err)) return false;
	}
	for (u32 p = 0; p < passes; ++p) {
		if (++sd->epoch >= (1u << 30)) {
			// Epoch space exhausted: clear the status words once and restart.
			if (hip_failed(hipStreamSynchronize(st), err, "hipStreamSynchronize") ||
				hip_failed(hipMemset(sd->status.ptr, 0, sd->status.size), err, "hipMemset"))
				return false;
			sd->epoch = 1;
		}
		Launch L(q, CLO_SORT_SATRADIX_KNAME_SCATTER);
		satradix_scatter<K, KV, ITEMS><<<static_cast<u32>(tiles), kBlock, 0, st>>>(b, meta,
			meta->tickets, status, p, nn, bits, sd->epoch);
		if (!L.done(err)) return false;
	}
	{
		Launch L(q, CLO_SORT_SATRADIX_KNAME_COPY);
		u32 grid = static_cast<u32>(div_ceil(n, 256 * 8));
		if (grid > 4096) grid = 4096;
		satradix_copy<K, KV><<<grid, 256, 0, st>>>(b, meta, nn);
		if (!L.done(err)) return false;
	}
	sd->last_queue = q;
	return true;
}

template <bool KV>
bool dispatch(SatradixData* sd, CCLQueue* q, size_t ksize, void* kin, u32* vin, void* kout,
	u32* vout, size_t n, GError** err) {
	switch (ksize) {
	case 1:
		return launch_all<u8, KV, kItemsDefault>(sd, q, static_cast<u8*>(kin), vin,
			static_cast<u8*>(kout), vout, n, err);
	case 2:
		return launch_all<u16, KV, kItemsDefault>(sd, q, static_cast<u16*>(kin), vin,
			static_cast<u16*>(kout), vout, n, err);
	case 4:
		return launch_all<u32, KV, kItemsDefault>(sd, q, static_cast<u32*>(kin), vin,
			static_cast<u32*>(kout), vout, n, err);
	case 8:
		return launch_all<u64, KV, kItemsDefault>(sd, q, static_cast<u64*>(kin), vin,
			static_cast<u64*>(kout), vout, n, err);
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
	for (auto& o : opts) {
		if (o.key == "radix") {
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
	return "satradix (gfx950: histogram + offsets + onesweep scatter)";
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
	switch (i) {
	case CLO_SORT_SATRADIX_KIDX_HISTOGRAM:
		return kHistLds * 4;
	case CLO_SORT_SATRADIX_KIDX_OFFSETS:
		return (kWaves + kMaxPasses) * 4;
	case CLO_SORT_SATRADIX_KIDX_SCATTER:
		return kBlock * kItemsDefault * (ks > 4 ? ks : 4) + kWaves * kRadixMax * 4 +
			2 * kRadixMax * 4 + kWaves * 4 + 4;
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

cl_uint clo_satradix_last_passes(CloSort* sorter, GError** err) {
	SatradixData* sd = sdata(sorter);
	if (!sd || !sd->meta.ptr || !sd->last_queue) return 0;
	u32 executed = 0;
	Meta* meta = static_cast<Meta*>(sd->meta.ptr);
	if (hip_failed(hipMemcpyAsync(&executed, &meta->plan.executed, sizeof(u32),
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


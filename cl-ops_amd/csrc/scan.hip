// Scanner factory (CL_Ops' src/cl_ops/scan/clo_scan_abstract.c:74-567) and
// the "blelloch" exclusive scan as one single-pass decoupled look-back kernel
// (Merrill & Garland, "Single-pass Parallel Prefix Scan with Decoupled
// Look-back", NVIDIA TR 2016) for gfx950.
//
// Result contract (clo_scan_blelloch.cl:49-211): out[i] = sum_{j<i} (SUM)in[j]
// with the element converted to the sum type as C does (sign extension for
// signed elements) and the sum taken modulo 2^(8*sizeof(SUM)).
//
// One launch reads the input once and writes the output once.  Tiles take a
// ticket from an atomic counter, so a tile's predecessors are always running
// or done; the tile's aggregate and inclusive prefix are published in 64-bit
// status words tagged with a per-launch epoch (no clearing between launches).

#include <cstring>

#include "clo_internal.hpp"
#include "dev_common.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr u32 kScanTile = kScanBlock * kScanItems;

__host__ __device__ constexpr u32 pad(u32 i) { return i + (i >> 5); }

// Status of 32-bit sums: one word per tile.  64-bit sums: two words (low and
// high halves) with the same flag; a reader accepts them only when both carry
// the current epoch and the same flag.
template <typename Acc>
struct Status;

template <>
struct Status<u32> {
	static constexpr int kWords = 1;
	__device__ static void store(u64* s, u32 tile, u32 epoch, u32 flag, u32 v) {
		status_store(&s[tile], epoch, flag, v);
	}
	__device__ static u32 load(const u64* s, u32 tile, u32 epoch, u32* v) {
		u64 w = status_load(&s[tile]);
		*v = static_cast<u32>(w);
		return status_flag(w, epoch);
	}
};

template <>
struct Status<u64> {
	static constexpr int kWords = 2;
	__device__ static void store(u64* s, u32 tile, u32 epoch, u32 flag, u64 v) {
		status_store(&s[2 * static_cast<size_t>(tile)], epoch, flag, static_cast<u32>(v));
		status_store(&s[2 * static_cast<size_t>(tile) + 1], epoch, flag,
			static_cast<u32>(v >> 32));
	}
	__device__ static u32 load(const u64* s, u32 tile, u32 epoch, u64* v) {
		u64 lo = status_load(&s[2 * static_cast<size_t>(tile)]);
		u64 hi = status_load(&s[2 * static_cast<size_t>(tile) + 1]);
		u32 f = status_flag(lo, epoch);
		if (f != status_flag(hi, epoch)) return 0;
		*v = (hi << 32) | (lo & 0xFFFFFFFFull);
		return f;
	}
};

// Exclusive prefix of this tile over all earlier tiles, computed by wave 0:
// each round inspects the 64 predecessors below `hi` at once.
template <typename Acc>
__device__ Acc lookback(const u64* status, u32 tile, u32 epoch) {
	const u32 lane = lane_id();
	Acc excl = 0;
	int64_t top = static_cast<int64_t>(tile) - 1;
	for (;;) {
		const int64_t idx = top - lane;
		Acc v = 0;
		u32 f = kFlagPrefix;  // below tile 0: an empty prefix
		if (idx >= 0) f = Status<Acc>::load(status, static_cast<u32>(idx), epoch, &v);
		const u64 pmask = ballot(f == kFlagPrefix);
		const u32 lim = pmask ? static_cast<u32>(__builtin_ctzll(pmask)) : 63u;
		const u64 need = lim >= 63 ? ~0ull : ((2ull << lim) - 1ull);
		if (ballot(f == 0) & need) {
			__builtin_amdgcn_s_sleep(1);
			continue;
		}
		excl += wave_sum(lane <= lim ? v : Acc(0));
		if (pmask) return excl;
		top -= kWave;
	}
}

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

// Exclusive block scan of (flag, value) pairs; *total = all threads combined.
template <typename Acc>
__device__ SegPair<Acc> block_exclusive_segscan(SegPair<Acc> mine, SegPair<Acc>* s_w,
	SegPair<Acc>* total) {
	constexpr int W = kScanBlock / kWave;
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

// SEG = false: plain exclusive scan.  SEG = true: the sum restarts at every
// segment start off[j] (the per-segment exclusive scans of a batch).
template <typename In, typename Acc, int OUTW, bool SEG>
__global__ __launch_bounds__(kScanBlock) void blelloch_lookback_scan(const In* __restrict__ in,
	void* __restrict__ out, u64 n, u64* __restrict__ status, u32* __restrict__ ticket,
	u32 ticket_base, u32 epoch, const u32* __restrict__ off, u32 nseg) {
	__shared__ Acc s_t[pad(kScanTile) + 1];
	__shared__ Acc s_wsum[kScanBlock / kWave];
	__shared__ SegPair<Acc> s_pw[kScanBlock / kWave];
	__shared__ u32 s_head[SEG ? kScanTile / 32 : 1];
	__shared__ Acc s_prefix;
	__shared__ u32 s_tile;
	__shared__ u32 s_j0;
	const u32 tid = threadIdx.x;
	if (tid == 0) s_tile = atomicAdd(ticket, 1u) - ticket_base;
	if (SEG)
		for (u32 i = tid; i < kScanTile / 32; i += kScanBlock) s_head[i] = 0;
	__syncthreads();
	const u32 tile = s_tile;
	const u64 base = static_cast<u64>(tile) * kScanTile;
	if (SEG) {
		if (tid == 0) s_j0 = first_at_or_after(off, nseg, base);
		__syncthreads();
		for (u32 j = s_j0 + tid; j < nseg; j += kScanBlock) {
			const u64 o = off[j];
			if (o >= base + kScanTile) break;
			const u32 p = static_cast<u32>(o - base);
			atomicOr(&s_head[p >> 5], 1u << (p & 31));
		}
	}

	// Striped (coalesced) load, converted to the sum type.
#pragma unroll
	for (int j = 0; j < kScanItems; ++j) {
		const u32 i = j * kScanBlock + tid;
		const u64 g = base + i;
		s_t[pad(i)] = g < n ? static_cast<Acc>(in[g]) : Acc(0);
	}
	__syncthreads();

	// Blocked: thread tid owns tile elements [tid*ITEMS, (tid+1)*ITEMS).
	Acc x[kScanItems];
	u32 heads = 0;  // bit j: item j starts a segment
	Acc run = 0;
	u32 any_head = 0;
#pragma unroll
	for (int j = 0; j < kScanItems; ++j) {
		x[j] = s_t[pad(tid * kScanItems + j)];
		if (SEG) {
			const u32 p = tid * kScanItems + j;
			if ((s_head[p >> 5] >> (p & 31)) & 1u) {
				heads |= 1u << j;
				run = 0;
				any_head = 1;
			}
		}
		run += x[j];
	}
	Acc tile_total;
	Acc tpre;
	u32 carry_known = 0;  // SEG: a segment starts before this thread's items in the tile
	u32 tile_head = 0;
	if (SEG) {
		SegPair<Acc> tot;
		const SegPair<Acc> ex = block_exclusive_segscan<Acc>({any_head, run}, s_pw, &tot);
		tpre = ex.v;
		carry_known = ex.f;
		tile_total = tot.v;
		tile_head = tot.f;
	} else {
		tpre = block_exclusive_sum<kScanBlock>(run, s_wsum, &tile_total);
	}

	if (tile == 0) {
		if (tid == 0) {
			Status<Acc>::store(status, 0, epoch, kFlagPrefix, tile_total);
			s_prefix = 0;
		}
	} else if (tid < kWave) {
		// A tile containing a segment start knows its carry-out (the sum after
		// its last start) without looking back: publish it as a prefix.
		if (tid == 0)
			Status<Acc>::store(status, tile, epoch, tile_head ? kFlagPrefix : kFlagAggregate,
				tile_total);
		const Acc excl = lookback<Acc>(status, tile, epoch);
		if (tid == 0) {
			if (!tile_head) Status<Acc>::store(status, tile, epoch, kFlagPrefix, excl + tile_total);
			s_prefix = excl;
		}
	}
	__syncthreads();

	Acc acc = carry_known ? tpre : s_prefix + tpre;
#pragma unroll
	for (int j = 0; j < kScanItems; ++j) {
		if (SEG && ((heads >> j) & 1u)) acc = 0;
		s_t[pad(tid * kScanItems + j)] = acc;
		acc += x[j];
	}
	__syncthreads();

#pragma unroll
	for (int j = 0; j < kScanItems; ++j) {
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

struct BlellochData {
	Scratch status;
	Scratch meta;  // the ticket counter; never reallocated after creation
	u32 epoch = 0;
	u32 issued = 0;
	// Set when a launch failed after tickets were reserved: the counter no
	// longer matches `issued` and is cleared before the next launch.
	bool dirty = false;
	BlellochData() {
		status.zero_on_alloc = true;
		meta.zero_on_alloc = true;
	}
};

struct ScanLaunch {
	const void* in;
	void* out;
	u64 n;
	u64* status;
	u32* ticket;
	u32 base;
	u32 epoch;
	u32 tiles;
	const u32* off;  // segment starts (segmented scan) or nullptr
	u32 nseg;
	hipStream_t st;
};

template <typename In, typename Acc, int OUTW>
void launch_scan(const ScanLaunch& a) {
	if (a.off)
		blelloch_lookback_scan<In, Acc, OUTW, true><<<a.tiles, kScanBlock, 0, a.st>>>(
			static_cast<const In*>(a.in), a.out, a.n, a.status, a.ticket, a.base, a.epoch, a.off,
			a.nseg);
	else
		blelloch_lookback_scan<In, Acc, OUTW, false><<<a.tiles, kScanBlock, 0, a.st>>>(
			static_cast<const In*>(a.in), a.out, a.n, a.status, a.ticket, a.base, a.epoch, nullptr,
			0);
}

template <typename In>
bool launch_for_in(size_t sum_size, const ScanLaunch& a) {
	switch (sum_size) {
	case 1: launch_scan<In, u32, 1>(a); return true;
	case 2: launch_scan<In, u32, 2>(a); return true;
	case 4: launch_scan<In, u32, 4>(a); return true;
	case 8: launch_scan<In, u64, 8>(a); return true;
	}
	return false;
}

const char* blelloch_init(CloScan* scanner, const char* options, GError** err) {
	// clo_scan_blelloch.c:43-45: any option is an error.
	if (options && strlen(options) > 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "Invalid options for blelloch scan.");
		return nullptr;
	}
	scanner->data = new BlellochData();
	return "blelloch (gfx950: single-pass decoupled look-back)";
}

void blelloch_finalize(CloScan* scanner) {
	delete static_cast<BlellochData*>(scanner->data);
	scanner->data = nullptr;
}

CCLEvent* scan_common(CloScan* scanner, CCLQueue* cq_exec, CCLBuffer* data_in,
	CCLBuffer* data_out, size_t numel, const u32* off, u32 nseg, GError** err) {
	BlellochData* bd = static_cast<BlellochData*>(scanner->data);
	if (numel == 0) return record_event(cq_exec, "blelloch_done", err);
	const size_t tiles = div_ceil(numel, kScanTile);
	if (tiles > 0x7FFFFFFFull) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "numel too large");
		return nullptr;
	}
	const size_t words = Status<u64>::kWords;
	if (!bd->status.ensure(tiles * words * sizeof(u64), err)) return nullptr;
	if (!bd->meta.ensure(256, err)) return nullptr;
	if (bd->dirty) {
		if (hip_failed(hipStreamSynchronize(cq_exec->stream), err, "hipStreamSynchronize") ||
			hip_failed(hipMemset(bd->meta.ptr, 0, bd->meta.size), err, "hipMemset"))
			return nullptr;
		bd->issued = 0;
		bd->dirty = false;
	}
	if (++bd->epoch >= (1u << 30)) {
		if (hip_failed(hipStreamSynchronize(cq_exec->stream), err, "hipStreamSynchronize") ||
			hip_failed(hipMemset(bd->status.ptr, 0, bd->status.size), err, "hipMemset"))
			return nullptr;
		bd->epoch = 1;
	}
	ScanLaunch a;
	a.base = bd->issued;
	bd->issued += static_cast<u32>(tiles);
	a.status = static_cast<u64*>(bd->status.ptr);
	a.ticket = static_cast<u32*>(bd->meta.ptr);
	a.in = data_in->ptr;
	a.out = data_out->ptr;
	a.n = numel;
	a.epoch = bd->epoch;
	a.tiles = static_cast<u32>(tiles);
	a.off = off;
	a.nseg = nseg;
	a.st = cq_exec->stream;
	const size_t ss = clo_type_sizeof(scanner->sum_type);
	Launch L(cq_exec, off ? CLO_SCAN_BLELLOCH_KNAME_SEGMENTED : CLO_SCAN_BLELLOCH_KNAME_LOOKBACK);
	bool ok = false;
	switch (scanner->elem_type) {
	case CLO_CHAR: ok = launch_for_in<int8_t>(ss, a); break;
	case CLO_UCHAR: ok = launch_for_in<u8>(ss, a); break;
	case CLO_SHORT: ok = launch_for_in<int16_t>(ss, a); break;
	case CLO_USHORT: ok = launch_for_in<u16>(ss, a); break;
	case CLO_INT: ok = launch_for_in<int32_t>(ss, a); break;
	case CLO_UINT: ok = launch_for_in<u32>(ss, a); break;
	case CLO_LONG: ok = launch_for_in<int64_t>(ss, a); break;
	case CLO_ULONG: ok = launch_for_in<u64>(ss, a); break;
	default: break;
	}
	if (!ok) {
		bd->issued = a.base;
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "unsupported scan types");
		return nullptr;
	}
	if (!L.done(err)) {
		bd->dirty = true;
		return nullptr;
	}
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
	return i == 0 ? CLO_SCAN_BLELLOCH_KNAME_LOOKBACK : CLO_SCAN_BLELLOCH_KNAME_SEGMENTED;
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
	return (kScanTile + kScanTile / 32 + 1 + kScanBlock / kWave + 1) * acc + 4;
}

}  // namespace

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
	ccl_context_destroy(scan->ctx);
	delete scan;
}

CCLEvent* clo_scan_with_device_data(CloScan* scanner, CCLQueue* cq_exec, CCLQueue* cq_comm,
	CCLBuffer* data_in, CCLBuffer* data_out, size_t numel, size_t lws_max, GError** err) {
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
	if (!use_device(cq_exec->ctx, err)) return nullptr;
	return scanner->impl.scan_with_device_data(scanner, cq_exec, cq_comm, data_in, data_out,
		numel, lws_max, err);
}

CCLEvent* clo_scan_segmented_with_device_data(CloScan* scanner, CCLQueue* cq_exec,
	CCLQueue* cq_comm, CCLBuffer* data_in, CCLBuffer* data_out, CCLBuffer* offsets, size_t nseg,
	size_t numel, GError** err) {
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
	if (err && *err) return CL_FALSE;
	if (!scanner || !data_in || !data_out) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL scanner or host pointer");
		return CL_FALSE;
	}
	const size_t ib = numel * clo_type_sizeof(scanner->elem_type);
	const size_t ob = numel * clo_type_sizeof(scanner->sum_type);
	CCLQueue* own = nullptr;
	CCLBuffer *din = nullptr, *dout = nullptr;
	CCLEvent* evt = nullptr;
	CCLEventWaitList ewl = nullptr;
	cl_bool ok = CL_FALSE;
	if (!cq_exec) {
		own = ccl_queue_new(scanner->ctx, nullptr, 0, err);
		if (!own) return CL_FALSE;
		cq_exec = own;
	}
	if (!cq_comm) cq_comm = cq_exec;
	din = ccl_buffer_new(scanner->ctx, CL_MEM_READ_ONLY, ib, nullptr, err);
	dout = din ? ccl_buffer_new(scanner->ctx, CL_MEM_WRITE_ONLY, ob, nullptr, err) : nullptr;
	if (!din || !dout) goto done;
	if (!ccl_buffer_enqueue_write(din, cq_comm, CL_TRUE, 0, ib, data_in, nullptr, err)) goto done;
	evt = clo_scan_with_device_data(scanner, cq_exec, cq_comm, din, dout, numel, lws_max, err);
	if (!evt) goto done;
	ccl_event_wait_list_add(&ewl, evt, nullptr);
	if (!ccl_buffer_enqueue_read(dout, cq_comm, CL_TRUE, 0, ob, data_out, &ewl, err)) goto done;
	ok = CL_TRUE;
done:
	ccl_event_wait_list_clear(&ewl);
	if (din) ccl_buffer_destroy(din);
	if (dout) ccl_buffer_destroy(dout);
	if (own) ccl_queue_destroy(own);
	return ok;
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

}  // extern "C"

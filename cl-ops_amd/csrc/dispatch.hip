// Request dispatcher: configuration table, GPU feature kernel, exhaustive
// tuner, persisted tuner table and the learned cost model's batched choice.
// No counterpart in CL_Ops.  Every configuration computes the same result;
// correctness never depends on the choice.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "clo_internal.hpp"
#include "dev_common.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

constexpr u32 kCfgs = CLO_DISPATCH_NUM_CONFIGS;
constexpr u32 kFeat = CLO_DISPATCH_NUM_FEATURES;
constexpr u32 kSample = 4096;
constexpr size_t kClassKeys = static_cast<size_t>(1) << 21;  // work-class timing size

struct CfgInfo {
	const char* name;
	int op;
	const char* algo;     // sort configurations
	const char* options;  // sort configurations
	u32 items;            // scan configurations
	size_t max_n;         // 0 = any
	u32 segtile;          // segmented configurations: items per thread of the tile
	u32 seggrid;          // segmented configurations: resident blocks per CU
};

#define SEGCFG(t, g) \
	{"segmented tile 256x" #t ", " #g "/CU", CLO_OP_SEGSORT, "satradix", nullptr, 0, 0, t, g}
constexpr CfgInfo kCfg[kCfgs] = {
	{"satradix 256x32", CLO_OP_SORT, "satradix", "wg=256,ipt=32", 0, 0, 0, 0},
	{"satradix 512x16", CLO_OP_SORT, "satradix", "wg=512,ipt=16", 0, 0, 0, 0},
	{"satradix 256x16", CLO_OP_SORT, "satradix", "wg=256,ipt=16", 0, 0, 0, 0},
	{"abitonic", CLO_OP_SORT, "abitonic", nullptr, 0, 0, 0, 0},
	{"segment-local sort", CLO_OP_SORT, nullptr, nullptr, 0, 65536, 0, 0},
	{"gselect", CLO_OP_SORT, "gselect", nullptr, 0, 4096, 0, 0},
	{"blelloch 8 items", CLO_OP_SCAN, nullptr, nullptr, 8, 0, 0, 0},
	{"blelloch 16 items", CLO_OP_SCAN, nullptr, nullptr, 16, 0, 0, 0},
	{"blelloch 32 items", CLO_OP_SCAN, nullptr, nullptr, 32, 0, 0, 0},
	{"satradix 4-bit digits", CLO_OP_SORT, "satradix", "radix=16", 0, 0, 0, 0},
	{"satradix 6-bit digits", CLO_OP_SORT, "satradix", "radix=64", 0, 0, 0, 0},
	{"satradix 7-bit digits", CLO_OP_SORT, "satradix", "radix=128", 0, 0, 0, 0},
	SEGCFG(8, 2), SEGCFG(8, 3), SEGCFG(8, 4), SEGCFG(8, 5),
	SEGCFG(16, 2), SEGCFG(16, 3), SEGCFG(16, 4), SEGCFG(16, 5),
	SEGCFG(32, 2), SEGCFG(32, 3), SEGCFG(32, 4), SEGCFG(32, 5),
};
#undef SEGCFG
constexpr u32 kSegFirst = 12;  // first segmented configuration

constexpr u32 seg_cfg_index(u32 tile, u32 grid) {
	return kSegFirst + (tile == 8 ? 0u : tile == 16 ? 4u : 8u) + (grid - 2);
}
static_assert(kCfg[seg_cfg_index(seg_default_tile(4), kSegBlocksPerCuDefault)].segtile ==
		seg_default_tile(4) && kCfg[seg_cfg_index(seg_default_tile(8),
		kSegBlocksPerCuDefault)].seggrid == kSegBlocksPerCuDefault,
	"segmented configuration table");

// LDS of the feature statistics (one request per 256-thread block).
struct FeatSmem {
	u64 x[kSample];  // the sample (32-bit keys use the first half as u32)
	u32 h[256];
	float f[4];
	float g[3][4];
	u32 u[3][4];
};

// Sample statistics of a request (see the feature list in clo_dispatch.h),
// computed by one block of 256 threads.
template <typename K>
__device__ void features_block(const K* __restrict__ x, u64 n, u32 op,
	const u32* __restrict__ off, u64 nseg, float* __restrict__ out, FeatSmem& sm) {
	constexpr u32 kPer = kSample / 256;
	K* s_x = reinterpret_cast<K*>(sm.x);
	u32* s_h = sm.h;
	float* s_f = sm.f;
	float (*s_g)[4] = sm.g;
	u32 (*s_u)[4] = sm.u;
	const u32 tid = threadIdx.x;
	// Segment lengths (one segment [0, n) when off is NULL): mean log2(len+1)
	// and the share of keys in segments longer than 4096, over up to 4096
	// evenly spaced segments, clamped as the segmented sort clamps them.
	{
		const u64 ns = off ? nseg : 1;
		const u32 S2 = ns < kSample ? static_cast<u32>(ns) : kSample;
		float lg = 0.f, tot = 0.f, big = 0.f;
		for (u32 i = tid; i < S2; i += 256) {
			const u64 j = static_cast<u64>(i) * ns / S2;
			u64 len = n;
			if (off) {
				u64 end = j + 1 < ns ? off[j + 1] : n;
				if (end > n) end = n;
				const u64 st = off[j] < end ? off[j] : end;
				len = end - st;
			}
			lg += log2f(static_cast<float>(len) + 1.f);
			tot += static_cast<float>(len);
			big += len > 4096 ? static_cast<float>(len) : 0.f;
		}
		lg = wave_sum(lg);
		tot = wave_sum(tot);
		big = wave_sum(big);
		if (lane_id() == 0) {
			s_g[0][wave_id()] = lg;
			s_g[1][wave_id()] = tot;
			s_g[2][wave_id()] = big;
		}
		__syncthreads();
		if (tid == 0) {
			const float L = s_g[0][0] + s_g[0][1] + s_g[0][2] + s_g[0][3];
			const float T = s_g[1][0] + s_g[1][1] + s_g[1][2] + s_g[1][3];
			const float B = s_g[2][0] + s_g[2][1] + s_g[2][2] + s_g[2][3];
			out[8] = S2 ? L / static_cast<float>(S2) / 16.f : 0.f;
			out[9] = T > 0.f ? B / T : 0.f;
		}
	}
	const u32 S = n < kSample ? static_cast<u32>(n) : kSample;
	// The sample and its right neighbours, every load in flight at once.
	K a[kPer], b[kPer];
#pragma unroll
	for (u32 r = 0; r < kPer; ++r) {
		const u32 i = r * 256 + tid;
		const u64 p = static_cast<u64>(i) * n / (S ? S : 1);
		a[r] = i < S ? x[p] : K(0);
		b[r] = (i < S && p + 1 < n) ? x[p + 1] : K(0);
	}
	u32 asc = 0, dsc = 0, pairs = 0;
#pragma unroll
	for (u32 r = 0; r < kPer; ++r) {
		const u32 i = r * 256 + tid;
		if (i < S) s_x[i] = a[r];
		const u64 p = static_cast<u64>(i) * n / (S ? S : 1);
		if (i < S && p + 1 < n) {
			asc += a[r] < b[r] ? 1u : 0u;
			dsc += a[r] > b[r] ? 1u : 0u;
			++pairs;
		}
	}
	asc = wave_sum(asc);
	dsc = wave_sum(dsc);
	pairs = wave_sum(pairs);
	if (lane_id() == 0) {
		s_u[0][wave_id()] = asc;
		s_u[1][wave_id()] = dsc;
		s_u[2][wave_id()] = pairs;
	}
	float esum = 0.f, emax = 0.f;
	u32 constant = 0;
	for (u32 bt = 0; bt < sizeof(K); ++bt) {
		s_h[tid] = 0;
		__syncthreads();
		for (u32 i = tid; i < S; i += 256)
			atomicAdd(&s_h[static_cast<u32>(s_x[i] >> (8 * bt)) & 255u], 1u);
		__syncthreads();
		const u32 c = s_h[tid];
		float e = 0.f;
		if (c) {
			const float pr = static_cast<float>(c) / static_cast<float>(S);
			e = -pr * log2f(pr);
		}
		e = wave_sum(e);
		if (lane_id() == 0) s_f[wave_id()] = e;
		__syncthreads();
		if (tid == 0) {
			const float H = s_f[0] + s_f[1] + s_f[2] + s_f[3];
			esum += H;
			emax = fmaxf(emax, H);
			constant += H < 1e-6f ? 1u : 0u;
		}
		__syncthreads();
	}
	if (tid == 0) {
		const u32 A = s_u[0][0] + s_u[0][1] + s_u[0][2] + s_u[0][3];
		const u32 D = s_u[1][0] + s_u[1][1] + s_u[1][2] + s_u[1][3];
		const u32 P = s_u[2][0] + s_u[2][1] + s_u[2][2] + s_u[2][3];
		out[0] = log2f(static_cast<float>(n)) / 32.f;
		out[1] = static_cast<float>(sizeof(K)) / 8.f;
		out[2] = esum / static_cast<float>(sizeof(K)) / 8.f;
		out[3] = emax / 8.f;
		out[4] = static_cast<float>(constant) / static_cast<float>(sizeof(K));
		out[5] = P ? static_cast<float>(A) / static_cast<float>(P) : 0.f;
		out[6] = P ? static_cast<float>(D) / static_cast<float>(P) : 0.f;
		out[7] = static_cast<float>(op);
	}
}

template <typename K>
__global__ __launch_bounds__(256) void dispatch_features(const K* __restrict__ x, u64 n, u32 op,
	const u32* __restrict__ off, u64 nseg, float* __restrict__ out) {
	__shared__ FeatSmem sm;
	features_block<K>(x, n, op, off, nseg, out, sm);
}

// Batched features: block b computes request b's features.
struct FeatDesc {
	const void* ptr;
	u64 n;
	u32 width;
	u32 op;
};

__global__ __launch_bounds__(256) void dispatch_features_batch(const FeatDesc* __restrict__ desc,
	float* __restrict__ out) {
	__shared__ FeatSmem sm;
	const FeatDesc dd = desc[blockIdx.x];
	float* o = out + static_cast<size_t>(blockIdx.x) * kFeat;
	if (dd.width == 8)
		features_block<u64>(static_cast<const u64*>(dd.ptr), dd.n, dd.op, nullptr, 0, o, sm);
	else
		features_block<u32>(static_cast<const u32*>(dd.ptr), dd.n, dd.op, nullptr, 0, o, sm);
}

struct Bucket {
	double sum[kCfgs] = {};
	u32 cnt[kCfgs] = {};
};

using BucketKey = std::tuple<int, int, int, int>;

int order_class(const float* f) {
	if (f[5] > 0.9f) return 1;   // ascending
	if (f[6] > 0.9f) return 2;   // descending
	if (f[3] < 0.55f) return 3;  // few distinct values (low byte entropy)
	return 0;
}

BucketKey bucket_key(CloOp op, size_t n, size_t width, const float* f) {
	int lg = 0;
	while ((static_cast<size_t>(2) << lg) <= n) ++lg;
	return BucketKey(static_cast<int>(op), static_cast<int>(width), lg, order_class(f));
}

bool valid_width(size_t w) { return w == 4 || w == 8; }

}  // namespace

struct clo_dispatch {
	CCLContext* ctx;
	u32 cus = 256;  // compute units of the device
	CloSort* sort[2][kCfgs] = {};
	CloScan* scan[2] = {};
	CCLBuffer* off1 = nullptr;  // one segment starting at 0
	Scratch feat;
	CloMlp* mlp = nullptr;
	float margin = 0.f;  // model policy: least predicted gain (log2) to leave the fixed config
	CloMlp* seg_mlp = nullptr;  // per-segment work-class cost model
	// Work-class timing data: keys in/out and offsets of equal segments.
	CCLBuffer* cls_in = nullptr;
	CCLBuffer* cls_out = nullptr;
	CCLBuffer* cls_off = nullptr;
	// Batched features: request descriptors, staged in pinned host memory.
	FeatDesc* h_desc = nullptr;
	size_t h_desc_cap = 0;
	Scratch desc;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	std::map<BucketKey, Bucket> table;
};

namespace {

CloSort* sorter_for(CloDispatch* d, u32 cfg, size_t width, GError** err) {
	const int w = width == 8 ? 1 : 0;
	const u32 slot = cfg == 4 ? 0 : cfg;  // the segment-local sort uses a satradix sorter
	if (!d->sort[w][slot]) {
		CloType t = width == 8 ? CLO_ULONG : CLO_UINT;
		char opts[64];
		const char* o = kCfg[slot].options;
		if (kCfg[slot].segtile) {
			snprintf(opts, sizeof opts, "segtile=%u,seggrid=%u", kCfg[slot].segtile,
				kCfg[slot].seggrid * d->cus);
			o = opts;
		}
		d->sort[w][slot] = clo_sort_new(kCfg[slot].algo, o, d->ctx, &t, nullptr, nullptr, nullptr,
			nullptr, err);
	}
	return d->sort[w][slot];
}

// Least length of length bin b (the segmented sort's seg_bin), 0 for the
// bins no length falls in (1, 2, 3, 5, 7).
size_t bin_length(u32 b) {
	if (b == 0) return 1;
	const u32 lg = b / 4, sub = b % 4;
	if (lg >= 2) return static_cast<size_t>(4 + sub) << (lg - 2);
	if (b == 4) return 2;
	if (b == 6) return 3;
	return 0;
}

// Packs one class digit per bin into the sorter's two words.
void pack_classes(const unsigned char* cls, u64 out[2]) {
	out[0] = out[1] = 0;
	if (!cls) return;
	for (u32 b = 0; b < CLO_DISPATCH_SEG_BINS; ++b)
		out[b / 32] |= static_cast<u64>(cls[b] > 2 ? 2 : cls[b]) << (2 * (b % 32));
}

__global__ void equal_offsets(u32* __restrict__ off, u32 nseg, u32 len) {
	const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j < nseg) off[j] = j * len;
}

// Times `run` (which enqueues one call on cq): once untimed, then reps timed
// with events; the least time in ms.
template <typename F>
bool time_best(CloDispatch* d, CCLQueue* cq, u32 reps, F run, float* best, GError** err) {
	if (!run()) return false;
	for (u32 r = 0; r < reps; ++r) {
		float ms = 0.f;
		if (hip_failed(hipEventRecord(d->e0, cq->stream), err, "hipEventRecord") || !run() ||
			hip_failed(hipEventRecord(d->e1, cq->stream), err, "hipEventRecord") ||
			hip_failed(hipEventSynchronize(d->e1), err, "hipEventSynchronize") ||
			hip_failed(hipEventElapsedTime(&ms, d->e0, d->e1), err, "hipEventElapsedTime"))
			return false;
		if (r == 0 || ms < *best) *best = ms;
	}
	return true;
}

CloScan* scanner_for(CloDispatch* d, size_t width, GError** err) {
	const int w = width == 8 ? 1 : 0;
	if (!d->scan[w]) {
		const CloType t = width == 8 ? CLO_ULONG : CLO_UINT;
		d->scan[w] = clo_scan_new("blelloch", nullptr, d->ctx, t, t, nullptr, err);
	}
	return d->scan[w];
}

}  // namespace

extern "C" {

CloDispatch* clo_dispatch_new(CCLContext* ctx, GError** err) {
	if (err && *err) return nullptr;
	if (!ctx) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL context");
		return nullptr;
	}
	if (!use_device(ctx, err)) return nullptr;
	CloDispatch* d = new clo_dispatch();
	d->ctx = ctx;
	ccl_context_ref(ctx);
	int cus = 0;
	if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->dev) ==
			hipSuccess && cus > 0)
		d->cus = static_cast<u32>(cus);
	u32 zero = 0;
	d->off1 = ccl_buffer_new(ctx, CL_MEM_READ_ONLY, 4, &zero, err);
	if (!d->off1 || hip_failed(hipEventCreate(&d->e0), err, "hipEventCreate") ||
		hip_failed(hipEventCreate(&d->e1), err, "hipEventCreate")) {
		clo_dispatch_destroy(d);
		return nullptr;
	}
	return d;
}

void clo_dispatch_destroy(CloDispatch* d) {
	if (!d) return;
	(void)hipSetDevice(d->ctx->dev);
	for (auto& row : d->sort)
		for (CloSort* s : row) clo_sort_destroy(s);
	for (CloScan* s : d->scan) clo_scan_destroy(s);
	if (d->off1) ccl_buffer_destroy(d->off1);
	for (CCLBuffer* b : {d->cls_in, d->cls_out, d->cls_off})
		if (b) ccl_buffer_destroy(b);
	if (d->e0) (void)hipEventDestroy(d->e0);
	if (d->e1) (void)hipEventDestroy(d->e1);
	d->feat.release();
	d->desc.release();
	if (d->h_desc) (void)hipHostFree(d->h_desc);
	ccl_context_destroy(d->ctx);
	delete d;
}

const char* clo_dispatch_config_name(cl_uint cfg) {
	return cfg < kCfgs ? kCfg[cfg].name : nullptr;
}

cl_bool clo_dispatch_config_feasible(cl_uint cfg, CloOp op, size_t n, size_t width) {
	if (cfg >= kCfgs || kCfg[cfg].op != static_cast<int>(op) || !valid_width(width)) return CL_FALSE;
	if (kCfg[cfg].segtile == 32 && width != 4) return CL_FALSE;  // LDS holds 16 x 8-byte keys
	return (kCfg[cfg].max_n == 0 || n <= kCfg[cfg].max_n) ? CL_TRUE : CL_FALSE;
}

cl_uint clo_dispatch_fixed_config(CloOp op, size_t n, size_t width) {
	if (op == CLO_OP_SCAN) return 7;
	if (op == CLO_OP_SEGSORT) return seg_cfg_index(seg_default_tile(width), kSegBlocksPerCuDefault);
	// satradix's default tile shape for keys only
	if (n <= (static_cast<size_t>(1) << 20)) return 2;
	return width == 8 ? 1 : 0;
}

void clo_dispatch_model_row(const float* features, cl_uint cfg, float* row) {
	for (u32 i = 0; i < kFeat; ++i) row[i] = features[i];
	for (u32 c = 0; c < kCfgs; ++c) row[kFeat + c] = c == cfg ? 1.f : 0.f;
}

namespace {

cl_bool features_impl(CloDispatch* d, CCLQueue* cq, CloOp op, CCLBuffer* data,
	CCLBuffer* offsets, size_t nseg, size_t n, size_t width, float* features, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!d || !cq || !data || !features || !valid_width(width) || n == 0 ||
		data->size < n * width || (offsets && (nseg == 0 || offsets->size < nseg * 4))) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad feature request");
		return CL_FALSE;
	}
	const u32* off = offsets ? static_cast<const u32*>(offsets->ptr) : nullptr;
	if (!use_device(d->ctx, err) || !d->feat.ensure(kFeat * 4, err)) return CL_FALSE;
	{
		Launch L(cq, "dispatch_features");
		float* out = static_cast<float*>(d->feat.ptr);
		if (width == 8)
			dispatch_features<u64><<<1, 256, 0, cq->stream>>>(static_cast<const u64*>(data->ptr),
				n, static_cast<u32>(op), off, nseg, out);
		else
			dispatch_features<u32><<<1, 256, 0, cq->stream>>>(static_cast<const u32*>(data->ptr),
				n, static_cast<u32>(op), off, nseg, out);
		if (!L.done(err)) return CL_FALSE;
	}
	return (hip_failed(hipMemcpyAsync(features, d->feat.ptr, kFeat * 4, hipMemcpyDeviceToHost,
				cq->stream), err, "hipMemcpyAsync D2H") ||
			   hip_failed(hipStreamSynchronize(cq->stream), err, "hipStreamSynchronize"))
		? CL_FALSE
		: CL_TRUE;
}

}  // namespace

cl_bool clo_dispatch_features(CloDispatch* d, CCLQueue* cq, CloOp op, CCLBuffer* data, size_t n,
	size_t width, float* features, GError** err) {
	if (op == CLO_OP_SEGSORT) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"segmented requests take clo_dispatch_features_segmented");
		return CL_FALSE;
	}
	return features_impl(d, cq, op, data, nullptr, 0, n, width, features, err);
}

cl_bool clo_dispatch_features_batch(CloDispatch* d, CCLQueue* cq, size_t count, const int* ops,
	CCLBuffer* const* data, const size_t* ns, const size_t* widths, float* features,
	GError** err) {
	ApiRange api_range("clo_dispatch_features_batch");
	if (err && *err) return CL_FALSE;
	if (!d || !cq || (count && (!ops || !data || !ns || !widths || !features)) ||
		count > 0x7FFFFFFFu) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad batched feature request");
		return CL_FALSE;
	}
	if (count == 0) return CL_TRUE;
	for (size_t i = 0; i < count; ++i)
		if (!data[i] || !valid_width(widths[i]) || ns[i] == 0 ||
			data[i]->size < ns[i] * widths[i] || (ops[i] != CLO_OP_SORT && ops[i] != CLO_OP_SCAN)) {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad feature request %zu of the batch", i);
			return CL_FALSE;
		}
	if (!use_device(d->ctx, err)) return CL_FALSE;
	if (d->h_desc_cap < count) {
		if (d->h_desc) (void)hipHostFree(d->h_desc);
		d->h_desc = nullptr;
		d->h_desc_cap = 0;
		if (hip_failed(hipHostMalloc(reinterpret_cast<void**>(&d->h_desc),
				count * sizeof(FeatDesc), hipHostMallocDefault), err, "hipHostMalloc"))
			return CL_FALSE;
		d->h_desc_cap = count;
	}
	for (size_t i = 0; i < count; ++i)
		d->h_desc[i] = {data[i]->ptr, static_cast<u64>(ns[i]), static_cast<u32>(widths[i]),
			static_cast<u32>(ops[i])};
	if (!d->desc.ensure(count * sizeof(FeatDesc), err) ||
		!d->feat.ensure(count * kFeat * sizeof(float), err))
		return CL_FALSE;
	if (hip_failed(hipMemcpyAsync(d->desc.ptr, d->h_desc, count * sizeof(FeatDesc),
			hipMemcpyHostToDevice, cq->stream), err, "hipMemcpyAsync H2D"))
		return CL_FALSE;
	{
		Launch L(cq, "dispatch_features_batch");
		dispatch_features_batch<<<static_cast<u32>(count), 256, 0, cq->stream>>>(
			static_cast<const FeatDesc*>(d->desc.ptr), static_cast<float*>(d->feat.ptr));
		if (!L.done(err)) return CL_FALSE;
	}
	return (hip_failed(hipMemcpyAsync(features, d->feat.ptr, count * kFeat * sizeof(float),
				hipMemcpyDeviceToHost, cq->stream), err, "hipMemcpyAsync D2H") ||
			   hip_failed(hipStreamSynchronize(cq->stream), err, "hipStreamSynchronize"))
		? CL_FALSE
		: CL_TRUE;
}

cl_bool clo_dispatch_features_segmented(CloDispatch* d, CCLQueue* cq, CCLBuffer* data,
	CCLBuffer* offsets, size_t nseg, size_t n, size_t width, float* features, GError** err) {
	if (!offsets) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "segmented features need offsets");
		return CL_FALSE;
	}
	return features_impl(d, cq, CLO_OP_SEGSORT, data, offsets, nseg, n, width, features, err);
}

CCLEvent* clo_dispatch_run(CloDispatch* d, CCLQueue* cq, cl_uint cfg, CloOp op, CCLBuffer* in,
	CCLBuffer* out, size_t n, size_t width, GError** err) {
	ApiRange api_range("clo_dispatch_run");
	if (err && *err) return nullptr;
	if (!d || !cq || !in || !out || !clo_dispatch_config_feasible(cfg, op, n, width)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "configuration %u not feasible for this request",
			cfg);
		return nullptr;
	}
	if (in->size < n * width || out->size < n * width) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "buffers smaller than the request");
		return nullptr;
	}
	if (op == CLO_OP_SEGSORT) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "segmented requests take clo_dispatch_run_segmented");
		return nullptr;
	}
	if (op == CLO_OP_SCAN) {
		CloScan* s = scanner_for(d, width, err);
		if (!s) return nullptr;
		return clo_blelloch_scan_items(s, cq, in, out, n, kCfg[cfg].items, err);
	}
	CloSort* s = sorter_for(d, cfg, width, err);
	if (!s) return nullptr;
	if (cfg == 4)
		return clo_sort_segmented_with_device_data(s, cq, nullptr, in, nullptr, out, nullptr,
			d->off1, 1, n, err);
	return clo_sort_with_device_data(s, cq, nullptr, in, out, n, 0, err);
}

cl_bool clo_dispatch_tune(CloDispatch* d, CCLQueue* cq, CloOp op, CCLBuffer* in, CCLBuffer* out,
	size_t n, size_t width, cl_uint reps, float* times_ms, GError** err) {
	ApiRange api_range("clo_dispatch_tune");
	if (err && *err) return CL_FALSE;
	if (!d || !cq || !times_ms || reps == 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad tuner arguments");
		return CL_FALSE;
	}
	for (u32 c = 0; c < kCfgs; ++c) {
		times_ms[c] = -1.f;
		if (!clo_dispatch_config_feasible(c, op, n, width)) continue;
		float best = 0.f;
		if (!time_best(d, cq, reps,
				[&] { return clo_dispatch_run(d, cq, c, op, in, out, n, width, err) != nullptr; },
				&best, err))
			return CL_FALSE;
		times_ms[c] = best;
	}
	return CL_TRUE;
}

CCLEvent* clo_dispatch_run_segmented(CloDispatch* d, CCLQueue* cq, cl_uint cfg, CCLBuffer* in,
	CCLBuffer* out, CCLBuffer* offsets, size_t nseg, size_t n, size_t width,
	const unsigned char* classes, GError** err) {
	ApiRange api_range("clo_dispatch_run_segmented");
	if (err && *err) return nullptr;
	if (!d || !cq || !in || !out || !offsets ||
		!clo_dispatch_config_feasible(cfg, CLO_OP_SEGSORT, n, width)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"configuration %u not feasible for this segmented request", cfg);
		return nullptr;
	}
	if (in->size < n * width || out->size < n * width) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "buffers smaller than the request");
		return nullptr;
	}
	CloSort* s = sorter_for(d, cfg, width, err);
	if (!s) return nullptr;
	pack_classes(classes, s->seg_cls);
	return clo_sort_segmented_with_device_data(s, cq, nullptr, in, nullptr, out, nullptr,
		offsets, nseg, n, err);
}

cl_bool clo_dispatch_tune_segmented(CloDispatch* d, CCLQueue* cq, CCLBuffer* in, CCLBuffer* out,
	CCLBuffer* offsets, size_t nseg, size_t n, size_t width, cl_uint reps, float* times_ms,
	GError** err) {
	ApiRange api_range("clo_dispatch_tune_segmented");
	if (err && *err) return CL_FALSE;
	if (!d || !cq || !times_ms || reps == 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad tuner arguments");
		return CL_FALSE;
	}
	for (u32 c = 0; c < kCfgs; ++c) {
		times_ms[c] = -1.f;
		if (!clo_dispatch_config_feasible(c, CLO_OP_SEGSORT, n, width)) continue;
		float best = 0.f;
		if (!time_best(d, cq, reps, [&] {
				return clo_dispatch_run_segmented(d, cq, c, in, out, offsets, nseg, n, width,
					nullptr, err) != nullptr;
			}, &best, err))
			return CL_FALSE;
		times_ms[c] = best;
	}
	return CL_TRUE;
}

cl_uint clo_dispatch_seg_bin(size_t len) {
	return seg_len_bin(len > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<u32>(len));
}

size_t clo_dispatch_seg_bin_length(cl_uint bin) {
	return bin < CLO_DISPATCH_SEG_BINS ? bin_length(bin) : 0;
}

cl_bool clo_dispatch_tune_segment_class(CloDispatch* d, CCLQueue* cq, cl_uint cfg, size_t width,
	size_t len, cl_uint reps, float* ns_per_key, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!d || !cq || !ns_per_key || reps == 0 || len == 0 || len > (1u << 16) ||
		!clo_dispatch_config_feasible(cfg, CLO_OP_SEGSORT, len, width)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad work-class timing request");
		return CL_FALSE;
	}
	// About 2^21 keys in equal segments of len keys.
	const size_t nseg = len >= kClassKeys ? 1 : kClassKeys / len;
	const size_t n = nseg * len;
	if (!use_device(d->ctx, err)) return CL_FALSE;
	if (!d->cls_in) {
		const size_t bytes = (kClassKeys + (1u << 16)) * 8;
		d->cls_in = ccl_buffer_new(d->ctx, CL_MEM_READ_WRITE, bytes, nullptr, err);
		d->cls_out = d->cls_in ? ccl_buffer_new(d->ctx, CL_MEM_READ_WRITE, bytes, nullptr, err)
			: nullptr;
		d->cls_off = d->cls_out ? ccl_buffer_new(d->ctx, CL_MEM_READ_WRITE, kClassKeys * 4,
			nullptr, err) : nullptr;
		if (!d->cls_off) return CL_FALSE;
	}
	if (!clo_util_generate(cq, d->cls_in, n, width == 8 ? CLO_GEN_U64 : CLO_GEN_U32_HI,
			0x5E6C1A55u + len, 0, 0, err))
		return CL_FALSE;
	{
		Launch L(cq, "dispatch_equal_offsets");
		equal_offsets<<<static_cast<u32>(div_ceil(nseg, 256)), 256, 0, cq->stream>>>(
			static_cast<u32*>(d->cls_off->ptr), static_cast<u32>(nseg), static_cast<u32>(len));
		if (!L.done(err)) return CL_FALSE;
	}
	const u32 CAP = 256 * kCfg[cfg].segtile, W = CAP / 2;
	for (u32 c = 0; c < 3; ++c) {
		ns_per_key[c] = -1.f;
		if ((c == 0 && len > W) || (c == 1 && len > CAP)) continue;
		unsigned char cls[CLO_DISPATCH_SEG_BINS];
		memset(cls, static_cast<int>(c), sizeof cls);
		float best = 0.f;
		if (!time_best(d, cq, reps, [&] {
				return clo_dispatch_run_segmented(d, cq, cfg, d->cls_in, d->cls_out, d->cls_off,
					nseg, n, width, cls, err) != nullptr;
			}, &best, err))
			return CL_FALSE;
		ns_per_key[c] = best * 1e6f / static_cast<float>(n);
	}
	return CL_TRUE;
}

void clo_dispatch_segment_row(size_t len, size_t width, cl_uint cfg, cl_uint cls, float* row) {
	row[0] = std::log2(static_cast<float>(len) + 1.f) / 16.f;
	row[1] = static_cast<float>(width) / 8.f;
	row[2] = cfg < kCfgs ? static_cast<float>(kCfg[cfg].segtile) / 32.f : 0.f;
	row[3] = cfg < kCfgs ? static_cast<float>(kCfg[cfg].seggrid) / 4.f : 0.f;
	for (u32 c = 0; c < 3; ++c) row[4 + c] = c == cls ? 1.f : 0.f;
}

cl_bool clo_dispatch_set_segment_model(CloDispatch* d, CloMlp* mlp, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!d || (mlp && clo_mlp_num_inputs(mlp) != CLO_DISPATCH_SEG_INPUTS)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "segment model must take %d inputs",
			CLO_DISPATCH_SEG_INPUTS);
		return CL_FALSE;
	}
	d->seg_mlp = mlp;
	return CL_TRUE;
}

cl_bool clo_dispatch_segment_classes(CloDispatch* d, CCLQueue* cq, cl_uint cfg, size_t width,
	float log2_margin, unsigned char* classes, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!d || !classes || !clo_dispatch_config_feasible(cfg, CLO_OP_SEGSORT, 1, width)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad segment class request");
		return CL_FALSE;
	}
	memset(classes, 0, CLO_DISPATCH_SEG_BINS);
	if (!d->seg_mlp) return CL_TRUE;
	if (!cq) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "the segment model needs a queue");
		return CL_FALSE;
	}
	const u32 CAP = 256 * kCfg[cfg].segtile, W = CAP / 2;
	std::vector<float> rows;
	std::vector<std::pair<u32, u32>> who;  // (bin, class)
	for (u32 b = 0; b < CLO_DISPATCH_SEG_BINS; ++b) {
		const size_t L = bin_length(b);
		if (L == 0 || L > CAP) continue;  // empty bin, or large by length alone
		for (u32 c = (L > W ? 1u : 0u); c < 3; ++c) {
			rows.resize(rows.size() + CLO_DISPATCH_SEG_INPUTS);
			clo_dispatch_segment_row(L, width, cfg, c, rows.data() + rows.size() -
				CLO_DISPATCH_SEG_INPUTS);
			who.emplace_back(b, c);
		}
	}
	std::vector<float> pred(who.size());
	if (!who.empty() && !clo_mlp_predict(d->seg_mlp, cq, rows.data(), who.size(), pred.data(), err))
		return CL_FALSE;
	// Per bin: the least feasible class unless another is predicted cheaper
	// by more than the margin.
	for (size_t i = 0; i < who.size();) {
		const u32 b = who[i].first;
		const float base = pred[i];
		u32 best = who[i].second;
		float bp = base;
		size_t k = i;
		for (; k < who.size() && who[k].first == b; ++k)
			if (pred[k] < bp) {
				bp = pred[k];
				best = who[k].second;
			}
		if (bp < base - log2_margin) classes[b] = static_cast<unsigned char>(best);
		i = k;
	}
	return CL_TRUE;
}

void clo_dispatch_table_add(CloDispatch* d, CloOp op, size_t n, size_t width,
	const float* features, const float* times_ms) {
	if (!d || !features || !times_ms) return;
	Bucket& b = d->table[bucket_key(op, n, width, features)];
	for (u32 c = 0; c < kCfgs; ++c) {
		if (times_ms[c] <= 0.f) continue;
		b.sum[c] += std::log2(static_cast<double>(times_ms[c]) * 1000.0);
		b.cnt[c] += 1;
	}
}

int clo_dispatch_table_choose(CloDispatch* d, CloOp op, size_t n, size_t width,
	const float* features) {
	if (!d || !features) return -1;
	auto it = d->table.find(bucket_key(op, n, width, features));
	if (it == d->table.end()) return -1;
	int best = -1;
	double bt = 0;
	for (u32 c = 0; c < kCfgs; ++c) {
		if (!it->second.cnt[c] || !clo_dispatch_config_feasible(c, op, n, width)) continue;
		const double m = it->second.sum[c] / it->second.cnt[c];
		if (best < 0 || m < bt) {
			best = static_cast<int>(c);
			bt = m;
		}
	}
	return best;
}

cl_bool clo_dispatch_table_save(CloDispatch* d, const char* path, GError** err) {
	if (err && *err) return CL_FALSE;
	FILE* f = (d && path) ? fopen(path, "w") : nullptr;
	if (!f) {
		set_error(err, kQuarkClo, CLO_ERROR_STREAM_WRITE, "cannot open %s", path ? path : "(null)");
		return CL_FALSE;
	}
	fprintf(f, "CLOTUNE01\n");
	for (const auto& kv : d->table)
		for (u32 c = 0; c < kCfgs; ++c)
			if (kv.second.cnt[c])
				fprintf(f, "%d %d %d %d %u %u %.17g\n", std::get<0>(kv.first), std::get<1>(kv.first),
					std::get<2>(kv.first), std::get<3>(kv.first), c, kv.second.cnt[c],
					kv.second.sum[c]);
	const bool ok = fclose(f) == 0;
	if (!ok) set_error(err, kQuarkClo, CLO_ERROR_STREAM_WRITE, "cannot write %s", path);
	return ok ? CL_TRUE : CL_FALSE;
}

cl_bool clo_dispatch_table_load(CloDispatch* d, const char* path, GError** err) {
	if (err && *err) return CL_FALSE;
	FILE* f = (d && path) ? fopen(path, "r") : nullptr;
	if (!f) {
		set_error(err, kQuarkClo, CLO_ERROR_OPENFILE, "cannot open %s", path ? path : "(null)");
		return CL_FALSE;
	}
	char magic[16] = {};
	bool ok = fscanf(f, "%15s", magic) == 1 && strcmp(magic, "CLOTUNE01") == 0;
	std::map<BucketKey, Bucket> t;
	int op, w, lg, cls;
	unsigned cfg, cnt;
	double sum;
	while (ok) {
		const int r = fscanf(f, "%d %d %d %d %u %u %lg", &op, &w, &lg, &cls, &cfg, &cnt, &sum);
		if (r == EOF) break;
		if (r != 7 || cfg >= kCfgs) {
			ok = false;
			break;
		}
		Bucket& b = t[BucketKey(op, w, lg, cls)];
		b.cnt[cfg] = cnt;
		b.sum[cfg] = sum;
	}
	fclose(f);
	if (!ok) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "%s is not a CLOTUNE01 table", path);
		return CL_FALSE;
	}
	d->table.swap(t);
	return CL_TRUE;
}

cl_bool clo_dispatch_set_model(CloDispatch* d, CloMlp* mlp, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!d || (mlp && clo_mlp_num_inputs(mlp) != CLO_DISPATCH_MODEL_INPUTS)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "model must take %d inputs",
			CLO_DISPATCH_MODEL_INPUTS);
		return CL_FALSE;
	}
	d->mlp = mlp;
	return CL_TRUE;
}

void clo_dispatch_set_margin(CloDispatch* d, float log2_margin) {
	if (d) d->margin = log2_margin > 0.f ? log2_margin : 0.f;
}

cl_bool clo_dispatch_choose(CloDispatch* d, CCLQueue* cq, CloPolicy policy, size_t count,
	const int* ops, const size_t* ns, const size_t* widths, const float* features, int* cfg_out,
	GError** err) {
	ApiRange api_range("clo_dispatch_choose");
	if (err && *err) return CL_FALSE;
	if (!d || !ops || !ns || !widths || !features || !cfg_out) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad choose arguments");
		return CL_FALSE;
	}
	for (size_t r = 0; r < count; ++r) {
		const CloOp op = static_cast<CloOp>(ops[r]);
		int c = -1;
		if (policy != CLO_POLICY_FIXED)
			c = clo_dispatch_table_choose(d, op, ns[r], widths[r], features + r * kFeat);
		cfg_out[r] = c >= 0 ? c : static_cast<int>(clo_dispatch_fixed_config(op, ns[r], widths[r]));
	}
	if (policy != CLO_POLICY_MODEL || !d->mlp || count == 0) return CL_TRUE;
	if (!cq) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "the model policy needs a queue");
		return CL_FALSE;
	}
	// One row per (request, feasible configuration); one inference launch.
	std::vector<float> rows;
	std::vector<std::pair<size_t, u32>> who;
	rows.reserve(count * 6 * CLO_DISPATCH_MODEL_INPUTS);
	for (size_t r = 0; r < count; ++r)
		for (u32 c = 0; c < kCfgs; ++c)
			if (clo_dispatch_config_feasible(c, static_cast<CloOp>(ops[r]), ns[r], widths[r])) {
				rows.resize(rows.size() + CLO_DISPATCH_MODEL_INPUTS);
				clo_dispatch_model_row(features + r * kFeat, c,
					rows.data() + rows.size() - CLO_DISPATCH_MODEL_INPUTS);
				who.emplace_back(r, c);
			}
	if (who.empty()) return CL_TRUE;
	std::vector<float> pred(who.size());
	if (!clo_mlp_predict(d->mlp, cq, rows.data(), who.size(), pred.data(), err)) return CL_FALSE;
	std::vector<float> best(count, INFINITY), fixed_pred(count, INFINITY);
	std::vector<int> arg(count, -1);
	for (size_t i = 0; i < who.size(); ++i) {
		const size_t r = who[i].first;
		if (who[i].second == clo_dispatch_fixed_config(static_cast<CloOp>(ops[r]), ns[r], widths[r]))
			fixed_pred[r] = pred[i];
		if (pred[i] < best[r]) {
			best[r] = pred[i];
			arg[r] = static_cast<int>(who[i].second);
		}
	}
	size_t moved = 0;
	for (size_t r = 0; r < count; ++r) {
		if (arg[r] < 0) continue;
		// Leave the fixed configuration only for a predicted gain above the margin.
		const int fx = static_cast<int>(clo_dispatch_fixed_config(static_cast<CloOp>(ops[r]),
			ns[r], widths[r]));
		cfg_out[r] = best[r] < fixed_pred[r] - d->margin ? arg[r] : fx;
		moved += cfg_out[r] != fx ? 1 : 0;
	}
	CLO_LOG(1, "dispatch: model policy over %zu requests (%zu rows, one inference launch): "
		"%zu leave the fixed configuration", count, who.size(), moved);
	return CL_TRUE;
}

}  // extern "C"

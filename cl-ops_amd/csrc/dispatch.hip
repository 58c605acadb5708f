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

struct CfgInfo {
	const char* name;
	int op;
	const char* algo;     // sort configurations
	const char* options;  // sort configurations
	u32 items;            // scan configurations
	size_t max_n;         // 0 = any
};

const CfgInfo kCfg[kCfgs] = {
	{"satradix 256x32", CLO_OP_SORT, "satradix", "wg=256,ipt=32", 0, 0},
	{"satradix 512x16", CLO_OP_SORT, "satradix", "wg=512,ipt=16", 0, 0},
	{"satradix 256x16", CLO_OP_SORT, "satradix", "wg=256,ipt=16", 0, 0},
	{"abitonic", CLO_OP_SORT, "abitonic", nullptr, 0, 0},
	{"segment-local sort", CLO_OP_SORT, nullptr, nullptr, 0, 65536},
	{"gselect", CLO_OP_SORT, "gselect", nullptr, 0, 4096},
	{"blelloch 8 items", CLO_OP_SCAN, nullptr, nullptr, 8, 0},
	{"blelloch 16 items", CLO_OP_SCAN, nullptr, nullptr, 16, 0},
	{"blelloch 32 items", CLO_OP_SCAN, nullptr, nullptr, 32, 0},
};

// Sample statistics of a request (see the feature list in clo_dispatch.h).
template <typename K>
__global__ __launch_bounds__(256) void dispatch_features(const K* __restrict__ x, u64 n, u32 op,
	float* __restrict__ out) {
	constexpr u32 kPer = kSample / 256;
	__shared__ K s_x[kSample];
	__shared__ u32 s_h[256];
	__shared__ float s_f[4];
	__shared__ u32 s_u[3][4];
	const u32 tid = threadIdx.x;
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
	CloSort* sort[2][kCfgs] = {};
	CloScan* scan[2] = {};
	CCLBuffer* off1 = nullptr;  // one segment starting at 0
	Scratch feat;
	CloMlp* mlp = nullptr;
	float margin = 0.f;  // model policy: least predicted gain (log2) to leave the fixed config
	hipEvent_t e0 = nullptr, e1 = nullptr;
	std::map<BucketKey, Bucket> table;
};

namespace {

CloSort* sorter_for(CloDispatch* d, u32 cfg, size_t width, GError** err) {
	const int w = width == 8 ? 1 : 0;
	const u32 slot = cfg == 4 ? 0 : cfg;  // the segment-local sort uses a satradix sorter
	if (!d->sort[w][slot]) {
		CloType t = width == 8 ? CLO_ULONG : CLO_UINT;
		d->sort[w][slot] = clo_sort_new(kCfg[slot].algo, kCfg[slot].options, d->ctx, &t, nullptr,
			nullptr, nullptr, nullptr, err);
	}
	return d->sort[w][slot];
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
	if (d->e0) (void)hipEventDestroy(d->e0);
	if (d->e1) (void)hipEventDestroy(d->e1);
	d->feat.release();
	ccl_context_destroy(d->ctx);
	delete d;
}

const char* clo_dispatch_config_name(cl_uint cfg) {
	return cfg < kCfgs ? kCfg[cfg].name : nullptr;
}

cl_bool clo_dispatch_config_feasible(cl_uint cfg, CloOp op, size_t n, size_t width) {
	if (cfg >= kCfgs || kCfg[cfg].op != static_cast<int>(op) || !valid_width(width)) return CL_FALSE;
	return (kCfg[cfg].max_n == 0 || n <= kCfg[cfg].max_n) ? CL_TRUE : CL_FALSE;
}

cl_uint clo_dispatch_fixed_config(CloOp op, size_t n, size_t width) {
	if (op == CLO_OP_SCAN) return 7;
	// satradix's default tile shape for keys only
	if (n <= (static_cast<size_t>(1) << 20)) return 2;
	return width == 8 ? 1 : 0;
}

void clo_dispatch_model_row(const float* features, cl_uint cfg, float* row) {
	for (u32 i = 0; i < kFeat; ++i) row[i] = features[i];
	for (u32 c = 0; c < kCfgs; ++c) row[kFeat + c] = c == cfg ? 1.f : 0.f;
}

cl_bool clo_dispatch_features(CloDispatch* d, CCLQueue* cq, CloOp op, CCLBuffer* data, size_t n,
	size_t width, float* features, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!d || !cq || !data || !features || !valid_width(width) || n == 0 ||
		data->size < n * width) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad feature request");
		return CL_FALSE;
	}
	if (!use_device(d->ctx, err) || !d->feat.ensure(kFeat * 4, err)) return CL_FALSE;
	{
		Launch L(cq, "dispatch_features");
		float* out = static_cast<float*>(d->feat.ptr);
		if (width == 8)
			dispatch_features<u64><<<1, 256, 0, cq->stream>>>(static_cast<const u64*>(data->ptr),
				n, static_cast<u32>(op), out);
		else
			dispatch_features<u32><<<1, 256, 0, cq->stream>>>(static_cast<const u32*>(data->ptr),
				n, static_cast<u32>(op), out);
		if (!L.done(err)) return CL_FALSE;
	}
	return (hip_failed(hipMemcpyAsync(features, d->feat.ptr, kFeat * 4, hipMemcpyDeviceToHost,
				cq->stream), err, "hipMemcpyAsync D2H") ||
			   hip_failed(hipStreamSynchronize(cq->stream), err, "hipStreamSynchronize"))
		? CL_FALSE
		: CL_TRUE;
}

CCLEvent* clo_dispatch_run(CloDispatch* d, CCLQueue* cq, cl_uint cfg, CloOp op, CCLBuffer* in,
	CCLBuffer* out, size_t n, size_t width, GError** err) {
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
	if (err && *err) return CL_FALSE;
	if (!d || !cq || !times_ms || reps == 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bad tuner arguments");
		return CL_FALSE;
	}
	for (u32 c = 0; c < kCfgs; ++c) {
		times_ms[c] = -1.f;
		if (!clo_dispatch_config_feasible(c, op, n, width)) continue;
		if (!clo_dispatch_run(d, cq, c, op, in, out, n, width, err)) return CL_FALSE;
		float best = 0.f;
		for (u32 r = 0; r < reps; ++r) {
			float ms = 0.f;
			if (hip_failed(hipEventRecord(d->e0, cq->stream), err, "hipEventRecord") ||
				!clo_dispatch_run(d, cq, c, op, in, out, n, width, err) ||
				hip_failed(hipEventRecord(d->e1, cq->stream), err, "hipEventRecord") ||
				hip_failed(hipEventSynchronize(d->e1), err, "hipEventSynchronize") ||
				hip_failed(hipEventElapsedTime(&ms, d->e0, d->e1), err, "hipEventElapsedTime"))
				return CL_FALSE;
			if (r == 0 || ms < best) best = ms;
		}
		times_ms[c] = best;
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
	for (size_t r = 0; r < count; ++r) {
		if (arg[r] < 0) continue;
		// Leave the fixed configuration only for a predicted gain above the margin.
		cfg_out[r] = best[r] < fixed_pred[r] - d->margin
			? arg[r] : static_cast<int>(clo_dispatch_fixed_config(static_cast<CloOp>(ops[r]),
				ns[r], widths[r]));
	}
	return CL_TRUE;
}

}  // extern "C"

// FP32 multilayer perceptron used as the learned cost model of the dispatcher
// (no counterpart in CL_Ops).  Shape: n_in -> 32 (ReLU) -> 32 (ReLU) -> 1.
//
// Parameters, one flat float array:
//   W1[32][n_in], b1[32], W2[32][32], b2[32], w3[32], b3
// Loss over a set of rows: L = (1/rows) * sum_r 0.5 * (y_r - t_r)^2.
//
// Kernels (hand-written; the work is tiny and irregular, so plain FMA):
//   mlp_grad      one block of kRows samples: forward, backward, and the
//                 block's gradient sums (LDS-staged activations, fixed order)
//   mlp_reduce    sums the per-block gradient rows in block order
//   mlp_adam      Adam step on every parameter
//   mlp_predict   batched inference, one thread per row
// The reductions run in a fixed order, so a gradient is bit-reproducible.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "clo_internal.hpp"
#include "dev_common.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

constexpr int kH = 32;       // hidden units
constexpr int kMaxIn = 48;   // input features
constexpr int kRows = 128;   // samples per gradient block
constexpr int kPad = kH + 1;
constexpr u32 kMaxParams = kH * kMaxIn + kH + kH * kH + kH + kH + 1;

struct Layout {
	u32 in, w1, b1, w2, b2, w3, b3, count;
};

__host__ __device__ inline Layout layout(u32 in) {
	Layout L;
	L.in = in;
	L.w1 = 0;
	L.b1 = kH * in;
	L.w2 = L.b1 + kH;
	L.b2 = L.w2 + kH * kH;
	L.w3 = L.b2 + kH;
	L.b3 = L.w3 + kH;
	L.count = L.b3 + 1;
	return L;
}

// Forward pass of one row held in registers; returns y and fills z1/z2.
__device__ __forceinline__ float forward_row(const float* __restrict__ p, const Layout& L,
	const float* x, float (&z1)[kH], float (&z2)[kH]) {
#pragma unroll
	for (int j = 0; j < kH; ++j) {
		float a = p[L.b1 + j];
#pragma unroll
		for (int i = 0; i < kMaxIn; ++i)
			if (static_cast<u32>(i) < L.in) a = fmaf(p[L.w1 + j * L.in + i], x[i], a);
		z1[j] = a;
	}
#pragma unroll
	for (int j = 0; j < kH; ++j) {
		float a = p[L.b2 + j];
#pragma unroll
		for (int k = 0; k < kH; ++k) a = fmaf(p[L.w2 + j * kH + k], fmaxf(z1[k], 0.f), a);
		z2[j] = a;
	}
	float y = p[L.b3];
#pragma unroll
	for (int j = 0; j < kH; ++j) y = fmaf(p[L.w3 + j], fmaxf(z2[j], 0.f), y);
	return y;
}

// part[block][0..count-1] = gradient sums of this block's rows (already
// scaled by `scale` = 1/rows); part[block][count] = its loss sum.
__global__ __launch_bounds__(kRows) void mlp_grad(const float* __restrict__ X,
	const float* __restrict__ T, u32 rows, u32 in, const float* __restrict__ prm,
	float* __restrict__ part, float scale) {
	__shared__ float s_p[kMaxParams];
	__shared__ float s_x[kRows][kMaxIn + 1];
	__shared__ float s_h1[kRows][kPad];
	__shared__ float s_h2[kRows][kPad];
	__shared__ float s_d1[kRows][kPad];
	__shared__ float s_d2[kRows][kPad];
	__shared__ float s_dy[kRows];
	__shared__ float s_loss[kRows / kWave];
	const Layout L = layout(in);
	const u32 tid = threadIdx.x;
	for (u32 i = tid; i < L.count; i += kRows) s_p[i] = prm[i];
	const u32 r = blockIdx.x * kRows + tid;
	const bool valid = r < rows;
	float x[kMaxIn];
#pragma unroll
	for (int i = 0; i < kMaxIn; ++i) {
		x[i] = (valid && static_cast<u32>(i) < in) ? X[static_cast<size_t>(r) * in + i] : 0.f;
		s_x[tid][i] = x[i];
	}
	__syncthreads();
	float z1[kH], z2[kH];
	const float y = forward_row(s_p, L, x, z1, z2);
	const float e = valid ? y - T[r] : 0.f;
	const float dy = e * scale;
	float d2[kH];
#pragma unroll
	for (int j = 0; j < kH; ++j) d2[j] = z2[j] > 0.f ? dy * s_p[L.w3 + j] : 0.f;
#pragma unroll
	for (int k = 0; k < kH; ++k) {
		float a = 0.f;
#pragma unroll
		for (int j = 0; j < kH; ++j) a = fmaf(s_p[L.w2 + j * kH + k], d2[j], a);
		s_d1[tid][k] = z1[k] > 0.f ? a : 0.f;
		s_h1[tid][k] = fmaxf(z1[k], 0.f);
		s_h2[tid][k] = fmaxf(z2[k], 0.f);
		s_d2[tid][k] = d2[k];
	}
	s_dy[tid] = dy;
	float lsum = wave_sum(0.5f * e * e * scale);
	if (lane_id() == 0) s_loss[wave_id()] = lsum;
	__syncthreads();
	float* __restrict__ out = part + static_cast<size_t>(blockIdx.x) * (L.count + 1);
	for (u32 q = tid; q < L.count; q += kRows) {
		float g = 0.f;
		if (q < L.b1) {
			const u32 j = q / in, i = q % in;
			for (int t = 0; t < kRows; ++t) g = fmaf(s_d1[t][j], s_x[t][i], g);
		} else if (q < L.w2) {
			const u32 j = q - L.b1;
			for (int t = 0; t < kRows; ++t) g += s_d1[t][j];
		} else if (q < L.b2) {
			const u32 j = (q - L.w2) / kH, k = (q - L.w2) % kH;
			for (int t = 0; t < kRows; ++t) g = fmaf(s_d2[t][j], s_h1[t][k], g);
		} else if (q < L.w3) {
			const u32 j = q - L.b2;
			for (int t = 0; t < kRows; ++t) g += s_d2[t][j];
		} else if (q < L.b3) {
			const u32 j = q - L.w3;
			for (int t = 0; t < kRows; ++t) g = fmaf(s_dy[t], s_h2[t][j], g);
		} else {
			for (int t = 0; t < kRows; ++t) g += s_dy[t];
		}
		out[q] = g;
	}
	if (tid == 0) {
		float l = 0.f;
		for (int w = 0; w < kRows / kWave; ++w) l += s_loss[w];
		out[L.count] = l;
	}
}

// grad[q] = sum over blocks b (in order) of part[b][q], q <= count.
__global__ __launch_bounds__(256) void mlp_reduce(const float* __restrict__ part, u32 blocks,
	u32 width, float* __restrict__ grad) {
	const u32 q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= width) return;
	float g = 0.f;
	for (u32 b = 0; b < blocks; ++b) g += part[static_cast<size_t>(b) * width + q];
	grad[q] = g;
}

// Adam (Kingma & Ba 2015) with bias correction folded into the host-computed
// step size: w -= lr_t * m / (sqrt(v) + eps_t).
__global__ __launch_bounds__(256) void mlp_adam(float* __restrict__ w, float* __restrict__ m,
	float* __restrict__ v, const float* __restrict__ g, u32 count, float b1, float b2, float lr_t,
	float eps_t) {
	const u32 q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= count) return;
	const float gq = g[q];
	const float mq = b1 * m[q] + (1.f - b1) * gq;
	const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
	m[q] = mq;
	v[q] = vq;
	w[q] -= lr_t * mq / (sqrtf(vq) + eps_t);
}

__global__ __launch_bounds__(256) void mlp_predict(const float* __restrict__ X, u32 rows, u32 in,
	const float* __restrict__ prm, float* __restrict__ out) {
	__shared__ float s_p[kMaxParams];
	const Layout L = layout(in);
	for (u32 i = threadIdx.x; i < L.count; i += blockDim.x) s_p[i] = prm[i];
	__syncthreads();
	const u32 r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= rows) return;
	float x[kMaxIn];
#pragma unroll
	for (int i = 0; i < kMaxIn; ++i)
		x[i] = static_cast<u32>(i) < in ? X[static_cast<size_t>(r) * in + i] : 0.f;
	float z1[kH], z2[kH];
	out[r] = forward_row(s_p, L, x, z1, z2);
}

// splitmix64 finaliser (Steele, Lea & Flood 2014) for the weight init.
inline u64 mix64(u64 z) {
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

}  // namespace

struct clo_mlp {
	CCLContext* ctx;
	Layout L;
	float* w = nullptr;  // device parameters
	float* m = nullptr;  // Adam moments
	float* v = nullptr;
	u32 step = 0;
	Scratch x, t, part, grad, out;
};

namespace {

bool upload(Scratch& s, const void* host, size_t bytes, CCLQueue* q, GError** err) {
	if (!s.ensure(bytes, err)) return false;
	return !hip_failed(hipMemcpyAsync(s.ptr, host, bytes, hipMemcpyHostToDevice, q->stream), err,
		"hipMemcpyAsync H2D");
}

// Gradient of the mean loss over X/T already on the device (in m->x, m->t).
bool grad_device(CloMlp* m, CCLQueue* q, u32 rows, GError** err) {
	const u32 blocks = static_cast<u32>(div_ceil(rows, kRows));
	const u32 width = m->L.count + 1;
	if (!m->part.ensure(static_cast<size_t>(blocks) * width * 4, err)) return false;
	if (!m->grad.ensure(static_cast<size_t>(width) * 4, err)) return false;
	{
		Launch Lk(q, "mlp_grad");
		mlp_grad<<<blocks, kRows, 0, q->stream>>>(static_cast<const float*>(m->x.ptr),
			static_cast<const float*>(m->t.ptr), rows, m->L.in, m->w,
			static_cast<float*>(m->part.ptr), 1.0f / static_cast<float>(rows));
		if (!Lk.done(err)) return false;
	}
	{
		Launch Lk(q, "mlp_reduce");
		mlp_reduce<<<div_ceil(width, 256), 256, 0, q->stream>>>(
			static_cast<const float*>(m->part.ptr), blocks, width, static_cast<float*>(m->grad.ptr));
		if (!Lk.done(err)) return false;
	}
	return true;
}

bool check_rows(CloMlp* m, CCLQueue* q, size_t rows, GError** err) {
	if (!m || !q) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL mlp or queue");
		return false;
	}
	if (rows == 0 || rows > 0x7FFFFFFFull) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "rows must be in 1..2^31-1");
		return false;
	}
	return use_device(m->ctx, err);
}

}  // namespace

extern "C" {

CloMlp* clo_mlp_new(CCLContext* ctx, cl_uint n_in, cl_ulong seed, GError** err) {
	if (err && *err) return nullptr;
	if (!ctx || n_in == 0 || n_in > kMaxIn) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "clo_mlp_new: n_in must be 1..%d", kMaxIn);
		return nullptr;
	}
	if (!use_device(ctx, err)) return nullptr;
	CloMlp* m = new clo_mlp();
	m->ctx = ctx;
	ccl_context_ref(ctx);
	m->L = layout(n_in);
	const size_t bytes = static_cast<size_t>(m->L.count) * 4;
	if (hip_failed(dev_malloc(reinterpret_cast<void**>(&m->w), bytes), err, "hipMalloc") ||
		hip_failed(dev_malloc(reinterpret_cast<void**>(&m->m), bytes), err, "hipMalloc") ||
		hip_failed(dev_malloc(reinterpret_cast<void**>(&m->v), bytes), err, "hipMalloc")) {
		clo_mlp_destroy(m);
		return nullptr;
	}
	// He-uniform weights (bound sqrt(6 / fan_in)), zero biases.
	std::vector<float> p(m->L.count, 0.f);
	u64 ctr = seed;
	auto uni = [&]() {
		ctr += 0x9E3779B97F4A7C15ull;
		return static_cast<float>((mix64(ctr) >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f;
	};
	const float a1 = std::sqrt(6.f / n_in), a2 = std::sqrt(6.f / kH);
	for (u32 i = 0; i < kH * n_in; ++i) p[m->L.w1 + i] = a1 * uni();
	for (u32 i = 0; i < kH * kH; ++i) p[m->L.w2 + i] = a2 * uni();
	for (u32 i = 0; i < kH; ++i) p[m->L.w3 + i] = a2 * uni();
	if (!clo_mlp_set_params(m, p.data(), err)) {
		clo_mlp_destroy(m);
		return nullptr;
	}
	return m;
}

void clo_mlp_destroy(CloMlp* m) {
	if (!m) return;
	(void)hipSetDevice(m->ctx->dev);
	if (m->w) (void)hipFree(m->w);
	if (m->m) (void)hipFree(m->m);
	if (m->v) (void)hipFree(m->v);
	m->x.release();
	m->t.release();
	m->part.release();
	m->grad.release();
	m->out.release();
	ccl_context_destroy(m->ctx);
	delete m;
}

cl_uint clo_mlp_num_params(CloMlp* m) { return m ? m->L.count : 0; }
cl_uint clo_mlp_num_inputs(CloMlp* m) { return m ? m->L.in : 0; }
cl_uint clo_mlp_num_hidden(CloMlp* m) { return m ? static_cast<cl_uint>(kH) : 0; }

cl_bool clo_mlp_get_params(CloMlp* m, float* out, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!m || !out) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL mlp or output");
		return CL_FALSE;
	}
	if (!use_device(m->ctx, err)) return CL_FALSE;
	return hip_failed(hipMemcpy(out, m->w, m->L.count * 4, hipMemcpyDeviceToHost), err,
		"hipMemcpy D2H") ? CL_FALSE : CL_TRUE;
}

// Also resets the Adam state.
cl_bool clo_mlp_set_params(CloMlp* m, const float* in, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!m || !in) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL mlp or input");
		return CL_FALSE;
	}
	if (!use_device(m->ctx, err)) return CL_FALSE;
	const size_t bytes = m->L.count * 4;
	if (hip_failed(hipMemcpy(m->w, in, bytes, hipMemcpyHostToDevice), err, "hipMemcpy H2D") ||
		hip_failed(hipMemset(m->m, 0, bytes), err, "hipMemset") ||
		hip_failed(hipMemset(m->v, 0, bytes), err, "hipMemset") ||
		// the fills run on the null stream, which queue streams do not wait for
		hip_failed(hipStreamSynchronize(nullptr), err, "hipStreamSynchronize"))
		return CL_FALSE;
	m->step = 0;
	return CL_TRUE;
}

cl_bool clo_mlp_gradient(CloMlp* m, CCLQueue* q, const float* X, const float* t, size_t rows,
	float* grad, float* loss, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!check_rows(m, q, rows, err)) return CL_FALSE;
	if (!X || !t || !grad) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL host array");
		return CL_FALSE;
	}
	if (!upload(m->x, X, rows * m->L.in * 4, q, err) || !upload(m->t, t, rows * 4, q, err) ||
		!grad_device(m, q, static_cast<u32>(rows), err))
		return CL_FALSE;
	std::vector<float> g(m->L.count + 1);
	if (hip_failed(hipMemcpyAsync(g.data(), m->grad.ptr, g.size() * 4, hipMemcpyDeviceToHost,
			q->stream), err, "hipMemcpyAsync D2H") ||
		hip_failed(hipStreamSynchronize(q->stream), err, "hipStreamSynchronize"))
		return CL_FALSE;
	memcpy(grad, g.data(), m->L.count * 4);
	if (loss) *loss = g[m->L.count];
	return CL_TRUE;
}

cl_bool clo_mlp_train(CloMlp* m, CCLQueue* q, const float* X, const float* t, size_t rows,
	cl_uint steps, float lr, float* final_loss, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!check_rows(m, q, rows, err)) return CL_FALSE;
	if (!X || !t || !(lr > 0.f)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL host array or lr <= 0");
		return CL_FALSE;
	}
	if (!upload(m->x, X, rows * m->L.in * 4, q, err) || !upload(m->t, t, rows * 4, q, err))
		return CL_FALSE;
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	const u32 cnt = m->L.count;
	for (cl_uint s = 0; s < steps; ++s) {
		if (!grad_device(m, q, static_cast<u32>(rows), err)) return CL_FALSE;
		++m->step;
		const double c1 = 1.0 - std::pow(0.9, m->step), c2 = 1.0 - std::pow(0.999, m->step);
		// lr * mhat / (sqrt(vhat) + eps) == lr_t * m / (sqrt(v) + eps_t)
		const float lr_t = static_cast<float>(lr * std::sqrt(c2) / c1);
		const float eps_t = static_cast<float>(eps * std::sqrt(c2));
		Launch Lk(q, "mlp_adam");
		mlp_adam<<<div_ceil(cnt, 256), 256, 0, q->stream>>>(m->w, m->m, m->v,
			static_cast<const float*>(m->grad.ptr), cnt, b1, b2, lr_t, eps_t);
		if (!Lk.done(err)) return CL_FALSE;
	}
	if (final_loss) {
		if (!grad_device(m, q, static_cast<u32>(rows), err)) return CL_FALSE;
		float l = 0.f;
		if (hip_failed(hipMemcpyAsync(&l, static_cast<float*>(m->grad.ptr) + cnt, 4,
				hipMemcpyDeviceToHost, q->stream), err, "hipMemcpyAsync D2H") ||
			hip_failed(hipStreamSynchronize(q->stream), err, "hipStreamSynchronize"))
			return CL_FALSE;
		*final_loss = l;
	}
	return CL_TRUE;
}

CCLEvent* clo_mlp_predict_device(CloMlp* m, CCLQueue* q, CCLBuffer* X, size_t rows,
	CCLBuffer* out, GError** err) {
	if (err && *err) return nullptr;
	if (!check_rows(m, q, rows, err)) return nullptr;
	if (!X || !out || X->size < rows * m->L.in * 4 || out->size < rows * 4) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "feature or output buffer too small");
		return nullptr;
	}
	Launch Lk(q, "mlp_predict");
	mlp_predict<<<div_ceil(rows, 256), 256, 0, q->stream>>>(static_cast<const float*>(X->ptr),
		static_cast<u32>(rows), m->L.in, m->w, static_cast<float*>(out->ptr));
	if (!Lk.done(err)) return nullptr;
	return record_event(q, "mlp_predict", err);
}

cl_bool clo_mlp_predict(CloMlp* m, CCLQueue* q, const float* X, size_t rows, float* out,
	GError** err) {
	if (err && *err) return CL_FALSE;
	if (!check_rows(m, q, rows, err)) return CL_FALSE;
	if (!X || !out) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL host array");
		return CL_FALSE;
	}
	if (!upload(m->x, X, rows * m->L.in * 4, q, err) || !m->out.ensure(rows * 4, err))
		return CL_FALSE;
	{
		Launch Lk(q, "mlp_predict");
		mlp_predict<<<div_ceil(rows, 256), 256, 0, q->stream>>>(
			static_cast<const float*>(m->x.ptr), static_cast<u32>(rows), m->L.in, m->w,
			static_cast<float*>(m->out.ptr));
		if (!Lk.done(err)) return CL_FALSE;
	}
	return (hip_failed(hipMemcpyAsync(out, m->out.ptr, rows * 4, hipMemcpyDeviceToHost, q->stream),
				err, "hipMemcpyAsync D2H") ||
			   hip_failed(hipStreamSynchronize(q->stream), err, "hipStreamSynchronize"))
		? CL_FALSE
		: CL_TRUE;
}

// File format (little-endian): "CLOMLP01", u32 n_in, u32 hidden, u32 count,
// then count float32 parameters.
cl_bool clo_mlp_save(CloMlp* m, const char* path, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!m || !path) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL mlp or path");
		return CL_FALSE;
	}
	std::vector<float> p(m->L.count);
	if (!clo_mlp_get_params(m, p.data(), err)) return CL_FALSE;
	FILE* f = fopen(path, "wb");
	if (!f) {
		set_error(err, kQuarkClo, CLO_ERROR_STREAM_WRITE, "cannot open %s", path);
		return CL_FALSE;
	}
	const u32 hdr[3] = {m->L.in, static_cast<u32>(kH), m->L.count};
	bool ok = fwrite("CLOMLP01", 1, 8, f) == 8 && fwrite(hdr, 4, 3, f) == 3 &&
		fwrite(p.data(), 4, p.size(), f) == p.size();
	ok = (fclose(f) == 0) && ok;
	if (!ok) set_error(err, kQuarkClo, CLO_ERROR_STREAM_WRITE, "cannot write %s", path);
	return ok ? CL_TRUE : CL_FALSE;
}

CloMlp* clo_mlp_load(CCLContext* ctx, const char* path, GError** err) {
	if (err && *err) return nullptr;
	FILE* f = path ? fopen(path, "rb") : nullptr;
	if (!f) {
		set_error(err, kQuarkClo, CLO_ERROR_OPENFILE, "cannot open %s", path ? path : "(null)");
		return nullptr;
	}
	char magic[8];
	u32 hdr[3];
	bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "CLOMLP01", 8) == 0 &&
		fread(hdr, 4, 3, f) == 3 && hdr[0] >= 1 && hdr[0] <= kMaxIn && hdr[1] == kH &&
		hdr[2] == layout(hdr[0]).count;
	std::vector<float> p(ok ? hdr[2] : 0);
	ok = ok && fread(p.data(), 4, p.size(), f) == p.size();
	fclose(f);
	if (!ok) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "%s is not a CLOMLP01 file of this shape", path);
		return nullptr;
	}
	CloMlp* m = clo_mlp_new(ctx, hdr[0], 0, err);
	if (!m) return nullptr;
	if (!clo_mlp_set_params(m, p.data(), err)) {
		clo_mlp_destroy(m);
		return nullptr;
	}
	return m;
}

}  // extern "C"

// gselect: stable O(n^2) rank sort (CL_Ops' src/cl_ops/sort/clo_sort_gselect.
// {c:31-243,cl:38-57}) for gfx950.
//
// out[pos(i)] = in[i] with pos(i) = #{j : key_j sorts before key_i}
//                                 + #{j < i : key_j == key_i},
// which is the stable sort.  Each block ranks 256 elements against the whole
// input streamed through LDS in tiles.  Order and signedness use the same XOR
// mask as the bitonic kernels (compare unsigned words ascending).

#include <cstring>

#include "clo_internal.hpp"
#include "dev_common.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

constexpr int kGsBlock = 256;
constexpr int kGsTile = 1024;

template <typename T, bool KV>
__global__ __launch_bounds__(kGsBlock) void gselect_rank(const T* __restrict__ in,
	const u32* __restrict__ vin, T* __restrict__ out, u32* __restrict__ vout, u32 n, T xmask) {
	__shared__ T s_k[kGsTile];
	const u32 i = blockIdx.x * kGsBlock + threadIdx.x;
	const bool mine = i < n;
	const T ki = mine ? static_cast<T>(in[i] ^ xmask) : T(0);
	u32 pos = 0;
	for (u32 base = 0; base < n; base += kGsTile) {
		const u32 cnt = (n - base) < kGsTile ? (n - base) : kGsTile;
		__syncthreads();
		for (u32 j = threadIdx.x; j < cnt; j += kGsBlock) s_k[j] = static_cast<T>(in[base + j] ^ xmask);
		__syncthreads();
		if (mine) {
			for (u32 j = 0; j < cnt; ++j) {
				const T kj = s_k[j];
				pos += (kj < ki) || (kj == ki && base + j < i);
			}
		}
	}
	if (mine) {
		out[pos] = in[i];
		if (KV) vout[pos] = vin[i];
	}
}

struct GselectData {
	Scratch tmp_keys;
	Scratch tmp_vals;
};

template <typename T>
T xmask_for(const CloSort* s) {
	T m = 0;
	if (type_is_signed(s->elem_type)) m ^= static_cast<T>(T(1) << (sizeof(T) * 8 - 1));
	if (s->descending) m ^= static_cast<T>(~T(0));
	return m;
}

template <typename T, bool KV>
bool launch_rank(CloSort* sorter, CCLQueue* q, const void* in, const u32* vin, void* out,
	u32* vout, u32 n, GError** err) {
	const u32 grid = (n + kGsBlock - 1) / kGsBlock;
	Launch L(q, CLO_SORT_GSELECT_KNAME);
	gselect_rank<T, KV><<<grid, kGsBlock, 0, q->stream>>>(static_cast<const T*>(in), vin,
		static_cast<T*>(out), vout, n, xmask_for<T>(sorter));
	return L.done(err);
}

template <bool KV>
bool rank_dispatch(CloSort* sorter, CCLQueue* q, const void* in, const u32* vin, void* out,
	u32* vout, u32 n, GError** err) {
	switch (clo_type_sizeof(sorter->elem_type)) {
	case 1: return launch_rank<u8, KV>(sorter, q, in, vin, out, vout, n, err);
	case 2: return launch_rank<u16, KV>(sorter, q, in, vin, out, vout, n, err);
	case 4: return launch_rank<u32, KV>(sorter, q, in, vin, out, vout, n, err);
	case 8: return launch_rank<u64, KV>(sorter, q, in, vin, out, vout, n, err);
	}
	return false;
}

CCLEvent* gselect_common(CloSort* sorter, CCLQueue* cq_exec, CCLBuffer* kin, CCLBuffer* vin,
	CCLBuffer* kout, CCLBuffer* vout, size_t numel, GError** err) {
	GselectData* gd = static_cast<GselectData*>(sorter->data);
	const size_t es = clo_type_sizeof(sorter->elem_type);
	if (numel == 0) return record_event(cq_exec, "gselect_done", err);
	const bool copy_back = (kout == nullptr || kout->ptr == kin->ptr);
	void* ko = nullptr;
	u32* vo = nullptr;
	if (copy_back) {
		// clo_sort_gselect.c:82-100: rank into a temporary, then copy back.
		if (!gd->tmp_keys.ensure(numel * es, err)) return nullptr;
		if (vin && !gd->tmp_vals.ensure(numel * 4, err)) return nullptr;
		ko = gd->tmp_keys.ptr;
		vo = vin ? static_cast<u32*>(gd->tmp_vals.ptr) : nullptr;
	} else {
		ko = kout->ptr;
		vo = vin ? static_cast<u32*>(vout->ptr) : nullptr;
	}
	const u32 n = static_cast<u32>(numel);
	bool ok = vin ? rank_dispatch<true>(sorter, cq_exec, kin->ptr, static_cast<u32*>(vin->ptr),
		ko, vo, n, err) : rank_dispatch<false>(sorter, cq_exec, kin->ptr, nullptr, ko, nullptr, n,
		err);
	if (!ok) return nullptr;
	if (copy_back) {
		if (hip_failed(hipMemcpyAsync(kin->ptr, ko, numel * es, hipMemcpyDeviceToDevice,
				cq_exec->stream), err, "gselect_copy"))
			return nullptr;
		if (vin && hip_failed(hipMemcpyAsync(vin->ptr, vo, numel * 4, hipMemcpyDeviceToDevice,
				cq_exec->stream), err, "gselect_copy"))
			return nullptr;
	}
	return record_event(cq_exec, "gselect_done", err);
}

const char* gselect_init(CloSort* sorter, const char* options, GError** err) {
	(void)options;  // clo_sort_gselect.c ignores options
	(void)err;
	sorter->data = new GselectData();
	return "gselect (gfx950: LDS-tiled stable rank)";
}

void gselect_finalize(CloSort* sorter) {
	delete static_cast<GselectData*>(sorter->data);
	sorter->data = nullptr;
}

CCLEvent* gselect_sort(CloSort* sorter, CCLQueue* cq_exec, CCLQueue* cq_comm, CCLBuffer* in,
	CCLBuffer* out, size_t numel, size_t lws_max, GError** err) {
	(void)cq_comm;
	(void)lws_max;
	return gselect_common(sorter, cq_exec, in, nullptr, out, nullptr, numel, err);
}

cl_uint gselect_num_kernels(CloSort*, GError**) { return 1; }
const char* gselect_kernel_name(CloSort*, cl_uint i, GError** err) {
	if (i != 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "kernel index %u out of range", i);
		return nullptr;
	}
	return CLO_SORT_GSELECT_KNAME;
}
size_t gselect_localmem(CloSort* sorter, cl_uint i, size_t, size_t, GError** err) {
	if (i != 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "kernel index %u out of range", i);
		return 0;
	}
	return kGsTile * clo_type_sizeof(sorter->elem_type);
}

}  // namespace

CCLEvent* clo_gselect_sort_kv(CloSort* sorter, CCLQueue* cq_exec, CCLQueue* cq_comm,
	CCLBuffer* keys_in, CCLBuffer* vals_in, CCLBuffer* keys_out, CCLBuffer* vals_out, size_t numel,
	GError** err) {
	(void)cq_comm;
	return gselect_common(sorter, cq_exec, keys_in, vals_in, keys_out, vals_out, numel, err);
}

// Host-pass only: the vtable holds host function addresses.
#if !defined(__HIP_DEVICE_COMPILE__)
extern "C" const CloSortImplDef clo_sort_gselect_def = {"gselect", CL_FALSE, gselect_init,
	gselect_finalize, gselect_sort, gselect_num_kernels, gselect_kernel_name, gselect_localmem};
#endif


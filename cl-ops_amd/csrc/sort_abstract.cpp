// Sorter factory, dispatch and host-data paths
// (CL_Ops' src/cl_ops/sort/clo_sort_abstract.c:91-629).

#include <cstring>

#include "clo_internal.hpp"

using namespace clo;

extern "C" {

CloSort* clo_sort_new(const char* type, const char* options, CCLContext* ctx,
	CloType* elem_type, CloType* key_type, const char* compare, const char* get_key,
	const char* compiler_opts, GError** err) {
	if (err && *err) return nullptr;
	if (!ctx || !elem_type) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "clo_sort_new: NULL context or element type");
		return nullptr;
	}
	// clo_sort_abstract.c:102-121: search the implementation by name.
	const CloSortImplDef* defs[] = {&clo_sort_sbitonic_def, &clo_sort_abitonic_def,
		&clo_sort_gselect_def, &clo_sort_satradix_def};
	const CloSortImplDef* impl = nullptr;
	for (const CloSortImplDef* d : defs)
		if (type && strcmp(type, d->name) == 0) impl = d;
	if (!impl) {
		set_error(err, kQuarkClo, CLO_ERROR_IMPL_NOT_FOUND,
			"The requested sort implementation, '%s', was not found.", type ? type : "(null)");
		return nullptr;
	}
	CloType et = *elem_type;
	CloType kt = key_type ? *key_type : et;
	if (!clo_type_get_name(et) || !clo_type_get_name(kt)) {
		set_error(err, kQuarkClo, CLO_ERROR_UNKNOWN_TYPE, "Unknown element or key type.");
		return nullptr;
	}
	CloSort* s = new clo_sort();
	s->impl = *impl;
	s->ctx = ctx;
	ccl_context_ref(ctx);
	s->elem_type = et;
	s->key_type = kt;
	s->descending = 0;
	s->data = nullptr;
	s->last_queue = nullptr;
	s->seg_tile = 0;
	s->seg_grid = 0;
	s->seg_cls[0] = s->seg_cls[1] = 0;
	s->seg_msd = -1;
	s->seg_cache = nullptr;
	s->snip = nullptr;
	s->host_cache = nullptr;
	GError* e = nullptr;
	const char* src = s->impl.init(s, options, &e);
	if (!src || e) {
		if (e) {
			if (err && !*err) *err = e;
			else clo_error_free(e);
		} else {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS, "initialisation of %s failed", type);
		}
		clo_sort_destroy(s);
		return nullptr;
	}
	// clo_sort_abstract.c:140-186: the sorter's kernels with the element
	// type, key type, compare and key-getter macros.  Integer elements sorted
	// by themselves, ascending or descending, run the ahead-of-time kernels;
	// everything else is compiled now.
	if (snippet_sorter_needed(s, compare, get_key)) {
		if (!snippet_sorter_init(s, compare, get_key, compiler_opts, err)) {
			clo_sort_destroy(s);
			return nullptr;
		}
	} else if (strcmp(s->impl.name, "satradix") != 0) {
		s->descending = compare_kind(compare);
	}
	return s;
}

void clo_sort_destroy(CloSort* sorter) {
	if (!sorter) return;
	if (sorter->impl.finalize) sorter->impl.finalize(sorter);
	if (sorter->seg_cache) clo_satradix_segsort_cache_free(sorter->seg_cache);
	if (sorter->snip) snippet_sorter_free(sorter);
	if (sorter->host_cache) host_cache_free(sorter->host_cache);
	ccl_context_destroy(sorter->ctx);
	delete sorter;
}

static bool check_device_args(CloSort* s, CCLQueue* cq, CCLBuffer* in, CCLBuffer* out,
	size_t numel, size_t elem, GError** err) {
	if (!s || !cq || !in) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL sorter, queue or input buffer");
		return false;
	}
	// 32-bit element indices, as the reference's kernels (clo_nlpo2 caps n at 2^31).
	if (numel > 0x80000000ull) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "numel %zu exceeds 2^31", numel);
		return false;
	}
	if (in->size < numel * elem || (out && out->size < numel * elem)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"buffer smaller than numel (%zu) elements of %zu bytes", numel, elem);
		return false;
	}
	return use_device(cq->ctx, err);
}

// True when the byte ranges [a, a + an) and [b, b + bn) share a byte.
static bool ranges_overlap(const CCLBuffer* a, size_t an, const CCLBuffer* b, size_t bn) {
	if (!a || !b || !an || !bn) return false;
	const uintptr_t x = reinterpret_cast<uintptr_t>(a->ptr);
	const uintptr_t y = reinterpret_cast<uintptr_t>(b->ptr);
	return x < y + bn && y < x + an;
}

// Where the output of `bytes` bytes goes relative to the input: 1 = in place
// (out NULL or the same address), 0 = a disjoint buffer, -1 = a partial
// overlap, which no algorithm here can sort through.
static int placement(const CCLBuffer* in, const CCLBuffer* out, size_t bytes) {
	if (!out || out->ptr == in->ptr) return 1;
	return ranges_overlap(in, bytes, out, bytes) ? -1 : 0;
}

// Key-value buffers: keys and values are both sorted in place or both out of
// place, and no two of the distinct buffers share a byte.  On success, an
// in-place call has keys_out and vals_out set to NULL.
static bool check_kv_placement(CCLBuffer* keys_in, CCLBuffer* vals_in, CCLBuffer** keys_out,
	CCLBuffer** vals_out, size_t kbytes, size_t vbytes, GError** err) {
	const int pk = placement(keys_in, *keys_out, kbytes);
	const int pv = placement(vals_in, *vals_out, vbytes);
	if (pk < 0 || pv < 0 || pk != pv) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"key-value sort: keys and values must both be sorted in place (output NULL or "
			"equal to the input) or both into buffers disjoint from the inputs");
		return false;
	}
	if (pk == 1) {
		*keys_out = nullptr;
		*vals_out = nullptr;
	}
	const CCLBuffer* kb[2] = {keys_in, *keys_out};
	const CCLBuffer* vb[2] = {vals_in, *vals_out};
	bool clash = false;
	for (const CCLBuffer* k : kb)
		for (const CCLBuffer* v : vb) clash = clash || ranges_overlap(k, kbytes, v, vbytes);
	if (clash) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"key-value sort: key and value buffers overlap");
		return false;
	}
	return true;
}

CCLEvent* clo_sort_with_device_data(CloSort* sorter, CCLQueue* cq_exec, CCLQueue* cq_comm,
	CCLBuffer* data_in, CCLBuffer* data_out, size_t numel, size_t lws_max, GError** err) {
	ApiRange api_range("clo_sort_with_device_data");
	if (err && *err) return nullptr;
	if (!check_device_args(sorter, cq_exec, data_in, data_out, numel,
			sorter ? clo_type_sizeof(sorter->elem_type) : 0, err))
		return nullptr;
	if (placement(data_in, data_out, numel * clo_type_sizeof(sorter->elem_type)) < 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"data_out overlaps data_in without being the same buffer");
		return nullptr;
	}
	if (sorter->snip)
		return snippet_sort(sorter, cq_exec, data_in, nullptr, data_out, nullptr, numel, err);
	return sorter->impl.sort_with_device_data(sorter, cq_exec, cq_comm, data_in, data_out,
		numel, lws_max, err);
}

cl_bool clo_sort_with_host_data(CloSort* sorter, CCLQueue* cq_exec, CCLQueue* cq_comm,
	void* data_in, void* data_out, size_t numel, size_t lws_max, GError** err) {
	ApiRange api_range("clo_sort_with_host_data");
	(void)lws_max;
	if (err && *err) return CL_FALSE;
	if (!sorter || !data_in || !data_out) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL sorter or host pointer");
		return CL_FALSE;
	}
	// clo_sort_abstract.c:333-341: internal queue when cq_exec is NULL.
	CCLQueue* own = nullptr;
	if (!cq_exec) {
		own = ccl_queue_new(sorter->ctx, nullptr, 0, err);
		if (!own) return CL_FALSE;
		cq_exec = own;
	}
	if (!cq_comm) cq_comm = cq_exec;
	const bool ok = sort_host_path(sorter, cq_exec, cq_comm, data_in, nullptr, data_out, numel,
		err);
	if (own) ccl_queue_destroy(own);
	return ok ? CL_TRUE : CL_FALSE;
}

CCLEvent* clo_sort_kv_with_device_data(CloSort* sorter, CCLQueue* cq_exec, CCLQueue* cq_comm,
	CCLBuffer* keys_in, CCLBuffer* vals_in, CCLBuffer* keys_out, CCLBuffer* vals_out,
	size_t numel, size_t lws_max, GError** err) {
	ApiRange api_range("clo_sort_kv_with_device_data");
	(void)lws_max;
	if (err && *err) return nullptr;
	if (!check_device_args(sorter, cq_exec, keys_in, keys_out, numel,
			sorter ? clo_type_sizeof(sorter->elem_type) : 0, err))
		return nullptr;
	if (!vals_in || vals_in->size < numel * 4 || (vals_out && vals_out->size < numel * 4)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"KV sort needs uint32 value buffers of numel elements");
		return nullptr;
	}
	if (!check_kv_placement(keys_in, vals_in, &keys_out, &vals_out,
			numel * clo_type_sizeof(sorter->elem_type), numel * 4, err))
		return nullptr;
	if (sorter->snip && (strcmp(sorter->impl.name, "satradix") == 0 ||
			strcmp(sorter->impl.name, "gselect") == 0))
		return snippet_sort(sorter, cq_exec, keys_in, vals_in, keys_out, vals_out, numel, err);
	if (strcmp(sorter->impl.name, "satradix") == 0)
		return clo_satradix_sort_kv(sorter, cq_exec, cq_comm, keys_in, vals_in, keys_out,
			vals_out, numel, err);
	if (strcmp(sorter->impl.name, "gselect") == 0)
		return clo_gselect_sort_kv(sorter, cq_exec, cq_comm, keys_in, vals_in, keys_out,
			vals_out, numel, err);
	set_error(err, kQuarkClo, CLO_ERROR_ARGS,
		"'%s' is not a stable sort; key-value sorting needs satradix or gselect",
		sorter->impl.name);
	return nullptr;
}

cl_bool clo_sort_kv_with_host_data(CloSort* sorter, CCLQueue* cq_exec, CCLQueue* cq_comm,
	void* keys, void* vals, size_t numel, size_t lws_max, GError** err) {
	ApiRange api_range("clo_sort_kv_with_host_data");
	(void)lws_max;
	if (err && *err) return CL_FALSE;
	if (!sorter || !keys || !vals) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL sorter or host pointer");
		return CL_FALSE;
	}
	CCLQueue* own = nullptr;
	if (!cq_exec) {
		own = ccl_queue_new(sorter->ctx, nullptr, 0, err);
		if (!own) return CL_FALSE;
		cq_exec = own;
	}
	if (!cq_comm) cq_comm = cq_exec;
	const bool ok = sort_host_path(sorter, cq_exec, cq_comm, keys, vals, keys, numel, err);
	if (own) ccl_queue_destroy(own);
	return ok ? CL_TRUE : CL_FALSE;
}

CCLEvent* clo_sort_segmented_with_device_data(CloSort* sorter, CCLQueue* cq_exec,
	CCLQueue* cq_comm, CCLBuffer* keys_in, CCLBuffer* vals_in, CCLBuffer* keys_out,
	CCLBuffer* vals_out, CCLBuffer* offsets, size_t nseg, size_t numel, GError** err) {
	ApiRange api_range("clo_sort_segmented_with_device_data");
	(void)cq_comm;
	if (err && *err) return nullptr;
	if (!check_device_args(sorter, cq_exec, keys_in, keys_out, numel,
			sorter ? clo_type_sizeof(sorter->elem_type) : 0, err))
		return nullptr;
	if (strcmp(sorter->impl.name, "satradix") != 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"segmented sorting is provided by satradix sorters, not '%s'", sorter->impl.name);
		return nullptr;
	}
	if (sorter->snip) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"segmented sorting runs the built-in integer keys only, not key snippets");
		return nullptr;
	}
	if (!offsets || offsets->size < nseg * 4 || nseg > 0xFFFFFFFFull ||
		(vals_in && (vals_in->size < numel * 4 || (vals_out && vals_out->size < numel * 4))) ||
		(!vals_in && vals_out)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"segmented sort needs nseg uint32 offsets and, for key-value data, uint32 value "
			"buffers of numel elements");
		return nullptr;
	}
	const size_t kbytes = numel * clo_type_sizeof(sorter->elem_type);
	if (vals_in) {
		if (!check_kv_placement(keys_in, vals_in, &keys_out, &vals_out, kbytes, numel * 4, err))
			return nullptr;
	} else if (placement(keys_in, keys_out, kbytes) < 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"keys_out overlaps keys_in without being the same buffer");
		return nullptr;
	}
	uint32_t tile = sorter->seg_tile ? sorter->seg_tile
		: clo::seg_default_tile(clo_type_sizeof(sorter->elem_type), vals_in != nullptr);
	return clo_satradix_sort_segmented(sorter, cq_exec, keys_in, vals_in, keys_out, vals_out,
		offsets, nseg, numel, tile, sorter->seg_grid, &sorter->seg_cache, err);
}

cl_uint clo_sort_get_last_passes(CloSort* sorter, GError** err) {
	if (!sorter) return 0;
	if (strcmp(sorter->impl.name, "satradix") != 0) return 0;
	return clo_satradix_last_passes(sorter, err);
}

CCLContext* clo_sort_get_context(CloSort* sorter) { return sorter ? sorter->ctx : nullptr; }
CCLProgram* clo_sort_get_program(CloSort* sorter) {
	(void)sorter;
	return nullptr;  // no runtime-compiled program: kernels are built ahead of time
}
CloType clo_sort_get_element_type(CloSort* sorter) { return sorter->elem_type; }
size_t clo_sort_get_element_size(CloSort* sorter) { return clo_type_sizeof(sorter->elem_type); }
CloType clo_sort_get_key_type(CloSort* sorter) { return sorter->key_type; }
const char* clo_sort_get_source(CloSort* sorter) {
	return sorter ? snippet_sorter_source(sorter) : nullptr;
}
size_t clo_sort_get_key_size(CloSort* sorter) { return clo_type_sizeof(sorter->key_type); }
void* clo_sort_get_data(CloSort* sorter) { return sorter ? sorter->data : nullptr; }
void clo_sort_set_data(CloSort* sorter, void* data) {
	if (sorter) sorter->data = data;
}

cl_uint clo_sort_get_num_kernels(CloSort* sorter, GError** err) {
	if (!sorter) return 0;
	return sorter->impl.get_num_kernels(sorter, err);
}

const char* clo_sort_get_kernel_name(CloSort* sorter, cl_uint i, GError** err) {
	if (!sorter) return nullptr;
	return sorter->impl.get_kernel_name(sorter, i, err);
}

size_t clo_sort_get_localmem_usage(CloSort* sorter, cl_uint i, size_t lws_max, size_t numel,
	GError** err) {
	if (!sorter) return 0;
	return sorter->impl.get_localmem_usage(sorter, i, lws_max, numel, err);
}

}  // extern "C"

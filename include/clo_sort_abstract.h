/*
 * clo_sort_abstract.h - sorter factory and vtable.
 *
 * Mirrors CL_Ops' src/cl_ops/sort/clo_sort_abstract.in.h:30-170: same
 * struct names, vtable fields and function signatures.  The vtable's `init`
 * returns a description string instead of OpenCL source (kernels are
 * compiled ahead of time for gfx950).
 */
#ifndef CLO_SORT_ABSTRACT_H_
#define CLO_SORT_ABSTRACT_H_

#include "clo_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* clo_sort_abstract.in.h:30 */
#define CLO_SORT_IMPLS "sbitonic, abitonic, gselect, satradix"

/* clo_sort_abstract.in.h:43-110 */
typedef struct clo_sort_impl_def {
	const char* name;
	cl_bool in_place;
	const char* (*init)(CloSort* sorter, const char* options, GError** err);
	void (*finalize)(CloSort* sorter);
	CCLEvent* (*sort_with_device_data)(CloSort* sorter, CCLQueue* cq_exec,
		CCLQueue* cq_comm, CCLBuffer* data_in, CCLBuffer* data_out,
		size_t numel, size_t lws_max, GError** err);
	cl_uint (*get_num_kernels)(CloSort* sorter, GError** err);
	const char* (*get_kernel_name)(CloSort* sorter, cl_uint i, GError** err);
	size_t (*get_localmem_usage)(CloSort* sorter, cl_uint i, size_t lws_max,
		size_t numel, GError** err);
} CloSortImplDef;

/* clo_sort_abstract.c:91-207.  Arguments as in CL_Ops: element type, key
 * type (NULL: the element type), the CLO_SORT_COMPARE(a, b) snippet (NULL:
 * "((a) > (b))", ascending), the CLO_SORT_KEY_GET(x) snippet (NULL: "(x)")
 * and compiler options.  Integer elements sorted by themselves, ascending or
 * descending ("((a) < (b))"), run precompiled kernels.  Any other snippet,
 * key type or element type (half, float, double) is compiled with hipRTC
 * when the sorter is created, from the HIP source that clo_sort_get_source()
 * returns, with the four macros defined as CL_Ops defines them:
 *   - sbitonic / abitonic run CL_Ops' bitonic network exactly (ties come out
 *     in its order); sizes that are not powers of two sort as if padded with
 *     elements that order after all others;
 *   - gselect ranks with CL_Ops' formula (stable for a strict order);
 *   - satradix orders stably by the bits of KEY_GET(x), which must be an
 *     integer type; it ignores `compare`, as CL_Ops' radix kernels do.
 * Snippets may use OpenCL C type names (uint, ulong, half, ...) and the
 * as_uint / as_int / as_ulong / as_long / as_ushort / as_short / as_float /
 * as_double reinterpretations.  compiler_opts pass to hipRTC ("-D..."),
 * except OpenCL-only "-cl-*" options, which are dropped.  A snippet that does
 * not compile fails with CLO_ERROR_ARGS and the compiler log.  Segmented
 * sorting takes precompiled kernels only. */
CloSort* clo_sort_new(const char* type, const char* options,
	CCLContext* ctx, CloType* elem_type, CloType* key_type,
	const char* compare, const char* get_key, const char* compiler_opts,
	GError** err);

/* clo_sort_abstract.c:214-246 */
void clo_sort_destroy(CloSort* sorter);

/* clo_sort_abstract.c:256-273.  Asynchronous: returns the event recorded
 * after the last kernel on cq_exec.  data_out == NULL sorts data_in in
 * place; otherwise the sorted data goes to data_out and data_in is left
 * unchanged.  Any numel (not only powers of two). */
CCLEvent* clo_sort_with_device_data(CloSort* sorter, CCLQueue* cq_exec,
	CCLQueue* cq_comm, CCLBuffer* data_in, CCLBuffer* data_out,
	size_t numel, size_t lws_max, GError** err);

/* clo_sort_abstract.c:296-418.  Blocking host-to-host sort. */
cl_bool clo_sort_with_host_data(CloSort* sorter, CCLQueue* cq_exec,
	CCLQueue* cq_comm, void* data_in, void* data_out, size_t numel,
	size_t lws_max, GError** err);

/* clo_sort_abstract.c:428-629 */
CCLContext* clo_sort_get_context(CloSort* sorter);
CCLProgram* clo_sort_get_program(CloSort* sorter);
CloType clo_sort_get_element_type(CloSort* sorter);
size_t clo_sort_get_element_size(CloSort* sorter);
CloType clo_sort_get_key_type(CloSort* sorter);
size_t clo_sort_get_key_size(CloSort* sorter);
void* clo_sort_get_data(CloSort* sorter);
void clo_sort_set_data(CloSort* sorter, void* data);
cl_uint clo_sort_get_num_kernels(CloSort* sorter, GError** err);
const char* clo_sort_get_kernel_name(CloSort* sorter, cl_uint i, GError** err);
size_t clo_sort_get_localmem_usage(CloSort* sorter, cl_uint i,
	size_t lws_max, size_t numel, GError** err);

/* ---- additions (no counterpart in CL_Ops) --------------------------------
 * Key-value sort: keys of the sorter's element type, 32-bit payloads.
 * The result is the stable sort of the pairs by key (payloads of equal keys
 * keep their input order).  keys and values are sorted both in place
 * (output NULL or equal to the input buffer) or both into buffers disjoint
 * from every input; a mixed call, or buffers that partly overlap, fail with
 * CLO_ERROR_ARGS before anything is enqueued.  Supported by "satradix"
 * (radix), "gselect" (rank) and,
 * through a stable rank, none other: bitonic sorts reject KV with
 * CLO_ERROR_ARGS because they are not stable. */
CCLEvent* clo_sort_kv_with_device_data(CloSort* sorter, CCLQueue* cq_exec,
	CCLQueue* cq_comm, CCLBuffer* keys_in, CCLBuffer* vals_in,
	CCLBuffer* keys_out, CCLBuffer* vals_out, size_t numel, size_t lws_max,
	GError** err);
cl_bool clo_sort_kv_with_host_data(CloSort* sorter, CCLQueue* cq_exec,
	CCLQueue* cq_comm, void* keys, void* vals, size_t numel, size_t lws_max,
	GError** err);

/* Batched segmented sort (satradix sorters, 32- or 64-bit keys): every
 * segment [offsets[j], offsets[j+1]) -- the last one ends at numel -- is
 * stable-sorted by key, all segments in one persistent launch.  `offsets` is a
 * device buffer of nseg uint32 segment starts, non-decreasing, offsets[0] = 0.
 * vals_in may be NULL (keys only).  keys_out == NULL sorts in place (the
 * key-value placement rules of clo_sort_kv_with_device_data apply).
 * Malformed offsets (decreasing, or beyond numel) are clamped into the
 * arrays: the order is then unspecified but no access leaves the buffers.
 * Sorter options tune it: segtile=8|16|32 (LDS tile = 256*segtile elements;
 * 8|16 for 64-bit keys; default 8, and 16 for 32-bit key-value sorts),
 * seggrid=<blocks> (0 = four per compute unit, two for 64-bit key-value
 * sorts), segclass=<up to 64 digits 0|1|2> (the least work class of each
 * segment length bin, see clo_dispatch.h; default all 0), segmsd=0|1 (1,
 * the default: keys alone are sorted by MSD bucketing, whose order of equal
 * keys is not the input order but whose output words are the same, since
 * equal keys are equal bit patterns; key-value sorts of 32-bit keys are
 * bucketed as 64-bit (key, input position) composites, which keeps the
 * stable order, with tile 8 unless segtile=16; 0: stable digit passes
 * only).  No option changes the result. */
CCLEvent* clo_sort_segmented_with_device_data(CloSort* sorter, CCLQueue* cq_exec,
	CCLQueue* cq_comm, CCLBuffer* keys_in, CCLBuffer* vals_in, CCLBuffer* keys_out,
	CCLBuffer* vals_out, CCLBuffer* offsets, size_t nseg, size_t numel, GError** err);

/* HIP source of the sorter's run-time compiled kernels (the macros of its
 * snippets followed by the kernels), NULL when the sorter runs the
 * precompiled ones.  Owned by the sorter. */
const char* clo_sort_get_source(CloSort* sorter);

/* Number of digit passes the last satradix call executed (passes whose digit
 * was the same for every key are skipped).  Blocks on the queue of that
 * call.  0 for the other algorithms. */
cl_uint clo_sort_get_last_passes(CloSort* sorter, GError** err);

#ifdef __cplusplus
}
#endif

#endif /* CLO_SORT_ABSTRACT_H_ */

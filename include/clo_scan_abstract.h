/*
 * clo_scan_abstract.h - scanner factory and vtable.
 *
 * Mirrors CL_Ops' src/cl_ops/scan/clo_scan_abstract.in.h:41-162.
 */
#ifndef CLO_SCAN_ABSTRACT_H_
#define CLO_SCAN_ABSTRACT_H_

#include "clo_common.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLO_SCAN_IMPLS "blelloch"

/* clo_scan_abstract.in.h:41-103 */
typedef struct clo_scan_impl_def {
	const char* name;
	const char* (*init)(CloScan* scanner, const char* options, GError** err);
	void (*finalize)(CloScan* scan);
	CCLEvent* (*scan_with_device_data)(CloScan* scanner, CCLQueue* cq_exec,
		CCLQueue* cq_comm, CCLBuffer* data_in, CCLBuffer* data_out,
		size_t numel, size_t lws_max, GError** err);
	cl_uint (*get_num_kernels)(CloScan* scanner, GError** err);
	const char* (*get_kernel_name)(CloScan* scanner, cl_uint i, GError** err);
	size_t (*get_localmem_usage)(CloScan* scanner, cl_uint i, size_t lws_max,
		size_t numel, GError** err);
} CloScanImplDef;

/* clo_scan_abstract.c:74-168.  Integer element and sum types only
 * (char..ulong); floating-point types fail with CLO_ERROR_ARGS. */
CloScan* clo_scan_new(const char* type, const char* options,
	CCLContext* ctx, CloType elem_type, CloType sum_type,
	const char* compiler_opts, GError** err);
void clo_scan_destroy(CloScan* scan);

/* Exclusive scan: data_out[i] = sum_{j<i} (SUM_TYPE) data_in[j], modulo
 * 2^(8*sizeof(SUM_TYPE)).  Asynchronous (returns the last event).  data_out
 * may be data_in itself when the element and sum types have the same width
 * (an in-place scan); any other overlap fails with CLO_ERROR_ARGS. */
CCLEvent* clo_scan_with_device_data(CloScan* scanner, CCLQueue* cq_exec,
	CCLQueue* cq_comm, CCLBuffer* data_in, CCLBuffer* data_out,
	size_t numel, size_t lws_max, GError** err);
/* Batched segmented exclusive scan (addition): the sum restarts at every
 * segment start offsets[j] (device buffer of nseg non-decreasing uint32
 * starts), giving the exclusive scan of each segment, all in one launch.
 * Blelloch scanners only. */
CCLEvent* clo_scan_segmented_with_device_data(CloScan* scanner, CCLQueue* cq_exec,
	CCLQueue* cq_comm, CCLBuffer* data_in, CCLBuffer* data_out, CCLBuffer* offsets, size_t nseg,
	size_t numel, GError** err);

cl_bool clo_scan_with_host_data(CloScan* scanner, CCLQueue* cq_exec,
	CCLQueue* cq_comm, void* data_in, void* data_out, size_t numel,
	size_t lws_max, GError** err);

CCLContext* clo_scan_get_context(CloScan* scanner);
CCLProgram* clo_scan_get_program(CloScan* scanner);
CloType clo_scan_get_elem_type(CloScan* scanner);
size_t clo_scan_get_element_size(CloScan* scanner);
CloType clo_scan_get_sum_type(CloScan* scanner);
size_t clo_scan_get_sum_size(CloScan* scanner);
void* clo_scan_get_data(CloScan* scanner);
void clo_scan_set_data(CloScan* scanner, void* data);
cl_uint clo_scan_get_num_kernels(CloScan* scanner, GError** err);
const char* clo_scan_get_kernel_name(CloScan* scanner, cl_uint i, GError** err);
size_t clo_scan_get_localmem_usage(CloScan* scanner, cl_uint i,
	size_t lws_max, size_t numel, GError** err);

/* No-allocation variant of clo_scan_with_device_data.  No counterpart in
 * CL_Ops: it is for callers that cannot allocate device memory while they
 * run (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same result as clo_scan_with_device_data (same kernels), but
 * its workspace is a 1 MiB static device array that each GPU holds from the
 * library's load until the process ends, so the call allocates nothing on
 * the device or the host: no hipMalloc/hipFree, no event (it returns
 * CL_TRUE/CL_FALSE; wait on the queue), and no profiling records.  numel is
 * limited to clo_scan_noalloc_max_numel(scanner) (355 M elements with 64-bit
 * sums, 532 M with 32-bit sums); larger calls fail with CLO_ERROR_ARGS.
 * Every call on one GPU shares that GPU's array: calls that overlap in time
 * on one GPU (two queues, two host threads) race and must be serialised by
 * the caller. */
cl_bool clo_scan_with_device_data_noalloc(CloScan* scanner, CCLQueue* cq_exec,
	CCLBuffer* data_in, CCLBuffer* data_out, size_t numel, GError** err);
/* Largest numel clo_scan_with_device_data_noalloc accepts for the scanner's
 * sum type. */
size_t clo_scan_noalloc_max_numel(CloScan* scanner);

#ifdef __cplusplus
}
#endif
#endif

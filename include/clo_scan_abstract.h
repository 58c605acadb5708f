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

#ifdef __cplusplus
}
#endif
#endif

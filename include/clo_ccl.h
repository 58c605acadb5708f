/*
 * clo_ccl.h - thin HIP layer that stands where cf4ocl2 stands for CL_Ops.
 *
 * CL_Ops' public API (src/cl_ops/sort/clo_sort_abstract.in.h:116-170,
 * src/cl_ops/scan/clo_scan_abstract.in.h:109-162,
 * src/cl_ops/rng/clo_rng.in.h:97-111) passes cf4ocl2 handles (CCLContext,
 * CCLQueue, CCLBuffer, CCLEvent, CCLProgram) and GLib's GError.  This header
 * declares opaque handles with the same names over HIP objects, so callers
 * keep their signatures:
 *
 *   CCLContext  -> a HIP device ordinal + reference count
 *   CCLQueue    -> a hipStream_t (+ optional per-kernel profiling records)
 *   CCLBuffer   -> a device pointer + size + reference count
 *   CCLEvent    -> a hipEvent_t owned by the queue that recorded it
 *   CCLProgram  -> a placeholder: kernels are compiled ahead of time
 *
 * Only plain C types appear here; no HIP, OpenCL or GLib header is needed.
 */
#ifndef CLO_CCL_H_
#define CLO_CCL_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- OpenCL scalar typedefs used by the CL_Ops signatures ---------------- */
#ifndef __OPENCL_CL_H
typedef int32_t cl_int;
typedef uint32_t cl_uint;
typedef int64_t cl_long;
typedef uint64_t cl_ulong;
typedef cl_uint cl_bool;
typedef cl_ulong cl_mem_flags;
typedef cl_ulong cl_command_queue_properties;
#define CL_FALSE 0
#define CL_TRUE 1
#define CL_MEM_READ_WRITE (1 << 0)
#define CL_MEM_WRITE_ONLY (1 << 1)
#define CL_MEM_READ_ONLY (1 << 2)
#define CL_QUEUE_PROFILING_ENABLE (1 << 1)
#endif

/* ---- GLib-compatible error record --------------------------------------- */
/* Same field layout as GLib's GError (domain quark, code, message), so code
 * written against GLib reads the fields unchanged.  Free with
 * clo_error_free(). */
#ifndef __G_LIB_H__
typedef uint32_t GQuark;
typedef char gchar;
typedef struct _GError {
	GQuark domain;
	int code;
	char* message;
} GError;
#endif

void clo_error_free(GError* err);
/* Frees *err (if any) and sets it to NULL, like g_clear_error(). */
void clo_error_clear(GError** err);

/* Error domain carrying a hipError_t as the code (errors from the HIP
 * runtime, e.g. out of memory). */
GQuark ccl_error_quark(void);
#define CCL_ERROR ccl_error_quark()

/* ---- opaque handles ------------------------------------------------------ */
typedef struct ccl_context CCLContext;
typedef struct ccl_queue CCLQueue;
typedef struct ccl_buffer CCLBuffer;
typedef struct ccl_event CCLEvent;
typedef struct ccl_program CCLProgram;
typedef struct ccl_device CCLDevice;
typedef struct ccl_event_wait_list* CCLEventWaitList;

/* Context on HIP device `dev` (ordinal).  ccl_context_new_any() takes the
 * current device. */
CCLContext* ccl_context_new_from_device_index(int dev, GError** err);
CCLContext* ccl_context_new_any(GError** err);
void ccl_context_ref(CCLContext* ctx);
void ccl_context_destroy(CCLContext* ctx);
int ccl_context_get_device_index(CCLContext* ctx);

/* Queue = new HIP stream (default priority) on the context's device.
 * `dev` is accepted for signature compatibility and must be NULL.
 * CL_QUEUE_PROFILING_ENABLE makes the library record a start/end hipEvent
 * pair around every kernel it launches on the queue (see ccl_prof_*). */
CCLQueue* ccl_queue_new(CCLContext* ctx, CCLDevice* dev,
	cl_command_queue_properties properties, GError** err);
/* Wraps an existing hipStream_t (passed as void*); the stream is not
 * destroyed with the queue.  NULL wraps the null stream. */
CCLQueue* ccl_queue_new_wrap(CCLContext* ctx, void* hip_stream,
	cl_command_queue_properties properties, GError** err);
void* ccl_queue_get_stream(CCLQueue* cq);
CCLContext* ccl_queue_get_context(CCLQueue* cq, GError** err);
cl_bool ccl_queue_finish(CCLQueue* cq, GError** err);
/* Releases the events the queue has recorded that have completed
 * (cf4ocl's ccl_queue_gc); does not wait.  Those events are invalid
 * afterwards.  For long-running callers that enqueue many calls on one
 * queue.  Does nothing while the queue is being captured. */
void ccl_queue_gc(CCLQueue* cq);
/* Makes the queue's later work wait for every event in the list (events of
 * any queue of the same context), without blocking the host; clears the
 * list.  Returns a marker event. */
CCLEvent* ccl_enqueue_barrier(CCLQueue* cq, CCLEventWaitList* ewl, GError** err);
void ccl_queue_destroy(CCLQueue* cq);
/* Records a marker event on the queue (for caller-side timing). */
CCLEvent* ccl_queue_enqueue_marker(CCLQueue* cq, GError** err);

/* Device buffer of `size` bytes.  If host_ptr is non-NULL its contents are
 * copied in (synchronously), like CL_MEM_COPY_HOST_PTR. */
CCLBuffer* ccl_buffer_new(CCLContext* ctx, cl_mem_flags flags, size_t size,
	void* host_ptr, GError** err);
/* Wraps caller-owned device memory; it is not freed with the buffer. */
CCLBuffer* ccl_buffer_new_wrap(CCLContext* ctx, void* dev_ptr, size_t size);
void ccl_buffer_ref(CCLBuffer* buf);
void ccl_buffer_destroy(CCLBuffer* buf);
void* ccl_buffer_get_device_ptr(CCLBuffer* buf);
size_t ccl_buffer_get_size(CCLBuffer* buf);
CCLEvent* ccl_buffer_enqueue_write(CCLBuffer* buf, CCLQueue* cq,
	cl_bool blocking, size_t offset, size_t size, const void* ptr,
	CCLEventWaitList* ewl, GError** err);
CCLEvent* ccl_buffer_enqueue_read(CCLBuffer* buf, CCLQueue* cq,
	cl_bool blocking, size_t offset, size_t size, void* ptr,
	CCLEventWaitList* ewl, GError** err);
CCLEvent* ccl_buffer_enqueue_copy(CCLBuffer* src, CCLBuffer* dst,
	CCLQueue* cq, size_t src_offset, size_t dst_offset, size_t size,
	CCLEventWaitList* ewl, GError** err);

/* Events.  Events returned by the library belong to the queue that
 * recorded them and live until the queue is destroyed. */
void ccl_event_set_name(CCLEvent* evt, const char* name);
const char* ccl_event_get_name(CCLEvent* evt);
/* Adds NULL-terminated events to a wait list (created on first use). */
void ccl_event_wait_list_add(CCLEventWaitList* ewl, CCLEvent* evt, ...);
void ccl_event_wait_list_clear(CCLEventWaitList* ewl);
/* Blocks the host until every event in the list completed; clears it. */
cl_bool ccl_event_wait(CCLEventWaitList* ewl, GError** err);
/* Milliseconds between two completed events (hipEventElapsedTime). */
double ccl_event_elapsed_ms(CCLEvent* start, CCLEvent* end, GError** err);

/* Per-kernel timing on a queue created with CL_QUEUE_PROFILING_ENABLE
 * (the counterpart of cf4ocl's CCLProf over cq_exec). */
void ccl_prof_reset(CCLQueue* cq);
/* Sum (ms) and count of the kernels whose name starts with `prefix`
 * (NULL or "" = all kernels) recorded since the last reset.  Waits for the
 * queue.  Returns the number of matching launches. */
size_t ccl_prof_get(CCLQueue* cq, const char* prefix, double* total_ms,
	GError** err);
/* i-th distinct kernel name recorded since the last reset, or NULL. */
const char* ccl_prof_get_name(CCLQueue* cq, size_t i);
/* Records events only for kernels whose name starts with `prefix` (NULL or
 * "" = every kernel, the default), so that timing one kernel adds no event
 * pairs around the others.  No counterpart in cf4ocl, whose profiling records
 * every event of the queue. */
void ccl_prof_set_filter(CCLQueue* cq, const char* prefix);

/* Test hook: the next `count` device allocations made by the library fail
 * with hipErrorOutOfMemory (0 disables). */
void ccl_debug_fail_next_allocs(int count);

#ifdef __cplusplus
}
#endif

#endif /* CLO_CCL_H_ */

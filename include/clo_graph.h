/*
 * clo_graph.h - capture of enqueued work into a replayable HIP graph.
 *
 * No counterpart in CL_Ops (an addition for batched and repeated calls): a
 * sequence of clo_sort_* / clo_scan_* / clo_rng_fill / clo_util_* calls on
 * one queue is recorded once (HIP stream capture) and relaunched as one graph,
 * which removes the per-kernel launch cost of the pass sequences.
 *
 * Rules:
 *  - Run each call once on the same sorter/scanner with the same or a larger
 *    size before capturing it: device scratch is sized on the eager call, and
 *    a call that would need to grow it while capturing fails with
 *    CLO_ERROR_ARGS (nothing is allocated during capture).
 *  - Calls made while capturing return a placeholder event (no hipEvent
 *    behind it; waiting on it is a no-op); blocking calls (the *_host_data
 *    functions, clo_sort_get_last_passes, clo_util_multiset_hash,
 *    clo_util_check_sorted, ccl_queue_finish) must not be made while
 *    capturing.  Profiling records are not taken for captured kernels.
 *  - The graph reads and writes the buffers named at capture time; replays
 *    recompute everything from those buffers (data-dependent choices such as
 *    skipped digit passes are made on the device at every replay).  Caller
 *    buffers must outlive the graph.
 *  - The graph holds the device scratch of every sorter/scanner/dispatcher
 *    used inside the capture: destroying such an object leaves its scratch
 *    to the graph (freed by clo_graph_destroy), and a later eager call on
 *    the object that would need larger scratch fails with CLO_ERROR_ARGS
 *    until the graph is destroyed (calls that fit the scratch run normally
 *    and may overwrite it; a replay recomputes it).
 */
#ifndef CLO_GRAPH_H_
#define CLO_GRAPH_H_

#include "clo_ccl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clo_graph CloGraph;

/* Starts capturing the work enqueued on cq (by the calling thread). */
cl_bool clo_graph_begin_capture(CCLQueue* cq, GError** err);
/* Ends the capture started on cq and instantiates it.  NULL on error (the
 * capture is closed either way). */
CloGraph* clo_graph_end_capture(CCLQueue* cq, GError** err);
/* Launches the graph on cq (asynchronous); returns an event recorded after it. */
CCLEvent* clo_graph_launch(CloGraph* graph, CCLQueue* cq, GError** err);
/* Number of nodes (kernels, copies) in the graph. */
size_t clo_graph_get_num_nodes(CloGraph* graph);
void clo_graph_destroy(CloGraph* graph);

#ifdef __cplusplus
}
#endif
#endif

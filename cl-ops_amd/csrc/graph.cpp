// Graph capture of enqueued library calls (HIP stream capture + instantiate).
// No counterpart in CL_Ops.

#include "clo_internal.hpp"

using namespace clo;

struct clo_graph {
	hipGraph_t graph;
	hipGraphExec_t exec;
	CCLContext* ctx;
	size_t nodes;
};

extern "C" {

cl_bool clo_graph_begin_capture(CCLQueue* cq, GError** err) {
	if (err && *err) return CL_FALSE;
	if (!cq) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL queue");
		return CL_FALSE;
	}
	if (cq->capturing) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "queue is already being captured");
		return CL_FALSE;
	}
	if (!use_device(cq->ctx, err)) return CL_FALSE;
	// Thread-local mode: allocation and synchronisation calls on this thread
	// fail while the capture is open instead of silently escaping the graph.
	if (hip_failed(hipStreamBeginCapture(cq->stream, hipStreamCaptureModeThreadLocal), err,
			"hipStreamBeginCapture"))
		return CL_FALSE;
	cq->capturing = true;
	++g_capture_depth;
	return CL_TRUE;
}

CloGraph* clo_graph_end_capture(CCLQueue* cq, GError** err) {
	if (!cq || !cq->capturing) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "queue is not being captured");
		return nullptr;
	}
	(void)use_device(cq->ctx, nullptr);
	hipGraph_t g = nullptr;
	hipError_t e = hipStreamEndCapture(cq->stream, &g);
	cq->capturing = false;
	--g_capture_depth;
	if (err && *err) {  // a call inside the capture failed: drop the graph
		if (g) (void)hipGraphDestroy(g);
		return nullptr;
	}
	if (hip_failed(e, err, "hipStreamEndCapture")) {
		if (g) (void)hipGraphDestroy(g);
		return nullptr;
	}
	size_t nodes = 0;
	if (hip_failed(hipGraphGetNodes(g, nullptr, &nodes), err, "hipGraphGetNodes")) {
		(void)hipGraphDestroy(g);
		return nullptr;
	}
	hipGraphExec_t x = nullptr;  // stays null for an empty capture (launch = no-op)
	if (nodes > 0 &&
		hip_failed(hipGraphInstantiate(&x, g, nullptr, nullptr, 0), err, "hipGraphInstantiate")) {
		(void)hipGraphDestroy(g);
		return nullptr;
	}
	CloGraph* cg = new clo_graph();
	cg->graph = g;
	cg->exec = x;
	cg->ctx = cq->ctx;
	ccl_context_ref(cq->ctx);
	cg->nodes = nodes;
	return cg;
}

CCLEvent* clo_graph_launch(CloGraph* graph, CCLQueue* cq, GError** err) {
	if (err && *err) return nullptr;
	if (!graph || !cq) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL graph or queue");
		return nullptr;
	}
	if (cq->capturing) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "cannot launch a graph into a capture");
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	if (graph->exec &&
		hip_failed(hipGraphLaunch(graph->exec, cq->stream), err, "hipGraphLaunch"))
		return nullptr;
	return record_event(cq, "graph", err);
}

size_t clo_graph_get_num_nodes(CloGraph* graph) { return graph ? graph->nodes : 0; }

void clo_graph_destroy(CloGraph* graph) {
	if (!graph) return;
	(void)hipSetDevice(graph->ctx->dev);
	if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
	(void)hipGraphDestroy(graph->graph);
	ccl_context_destroy(graph->ctx);
	delete graph;
}

}  // extern "C"

// Graph capture of enqueued library calls (HIP stream capture + instantiate).
// No counterpart in CL_Ops.

#include "clo_internal.hpp"

using namespace clo;

struct clo_graph {
	hipGraph_t graph;
	hipGraphExec_t exec;
	CCLContext* ctx;
	size_t nodes;
	std::vector<DevBlock*> held;  // library scratch the captured kernels use
};

// The distinct blocks logged since the capture on cq began, each with a new
// reference for the graph.  When no capture remains open on this thread the
// log drops its own references and is cleared.  With overlapping captures a
// graph may hold a block that only another capture used: never too few.
static std::vector<DevBlock*> take_capture_blocks(CCLQueue* cq) {
	std::vector<DevBlock*> out;
	for (size_t i = cq->capture_mark; i < g_capture_blocks.size(); ++i) {
		DevBlock* b = g_capture_blocks[i];
		bool seen = false;
		for (DevBlock* o : out) seen = seen || o == b;
		if (!seen) {
			b->refs.fetch_add(1);
			out.push_back(b);
		}
	}
	if (g_capture_depth == 0) {
		for (DevBlock* b : g_capture_blocks) devblock_unref(b);
		g_capture_blocks.clear();
	}
	return out;
}

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
	// RNG seeds whose initialisation waits for their first fill are
	// initialised now, outside the graph: a replay must not re-seed them.
	if (!rng_init_pending_on(cq, err)) return CL_FALSE;
	// Thread-local mode: allocation and synchronisation calls on this thread
	// fail while the capture is open instead of silently escaping the graph.
	if (hip_failed(hipStreamBeginCapture(cq->stream, hipStreamCaptureModeThreadLocal), err,
			"hipStreamBeginCapture"))
		return CL_FALSE;
	cq->capturing = true;
	cq->capture_mark = g_capture_blocks.size();
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
	std::vector<DevBlock*> held = take_capture_blocks(cq);
	auto drop = [&]() {
		if (g) (void)hipGraphDestroy(g);
		for (DevBlock* b : held) devblock_unref(b);
	};
	if (err && *err) {  // a call inside the capture failed: drop the graph
		drop();
		return nullptr;
	}
	if (hip_failed(e, err, "hipStreamEndCapture")) {
		drop();
		return nullptr;
	}
	size_t nodes = 0;
	if (hip_failed(hipGraphGetNodes(g, nullptr, &nodes), err, "hipGraphGetNodes")) {
		drop();
		return nullptr;
	}
	hipGraphExec_t x = nullptr;  // stays null for an empty capture (launch = no-op)
	if (nodes > 0 &&
		hip_failed(hipGraphInstantiate(&x, g, nullptr, nullptr, 0), err, "hipGraphInstantiate")) {
		drop();
		return nullptr;
	}
	CloGraph* cg = new clo_graph();
	cg->graph = g;
	cg->exec = x;
	cg->ctx = cq->ctx;
	ccl_context_ref(cq->ctx);
	cg->nodes = nodes;
	cg->held = std::move(held);
	return cg;
}

CCLEvent* clo_graph_launch(CloGraph* graph, CCLQueue* cq, GError** err) {
	ApiRange api_range("clo_graph_launch");
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
	for (DevBlock* b : graph->held) devblock_unref(b);
	ccl_context_destroy(graph->ctx);
	delete graph;
}

}  // extern "C"

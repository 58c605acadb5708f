// Thin HIP layer behind the CCL* handles (the role cf4ocl2 plays for CL_Ops)
// and the common module (CL_Ops' src/cl_ops/common/clo_common.c).

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "clo_internal.hpp"

namespace clo {

static std::atomic<int> g_fail_allocs{0};

void set_error(GError** err, GQuark domain, int code, const char* fmt, ...) {
	if (err == nullptr || *err != nullptr) return;
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	GError* e = static_cast<GError*>(malloc(sizeof(GError)));
	if (!e) return;
	e->domain = domain;
	e->code = code;
	e->message = strdup(buf);
	*err = e;
}

bool hip_failed(hipError_t e, GError** err, const char* what) {
	if (e == hipSuccess) return false;
	set_error(err, kQuarkCcl, static_cast<int>(e), "%s: %s (%d)", what,
		hipGetErrorString(e), static_cast<int>(e));
	return true;
}

hipError_t dev_malloc(void** p, size_t bytes) {
	int left = g_fail_allocs.load();
	while (left > 0) {
		if (g_fail_allocs.compare_exchange_weak(left, left - 1)) {
			*p = nullptr;
			return hipErrorOutOfMemory;
		}
	}
	if (bytes == 0) bytes = 1;
	return hipMalloc(p, bytes);
}

bool use_device(CCLContext* ctx, GError** err) {
	int cur = -1;
	if (hipGetDevice(&cur) == hipSuccess && cur == ctx->dev) return true;
	return !hip_failed(hipSetDevice(ctx->dev), err, "hipSetDevice");
}

thread_local int g_capture_depth = 0;

CCLEvent* record_event(CCLQueue* q, const char* name, GError** err) {
	if (q->capturing) {
		if (!q->capture_event) {
			q->capture_event = new ccl_event();
			q->capture_event->ev = nullptr;
			q->capture_event->name = "captured";
		}
		return q->capture_event;
	}
	CCLEvent* e = new ccl_event();
	if (hip_failed(hipEventCreate(&e->ev), err, "hipEventCreate")) {
		delete e;
		return nullptr;
	}
	if (hip_failed(hipEventRecord(e->ev, q->stream), err, "hipEventRecord")) {
		(void)hipEventDestroy(e->ev);
		delete e;
		return nullptr;
	}
	if (name) e->name = name;
	q->events.push_back(e);
	return e;
}

thread_local bool g_launch_quiet = false;

namespace {
std::atomic<int> g_log_override{-1};
int env_log_level() {
	static const int lvl = [] {
		const char* e = getenv("CLO_LOG");
		return e ? atoi(e) : 0;
	}();
	return lvl;
}
}  // namespace

int log_level() {
	const int o = g_log_override.load(std::memory_order_relaxed);
	return o >= 0 ? o : env_log_level();
}

void log_msg(const char* fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	fprintf(stderr, "clo: %s\n", buf);
}

ApiRange::ApiRange(const char* name) { roctxRangePushA(name); }
ApiRange::~ApiRange() { roctxRangePop(); }

Launch::Launch(CCLQueue* q_, const char* name_) : q(q_), name(name_), rec(nullptr) {
	CLO_LOG(2, "launch %s", name_);
	(void)hipGetLastError();  // clear stale launch errors
	if (!q->prof || q->capturing || g_launch_quiet) return;
	if (!q->prof_filter.empty() &&
		strncmp(name, q->prof_filter.c_str(), q->prof_filter.size()) != 0)
		return;
	ccl_prof_rec r;
	r.name = name;
	bool ok = true;
	for (int k = 0; k < 2; ++k) {
		hipEvent_t ev;
		if (!q->prof_pool.empty()) {
			ev = q->prof_pool.back();
			q->prof_pool.pop_back();
		} else if (hipEventCreate(&ev) != hipSuccess) {
			ok = false;
			break;
		}
		(k == 0 ? r.start : r.end) = ev;
	}
	if (!ok) return;
	if (hipEventRecord(r.start, q->stream) != hipSuccess) {
		// No record for this launch: the events go back to the pool and the
		// launch's error check still runs in done().
		q->prof_pool.push_back(r.start);
		q->prof_pool.push_back(r.end);
		return;
	}
	q->prof_recs.push_back(r);
	rec = &q->prof_recs.back();
}

bool Launch::done(GError** err) {
	hipError_t e = hipGetLastError();
	if (q->prof && rec) {
		const hipError_t re = hipEventRecord(q->prof_recs.back().end, q->stream);
		if (e == hipSuccess) e = re;
	}
	if (e != hipSuccess) {
		set_error(err, kQuarkCcl, static_cast<int>(e), "launch of %s: %s", name,
			hipGetErrorString(e));
		return false;
	}
	return true;
}

thread_local std::vector<DevBlock*> g_capture_blocks;

void devblock_unref(DevBlock* b) {
	if (b && b->refs.fetch_sub(1) == 1) {
		(void)hipFree(b->ptr);
		delete b;
	}
}

bool Scratch::ensure(size_t bytes, GError** err) {
	if (bytes <= size && ptr) {
		// Used inside a capture: logged (the log holds a reference until the
		// last open capture ends), and the graph will hold the allocation.
		if (g_capture_depth > 0 && (g_capture_blocks.empty() || g_capture_blocks.back() != blk)) {
			blk->refs.fetch_add(1);
			g_capture_blocks.push_back(blk);
		}
		return true;
	}
	if (g_capture_depth > 0) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"device scratch of %zu bytes needed during graph capture: run the same call once "
			"before capturing it", bytes);
		return false;
	}
	if (blk && blk->refs.load() > 1) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"device scratch would grow to %zu bytes while a captured graph uses it: destroy "
			"the graph first, or size the scratch with the largest call before capturing",
			bytes);
		return false;
	}
	release();
	size_t want = bytes + bytes / 8 + 256;  // headroom against regrowth
	void* p = nullptr;
	if (hip_failed(dev_malloc(&p, want), err, "device scratch allocation"))
		return false;
	// hipMemset runs on the null stream, which the library's non-blocking
	// queues do not wait for, and may return before the fill is done: wait
	// for it here, or a kernel on a queue could read the memory unzeroed.
	if (zero_on_alloc &&
		(hip_failed(hipMemset(p, 0, want), err, "hipMemset") ||
			hip_failed(hipStreamSynchronize(nullptr), err, "hipStreamSynchronize"))) {
		(void)hipFree(p);
		return false;
	}
	blk = new DevBlock();
	blk->ptr = p;
	blk->refs.store(1);
	ptr = p;
	size = want;
	return true;
}

void Scratch::release() {
	devblock_unref(blk);
	blk = nullptr;
	ptr = nullptr;
	size = 0;
}

bool parse_options(const char* options, std::vector<Opt>& out,
	const char* algo, GError** err) {
	out.clear();
	if (!options) return true;
	std::string s(options);
	size_t pos = 0;
	while (pos <= s.size()) {
		size_t comma = s.find(',', pos);
		if (comma == std::string::npos) comma = s.size();
		std::string tok = s.substr(pos, comma - pos);
		pos = comma + 1;
		if (tok.empty()) continue;
		size_t eq = tok.find('=');
		if (eq == std::string::npos) {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS,
				"Invalid option '%s' for %s.", tok.c_str(), algo);
			return false;
		}
		out.push_back({tok.substr(0, eq), tok.substr(eq + 1)});
	}
	return true;
}

static std::string strip_ws(const char* s) {
	std::string r;
	for (; s && *s; ++s)
		if (*s != ' ' && *s != '\t' && *s != '\n') r += *s;
	return r;
}

int compare_kind(const char* compare) {
	if (compare == nullptr) return 0;
	std::string c = strip_ws(compare);
	if (c == "((a)>(b))" || c == "(a)>(b)" || c == "a>b" || c == "(a>b)") return 0;
	if (c == "((a)<(b))" || c == "(a)<(b)" || c == "a<b" || c == "(a<b)") return 1;
	return -1;
}

bool get_key_is_identity(const char* get_key) {
	if (get_key == nullptr) return true;
	std::string g = strip_ws(get_key);
	return g.empty() || g == "(x)" || g == "x";
}

bool type_is_integer(CloType t) { return t >= CLO_CHAR && t <= CLO_ULONG; }
bool type_is_signed(CloType t) {
	return t == CLO_CHAR || t == CLO_SHORT || t == CLO_INT || t == CLO_LONG;
}

}  // namespace clo

using namespace clo;

extern "C" {

/* ---------------------------- errors ------------------------------------ */

void clo_error_free(GError* err) {
	if (!err) return;
	free(err->message);
	free(err);
}

void clo_error_clear(GError** err) {
	if (err && *err) {
		clo_error_free(*err);
		*err = nullptr;
	}
}

GQuark ccl_error_quark(void) { return kQuarkCcl; }
GQuark clo_error_quark(void) { return kQuarkClo; }

void ccl_debug_fail_next_allocs(int count) { g_fail_allocs.store(count < 0 ? 0 : count); }

/* ---------------------------- context ----------------------------------- */

CCLContext* ccl_context_new_from_device_index(int dev, GError** err) {
	int count = 0;
	if (hip_failed(hipGetDeviceCount(&count), err, "hipGetDeviceCount")) return nullptr;
	if (dev < 0 || dev >= count) {
		set_error(err, kQuarkCcl, static_cast<int>(hipErrorInvalidDevice),
			"device %d not present (%d devices)", dev, count);
		return nullptr;
	}
	if (hip_failed(hipSetDevice(dev), err, "hipSetDevice")) return nullptr;
	CCLContext* c = new ccl_context();
	c->dev = dev;
	c->refc = 1;
	return c;
}

CCLContext* ccl_context_new_any(GError** err) {
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) dev = 0;
	return ccl_context_new_from_device_index(dev, err);
}

void ccl_context_ref(CCLContext* ctx) {
	if (ctx) ++ctx->refc;
}

void ccl_context_destroy(CCLContext* ctx) {
	if (ctx && --ctx->refc == 0) delete ctx;
}

int ccl_context_get_device_index(CCLContext* ctx) { return ctx ? ctx->dev : -1; }

/* ---------------------------- queues ------------------------------------ */

static CCLQueue* queue_make(CCLContext* ctx, hipStream_t s, bool owns,
	cl_command_queue_properties props) {
	CCLQueue* q = new ccl_queue();
	q->ctx = ctx;
	ccl_context_ref(ctx);
	q->stream = s;
	q->owns_stream = owns;
	q->prof = (props & CL_QUEUE_PROFILING_ENABLE) != 0;
	return q;
}

CCLQueue* ccl_queue_new(CCLContext* ctx, CCLDevice* dev,
	cl_command_queue_properties properties, GError** err) {
	(void)dev;
	if (!ctx) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "ccl_queue_new: NULL context");
		return nullptr;
	}
	if (!use_device(ctx, err)) return nullptr;
	hipStream_t s;
	if (hip_failed(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), err,
			"hipStreamCreate"))
		return nullptr;
	return queue_make(ctx, s, true, properties);
}

CCLQueue* ccl_queue_new_wrap(CCLContext* ctx, void* hip_stream,
	cl_command_queue_properties properties, GError** err) {
	if (!ctx) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "ccl_queue_new_wrap: NULL context");
		return nullptr;
	}
	return queue_make(ctx, static_cast<hipStream_t>(hip_stream), false, properties);
}

void* ccl_queue_get_stream(CCLQueue* cq) { return cq ? cq->stream : nullptr; }

CCLContext* ccl_queue_get_context(CCLQueue* cq, GError** err) {
	if (!cq) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL queue");
		return nullptr;
	}
	return cq->ctx;
}

cl_bool ccl_queue_finish(CCLQueue* cq, GError** err) {
	if (!cq) return CL_FALSE;
	if (!use_device(cq->ctx, err)) return CL_FALSE;
	return hip_failed(hipStreamSynchronize(cq->stream), err, "hipStreamSynchronize")
		? CL_FALSE : CL_TRUE;
}

void ccl_queue_gc(CCLQueue* cq) {
	if (!cq || cq->capturing) return;
	(void)hipSetDevice(cq->ctx->dev);
	size_t keep = 0;
	for (CCLEvent* e : cq->events) {
		if (e->ev && hipEventQuery(e->ev) != hipSuccess) {
			cq->events[keep++] = e;  // still pending (or failed): keep it
			continue;
		}
		if (e->ev) (void)hipEventDestroy(e->ev);
		delete e;
	}
	cq->events.resize(keep);
}

void ccl_queue_destroy(CCLQueue* cq) {
	if (!cq) return;
	(void)hipSetDevice(cq->ctx->dev);
	(void)hipStreamSynchronize(cq->stream);
	for (CCLEvent* e : cq->events) {
		(void)hipEventDestroy(e->ev);
		delete e;
	}
	for (auto& r : cq->prof_recs) {
		(void)hipEventDestroy(r.start);
		(void)hipEventDestroy(r.end);
	}
	for (hipEvent_t e : cq->prof_pool) (void)hipEventDestroy(e);
	delete cq->capture_event;
	if (cq->owns_stream) (void)hipStreamDestroy(cq->stream);
	ccl_context_destroy(cq->ctx);
	delete cq;
}

static bool wait_list_on_stream(CCLEventWaitList* ewl, hipStream_t s, GError** err);

CCLEvent* ccl_enqueue_barrier(CCLQueue* cq, CCLEventWaitList* ewl, GError** err) {
	if (!cq) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL queue");
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	if (!wait_list_on_stream(ewl, cq->stream, err)) return nullptr;
	return record_event(cq, "barrier", err);
}

CCLEvent* ccl_queue_enqueue_marker(CCLQueue* cq, GError** err) {
	if (!cq) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL queue");
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	return record_event(cq, "marker", err);
}

/* ---------------------------- buffers ----------------------------------- */

CCLBuffer* ccl_buffer_new(CCLContext* ctx, cl_mem_flags flags, size_t size,
	void* host_ptr, GError** err) {
	(void)flags;
	if (!ctx) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "ccl_buffer_new: NULL context");
		return nullptr;
	}
	if (!use_device(ctx, err)) return nullptr;
	void* p = nullptr;
	if (hip_failed(dev_malloc(&p, size), err, "hipMalloc")) return nullptr;
	// From pageable memory hipMemcpy may return before its DMA lands, and the
	// queues (non-blocking streams) do not wait for the null stream: wait.
	if (host_ptr && size &&
		(hip_failed(hipMemcpy(p, host_ptr, size, hipMemcpyHostToDevice), err,
			"hipMemcpy H2D") ||
			hip_failed(hipStreamSynchronize(nullptr), err, "hipStreamSynchronize"))) {
		(void)hipFree(p);
		return nullptr;
	}
	CCLBuffer* b = new ccl_buffer();
	b->ctx = ctx;
	ccl_context_ref(ctx);
	b->ptr = p;
	b->size = size;
	b->owns = true;
	b->refc = 1;
	return b;
}

CCLBuffer* ccl_buffer_new_wrap(CCLContext* ctx, void* dev_ptr, size_t size) {
	if (!ctx) return nullptr;
	CCLBuffer* b = new ccl_buffer();
	b->ctx = ctx;
	ccl_context_ref(ctx);
	b->ptr = dev_ptr;
	b->size = size;
	b->owns = false;
	b->refc = 1;
	return b;
}

void ccl_buffer_ref(CCLBuffer* buf) {
	if (buf) ++buf->refc;
}

void ccl_buffer_destroy(CCLBuffer* buf) {
	if (!buf || --buf->refc > 0) return;
	if (buf->owns && buf->ptr) {
		(void)hipSetDevice(buf->ctx->dev);
		(void)hipFree(buf->ptr);
	}
	ccl_context_destroy(buf->ctx);
	delete buf;
}

void* ccl_buffer_get_device_ptr(CCLBuffer* buf) { return buf ? buf->ptr : nullptr; }
size_t ccl_buffer_get_size(CCLBuffer* buf) { return buf ? buf->size : 0; }

static bool wait_list_on_stream(CCLEventWaitList* ewl, hipStream_t s, GError** err) {
	if (!ewl || !*ewl) return true;
	for (CCLEvent* e : (*ewl)->evts)
		if (e && e->ev && hip_failed(hipStreamWaitEvent(s, e->ev, 0), err, "hipStreamWaitEvent"))
			return false;
	ccl_event_wait_list_clear(ewl);
	return true;
}

static CCLEvent* copy_common(CCLQueue* cq, void* dst, const void* src, size_t size,
	hipMemcpyKind kind, cl_bool blocking, CCLEventWaitList* ewl, const char* name,
	GError** err) {
	if (!cq) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "%s: NULL queue", name);
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	if (!wait_list_on_stream(ewl, cq->stream, err)) return nullptr;
	if (size && hip_failed(hipMemcpyAsync(dst, src, size, kind, cq->stream), err, name))
		return nullptr;
	CCLEvent* e = record_event(cq, name, err);
	if (e && e->ev && blocking && hip_failed(hipEventSynchronize(e->ev), err, name)) return nullptr;
	return e;
}

static bool range_ok(CCLBuffer* buf, size_t offset, size_t size, GError** err) {
	if (!buf || offset > buf->size || size > buf->size - offset) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"buffer range [%zu, +%zu) outside a buffer of %zu bytes", offset, size,
			buf ? buf->size : 0);
		return false;
	}
	return true;
}

CCLEvent* ccl_buffer_enqueue_write(CCLBuffer* buf, CCLQueue* cq, cl_bool blocking,
	size_t offset, size_t size, const void* ptr, CCLEventWaitList* ewl, GError** err) {
	if (!range_ok(buf, offset, size, err)) return nullptr;
	return copy_common(cq, static_cast<char*>(buf->ptr) + offset, ptr, size,
		hipMemcpyHostToDevice, blocking, ewl, "write_buffer", err);
}

CCLEvent* ccl_buffer_enqueue_read(CCLBuffer* buf, CCLQueue* cq, cl_bool blocking,
	size_t offset, size_t size, void* ptr, CCLEventWaitList* ewl, GError** err) {
	if (!range_ok(buf, offset, size, err)) return nullptr;
	return copy_common(cq, ptr, static_cast<char*>(buf->ptr) + offset, size,
		hipMemcpyDeviceToHost, blocking, ewl, "read_buffer", err);
}

CCLEvent* ccl_buffer_enqueue_copy(CCLBuffer* src, CCLBuffer* dst, CCLQueue* cq,
	size_t src_offset, size_t dst_offset, size_t size, CCLEventWaitList* ewl,
	GError** err) {
	if (!range_ok(src, src_offset, size, err) || !range_ok(dst, dst_offset, size, err))
		return nullptr;
	return copy_common(cq, static_cast<char*>(dst->ptr) + dst_offset,
		static_cast<char*>(src->ptr) + src_offset, size, hipMemcpyDeviceToDevice,
		CL_FALSE, ewl, "copy_buffer", err);
}

/* ---------------------------- events ------------------------------------ */

void ccl_event_set_name(CCLEvent* evt, const char* name) {
	if (evt) evt->name = name ? name : "";
}

const char* ccl_event_get_name(CCLEvent* evt) { return evt ? evt->name.c_str() : nullptr; }

void ccl_event_wait_list_add(CCLEventWaitList* ewl, CCLEvent* evt, ...) {
	if (!ewl) return;
	if (!*ewl) *ewl = new ccl_event_wait_list();
	va_list ap;
	va_start(ap, evt);
	for (CCLEvent* e = evt; e != nullptr; e = va_arg(ap, CCLEvent*)) (*ewl)->evts.push_back(e);
	va_end(ap);
}

void ccl_event_wait_list_clear(CCLEventWaitList* ewl) {
	if (ewl && *ewl) {
		delete *ewl;
		*ewl = nullptr;
	}
}

cl_bool ccl_event_wait(CCLEventWaitList* ewl, GError** err) {
	if (!ewl || !*ewl) return CL_TRUE;
	bool ok = true;
	for (CCLEvent* e : (*ewl)->evts)
		if (e && e->ev && hip_failed(hipEventSynchronize(e->ev), err, "hipEventSynchronize")) {
			ok = false;
			break;
		}
	ccl_event_wait_list_clear(ewl);
	return ok ? CL_TRUE : CL_FALSE;
}

double ccl_event_elapsed_ms(CCLEvent* start, CCLEvent* end, GError** err) {
	if (!start || !end || !start->ev || !end->ev) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL event (or one recorded during capture)");
		return -1.0;
	}
	if (hip_failed(hipEventSynchronize(end->ev), err, "hipEventSynchronize")) return -1.0;
	float ms = 0.f;
	if (hip_failed(hipEventElapsedTime(&ms, start->ev, end->ev), err, "hipEventElapsedTime"))
		return -1.0;
	return ms;
}

/* ---------------------------- profiling --------------------------------- */

void ccl_prof_set_filter(CCLQueue* cq, const char* prefix) {
	if (cq) cq->prof_filter = prefix ? prefix : "";
}

void ccl_prof_reset(CCLQueue* cq) {
	if (!cq) return;
	// A void call: a failed synchronisation resurfaces in the queue's next
	// checked call.
	(void)hipStreamSynchronize(cq->stream);
	for (auto& r : cq->prof_recs) {
		cq->prof_pool.push_back(r.start);
		cq->prof_pool.push_back(r.end);
	}
	cq->prof_recs.clear();
	cq->prof_names.clear();
}

size_t ccl_prof_get(CCLQueue* cq, const char* prefix, double* total_ms, GError** err) {
	if (total_ms) *total_ms = 0.0;
	if (!cq) return 0;
	if (hip_failed(hipStreamSynchronize(cq->stream), err, "hipStreamSynchronize")) return 0;
	size_t plen = prefix ? strlen(prefix) : 0;
	size_t n = 0;
	double tot = 0.0;
	for (auto& r : cq->prof_recs) {
		if (plen && r.name.compare(0, plen, prefix) != 0) continue;
		float ms = 0.f;
		if (hipEventElapsedTime(&ms, r.start, r.end) == hipSuccess) tot += ms;
		++n;
	}
	if (total_ms) *total_ms = tot;
	return n;
}

const char* ccl_prof_get_name(CCLQueue* cq, size_t i) {
	if (!cq) return nullptr;
	cq->prof_names.clear();
	for (auto& r : cq->prof_recs) {
		bool seen = false;
		for (auto& s : cq->prof_names)
			if (s == r.name) seen = true;
		if (!seen) cq->prof_names.push_back(r.name);
	}
	return i < cq->prof_names.size() ? cq->prof_names[i].c_str() : nullptr;
}

/* ---------------------------- common module ----------------------------- */

struct clo_type_info {
	const char* name;
	size_t size;
};

// clo_common.c:54-68: OpenCL type names and sizes, indexed by CloType.
static const CloTypeInfo kTypes[] = {
	{"char", 1}, {"uchar", 1}, {"short", 2}, {"ushort", 2}, {"int", 4}, {"uint", 4},
	{"long", 8}, {"ulong", 8}, {"half", 2}, {"float", 4}, {"double", 8}, {nullptr, 0}};

const char* clo_type_get_name(CloType type) {
	if (static_cast<int>(type) < 0 || type > CLO_DOUBLE) return nullptr;
	return kTypes[type].name;
}

size_t clo_type_sizeof(CloType type) {
	if (static_cast<int>(type) < 0 || type > CLO_DOUBLE) return 0;
	return kTypes[type].size;
}

CloType clo_type_by_name(const char* name, GError** err) {
	for (int i = 0; kTypes[i].name; ++i)
		if (name && strcmp(name, kTypes[i].name) == 0) return static_cast<CloType>(i);
	set_error(err, kQuarkClo, CLO_ERROR_UNKNOWN_TYPE, "Unknown type '%s'",
		name ? name : "(null)");
	return static_cast<CloType>(-1);
}

// clo_common.c:141-198.  Same results as the reference's bit tricks,
// including nlpo2(0) == 0 and nlpo2(x > 2^31) == 0 (32-bit wrap).
void clo_set_log_level(int level) { g_log_override.store(level < 0 ? -1 : level); }
int clo_get_log_level(void) { return log_level(); }

unsigned int clo_nlpo2(unsigned int x) {
	if ((x & (x - 1)) == 0) return x;
	if (x > 0x80000000u) return 0u;
	return 1u << (32 - __builtin_clz(x));  // x > 2, not a power of 2
}

unsigned int clo_ones32(unsigned int x) { return static_cast<unsigned int>(__builtin_popcount(x)); }

unsigned int clo_tzc(int x) {
	unsigned int u = static_cast<unsigned int>(x);
	return clo_ones32((u & (0u - u)) - 1u);
}

unsigned int clo_sum(unsigned int x) {
	// Closed form of the reference's recursion 0 + 1 + ... + x (mod 2^32).
	unsigned long long v = static_cast<unsigned long long>(x) * (x + 1ull) / 2ull;
	return static_cast<unsigned int>(v);
}

void clo_print_to_null(const gchar* string) { (void)string; }

}  // extern "C"

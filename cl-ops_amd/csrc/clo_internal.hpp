// Internal definitions shared by the library's translation units: the
// layouts behind the opaque CCL* handles, error helpers, the launch wrapper
// that names and (optionally) times every kernel, and device scratch caches.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "cl_ops.h"

struct ccl_context {
	int dev;
	int refc;
};

struct ccl_event {
	hipEvent_t ev;
	std::string name;
};

struct ccl_prof_rec {
	std::string name;
	hipEvent_t start;
	hipEvent_t end;
};

struct ccl_queue {
	CCLContext* ctx;
	hipStream_t stream;
	bool owns_stream;
	bool prof;
	bool capturing = false;                // between clo_graph_begin/end_capture
	size_t capture_mark = 0;               // g_capture_blocks size at begin_capture
	ccl_event* capture_event = nullptr;   // returned for work enqueued while capturing
	std::vector<ccl_event*> events;       // every event this queue recorded
	std::vector<ccl_prof_rec> prof_recs;  // used records since last reset
	std::vector<hipEvent_t> prof_pool;    // spare events for prof records
	std::vector<std::string> prof_names;  // scratch for ccl_prof_get_name
	std::string prof_filter;              // record only kernels with this name prefix
};

struct ccl_buffer {
	CCLContext* ctx;
	void* ptr;
	size_t size;
	bool owns;
	int refc;
};

struct ccl_event_wait_list {
	std::vector<CCLEvent*> evts;
};

struct ccl_program {
	int unused;
};

namespace clo {

enum : GQuark { kQuarkClo = 0x636c6f01u, kQuarkCcl = 0x63636c01u };

void set_error(GError** err, GQuark domain, int code, const char* fmt, ...)
	__attribute__((format(printf, 4, 5)));

// Sets a CCL_ERROR (code = hipError_t) and returns true if `e` is an error.
bool hip_failed(hipError_t e, GError** err, const char* what);

// Device allocation honouring the allocation-failure test hook.
hipError_t dev_malloc(void** p, size_t bytes);

// Records a new event on the queue (owned by the queue).  While the queue is
// being captured into a graph no event is recorded: the queue's placeholder
// event (no hipEvent behind it) is returned instead.
CCLEvent* record_event(CCLQueue* q, const char* name, GError** err);

// Number of graph captures open on the calling thread.  Device scratch is
// never allocated while one is open (Scratch::ensure fails instead).
extern thread_local int g_capture_depth;

// Activates the context's device for the calling thread.
bool use_device(CCLContext* ctx, GError** err);

// Set by entry points that must not create anything on the host or device
// while they run (the no-allocation scan): launches then take no profiling
// records even on a profiling queue.
extern thread_local bool g_launch_quiet;

// Kernel launch wrapper: records start/end events when the queue profiles,
// and converts a launch failure into a GError.
// Debug log of the library's decisions on stderr ("clo: ..."), as CL_Ops
// logs them with g_debug (clo_sort_satradix.c:171,192,
// clo_sort_abitonic.c:398-407, clo_scan_blelloch.c:160): level 1 one line
// per call with the chosen path and sizes, level 2 also every kernel launch.
// The level comes from clo_set_log_level, else the CLO_LOG environment
// variable, else 0 (off).
int log_level();
void log_msg(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
#define CLO_LOG(level, ...)                                      \
	do {                                                         \
		if (::clo::log_level() >= (level)) ::clo::log_msg(__VA_ARGS__); \
	} while (0)

// Enqueues on cq the deferred DEV_GID seed initialisation of every RNG on
// cq's device that has not filled yet (rng.hip).
bool rng_init_pending_on(CCLQueue* cq, GError** err);

// A named range around one public call for tracing tools (roctx, shown by
// rocprofv3 --marker-trace); a no-op when no tool is attached.
struct ApiRange {
	explicit ApiRange(const char* name);
	~ApiRange();
	ApiRange(const ApiRange&) = delete;
	ApiRange& operator=(const ApiRange&) = delete;
};

struct Launch {
	CCLQueue* q;
	const char* name;
	ccl_prof_rec* rec;
	Launch(CCLQueue* q_, const char* name_);
	bool done(GError** err);  // after the launch; false on error
};

// A device allocation shared by the Scratch that made it and by every graph
// captured over calls that used it; freed when the last holder lets go.
struct DevBlock {
	void* ptr;
	std::atomic<int> refs;
};
void devblock_unref(DevBlock* b);

// Blocks used by library calls while a graph capture is open on the calling
// thread; clo_graph_end_capture hands them to the graph (graph.cpp).
extern thread_local std::vector<DevBlock*> g_capture_blocks;

// A growable device scratch buffer owned by a sorter/scanner.  Growing frees
// the previous allocation (hipFree synchronises with the device, so kernels
// still reading it have finished).  A scratch used inside a graph capture is
// held by that graph: it then cannot grow (ensure fails with CLO_ERROR_ARGS
// until the graph is destroyed), and releasing it leaves the memory to the
// graph, which frees it when destroyed.
struct Scratch {
	void* ptr = nullptr;
	size_t size = 0;
	bool zero_on_alloc = false;
	DevBlock* blk = nullptr;
	bool ensure(size_t bytes, GError** err);
	void release();
	~Scratch() { release(); }
};

inline size_t div_ceil(size_t a, size_t b) { return (a + b - 1) / b; }

// Option string "k=v,k=v".  Returns false (and sets CLO_ERROR_ARGS) on a
// token without exactly one '='.
struct Opt {
	std::string key, value;
};
bool parse_options(const char* options, std::vector<Opt>& out,
	const char* algo, GError** err);

// Normalises a compare snippet: 0 = ascending ("((a) > (b))" or NULL),
// 1 = descending ("((a) < (b))"), -1 = unsupported.
int compare_kind(const char* compare);
// True for NULL, "", "(x)", "x".
bool get_key_is_identity(const char* get_key);

bool type_is_integer(CloType t);
bool type_is_signed(CloType t);

// Segmented sort defaults: items per thread of the 256-thread tile and
// blocks per CU of the persistent launch, measured on 32768 log-uniform
// segments of up to 65536 keys.  Keys alone (medium and large segments on the
// MSD bucketing kernel): tile 8, 2.30 ms for 32-bit keys and 3.40 ms for
// 64-bit keys, against 2.52-2.71 and 6.5 ms on the digit passes (tile 16,
// the tuner's best there).  Key-value sorts (digit passes): tile 16 for
// 32-bit keys (4.28 ms against 5.05 for tile 8), tile 8 for 64-bit keys
// (tile 16 spills registers).  The dispatcher's fixed segmented
// configuration (keys only) is the same.
constexpr uint32_t seg_default_tile(size_t key_bytes, bool kv = false) {
	return (kv && key_bytes == 4) ? 16u : 8u;
}
constexpr uint32_t kSegBlocksPerCuDefault = 4;
// 64-bit key-value sorts: at most two blocks per CU (three resident blocks,
// each with a large segment's keys and values in flight, measured 16% slower).
constexpr uint32_t kSegBlocksPerCuKv64 = 2;

}  // namespace clo

// Internal accessors of the sorter shared by the algorithm files.
struct clo_sort {
	CloSortImplDef impl;
	CCLContext* ctx;
	CloType elem_type;
	CloType key_type;
	int descending;
	void* data;
	// Set by satradix: device word with the executed pass count of the last
	// call, and the queue it ran on.
	CCLQueue* last_queue;
	// Segmented sort: tuning (segtile, seggrid, segclass options) and its
	// scratch.  seg_cls: least work class per length bin, two bits per bin
	// (0 chunk, 1 medium, 2 large; all zero = by length alone).
	uint32_t seg_tile;
	uint32_t seg_grid;
	uint64_t seg_cls[2];
	// Keys-only segmented sorts: 1 = MSD bucketing kernel for medium and large
	// segments, 0 = digit passes only, -1 = default (1).
	int32_t seg_msd;
	void* seg_cache;
	// Run-time compiled kernels for user compare / key snippets (NULL when
	// the ahead-of-time kernels serve the sorter).
	void* snip;
	// Device buffers of the host-data path, kept between calls.
	void* host_cache;
};

struct clo_scan {
	CloScanImplDef impl;
	CCLContext* ctx;
	CloType elem_type;
	CloType sum_type;
	void* data;
	void* host_cache = nullptr;  // device buffers of the host-data path
};

// KV entry points of the algorithms that support them.
CCLEvent* clo_satradix_sort_kv(CloSort* sorter, CCLQueue* cq_exec,
	CCLQueue* cq_comm, CCLBuffer* keys_in, CCLBuffer* vals_in,
	CCLBuffer* keys_out, CCLBuffer* vals_out, size_t numel, GError** err);
CCLEvent* clo_gselect_sort_kv(CloSort* sorter, CCLQueue* cq_exec,
	CCLQueue* cq_comm, CCLBuffer* keys_in, CCLBuffer* vals_in,
	CCLBuffer* keys_out, CCLBuffer* vals_out, size_t numel, GError** err);
cl_uint clo_satradix_last_passes(CloSort* sorter, GError** err);
CCLEvent* clo_satradix_sort_segmented(CloSort* sorter, CCLQueue* cq, CCLBuffer* kin,
	CCLBuffer* vin, CCLBuffer* kout, CCLBuffer* vout, CCLBuffer* offsets, size_t nseg,
	size_t numel, uint32_t tile_items, uint32_t grid, void** cache, GError** err);
void clo_satradix_segsort_cache_free(void* cache);
// Stable satradix sort of (key word, u32 payload) pairs with the sorter's
// radix settings; ksize is 1, 2, 4 or 8.  keys_out/vals_out NULL: in place.
bool clo_satradix_sort_pairs(CloSort* sorter, CCLQueue* q, size_t ksize, void* keys,
	uint32_t* vals, void* keys_out, uint32_t* vals_out, size_t n, GError** err);

namespace clo {
// hipRTC (rtc.cpp): the functions `names` of `src` compiled for the device of
// ctx, one compilation per distinct source and options in the process.  A
// source that does not compile fails with CLO_ERROR_ARGS naming `what` and
// quoting the compiler log.
bool rtc_functions(CCLContext* ctx, const std::string& src, const char* compiler_opts,
	const char* what, const char* const* names, int nnames, hipFunction_t* out, GError** err);
// Launches a run-time compiled kernel over `items` work-items in blocks of
// 256 on q, profiled under `name`.
bool rtc_launch(CCLQueue* q, const char* name, hipFunction_t f, size_t items, void** args,
	GError** err);

// Host-data paths through pinned staging (host_path.cpp).  Blocking.
// sort: keys (and optional u32 vals) in place on the host, except that the
// sorted keys go to keys_out.  scan: `in` -> `out`, chunk by chunk.
bool sort_host_path(CloSort* s, CCLQueue* cq_exec, CCLQueue* cq_comm, void* keys, void* vals,
	void* keys_out, size_t numel, GError** err);
bool scan_host_path(CloScan* sc, CCLQueue* cq_exec, CCLQueue* cq_comm, const void* in, void* out,
	size_t numel, GError** err);
void host_cache_free(void* cache);

// Sorters with user snippets (snippet.cpp).  needed(): the snippets, types
// or key getter are beyond the ahead-of-time kernels.  init(): compiles the
// run-time kernels for the sorter's device.  sort(): the algorithm of
// s->impl.name on them; vin/vout are an optional u32 payload (stable
// algorithms only).
bool snippet_sorter_needed(const CloSort* s, const char* compare, const char* get_key);
bool snippet_sorter_init(CloSort* s, const char* compare, const char* get_key,
	const char* compiler_opts, GError** err);
void snippet_sorter_free(CloSort* s);
const char* snippet_sorter_source(const CloSort* s);
CCLEvent* snippet_sort(CloSort* s, CCLQueue* q, CCLBuffer* kin, CCLBuffer* vin, CCLBuffer* kout,
	CCLBuffer* vout, size_t numel, GError** err);
}  // namespace clo

// One chunk of a plain blelloch scan of a longer array: the running total
// before the chunk is read from the device word `chain` (sum type) and the
// total after it is left there.  Chunks must run in order on one queue.
bool clo_blelloch_scan_chunk(CloScan* scanner, CCLQueue* cq, const void* in, void* out,
	size_t numel, void* chain, GError** err);

// Plain blelloch scan with 8, 16 (the default) or 32 items per thread (8 and
// 32 for unsigned 32/64-bit elements only).
CCLEvent* clo_blelloch_scan_items(CloScan* scanner, CCLQueue* cq, CCLBuffer* in,
	CCLBuffer* out, size_t numel, uint32_t items, GError** err);

// Host-data paths: clo_sort_with_host_data, clo_sort_kv_with_host_data and
// clo_scan_with_host_data (CL_Ops' sort/clo_sort_abstract.c:296-418 and
// scan/clo_scan_abstract.c:255-362) through pinned staging memory.
//
// Pageable memory cannot be read by the copy engines directly, so each
// direction stages through a ring of pinned slots: the host copies chunk i
// into a slot (split across a few threads) while the copy engine moves chunk
// i-1 on to the device.  Sorts copy in on cq_comm, sort on cq_exec and copy
// out on a second internal stream.  Scans run chunk by chunk: copy-in of a
// chunk on cq_comm, its scan on cq_exec (the running total crosses chunks
// on the device), its copy-out on the internal stream, so the two copy
// directions and the scans overlap.

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "clo_internal.hpp"
#include "copy_pool.hpp"

using namespace clo;

namespace {

constexpr size_t kChunk = 8u << 20;  // bytes per pinned slot
constexpr int kSlots = 3;            // slots per direction
constexpr int kMaxDevices = 64;

// ---- pinned slots per device -------------------------------------------------

enum Dir { kIn = 0, kOut = 1 };

struct PinnedPool {
	std::mutex mu;  // held for a whole host-data call
	bool ready = false;
	char* slot[2][kSlots] = {};
	hipEvent_t ev[2][kSlots] = {};  // last copy through each slot
	bool used[2][kSlots] = {};
	hipEvent_t step[kSlots] = {};   // scan chunk done (exec -> out stream)
	hipEvent_t done = nullptr;      // sort done (exec -> out stream)
	hipStream_t out = nullptr;      // device-to-host copies

	// Allocates the slots on first use.  They stay allocated until the
	// process ends: 2 * kSlots * kChunk bytes of pinned memory per device.
	bool ensure(GError** err) {
		if (ready) return true;
		for (int d = 0; d < 2; ++d)
			for (int i = 0; i < kSlots; ++i) {
				if (hip_failed(hipHostMalloc(reinterpret_cast<void**>(&slot[d][i]), kChunk),
						err, "hipHostMalloc") ||
					hip_failed(hipEventCreateWithFlags(&ev[d][i], hipEventDisableTiming), err,
						"hipEventCreate"))
					return false;
			}
		for (int i = 0; i < kSlots; ++i)
			if (hip_failed(hipEventCreateWithFlags(&step[i], hipEventDisableTiming), err,
					"hipEventCreate"))
				return false;
		if (hip_failed(hipEventCreateWithFlags(&done, hipEventDisableTiming), err,
				"hipEventCreate") ||
			hip_failed(hipStreamCreateWithFlags(&out, hipStreamNonBlocking), err,
				"hipStreamCreate"))
			return false;
		ready = true;
		return true;
	}
};

PinnedPool& pool(int dev) {
	static PinnedPool* pools = new PinnedPool[kMaxDevices];
	return pools[dev];
}

// Waits for everything a call enqueued before the pool is released, also on
// the error paths (slots may still be in flight).
struct Drain {
	hipStream_t s[3];
	~Drain() {
		for (hipStream_t x : s)
			if (x) (void)hipStreamSynchronize(x);
	}
};

// Host -> device through the in-slots on stream st; returns once the last
// chunk is enqueued (st orders whatever follows).
bool stage_in(PinnedPool& P, hipStream_t st, void* dev, const void* host, size_t bytes,
	GError** err) {
	char* d = static_cast<char*>(dev);
	const char* h = static_cast<const char*>(host);
	int i = 0;
	for (size_t off = 0; off < bytes; off += kChunk, i = (i + 1) % kSlots) {
		const size_t len = std::min(kChunk, bytes - off);
		if (P.used[kIn][i] &&
			hip_failed(hipEventSynchronize(P.ev[kIn][i]), err, "hipEventSynchronize"))
			return false;
		CopyPool::get().copy(P.slot[kIn][i], h + off, len);
		if (hip_failed(hipMemcpyAsync(d + off, P.slot[kIn][i], len, hipMemcpyHostToDevice, st),
				err, "hipMemcpyAsync H2D") ||
			hip_failed(hipEventRecord(P.ev[kIn][i], st), err, "hipEventRecord"))
			return false;
		P.used[kIn][i] = true;
	}
	return true;
}

// Device -> host through the out-slots on the pool's out stream (ordered
// after whatever that stream already waits on); returns with the data in
// host memory.
bool stage_out(PinnedPool& P, void* host, const void* dev, size_t bytes, GError** err) {
	const size_t nch = div_ceil(bytes, kChunk);
	char* h = static_cast<char*>(host);
	const char* d = static_cast<const char*>(dev);
	size_t issued = 0;
	for (size_t c = 0; c < nch; ++c) {
		for (; issued < nch && issued < c + kSlots; ++issued) {
			const int i = static_cast<int>(issued % kSlots);
			const size_t off = issued * kChunk, len = std::min(kChunk, bytes - off);
			if (hip_failed(hipMemcpyAsync(P.slot[kOut][i], d + off, len, hipMemcpyDeviceToHost,
					P.out), err, "hipMemcpyAsync D2H") ||
				hip_failed(hipEventRecord(P.ev[kOut][i], P.out), err, "hipEventRecord"))
				return false;
		}
		const int i = static_cast<int>(c % kSlots);
		const size_t off = c * kChunk, len = std::min(kChunk, bytes - off);
		if (hip_failed(hipEventSynchronize(P.ev[kOut][i]), err, "hipEventSynchronize"))
			return false;
		CopyPool::get().copy(h + off, P.slot[kOut][i], len);
	}
	return true;
}

}  // namespace

namespace clo {

struct HostCache {
	Scratch a, b, c, d, chain;
};

void host_cache_free(void* cache) { delete static_cast<HostCache*>(cache); }

static HostCache* cache_of(void** slot) {
	if (!*slot) *slot = new HostCache();
	return static_cast<HostCache*>(*slot);
}

bool sort_host_path(CloSort* s, CCLQueue* cq_exec, CCLQueue* cq_comm, void* keys, void* vals,
	void* keys_out, size_t numel, GError** err) {
	const size_t kb = numel * clo_type_sizeof(s->elem_type), vb = vals ? numel * 4 : 0;
	if (!use_device(s->ctx, err)) return false;
	PinnedPool& P = pool(s->ctx->dev);
	std::lock_guard<std::mutex> lock(P.mu);
	if (!P.ensure(err)) return false;
	HostCache* hc = cache_of(&s->host_cache);
	// Out-of-place keys sorts need a second device buffer (the in-place
	// algorithms take NULL for data_out, clo_sort_abstract.c:353-367).
	const bool oop = !vals && !s->impl.in_place;
	if (!hc->a.ensure(std::max<size_t>(kb, 1), err) || (vals && !hc->b.ensure(vb, err)) ||
		(oop && !hc->c.ensure(std::max<size_t>(kb, 1), err)))
		return false;
	Drain drain{{cq_comm->stream, cq_exec->stream, P.out}};
	if (!stage_in(P, cq_comm->stream, hc->a.ptr, keys, kb, err) ||
		(vals && !stage_in(P, cq_comm->stream, hc->b.ptr, vals, vb, err)) ||
		hip_failed(hipEventRecord(P.done, cq_comm->stream), err, "hipEventRecord") ||
		hip_failed(hipStreamWaitEvent(cq_exec->stream, P.done, 0), err, "hipStreamWaitEvent"))
		return false;
	CCLBuffer* dk = ccl_buffer_new_wrap(s->ctx, hc->a.ptr, kb);
	CCLBuffer* dv = vals ? ccl_buffer_new_wrap(s->ctx, hc->b.ptr, vb) : nullptr;
	CCLBuffer* dout = oop ? ccl_buffer_new_wrap(s->ctx, hc->c.ptr, kb) : nullptr;
	CCLEvent* evt = vals
		? clo_sort_kv_with_device_data(s, cq_exec, cq_comm, dk, dv, nullptr, nullptr, numel, 0,
			err)
		: clo_sort_with_device_data(s, cq_exec, cq_comm, dk, dout, numel, 0, err);
	ccl_buffer_destroy(dk);
	if (dv) ccl_buffer_destroy(dv);
	if (dout) ccl_buffer_destroy(dout);
	if (!evt ||
		hip_failed(hipEventRecord(P.done, cq_exec->stream), err, "hipEventRecord") ||
		hip_failed(hipStreamWaitEvent(P.out, P.done, 0), err, "hipStreamWaitEvent"))
		return false;
	return stage_out(P, keys_out, oop ? hc->c.ptr : hc->a.ptr, kb, err) &&
		(!vals || stage_out(P, vals, hc->b.ptr, vb, err));
}

bool scan_host_path(CloScan* sc, CCLQueue* cq_exec, CCLQueue* cq_comm, const void* in, void* out,
	size_t numel, GError** err) {
	const size_t es = clo_type_sizeof(sc->elem_type), ss = clo_type_sizeof(sc->sum_type);
	if (!use_device(sc->ctx, err)) return false;
	PinnedPool& P = pool(sc->ctx->dev);
	std::lock_guard<std::mutex> lock(P.mu);
	if (!P.ensure(err)) return false;
	HostCache* hc = cache_of(&sc->host_cache);
	if (!hc->a.ensure(std::max<size_t>(numel * es, 1), err) ||
		!hc->b.ensure(std::max<size_t>(numel * ss, 1), err) || !hc->chain.ensure(8, err))
		return false;
	Drain drain{{cq_comm->stream, cq_exec->stream, P.out}};
	if (hip_failed(hipMemsetAsync(hc->chain.ptr, 0, 8, cq_exec->stream), err, "hipMemsetAsync"))
		return false;
	// Chunks of E elements: one slot holds a chunk of input or of output.
	const size_t E = kChunk / std::max(es, ss);
	const size_t nch = div_ceil(numel, E);
	const char* hin = static_cast<const char*>(in);
	char* hout = static_cast<char*>(out);
	char* din = static_cast<char*>(hc->a.ptr);
	char* dout = static_cast<char*>(hc->b.ptr);
	auto drain_out = [&](size_t c) {
		const int i = static_cast<int>(c % kSlots);
		const size_t len = std::min(E, numel - c * E) * ss;
		if (hip_failed(hipEventSynchronize(P.ev[kOut][i]), err, "hipEventSynchronize"))
			return false;
		CopyPool::get().copy(hout + c * E * ss, P.slot[kOut][i], len);
		return true;
	};
	for (size_t c = 0; c < nch; ++c) {
		const int i = static_cast<int>(c % kSlots);
		const size_t cnt = std::min(E, numel - c * E);
		// in: host -> slot -> device on cq_comm
		if (P.used[kIn][i] &&
			hip_failed(hipEventSynchronize(P.ev[kIn][i]), err, "hipEventSynchronize"))
			return false;
		CopyPool::get().copy(P.slot[kIn][i], hin + c * E * es, cnt * es);
		if (hip_failed(hipMemcpyAsync(din + c * E * es, P.slot[kIn][i], cnt * es,
				hipMemcpyHostToDevice, cq_comm->stream), err, "hipMemcpyAsync H2D") ||
			hip_failed(hipEventRecord(P.ev[kIn][i], cq_comm->stream), err, "hipEventRecord"))
			return false;
		P.used[kIn][i] = true;
		// scan on cq_exec once the chunk is on the device
		if (hip_failed(hipStreamWaitEvent(cq_exec->stream, P.ev[kIn][i], 0), err,
				"hipStreamWaitEvent") ||
			!clo_blelloch_scan_chunk(sc, cq_exec, din + c * E * es, dout + c * E * ss, cnt,
				hc->chain.ptr, err) ||
			hip_failed(hipEventRecord(P.step[i], cq_exec->stream), err, "hipEventRecord") ||
			hip_failed(hipStreamWaitEvent(P.out, P.step[i], 0), err, "hipStreamWaitEvent"))
			return false;
		// out: device -> slot on the out stream; the slot's previous chunk
		// goes to host memory first
		if (c >= static_cast<size_t>(kSlots) && !drain_out(c - kSlots)) return false;
		if (hip_failed(hipMemcpyAsync(P.slot[kOut][i], dout + c * E * ss, cnt * ss,
				hipMemcpyDeviceToHost, P.out), err, "hipMemcpyAsync D2H") ||
			hip_failed(hipEventRecord(P.ev[kOut][i], P.out), err, "hipEventRecord"))
			return false;
	}
	for (size_t c = nch > static_cast<size_t>(kSlots) ? nch - kSlots : 0; c < nch; ++c)
		if (!drain_out(c)) return false;
	return true;
}

}  // namespace clo

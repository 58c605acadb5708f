// RNG objects (CL_Ops' src/cl_ops/rng/clo_rng.c:35-481) and their gfx950
// kernels: device seed initialisation (clo_rng_init.cl:27-60) and the
// register-resident multi-draw fill.

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "clo_internal.hpp"
#include "clo_rng_device.h"
#include "dev_common.hpp"

using namespace clo;
using namespace clo::dev;

namespace {

// Generated at build time from include/clo_rng_device.h (the @@section
// blocks): kSrcLcg, kSrcXorshift64, kSrcXorshift128, kSrcMwc64x,
// kSrcParkmiller, kSrcTauslcg, kSrcApi.
#include "rng_sources.inc"

enum GenId { kLcg = 0, kXs64 = 1, kXs128 = 2, kMwc = 3, kPm = 4, kTaus = 5 };

// Seed hash of the DEV_GID initialisation (clo_rng_init.cl:27-38).
enum HashId { kHashNone = 0, kHashKnuth = 1, kHashXs1 = 2 };

__device__ __forceinline__ u64 apply_hash(u64 x, u32 hash) {
	if (hash == kHashKnuth) {
		x = (x * 2654435761ull) % 0x100000000ull;
	} else if (hash == kHashXs1) {
		x = ((x >> 16) ^ x) * 0x45d9f3bull;
		x = ((x >> 16) ^ x) * 0x45d9f3bull;
		x = ((x >> 16) ^ x);
	}
	return x;
}

template <typename NS_STATE, NS_STATE (*CONV)(unsigned long long)>
__global__ __launch_bounds__(256) void rng_init_kernel(NS_STATE* seeds, u64 count, u64 main_seed,
	u32 hash) {
	const u64 gid = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (gid >= count) return;
	u64 seed = gid + main_seed;
	seed = apply_hash(seed, hash);
	seeds[gid] = CONV(seed);
}

// out[r * G + gid] for r < draws; the state is read once and written once.
// INIT: the first fill of DEV_GID seeds, whose initialisation
// (clo_rng_init.cl:47-60) was deferred to here: the state comes from
// CONV(hash(gid + main_seed)) instead of memory, so the states are written
// once instead of written, read and written again.
template <typename S, unsigned int (*STEP)(S&), S (*CONV)(unsigned long long), bool INIT>
__global__ __launch_bounds__(256) void rng_fill_kernel(S* __restrict__ states, u32* __restrict__ out,
	u64 G, u32 draws, u32 shift, u32 maxint, u64 main_seed, u32 hash) {
	const u64 gid = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (gid >= G) return;
	S s;
	if constexpr (INIT) s = CONV(apply_hash(gid + main_seed, hash));
	else s = states[gid];
	u32* o = out + gid;
	if (maxint == 0) {
		for (u32 r = 0; r < draws; ++r) {
			u32 v = STEP(s) >> shift;
			__builtin_nontemporal_store(v, o);
			o += G;
		}
	} else {
		for (u32 r = 0; r < draws; ++r) {
			u32 v = STEP(s) % maxint;
			__builtin_nontemporal_store(v, o);
			o += G;
		}
	}
	states[gid] = s;
}

struct Info {
	const char* name;
	size_t seed_size;
	const char* src;
};

const Info kInfos[] = {
	{"lcg", 8, kSrcLcg},
	{"xorshift64", 8, kSrcXorshift64},
	{"xorshift128", 16, kSrcXorshift128},
	{"mwc64x", 8, kSrcMwc64x},
	{"parkmiller", 4, kSrcParkmiller},
	{"tauslcg", 16, kSrcTauslcg},
};

// MT19937 (Matsumoto & Nishimura 1998) with init_genrand seeding: the
// generator behind GLib's g_rand_new_with_seed()/g_rand_int(), which CL_Ops
// uses for HOST_MT seeds (clo_rng.c:196-203).
struct Mt19937 {
	u32 mt[624];
	int mti;
	explicit Mt19937(u32 seed) {
		mt[0] = seed;
		for (mti = 1; mti < 624; ++mti)
			mt[mti] = 1812433253u * (mt[mti - 1] ^ (mt[mti - 1] >> 30)) + static_cast<u32>(mti);
	}
	u32 next() {
		if (mti >= 624) {
			for (int k = 0; k < 624; ++k) {
				u32 y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
				mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
			}
			mti = 0;
		}
		u32 y = mt[mti++];
		y ^= (y >> 11);
		y ^= (y << 7) & 0x9d2c5680u;
		y ^= (y << 15) & 0xefc60000u;
		y ^= (y >> 18);
		return y;
	}
};

}  // namespace

struct clo_rng {
	int gen;
	CCLBuffer* seeds;
	size_t count;
	std::string src;
	// DEV_GID seeds with a predefined hash are initialised by the first fill
	// (or when the seeds are first read, or a graph capture begins on the
	// device): until then `pending` and the seeds buffer holds no states.
	bool pending = false;
	u64 init_seed = 0;
	u32 init_hash = 0;
};

namespace {
std::mutex g_pending_mu;
std::vector<CloRng*> g_pending;  // RNGs whose DEV_GID initialisation is deferred

void pending_remove(CloRng* r) {
	std::lock_guard<std::mutex> lk(g_pending_mu);
	g_pending.erase(std::remove(g_pending.begin(), g_pending.end(), r), g_pending.end());
}
template <typename S, unsigned int (*STEP)(S&), S (*CONV)(unsigned long long)>
void launch_fill(bool init, u32 g, hipStream_t s, void* st, u32* o, u64 G, u32 draws, u32 shift,
	u32 maxint, const CloRng* rng) {
	if (init)
		rng_fill_kernel<S, STEP, CONV, true><<<g, 256, 0, s>>>(static_cast<S*>(st), o, G, draws,
			shift, maxint, rng->init_seed, rng->init_hash);
	else
		rng_fill_kernel<S, STEP, CONV, false><<<g, 256, 0, s>>>(static_cast<S*>(st), o, G, draws,
			shift, maxint, 0, 0);
}
}  // namespace

extern "C" {

// Host-pass only: the vtable holds host function addresses.
#if !defined(__HIP_DEVICE_COMPILE__)
const struct clo_rng_info clo_rng_infos[] = {
	{"lcg", kSrcLcg, 8},
	{"xorshift64", kSrcXorshift64, 8},
	{"xorshift128", kSrcXorshift128, 16},
	{"mwc64x", kSrcMwc64x, 8},
	{"parkmiller", kSrcParkmiller, 4},
	{"tauslcg", kSrcTauslcg, 16},
	{nullptr, nullptr, 0},
};
#endif


// The device source of generator `gen` that clo_rng_get_source() returns:
// its section, the draw API, and the API bound to it.
static std::string gen_source(int gen) {
	return std::string(kInfos[gen].src) + kSrcApi + "CLO_RNG_DEFINE_API(clo_rng_" +
		kInfos[gen].name + ")\nusing namespace clo_rng_" + kInfos[gen].name + ";\n";
}

static bool launch_init(int gen, void* seeds, u64 count, u64 main_seed, u32 hash, CCLQueue* cq,
	GError** err) {
	u64 g = (count + 255) / 256;
	Launch L(cq, "clo_rng_init");
	hipStream_t st = cq->stream;
	switch (gen) {
	case kLcg:
		rng_init_kernel<clo_rng_lcg::clo_statetype, clo_rng_lcg::clo_ulong2statetype>
			<<<static_cast<u32>(g), 256, 0, st>>>(static_cast<clo_rng_lcg::clo_statetype*>(seeds),
			count, main_seed, hash);
		break;
	case kXs64:
		rng_init_kernel<clo_rng_xorshift64::clo_statetype, clo_rng_xorshift64::clo_ulong2statetype>
			<<<static_cast<u32>(g), 256, 0, st>>>(
			static_cast<clo_rng_xorshift64::clo_statetype*>(seeds), count, main_seed, hash);
		break;
	case kXs128:
		rng_init_kernel<clo_rng_xorshift128::clo_statetype, clo_rng_xorshift128::clo_ulong2statetype>
			<<<static_cast<u32>(g), 256, 0, st>>>(
			static_cast<clo_rng_xorshift128::clo_statetype*>(seeds), count, main_seed, hash);
		break;
	case kMwc:
		rng_init_kernel<clo_rng_mwc64x::clo_statetype, clo_rng_mwc64x::clo_ulong2statetype>
			<<<static_cast<u32>(g), 256, 0, st>>>(
			static_cast<clo_rng_mwc64x::clo_statetype*>(seeds), count, main_seed, hash);
		break;
	case kPm:
		rng_init_kernel<clo_rng_parkmiller::clo_statetype, clo_rng_parkmiller::clo_ulong2statetype>
			<<<static_cast<u32>(g), 256, 0, st>>>(
			static_cast<clo_rng_parkmiller::clo_statetype*>(seeds), count, main_seed, hash);
		break;
	case kTaus:
		rng_init_kernel<clo_rng_tauslcg::clo_statetype, clo_rng_tauslcg::clo_ulong2statetype>
			<<<static_cast<u32>(g), 256, 0, st>>>(
			static_cast<clo_rng_tauslcg::clo_statetype*>(seeds), count, main_seed, hash);
		break;
	}
	return L.done(err);
}

}  // extern "C"

namespace clo {

// Deferred DEV_GID initialisations of RNGs on cq's device, enqueued on cq
// (a graph capture must not replay an initialisation; clo_graph_begin_capture
// calls this first).
bool rng_init_pending_on(CCLQueue* cq, GError** err) {
	std::vector<CloRng*> todo;
	{
		std::lock_guard<std::mutex> lk(g_pending_mu);
		for (CloRng* r : g_pending)
			if (r->seeds->ctx->dev == cq->ctx->dev) todo.push_back(r);
	}
	for (CloRng* r : todo) {
		if (!launch_init(r->gen, r->seeds->ptr, r->count, r->init_seed, r->init_hash, cq, err))
			return false;
		r->pending = false;
		pending_remove(r);
	}
	return true;
}

}  // namespace clo

extern "C" {

// DEV_GID initialisation for a seed hash other than the predefined ones:
// clo_rng_init.cl's kernel with `#define CLO_RNG_HASH(x) <hash>` prepended
// (clo_rng.c:101-109), compiled at run time after the generator's device
// source.  KNUTH and XS1 stay available to the snippet.
static const char kInitRtc[] =
	"#define KNUTH(x) x = ((x * 2654435761) % 0x100000000)\n"
	"#define XS1(x) x = ((x >> 16) ^ x) * 0x45d9f3b; x = ((x >> 16) ^ x) * 0x45d9f3b; "
	"x = ((x >> 16) ^ x);\n"
	"extern \"C\" __global__ __launch_bounds__(256) void clo_rng_init(clo_statetype* seeds,\n"
	"	ulong count, ulong main_seed) {\n"
	"	const ulong gid = (ulong)blockIdx.x * 256 + threadIdx.x;\n"
	"	if (gid >= count) return;\n"
	"	ulong seed = gid + main_seed;\n"
	"	CLO_RNG_HASH(seed);\n"
	"	seeds[gid] = clo_ulong2statetype(seed);\n"
	"}\n";

static std::string strip(const char* s) {
	std::string r;
	for (; s && *s; ++s)
		if (*s != ' ' && *s != '\t' && *s != '\n') r += *s;
	return r;
}

CloRng* clo_rng_new(const char* type, CloRngSeedType seed_type, void* seeds, size_t seeds_count,
	cl_ulong main_seed, const char* hash, CCLContext* ctx, CCLQueue* cq, GError** err) {
	ApiRange api_range("clo_rng_new");
	if (err && *err) return nullptr;
	if (!ctx) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "clo_rng_new: NULL context");
		return nullptr;
	}
	int gen = -1;
	for (int i = 0; i < static_cast<int>(sizeof(kInfos) / sizeof(kInfos[0])); ++i)
		if (type && strcmp(type, kInfos[i].name) == 0) gen = i;
	if (gen < 0) {
		set_error(err, kQuarkClo, CLO_ERROR_IMPL_NOT_FOUND,
			"The requested RNG implementation, '%s', was not found.", type ? type : "(null)");
		return nullptr;
	}
	const size_t ss = kInfos[gen].seed_size;
	if (!use_device(ctx, err)) return nullptr;
	CCLBuffer* buf = nullptr;
	CCLQueue* own = nullptr;
	switch (seed_type) {
	case CLO_RNG_SEED_DEV_GID: {
		// clo_rng.c:85-153.
		std::string h = strip(hash);
		// The predefined hashes run the precompiled kernel; any other
		// snippet is compiled with the generator's source.
		int hid = -1;
		if (h.empty() || h == "x" || h == "x=x") hid = kHashNone;
		else if (h == "KNUTH(x)") hid = kHashKnuth;
		else if (h == "XS1(x)") hid = kHashXs1;
		hipFunction_t init_fn = nullptr;
		if (hid < 0) {
			const std::string src = std::string("#define CLO_RNG_HASH(x) ") + hash + "\n" +
				gen_source(gen) + kInitRtc;
			const char* name = "clo_rng_init";
			if (!rtc_functions(ctx, src, nullptr, "seed hash snippet", &name, 1, &init_fn, err))
				return nullptr;
		}
		if (seeds_count == 0) {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS, "seeds_count must be positive");
			return nullptr;
		}
		if (!cq) {
			own = ccl_queue_new(ctx, nullptr, 0, err);
			if (!own) return nullptr;
		}
		CCLQueue* iq = cq ? cq : own;
		buf = ccl_buffer_new(ctx, CL_MEM_READ_WRITE, seeds_count * ss, nullptr, err);
		bool launched = false;
		if (buf && init_fn) {
			void* p = buf->ptr;
			u64 cnt = seeds_count, ms = main_seed;
			void* args[] = {&p, &cnt, &ms};
			launched = rtc_launch(iq, "clo_rng_init", init_fn, seeds_count, args, err);
		} else if (buf) {
			// Predefined hash: the initialisation is deferred to the first
			// fill, which computes the states instead of reading them.
			CloRng* r = new clo_rng();
			r->gen = gen;
			r->seeds = buf;
			r->count = seeds_count;
			r->src = gen_source(gen);
			r->pending = true;
			r->init_seed = main_seed;
			r->init_hash = static_cast<u32>(hid);
			{
				std::lock_guard<std::mutex> lk(g_pending_mu);
				g_pending.push_back(r);
			}
			if (own) ccl_queue_destroy(own);
			CLO_LOG(1, "rng new: %s, %zu work-items, DEV_GID seeds (hash %d) deferred to the "
				"first fill", kInfos[gen].name, seeds_count, hid);
			return r;
		}
		if (!launched) {
			if (buf) ccl_buffer_destroy(buf);
			buf = nullptr;
		}
		if (own) {
			if (buf) ccl_queue_finish(own, err);
			ccl_queue_destroy(own);
		}
		if (!buf) return nullptr;
		break;
	}
	case CLO_RNG_SEED_HOST_MT: {
		// clo_rng.c:168-234: word i = i-th output of MT19937((uint32) main_seed).
		if (seeds_count == 0) {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS, "seeds_count must be positive");
			return nullptr;
		}
		const size_t bytes = seeds_count * ss;
		u32* host = nullptr;
		if (hip_failed(hipHostMalloc(reinterpret_cast<void**>(&host), bytes), err, "hipHostMalloc"))
			return nullptr;
		Mt19937 mt(static_cast<u32>(main_seed));
		for (size_t i = 0; i < bytes / 4; ++i) host[i] = mt.next();
		buf = ccl_buffer_new(ctx, CL_MEM_READ_WRITE, bytes, nullptr, err);
		if (buf && hip_failed(hipMemcpy(buf->ptr, host, bytes, hipMemcpyHostToDevice), err,
				"hipMemcpy H2D")) {
			ccl_buffer_destroy(buf);
			buf = nullptr;
		}
		(void)hipHostFree(host);
		if (!buf) return nullptr;
		break;
	}
	case CLO_RNG_SEED_EXT_DEV: {
		// clo_rng.c:321-338: reference the caller's buffer.
		CCLBuffer* ext = static_cast<CCLBuffer*>(seeds);
		if (!ext || ext->size < seeds_count * ss || seeds_count == 0) {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS,
				"EXT_DEV seeds must be a device buffer of seeds_count states");
			return nullptr;
		}
		ccl_buffer_ref(ext);
		buf = ext;
		break;
	}
	case CLO_RNG_SEED_EXT_HOST: {
		// clo_rng.c:340-357: blocking copy of the caller's host seeds.
		if (!seeds || seeds_count == 0) {
			set_error(err, kQuarkClo, CLO_ERROR_ARGS,
				"EXT_HOST seeds must be a host array of seeds_count states");
			return nullptr;
		}
		buf = ccl_buffer_new(ctx, CL_MEM_READ_WRITE, seeds_count * ss, seeds, err);
		if (!buf) return nullptr;
		break;
	}
	default:
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "Unknown seed type %d", seed_type);
		return nullptr;
	}
	CloRng* r = new clo_rng();
	r->gen = gen;
	r->seeds = buf;
	r->count = seeds_count;
	r->src = gen_source(gen);
	return r;
}

void clo_rng_destroy(CloRng* rng) {
	if (!rng) return;
	if (rng->pending) pending_remove(rng);
	ccl_buffer_destroy(rng->seeds);
	delete rng;
}

const char* clo_rng_get_source(CloRng* rng) { return rng ? rng->src.c_str() : nullptr; }

CCLBuffer* clo_rng_get_device_seeds(CloRng* rng) {
	if (!rng) return nullptr;
	if (rng->pending) {
		// The caller may read the buffer at once: initialise it now, blocking.
		GError* e = nullptr;
		CCLQueue* q = ccl_queue_new(rng->seeds->ctx, nullptr, 0, &e);
		if (q && launch_init(rng->gen, rng->seeds->ptr, rng->count, rng->init_seed,
				rng->init_hash, q, &e) && ccl_queue_finish(q, &e)) {
			rng->pending = false;
			pending_remove(rng);
		}
		if (q) ccl_queue_destroy(q);
		if (e) {
			fprintf(stderr, "clo_rng_get_device_seeds: %s\n", e->message ? e->message : "error");
			clo_error_free(e);
			return nullptr;
		}
	}
	return rng->seeds;
}
size_t clo_rng_get_size(CloRng* rng) {
	return rng ? rng->count * kInfos[rng->gen].seed_size : 0;
}

CCLEvent* clo_rng_fill(CloRng* rng, CCLQueue* cq, CCLBuffer* out, cl_uint draws, cl_uint bits,
	cl_uint maxint, GError** err) {
	ApiRange api_range("clo_rng_fill");
	if (err && *err) return nullptr;
	if (!rng || !cq || !out) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "NULL rng, queue or buffer");
		return nullptr;
	}
	if (maxint == 0 && (bits < 1 || bits > 32)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "bits must be in 1..32");
		return nullptr;
	}
	const u64 G = rng->count;
	if (out->size < static_cast<size_t>(draws) * G * 4) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS, "output buffer smaller than draws * G words");
		return nullptr;
	}
	if (!use_device(cq->ctx, err)) return nullptr;
	CLO_LOG(1, "rng fill: generator %d, %llu work-items x %u draws, %s", rng->gen,
		static_cast<unsigned long long>(G), draws, maxint ? "next %% maxint" : "next >> (32 - bits)");
	if (draws == 0) return record_event(cq, "clo_rng_fill", err);
	const u32 shift = 32 - bits;
	const u32 g = static_cast<u32>((G + 255) / 256);
	void* st = rng->seeds->ptr;
	u32* o = static_cast<u32*>(out->ptr);
	hipStream_t s = cq->stream;
	const bool init = rng->pending;
	Launch L(cq, init ? "clo_rng_fill_init" : "clo_rng_fill");
	switch (rng->gen) {
	case kLcg:
		launch_fill<clo_rng_lcg::clo_statetype, clo_rng_lcg::clo_rng_step,
			clo_rng_lcg::clo_ulong2statetype>(init, g, s, st, o, G, draws, shift, maxint, rng);
		break;
	case kXs64:
		launch_fill<clo_rng_xorshift64::clo_statetype, clo_rng_xorshift64::clo_rng_step,
			clo_rng_xorshift64::clo_ulong2statetype>(init, g, s, st, o, G, draws, shift, maxint,
			rng);
		break;
	case kXs128:
		launch_fill<clo_rng_xorshift128::clo_statetype, clo_rng_xorshift128::clo_rng_step,
			clo_rng_xorshift128::clo_ulong2statetype>(init, g, s, st, o, G, draws, shift, maxint,
			rng);
		break;
	case kMwc:
		launch_fill<clo_rng_mwc64x::clo_statetype, clo_rng_mwc64x::clo_rng_step,
			clo_rng_mwc64x::clo_ulong2statetype>(init, g, s, st, o, G, draws, shift, maxint, rng);
		break;
	case kPm:
		launch_fill<clo_rng_parkmiller::clo_statetype, clo_rng_parkmiller::clo_rng_step,
			clo_rng_parkmiller::clo_ulong2statetype>(init, g, s, st, o, G, draws, shift, maxint,
			rng);
		break;
	case kTaus:
		launch_fill<clo_rng_tauslcg::clo_statetype, clo_rng_tauslcg::clo_rng_step,
			clo_rng_tauslcg::clo_ulong2statetype>(init, g, s, st, o, G, draws, shift, maxint, rng);
		break;
	}
	if (!L.done(err)) return nullptr;
	if (init) {
		rng->pending = false;
		pending_remove(rng);
	}
	return record_event(cq, "clo_rng_fill", err);
}

}  // extern "C"

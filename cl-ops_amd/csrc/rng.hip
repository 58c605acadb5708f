// RNG objects (CL_Ops' src/cl_ops/rng/clo_rng.c:35-481) and their gfx950
// kernels: device seed initialisation (clo_rng_init.cl:27-60) and the
// register-resident multi-draw fill.

#include <cstring>
#include <string>

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
template <typename S, unsigned int (*STEP)(S&)>
__global__ __launch_bounds__(256) void rng_fill_kernel(S* __restrict__ states, u32* __restrict__ out,
	u64 G, u32 draws, u32 shift, u32 maxint) {
	const u64 gid = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (gid >= G) return;
	S s = states[gid];
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
};

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
			launched = launch_init(gen, buf->ptr, seeds_count, main_seed,
				static_cast<u32>(hid), iq, err);
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
	ccl_buffer_destroy(rng->seeds);
	delete rng;
}

const char* clo_rng_get_source(CloRng* rng) { return rng ? rng->src.c_str() : nullptr; }
CCLBuffer* clo_rng_get_device_seeds(CloRng* rng) { return rng ? rng->seeds : nullptr; }
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
	Launch L(cq, "clo_rng_fill");
	switch (rng->gen) {
	case kLcg:
		rng_fill_kernel<clo_rng_lcg::clo_statetype, clo_rng_lcg::clo_rng_step><<<g, 256, 0, s>>>(
			static_cast<clo_rng_lcg::clo_statetype*>(st), o, G, draws, shift, maxint);
		break;
	case kXs64:
		rng_fill_kernel<clo_rng_xorshift64::clo_statetype, clo_rng_xorshift64::clo_rng_step>
			<<<g, 256, 0, s>>>(static_cast<clo_rng_xorshift64::clo_statetype*>(st), o, G, draws,
			shift, maxint);
		break;
	case kXs128:
		rng_fill_kernel<clo_rng_xorshift128::clo_statetype, clo_rng_xorshift128::clo_rng_step>
			<<<g, 256, 0, s>>>(static_cast<clo_rng_xorshift128::clo_statetype*>(st), o, G, draws,
			shift, maxint);
		break;
	case kMwc:
		rng_fill_kernel<clo_rng_mwc64x::clo_statetype, clo_rng_mwc64x::clo_rng_step>
			<<<g, 256, 0, s>>>(static_cast<clo_rng_mwc64x::clo_statetype*>(st), o, G, draws,
			shift, maxint);
		break;
	case kPm:
		rng_fill_kernel<clo_rng_parkmiller::clo_statetype, clo_rng_parkmiller::clo_rng_step>
			<<<g, 256, 0, s>>>(static_cast<clo_rng_parkmiller::clo_statetype*>(st), o, G, draws,
			shift, maxint);
		break;
	case kTaus:
		rng_fill_kernel<clo_rng_tauslcg::clo_statetype, clo_rng_tauslcg::clo_rng_step>
			<<<g, 256, 0, s>>>(static_cast<clo_rng_tauslcg::clo_statetype*>(st), o, G, draws,
			shift, maxint);
		break;
	}
	if (!L.done(err)) return nullptr;
	return record_event(cq, "clo_rng_fill", err);
}

}  // extern "C"

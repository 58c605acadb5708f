// Sorters built from user compare / key snippets.  CL_Ops compiles every
// sorter from source with the element type, key type, compare and key-getter
// macros prepended (src/cl_ops/sort/clo_sort_abstract.c:140-186); here the
// built-in types and orders use the ahead-of-time kernels, and anything else
// compiles snippet_kernels.rtc with those macros through hipRTC for the
// context's device when the sorter is created.

#include <cstring>
#include <string>

#include "clo_internal.hpp"

using namespace clo;

namespace {

#include "snippet_kernels.inc"  // kSnippetKernels: the text of snippet_kernels.rtc

constexpr unsigned kTile = 2048;  // CLO_SNIP_TILE

// Kernel handles of one compiled snippet program.
// kGroup + 4 * pad + (q - 1): q fused global steps, with or without flags.
enum Fn {
	kGroup = 0, kLocal0 = 8, kLocal1, kPad, kGselect, kRadixKeys, kRadixKeysElem, kNarrow,
	kGather, kNumFns
};
const char* const kFnNames[kNumFns] = {"clo_snip_bitonic_q1_0", "clo_snip_bitonic_q2_0",
	"clo_snip_bitonic_q3_0", "clo_snip_bitonic_q4_0", "clo_snip_bitonic_q1_1",
	"clo_snip_bitonic_q2_1", "clo_snip_bitonic_q3_1", "clo_snip_bitonic_q4_1",
	"clo_snip_bitonic_local0", "clo_snip_bitonic_local1", "clo_snip_pad_flags",
	"clo_snip_gselect", "clo_snip_radix_keys", "clo_snip_radix_keys_elem", "clo_snip_narrow",
	"clo_snip_gather"};

struct SnipSorter {
	hipFunction_t fn[kNumFns];
	std::string source;
	Scratch a, b, c, d;
};

const char* bits_type(size_t size) {
	switch (size) {
	case 1: return "uchar";
	case 2: return "ushort";
	case 4: return "uint";
	case 8: return "ulong";
	}
	return nullptr;
}

bool launch(CCLQueue* q, const char* name, hipFunction_t f, size_t items, void** args,
	GError** err) {
	return rtc_launch(q, name, f, items, args, err);
}

bool copy(CCLQueue* q, void* dst, const void* src, size_t bytes, GError** err) {
	return !hip_failed(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, q->stream), err,
		"hipMemcpyAsync");
}

// The network of sbitonic.cl / abitonic.cl on npow2 slots of `d`.  Stages
// up to log2(tile) run in one LDS launch; each later stage runs its steps
// with pair distance >= tile up to four per launch (in registers) and the
// rest in one LDS launch.
// Profiling names: sbitonic launches all carry the sbitonic kernel name;
// abitonic launches carry the names of the precompiled kernels of the same
// role (tile sort, local steps, 1-4 fused global steps).
bool bitonic_network(const SnipSorter* ss, CCLQueue* q, bool advanced, void* d,
	unsigned char* pad, uint32_t npow2, GError** err) {
	static const char* const gname[5] = {"", CLO_SORT_ABITONIC_KNAME_ANY,
		CLO_SORT_ABITONIC_KNAME_PRIV_2S4V, CLO_SORT_ABITONIC_KNAME_PRIV_3S8V,
		CLO_SORT_ABITONIC_KNAME_PRIV_4S16V};
	const int P = pad ? 1 : 0;
	uint32_t tile = npow2 < kTile ? npow2 : kTile;
	const uint32_t stages = static_cast<uint32_t>(__builtin_ctz(npow2));
	const uint32_t tstages = static_cast<uint32_t>(__builtin_ctz(tile));
	const uint32_t nblocks = npow2 / tile;
	auto local = [&](uint32_t lo, uint32_t hi, uint32_t first) {
		void* args[] = {&d, &pad, &lo, &hi, &first, &tile};
		const size_t items = static_cast<size_t>(nblocks) * 256;  // one block per tile
		const char* name = !advanced ? CLO_SORT_SBITONIC_KNAME
			: lo == 1 ? CLO_SORT_ABITONIC_KNAME_TILESORT : CLO_SORT_ABITONIC_KNAME_LOCAL;
		return launch(q, name, ss->fn[kLocal0 + P], items, args, err);
	};
	if (!local(1, tstages, 1)) return false;
	for (uint32_t stage = tstages + 1; stage <= stages; ++stage) {
		for (uint32_t step = stage; step > tstages;) {
			const uint32_t fused = step - tstages < 4 ? step - tstages : 4;
			uint32_t st = stage, hi = step, items = npow2 >> fused;
			void* args[] = {&d, &pad, &st, &hi, &items};
			const char* name = advanced ? gname[fused] : CLO_SORT_SBITONIC_KNAME;
			if (!launch(q, name, ss->fn[kGroup + 4 * P + fused - 1], items, args, err))
				return false;
			step -= fused;
		}
		if (!local(stage, stage, tstages)) return false;
	}
	return true;
}

}  // namespace

namespace clo {

bool snippet_sorter_needed(const CloSort* s, const char* compare, const char* get_key) {
	const bool radix = strcmp(s->impl.name, "satradix") == 0;
	if (!type_is_integer(s->elem_type) || s->key_type != s->elem_type ||
		!get_key_is_identity(get_key))
		return true;
	// satradix orders by the key alone (satradix.cl never uses the compare
	// macro); the comparison sorts have ahead-of-time ascending and
	// descending kernels.
	return !radix && compare_kind(compare) < 0;
}

bool snippet_sorter_init(CloSort* s, const char* compare, const char* get_key,
	const char* compiler_opts, GError** err) {
	const bool radix = strcmp(s->impl.name, "satradix") == 0;
	if (radix && !type_is_integer(s->key_type)) {
		set_error(err, kQuarkClo, CLO_ERROR_ARGS,
			"satradix needs an integer key type (char..ulong); '%s' has no radix digits.",
			clo_type_get_name(s->key_type) ? clo_type_get_name(s->key_type) : "?");
		return false;
	}
	std::string src;
	src += "#define CLO_SORT_ELEM_TYPE ";
	src += clo_type_get_name(s->elem_type);
	src += "\n#define CLO_SORT_KEY_TYPE ";
	src += clo_type_get_name(s->key_type);
	src += "\n#define CLO_SORT_COMPARE(a, b) ";
	src += compare ? compare : "((a) > (b))";
	src += "\n#define CLO_SORT_KEY_GET(x) ";
	src += get_key ? get_key : "(x)";
	src += "\n#define CLO_SNIP_KEY_BITS ";
	src += bits_type(clo_type_sizeof(s->key_type));
	src += "\n";
	src += kSnippetKernels;
	auto* ss = new SnipSorter();
	if (!rtc_functions(s->ctx, src, compiler_opts, "compare/get_key snippets", kFnNames,
			kNumFns, ss->fn, err)) {
		delete ss;
		return false;
	}
	ss->source = std::move(src);
	s->snip = ss;
	return true;
}

void snippet_sorter_free(CloSort* s) {
	delete static_cast<SnipSorter*>(s->snip);
	s->snip = nullptr;
}

const char* snippet_sorter_source(const CloSort* s) {
	return s->snip ? static_cast<SnipSorter*>(s->snip)->source.c_str() : nullptr;
}

CCLEvent* snippet_sort(CloSort* s, CCLQueue* q, CCLBuffer* kin, CCLBuffer* vin, CCLBuffer* kout,
	CCLBuffer* vout, size_t numel, GError** err) {
	auto* ss = static_cast<SnipSorter*>(s->snip);
	const size_t es = clo_type_sizeof(s->elem_type);
	const char* algo = s->impl.name;
	const std::string done_name = std::string(algo) + "_done";
	if (kout && kout->ptr == kin->ptr) kout = nullptr;
	if (vout && vin && vout->ptr == vin->ptr) vout = nullptr;
	const uint32_t n = static_cast<uint32_t>(numel);
	if (numel <= 1) {
		if (kout && numel == 1) {
			if (!copy(q, kout->ptr, kin->ptr, es, err)) return nullptr;
			if (vout && !copy(q, vout->ptr, vin->ptr, 4, err)) return nullptr;
		}
		return record_event(q, done_name.c_str(), err);
	}
	if (strcmp(algo, "sbitonic") == 0 || strcmp(algo, "abitonic") == 0) {
		const bool advanced = strcmp(algo, "abitonic") == 0;
		void* d = kin->ptr;
		if (kout) {
			// sbitonic.c:84-96: copy, then sort the copy.
			if (!copy(q, kout->ptr, kin->ptr, numel * es, err)) return nullptr;
			d = kout->ptr;
		}
		const uint32_t npow2 = static_cast<uint32_t>(clo_nlpo2(numel));
		if (npow2 == n) {
			if (!bitonic_network(ss, q, advanced, d, nullptr, n, err)) return nullptr;
		} else {
			if (!ss->a.ensure(static_cast<size_t>(npow2) * es, err) ||
				!ss->b.ensure(npow2, err) || !copy(q, ss->a.ptr, d, numel * es, err))
				return nullptr;
			unsigned char* pad = static_cast<unsigned char*>(ss->b.ptr);
			uint32_t nn = n, np = npow2;
			void* args[] = {&pad, &nn, &np};
			if (!launch(q, "bitonic_pad", ss->fn[kPad], npow2, args, err) ||
				!bitonic_network(ss, q, advanced, ss->a.ptr, pad, npow2, err) ||
				!copy(q, d, ss->a.ptr, numel * es, err))
				return nullptr;
		}
	} else if (strcmp(algo, "gselect") == 0) {
		void* ko;
		void* vo = nullptr;
		if (kout) {
			ko = kout->ptr;
			vo = vin ? vout->ptr : nullptr;
		} else {
			// clo_sort_gselect.c:82-100: rank into a temporary, copy back.
			if (!ss->a.ensure(numel * es, err) || (vin && !ss->b.ensure(numel * 4, err)))
				return nullptr;
			ko = ss->a.ptr;
			vo = vin ? ss->b.ptr : nullptr;
		}
		void* ki = kin->ptr;
		void* vi = vin ? vin->ptr : nullptr;
		uint32_t nn = n;
		void* args[] = {&ki, &ko, &vi, &vo, &nn};
		if (!launch(q, CLO_SORT_GSELECT_KNAME, ss->fn[kGselect], numel, args, err)) return nullptr;
		if (!kout) {
			if (!copy(q, kin->ptr, ko, numel * es, err)) return nullptr;
			if (vin && !copy(q, vin->ptr, vo, numel * 4, err)) return nullptr;
		}
	} else {
		const size_t ks = clo_type_sizeof(s->key_type);
		void* ki = kin->ptr;
		uint32_t nn = n;
		if (!vin && es <= 4) {
			// Keys only, elements of at most 4 bytes: sort (KEY_GET bits,
			// element) pairs; the payload is the element.
			if (!ss->a.ensure(numel * ks, err) || !ss->b.ensure(numel * 4, err) ||
				!ss->c.ensure(numel * ks, err))
				return nullptr;
			void* keys = ss->a.ptr;
			void* pay = ss->b.ptr;
			void* kargs[] = {&ki, &keys, &pay, &nn};
			if (!launch(q, "satradix_keys", ss->fn[kRadixKeysElem], numel, kargs, err))
				return nullptr;
			void* dst = kout ? kout->ptr : kin->ptr;  // kin was read into the payload
			if (es == 4) {
				if (!clo_satradix_sort_pairs(s, q, ks, keys, static_cast<uint32_t*>(pay),
						ss->c.ptr, static_cast<uint32_t*>(dst), numel, err))
					return nullptr;
			} else {
				if (!clo_satradix_sort_pairs(s, q, ks, keys, static_cast<uint32_t*>(pay),
						nullptr, nullptr, numel, err))
					return nullptr;
				void* nargs[] = {&pay, &dst, &nn};
				if (!launch(q, "satradix_narrow", ss->fn[kNarrow], numel, nargs, err))
					return nullptr;
			}
		} else {
			// Radix-sort (KEY_GET bits, index) pairs, then gather elements
			// (and payloads) by index.
			if (!ss->a.ensure(numel * ks, err) || !ss->b.ensure(numel * 4, err))
				return nullptr;
			void* keys = ss->a.ptr;
			void* idx = ss->b.ptr;
			void* kargs[] = {&ki, &keys, &idx, &nn};
			if (!launch(q, "satradix_keys", ss->fn[kRadixKeys], numel, kargs, err))
				return nullptr;
			if (!clo_satradix_sort_pairs(s, q, ks, keys, static_cast<uint32_t*>(idx), nullptr,
					nullptr, numel, err))
				return nullptr;
			void* ko;
			void* vo = nullptr;
			if (kout) {
				ko = kout->ptr;
				vo = vin ? vout->ptr : nullptr;
			} else {
				if (!ss->c.ensure(numel * es, err) || (vin && !ss->d.ensure(numel * 4, err)))
					return nullptr;
				ko = ss->c.ptr;
				vo = vin ? ss->d.ptr : nullptr;
			}
			void* vi = vin ? vin->ptr : nullptr;
			void* gargs[] = {&ki, &idx, &ko, &vi, &vo, &nn};
			if (!launch(q, "satradix_gather", ss->fn[kGather], numel, gargs, err))
				return nullptr;
			if (!kout) {
				if (!copy(q, kin->ptr, ko, numel * es, err)) return nullptr;
				if (vin && !copy(q, vin->ptr, vo, numel * 4, err)) return nullptr;
			}
		}
	}
	return record_event(q, done_name.c_str(), err);
}

}  // namespace clo

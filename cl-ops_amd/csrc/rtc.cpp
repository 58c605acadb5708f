// Run-time compilation of kernels built from caller snippets (sort compare /
// key getters, RNG seed hashes) with hipRTC, for the device of a context.
// CL_Ops builds these programs from source on every object creation
// (src/cl_ops/sort/clo_sort_abstract.c:140-186, src/cl_ops/rng/clo_rng.c:
// 101-123); here one compilation per distinct (device, options, source)
// serves every object of the process.

#include <hip/hiprtc.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "clo_internal.hpp"

namespace {

struct RtcModule {
	hipModule_t mod = nullptr;
	std::map<std::string, hipFunction_t> fns;
};

std::mutex g_mu;
// Modules stay loaded until the process ends (hipModuleUnload at exit would
// race the runtime's own teardown).
std::map<std::string, RtcModule*>& cache() {
	static auto* m = new std::map<std::string, RtcModule*>();
	return *m;
}

bool rtc_failed(hiprtcResult r, GError** err, const char* what) {
	if (r == HIPRTC_SUCCESS) return false;
	clo::set_error(err, clo::kQuarkClo, CLO_ERROR_LIBRARY, "%s: %s", what,
		hiprtcGetErrorString(r));
	return true;
}

}  // namespace

namespace clo {

bool rtc_functions(CCLContext* ctx, const std::string& src, const char* compiler_opts,
	const char* what, const char* const* names, int nnames, hipFunction_t* out, GError** err) {
	if (!use_device(ctx, err)) return false;
	hipDeviceProp_t prop;
	if (hip_failed(hipGetDeviceProperties(&prop, ctx->dev), err, "hipGetDeviceProperties"))
		return false;
	std::string arch = prop.gcnArchName;
	arch = arch.substr(0, arch.find(':'));
	std::vector<std::string> opts = {"-O3", "--offload-arch=" + arch,
		// no packed FP32 instructions in any kernel of this library
		"-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
	// Caller options pass through (macro definitions, include paths); the
	// OpenCL C-only "-cl-*" options have no meaning for HIP and are dropped.
	for (const char* p = compiler_opts; p && *p;) {
		while (*p == ' ' || *p == '\t' || *p == '\n') ++p;
		const char* q = p;
		while (*q && *q != ' ' && *q != '\t' && *q != '\n') ++q;
		if (q > p) {
			std::string tok(p, q);
			if (tok.compare(0, 4, "-cl-") != 0) opts.push_back(tok);
		}
		p = q;
	}
	std::string key = std::to_string(ctx->dev) + '\n';
	for (auto& o : opts) key += o + '\n';
	key += src;

	std::lock_guard<std::mutex> lock(g_mu);
	RtcModule* m = nullptr;
	auto it = cache().find(key);
	if (it != cache().end()) {
		m = it->second;
	} else {
		hiprtcProgram prog;
		if (rtc_failed(hiprtcCreateProgram(&prog, src.c_str(), "clo_snippets.hip", 0, nullptr,
				nullptr), err, "hiprtcCreateProgram"))
			return false;
		std::vector<const char*> argv;
		for (auto& o : opts) argv.push_back(o.c_str());
		hiprtcResult rc = hiprtcCompileProgram(prog, static_cast<int>(argv.size()), argv.data());
		if (rc != HIPRTC_SUCCESS) {
			size_t ls = 0;
			std::string log;
			if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
				log.resize(ls);
				hiprtcGetProgramLog(prog, &log[0]);
				log.resize(strlen(log.c_str()));
				if (log.size() > 1500) log = "..." + log.substr(log.size() - 1500);
			}
			hiprtcDestroyProgram(&prog);
			set_error(err, kQuarkClo, CLO_ERROR_ARGS, "The %s did not compile: %s", what,
				log.c_str());
			return false;
		}
		size_t cs = 0;
		std::vector<char> code;
		rc = hiprtcGetCodeSize(prog, &cs);
		if (rc == HIPRTC_SUCCESS) {
			code.resize(cs);
			rc = hiprtcGetCode(prog, code.data());
		}
		hiprtcDestroyProgram(&prog);
		if (rtc_failed(rc, err, "hiprtcGetCode")) return false;
		m = new RtcModule();
		if (hip_failed(hipModuleLoadData(&m->mod, code.data()), err, "hipModuleLoadData")) {
			delete m;
			return false;
		}
		cache()[key] = m;
	}
	for (int i = 0; i < nnames; ++i) {
		auto f = m->fns.find(names[i]);
		if (f == m->fns.end()) {
			hipFunction_t fn;
			if (hip_failed(hipModuleGetFunction(&fn, m->mod, names[i]), err, names[i]))
				return false;
			f = m->fns.emplace(names[i], fn).first;
		}
		out[i] = f->second;
	}
	return true;
}

bool rtc_launch(CCLQueue* q, const char* name, hipFunction_t f, size_t items, void** args,
	GError** err) {
	const size_t grid = div_ceil(items, 256);
	Launch L(q, name);
	if (hip_failed(hipModuleLaunchKernel(f, static_cast<unsigned>(grid), 1, 1, 256, 1, 1, 0,
			q->stream, args, nullptr), err, name))
		return false;
	return L.done(err);
}

}  // namespace clo

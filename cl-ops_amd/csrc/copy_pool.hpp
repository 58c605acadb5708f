// Parallel host memcpy for the host-data paths: a copy of at least 1 MiB
// is split into equal 64-byte-aligned parts, the calling thread copies the
// first and a few resident worker threads the others.  Pure C++ (no HIP), so
// the sanitizer drivers build it on its own.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace clo {

class CopyPool {
public:
	static CopyPool& get() {
		static CopyPool* p = new CopyPool();  // workers live until the process ends
		return *p;
	}
	void copy(void* dst, const void* src, size_t bytes) {
		if (nworkers_ == 0 || bytes < (1u << 20)) {
			memcpy(dst, src, bytes);
			return;
		}
		std::lock_guard<std::mutex> call(call_mu_);
		const size_t parts = static_cast<size_t>(nworkers_) + 1;
		// Shares of ceil(bytes / parts) rounded up to 64 bytes, so that the
		// parts cover every byte.
		const size_t per = ((bytes + parts - 1) / parts + 63) & ~static_cast<size_t>(63);
		char* d = static_cast<char*>(dst);
		const char* s = static_cast<const char*>(src);
		{
			std::lock_guard<std::mutex> lk(mu_);
			for (int i = 0; i < nworkers_; ++i) {
				const size_t off = per * static_cast<size_t>(i + 1);
				parts_[i].d = d + off;
				parts_[i].s = s + off;
				parts_[i].n = off < bytes ? std::min(per, bytes - off) : 0;
			}
			pending_ = nworkers_;
			++gen_;
		}
		cv_.notify_all();
		memcpy(d, s, std::min(per, bytes));
		std::unique_lock<std::mutex> lk(mu_);
		done_.wait(lk, [&] { return pending_ == 0; });
	}

private:
	struct Part {
		char* d;
		const char* s;
		size_t n;
	};
	CopyPool() {
		const unsigned hc = std::thread::hardware_concurrency();
		nworkers_ = hc >= 16 ? 7 : (hc >= 8 ? 3 : (hc >= 4 ? 1 : 0));
		parts_.resize(static_cast<size_t>(nworkers_));
		for (int i = 0; i < nworkers_; ++i) std::thread(&CopyPool::worker, this, i).detach();
	}
	void worker(int i) {
		uint64_t seen = 0;
		std::unique_lock<std::mutex> lk(mu_);
		for (;;) {
			cv_.wait(lk, [&] { return gen_ != seen; });
			seen = gen_;
			const Part p = parts_[static_cast<size_t>(i)];
			lk.unlock();
			if (p.n) memcpy(p.d, p.s, p.n);
			lk.lock();
			if (--pending_ == 0) done_.notify_one();
		}
	}
	std::mutex call_mu_, mu_;
	std::condition_variable cv_, done_;
	std::vector<Part> parts_;
	int nworkers_ = 0;
	int pending_ = 0;
	uint64_t gen_ = 0;
};

}  // namespace clo

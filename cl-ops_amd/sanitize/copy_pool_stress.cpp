// Stress driver for the host copy pool (csrc/copy_pool.hpp) under
// ThreadSanitizer and AddressSanitizer (built by `make -C cl-ops_amd
// sanitize`; CPU only).  Several caller threads copy buffers at once, most of
// them large enough to be split across the pool's workers; every byte of
// every copy is checked, and a guard region after each destination must
// stay untouched.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../csrc/copy_pool.hpp"

int main() {
	constexpr int kCallers = 4, kRounds = 24, kGuard = 64;
	std::atomic<int> bad{0};
	std::vector<std::thread> callers;
	for (int c = 0; c < kCallers; ++c)
		callers.emplace_back([&bad, c] {
			std::mt19937_64 rng(0xC0B1 + c);
			for (int r = 0; r < kRounds; ++r) {
				// >= 1 MiB is split across the workers; every fifth copy is
				// below that and takes the plain memcpy
				size_t n = (1u << 20) + rng() % (3u << 20);
				if (r % 5 == 0) n = rng() % 4096;
				std::vector<unsigned char> src(n), dst(n + kGuard, 0xA5);
				for (size_t i = 0; i < n; ++i) src[i] = static_cast<unsigned char>(i * 131 + c + r);
				clo::CopyPool::get().copy(dst.data(), src.data(), n);
				if (n && std::memcmp(dst.data(), src.data(), n) != 0) ++bad;
				for (size_t i = n; i < n + kGuard; ++i)
					if (dst[i] != 0xA5) {
						++bad;
						break;
					}
			}
		});
	for (auto& t : callers) t.join();
	std::printf("copy pool stress: %d bad copies\n", bad.load());
	return bad.load() ? 1 : 0;
}

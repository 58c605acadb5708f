// Block-level stable digit ranking for 256-thread blocks (wave64 ballots),
// shared by the segmented sort kernels.
#pragma once

#include "dev_common.hpp"

namespace clo {
namespace dev {

constexpr int kRB = 256;            // threads per block == digit bins
constexpr int kRBWaves = kRB / kWave;

struct RankSmem {
	u32 whist[kRBWaves][256];
	u32 tstart[256];
	u32 total[256];
	u32 wsum[kRBWaves];
};

// Stable block-wide rank of IPT items per thread held in wave-striped order
// (item j of lane l in wave w is element w*64*IPT + j*64 + l).  `dig` are
// digits below 2^bits; invalid items do not participate.  On return pos[j] is
// the item's position when the block's valid items are ordered by digit,
// ties in element order; sm.tstart[d] / sm.total[d] hold the block's digit
// starts and counts.  Contains barriers: call from uniform control flow.
// FULL: every item of the block is valid (`valid` is not read).
template <int IPT, bool FULL = false>
__device__ __forceinline__ void block_rank(const u32 (&dig)[IPT], const bool (&valid)[IPT],
	u32 (&pos)[IPT], u32 bits, RankSmem& sm) {
	const u32 tid = threadIdx.x;
	const u32 lane = lane_id();
	const u32 w = wave_id();
	for (u32 i = lane; i < 256; i += kWave) sm.whist[w][i] = 0;
	__syncthreads();
#pragma unroll
	for (int j = 0; j < IPT; ++j) {
		const u32 d = dig[j];
		const u64 peers = FULL ? match_peers_full<8>(d, bits) : match_peers<8>(d, bits, valid[j]);
		const u32 below = count_below(peers);
		u32 base = 0;
		if (FULL || valid[j]) {
			base = sm.whist[w][d];
			if (below == 0) sm.whist[w][d] = base + static_cast<u32>(__popcll(peers));
		}
		pos[j] = base + below;
	}
	__syncthreads();
	u32 run = 0;
#pragma unroll
	for (int ww = 0; ww < kRBWaves; ++ww) {
		const u32 c = sm.whist[ww][tid];
		sm.whist[ww][tid] = run;
		run += c;
	}
	sm.total[tid] = run;
	const u32 ts = block_exclusive_sum<kRB>(run, sm.wsum, static_cast<u32*>(nullptr));
	sm.tstart[tid] = ts;
	__syncthreads();
#pragma unroll
	for (int j = 0; j < IPT; ++j)
		if (FULL || valid[j]) pos[j] += sm.tstart[dig[j]] + sm.whist[w][dig[j]];
}

}  // namespace dev
}  // namespace clo

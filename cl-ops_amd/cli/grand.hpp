// GLib's GRand as the reference's benchmarks draw their inputs with it
// (src/benchmarks/clo_sort_bench.c:145,190-193 and clo_scan_bench.c:163,
// 219-224: one generator seeded with the bench seed, drawn in order across
// sizes and runs), restated from GLib's published algorithm (GLib 2.72,
// glib/grand.c, default random version 2.2):
//   - MT19937 (Matsumoto & Nishimura 1998), seeded with init_genrand:
//     mt[0] = seed, mt[i] = 1812433253 * (mt[i-1] ^ (mt[i-1] >> 30)) + i;
//   - g_rand_double: (w1 * 2^-32 + w2) * 2^-32 from two words, drawn again if
//     rounding reaches 1.0;
//   - g_rand_boolean: bit 15 of one word;
//   - g_rand_int_range(begin, end): dist = end - begin (unsigned), words above
//     the largest multiple of dist below 2^32 are rejected, then word % dist;
//   - g_rand_double_range(begin, end): r * end - (r - 1) * begin.
// clo_bench_rand (src/benchmarks/clo_bench.c:67-142) maps the draws to each
// element type.  tests/golden/glib_grand.json pins every function here
// against the system's libglib.

#pragma once

#include <cfloat>
#include <climits>
#include <cstdint>
#include <cstring>

#include "cl_ops.h"

namespace cli {

class GRand {
public:
	explicit GRand(uint32_t seed) {
		mt_[0] = seed;
		for (int i = 1; i < kN; ++i)
			mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + static_cast<uint32_t>(i);
		mti_ = kN;
	}

	uint32_t next_u32() {
		if (mti_ >= kN) refill();
		uint32_t y = mt_[mti_++];
		y ^= y >> 11;
		y ^= (y << 7) & 0x9D2C5680u;
		y ^= (y << 15) & 0xEFC60000u;
		y ^= y >> 18;
		return y;
	}

	double next_double() {
		const double kTransform = 2.3283064365386962890625e-10;  // 2^-32
		for (;;) {
			double r = next_u32() * kTransform;
			r = (r + next_u32()) * kTransform;
			if (r < 1.0) return r;
		}
	}

	bool next_boolean() { return (next_u32() & (1u << 15)) != 0; }

	int32_t int_range(int32_t begin, int32_t end) {
		const uint32_t dist = static_cast<uint32_t>(end) - static_cast<uint32_t>(begin);
		uint32_t r = 0;
		if (dist != 0) {
			uint32_t maxvalue;
			if (dist <= 0x80000000u) {
				uint32_t leftover = (0x80000000u % dist) * 2;
				if (leftover >= dist) leftover -= dist;
				maxvalue = 0xFFFFFFFFu - leftover;
			} else {
				maxvalue = dist - 1;
			}
			do r = next_u32();
			while (r > maxvalue);
			r %= dist;
		}
		return static_cast<int32_t>(static_cast<uint32_t>(begin) + r);
	}

	double double_range(double begin, double end) {
		const double r = next_double();
		return r * end - (r - 1) * begin;
	}

private:
	static constexpr int kN = 624, kM = 397;

	void refill() {
		for (int i = 0; i < kN; ++i) {
			const uint32_t y = (mt_[i] & 0x80000000u) | (mt_[(i + 1) % kN] & 0x7FFFFFFFu);
			mt_[i] = mt_[(i + kM) % kN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908B0DFu : 0u);
		}
		mti_ = 0;
	}

	uint32_t mt_[kN];
	int mti_;
};

// One element of type t from the generator, as clo_bench_rand writes it
// (clo_bench.c:67-142).  Where the reference's expression leaves the order
// of its two draws unspecified (long: the double and the sign boolean), the
// double is drawn first.  half takes g_rand_double_range(-1024, 1024)
// converted to an unsigned 16-bit pattern by truncation toward zero and
// wrap-around (the reference's cl_half is cl_ushort, and the conversion of a
// negative double to it is undefined in C).  Returns false for an unknown
// type.
inline bool bench_rand(GRand& g, CloType t, void* dst) {
	switch (t) {
	case CLO_CHAR: {
		const int8_t v = static_cast<int8_t>(g.int_range(SCHAR_MIN, SCHAR_MAX));
		memcpy(dst, &v, 1);
		return true;
	}
	case CLO_UCHAR: {
		const uint8_t v = static_cast<uint8_t>(g.int_range(0, UCHAR_MAX));
		memcpy(dst, &v, 1);
		return true;
	}
	case CLO_SHORT: {
		const int16_t v = static_cast<int16_t>(g.int_range(SHRT_MIN, SHRT_MAX));
		memcpy(dst, &v, 2);
		return true;
	}
	case CLO_USHORT: {
		const uint16_t v = static_cast<uint16_t>(g.int_range(0, USHRT_MAX));
		memcpy(dst, &v, 2);
		return true;
	}
	case CLO_INT: {
		const int32_t v = g.int_range(INT32_MIN, INT32_MAX);
		memcpy(dst, &v, 4);
		return true;
	}
	case CLO_UINT: {
		const uint32_t v = static_cast<uint32_t>(g.next_double() * 4294967295.0);
		memcpy(dst, &v, 4);
		return true;
	}
	case CLO_LONG: {
		const double r = g.next_double();
		const double m = g.next_boolean() ? -9223372036854775808.0 : 9223372036854775807.0;
		const int64_t v = static_cast<int64_t>(r * m);
		memcpy(dst, &v, 8);
		return true;
	}
	case CLO_ULONG: {
		const uint64_t v = static_cast<uint64_t>(g.next_double() * 18446744073709551615.0);
		memcpy(dst, &v, 8);
		return true;
	}
	case CLO_FLOAT: {
		const float v = static_cast<float>(g.double_range(FLT_MIN, FLT_MAX));
		memcpy(dst, &v, 4);
		return true;
	}
	case CLO_DOUBLE: {
		const double v = g.double_range(DBL_MIN, DBL_MAX);
		memcpy(dst, &v, 8);
		return true;
	}
	case CLO_HALF: {
		const double d = g.double_range(-1024.0, 1024.0);
		const uint16_t v = static_cast<uint16_t>(static_cast<int32_t>(d));
		memcpy(dst, &v, 2);
		return true;
	}
	default:
		return false;
	}
}

}  // namespace cli

/*
 * The random stream of smina's search: boost::mt19937 (typedef rng,
 * src/lib/random.h:33) and the two Boost distributions the search draws from
 * (src/lib/random.cpp:28-50):
 *   random_fl(a, b)  = boost::uniform_real<double>: u32 / 2^32 * (b - a) + a,
 *                      drawn again while the result is not below b;
 *   random_int(a, b) = boost::uniform_int<int>: bucket rejection over the
 *                      2^32 - 1 wide engine range.
 * The reference restates the same published algorithms for its device code
 * at src/gpu/gpu_rng.cuh:28-92.
 *
 * The state is 624 words plus a position.  Host code seeds it (standard MT
 * seeding, as boost::mt19937(seed)) and draws the chain's initial
 * configuration; the device then continues the same stream.
 */
#ifndef SGPU_RNG_H
#define SGPU_RNG_H

#include <stdint.h>
#include "pmath.h"

namespace sgpu
{

constexpr int MT_N = 624;
constexpr int MT_M = 397;

// Word layout of a chain's generator in memory: x[0..623], then the position.
struct MTState
{
	uint32_t x[MT_N];
	uint32_t i;
};

SGPU_HD void mt_seed(MTState& s, uint32_t seed)
{
	s.x[0] = seed;
	for (int k = 1; k < MT_N; k++)
		s.x[k] = 1812433253u * (s.x[k - 1] ^ (s.x[k - 1] >> 30)) + uint32_t(k);
	s.i = MT_N;
}

// one regeneration of the 624 words; word k of the new state depends on
// old words k, k+1 and (k+397 mod 624), which is new for k >= 227
SGPU_HD uint32_t mt_mix(uint32_t xk, uint32_t xk1, uint32_t xkm)
{
	const uint32_t y = (xk & 0x80000000u) | (xk1 & 0x7fffffffu);
	return xkm ^ (y >> 1) ^ ((xk1 & 1u) ? 0x9908b0dfu : 0u);
}

SGPU_HD void mt_regenerate(MTState& s)
{
	for (int k = 0; k < MT_N; k++)
		s.x[k] = mt_mix(s.x[k], s.x[(k + 1) % MT_N], s.x[(k + MT_M) % MT_N]);
	s.i = 0;
}

SGPU_HD uint32_t mt_temper(uint32_t z)
{
	z ^= z >> 11;
	z ^= (z << 7) & 0x9d2c5680u;
	z ^= (z << 15) & 0xefc60000u;
	z ^= z >> 18;
	return z;
}

SGPU_HD uint32_t mt_next(MTState& s)
{
	if (s.i >= MT_N)
		mt_regenerate(s);
	return mt_temper(s.x[s.i++]);
}

// boost::uniform_real<double>(a, b) on a 32-bit integral engine
SGPU_HD double uniform_from_word(uint32_t w, double a, double b)
{
	const double numerator = double(w);
	const double divisor = 4294967295.0 + 1.0;
	return numerator / divisor * (b - a) + a;
}

// boost::uniform_int<int>(a, b): bucket size and acceptance for one draw
struct IntRange
{
	uint32_t range;
	uint32_t bucket;
};
SGPU_HD IntRange int_range(int a, int b)
{
	IntRange r;
	r.range = uint32_t(b) - uint32_t(a);
	const uint32_t brange = 0xffffffffu;
	if (r.range == 0)
	{
		r.bucket = 1;
		return r;
	}
	r.bucket = brange / (r.range + 1u);
	if (brange % (r.range + 1u) == r.range)
		++r.bucket;
	return r;
}

template<typename Gen>
SGPU_HD double random_fl(double a, double b, Gen& g)
{
	for (;;)
	{
		const double r = uniform_from_word(g.next(), a, b);
		if (r < b)
			return r;
	}
}

template<typename Gen>
SGPU_HD int random_int(int a, int b, Gen& g)
{
	const IntRange r = int_range(a, b);
	if (r.range == 0)
		return a;
	for (;;)
	{
		const uint32_t v = g.next() / r.bucket;
		if (v <= r.range)
			return int(v + uint32_t(a));
	}
}

// a generator over an MTState held by the caller
struct MTRef
{
	MTState* s;
	SGPU_HD uint32_t next() { return mt_next(*s); }
};

} // namespace sgpu

#endif

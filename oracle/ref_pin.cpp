// Test infrastructure (not part of the product): pins this build's portable
// math and random stream to the reference's own headers, compiled from where
// they lie in the reference tree by oracle/Makefile's `ref` target into
// oracle/_ref/.  Only tests/test_ref_pin.py loads the result.
//
//   src/lib/portable_math.h  portable_math::{sin, cos, exp, acos}, the
//                            functions smina's SMINA_PORTABLE_MATH build
//                            and its double-precision GPU kernels share
//   src/gpu/gpu_rng.cuh      gpusmina::{mt_next, random_fl, random_int},
//                            the reference's device restatement of
//                            boost::mt19937 and Boost's uniform draws
//
// Each comparison runs here, in C++, over many arguments and returns the
// number of results whose bits differ, with the first such argument.
#include <cstdint>
#include <cstring>

#include "lib/portable_math.h"
#include "gpu/gpu_rng.cuh"

#include "../gpu-smina_amd/csrc/pmath.h"
#include "../gpu-smina_amd/csrc/rng.h"

namespace
{

// splitmix64: a fixed, seedable stream of 64-bit words for the arguments
struct Words
{
	uint64_t s;
	uint64_t next()
	{
		uint64_t z = (s += 0x9e3779b97f4a7c15ull);
		z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
		z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
		return z ^ (z >> 31);
	}
	// uniform in [lo, hi) with 53 random bits
	double in(double lo, double hi) { return lo + (hi - lo) * (double(next() >> 11) * 0x1.0p-53); }
};

uint64_t bits_of(double x)
{
	uint64_t b;
	std::memcpy(&b, &x, sizeof b);
	return b;
}

// the reference generator seeded as boost::mt19937(seed) seeds itself
void seed_ref(gpusmina::MTState& s, uint32_t seed)
{
	s.x[0] = seed;
	for (int k = 1; k < gpusmina::MT_N; k++)
		s.x[k] = 1812433253u * (s.x[k - 1] ^ (s.x[k - 1] >> 30)) + uint32_t(k);
	s.i = gpusmina::MT_N;
}

} // namespace

extern "C" {

// fn: 0 sin, 1 cos, 2 exp, 3 acos; n arguments uniform in [lo, hi) from the
// stream `seed`; returns the count of differing results and the first
// differing argument in *first (untouched if none)
long long pin_math(int fn, double lo, double hi, long long n, uint64_t seed, double* first)
{
	Words w{seed};
	long long bad = 0;
	for (long long i = 0; i < n; i++)
	{
		const double x = w.in(lo, hi);
		double a = 0, b = 0;
		switch (fn)
		{
		case 0:
			a = portable_math::sin(x);
			b = pmath::sin(x);
			break;
		case 1:
			a = portable_math::cos(x);
			b = pmath::cos(x);
			break;
		case 2:
			a = portable_math::exp(x);
			b = pmath::exp(x);
			break;
		default:
			a = portable_math::acos(x);
			b = pmath::acos(x);
			break;
		}
		if (bits_of(a) != bits_of(b))
		{
			if (bad == 0 && first)
				*first = x;
			bad++;
		}
	}
	return bad;
}

// the same on caller-given arguments (exact special points)
long long pin_math_at(int fn, const double* x, long long n, double* ref_out, double* own_out)
{
	long long bad = 0;
	for (long long i = 0; i < n; i++)
	{
		double a, b;
		switch (fn)
		{
		case 0:
			a = portable_math::sin(x[i]);
			b = pmath::sin(x[i]);
			break;
		case 1:
			a = portable_math::cos(x[i]);
			b = pmath::cos(x[i]);
			break;
		case 2:
			a = portable_math::exp(x[i]);
			b = pmath::exp(x[i]);
			break;
		default:
			a = portable_math::acos(x[i]);
			b = pmath::acos(x[i]);
			break;
		}
		ref_out[i] = a;
		own_out[i] = b;
		bad += bits_of(a) != bits_of(b) && !(a != a && b != b);
	}
	return bad;
}

// kind: 0 raw words, 1 random_fl(a, b), 2 random_int(ia, ib); n draws from
// generators seeded with `seed`; returns the count of differing draws
long long pin_rng(int kind, uint32_t seed, long long n, double a, double b, int ia, int ib)
{
	gpusmina::MTState r;
	seed_ref(r, seed);
	sgpu::MTState m;
	sgpu::mt_seed(m, seed);
	sgpu::MTRef g{&m};
	long long bad = 0;
	for (long long i = 0; i < n; i++)
	{
		switch (kind)
		{
		case 0:
			bad += gpusmina::mt_next(r) != sgpu::mt_next(m);
			break;
		case 1:
			bad += bits_of(gpusmina::random_fl(a, b, r)) != bits_of(sgpu::random_fl(a, b, g));
			break;
		default:
			bad += gpusmina::random_int(ia, ib, r) != sgpu::random_int(ia, ib, g);
			break;
		}
	}
	return bad;
}

} // extern "C"

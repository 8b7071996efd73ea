/*
 * Portable double-precision sin, cos, exp and acos, compiled identically for
 * the host (g++) and for gfx950 (hipcc), so that the FP64 docking search gives
 * the same bits on the GPU as a CPU smina build with SMINA_PORTABLE_MATH.
 *
 * Method: Sun fdlibm 5.3 (freely distributable): the minimax kernels of
 * k_sin.c and k_cos.c, e_exp.c and e_acos.c with their published
 * coefficients.  The variant is the one smina's portable build uses
 * (src/lib/portable_math.h:89-264), which departs from stock fdlibm:
 *   - the kernels take no tail word (the reduced argument is one double);
 *   - cos on 0.3 <= |x| <= 0.78125 splits off qx = (x*x)/4 with its low
 *     word cleared (src/lib/portable_math.h:112-128);
 *   - reduction by pi/2 is one Cody-Waite step with a 33-bit head and its
 *     tail, the quotient rounded to nearest (:131-141), with no second or
 *     third refinement;
 *   - exp reduces every |x| > ln2/2 by k = (int)(x/ln2 +- 1/2) and scales
 *     by 2^k through multiplications (:62-86, :232-264); acos has no tiny-x
 *     shortcut (:191-230).
 * These choices fix the result bits; changing any of them changes sin or cos
 * by one ulp on a sizeable share of arguments.
 *
 * Domain: the search calls sin and cos with |x| <= pi; the reduction is good
 * to full accuracy for |x| up to about 2^20 (its 33-bit head times the
 * quotient is exact).
 *
 * Every expression is written so that no fused multiply-add may change it:
 * build with -ffp-contract=off where the bits matter.
 */
#ifndef SGPU_PMATH_H
#define SGPU_PMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define SGPU_HD __host__ __device__ inline
#else
#define SGPU_HD inline
#endif

namespace pmath
{

SGPU_HD uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
SGPU_HD double from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }
SGPU_HD double magnitude(double x) { return from_bits(bits(x) & 0x7fffffffffffffffull); }
// x with its low 32 mantissa bits cleared
SGPU_HD double high_part(double x) { return from_bits(bits(x) & 0xffffffff00000000ull); }
// 2^e for a normal exponent e
SGPU_HD double pow2(int e) { return from_bits(uint64_t(e + 1023) << 52); }

// x * 2^k by at most three exact power-of-two factors (results that are
// subnormal round once, in the last factor)
SGPU_HD double ldexp_mul(double x, int k)
{
	if (k > 1023)
	{
		x = x * pow2(1023);
		k -= 1023;
	}
	else if (k < -1022)
	{
		x = x * pow2(-1000);
		k += 1000;
		if (k < -1022)
		{
			x = x * pow2(-1000);
			k += 1000;
		}
	}
	return x * pow2(k);
}

// nearest integer, ties to even, by the 1.5 * 2^52 shift (|x| < 2^51)
SGPU_HD double round_even(double x)
{
	const double shift = 6755399441055744.0;
	return x >= 0 ? (x + shift) - shift : -((shift - x) - shift);
}

// sin on [-pi/4, pi/4]: x + x^3 (S1 + x^2 P(x^2))
SGPU_HD double sin_kernel(double x)
{
	const double s1 = -1.66666666666666324348e-01, s2 = 8.33333333332248946124e-03,
			s3 = -1.98412698298579493134e-04, s4 = 2.75573137070700676789e-06,
			s5 = -2.50507602534068634195e-08, s6 = 1.58969099521155010221e-10;
	const double z = x * x;
	const double x3 = z * x;
	const double poly = s2 + z * (s3 + z * (s4 + z * (s5 + z * s6)));
	return x + x3 * (s1 + z * poly);
}

// cos on [-pi/4, pi/4]: 1 - x^2/2 + x^4 Q(x^2); above |x| = 0.3 the 1 - x^2/2
// part is split at the dyadic qx so that 1 - qx is exact
SGPU_HD double cos_kernel(double x)
{
	const double c1 = 4.16666666666666019037e-02, c2 = -1.38888888888741095749e-03,
			c3 = 2.48015872894767294178e-05, c4 = -2.75573143513906633035e-07,
			c5 = 2.08757232129817482790e-09, c6 = -1.13596475577881948265e-11;
	const double z = x * x;
	const double zq = z * (c1 + z * (c2 + z * (c3 + z * (c4 + z * (c5 + z * c6)))));
	const double ax = magnitude(x);
	if (ax < 0.3)
		return 1.0 - (0.5 * z - z * zq);
	const double qx = ax > 0.78125 ? 0.28125 : high_part(z * 0.25);
	const double half_z = 0.5 * z - qx;
	const double one_minus_q = 1.0 - qx;
	return one_minus_q - (half_z - z * zq);
}

// x = q pi/2 + r: returns r, and q through *quadrant
SGPU_HD double reduce_half_pi(double x, int* quadrant)
{
	const double two_over_pi = 6.36619772367581382433e-01;
	const double half_pi_head = 1.57079632673412561417e+00; // 33 leading bits
	const double half_pi_tail = 6.07710050650619224932e-11; // pi/2 - head
	const double q = round_even(x * two_over_pi);
	*quadrant = int(q);
	return (x - q * half_pi_head) - q * half_pi_tail;
}

SGPU_HD double sin(double x)
{
	if (magnitude(x) <= 7.85398163397448278999e-01)
		return sin_kernel(x);
	int q = 0;
	const double r = reduce_half_pi(x, &q);
	const int quad = q & 3;
	const double v = (quad & 1) ? cos_kernel(r) : sin_kernel(r);
	return (quad & 2) ? -v : v;
}

SGPU_HD double cos(double x)
{
	if (magnitude(x) <= 7.85398163397448278999e-01)
		return cos_kernel(x);
	int q = 0;
	const double r = reduce_half_pi(x, &q);
	const int quad = q & 3;
	const double v = (quad & 1) ? sin_kernel(r) : cos_kernel(r);
	return (quad == 1 || quad == 2) ? -v : v;
}

// e^x = 2^k e^r, r = x - k ln2 in two parts; e^r by the Remez rational form
SGPU_HD double exp(double x)
{
	const double ln2_head = 6.93147180369123816490e-01, ln2_tail = 1.90821492927058770002e-10;
	const double inv_ln2 = 1.44269504088896338700e+00;
	const double p1 = 1.66666666666666019037e-01, p2 = -2.77777777770155933842e-03,
			p3 = 6.61375632143793436117e-05, p4 = -1.65339022054652515390e-06,
			p5 = 4.13813679705723846039e-08;
	if (!(x == x))
		return x;
	if (x > 7.09782712893383973096e+02)
		return from_bits(0x7ff0000000000000ull);
	if (x < -7.45133219101941108420e+02)
		return 0.0;
	const double ax = magnitude(x);
	int k = 0;
	double head = x, tail = 0.0;
	if (ax > 0.5 * 6.93147180559945286227e-01)
	{
		k = int(inv_ln2 * x + (x < 0 ? -0.5 : 0.5));
		const double kd = double(k);
		head = x - kd * ln2_head;
		tail = kd * ln2_tail;
		x = head - tail;
	}
	else if (ax < 3.7252902984e-09)
	{
		return 1.0 + x;
	}
	const double t = x * x;
	const double c = x - t * (p1 + t * (p2 + t * (p3 + t * (p4 + t * p5))));
	if (k == 0)
		return 1.0 - ((x * c) / (c - 2.0) - x);
	const double y = 1.0 - ((tail - (x * c) / (2.0 - c)) - head);
	return ldexp_mul(y, k);
}

// acos: |x| < 1/2 as pi/2 - asin(x) with asin(x) = x + x R(x^2); otherwise
// through z = (1 -+ x)/2 and s = sqrt(z), with s split into its high part
// and a correction for x > 1/2
SGPU_HD double acos(double x)
{
	const double pi = 3.14159265358979311600e+00;
	const double half_pi_hi = 1.57079632679489655800e+00, half_pi_lo = 6.12323399573676603587e-17;
	const double a0 = 1.66666666666666657415e-01, a1 = -3.25565818622400915405e-01,
			a2 = 2.01212532134862925881e-01, a3 = -4.00555345006794114027e-02,
			a4 = 7.91534994289814532176e-04, a5 = 3.47933107596021167570e-05;
	const double b1 = -2.40339491173441421878e+00, b2 = 2.02094576023350569471e+00,
			b3 = -6.88283971605453293030e-01, b4 = 7.70381505559019352791e-02;
	const double ax = magnitude(x);
	if (ax >= 1.0)
	{
		if (x == 1.0)
			return 0.0;
		if (x == -1.0)
			return pi;
		return (x - x) / (x - x);
	}
	if (ax < 0.5)
	{
		const double z = x * x;
		const double num = z * (a0 + z * (a1 + z * (a2 + z * (a3 + z * (a4 + z * a5)))));
		const double den = 1.0 + z * (b1 + z * (b2 + z * (b3 + z * b4)));
		return half_pi_hi - (x - (half_pi_lo - x * (num / den)));
	}
	const double z = x < 0 ? (1.0 + x) * 0.5 : (1.0 - x) * 0.5;
	const double num = z * (a0 + z * (a1 + z * (a2 + z * (a3 + z * (a4 + z * a5)))));
	const double den = 1.0 + z * (b1 + z * (b2 + z * (b3 + z * b4)));
	const double s = __builtin_sqrt(z);
	if (x < 0)
	{
		const double w = (num / den) * s - half_pi_lo;
		return pi - 2.0 * (s + w);
	}
	const double s_hi = high_part(s);
	const double corr = (z - s_hi * s_hi) / (s + s_hi);
	const double w = (num / den) * s + corr;
	return 2.0 * (s_hi + w);
}

} // namespace pmath

#endif

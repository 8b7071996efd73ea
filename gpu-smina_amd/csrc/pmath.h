/*
 * Portable double-precision sin, cos, exp and acos, compiled identically for
 * the host (g++) and for gfx950 (hipcc), so that the FP64 docking search gives
 * the same bits on the GPU as the CPU restatement of smina's search.
 *
 * Method: Sun fdlibm 5.3 (freely distributable; the published algorithms of
 * k_sin.c, k_cos.c, e_rem_pio2.c (medium-argument Cody-Waite path), e_exp.c
 * and e_acos.c with their published minimax coefficients).  The reference
 * switches smina's search to such functions for its bit-exact validation
 * build (src/lib/portable_math.h:1-34, used by src/lib/monte_carlo.cpp:28-35
 * and src/lib/quaternion.cpp:25-32).
 *
 * Domain: sin/cos are exact to fdlibm accuracy for |x| < 2^19 * pi/2 (the
 * search only calls them with |x| <= pi); larger arguments are reduced with
 * the same three-stage Cody-Waite scheme, which loses accuracy there.
 *
 * Every expression is written so that no fused multiply-add may change it:
 * build with -ffp-contract=off.
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
SGPU_HD int32_t hi_word(double x) { return int32_t(bits(x) >> 32); }
SGPU_HD uint32_t lo_word(double x) { return uint32_t(bits(x)); }
SGPU_HD double make(int32_t hi, uint32_t lo)
{
	return from_bits((uint64_t(uint32_t(hi)) << 32) | uint64_t(lo));
}
SGPU_HD double with_hi(double x, int32_t hi) { return make(hi, lo_word(x)); }

// k_sin.c: sin on [-pi/4, pi/4]; y is the tail of x, iy = 0 means y = 0
SGPU_HD double kernel_sin(double x, double y, int iy)
{
	const double S1 = -1.66666666666666324348e-01;
	const double S2 = 8.33333333332248946124e-03;
	const double S3 = -1.98412698298579493134e-04;
	const double S4 = 2.75573137070700676789e-06;
	const double S5 = -2.50507602534068634195e-08;
	const double S6 = 1.58969099521155010221e-10;
	const int32_t ix = hi_word(x) & 0x7fffffff;
	if (ix < 0x3e400000) // |x| < 2^-27
		if (int(x) == 0)
			return x;
	const double z = x * x;
	const double v = z * x;
	const double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
	if (iy == 0)
		return x + v * (S1 + z * r);
	return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

// k_cos.c: cos on [-pi/4, pi/4]
SGPU_HD double kernel_cos(double x, double y)
{
	const double C1 = 4.16666666666666019037e-02;
	const double C2 = -1.38888888888741095749e-03;
	const double C3 = 2.48015872894767294178e-05;
	const double C4 = -2.75573143513906633035e-07;
	const double C5 = 2.08757232129817482790e-09;
	const double C6 = -1.13596475577881948265e-11;
	const int32_t ix = hi_word(x) & 0x7fffffff;
	if (ix < 0x3e400000) // |x| < 2^-27
		if (int(x) == 0)
			return 1.0;
	const double z = x * x;
	const double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
	if (ix < 0x3FD33333) // |x| < 0.3
		return 1.0 - (0.5 * z - (z * r - x * y));
	double qx;
	if (ix > 0x3fe90000) // |x| > 0.78125
		qx = 0.28125;
	else
		qx = make(ix - 0x00200000, 0); // x/4
	const double hz = 0.5 * z - qx;
	const double a = 1.0 - qx;
	return a - (hz - (z * r - x * y));
}

// e_rem_pio2.c, medium path: x = n*pi/2 + (y0 + y1), |y0 + y1| <= pi/4
SGPU_HD int rem_pio2(double x, double& y0, double& y1)
{
	const double invpio2 = 6.36619772367581382433e-01;
	const double pio2_1 = 1.57079632673412561417e+00;
	const double pio2_1t = 6.07710050650619224932e-11;
	const double pio2_2 = 6.07710050630396597660e-11;
	const double pio2_2t = 2.02226624879595063154e-21;
	const double pio2_3 = 2.02226624871116645580e-21;
	const double pio2_3t = 8.47842766036889956997e-32;
	const int32_t hx = hi_word(x);
	const int32_t ix = hx & 0x7fffffff;
	const double t = x < 0 ? -x : x;
	const int n = int(t * invpio2 + 0.5);
	const double fn = double(n);
	double r = t - fn * pio2_1;
	double w = fn * pio2_1t; // first round, good to 85 bits
	const int32_t j = ix >> 20;
	y0 = r - w;
	int32_t i = j - ((hi_word(y0) >> 20) & 0x7ff);
	if (i > 16)
	{ // second iteration, good to 118 bits
		double tt = r;
		w = fn * pio2_2;
		r = tt - w;
		w = fn * pio2_2t - ((tt - r) - w);
		y0 = r - w;
		i = j - ((hi_word(y0) >> 20) & 0x7ff);
		if (i > 49)
		{ // third iteration, 151 bits
			tt = r;
			w = fn * pio2_3;
			r = tt - w;
			w = fn * pio2_3t - ((tt - r) - w);
			y0 = r - w;
		}
	}
	y1 = (r - y0) - w;
	if (hx < 0)
	{
		y0 = -y0;
		y1 = -y1;
		return -n;
	}
	return n;
}

SGPU_HD double sin(double x)
{
	const int32_t ix = hi_word(x) & 0x7fffffff;
	if (ix <= 0x3fe921fb)
		return kernel_sin(x, 0.0, 0);
	if (ix >= 0x7ff00000)
		return x - x; // NaN
	double y0, y1;
	const int n = rem_pio2(x, y0, y1);
	switch (n & 3)
	{
	case 0:
		return kernel_sin(y0, y1, 1);
	case 1:
		return kernel_cos(y0, y1);
	case 2:
		return -kernel_sin(y0, y1, 1);
	default:
		return -kernel_cos(y0, y1);
	}
}

SGPU_HD double cos(double x)
{
	const int32_t ix = hi_word(x) & 0x7fffffff;
	if (ix <= 0x3fe921fb)
		return kernel_cos(x, 0.0);
	if (ix >= 0x7ff00000)
		return x - x; // NaN
	double y0, y1;
	const int n = rem_pio2(x, y0, y1);
	switch (n & 3)
	{
	case 0:
		return kernel_cos(y0, y1);
	case 1:
		return -kernel_sin(y0, y1, 1);
	case 2:
		return -kernel_cos(y0, y1);
	default:
		return kernel_sin(y0, y1, 1);
	}
}

// e_exp.c
SGPU_HD double exp(double x)
{
	const double o_threshold = 7.09782712893383973096e+02;
	const double u_threshold = -7.45133219101941108420e+02;
	const double ln2HI = 6.93147180369123816490e-01;
	const double ln2LO = 1.90821492927058770002e-10;
	const double invln2 = 1.44269504088896338700e+00;
	const double P1 = 1.66666666666666019037e-01;
	const double P2 = -2.77777777770155933842e-03;
	const double P3 = 6.61375632143793436117e-05;
	const double P4 = -1.65339022054652515390e-06;
	const double P5 = 4.13813679705723846039e-08;
	const double twom1000 = 9.33263618503218878990e-302; // 2^-1000
	int32_t hx = hi_word(x);
	const int xsb = (hx >> 31) & 1; // sign bit
	hx &= 0x7fffffff;
	if (hx >= 0x40862E42)
	{ // |x| >= 709.78...
		if (hx >= 0x7ff00000)
		{
			if (((hx & 0xfffff) | lo_word(x)) != 0)
				return x + x; // NaN
			return xsb == 0 ? x : 0.0; // exp(+-inf) = {inf, 0}
		}
		if (x > o_threshold)
			return from_bits(0x7ff0000000000000ull); // overflow
		if (x < u_threshold)
			return 0.0; // underflow
	}
	double hi = 0, lo = 0;
	int k = 0;
	if (hx > 0x3fd62e42)
	{ // |x| > 0.5 ln2
		if (hx < 0x3FF0A2B2)
		{ // and |x| < 1.5 ln2
			hi = xsb ? x + ln2HI : x - ln2HI;
			lo = xsb ? -ln2LO : ln2LO;
			k = 1 - xsb - xsb;
		}
		else
		{
			k = int(invln2 * x + (xsb ? -0.5 : 0.5));
			const double t = double(k);
			hi = x - t * ln2HI; // t*ln2HI is exact here
			lo = t * ln2LO;
		}
		x = hi - lo;
	}
	else if (hx < 0x3e300000)
	{ // |x| < 2^-28
		return 1.0 + x;
	}
	else
		k = 0;
	const double t = x * x;
	const double c = x - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
	if (k == 0)
		return 1.0 - ((x * c) / (c - 2.0) - x);
	double y = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
	if (k >= -1021)
		return with_hi(y, hi_word(y) + (k << 20));
	y = with_hi(y, hi_word(y) + ((k + 1000) << 20));
	return y * twom1000;
}

// e_acos.c
SGPU_HD double acos(double x)
{
	const double pi = 3.14159265358979311600e+00;
	const double pio2_hi = 1.57079632679489655800e+00;
	const double pio2_lo = 6.12323399573676603587e-17;
	const double pS0 = 1.66666666666666657415e-01;
	const double pS1 = -3.25565818622400915405e-01;
	const double pS2 = 2.01212532134862925881e-01;
	const double pS3 = -4.00555345006794114027e-02;
	const double pS4 = 7.91534994289814532176e-04;
	const double pS5 = 3.47933107596021167570e-05;
	const double qS1 = -2.40339491173441421878e+00;
	const double qS2 = 2.02094576023350569471e+00;
	const double qS3 = -6.88283971605453293030e-01;
	const double qS4 = 7.70381505559019352791e-02;
	const int32_t hx = hi_word(x);
	const int32_t ix = hx & 0x7fffffff;
	if (ix >= 0x3ff00000)
	{ // |x| >= 1
		if (((ix - 0x3ff00000) | int32_t(lo_word(x))) == 0)
		{ // |x| == 1
			if (hx > 0)
				return 0.0;
			return pi + 2.0 * pio2_lo;
		}
		return (x - x) / (x - x); // NaN
	}
	if (ix < 0x3fe00000)
	{ // |x| < 0.5
		if (ix <= 0x3c600000)
			return pio2_hi + pio2_lo; // |x| < 2^-57
		const double z = x * x;
		const double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
		const double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
		const double r = p / q;
		return pio2_hi - (x - (pio2_lo - x * r));
	}
	if (hx < 0)
	{ // x < -0.5
		const double z = (1.0 + x) * 0.5;
		const double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
		const double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
		const double s = __builtin_sqrt(z);
		const double r = p / q;
		const double w = r * s - pio2_lo;
		return pi - 2.0 * (s + w);
	}
	// x > 0.5
	const double z = (1.0 - x) * 0.5;
	const double s = __builtin_sqrt(z);
	const double df = make(hi_word(s), 0);
	const double c = (z - df * df) / (s + df);
	const double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
	const double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
	const double r = p / q;
	const double w = r * s + c;
	return 2.0 * (df + w);
}

} // namespace pmath

#endif

/*
 * Spline::eval_deriv (src/lib/splines.h:103-121) over the spline tables of
 * precalculate_splines (src/lib/precalculate.h:313-455): per type pair, nseg
 * cubic segments of `fraction` A, coefficients (x, a, b, c, d) in FP64.  A
 * pair whose terms vanish has cutoff 0 and evaluates to zero.
 */
#ifndef SGPU_SPLINE_H
#define SGPU_SPLINE_H

#include "pmath.h"

namespace sgpu
{

SGPU_HD void spline_eval(const double* spl, const double* cut, double fraction, int nseg, int tri, double r,
		double& val, double& der)
{
	val = 0;
	der = 0;
	if (r < cut[tri])
	{
		unsigned index = unsigned(r / fraction);
		index = index < unsigned(nseg) ? index : unsigned(nseg - 1);
		const double* s = spl + (size_t(tri) * nseg + index) * 5;
		const double lx = r - s[0];
		val = ((s[1] * lx + s[2]) * lx + s[3]) * lx + s[4];
		der = (3 * s[1] * lx + 2 * s[2]) * lx + s[3];
	}
}

} // namespace sgpu

#endif

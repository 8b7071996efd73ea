// The trilinear map term (grid::evaluate_aux, src/lib/grid.cpp:111-236) on
// corner-packed cells: the receptor maps of the search and the user grid
// (--user_grid) alike.
#ifndef SGPU_GRIDTERM_H
#define SGPU_GRIDTERM_H

#include "wave.h"

namespace sgpu
{

// ------------------------------------------------------- receptor map term
// grid::evaluate_aux (src/lib/grid.cpp:111-236) on a corner-packed cell, in
// two halves so that callers can overlap the gather with other work:
// grid_fetch locates the cell and loads its 8 corners, grid_finish
// interpolates, curls and adds the out-of-box penalty
template<typename T>
struct GridSample
{
	T c[8];     // f000 f100 f010 f110 f001 f101 f011 f111
	T s[3];     // position inside the cell
	T miss[3];  // distance outside the box along each axis (grid units)
	int region[3];
};

// the corners from `cells` (a map set laid out as g.cells: the receptor
// maps or, with charge-dependent terms, the charge maps)
template<typename T>
__device__ inline GridSample<T> grid_fetch(const GridPtrs<T>& g, const T* cells, int slot, V3<T> loc)
{
	GridSample<T> r;
	r.s[0] = (loc.x - T(g.init[0])) * T(g.factor[0]);
	r.s[1] = (loc.y - T(g.init[1])) * T(g.factor[1]);
	r.s[2] = (loc.z - T(g.init[2])) * T(g.factor[2]);
	int a[3];
	for (int i = 0; i < 3; i++)
	{
		const T dm1 = T(g.dim_fl_minus_1[i]);
		r.miss[i] = 0;
		if (r.s[i] < 0)
		{
			r.miss[i] = -r.s[i];
			r.region[i] = -1;
			a[i] = 0;
			r.s[i] = 0;
		}
		else if (r.s[i] >= dm1)
		{
			r.miss[i] = r.s[i] - dm1;
			r.region[i] = 1;
			a[i] = g.dim[i] - 2;
			r.s[i] = 1;
		}
		else
		{
			r.region[i] = 0;
			a[i] = int(r.s[i]);
			r.s[i] -= T(a[i]);
		}
		// in range for every finite location; keeps a NaN location's gather in bounds
		a[i] = a[i] < 0 ? 0 : (a[i] > g.dim[i] - 2 ? g.dim[i] - 2 : a[i]);
	}
	const size_t cell = size_t(a[0]) + size_t(g.dim[0] - 1) * (size_t(a[1]) + size_t(g.dim[1] - 1) * a[2]);
	const T* c = cells + (size_t(slot) * g.ncells + cell) * 8;
	for (int k = 0; k < 8; k++)
		r.c[k] = c[k];
	return r;
}

template<typename T>
__device__ inline GridSample<T> grid_fetch(const GridPtrs<T>& g, int slot, V3<T> loc)
{
	return grid_fetch(g, g.cells, slot, loc);
}

template<typename T>
__device__ inline T grid_finish(const GridPtrs<T>& g, const GridSample<T>& r, T v, T slope, V3<T>& deriv)
{
	const T penalty = slope * (r.miss[0] * T(g.factor_inv[0]) + r.miss[1] * T(g.factor_inv[1]) +
			r.miss[2] * T(g.factor_inv[2]));
	const T f000 = r.c[0], f100 = r.c[1], f010 = r.c[2], f110 = r.c[3];
	const T f001 = r.c[4], f101 = r.c[5], f011 = r.c[6], f111 = r.c[7];
	const T x = r.s[0], y = r.s[1], z = r.s[2];
	const T mx = 1 - x, my = 1 - y, mz = 1 - z;
	T f = f000 * mx * my * mz + f100 * x * my * mz + f010 * mx * y * mz + f110 * x * y * mz +
			f001 * mx * my * z + f101 * x * my * z + f011 * mx * y * z + f111 * x * y * z;
	// (-1) and 1 factors of the CPU expression are exact negations / no-ops
	const T x_g = -f000 * my * mz + f100 * my * mz + -f010 * y * mz + f110 * y * mz + -f001 * my * z +
			f101 * my * z + -f011 * y * z + f111 * y * z;
	const T y_g = (f000 * mx) * (-1) * mz + (f100 * x) * (-1) * mz + f010 * mx * mz + f110 * x * mz +
			(f001 * mx) * (-1) * z + (f101 * x) * (-1) * z + f011 * mx * z + f111 * x * z;
	const T z_g = (f000 * mx * my) * (-1) + (f100 * x * my) * (-1) + (f010 * mx * y) * (-1) +
			(f110 * x * y) * (-1) + f001 * mx * my + f101 * x * my + f011 * mx * y + f111 * x * y;
	V3<T> grad = v3(x_g, y_g, z_g);
	curl(f, grad, v);
	const T gr[3] = {grad.x, grad.y, grad.z};
	T d[3];
	for (int i = 0; i < 3; i++)
	{
		const T ge = r.region[i] == 0 ? gr[i] : T(0);
		d[i] = T(g.factor[i]) * ge + slope * T(r.region[i]);
	}
	deriv = v3(d[0], d[1], d[2]);
	return f + penalty;
}

template<typename T>
__device__ inline T grid_finish(const GridPtrs<T>& g, const GridSample<T>& r, T v, V3<T>& deriv)
{
	return grid_finish(g, r, v, T(g.slope), deriv);
}

template<typename T>
__device__ inline T grid_term(const GridPtrs<T>& g, int slot, V3<T> loc, T v, V3<T>& deriv)
{
	return grid_finish(g, grid_fetch(g, slot, loc), v, deriv);
}

// grid::evaluate's charge-dependent part (src/lib/grid.cpp:28-48): for an
// atom of charge q != 0, q times the charge map's term (value, derivative)
// added to the map term's
template<typename T>
__device__ inline void grid_charge_term(const GridPtrs<T>& g, int slot, V3<T> loc, T v, T q, T& e, V3<T>& deriv)
{
	V3<T> cd;
	const T ce = grid_finish(g, grid_fetch(g, g.qcells, slot, loc), v, cd);
	e += q * ce;
	deriv = v3(deriv.x + q * cd.x, deriv.y + q * cd.y, deriv.z + q * cd.z);
}

// evaluate_aux without a derivative (deriv NULL): the value curled by v
// plus the out-of-box penalty at `slope`
template<typename T>
__device__ inline T grid_value(const GridPtrs<T>& g, V3<T> loc, T v, T slope)
{
	const GridSample<T> r = grid_fetch(g, 0, loc);
	const T penalty = slope * (r.miss[0] * T(g.factor_inv[0]) + r.miss[1] * T(g.factor_inv[1]) +
			r.miss[2] * T(g.factor_inv[2]));
	const T f000 = r.c[0], f100 = r.c[1], f010 = r.c[2], f110 = r.c[3];
	const T f001 = r.c[4], f101 = r.c[5], f011 = r.c[6], f111 = r.c[7];
	const T x = r.s[0], y = r.s[1], z = r.s[2];
	const T mx = 1 - x, my = 1 - y, mz = 1 - z;
	T f = f000 * mx * my * mz + f100 * x * my * mz + f010 * mx * y * mz + f110 * x * y * mz +
			f001 * mx * my * z + f101 * x * my * z + f011 * mx * y * z + f111 * x * y * z;
	curl(f, v);
	return f + penalty;
}

} // namespace sgpu

#endif

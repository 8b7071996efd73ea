/*
 * Receptor map build on the device (the reference's build_grids_kernel,
 * src/gpu/gpu_device.cu:133-186, restating cache::populate,
 * src/lib/cache.cpp:116-210).
 *
 * 1. cell lists: for each 3 A cell, in receptor order, the heavy receptor
 *    atoms within the cutoff of both the box and the cell's brick.  The
 *    brick of a cell is [floor, ceil] (x 3 A) of the first probe coordinate
 *    falling into it, per dimension: that is the brick smina's szv_grid_cache
 *    builds (src/lib/szv_grid.h:100-130), degenerate where a probe lies on a
 *    3 A boundary.  Count pass, scan, ordered fill (one workgroup per cell,
 *    prefix sums keep receptor order).
 * 2. points: one thread per grid point (x fastest, so a wavefront's probes
 *    share a cell and read the same atoms), FP64 sums of the fast table over
 *    the cell's atoms for every needed ligand type.
 *    With a user grid (--user_grid), each point adds its value as
 *    cache::populate does (src/lib/cache.cpp:204-205).  With
 *    charge-dependent terms the point also sums the charge map
 *    (src/lib/cache.cpp:172-187): the map gets type-only plus |q_rec|
 *    components, the charge map the |q_lig| plus q_rec q_lig components
 *    that the evaluation scales by the ligand atom's charge.
 * 3. pack: point maps -> corner-packed cells (8 corner values per cell).
 */
#include <hip/hip_runtime.h>
#include "launch.h"
#include "atom_types.h"
#include "spline.h"
#include "gridterm.h"

namespace sgpu
{

constexpr int CELL_THREADS = 256;
constexpr int MAX_GRID_TYPES = 16;

__device__ inline double clamp_to(double lo, double hi, double x)
{
	if (x <= lo)
		return lo;
	else if (x >= hi)
		return hi;
	return x;
}
// brick_distance_sqr (src/lib/brick.h:26-40)
__device__ inline double brick_dist2(const double* lo, const double* hi, const double* v)
{
	const double c0 = clamp_to(lo[0], hi[0], v[0]);
	const double c1 = clamp_to(lo[1], hi[1], v[1]);
	const double c2 = clamp_to(lo[2], hi[2], v[2]);
	return (c0 - v[0]) * (c0 - v[0]) + (c1 - v[1]) * (c1 - v[1]) + (c2 - v[2]) * (c2 - v[2]);
}

__device__ inline bool cell_has(const GridBuildArgs& a, int cell, int atom)
{
	const int t = a.rec_type[atom];
	if (t >= NUM_ATOM_TYPES || is_hydrogen_type(t))
		return false;
	const double* v = a.rec_xyz + 3 * atom;
	if (!(brick_dist2(a.box_lo, a.box_hi, v) < a.cutoff_sqr))
		return false;
	const int cx = cell % a.cell_dim[0];
	const int cy = (cell / a.cell_dim[0]) % a.cell_dim[1];
	const int cz = cell / (a.cell_dim[0] * a.cell_dim[1]);
	const double lo[3] = {a.cell_lo[cx], a.cell_lo[a.cell_dim[0] + cy], a.cell_lo[a.cell_dim[0] + a.cell_dim[1] + cz]};
	const double hi[3] = {a.cell_hi[cx], a.cell_hi[a.cell_dim[0] + cy], a.cell_hi[a.cell_dim[0] + a.cell_dim[1] + cz]};
	return brick_dist2(lo, hi, v) < a.cutoff_sqr;
}

// block-wide exclusive prefix sum of one flag per thread
__device__ inline int block_scan(int flag, int& total)
{
	__shared__ int warp_sums[CELL_THREADS / 64];
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const unsigned long long mask = __ballot(flag);
	const int before = __popcll(mask & ((1ull << lane) - 1ull));
	if (lane == 0)
		warp_sums[w] = __popcll(mask);
	__syncthreads();
	int off = 0;
	total = 0;
	for (int k = 0; k < CELL_THREADS / 64; k++)
	{
		if (k < w)
			off += warp_sums[k];
		total += warp_sums[k];
	}
	__syncthreads();
	return off + before;
}

__global__ __launch_bounds__(CELL_THREADS) void cell_count_kernel(GridBuildArgs a)
{
	const int cell = blockIdx.x;
	int count = 0;
	for (int base = 0; base < a.rec_n; base += CELL_THREADS)
	{
		const int i = base + int(threadIdx.x);
		const int f = (i < a.rec_n && cell_has(a, cell, i)) ? 1 : 0;
		int total;
		block_scan(f, total);
		count += total;
	}
	if (threadIdx.x == 0)
		a.cell_start[cell + 1] = count;
}

// exclusive scan of the counts in place (cell_start[0] = 0), one block
__global__ __launch_bounds__(64) void cell_scan_kernel(int* start, int ncells)
{
	if (threadIdx.x == 0)
	{
		start[0] = 0;
		for (int c = 0; c < ncells; c++)
			start[c + 1] += start[c];
	}
}

__global__ __launch_bounds__(CELL_THREADS) void cell_fill_kernel(GridBuildArgs a)
{
	const int cell = blockIdx.x;
	int out = a.cell_start[cell];
	for (int base = 0; base < a.rec_n; base += CELL_THREADS)
	{
		const int i = base + int(threadIdx.x);
		const int f = (i < a.rec_n && cell_has(a, cell, i)) ? 1 : 0;
		int total;
		const int pos = block_scan(f, total);
		if (f)
			a.cell_atoms[out + pos] = i;
		out += total;
	}
}

__global__ __launch_bounds__(256) void grid_points_kernel(GridBuildArgs a, double* points, double* qpoints)
{
	const size_t npts = size_t(a.dim[0]) * a.dim[1] * a.dim[2];
	const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (p >= npts)
		return;
	const int x = int(p % a.dim[0]);
	const int y = int((p / a.dim[0]) % a.dim[1]);
	const int z = int(p / (size_t(a.dim[0]) * a.dim[1]));
	// grid::index_to_argument
	const double probe[3] = {a.init[0] + a.factor_inv[0] * x, a.init[1] + a.factor_inv[1] * y,
			a.init[2] + a.factor_inv[2] * z};
	int ix[3];
	for (int d = 0; d < 3; d++)
		ix[d] = int(__builtin_floor(probe[d] / 3.0)) - a.cell_off[d];
	const int cell = ix[0] + a.cell_dim[0] * (ix[1] + a.cell_dim[1] * ix[2]);
	double aff[MAX_GRID_TYPES], caff[MAX_GRID_TYPES];
	int tt[MAX_GRID_TYPES];
	for (int j = 0; j < MAX_GRID_TYPES; j++)
	{
		aff[j] = 0;
		caff[j] = 0;
		tt[j] = j < a.ntypes ? a.types[j] : 0;
	}
	for (int k = a.cell_start[cell]; k < a.cell_start[cell + 1]; k++)
	{
		const int i = a.cell_atoms[k];
		const double* v = a.rec_xyz + 3 * i;
		const double d0 = v[0] - probe[0], d1 = v[1] - probe[1], d2 = v[2] - probe[2];
		const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
		if (r2 <= a.cutoff_sqr)
		{
			const int t1 = a.rec_type[i];
			if (a.kind == 1)
			{
				// precalculate_splines::eval_fast: the spline's value
				const double r = __builtin_sqrt(r2);
				for (int j = 0; j < MAX_GRID_TYPES; j++)
					if (j < a.ntypes)
					{
						double val, der;
						spline_eval(a.spl, a.spl_cut, a.fraction, a.nseg, tri_index_permissive(t1, tt[j]), r, val, der);
						aff[j] += val;
					}
				continue;
			}
			const size_t idx = size_t(a.factor * r2);
			if (a.ncomp == 4)
			{
				// precalculate_linear::eval_fast_data of (receptor type,
				// ligand type): the components swapped when t1 > t2
				const double q = a.rec_q[i];
				const double aq = q < 0 ? -q : q;
				for (int j = 0; j < MAX_GRID_TYPES; j++)
					if (j < a.ntypes)
					{
						const size_t k = size_t(tri_index_permissive(t1, tt[j])) * a.table_n + idx;
						const double* qf = a.qfast + 3 * k;
						const bool in_order = t1 <= tt[j];
						const double c1 = in_order ? qf[0] : qf[1], c2 = in_order ? qf[1] : qf[0];
						aff[j] += a.fast[k] + c1 * aq;
						caff[j] += c2 + qf[2] * q;
					}
				continue;
			}
			for (int j = 0; j < MAX_GRID_TYPES; j++)
				if (j < a.ntypes)
					aff[j] += a.fast[size_t(tri_index_permissive(t1, tt[j])) * a.table_n + idx];
		}
	}
	// the user grid at (x, y, z) as a location in A: the reference's own
	// expression, kept as it is
	const double uge = a.has_ug ? grid_value(a.ug, v3(double(x), double(y), double(z)), 1000.0, a.ug_slope) : 0.0;
	for (int j = 0; j < MAX_GRID_TYPES; j++)
		if (j < a.ntypes)
		{
			points[size_t(j) * npts + p] = a.has_ug ? aff[j] + uge : aff[j];
			if (a.ncomp == 4)
				qpoints[size_t(j) * npts + p] = caff[j];
		}
}

template<typename T>
__global__ __launch_bounds__(256) void grid_pack_kernel(const double* points, int ntypes, int d0, int d1, int d2, T* cells)
{
	const size_t nc = size_t(d0 - 1) * (d1 - 1) * (d2 - 1);
	const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (k >= nc * ntypes)
		return;
	const int s = int(k / nc);
	const size_t c = k % nc;
	const int x = int(c % (d0 - 1));
	const int y = int((c / (d0 - 1)) % (d1 - 1));
	const int z = int(c / (size_t(d0 - 1) * (d1 - 1)));
	const double* m = points + size_t(s) * d0 * d1 * d2;
	auto at = [&](int i, int j, int l) { return m[size_t(i) + size_t(d0) * (size_t(j) + size_t(d1) * l)]; };
	T* out = cells + k * 8;
	out[0] = T(at(x, y, z));
	out[1] = T(at(x + 1, y, z));
	out[2] = T(at(x, y + 1, z));
	out[3] = T(at(x + 1, y + 1, z));
	out[4] = T(at(x, y, z + 1));
	out[5] = T(at(x + 1, y, z + 1));
	out[6] = T(at(x, y + 1, z + 1));
	out[7] = T(at(x + 1, y + 1, z + 1));
}

hipError_t launch_cell_count(const GridBuildArgs& a, hipStream_t st)
{
	const int ncells = a.cell_dim[0] * a.cell_dim[1] * a.cell_dim[2];
	hipLaunchKernelGGL(cell_count_kernel, dim3(ncells), dim3(CELL_THREADS), 0, st, a);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return e;
	hipLaunchKernelGGL(cell_scan_kernel, dim3(1), dim3(64), 0, st, a.cell_start, ncells);
	return hipGetLastError();
}

hipError_t launch_cell_fill(const GridBuildArgs& a, hipStream_t st)
{
	const int ncells = a.cell_dim[0] * a.cell_dim[1] * a.cell_dim[2];
	hipLaunchKernelGGL(cell_fill_kernel, dim3(ncells), dim3(CELL_THREADS), 0, st, a);
	return hipGetLastError();
}

hipError_t launch_grid_points(const GridBuildArgs& a, double* points, double* qpoints, hipStream_t st)
{
	if (a.ntypes > MAX_GRID_TYPES || (a.ncomp == 4 && (!qpoints || !a.qfast || !a.rec_q || a.kind != 0)))
		return hipErrorInvalidValue;
	const size_t npts = size_t(a.dim[0]) * a.dim[1] * a.dim[2];
	hipLaunchKernelGGL(grid_points_kernel, dim3(unsigned((npts + 255) / 256)), dim3(256), 0, st, a, points, qpoints);
	return hipGetLastError();
}

template<typename T>
hipError_t launch_grid_pack(const double* points, int ntypes, const int* dim, T* cells, hipStream_t st)
{
	const size_t nc = size_t(dim[0] - 1) * (dim[1] - 1) * (dim[2] - 1) * ntypes;
	hipLaunchKernelGGL(grid_pack_kernel<T>, dim3(unsigned((nc + 255) / 256)), dim3(256), 0, st, points, ntypes, dim[0],
			dim[1], dim[2], cells);
	return hipGetLastError();
}
template hipError_t launch_grid_pack<float>(const double*, int, const int*, float*, hipStream_t);
template hipError_t launch_grid_pack<double>(const double*, int, const int*, double*, hipStream_t);

} // namespace sgpu

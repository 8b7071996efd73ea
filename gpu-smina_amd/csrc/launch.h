// Host-side declarations of the kernel launches (one definition per
// precision in kernels_f32.hip / kernels_f64.hip / kernels_mixed.hip; grid
// build in grid.hip).
#ifndef SGPU_LAUNCH_H
#define SGPU_LAUNCH_H

#include <hip/hip_runtime.h>
#include "layout.h"

namespace sgpu
{

struct WavePlan;
template<typename T> struct EvalCtx;
struct McParams;
struct ChainLayout;

template<typename T>
hipError_t launch_eval(const BatchPtrs<T>& B, const EvalCtx<T>& ec, const WavePlan& plan, int n, const int* model,
		const T* conf, int conf_stride, T* e, T* g, int g_stride, T* x, int x_stride, T* scratch, int pstride,
		hipStream_t st);
// nblocks workgroup items of `waves` (1..8) chains of one ligand each,
// running the steps [ceil(steps x part_lo / parts), ceil(steps x part_hi /
// parts)) of every chain: a persistent grid of `grid` workgroups taking
// items from the work-queue counter *queue (cleared), or with queue null
// one workgroup per item (grid = nblocks); T is the
// evaluation's real type, S the search state's (kernels_mixed.hip: float,
// double)
template<typename T, typename S>
hipError_t launch_mc(const BatchPtrs<T>& B, const EvalCtx<T>& ec, const WavePlan& plan, const McParams& P,
		const ChainLayout& L, char* chains, int nblocks, int waves, const int* block_model, const int* slot_chain,
		const int* chain_steps, const int* chain_maxiters, int part_lo, int part_hi, int parts, int grid, int* queue,
		bool spec, hipStream_t st);
// workgroups of `waves` chains with `lds` bytes of LDS that one CU holds at
// once (the runtime's occupancy calculator: registers and LDS)
template<typename T, typename S>
// variant: 0 smina's search, 1 with QuickVina 2's history, 2 with the
// spline or charge-dependent tables (and the history); spec: the
// speculative-chain kernel (waves = its width)
int mc_blocks_per_cu(int waves, size_t lds, int variant = 0, bool persist = false, bool spec = false);
// per-lane private (scratch) bytes of a Monte Carlo kernel form: register
// spills, which the persistent form must not add
template<typename T, typename S>
size_t mc_private_bytes(int variant, bool persist);
template<typename T, typename S>
hipError_t launch_refine(const BatchPtrs<T>& B, const EvalCtx<T>& ec, const WavePlan& plan, int n, const int* model,
		S* conf, int conf_stride, const int* maxiters, int accurate, int early_term, S* e, T* x, int x_stride,
		T* scratch, int pstride, int64_t* evals, const double* boxes, int* tries, hipStream_t st);
template<typename T, typename S>
hipError_t launch_bfgs(const BatchPtrs<T>& B, const EvalCtx<T>& ec, const WavePlan& plan, int n, const int* model,
		S* conf, int conf_stride, const int* maxiters, int accurate, int early_term, S* e, T* scratch, int pstride,
		int64_t* evals, hipStream_t st);
// sgpu_eval_confs_noalloc: the caller's device arrays (double) and the
// layout of the library's scratch region for n confs of a batch
struct NoallocRequest
{
	int n;
	const int* model;
	const double* conf;
	int conf_stride;
	double* e;
	double* g;
	int change_stride;
	double* coords;
	int coord_stride;
	int* status;
};
// bits of the status word
constexpr int NOALLOC_ATOM_OVERLAP = 1; // the direct receptor term met an exact overlap
constexpr int NOALLOC_BAD_MODEL = 2;    // a model index outside the batch
struct NoallocLayout
{
	int cs, gs, xs, ps;
	size_t status, model, conf, e, g, x, pairs, bytes;
};
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
// n confs of a batch whose largest model has cmax conf entries, gmax change
// entries, nmax atoms and pmax pairs, evaluated in reals of `real` bytes
inline NoallocLayout noalloc_layout(int n, int cmax, int gmax, int nmax, int pmax, size_t real)
{
	NoallocLayout L{};
	L.cs = cmax;
	L.gs = gmax;
	L.xs = 3 * nmax;
	L.ps = pmax;
	const size_t nn = size_t(n > 0 ? n : 0);
	L.status = 0;
	L.model = 256;
	L.conf = L.model + align256(nn * sizeof(int));
	L.e = L.conf + align256(nn * size_t(L.cs) * real);
	L.g = L.e + align256(nn * real);
	L.x = L.g + align256(nn * size_t(L.gs) * real);
	L.pairs = L.x + align256(nn * size_t(L.xs) * real);
	L.bytes = L.pairs + align256(nn * size_t(L.ps) * real);
	return L;
}
template<typename T>
hipError_t launch_eval_noalloc(const BatchPtrs<T>& B, const EvalCtx<T>& ec, const WavePlan& plan, int nmodels,
		const NoallocRequest& q, unsigned char* scratch, size_t scratch_bytes, hipStream_t st);

template<typename T>
hipError_t launch_score(const BatchPtrs<T>& B, const EvalCtx<T>& ec, const WavePlan& plan, int mode, int n,
		const int* model, const T* conf, int conf_stride, T* e, T* scratch, int pstride, hipStream_t st);

// grid build: cell lists (per-dimension bricks lo/hi of each cell), then
// the point maps in FP64 sums, then the corner-packed cells
struct GridBuildArgs
{
	const double* rec_xyz;
	const int* rec_type;
	const double* rec_q;   // receptor partial charges (ncomp 4)
	int rec_n;
	const double* fast;
	const double* qfast;   // [378][n][3] charge components, ncomp 4
	int ncomp;             // 4: the point pass also writes the charge maps
	int table_n;
	double factor;
	double cutoff_sqr;
	int kind;              // 0: linear tables (fast), 1: splines
	const double* spl;
	const double* spl_cut;
	double fraction;
	int nseg;
	int cell_off[3];
	int cell_dim[3];
	const double* cell_lo; // [cell_dim[0] + cell_dim[1] + cell_dim[2]] per dimension
	const double* cell_hi;
	double box_lo[3];
	double box_hi[3];
	int* cell_start;       // [ncells + 1]
	int* cell_atoms;       // sized by the count pass
	int dim[3];            // grid points per axis
	double init[3];
	double factor_inv[3];
	const int* types;      // [ntypes] ligand types
	int ntypes;
	// the user grid, added at every point as cache::populate adds it
	// (src/lib/cache.cpp:204-205): evaluated at the point's integer
	// indices taken as a location, with the map slope and cap 1000
	GridPtrs<double> ug;
	int has_ug;
	double ug_slope;
};
hipError_t launch_cell_count(const GridBuildArgs& a, hipStream_t st);
hipError_t launch_cell_fill(const GridBuildArgs& a, hipStream_t st);
// points: [ntypes][npts]; qpoints: the charge maps, same layout (ncomp 4)
hipError_t launch_grid_points(const GridBuildArgs& a, double* points, double* qpoints, hipStream_t st);
template<typename T>
hipError_t launch_grid_pack(const double* points, int ntypes, const int* dim, T* cells, hipStream_t st);

} // namespace sgpu

#endif

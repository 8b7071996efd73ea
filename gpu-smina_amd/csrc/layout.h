/*
 * Device-memory layout of the docking search (HBM, gfx950).
 *
 * A batch of models is packed per model into one contiguous int blob and
 * one contiguous real blob (the reference's FlatModel / DevModels,
 * src/gpu/gpu_flatten_data.h:16-46, src/gpu/gpu_dev_structs.h:137).  A
 * workgroup copies its model's two blobs into LDS with coalesced loads once,
 * then every wave of the workgroup (one Monte Carlo chain each, all of the
 * same ligand) reads the model from LDS only.  ModelHdr holds the counts
 * and each array's offset inside the blobs.
 *
 * int blob:  atom type (movable then fixed atoms), atom frame, per-atom
 *            neighbour lists (start, packed codes), frames (parent, kind,
 *            atom range, conf / change index, children), tree levels, heavy
 *            atoms, trees, conf / change maps;
 * real blob: initial coordinates (movable then fixed), local coordinates,
 *            frame origins and axes (segment: relative to the parent),
 *            partial charges (movable then fixed).
 *
 * A neighbour code is one 32-bit word: the partner atom (bits 0-7), the
 * type-pair table index (8-16), "this atom is the pair's b" (17), "the pair
 * is an other pair" (18) and the pair index (19-31), so the pair loop reads
 * one LDS word per neighbour.  The pair-parallel loop of the FP32 search
 * reads the pairs themselves, two words each (coordinate offsets of a and
 * b; table index and "other pair").
 *
 * Receptor maps are stored per ligand type as corner-packed cells: for each
 * grid cell (x fastest), its 8 corner values in the order f000 f100 f010
 * f110 f001 f101 f011 f111, so one trilinear gather is one 32-byte (FP32)
 * or 64-byte (FP64) load.
 */
#ifndef SGPU_LAYOUT_H
#define SGPU_LAYOUT_H

#include <hip/hip_runtime.h>

namespace sgpu
{

constexpr int WAVE = 64;
constexpr int MAX_ATOMS = 255;    // movable + fixed atoms per model (8-bit partner field)
constexpr int MAX_PAIRS = 8191;   // 13-bit pair field
constexpr int MAX_FRAMES = 64;
constexpr int MAX_CHANGE = 64;
constexpr int MAX_CONF = 64;
constexpr int MAX_TREES = 16;
constexpr int MAX_POSES = 64;
constexpr int MAX_PATH = 32;
constexpr int MAX_CHAINS_PER_GROUP = 8; // waves of one Monte Carlo workgroup      // frames on a root-to-frame path (ancestor walk of set_conf)

// conf element kinds: position component, quaternion component a..d, torsion
enum ConfKind
{
	CK_POSITION = 0,
	CK_QUAT0 = 1, // CK_QUAT0 + component
	CK_TORSION = 5
};
// change element kinds: position, orientation component x..z, torsion
enum ChangeKind
{
	GK_POSITION = 0,
	GK_ORIENT0 = 1, // GK_ORIENT0 + component
	GK_TORSION = 4
};

__host__ __device__ inline int code_partner(int c) { return c & 255; }
__host__ __device__ inline int code_tri(int c) { return (c >> 8) & 511; }
__host__ __device__ inline bool code_is_b(int c) { return (c >> 17) & 1; }
__host__ __device__ inline bool code_other(int c) { return (c >> 18) & 1; }
__host__ __device__ inline int code_pair(int c) { return int(unsigned(c) >> 19); }
__host__ __device__ inline int make_code(int partner, int tri, bool is_b, bool other, int pair)
{
	return int(unsigned(partner) | (unsigned(tri) << 8) | (unsigned(is_b) << 17) | (unsigned(other) << 18) |
			(unsigned(pair) << 19));
}

// pair record of the pair-parallel loop: two words, the coordinate offsets
// 3a (bits 0-15) and 3b (16-31), then the table index (0-15) and "other
// pair" (16)
__host__ __device__ inline int prec_xa(int w0) { return w0 & 0xffff; }
__host__ __device__ inline int prec_xb(int w0) { return int(unsigned(w0) >> 16); }
__host__ __device__ inline int prec_tri(int w1) { return w1 & 0xffff; }
__host__ __device__ inline bool prec_other(int w1) { return (w1 >> 16) & 1; }

struct ModelHdr
{
	int natoms, ninflex, nframes, ntrees, nligands, npairs, nother;
	int conf_size, change_size, nheavy, nlevels, num_tors;
	int path_ok; // every root-to-frame path has at most MAX_PATH frames
	int iblob_off, iblob_len; // in the batch int blob
	int tblob_off, tblob_len; // in the batch real blob
	int pair_off;             // pair arrays (a, b, table index), global
	// offsets inside the int blob
	int o_type, o_afr, o_nbr, o_code, o_prec, o_fparent, o_fkind, o_fbegin, o_fend, o_fconf, o_fchange, o_fsend, o_fdepth, o_fpath, o_cstart,
			o_ccount, o_clist, o_lstart, o_lframes, o_heavy, o_troot, o_tconf, o_tchange, o_tlig, o_tntors,
			o_tab, o_tae, o_tpb, o_tpe, o_maps;
	// offsets inside the real blob
	int o_xyz0, o_local, o_forig, o_faxis, o_charge;
};

template<typename T>
struct BatchPtrs
{
	const ModelHdr* hdr;
	const int* iblob;
	const T* tblob;
	const int* pair_a;
	const int* pair_b;
	const int* pair_tri;
};

// scoring tables on the device: smooth as (e, dor) pairs per point; fast
// values kept in FP64 for the grid build and the final rescoring.  With
// charge-dependent terms (ncomp 4) smooth and fast hold the first (type
// only) component and qsmooth / qfast the other three (|q_a|, |q_b| and
// q_a q_b dependent) of the pair (min type, max type).
template<typename T>
struct TablePtrs
{
	const T* smooth; // [378][n][2]
	const double* fast; // [378][n]
	int ncomp;           // 1 or 4
	const T* qsmooth;    // [378][n][3][2] (ncomp 4)
	const double* qfast; // [378][n][3] (ncomp 4)
	int n;
	double factor;
	double cutoff_sqr;
	// precalculate_splines (kind 1; FP64 direct evaluation only)
	int kind;            // 0: linear tables, 1: splines
	const double* spl;   // [378][nseg][5]: x, a, b, c, d
	const double* spl_cut; // [378]: 0 for a zero spline
	double fraction;     // segment length
	int nseg;
};

// one receptor map set (cache back end)
template<typename T>
struct GridPtrs
{
	const T* cells;      // [type slot][cells][8]
	const T* qcells;     // charge maps, same slots (ncomp 4), else null
	const int* slot_of;  // [27] -> map slot or -1
	int dim[3];          // points per axis
	int ncells;          // (dim0-1)(dim1-1)(dim2-1)
	double init[3];
	double factor[3];
	double factor_inv[3];
	double dim_fl_minus_1[3];
	double slope;
};

// receptor for the direct (non_cache) back end: atoms and 3 A cells over
// the box holding, in receptor order, the atoms within the cutoff of the
// cell's brick
template<typename T>
struct ReceptorPtrs
{
	const T* xyz;        // [natoms*3] all atoms (H included)
	const int* type;
	const T* charge;     // [natoms] partial charges
	int natoms;
	const int* cell_start; // [ncells+1]
	const int* cell_atoms;
	int cell_off[3];
	int cell_dim[3];
	double begin[3];
	double end[3];
	int enabled[3];
};

} // namespace sgpu

#endif

/*
 * Device-memory layout of the docking search (HBM, gfx950).
 *
 * A batch of models is packed into flat arrays; ModelHdr gives each model's
 * counts and offsets (the reference's ModelDesc / DevModels,
 * src/gpu/gpu_defs.h:98-115, src/gpu/gpu_dev_structs.h:137).  All offsets are
 * element offsets into the batch arrays.  Per model:
 *   atoms      type (movable then fixed), frame, local coordinates, initial
 *              coordinates (movable then fixed), per-atom pair lists (CSR:
 *              code = pair << 1 | (atom is the pair's b));
 *   frames     parent, kind, atom range, conf / change index, children list,
 *              origin and axis (segment: relative to the parent);
 *   levels     frames grouped by tree depth (set_conf top-down, derivative
 *              bottom-up);
 *   pairs      a, b, type-pair table index; other pairs first, then the
 *              ligand-internal pairs;
 *   change map kind and conf index of every change (gradient) element;
 *   trees      root frame, conf / change offsets, ligand flag, torsions,
 *              atom range.
 *
 * Receptor maps are stored per ligand type as corner-packed cells: for each
 * grid cell (x fastest), its 8 corner values in the order f000 f100 f010
 * f110 f001 f101 f011 f111, so one trilinear gather is one 32-byte (FP32)
 * or 64-byte (FP64) load.
 */
#ifndef SGPU_LAYOUT_H
#define SGPU_LAYOUT_H

namespace sgpu
{

constexpr int WAVE = 64;
constexpr int MAX_ATOMS = 256;    // movable atoms per model
constexpr int MAX_FRAMES = 64;    // frames per model (one per lane)
constexpr int MAX_CHANGE = 64;    // change (gradient) size (one per lane)
constexpr int MAX_CONF = 64;      // conf size (one per lane)
constexpr int MAX_TREES = 16;
constexpr int MAX_LEVELS = 64;
constexpr int MAX_POSES = 64;

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

struct ModelHdr
{
	int natoms, ninflex, nframes, ntrees, nligands, npairs, nother;
	int conf_size, change_size, nheavy, nlevels, num_tors;
	int atom_off;   // atom_type (natoms+ninflex), atom_xyz0 (3*(natoms+ninflex))
	int mov_off;    // atom_frame, atom_local (3x), per movable atom
	int nbr_off;    // nbr_start: natoms+1 entries
	int code_off;   // nbr_code
	int frame_off;  // frame arrays
	int child_off;  // child_list
	int level_off;  // level_start: nlevels+1 entries
	int lvf_off;    // level_frames: nframes entries
	int pair_off;
	int heavy_off;  // heavy movable atom indices
	int tree_off;
	int map_off;    // conf_kind, conf_chg (conf_size each), change_kind, change_conf (change_size each)
};

template<typename T>
struct BatchPtrs
{
	const ModelHdr* hdr;
	const int* atom_type;
	const T* atom_xyz0;
	const int* atom_frame;
	const T* atom_local;
	const int* nbr_start;
	const int* nbr_code;
	const int* frame_parent;
	const int* frame_kind;
	const int* frame_begin;
	const int* frame_end;
	const int* frame_conf;
	const int* frame_change;
	const int* child_start;
	const int* child_count;
	const int* child_list;
	const T* frame_origin;
	const T* frame_axis;
	const int* level_start;
	const int* level_frames;
	const int* pair_a;
	const int* pair_b;
	const int* pair_tri;
	const int* heavy;
	const int* tree_root;
	const int* tree_conf;
	const int* tree_change;
	const int* tree_ligand;
	const int* tree_ntors;
	const int* tree_atom_begin;
	const int* tree_atom_end;
	const int* tree_pair_begin;
	const int* tree_pair_end;
	const int* maps;
};

// scoring tables on the device: smooth as (e, dor) pairs per point; fast
// values kept in FP64 for the grid build and the final rescoring
template<typename T>
struct TablePtrs
{
	const T* smooth; // [378][n][2]
	const double* fast; // [378][n]
	int n;
	double factor;
	double cutoff_sqr;
};

// one receptor map set (cache back end)
template<typename T>
struct GridPtrs
{
	const T* cells;      // [type slot][cells][8]
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

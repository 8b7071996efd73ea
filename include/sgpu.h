/*
 * sgpu: MI355X (gfx950) docking-search kernels behind a C ABI.
 *
 * This is the device layer under the host GpuDocking facade
 * (gpu-smina_amd/host/gpu_docking.h).  It replaces, one entry point each,
 * the reference's kernel-level interface src/gpu/gpu_kernels.h:22-126
 * (eval_confs, run_monte_carlo, run_bfgs, run_refine, residency) and the
 * device-memory owners of src/gpu/gpu_device.h:59-119 (DeviceTables,
 * DeviceReceptor, DeviceGrids::build, DeviceModels::upload).  Templates on
 * float|double become a precision argument.
 *
 * Conventions: every call returns an sgpu_status; sgpu_last_error() gives the
 * message of the last failure on a context.  Host arrays belong to the caller
 * and are copied; device objects are owned by the context and named by ids.
 * Calls on one context are serialised by the context.
 */
#ifndef SGPU_H
#define SGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sgpu_status
{
	SGPU_OK = 0,
	SGPU_E_INVALID = 1,      /* bad argument or inconsistent sizes */
	SGPU_E_HIP = 2,          /* a HIP runtime call failed */
	SGPU_E_OOM = 3,          /* device allocation failed */
	SGPU_E_ATOM_OVERLAP = 4, /* a ligand atom exactly on a receptor atom
	                            (src/lib/non_cache.cpp:170-172) */
	SGPU_E_NO_DEVICE = 5,
	SGPU_E_NOT_FOUND = 6     /* unknown id */
} sgpu_status;

typedef enum sgpu_precision
{
	SGPU_FP32 = 0, /* served search: FP32 geometry, energies and forces */
	SGPU_FP64 = 1, /* oracle-parity search, ordered like smina's CPU code */
	SGPU_MIXED = 2 /* FP32 evaluation (maps, pair tables, forces) with the
	                  search state in FP64: conf and root accumulation,
	                  BFGS Hessian update, line search and Armijo test,
	                  Metropolis energies.  Models only: its maps are FP32. */
} sgpu_precision;

enum
{
	SGPU_NUM_TYPES = 27,      /* smina atom types, src/lib/atom_constants.h:92-121 */
	SGPU_NUM_TYPE_PAIRS = 378 /* triangular pairs of them */
};

/* ---- scoring tables (precalculate_linear, src/lib/precalculate.h:206-237) */
typedef struct sgpu_tables
{
	int n;               /* control points per type pair (2051 at factor 32, cutoff 8) */
	double factor;       /* points per A^2 */
	double cutoff_sqr;
	const double* smooth; /* [378][n][2]: (e, dor) of the charge-independent component */
	const double* fast;   /* [378][n] */
	/* with charge-dependent terms (weighted_terms.h:51-57: four result
	 * components): 4, and the other three components (|q_a|, |q_b| and
	 * q_a q_b dependent, for the pair (min type, max type)) in qsmooth
	 * [378][n][3][2] and qfast [378][n][3]; else 1 and NULL */
	int ncomp;
	const double* qsmooth;
	const double* qfast;
} sgpu_tables;

/* ---- precalculate_splines (src/lib/precalculate.h:313-455): per type pair,
 * the cubic spline of the terms over nseg segments of `fraction` A */
typedef struct sgpu_splines
{
	int nseg;            /* factor x cutoff (80 at factor 10, cutoff 8) */
	double fraction;     /* cutoff / nseg */
	double cutoff_sqr;
	const double* coef;  /* [378][nseg][5]: x, a, b, c, d */
	const double* cut;   /* [378]: the spline's cutoff, 0 for a pair whose terms vanish */
} sgpu_splines;

/* ---- receptor (heavy, typed atoms; model::grid_atoms) */
typedef struct sgpu_receptor
{
	int natoms;
	const double* xyz; /* [3*natoms] */
	const int* type;   /* [natoms] smina types */
	const double* charge; /* [natoms] partial charges (NULL: all 0) */
} sgpu_receptor;

/* ---- a grid box, grid_dims (src/lib/grid_dim.h): n intervals per axis */
typedef struct sgpu_grid_dims
{
	double begin[3];
	double end[3];
	int n[3];
} sgpu_grid_dims;

/* ---- flat model (FlatModel, src/gpu/gpu_flatten_data.h:16-46), built by the
 * host facade from a model description.
 * Atoms 0..natoms-1 move (ligand atoms, then flexible side-chain atoms); atoms
 * natoms..natoms+ninflex-1 are the fixed atoms of flexible residues.
 * Frames are in preorder within each tree; trees are the ligands, then the
 * flexible side chains.  Conf layout per ligand: position[3], orientation
 * quaternion[4], torsions; per side chain: torsions.  Change layout per
 * ligand: force-like position[3], torque-like orientation[3], torsions. */
enum
{
	SGPU_FRAME_RIGID_ROOT = 0, /* rigid_body: a ligand root */
	SGPU_FRAME_FIRST_SEGMENT = 1, /* first_segment: a side-chain root */
	SGPU_FRAME_SEGMENT = 2     /* segment: one rotatable bond */
};
typedef struct sgpu_flat_model
{
	int natoms;
	int ninflex;
	int nframes;
	int nligands;
	int ntrees;
	int npairs;  /* other_pairs (list 2) first, then ligand-internal pairs (list 0) */
	int nother;
	int conf_size;
	int change_size;
	int num_tors; /* conf_independent_inputs::num_tors of the ligand(s) */
	const int* atom_type;     /* [natoms+ninflex] */
	const int* atom_frame;    /* [natoms] */
	const double* atom_local; /* [3*natoms] position relative to the frame origin */
	const double* atom_init;  /* [3*(natoms+ninflex)] model coordinates before the first set */
	const int* frame_parent;  /* [nframes] -1 for a tree root */
	const int* frame_kind;    /* [nframes] SGPU_FRAME_* */
	const int* frame_begin;   /* [nframes] atom range */
	const int* frame_end;
	const int* frame_tree;    /* [nframes] */
	const int* frame_conf;    /* [nframes] conf index: position (root) or torsion */
	const int* frame_change;  /* [nframes] change index: position (root) or torsion */
	const double* frame_origin; /* [3*nframes] root: initial origin; first segment:
	                               origin; segment: origin relative to the parent */
	const double* frame_axis;   /* [3*nframes] first segment: axis; segment: axis
	                               relative to the parent; root: unused */
	const int* pair_a;        /* [npairs] a < b; b may be a fixed atom */
	const int* pair_b;
	const int* pair_tri;      /* [npairs] triangular index of (min type, max type) */
	const double* atom_charge; /* [natoms+ninflex] partial charges (NULL: all 0) */
} sgpu_flat_model;

/* ---- search parameters (MCParams, src/gpu/gpu_mc_structs.h:21-34) */
typedef struct sgpu_mc_params
{
	double temperature;        /* 1.2 */
	double mutation_amplitude; /* 2 */
	double min_rmsd;           /* 1 */
	int num_saved_mins;        /* 20 (at most SGPU_MAX_POSES) */
	double hunt_cap[3];        /* (10, 10, 10) */
	double authentic_v[3];     /* (1000, 1000, 1000) */
	double cache_slope;        /* 1e6 */
	int accurate_line_search;  /* 0: fast_line_search */
	int early_term;
	/* Early stopping (not smina's search; EarlyStop and stop_agreeing,
	 * src/gpu/gpu_kernels.h:51-56, src/gpu/gpu_kernels_impl.cuh:607-657):
	 * with n = early_stop_checkpoints, the launches end at first/n,
	 * (first+1)/n, .., n/n of every chain's steps (first =
	 * early_stop_first); after each but the last, a model whose search units'
	 * best energies agree (at least max(2, ceil(fraction x units)) within eps
	 * of the best) stops.  0 checkpoints: off. */
	int early_stop_checkpoints;
	int early_stop_first;
	double early_stop_fraction; /* 0.5 */
	double early_stop_eps;      /* 0.05 kcal/mol */
	/* QuickVina 2's history check (not smina's search; MCParams::visited,
	 * src/gpu/gpu_mc.cuh:146-264, --gpu_qvina): a local optimization whose
	 * start is not "interesting" to the chain's last 5n visited points is
	 * skipped (n = degrees of freedom, at most 64; larger models keep no
	 * history).  0: off. */
	int visited;
} sgpu_mc_params;

enum { SGPU_MAX_POSES = 64 };

/* ---- chains (ChainInput, src/gpu/gpu_kernels.h:34-43) */
typedef struct sgpu_chains
{
	int nchains;
	const int* model;      /* [nchains] index into the model batch */
	const double* conf;    /* [nchains*conf_stride] initial conf */
	int conf_stride;
	const uint32_t* rng;   /* [nchains*625]: MT words then position */
	const int* num_steps;  /* [nchains] */
	const int* maxiters;   /* [nchains] BFGS iterations */
	const int* unit;       /* [nchains] search unit within its model, for early stopping (NULL: one per chain) */
} sgpu_chains;

/* ---- chain results (ChainResult, src/gpu/gpu_kernels.h:59-64): each
 * chain's output container, sorted by energy */
typedef struct sgpu_chain_results
{
	int max_poses;     /* poses per chain the arrays hold */
	int conf_stride;
	int coord_stride;  /* >= 3 * heavy movable atoms */
	int* count;        /* [nchains] */
	double* e;         /* [nchains*max_poses] */
	double* conf;      /* [nchains*max_poses*conf_stride] */
	double* coords;    /* [nchains*max_poses*coord_stride] heavy movable atoms */
	int64_t* evals;    /* [nchains] eval_deriv calls (may be NULL) */
} sgpu_chain_results;

typedef struct sgpu_ctx sgpu_ctx;

/* context: one per GPU; device = HIP ordinal */
sgpu_status sgpu_create(int device, sgpu_ctx** out);
void sgpu_destroy(sgpu_ctx* ctx);
const char* sgpu_last_error(const sgpu_ctx* ctx);
/* name and compute-unit count of the context's device */
sgpu_status sgpu_device_info(sgpu_ctx* ctx, char* name, int name_len, int* cus);
/* the device's free and total memory (hipMemGetInfo), for admission by
 * memory footprint */
sgpu_status sgpu_device_memory(sgpu_ctx* ctx, size_t* free_bytes, size_t* total_bytes);
/* Monte Carlo chains of this precision the device holds at once (one per
 * wave, as the occupancy calculator gives it from the search kernel's
 * registers): EngineBase::capacity, src/gpu/gpu_docking.cpp:223-226 */
sgpu_status sgpu_mc_capacity(sgpu_ctx* ctx, int precision, int* chains);

/* DeviceTables::upload (src/gpu/gpu_device.cu:71-90) */
sgpu_status sgpu_upload_tables(sgpu_ctx* ctx, const sgpu_tables* t);
/* the spline approximation (--approximation spline, the default of
 * --minimize; DeviceTables::upload's spline branch, src/gpu/gpu_flatten.cpp:
 * 275-325): replaces the context's tables for every later call (maps,
 * evaluation, search, refinement, rescoring).  The splines are evaluated in
 * FP64; the FP32 search rounds the results. */
sgpu_status sgpu_upload_splines(sgpu_ctx* ctx, const sgpu_splines* s);

/* DeviceReceptor::upload (src/gpu/gpu_device.cu:92-110) plus the 3 A cell
 * lists of flatten_receptor (src/gpu/gpu_flatten.cpp:405-445), built on the
 * device; cutoff_sqr as the tables' */
sgpu_status sgpu_upload_receptor(sgpu_ctx* ctx, const sgpu_receptor* r, int* receptor_id);

/* --user_grid (grid::init from a map file, src/lib/grid.cpp:77-108; the
 * reference's FlatUserGrid, src/gpu/gpu_device.h:95-99): a user-supplied
 * map added to the score.  data holds (n0+1)(n1+1)(n2+1) values, x fastest,
 * as grid::init stores them (the file's values times -scaling, the last
 * plane of each axis 0; sgpu_read_user_grid in sgpu_host.h builds them from
 * the file).  Once set, on this context:
 *   - maps built afterwards add it at every point as cache::populate does
 *     (src/lib/cache.cpp:204-205): evaluated at the point's integer indices
 *     taken as a location in A, with map_slope (the cache slope, 1e6) and
 *     cap 1000;
 *   - the direct receptor term adds it at each heavy movable atom with that
 *     evaluation's slope and cap 1000, value and derivative
 *     (src/lib/non_cache.cpp:180-184);
 *   - the final rescoring (sgpu_score mode 0) adds it at every ligand atom
 *     with slope and cap 1000 (model::eval, src/lib/model.cpp:953-964).
 * NULL removes it. */
typedef struct sgpu_user_grid
{
	sgpu_grid_dims gd;   /* the user grid's own box, n intervals per axis */
	const double* data;
	double map_slope;
} sgpu_user_grid;
sgpu_status sgpu_set_user_grid(sgpu_ctx* ctx, const sgpu_user_grid* ug);

/* DeviceGrids::build (src/gpu/gpu_device.cu:199-250) = cache::populate
 * (src/lib/cache.cpp:116-210): one map per listed ligand type, sums in FP64,
 * stored in `precision`.  slope is the cache slope (1e6). */
sgpu_status sgpu_build_grids(sgpu_ctx* ctx, int receptor_id, const sgpu_grid_dims* gd,
		const int* types, int ntypes, int precision, int* grid_id);
/* Another context on the same device uses the maps of src's grid src_grid
 * without a copy: the maps stay resident once per device for every context
 * that serves their receptor (serving: one context per concurrent launch).
 * The memory stays with src, so src is destroyed after every context that
 * shares its grids; releasing the grid in one context leaves the others'
 * use of it intact.  The sharer's own receptor id serves the direct
 * receptor term; the maps carry only what was built into them. */
sgpu_status sgpu_share_grid(sgpu_ctx* dst, sgpu_ctx* src, int src_grid, int* grid_id);
/* copy one type's map back, point layout x fastest, (n0+1)(n1+1)(n2+1) values */
sgpu_status sgpu_download_grid(sgpu_ctx* ctx, int grid_id, int type, double* out);
/* the same for the type's charge map (grid::chargedata, built when the
 * tables carry charge-dependent terms: the values the evaluation scales by
 * the ligand atom's charge, src/lib/cache.cpp:172-187) */
sgpu_status sgpu_download_charge_grid(sgpu_ctx* ctx, int grid_id, int type, double* out);
sgpu_status sgpu_release_grid(sgpu_ctx* ctx, int grid_id);

/* DeviceModels::upload (src/gpu/gpu_device.cu:307-423) */
sgpu_status sgpu_upload_models(sgpu_ctx* ctx, const sgpu_flat_model* models, int nmodels,
		int precision, int* batch_id);
sgpu_status sgpu_release_models(sgpu_ctx* ctx, int batch_id);

/* eval_confs (src/gpu/gpu_kernels.h:22-31): model::eval_deriv for each conf.
 * grid_id >= 0: cache back end (the docking search); grid_id < 0: direct
 * receptor (non_cache) back end in box gd with `slope`.  v: curl caps.
 * Outputs e[nconfs], g[nconfs*change_stride], coords[nconfs*coord_stride]
 * (all movable atoms; may be NULL). */
typedef struct sgpu_eval_request
{
	int batch_id;
	int grid_id;
	int receptor_id;
	const sgpu_grid_dims* gd; /* non_cache box */
	double slope;
	double v[3];
	int nconfs;
	const int* model;
	const double* conf;
	int conf_stride;
	double* e;
	double* g;
	int change_stride;
	double* coords;
	int coord_stride;
} sgpu_eval_request;
sgpu_status sgpu_eval_confs(sgpu_ctx* ctx, const sgpu_eval_request* req);

/* sgpu_eval_confs for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * It has no counterpart in the reference.  Inputs and outputs are the
 * caller's DEVICE arrays in double: model[nconfs], conf[nconfs][conf_stride],
 * e[nconfs], g[nconfs][change_stride] and coords[nconfs][coord_stride]
 * (g and coords may be NULL; entries past a model's own sizes are written as
 * 0).  The strides must cover the batch's largest model.  The staged confs,
 * the evaluation's outputs, the pair scratch and the status word live in one
 * static device array of SGPU_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends: a request needing more is refused (SGPU_E_INVALID).  The
 * call makes no device allocation.  The batch, tables and maps come from
 * earlier calls; for the direct receptor term (grid_id < 0) the receptor's
 * cells for the box must already exist (any sgpu_eval_confs, sgpu_run_refine
 * or sgpu_score call with that box builds them).
 * stream: a hipStream_t, or NULL for the context's stream.  status: a device
 * int that receives the status bits (1: an exact atom overlap, 2: a model
 * index out of range) and the call returns once its work is enqueued; with
 * status NULL the call waits and returns SGPU_E_ATOM_OVERLAP or
 * SGPU_E_INVALID for those.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, threads or contexts) race. */
#define SGPU_NOALLOC_SCRATCH_BYTES (32 * 1024 * 1024)
typedef struct sgpu_eval_request_dev
{
	int batch_id;
	int grid_id;
	int receptor_id;
	const sgpu_grid_dims* gd; /* host pointer: the direct term's box */
	double slope;
	double v[3];
	int nconfs;
	const int* model;
	const double* conf;
	int conf_stride;
	double* e;
	double* g;
	int change_stride;
	double* coords;
	int coord_stride;
	int* status;
	void* stream;
} sgpu_eval_request_dev;
sgpu_status sgpu_eval_confs_noalloc(sgpu_ctx* ctx, const sgpu_eval_request_dev* req);

/* run_monte_carlo (src/gpu/gpu_kernels.h:66-72): one chain per wavefront,
 * monte_carlo::operator() (src/lib/monte_carlo.cpp:95-130) for each chain */
sgpu_status sgpu_run_mc(sgpu_ctx* ctx, int batch_id, int grid_id, const sgpu_mc_params* p,
		const sgpu_chains* chains, sgpu_chain_results* out);

/* The launch-plan autotuner.  It has no counterpart in the reference,
 * whose threads-per-chain policy is a fixed rule
 * (src/gpu/gpu_kernels_impl.cuh:552-605).  A candidate is one plan for
 * every size class of these chains (by pair count) at once.  The first is
 * the default plan (per class the chains per workgroup with the most
 * resident waves per CU, the larger on ties, and the default form); the
 * others take a target of 8, 4, 2 or 1 chains per workgroup (each class
 * the largest count up to the target that its LDS plan and chains per
 * ligand allow) and one form, one-shot grid or persistent grid over a
 * device work queue.  Every chain runs from its start state with its steps
 * scaled by one factor, the longest chain's to probe_steps, under each
 * candidate, the classes concurrently as in a real launch (they share the
 * device, so a class's plan is only good together with the others').  Each
 * probe is timed with HIP events.  A candidate replaces the default only
 * when it is more than 3% faster; the chosen plans are kept by the context
 * for its later sgpu_run_mc calls of the same precision, kernel variant
 * and class.  The chains' trajectories do not depend on the plan.  The
 * report lists, per class, its plan under every candidate with the
 * candidate's time, and the choice. */
#define SGPU_TUNE_MAX_CANDIDATES 8
typedef struct sgpu_mc_tune_report
{
	int nclasses;
	int class_index[4];   /* size class: pairs <= 256, <= 512, <= 1024, more */
	int ncand[4];
	int cand_waves[4][SGPU_TUNE_MAX_CANDIDATES];
	int cand_persist[4][SGPU_TUNE_MAX_CANDIDATES];
	double cand_ms[4][SGPU_TUNE_MAX_CANDIDATES];
	int chosen_waves[4];
	int chosen_persist[4];
} sgpu_mc_tune_report;
sgpu_status sgpu_tune_mc(sgpu_ctx* ctx, int batch_id, int grid_id, const sgpu_mc_params* p,
		const sgpu_chains* chains, int probe_steps, sgpu_mc_tune_report* out);

/* run_refine (src/gpu/gpu_kernels.h:100-110): refine_structure
 * (src/main/main.cpp:165-191) of every pose on the non_cache back end with
 * box gd; conf is updated in place; e = max double if the pose stays out of
 * the box; coords of heavy movable atoms. */
typedef struct sgpu_refine_request
{
	int batch_id;
	int receptor_id;
	const sgpu_grid_dims* gd;
	double v[3];
	int accurate_line_search;
	int early_term;
	int nposes;
	const int* model;
	double* conf;
	int conf_stride;
	const int* maxiters;
	double* e;
	double* coords;
	int coord_stride;
	/* optional per-pose boxes (GpuDocking::MinimizeJob::gd,
	 * src/gpu/gpu_docking.h:101-108): [begin x y z, end x y z] per pose; NULL
	 * = gd for every pose.  Each box must lie inside gd, which spans the
	 * receptor cells. */
	const double* boxes;
	/* optional out [nposes]: the box-slope increases the refinement took
	 * (the last BFGS run used slope 10 x 10^tries; 5: the pose never got
	 * inside the box), or NULL */
	int* tries;
} sgpu_refine_request;
sgpu_status sgpu_run_refine(sgpu_ctx* ctx, const sgpu_refine_request* req);

/* run_bfgs (src/gpu/gpu_kernels.h:104-111): quasi_newton::operator()
 * (bfgs, src/lib/bfgs.h:31-282) on every conf in place, one wavefront each,
 * on the receptor maps (grid_id >= 0, map slope) or on the direct receptor
 * term (grid_id < 0: receptor_id's atoms in box gd, non_cache slope), with
 * caps v; e receives the minimised energies and evals (optional) the
 * evaluations each minimisation took. */
typedef struct sgpu_bfgs_request
{
	int batch_id;
	int grid_id;
	int receptor_id;
	const sgpu_grid_dims* gd;
	double slope;
	double v[3];
	int accurate_line_search;
	int early_term;
	int nconfs;
	const int* model;
	double* conf;
	int conf_stride;
	const int* maxiters;
	double* e;
	int64_t* evals;
} sgpu_bfgs_request;
sgpu_status sgpu_run_bfgs(sgpu_ctx* ctx, const sgpu_bfgs_request* req);

/* final rescoring of finish_docking (src/main/main.cpp:256-324) in FP64:
 * mode 0: model::eval (src/lib/model.cpp:931-953) with the non_cache back
 *         end (slope) and fast tables;
 * mode 1: model::eval_intramolecular (src/lib/model.cpp:1028-1083);
 * mode 2: model::eval_deriv's energy (src/lib/model.cpp:969-983) with the
 *         non_cache back end (slope) and the search's smooth tables: the
 *         function refine_structure minimises, in FP64 (the FP64 value of a
 *         refined pose's energy, for the error of an FP32 search). */
typedef struct sgpu_score_request
{
	int batch_id;
	int receptor_id;
	const sgpu_grid_dims* gd;
	double slope;
	double v[3];
	int mode;
	int nconfs;
	const int* model;
	const double* conf;
	int conf_stride;
	double* e;
} sgpu_score_request;
sgpu_status sgpu_score(sgpu_ctx* ctx, const sgpu_score_request* req);

/* timing of the last launch of each kind, for measurement: milliseconds of
 * the kernel measured with HIP events on the launch stream */
sgpu_status sgpu_last_kernel_ms(sgpu_ctx* ctx, const char* kind, double* ms);

/* Progress of the Monte Carlo search (smina's progress display,
 * src/lib/parallel_progress.h): sgpu_run_mc runs the chains' steps as
 * `chunks` consecutive launches (the chains resume from their device state)
 * and calls fn(user, fraction of steps done) after each, on the calling
 * thread.  chunks 1 (the default) runs one launch; fn may be NULL. */
typedef void (*sgpu_progress_fn)(void* user, double fraction);
sgpu_status sgpu_set_progress(sgpu_ctx* ctx, int chunks, sgpu_progress_fn fn, void* user);

#ifdef __cplusplus
}
#endif

#endif

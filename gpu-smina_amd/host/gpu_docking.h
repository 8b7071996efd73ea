/*
 * GpuDocking: the C++ facade of the docking path, mirroring the reference's
 * class of the same name (src/gpu/gpu_docking.h:51-114) over the sgpu C ABI.
 *
 *   parallel_mc::operator()  -> GpuDocking::dock_batch (search + merge)
 *   refine_structure         -> refine step of dock_batch
 *   finish_docking rescoring -> FP64 device rescoring in dock_batch
 *
 * A smina build keeps its option names: --gpu (FP32 served search),
 * --gpu_fp64 (FP64, bit-identical to the CPU search), --exhaustiveness,
 * --seed, --num_modes, --energy_range, --gpu_batch (ligands per launch).
 */
#ifndef SGPU_GPU_DOCKING_H
#define SGPU_GPU_DOCKING_H

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sgpu.h"
#include "../../include/sgpu_host.h"

namespace sgpu_host
{

// a conf with its energy and heavy-atom coordinates (output_type,
// src/lib/conf.h:400-406)
struct Pose
{
	std::vector<double> c;
	double e = 0;
	std::vector<double> coords;
	// the refinement's energy in the search precision, and the box-slope
	// increases it took (its last BFGS run used slope 10 x 10^tries)
	double e_search = 0;
	int tries = 0;
};

// smina's model after flattening (FlatModel, src/gpu/gpu_flatten_data.h)
struct FlatModel
{
	sgpu_flat_model view{};
	std::vector<int> atom_type, atom_frame, frame_parent, frame_kind, frame_begin, frame_end, frame_tree,
			frame_conf, frame_change, pair_a, pair_b, pair_tri;
	std::vector<double> atom_local, atom_init, frame_origin, frame_axis;
	std::vector<double> atom_charge; // empty: every charge 0
	std::vector<int> tree_ligand, tree_ntors, tree_conf;
	int nheavy = 0;
	double num_tors = 0;
	void publish();
};

FlatModel flatten(const sgpu_model_desc& d);
// model::movable_atoms_box (src/lib/model.cpp:677-707) on the parsed pose
sgpu_grid_dims movable_atoms_box(const FlatModel& m, double add_to_each_dimension, double granularity);
std::vector<double> initial_conf(const FlatModel& m);
// make_chain (src/gpu/gpu_search.cpp:19-31): rng(seed), conf::randomize in
// the box; the generator state (624 words + position) after the draws
void make_chain(const FlatModel& m, uint32_t seed, const sgpu_grid_dims& gd, double* conf, uint32_t* rng_state);

// search parameters of one ligand (setup_search, src/main/main.cpp:513-529)
struct SearchParams
{
	int num_steps = 0;
	int maxiters = 0;
	int num_saved_mins = 20;
	double min_rmsd = 1.0;
	int num_tasks = 8;
};
SearchParams setup_search(const FlatModel& m, int exhaustiveness, int num_modes = 9, int minimize_iters = 0);

// scoring tables: precalculate_linear(sf, factor) (src/lib/precalculate.h:206-237)
struct Tables
{
	int n = 0;
	double factor = 32;
	double cutoff_sqr = 64;
	std::vector<double> smooth, fast;
	// with charge-dependent terms: 4 components, the three charge-dependent
	// ones ([378][n][3][2] and [378][n][3])
	int ncomp = 1;
	std::vector<double> qsmooth, qfast;
};
Tables build_tables(const sgpu_scoring& sf, double factor);
// precalculate_splines(sf, factor) (src/lib/precalculate.h:313-455)
struct Splines
{
	int nseg = 0;
	double fraction = 0;
	double cutoff_sqr = 64;
	std::vector<double> coef, cut; // [378][nseg][5], [378]
};
Splines build_splines(const sgpu_scoring& sf, double factor);

void add_to_output_container(std::vector<Pose>& out, const Pose& t, double min_rmsd, size_t max_size);

// one scoring term between atom types t1 and t2 at distance r, unweighted
// (the term's own eval, src/lib/everything.h), its cutoff and its name as
// smina prints it
// a term's unweighted value for atoms of types t1, t2 and charges q1, q2
// at distance r (charge-dependent terms use the charges)
double term_value(const sgpu_term& t, int t1, int t2, double r, int atom_params, double q1 = 0, double q2 = 0);
bool charge_dependent(const sgpu_term& t);
double term_cutoff(const sgpu_term& t);
std::string term_name(const sgpu_term& t);
// model::movable_atoms_box (src/lib/model.cpp:677-707) of n atoms at xyz
sgpu_grid_dims movable_atoms_box_at(const double* xyz, int n, double add, double granularity);

// --custom_scoring (custom_terms::add_terms_from_file, src/lib/custom_terms.cpp:17-71):
// the file's pairwise terms, charge-independent and charge-dependent, each in
// file order, and its conf-independent terms, weights times scale; throws
// std::invalid_argument naming the line of an unknown or unsupported term
void parse_custom_scoring(const std::string& text, double scale, std::vector<sgpu_term>& ind,
		std::vector<sgpu_term>& dep, std::vector<sgpu_ci_term>& ci);

struct GpuSettings
{
	int device = 0;
	int precision = SGPU_FP32; // SGPU_FP32, SGPU_FP64 (--gpu_fp64) or SGPU_MIXED
	int exhaustiveness = 8;
	int seed = 1;
	int num_modes = 9;
	double energy_range = 3.0;
	// --force_cap: authentic_v of the refinement and the final scoring
	// (main_procedure, src/main/main.cpp:392); the chains keep monte_carlo's
	// own 1000 (src/lib/monte_carlo.cpp:96)
	double force_cap = 1000;
	// --accurate_line: minimization_params::BFGSAccurateLineSearch for
	// every BFGS, the chains' and the refinement's (src/main/main.cpp:1231-1234)
	bool accurate_line = false;
	// --min_rmsd_filter: remove_redundant's RMSD (src/main/main.cpp:276)
	double out_min_rmsd = 1;
	// --minimize_iters (0: setup_search's (25 + movable atoms) / 3) and
	// --minimize_early_term: minimization_params of every BFGS, the chains'
	// and the refinement's (src/main/main.cpp:513-522, :1059-1066)
	int minimize_iters = 0;
	bool early_term = false;
	int ligands_per_launch = 0; // 0: the whole batch in one launch
	// --approximation: LINEAR or SPLINE (the default of --minimize,
	// src/main/main.cpp:1225-1228); factor: points per A^2 (linear) or per
	// A (spline)
	enum { LINEAR, SPLINE };
	int approximation = LINEAR;
	double approx_factor = 32;
	// GPU-style search, not smina's (GpuSettings, src/gpu/gpu_docking.h:23-44).
	// chain_split: every Monte Carlo chain becomes this many chains with
	// 1/chain_split of its steps (0: automatic, auto_split); step_fraction:
	// the fraction of smina's steps (1: smina's search; 0: automatic,
	// auto_fraction); early_stop: stop a ligand once its exhaustiveness
	// units agree on its best energy (EARLY_STOP_FULL: only while the device
	// is full, chain split below 32)
	int chain_split = 1;
	double step_fraction = 1;
	// QuickVina 2's history check (--gpu_qvina): local optimizations that
	// would revisit known minima are skipped
	bool skip_visited = false;
	enum { EARLY_STOP_OFF, EARLY_STOP_FULL, EARLY_STOP_ALWAYS };
	int early_stop = EARLY_STOP_OFF;
	// two-pass screening (src/main/main.cpp:1563-1597): every ligand with
	// fast_first of the steps, then the full search for those within
	// fast_top_margin of the best fast_top of all ligands so far (0: one pass)
	double fast_top = 0;
	double fast_first = 0.0625;
	double fast_top_margin = 0.3;
	// progress of the Monte Carlo search (sgpu_set_progress): its steps in
	// this many launches, progress(progress_user, fraction) after each
	int progress_chunks = 1;
	sgpu_progress_fn progress = nullptr;
	void* progress_user = nullptr;
	// --user_grid: copied to the device when the docking object is made
	// (its maps, direct receptor term and rescoring then carry it)
	const sgpu_user_grid* user_grid = nullptr;
};

// the settings --fast selects (src/main/main.cpp:1451-1461, :1426-1431):
// automatic chain split and step fraction, QuickVina 2's history check, early
// stopping while the device is full, two-pass screening of the best 10%.
void apply_fast(GpuSettings& s);

// the GPU-style search of a launch of `tasks` exhaustiveness units on a
// device holding `capacity` chains: chains per unit, fraction of smina's
// steps, early-stopping checkpoints (0: none; the rule from `first` on).
// budget > 0: a first pass at that fraction, without early stopping.
struct SearchPlan
{
	int split = 1;
	double fraction = 1;
	int checkpoints = 0;
	int first = 2;
};
SearchPlan plan_search(const GpuSettings& s, size_t tasks, int capacity, double budget);
// two-pass screening: ligands whose first-pass score is at most this get the
// full search (the best fast_top of all first-pass scores so far, plus margin)
double full_search_cutoff(std::vector<double> history, double fast_top, double margin);

class GpuDocking
{
public:
	GpuDocking(const GpuSettings& s, const sgpu_scoring& sf, const sgpu_receptor& rec, const sgpu_grid_dims& gd);
	// the same on a device context that the caller owns and shares (serving:
	// several resident receptors per context)
	GpuDocking(const GpuSettings& s, const sgpu_scoring& sf, const sgpu_receptor& rec, const sgpu_grid_dims& gd,
			sgpu_ctx* shared);
	~GpuDocking();
	GpuDocking(const GpuDocking&) = delete;
	GpuDocking& operator=(const GpuDocking&) = delete;

	struct DockJob
	{
		const sgpu_model_desc* m;
		int exhaustiveness = 0; // 0: the settings' value
		std::vector<Pose> out; // final, rescored poses (finish_docking order)
		int64_t evals = 0;
		double seconds = 0;
		// FP32 and mixed searches: the top pose's refined energy in the
		// search precision against the FP64 value of the same function at
		// the same conf (model::eval_deriv on the non_cache back end, bit-
		// equal to the CPU's), |e_search - e_fp64|; -1 for an FP64 search
		// or a job without poses
		double fp64_error = -1;
		double e_search = 0, e_fp64 = 0;
	};
	// for each job: parallel_mc seeded with rng(seed) (do_search), refine of
	// every pose, FP64 rescoring, clustering (src/gpu/gpu_docking.cpp:756-831,
	// src/main/main.cpp:256-324)
	void dock_batch(std::vector<DockJob*>& jobs);
	// the launch-plan autotuner (sgpu_tune_mc) on the chains these jobs'
	// search would run, probe_steps Monte Carlo steps per candidate plan;
	// the context keeps the fastest plan per size class for later searches
	void tune(std::vector<DockJob*>& jobs, int probe_steps, sgpu_mc_tune_report& report);
	// the ligands of the last dock_batch that got the full search of a
	// two-pass screen
	size_t full_searches() const { return full_searches_; }

	// local optimization (--minimize / --local_only), always in FP64:
	// refine_structure of each job's start conf in the job's own box
	// (GpuDocking::minimize_batch, src/gpu/gpu_docking.h:96-109 and
	// src/gpu/gpu_docking.cpp:690-742; box and iterations as main.cpp sets
	// them, src/main/main.cpp:1216-1233, :1688-1700)
	struct MinimizeJob
	{
		const sgpu_model_desc* m;
		sgpu_grid_dims gd{};        // the job's box; n[0] == 0: movable_atoms_box(4, 0.375)
		int maxiters = 10000;       // --minimize_iters (0 in smina's options means 10000)
		bool accurate = true;       // --minimize uses the accurate line search
		std::vector<double> conf;   // in: start conf (empty: the parsed pose); out: refined
		double e = 0;               // energy of the refined conf (maxed if outside the box)
		std::vector<double> coords; // heavy movable atoms after refinement
		double rmsd = 0;            // heavy-atom RMSD from the parsed pose
	};
	// cap: every BFGS's force cap (authentic_v's three components;
	// main.cpp sets 10 for --minimize and keeps --force_cap's 1000 for
	// --local_only, src/main/main.cpp:1216-1220)
	void minimize_batch(std::vector<MinimizeJob*>& jobs, double cap);

	// receptor maps for every heavy-atom type of these models
	void prepare(const std::vector<const sgpu_model_desc*>& models);
	// receptor maps for these ligand atom types
	void prepare_types(const std::vector<int>& types) { ensure_grids(types); }
	// use the maps `store` holds (grid store_grid, covering `types`) instead of
	// building them here: one copy per device for every context serving them
	void adopt_grid(sgpu_ctx* store, int store_grid, const std::vector<int>& types);
	int grid_id() const { return grid_; }
	// device bytes of the maps built so far
	size_t map_bytes() const;
	sgpu_ctx* ctx() { return ctx_; }
	int receptor_id() const { return receptor_; }
	std::string description() const;

private:
	void ensure_grids(const std::vector<int>& types);
	// budget > 0: that fraction of smina's steps, without early stopping
	// (the first pass of a two-pass screen)
	void dock_sub(DockJob* const* jobs, size_t n, double budget);
	void dock_pass(std::vector<DockJob*>& jobs, double budget);
	void fp64_errors(DockJob* const* jobs, size_t nj, int batch, int cstride);
	GpuSettings s_;
	int capacity_ = 1;                // chains the device holds at once
	std::vector<double> top_history_; // two-pass: first-pass top scores so far
	size_t full_searches_ = 0;
	int tune_steps_ = 0;
	sgpu_mc_tune_report* tune_report_ = nullptr; // set during tune()
	sgpu_ctx* ctx_ = nullptr;
	bool owns_ctx_ = true;
	sgpu_scoring sf_;
	std::vector<sgpu_term> terms_;
	std::vector<sgpu_ci_term> ci_;
	sgpu_grid_dims gd_;
	int receptor_ = -1;
	int grid_ = -1;
	std::vector<int> grid_types_;
	bool shared_grid_ = false; // maps adopted from another context
};

} // namespace sgpu_host

#endif

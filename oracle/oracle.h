/*
 * TEST INFRASTRUCTURE ONLY.  The CPU oracle: a plain FP64 C++ restatement of
 * smina's docking search (the CPU code the reference's --gpu path replaces),
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker.  The product path never links or calls it.
 *
 * Each function cites the reference file:line it restates.  Built with
 * -ffp-contract=off and the project's portable sin/cos/exp/acos so that the
 * FP64 device search can be compared with it bit for bit.
 *
 * Parity pinning (tests/test_oracle_*.py): MT19937 against libstdc++'s
 * std::mt19937 and the C++ standard's known answer; uniform draws against
 * their algorithm's range/rejection invariants; portable math against glibc
 * within 1 ulp; terms and tables against closed forms; trees against the
 * zero-conf identity and rigid-motion invariants; eval_deriv gradients
 * against central finite differences; grids against brute-force summation;
 * BFGS against the exact minimum of a convex quadratic; MC against the
 * steps = 1 identity.
 * Parity unpinned: the normal sampler used for initial orientations
 * (Boost's normal_distribution cannot be run here; this oracle and the
 * product share one documented polar-method sampler instead).
 */
#ifndef SGPU_ORACLE_H
#define SGPU_ORACLE_H

#include <stdint.h>
#include "../include/sgpu.h"
#include "../include/sgpu_model.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* scoring + receptor + box; tables at factor 32 (precalculate_linear(wt, 32),
 * src/main/main.cpp:1411-1421) */
orc_ctx* orc_create(const sgpu_scoring* sf, const sgpu_receptor* rec, const sgpu_grid_dims* gd);
void orc_destroy(orc_ctx* c);
/* --user_grid: (n+1)^3 values, x fastest, as grid::init stores them; the
 * maps built afterwards, the non_cache term and model::eval carry it */
void orc_set_user_grid(orc_ctx* c, const sgpu_grid_dims* gd, const double* data, double map_slope);
const char* orc_error(orc_ctx* c);

/* table access: smooth[(tri*n + i)*2 + {0,1}], fast[tri*n + i] */
int orc_table_n(orc_ctx* c);
/* switch the context to precalculate_splines(sf, factor)
 * (src/lib/precalculate.h:313-455): every later evaluation, refinement and
 * map uses the splines.  coef [378][nseg][5] (x, a, b, c, d) and cut [378]
 * receive the tables when not NULL; returns nseg (factor x cutoff), or -1 */
int orc_use_splines(orc_ctx* c, double factor, double* coef, double* cut);
void orc_tables(orc_ctx* c, double* smooth, double* fast);
/* one charge-independent weighted sum of terms (weighted_terms::eval_fast) */
double orc_eval_terms(orc_ctx* c, int t1, int t2, double r);
/* the four weighted components (type only, |q_a|, |q_b|, q_a q_b) */
void orc_eval_components(orc_ctx* c, int t1, int t2, double r, double* out4);
/* weighted_terms::conf_independent: the final score of intermolecular energy e */
double orc_conf_independent(orc_ctx* c, double num_tors, double e);

/* cache::populate for the listed types; maps in point layout (x fastest) */
int orc_build_grids(orc_ctx* c, const int* types, int ntypes);
/* 1, or 4 with charge-dependent terms; their three extra components
 * (qsmooth [378][n][3][2], qfast [378][n][3]) and a type's charge map */
int orc_table_ncomp(orc_ctx* c);
void orc_qtables(orc_ctx* c, double* qsmooth, double* qfast);
int orc_charge_grid_values(orc_ctx* c, int type, double* out);
int orc_grid_values(orc_ctx* c, int type, double* out);

/* models */
int orc_add_model(orc_ctx* c, const sgpu_model_desc* d);
/* info[0..9]: movable atoms, all atoms, frames, pairs, other pairs, conf size,
 * change size, heavy movable atoms, ligands, trees */
void orc_model_info(orc_ctx* c, int model, int* info);
void orc_initial_conf(orc_ctx* c, int model, double* conf);
/* the pairs in evaluation order: a, b, list (0 ligand, 2 other) */
void orc_model_pairs(orc_ctx* c, int model, int* a, int* b, int* list);
double orc_num_tors(orc_ctx* c, int model);

/* model::eval_deriv (src/lib/model.cpp:969-983); backend 0 = cache (grids),
 * 1 = non_cache with `slope`.  coords may be NULL.  Returns e. */
double orc_eval_deriv(orc_ctx* c, int model, int backend, double slope, const double* v,
		const double* conf, double* g, double* coords);
/* model::eval (non_cache, fast tables) and eval_intramolecular */
double orc_eval(orc_ctx* c, int model, double slope, const double* v, const double* conf);
double orc_eval_intramolecular(orc_ctx* c, int model, const double* v, const double* conf);

/* bfgs (src/lib/bfgs.h:201-282) with the given back end; conf in/out;
 * returns f; *evals counts eval_deriv calls */
double orc_bfgs(orc_ctx* c, int model, int backend, double slope, const double* v,
		int maxiters, int accurate, int early_term, double* conf, int64_t* evals);

/* the same bfgs on f(x) = x'Ax/2 - b'x in R^n (x in/out); returns f */
double orc_bfgs_quadratic(int n, const double* A, const double* b, double* x, int maxiters, int accurate,
		int64_t* evals);

/* QuickVina 2's history check (src/gpu/gpu_mc.cuh:146-264): a history of n
 * degrees of freedom fed `count` points in order (xs [count][n], fs, gs
 * [count][n]; the ring keeps the last 5n), then whether (xv, fv, g) is
 * interesting: 1 or 0 */
int orc_visited_check(int n, int count, const double* xs, const double* fs, const double* gs, const double* xv,
		double fv, const double* g);

/* chain initialisation (make_chain / conf::randomize): seeds rng(seed),
 * randomizes a conf in the box, returns the generator state (625 words) */
void orc_randomize(orc_ctx* c, int model, uint32_t seed, double* conf, uint32_t* rng_state);

/* monte_carlo::operator() (src/lib/monte_carlo.cpp:95-130) on the cache back
 * end from a given start conf and generator state.  Output container sorted:
 * count, e[], conf[] (stride conf_size), coords[] (stride 3*heavy). */
int orc_mc_chain(orc_ctx* c, int model, const sgpu_mc_params* p, int num_steps, int maxiters,
		const double* conf0, const uint32_t* rng_state, int max_out, double* e, double* conf,
		double* coords, int64_t* evals);

/* refine_structure (src/main/main.cpp:165-191); conf in/out; returns e */
double orc_refine(orc_ctx* c, int model, const double* v, int maxiters, double* conf, double* coords);
/* refine_structure with the accurate line search when accurate != 0 */
double orc_refine_opts(orc_ctx* c, int model, const double* v, int maxiters, int accurate, double* conf,
		double* coords);

/* the whole docking of one ligand (with early stopping when
 * p->early_stop_checkpoints > 0, units of `split` chains): parallel_mc (src/lib/parallel_mc.cpp:80-98)
 * + refine + finish_docking (src/main/main.cpp:256-324), single-threaded or on
 * `threads` host threads.  With step_fraction < 1 or split > 1 (the GPU-style
 * search, src/gpu/gpu_docking.cpp:200-220): exhaustiveness x split chains,
 * each cut to max(1, ceil(steps x step_fraction / split)) steps.  The chains
 * use p->authentic_v (monte_carlo's own 1000, src/lib/monte_carlo.cpp:96);
 * the refinement and the final scoring use force_cap (main_procedure's
 * authentic_v, src/main/main.cpp:392; <= 0: p->authentic_v);
 * remove_redundant keeps modes out_min_rmsd apart (<= 0: 1); every BFGS
 * runs minimize_iters iterations (0: (25 + movable atoms) / 3).  Results
 * sorted: count, energies, confs. */
int orc_dock(orc_ctx* c, int model, const sgpu_mc_params* p, int seed, int exhaustiveness,
		int num_modes, double energy_range, int threads, double step_fraction, int split, int max_out, double* e,
		double* conf, int64_t* evals, double force_cap, double out_min_rmsd, int minimize_iters);

#ifdef __cplusplus
}
#endif

#endif

/*
 * Inputs of the docking path, as plain arrays: a typed, tree-structured
 * molecule (what smina's PDBQT parser builds into `model`, src/lib/model.h
 * and src/lib/parse_pdbqt.cpp:255-330), a receptor and a scoring function
 * (src/lib/custom_terms.cpp, the default Vina terms of
 * src/main/main.cpp:1347-1353).
 *
 * The host facade (gpu-smina_amd/host) flattens a description into an
 * sgpu_flat_model; the CPU oracle builds its own smina-like model from the
 * same description.
 */
#ifndef SGPU_MODEL_H
#define SGPU_MODEL_H

#ifdef __cplusplus
extern "C" {
#endif

/* A model: ligand(s) plus optional flexible side chains.
 * Atoms 0..nmovable-1 move: each tree's atoms are contiguous, frames of a
 * tree are in preorder and own contiguous atom ranges (smina's atom order).
 * Atoms nmovable..natoms-1 are fixed atoms of flexible residues ("inflex").
 * Frame origins are input positions of an atom: the first root atom for a
 * ligand root; the first atom of the branch (the moving end of the bond) for
 * a segment.  Each segment and side-chain root rotates about the bond from
 * axis_atom (on the parent side) to its origin atom. */
typedef struct sgpu_model_desc
{
	int natoms;
	int nmovable;
	const int* type;       /* [natoms] smina atom types */
	const double* xyz;     /* [3*natoms] */
	int nbonds;
	const int* bonds;      /* [2*nbonds] covalent bonds (atom indices) */
	int nframes;
	const int* frame_parent;      /* [nframes] -1 for a tree root */
	const int* frame_begin;       /* [nframes] */
	const int* frame_end;         /* [nframes] */
	const int* frame_origin_atom; /* [nframes] */
	const int* frame_axis_atom;   /* [nframes] -1 for a ligand root */
	int nligands;                 /* the first nligands trees are ligands */
	int ntrees;
	const int* tree_frame;        /* [ntrees] first frame of each tree */
	const double* charge;         /* [natoms] partial charges (NULL: all 0); only the
	                                 charge-dependent terms read them */
} sgpu_model_desc;

/* Scoring terms (src/lib/everything.h).  Parameters p[] in the order of the
 * term's name, e.g. gauss(o, w, c), repulsion(o, c), hydrophobic(g, b, c),
 * non_hydrophobic(g, b, c), non_dir_h_bond(g, b, c),
 * non_dir_anti_h_bond_quadratic(o, c), donor_donor_quadratic(o, c),
 * acceptor_acceptor_quadratic(o, c), vdw(i, j, s, cap, c),
 * non_dir_h_bond_lj(o, cap, c), electrostatic(i, cap, c),
 * ad4_solvation(d-sigma, s/q, c).  The charge-independent terms are summed
 * first, in list order, then the charge-dependent ones (weighted_terms::
 * eval_fast, src/lib/weighted_terms.cpp:52-62). */
typedef enum sgpu_term_kind
{
	SGPU_TERM_GAUSS = 0,
	SGPU_TERM_REPULSION = 1,
	SGPU_TERM_HYDROPHOBIC = 2,
	SGPU_TERM_NON_HYDROPHOBIC = 3,
	SGPU_TERM_NON_DIR_H_BOND = 4,
	SGPU_TERM_NON_DIR_ANTI_H_BOND_QUADRATIC = 5,
	SGPU_TERM_DONOR_DONOR_QUADRATIC = 6,
	SGPU_TERM_ACCEPTOR_ACCEPTOR_QUADRATIC = 7,
	SGPU_TERM_VDW = 8,
	SGPU_TERM_NON_DIR_H_BOND_LJ = 9,
	/* charge-dependent terms (src/lib/everything.h:64-156): four result
	 * components each, combined with the atoms' charges */
	SGPU_TERM_ELECTROSTATIC = 10,  /* electrostatic(i, cap, c): q_a q_b min(cap, 1 / r^i), i = 1 or 2 */
	SGPU_TERM_AD4_SOLVATION = 11   /* ad4_solvation(d-sigma, s/q, c) */
} sgpu_term_kind;

typedef struct sgpu_term
{
	int kind;
	double p[5];
	double weight;
} sgpu_term;

/* Conf-independent terms (src/lib/everything.h:730-856, :980-1002): the
 * final score x of a pose becomes f(x, num_tors) with weight w:
 *   NUM_TORS_DIV         smooth_div(x, 1 + 0.1 (w + 1) num_tors / 5)
 *   NUM_TORS_ADD         x + w num_tors
 *   NUM_TORS_SQR         x + 0.1 w num_tors^2 / 5
 *   NUM_TORS_SQRT        x + 0.1 w sqrt(num_tors) / sqrt(5)
 *   NUM_TORS_DIV_SIMPLE  smooth_div(x, 1 + w num_tors)
 *   CONSTANT             x + w */
typedef enum sgpu_ci_kind
{
	SGPU_CI_NUM_TORS_DIV = 0,
	SGPU_CI_NUM_TORS_ADD = 1,
	SGPU_CI_NUM_TORS_SQR = 2,
	SGPU_CI_NUM_TORS_SQRT = 3,
	SGPU_CI_NUM_TORS_DIV_SIMPLE = 4,
	SGPU_CI_CONSTANT = 5
} sgpu_ci_kind;

typedef struct sgpu_ci_term
{
	int kind;
	double weight;
} sgpu_ci_term;

/* atom parameter sets: smina's default table, or vinardo's
 * (src/lib/builtinscoring.cpp:6-62, :83) */
typedef enum sgpu_atom_params
{
	SGPU_ATOM_PARAMS_DEFAULT = 0,
	SGPU_ATOM_PARAMS_VINARDO = 1
} sgpu_atom_params;

typedef struct sgpu_scoring
{
	int nterms;
	const sgpu_term* terms; /* charge-independent terms, in weight order */
	int nci;
	const sgpu_ci_term* ci; /* conf-independent terms, applied in this order */
	int atom_params;        /* sgpu_atom_params */
} sgpu_scoring;

#ifdef __cplusplus
}
#endif

#endif

/*
 * The smina atom-type properties the docking search uses: X-Score radius and
 * the hydrophobe / donor / acceptor flags, plus the hydrogen and
 * heteroatom flags.  Values of the reference's default_data table
 * (src/lib/atom_constants.h:150-216, copied into the mutable data[] at start,
 * src/lib/model.cpp:29-44); the enum order is smina_atom_type::type
 * (src/lib/atom_constants.h:92-121).
 */
#ifndef SGPU_ATOM_TYPES_H
#define SGPU_ATOM_TYPES_H

#include "pmath.h"

namespace sgpu
{

enum AtomType
{
	T_Hydrogen = 0, T_PolarHydrogen, T_CAliph_H, T_CAliph_P, T_CArom_H, T_CArom_P,
	T_N_P, T_N_D, T_N_DA, T_N_A, T_O_P, T_O_D, T_O_DA, T_O_A, T_S_P, T_S_A,
	T_P_P, T_F_H, T_Cl_H, T_Br_H, T_I_H, T_Mg, T_Mn, T_Zn, T_Ca, T_Fe, T_Met,
	NUM_ATOM_TYPES
};

struct AtomTypeInfo
{
	double xs_radius;
	bool hydrophobe;
	bool donor;
	bool acceptor;
	bool hydrogen;
	bool heteroatom;
};

SGPU_HD AtomTypeInfo atom_type_info(int t)
{
	// xs_radius, xs hydrophobe, xs donor, xs acceptor, hydrogen, heteroatom
	switch (t)
	{
	case T_Hydrogen: return {0.0, false, false, false, true, false};
	case T_PolarHydrogen: return {0.0, false, false, false, true, false};
	case T_CAliph_H: return {1.9, true, false, false, false, false};
	case T_CAliph_P: return {1.9, false, false, false, false, false};
	case T_CArom_H: return {1.9, true, false, false, false, false};
	case T_CArom_P: return {1.9, false, false, false, false, false};
	case T_N_P: return {1.8, false, false, false, false, true};
	case T_N_D: return {1.8, false, true, false, false, true};
	case T_N_DA: return {1.8, false, true, true, false, true};
	case T_N_A: return {1.8, false, false, true, false, true};
	case T_O_P: return {1.7, false, false, false, false, true};
	case T_O_D: return {1.7, false, true, false, false, true};
	case T_O_DA: return {1.7, false, true, true, false, true};
	case T_O_A: return {1.7, false, false, true, false, true};
	case T_S_P: return {2.0, false, false, false, false, true};
	case T_S_A: return {2.0, false, false, false, false, true};
	case T_P_P: return {2.1, false, false, false, false, true};
	case T_F_H: return {1.5, true, false, false, false, true};
	case T_Cl_H: return {1.8, true, false, false, false, true};
	case T_Br_H: return {2.0, true, false, false, false, true};
	case T_I_H: return {2.2, true, false, false, false, true};
	case T_Mg: return {1.2, false, true, false, false, true};
	case T_Mn: return {1.2, false, true, false, false, true};
	case T_Zn: return {1.2, false, true, false, false, true};
	case T_Ca: return {1.2, false, true, false, false, true};
	case T_Fe: return {1.2, false, true, false, false, true};
	default: return {1.2, false, true, false, false, true}; // generic metal
	}
}

// vinardo's own atom parameters (src/lib/builtinscoring.cpp:6-62, installed
// in place of the default table when vinardo scoring is chosen): other
// X-Score radii, and aromatic non-hydrophobe carbon and sulfur acceptor
// count as hydrophobes
SGPU_HD AtomTypeInfo atom_type_info_vinardo(int t)
{
	// xs_radius, xs hydrophobe, xs donor, xs acceptor, hydrogen, heteroatom
	switch (t)
	{
	case T_Hydrogen: return {0.0, false, false, false, true, false};
	case T_PolarHydrogen: return {0.0, false, false, false, true, false};
	case T_CAliph_H: return {2.0, true, false, false, false, false};
	case T_CAliph_P: return {2.0, false, false, false, false, false};
	case T_CArom_H: return {1.9, true, false, false, false, false};
	case T_CArom_P: return {1.9, true, false, false, false, false};
	case T_N_P: return {1.7, false, false, false, false, true};
	case T_N_D: return {1.7, false, true, false, false, true};
	case T_N_DA: return {1.7, false, true, true, false, true};
	case T_N_A: return {1.7, false, false, true, false, true};
	case T_O_P: return {1.6, false, false, false, false, true};
	case T_O_D: return {1.6, false, true, false, false, true};
	case T_O_DA: return {1.6, false, true, true, false, true};
	case T_O_A: return {1.6, false, false, true, false, true};
	case T_S_P: return {2.0, false, false, false, false, true};
	case T_S_A: return {2.0, true, false, false, false, true};
	case T_P_P: return {2.1, false, false, false, false, true};
	case T_F_H: return {1.5, true, false, false, false, true};
	case T_Cl_H: return {1.8, true, false, false, false, true};
	case T_Br_H: return {2.0, true, false, false, false, true};
	case T_I_H: return {2.2, true, false, false, false, true};
	default: return {1.2, false, true, false, false, true}; // metals
	}
}

// the parameter set of a scoring function (sgpu_scoring::atom_params)
enum AtomParams
{
	ATOM_PARAMS_DEFAULT = 0,
	ATOM_PARAMS_VINARDO = 1
};
SGPU_HD AtomTypeInfo atom_type_info(int t, int params)
{
	return params == ATOM_PARAMS_VINARDO ? atom_type_info_vinardo(t) : atom_type_info(t);
}

// the AutoDock 4 solvation parameter and atomic volume of each type (the
// ad_solvation and ad_volume columns of the default table,
// src/lib/atom_constants.h:150-216), read by ad4_solvation
SGPU_HD double ad_solvation(int t)
{
	switch (t)
	{
	case T_Hydrogen:
	case T_PolarHydrogen: return 0.000510;
	case T_CAliph_H:
	case T_CAliph_P: return -0.001430;
	case T_CArom_H:
	case T_CArom_P: return -0.000520;
	case T_N_P:
	case T_N_D:
	case T_N_DA:
	case T_N_A: return -0.001620;
	case T_O_P:
	case T_O_D:
	case T_O_DA:
	case T_O_A: return -0.002510;
	case T_S_P:
	case T_S_A: return -0.002140;
	default: return -0.001100; // P, halogens, metals
	}
}
SGPU_HD double ad_volume(int t)
{
	switch (t)
	{
	case T_Hydrogen:
	case T_PolarHydrogen: return 0.0;
	case T_CAliph_H:
	case T_CAliph_P:
	case T_CArom_H:
	case T_CArom_P:
	case T_S_P:
	case T_S_A: return 33.510300;
	case T_N_P:
	case T_N_D:
	case T_N_DA:
	case T_N_A: return 22.449300;
	case T_O_P:
	case T_O_D:
	case T_O_DA:
	case T_O_A: return 17.157300;
	case T_P_P: return 38.792400;
	case T_F_H: return 15.448000;
	case T_Cl_H: return 35.823500;
	case T_Br_H: return 42.566100;
	case T_I_H: return 55.058500;
	case T_Mg: return 1.560000;
	case T_Mn: return 2.140000;
	case T_Zn: return 1.700000;
	case T_Ca: return 2.770000;
	case T_Fe: return 1.840000;
	default: return 22.449300; // generic metal
	}
}

SGPU_HD bool is_hydrogen_type(int t) { return t == T_Hydrogen || t == T_PolarHydrogen; }

// triangular_matrix_index (src/lib/triangular_matrix_index.h) of i <= j
SGPU_HD int tri_index(int i, int j) { return i + j * (j + 1) / 2; }
SGPU_HD int tri_index_permissive(int i, int j)
{
	return i <= j ? tri_index(i, j) : tri_index(j, i);
}

} // namespace sgpu

#endif

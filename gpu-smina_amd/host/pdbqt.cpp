// PDBQT input and output for the docking path, without OpenBabel.
//
// Reading follows smina's own PDBQT parser and model set-up:
//   - the record grammar of parse_pdbqt_rigid / parse_pdbqt_ligand_stream
//     (src/lib/parse_pdbqt.cpp:146-176, :179-268, :440-477): atom records,
//     ROOT / ENDROOT, nested BRANCH a b / ENDBRANCH a b, one TORSDOF;
//   - the atom order of postprocess_branch (src/lib/parse_pdbqt.cpp:
//     296-336): a frame's atoms in file order, each followed by the first
//     ("immobile") atoms of its child branches, which turn about the bond but
//     do not move; then the child frames depth first;
//   - branches whose moving atoms are all hydrogens merge into their parent
//     (parsing_struct::mobile_hydrogens_only, src/lib/parsing.h:200-215; the
//     default without --flex_hydrogens);
//   - AutoDock type names to smina types (string_to_smina_type,
//     src/lib/atom_constants.h:294-318);
//   - covalent bonds by distance between atoms whose distance cannot change
//     (model::assign_bonds, src/lib/model.cpp:448-527), then the X-Score
//     types from the bonds (model::assign_types, adjust_smina_type,
//     src/lib/model.cpp:552-566, src/lib/atom_constants.h:347-374).
// Writing follows context::writePDBQT and result_info::write
// (src/lib/model.cpp:710-750, src/lib/result_info.cpp:146-153): the file's
// lines with each atom's coordinates at the pose, one MODEL per pose.

#include <array>
#include <cmath>
#include <cstdio>
#include <functional>
#include <map>
#include <set>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../csrc/atom_types.h"
#include "gpu_docking.h"
#include "mol.h"

namespace sgpu_host
{
using namespace sgpu;
namespace
{

// per smina type (atom_constants.h default_data): AutoDock name, smina
// name, covalent radius, AutoDock heteroatom flag
struct TypeRow
{
	const char* ad;
	const char* name;
	double covalent;
	bool hetero;
};
const TypeRow kTypes[NUM_ATOM_TYPES] = {
		{"H", "Hydrogen", 0.37, false},
		{"HD", "PolarHydrogen", 0.37, false},
		{"C", "AliphaticCarbonXSHydrophobe", 0.77, false},
		{"C", "AliphaticCarbonXSNonHydrophobe", 0.77, false},
		{"A", "AromaticCarbonXSHydrophobe", 0.77, false},
		{"A", "AromaticCarbonXSNonHydrophobe", 0.77, false},
		{"N", "Nitrogen", 0.75, true},
		{"N", "NitrogenXSDonor", 0.75, true},
		{"NA", "NitrogenXSDonorAcceptor", 0.75, true},
		{"NA", "NitrogenXSAcceptor", 0.75, true},
		{"O", "Oxygen", 0.73, true},
		{"O", "OxygenXSDonor", 0.73, true},
		{"OA", "OxygenXSDonorAcceptor", 0.73, true},
		{"OA", "OxygenXSAcceptor", 0.73, true},
		{"S", "Sulfur", 1.02, true},
		{"SA", "SulfurAcceptor", 1.02, true},
		{"P", "Phosphorus", 1.06, true},
		{"F", "Fluorine", 0.71, true},
		{"Cl", "Chlorine", 0.99, true},
		{"Br", "Bromine", 1.14, true},
		{"I", "Iodine", 1.33, true},
		{"Mg", "Magnesium", 1.30, true},
		{"Mn", "Manganese", 1.39, true},
		{"Zn", "Zinc", 1.31, true},
		{"Ca", "Calcium", 1.74, true},
		{"Fe", "Iron", 1.25, true},
		{"M", "GenericMetal", 1.75, true},
};

double max_covalent_radius()
{
	double m = 0;
	for (const TypeRow& r : kTypes)
		m = std::max(m, r.covalent);
	return m;
}

struct ParseError : std::runtime_error
{
	ParseError(int line, const std::string& why) : std::runtime_error("line " + std::to_string(line) + ": " + why) {}
};

bool starts(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }

std::string trim(const std::string& s)
{
	size_t a = 0, b = s.size();
	while (a < b && std::isspace(static_cast<unsigned char>(s[a])))
		++a;
	while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1])))
		--b;
	return s.substr(a, b - a);
}

// string_to_smina_type: a name of up to 2 characters is an AutoDock type
// (the first table row with that name; Se reads as S; other metals as the
// generic metal), a longer one a smina type name; -1 if unknown
int type_of_name(const std::string& name)
{
	if (name.size() <= 2)
	{
		for (int t = 0; t < NUM_ATOM_TYPES; t++)
			if (name == kTypes[t].ad)
				return t;
		if (name == "Se")
			return type_of_name("S");
		static const char* metals[] = {"Cu", "Fe", "Na", "K", "Hg", "Co", "U", "Cd", "Ni"};
		for (const char* m : metals)
			if (name == m)
				return T_Met;
		return -1;
	}
	for (int t = 0; t < NUM_ATOM_TYPES; t++)
		if (name == kTypes[t].name)
			return t;
	return -1;
}

struct Atom
{
	int number = 0;
	int type = 0;
	double xyz[3] = {0, 0, 0};
	double charge = 0;
	int line = 0; // index into the file's lines
};

// columns (1-based, inclusive) of an atom record, blank-skipping like
// checked_convert_substring
double column_number(const std::string& s, size_t i, size_t j, int line, const char* what)
{
	if (j > s.size())
		throw ParseError(line, "ATOM syntax incorrect: the line is too short");
	while (i <= j && std::isspace(static_cast<unsigned char>(s[i - 1])))
		++i;
	const std::string sub = s.substr(i - 1, j - i + 1);
	char* end = nullptr;
	const double v = std::strtod(sub.c_str(), &end);
	if (sub.empty() || end == sub.c_str() || !trim(std::string(end)).empty())
		throw ParseError(line, std::string("ATOM syntax incorrect: \"") + sub + "\" is not a valid " + what);
	return v;
}

// parse_pdbqt_atom_string (src/lib/parse_pdbqt.cpp:103-122); idx: the
// record's 0-based line index
Atom parse_atom(const std::string& s, int idx)
{
	const int line = idx + 1;
	Atom a;
	const double number = column_number(s, 7, 11, line, "atom number");
	if (number < 0 || number != std::floor(number))
		throw ParseError(line, "ATOM syntax incorrect: bad atom number");
	a.number = int(number);
	for (int k = 0; k < 3; k++)
		a.xyz[k] = column_number(s, 31 + 8 * k, 38 + 8 * k, line, "coordinate");
	// the partial charge, columns 69-76 unless blank
	const std::string q = s.size() >= 69 ? s.substr(68, std::min<size_t>(8, s.size() - 68)) : std::string();
	if (!trim(q).empty())
		a.charge = column_number(s, 69, 76, line, "charge");
	if (s.size() < 78)
		throw ParseError(line, "ATOM syntax incorrect: no atom type");
	const std::string name = trim(s.substr(77));
	a.type = type_of_name(name);
	if (a.type < 0)
		throw ParseError(line, "\"" + name + "\" is not a valid AutoDock type (types are case-sensitive)");
	a.line = idx;
	return a;
}

bool ignorable(const std::string& s)
{
	return s.empty() || starts(s, "WARNING") || starts(s, "REMARK") || starts(s, "USER") || starts(s, "TER");
}

// parsing_struct: one frame level of the file
struct Level;
struct Node
{
	Atom a;
	std::vector<Level> children;
};
struct Level
{
	std::vector<Node> atoms;
	int immobile = -1; // the branch's first atom (BRANCH a b: atom b)
	bool essentially_empty() const
	{
		for (size_t i = 0; i < atoms.size(); i++)
			if (int(i) != immobile || !atoms[i].children.empty())
				return false;
		return true;
	}
	bool mobile_hydrogens_only() const
	{
		for (size_t i = 0; i < atoms.size(); i++)
		{
			if (!atoms[i].children.empty())
				return false;
			if (int(i) != immobile && !(atoms[i].a.type == T_Hydrogen || atoms[i].a.type == T_PolarHydrogen))
				return false;
		}
		return true;
	}
};

struct Reader
{
	std::vector<std::string> lines;
	size_t pos = 0;
	bool next(std::string& s)
	{
		if (pos >= lines.size())
			return false;
		s = lines[pos++];
		return true;
	}
	int line() const { return int(pos); } // 1-based number of the line just read
};

std::vector<std::string> split_lines(const char* text)
{
	std::vector<std::string> out;
	std::string cur;
	for (const char* p = text; *p; ++p)
	{
		if (*p == '\n')
		{
			if (!cur.empty() && cur.back() == '\r')
				cur.pop_back();
			out.push_back(cur);
			cur.clear();
		}
		else
			cur.push_back(*p);
	}
	if (!cur.empty())
		out.push_back(cur);
	return out;
}

void two_numbers(const std::string& s, const char* tag, int line, int& a, int& b)
{
	std::istringstream in(s.substr(std::strlen(tag)));
	long x = -1, y = -1;
	in >> x >> y;
	if (!in || x < 0 || y < 0)
		throw ParseError(line, "syntax error");
	a = int(x);
	b = int(y);
}

void parse_branch(Reader& r, Level& lv, int from, int to);

// BRANCH a b inside level lv (parse_pdbqt_branch_aux): the branch hangs
// from lv's atom numbered a, or merges into lv if only hydrogens move
void branch_into(Reader& r, const std::string& s, Level& lv)
{
	int from, to;
	two_numbers(s, "BRANCH", r.line(), from, to);
	for (Node& nd : lv.atoms)
		if (nd.a.number == from)
		{
			Level b;
			parse_branch(r, b, from, to);
			if (b.mobile_hydrogens_only())
				for (const Node& m : b.atoms)
					lv.atoms.push_back(Node{m.a, {}});
			else
				nd.children.push_back(std::move(b));
			return;
		}
	throw ParseError(r.line(), "no atom number " + std::to_string(from) + " in this branch");
}

void parse_branch(Reader& r, Level& lv, int from, int to)
{
	std::string s;
	while (r.next(s))
	{
		if (ignorable(s) && !starts(s, "TER"))
			continue;
		if (starts(s, "BRANCH"))
			branch_into(r, s, lv);
		else if (starts(s, "ENDBRANCH"))
		{
			int a, b;
			two_numbers(s, "ENDBRANCH", r.line(), a, b);
			if (a != from || b != to)
				throw ParseError(r.line(), "inconsistent branch numbers");
			if (lv.immobile < 0)
				throw ParseError(r.line(), "atom " + std::to_string(to) + " has not been found in this branch");
			return;
		}
		else if (starts(s, "ATOM  ") || starts(s, "HETATM"))
		{
			Atom a = parse_atom(s, int(r.pos) - 1);
			if (a.number == to)
				lv.immobile = int(lv.atoms.size());
			lv.atoms.push_back(Node{a, {}});
		}
		else if (starts(s, "MODEL"))
			throw ParseError(r.line(), "unexpected multi-MODEL input");
		else
			throw ParseError(r.line(), "unknown or inappropriate tag");
	}
	throw ParseError(r.line(), "missing ENDBRANCH");
}

// mobility of a pair (distance_type): can the distance change?
enum Mobility : unsigned char
{
	VARIABLE = 0,
	FIXED = 1,
	ROTOR = 2
};

// the model under construction: atoms in smina's order, frames, mobility
struct Builder
{
	std::vector<Atom> atoms;
	std::vector<int> frame_parent, frame_begin, frame_end, frame_origin, frame_axis;
	std::vector<std::vector<unsigned char>> mob;

	int push(const Atom& a)
	{
		atoms.push_back(a);
		return int(atoms.size()) - 1;
	}
	void grow()
	{
		const size_t n = atoms.size();
		for (auto& row : mob)
			row.resize(n, VARIABLE);
		mob.resize(n, std::vector<unsigned char>(n, VARIABLE));
	}
	void set(int i, int j, Mobility m)
	{
		mob[size_t(i)][size_t(j)] = m;
		mob[size_t(j)][size_t(i)] = m;
	}

	// postprocess_branch: this level's atoms (skipping its immobile atom,
	// already placed in the parent frame), each followed by the immobile
	// atoms of its children; axis_begin / axis_end are this frame's bond
	// atoms in the parent frame (-1 for a root)
	void frame(Level& lv, int parent, int axis_begin, int axis_end, int origin)
	{
		const int f = int(frame_parent.size());
		frame_parent.push_back(parent);
		frame_begin.push_back(int(atoms.size()));
		frame_end.push_back(0);
		frame_origin.push_back(origin);
		frame_axis.push_back(axis_begin);
		std::vector<std::vector<int>> child_axis(lv.atoms.size()); // per node: its children's immobile atoms
		std::vector<int> node_index(lv.atoms.size(), -1);
		for (size_t i = 0; i < lv.atoms.size(); i++)
		{
			Node& nd = lv.atoms[i];
			if (int(i) != lv.immobile)
				node_index[i] = push(nd.a);
			for (Level& c : nd.children)
				child_axis[i].push_back(c.atoms.empty() ? -1 : push(c.atoms[size_t(c.immobile)].a));
		}
		frame_end[size_t(f)] = int(atoms.size());
		grow();
		const int b = frame_begin[size_t(f)], e = frame_end[size_t(f)];
		for (int i = b; i < e; i++)
		{
			if (axis_begin >= 0)
				set(axis_begin, i, FIXED);
			if (axis_end >= 0)
				set(axis_end, i, FIXED);
		}
		if (axis_begin >= 0 && axis_end >= 0)
			set(axis_begin, axis_end, ROTOR);
		for (int i = b; i < e; i++)
			for (int j = i + 1; j < e; j++)
				set(i, j, FIXED);
		for (size_t i = 0; i < lv.atoms.size(); i++)
		{
			Node& nd = lv.atoms[i];
			// the parent atom of a child branch: this node's own index, or for
			// the immobile atom (placed by the parent) the axis end
			const int here = int(i) == lv.immobile ? axis_end : node_index[i];
			for (size_t k = 0; k < nd.children.size(); k++)
			{
				Level& c = nd.children[k];
				if (!c.essentially_empty())
					frame(c, f, here, child_axis[i][k], child_axis[i][k]);
			}
		}
	}
};

// model::assign_bonds over atoms whose mutual distance cannot change:
// bonded if closer than 1.1 x the sum of covalent radii and no third atom,
// immobile relative to both and within reach of the first, is closer to
// both than they are to each other
std::vector<int> assign_bonds(const std::vector<Atom>& atoms,
		const std::function<bool(int, int)>& fixed_pair, const std::vector<std::vector<int>>* neighbors)
{
	const double allowance = 1.1, maxr = max_covalent_radius();
	const int n = int(atoms.size());
	auto d2 = [&](int i, int j) {
		double s = 0;
		for (int k = 0; k < 3; k++)
			s += (atoms[size_t(i)].xyz[k] - atoms[size_t(j)].xyz[k]) * (atoms[size_t(i)].xyz[k] - atoms[size_t(j)].xyz[k]);
		return s;
	};
	std::vector<int> bonds;
	std::vector<int> relevant;
	for (int i = 0; i < n; i++)
	{
		const double ri = kTypes[atoms[size_t(i)].type].covalent;
		const double reach = allowance * (ri + maxr);
		relevant.clear();
		auto consider = [&](int j) {
			if (j != i && fixed_pair(i, j) && d2(i, j) < reach * reach)
				relevant.push_back(j);
		};
		if (neighbors)
			for (int j : (*neighbors)[size_t(i)])
				consider(j);
		else
			for (int j = 0; j < n; j++)
				consider(j);
		for (int j : relevant)
		{
			if (j <= i)
				continue;
			const double len = allowance * (ri + kTypes[atoms[size_t(j)].type].covalent);
			const double r2 = d2(i, j);
			if (!(r2 < len * len))
				continue;
			bool between = false;
			for (int c : relevant)
				if (c != j && fixed_pair(j, c) && d2(i, c) < r2 && d2(j, c) < r2)
				{
					between = true;
					break;
				}
			if (!between)
			{
				bonds.push_back(i);
				bonds.push_back(j);
			}
		}
	}
	return bonds;
}

// adjust_smina_type from the bonds: carbon is polar next to a heteroatom,
// nitrogen and oxygen donate with a polar hydrogen
void assign_types(std::vector<Atom>& atoms, const std::vector<int>& bonds)
{
	const size_t n = atoms.size();
	std::vector<char> to_hd(n, 0), to_hetero(n, 0);
	for (size_t k = 0; k + 1 < bonds.size(); k += 2)
	{
		const int a = bonds[k], b = bonds[k + 1];
		for (int pass = 0; pass < 2; pass++)
		{
			const int x = pass ? b : a, y = pass ? a : b;
			if (atoms[size_t(y)].type == T_PolarHydrogen)
				to_hd[size_t(x)] = 1;
			if (kTypes[atoms[size_t(y)].type].hetero)
				to_hetero[size_t(x)] = 1;
		}
	}
	for (size_t i = 0; i < n; i++)
	{
		int& t = atoms[i].type;
		const bool hd = to_hd[i], het = to_hetero[i];
		switch (t)
		{
		case T_CAliph_H:
		case T_CAliph_P: t = het ? T_CAliph_P : T_CAliph_H; break;
		case T_CArom_H:
		case T_CArom_P: t = het ? T_CArom_P : T_CArom_H; break;
		case T_N_D:
		case T_N_P: t = hd ? T_N_D : T_N_P; break;
		case T_N_DA:
		case T_N_A: t = hd ? T_N_DA : T_N_A; break;
		case T_O_D:
		case T_O_P: t = hd ? T_O_D : T_O_P; break;
		case T_O_DA:
		case T_O_A: t = hd ? T_O_DA : T_O_A; break;
		default: break;
		}
	}
}

sgpu_mol* finish(std::vector<Atom>& atoms, const std::vector<int>& bonds, std::vector<std::string>& lines)
{
	assign_types(atoms, bonds);
	sgpu_mol* m = new sgpu_mol;
	for (const Atom& a : atoms)
	{
		m->type.push_back(a.type);
		m->xyz.insert(m->xyz.end(), a.xyz, a.xyz + 3);
		m->charge.push_back(a.charge);
	}
	m->bonds = bonds;
	m->text = std::move(lines);
	m->text_atom.assign(m->text.size(), -1);
	for (size_t i = 0; i < atoms.size(); i++)
		m->text_atom[size_t(atoms[i].line)] = int(i);
	return m;
}

sgpu_mol* read_ligand(const char* text)
{
	Reader r;
	r.lines = split_lines(text);
	std::string s;
	Level root;
	bool in_root = false;
	// parse_pdbqt_root and its _aux: up to ROOT, then its atoms
	while (r.next(s))
	{
		if (ignorable(s))
			continue;
		if (starts(s, "ROOT"))
		{
			in_root = true;
			break;
		}
		if (starts(s, "MODEL"))
			throw ParseError(r.line(), "unexpected multi-MODEL input");
		throw ParseError(r.line(), "unknown or inappropriate tag");
	}
	if (!in_root)
		throw ParseError(r.line(), "no ROOT");
	bool end_root = false;
	while (r.next(s))
	{
		if (ignorable(s))
			continue;
		if (starts(s, "ATOM  ") || starts(s, "HETATM"))
			root.atoms.push_back(Node{parse_atom(s, int(r.pos) - 1), {}});
		else if (starts(s, "ENDROOT"))
		{
			end_root = true;
			break;
		}
		else if (starts(s, "MODEL"))
			throw ParseError(r.line(), "unexpected multi-MODEL input");
		else
			throw ParseError(r.line(), "unknown or inappropriate tag");
	}
	if (!end_root)
		throw ParseError(r.line(), "missing ENDROOT");
	// parse_pdbqt_aux: branches of the root and one TORSDOF
	bool torsdof = false;
	while (r.next(s))
	{
		if (ignorable(s))
			continue;
		if (starts(s, "BRANCH"))
			branch_into(r, s, root);
		else if (starts(s, "TORSDOF"))
		{
			if (torsdof)
				throw ParseError(r.line(), "TORSDOF can occur only once");
			std::istringstream in(s.substr(7));
			long v = -1;
			in >> v;
			if (!in || v < 0)
				throw ParseError(r.line(), "syntax error");
			torsdof = true;
		}
		else if (starts(s, "MODEL"))
			throw ParseError(r.line(), "unexpected multi-MODEL input");
		else
			throw ParseError(r.line(), "unknown or inappropriate tag");
	}
	if (root.atoms.empty())
		throw ParseError(r.line(), "no atoms in the ligand");
	if (!torsdof)
		throw ParseError(r.line(), "missing TORSDOF");
	// postprocess_ligand: the root frame's origin is its first atom
	Builder b;
	b.frame(root, -1, -1, -1, 0);
	b.grow();
	const auto fixed = [&](int i, int j) { return b.mob[size_t(i)][size_t(j)] != VARIABLE; };
	std::vector<int> bonds = assign_bonds(b.atoms, fixed, nullptr);
	sgpu_mol* m = finish(b.atoms, bonds, r.lines);
	m->nmovable = m->natoms();
	m->frame_parent = b.frame_parent;
	m->frame_begin = b.frame_begin;
	m->frame_end = b.frame_end;
	m->frame_origin_atom = b.frame_origin;
	m->frame_axis_atom = b.frame_axis;
	m->tree_frame = {0};
	m->nligands = 1;
	m->publish();
	return m;
}

sgpu_mol* read_receptor(const char* text)
{
	std::vector<std::string> lines = split_lines(text);
	std::vector<Atom> atoms;
	for (size_t i = 0; i < lines.size(); i++)
	{
		const std::string& s = lines[i];
		if (starts(s, "ATOM  ") || starts(s, "HETATM"))
			atoms.push_back(parse_atom(s, int(i)));
		else if (starts(s, "MODEL"))
			throw ParseError(int(i) + 1, "unexpected multi-MODEL input");
	}
	// every receptor pair is fixed; neighbours from 4 A cells (the bond
	// reach is at most 1.1 x 2 x 1.75 A)
	const int n = int(atoms.size());
	std::vector<std::vector<int>> nb(static_cast<size_t>(n));
	if (n > 0)
	{
		const double cell = 4.0;
		double lo[3] = {atoms[0].xyz[0], atoms[0].xyz[1], atoms[0].xyz[2]};
		for (const Atom& a : atoms)
			for (int k = 0; k < 3; k++)
				lo[k] = std::min(lo[k], a.xyz[k]);
		std::vector<std::array<int, 3>> key(static_cast<size_t>(n));
		std::map<std::array<int, 3>, std::vector<int>> grid;
		for (int i = 0; i < n; i++)
		{
			for (int k = 0; k < 3; k++)
				key[size_t(i)][k] = int(std::floor((atoms[size_t(i)].xyz[k] - lo[k]) / cell));
			grid[key[size_t(i)]].push_back(i);
		}
		for (int i = 0; i < n; i++)
			for (int dx = -1; dx <= 1; dx++)
				for (int dy = -1; dy <= 1; dy++)
					for (int dz = -1; dz <= 1; dz++)
					{
						const std::array<int, 3> c = {key[size_t(i)][0] + dx, key[size_t(i)][1] + dy, key[size_t(i)][2] + dz};
						auto it = grid.find(c);
						if (it != grid.end())
							for (int j : it->second)
								nb[size_t(i)].push_back(j);
					}
	}
	std::vector<int> bonds = assign_bonds(atoms, [](int, int) { return true; }, &nb);
	sgpu_mol* m = finish(atoms, bonds, lines);
	m->nmovable = 0;
	m->publish();
	return m;
}

// Flexible residues (parse_pdbqt_flex and postprocess_residue,
// src/lib/parse_pdbqt.cpp:278-294, :413-437): BEGIN_RES ... END_RES blocks,
// each a ROOT and its branches.  A residue's root atoms and the first atoms
// of their branches are fixed ("inflex"); each branch of a root atom is a
// side-chain tree whose first segment turns about the bond from the root
// atom to the branch's first atom.  References: >= 0 a movable atom of the
// side chains, <= -2 the fixed atom -2 - ref.
struct FlexBuilder
{
	std::vector<Atom> mov, fixed;
	std::vector<int> frame_parent, frame_begin, frame_end, frame_origin, frame_axis, tree_frame;
	std::vector<std::pair<int, int>> tied; // reference pairs whose distance cannot change

	int fixed_ref(const Atom& a)
	{
		fixed.push_back(a);
		return -2 - (int(fixed.size()) - 1);
	}
	int push(const Atom& a)
	{
		mov.push_back(a);
		return int(mov.size()) - 1;
	}
	// as Builder::frame, with references
	void frame(Level& lv, int parent, int axis_begin, int axis_end)
	{
		const int f = int(frame_parent.size());
		frame_parent.push_back(parent);
		frame_begin.push_back(int(mov.size()));
		frame_end.push_back(0);
		frame_origin.push_back(axis_end);
		frame_axis.push_back(axis_begin);
		std::vector<std::vector<int>> child_axis(lv.atoms.size());
		std::vector<int> node_index(lv.atoms.size(), -1);
		for (size_t i = 0; i < lv.atoms.size(); i++)
		{
			Node& nd = lv.atoms[i];
			if (int(i) != lv.immobile)
				node_index[i] = push(nd.a);
			for (Level& c : nd.children)
				child_axis[i].push_back(push(c.atoms[size_t(c.immobile)].a));
		}
		frame_end[size_t(f)] = int(mov.size());
		for (int i = frame_begin[size_t(f)]; i < frame_end[size_t(f)]; i++)
		{
			tied.push_back({axis_begin, i});
			tied.push_back({axis_end, i});
		}
		tied.push_back({axis_begin, axis_end});
		for (size_t i = 0; i < lv.atoms.size(); i++)
		{
			Node& nd = lv.atoms[i];
			const int here = int(i) == lv.immobile ? axis_end : node_index[i];
			for (size_t k = 0; k < nd.children.size(); k++)
				if (!nd.children[k].essentially_empty())
					frame(nd.children[k], f, here, child_axis[i][k]);
		}
	}
	void residue(Level& root)
	{
		std::vector<int> node_ref(root.atoms.size());
		std::vector<std::vector<int>> child_ref(root.atoms.size());
		for (size_t i = 0; i < root.atoms.size(); i++)
		{
			node_ref[i] = fixed_ref(root.atoms[i].a);
			for (Level& c : root.atoms[i].children)
				child_ref[i].push_back(fixed_ref(c.atoms[size_t(c.immobile)].a));
		}
		for (size_t i = 0; i < root.atoms.size(); i++)
			for (size_t k = 0; k < root.atoms[i].children.size(); k++)
			{
				Level& c = root.atoms[i].children[k];
				if (c.essentially_empty())
					continue;
				tree_frame.push_back(int(frame_parent.size()));
				frame(c, -1, node_ref[i], child_ref[i][k]);
			}
	}
};

// parse_pdbqt_aux for a residue: ROOT ... ENDROOT, then branches up to
// END_RES (no TORSDOF)
void parse_residue(Reader& r, Level& root)
{
	std::string s;
	bool in_root = false;
	while (r.next(s))
	{
		if (ignorable(s))
			continue;
		if (starts(s, "ROOT"))
		{
			in_root = true;
			break;
		}
		throw ParseError(r.line(), "unknown or inappropriate tag");
	}
	if (!in_root)
		throw ParseError(r.line(), "no ROOT in the residue");
	bool end_root = false;
	while (r.next(s))
	{
		if (ignorable(s))
			continue;
		if (starts(s, "ATOM  ") || starts(s, "HETATM"))
			root.atoms.push_back(Node{parse_atom(s, int(r.pos) - 1), {}});
		else if (starts(s, "ENDROOT"))
		{
			end_root = true;
			break;
		}
		else
			throw ParseError(r.line(), "unknown or inappropriate tag");
	}
	if (!end_root)
		throw ParseError(r.line(), "missing ENDROOT");
	while (r.next(s))
	{
		if (ignorable(s))
			continue;
		if (starts(s, "BRANCH"))
			branch_into(r, s, root);
		else if (starts(s, "END_RES"))
			return;
		else
			throw ParseError(r.line(), "unknown or inappropriate tag");
	}
	throw ParseError(r.line(), "missing END_RES");
}

// the ligand plus the side chains of a flexible-residue file, in the layout
// of the model description: the ligand's atoms and trees, the side chains'
// movable atoms and trees, then the fixed atoms.  Bonds of the side chains
// by distance (model::assign_bonds): within a frame, a frame and its axis
// atoms, and among the fixed atoms and the rigid receptor's atoms near them
// (the fixed atoms' types see the receptor).
sgpu_mol* read_flex(const sgpu_mol& lig, const sgpu_mol& rigid, const char* text)
{
	Reader r;
	r.lines = split_lines(text);
	FlexBuilder fb;
	std::string s;
	while (r.next(s))
	{
		if (s.empty() || starts(s, "WARNING") || starts(s, "REMARK") || starts(s, "USER"))
			continue;
		if (starts(s, "BEGIN_RES"))
		{
			Level root;
			parse_residue(r, root);
			fb.residue(root);
		}
		else if (starts(s, "MODEL"))
			throw ParseError(r.line(), "unexpected multi-MODEL input");
		else
			throw ParseError(r.line(), "unknown or inappropriate tag");
	}
	if (fb.tree_frame.empty())
		throw ParseError(r.line(), "no flexible side chain");
	const int M = int(fb.mov.size()), I = int(fb.fixed.size());
	// receptor atoms within 4 A of a fixed atom
	std::vector<Atom> atoms = fb.mov;
	atoms.insert(atoms.end(), fb.fixed.begin(), fb.fixed.end());
	const int nrig = rigid.natoms();
	for (int a = 0; a < nrig; a++)
	{
		bool near = false;
		for (const Atom& x : fb.fixed)
		{
			double d2 = 0;
			for (int k = 0; k < 3; k++)
				d2 += (rigid.xyz[3 * size_t(a) + size_t(k)] - x.xyz[k]) * (rigid.xyz[3 * size_t(a) + size_t(k)] - x.xyz[k]);
			near = near || d2 < 16.0;
		}
		if (near)
		{
			Atom ra;
			ra.type = rigid.type[size_t(a)];
			for (int k = 0; k < 3; k++)
				ra.xyz[k] = rigid.xyz[3 * size_t(a) + size_t(k)];
			ra.line = -1;
			atoms.push_back(ra);
		}
	}
	auto local = [&](int ref) { return ref >= 0 ? ref : M + (-2 - ref); };
	std::vector<int> frame_of(size_t(M), -1);
	for (size_t f = 0; f < fb.frame_parent.size(); f++)
		for (int a = fb.frame_begin[f]; a < fb.frame_end[f]; a++)
			frame_of[size_t(a)] = int(f);
	std::set<std::pair<int, int>> tied;
	for (auto& t : fb.tied)
	{
		const int a = local(t.first), b = local(t.second);
		tied.insert({std::min(a, b), std::max(a, b)});
	}
	const auto fixed_pair = [&](int i, int j) {
		const bool mi = i < M, mj = j < M;
		if (!mi && !mj)
			return true; // fixed-fixed, fixed-receptor, receptor-receptor
		if (mi && mj && frame_of[size_t(i)] == frame_of[size_t(j)])
			return true;
		if ((mi && j >= M + I) || (mj && i >= M + I))
			return false;
		return tied.count({std::min(i, j), std::max(i, j)}) > 0;
	};
	std::vector<int> all_bonds = assign_bonds(atoms, fixed_pair, nullptr);
	assign_types(atoms, all_bonds);
	sgpu_mol* m = new sgpu_mol(lig);
	m->text.clear();
	m->text_atom.clear();
	const int L = lig.natoms();
	auto final_index = [&](int loc) { return L + loc; }; // movable then fixed, in local order
	if (m->charge.size() != size_t(L))
		m->charge.assign(size_t(L), 0.0);
	for (int a = 0; a < M + I; a++)
	{
		m->type.push_back(atoms[size_t(a)].type);
		m->xyz.insert(m->xyz.end(), atoms[size_t(a)].xyz, atoms[size_t(a)].xyz + 3);
		m->charge.push_back(atoms[size_t(a)].charge);
	}
	for (size_t k = 0; k + 1 < all_bonds.size(); k += 2)
		if (all_bonds[k] < M + I && all_bonds[k + 1] < M + I)
		{
			m->bonds.push_back(final_index(all_bonds[k]));
			m->bonds.push_back(final_index(all_bonds[k + 1]));
		}
	const int f0 = int(lig.frame_parent.size());
	for (size_t f = 0; f < fb.frame_parent.size(); f++)
	{
		m->frame_parent.push_back(fb.frame_parent[f] < 0 ? -1 : f0 + fb.frame_parent[f]);
		m->frame_begin.push_back(L + fb.frame_begin[f]);
		m->frame_end.push_back(L + fb.frame_end[f]);
		m->frame_origin_atom.push_back(final_index(local(fb.frame_origin[f])));
		m->frame_axis_atom.push_back(final_index(local(fb.frame_axis[f])));
	}
	for (int t : fb.tree_frame)
		m->tree_frame.push_back(f0 + t);
	m->nmovable = L + M;
	m->publish();
	return m;
}

// set_conf (src/lib/tree.h:84-195) on the host: the movable atoms at conf
std::vector<double> pose_coords(const FlatModel& f, const double* conf)
{
	const int nf = int(f.frame_parent.size());
	std::vector<double> O(3 * size_t(nf)), R(9 * size_t(nf)), Q(4 * size_t(nf));
	auto to_r3 = [](const double* q, double* r) {
		const double a = q[0], b = q[1], c = q[2], d = q[3];
		r[0] = a * a + b * b - c * c - d * d;
		r[1] = 2 * (b * c - a * d);
		r[2] = 2 * (a * c + b * d);
		r[3] = 2 * (b * c + a * d);
		r[4] = a * a - b * b + c * c - d * d;
		r[5] = 2 * (c * d - a * b);
		r[6] = 2 * (b * d - a * c);
		r[7] = 2 * (a * b + c * d);
		r[8] = a * a - b * b - c * c + d * d;
	};
	auto normalize = [](double* q) {
		const double s = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
		for (int k = 0; k < 4; k++)
			q[k] /= s;
	};
	std::vector<double> x(f.atom_local.size());
	for (int fr = 0; fr < nf; fr++)
	{
		double* o = &O[3 * size_t(fr)];
		double* q = &Q[4 * size_t(fr)];
		const int ci = f.frame_conf[size_t(fr)];
		const int kind = f.frame_kind[size_t(fr)];
		if (kind == SGPU_FRAME_RIGID_ROOT)
		{
			for (int k = 0; k < 3; k++)
				o[k] = conf[ci + k];
			for (int k = 0; k < 4; k++)
				q[k] = conf[ci + 3 + k];
			normalize(q);
		}
		else
		{
			const int p = f.frame_parent[size_t(fr)];
			double axis[3];
			if (kind == SGPU_FRAME_SEGMENT)
			{
				const double* rp = &R[9 * size_t(p)];
				const double* ro = &f.frame_origin[3 * size_t(fr)];
				const double* ra = &f.frame_axis[3 * size_t(fr)];
				for (int k = 0; k < 3; k++)
				{
					o[k] = O[3 * size_t(p) + k] + rp[3 * k] * ro[0] + rp[3 * k + 1] * ro[1] + rp[3 * k + 2] * ro[2];
					axis[k] = rp[3 * k] * ra[0] + rp[3 * k + 1] * ra[1] + rp[3 * k + 2] * ra[2];
				}
			}
			else
				for (int k = 0; k < 3; k++)
				{
					o[k] = f.frame_origin[3 * size_t(fr) + k];
					axis[k] = f.frame_axis[3 * size_t(fr) + k];
				}
			const double h = conf[ci] / 2, s = std::sin(h);
			const double t[4] = {std::cos(h), s * axis[0], s * axis[1], s * axis[2]};
			if (kind == SGPU_FRAME_SEGMENT)
			{
				const double* u = &Q[4 * size_t(p)];
				q[0] = t[0] * u[0] - t[1] * u[1] - t[2] * u[2] - t[3] * u[3];
				q[1] = t[0] * u[1] + t[1] * u[0] + t[2] * u[3] - t[3] * u[2];
				q[2] = t[0] * u[2] - t[1] * u[3] + t[2] * u[0] + t[3] * u[1];
				q[3] = t[0] * u[3] + t[1] * u[2] - t[2] * u[1] + t[3] * u[0];
			}
			else
				for (int k = 0; k < 4; k++)
					q[k] = t[k];
			normalize(q);
		}
		double* r = &R[9 * size_t(fr)];
		to_r3(q, r);
		for (int a = f.frame_begin[size_t(fr)]; a < f.frame_end[size_t(fr)]; a++)
		{
			const double* l = &f.atom_local[3 * size_t(a)];
			for (int k = 0; k < 3; k++)
				x[3 * size_t(a) + k] = o[k] + r[3 * k] * l[0] + r[3 * k + 1] * l[1] + r[3 * k + 2] * l[2];
		}
	}
	return x;
}

// string_write_coord: %8.3f at a 1-based column
void write_coord(std::string& s, size_t col, double v)
{
	char buf[64];
	std::snprintf(buf, sizeof buf, "%8.3f", v);
	if (std::strlen(buf) != 8 || s.size() < col - 1 + 8)
		throw std::runtime_error("coordinate does not fit the PDBQT columns");
	for (int k = 0; k < 8; k++)
		s[col - 1 + size_t(k)] = buf[k];
}

} // namespace

std::string write_pdbqt_poses(const sgpu_mol& lig, int nposes, const double* conf, int conf_stride, const double* e,
		const double* rmsd)
{
	if (lig.text.empty())
		throw std::invalid_argument("the ligand was not read from PDBQT");
	FlatModel f = flatten(lig.desc);
	if (conf_stride < f.view.conf_size)
		throw std::invalid_argument("conf stride below the conf size");
	std::ostringstream out;
	for (int p = 0; p < nposes; p++)
	{
		const std::vector<double> x = pose_coords(f, conf + size_t(p) * conf_stride);
		out << "MODEL " << p + 1 << "\n";
		char buf[64];
		std::snprintf(buf, sizeof buf, "%.9g", double(float(e[p])));
		out << "REMARK minimizedAffinity " << buf << "\n";
		std::snprintf(buf, sizeof buf, "%.9g", double(float(rmsd ? rmsd[p] : -1.0)));
		out << "REMARK minimizedRMSD " << buf << "\n";
		for (size_t i = 0; i < lig.text.size(); i++)
		{
			std::string s = lig.text[i];
			const int a = lig.text_atom[i];
			if (a >= 0)
				for (int k = 0; k < 3; k++)
					write_coord(s, 31 + 8 * size_t(k), x[3 * size_t(a) + k]);
			else if (starts(s, "BEGIN_RES") || starts(s, "END_RES"))
				continue;
			out << s << "\n";
		}
		out << "ENDMDL\n";
	}
	return out.str();
}

namespace
{
// element symbols of the smina types (src/lib/atom_constants.h:150-216)
const char* element_of(int t)
{
	static const char* el[NUM_ATOM_TYPES] = {"H", "H", "C", "C", "C", "C", "N", "N", "N", "N", "O", "O", "O", "O",
			"S", "S", "P", "F", "Cl", "Br", "I", "Mg", "Mn", "Zn", "Ca", "Fe", "X"};
	return t >= 0 && t < NUM_ATOM_TYPES ? el[t] : "X";
}

// the ligand's atoms: those of its trees (flexible side chains and fixed
// atoms follow them)
int ligand_atoms(const sgpu_mol& m)
{
	int end = 0;
	for (int t = 0; t < m.nligands && t < int(m.tree_frame.size()); t++)
	{
		const int f1 = t + 1 < int(m.tree_frame.size()) ? m.tree_frame[size_t(t) + 1] : int(m.frame_parent.size());
		for (int f = m.tree_frame[size_t(t)]; f < f1; f++)
			end = std::max(end, m.frame_end[size_t(f)]);
	}
	return end;
}

// columns 7-15 of an atom's PDBQT line (model::ligand_atom_str,
// src/lib/model.cpp:845-865): its serial number and name
std::string atom_str(const sgpu_mol& m, int i)
{
	for (size_t k = 0; k < m.text.size(); k++)
		if (m.text_atom[k] == i)
			return m.text[k].size() > 6 ? m.text[k].substr(6, 9) : std::string();
	return std::string();
}

double brick_d2(const double* lo, const double* hi, const double* v)
{
	double s = 0;
	for (int k = 0; k < 3; k++)
	{
		const double c = v[k] <= lo[k] ? lo[k] : v[k] >= hi[k] ? hi[k] : v[k];
		s += (c - v[k]) * (c - v[k]);
	}
	return s;
}
} // namespace

// --atom_terms: each ligand atom's share of every term (weighted) at the
// pose, as result_info::setAtomValues and writeAtomValues write it
// (src/lib/result_info.cpp:41-76) from terms::evale_robust
// (src/lib/terms.cpp:235-307): the receptor's heavy atoms and the model's
// side-chain and fixed heavy atoms within the terms' largest cutoff of the
// movable atoms' box; each term exact at the pair's distance (no tables)
// where the distance is below its own cutoff, charge-dependent terms with
// the atoms' partial charges.  The columns follow the terms' internal order
// (terms::eval_additive_aux, src/lib/terms.cpp:202-230): the
// charge-independent terms, then the charge-dependent ones.
std::string atom_terms(const sgpu_scoring& sf, const sgpu_mol& rec, const sgpu_mol& lig, const double* conf)
{
	std::vector<int> order;
	for (int pass = 0; pass < 2; pass++)
		for (int k = 0; k < sf.nterms; k++)
			if (charge_dependent(sf.terms[k]) == (pass == 1))
				order.push_back(k);
	auto q_of = [](const sgpu_mol& m, int i) { return m.charge.empty() ? 0.0 : m.charge[size_t(i)]; };
	FlatModel f = flatten(lig.desc);
	const std::vector<double> x = pose_coords(f, conf);
	const int nmov = lig.nmovable, nall = lig.natoms(), nlig = ligand_atoms(lig);
	std::vector<double> xa(3 * size_t(nall));
	for (int i = 0; i < nall; i++)
		for (int k = 0; k < 3; k++)
			xa[3 * size_t(i) + k] = i < nmov ? x[3 * size_t(i) + k] : lig.xyz[3 * size_t(i) + k];
	double maxc = 0;
	for (int k = 0; k < sf.nterms; k++)
		maxc = std::max(maxc, term_cutoff(sf.terms[k]));
	const double maxc2 = maxc * maxc;
	const sgpu_grid_dims box = movable_atoms_box_at(xa.data(), nmov, 0, 0.375);
	auto usable = [](int t) { return t < NUM_ATOM_TYPES && !is_hydrogen_type(t); };
	// (coordinates, type) of the relevant atoms: receptor first, then the
	// model's non-ligand atoms
	std::vector<const double*> rx;
	std::vector<int> rt;
	std::vector<double> rq;
	for (int j = 0; j < rec.natoms(); j++)
		if (usable(rec.type[size_t(j)]) && brick_d2(box.begin, box.end, &rec.xyz[3 * size_t(j)]) < maxc2)
		{
			rx.push_back(&rec.xyz[3 * size_t(j)]);
			rt.push_back(rec.type[size_t(j)]);
			rq.push_back(q_of(rec, j));
		}
	for (int j = nlig; j < nall; j++)
		if (usable(lig.type[size_t(j)]) && brick_d2(box.begin, box.end, &xa[3 * size_t(j)]) < maxc2)
		{
			rx.push_back(&xa[3 * size_t(j)]);
			rt.push_back(lig.type[size_t(j)]);
			rq.push_back(q_of(lig, j));
		}
	std::ostringstream out;
	out << "atomid el pos";
	for (int k : order)
		out << " " << term_name(sf.terms[k]);
	out << "\n";
	for (int i = 0; i < nlig; i++)
	{
		std::vector<double> v(size_t(sf.nterms), 0.0);
		const double* xi = &xa[3 * size_t(i)];
		if (usable(lig.type[size_t(i)]))
			for (size_t j = 0; j < rx.size(); j++)
			{
				const double d2 = (xi[0] - rx[j][0]) * (xi[0] - rx[j][0]) + (xi[1] - rx[j][1]) * (xi[1] - rx[j][1]) +
						(xi[2] - rx[j][2]) * (xi[2] - rx[j][2]);
				if (d2 > maxc2)
					continue;
				const double d = std::sqrt(d2);
				for (int k = 0; k < sf.nterms; k++)
					if (d < term_cutoff(sf.terms[k]))
						v[size_t(k)] += term_value(sf.terms[k], lig.type[size_t(i)], rt[j], d, sf.atom_params,
								q_of(lig, i), rq[j]);
			}
		out << atom_str(lig, i) << " <" << xi[0] << "," << xi[1] << "," << xi[2] << ">";
		for (int k : order)
			out << " " << v[size_t(k)] * sf.terms[k].weight;
		out << "\n";
	}
	out << "END\n";
	return out.str();
}

// SDF poses (result_info::write's native SDF branch,
// src/lib/result_info.cpp:139-163): per pose a molfile of the ligand's atoms
// at the pose with the model's covalent bonds (single bonds: orders are not
// perceived without OpenBabel), then the minimizedAffinity and
// minimizedRMSD data items to 5 decimals, the atomic_interaction_terms item
// when asked for, and $$$$
std::string write_sdf_poses(const sgpu_mol& lig, int nposes, const double* conf, int conf_stride, const double* e,
		const double* rmsd, const sgpu_scoring* sf, const sgpu_mol* rec)
{
	FlatModel f = flatten(lig.desc);
	if (conf_stride < f.view.conf_size)
		throw std::invalid_argument("conf stride below the conf size");
	const int nlig = ligand_atoms(lig);
	if (nlig > 999)
		throw std::invalid_argument("SDF V2000 holds at most 999 atoms");
	std::string name;
	for (const std::string& l : lig.text)
		if (l.compare(0, 15, "REMARK  Name = ") == 0)
			name = l.substr(15);
	std::vector<std::pair<int, int>> bonds;
	for (size_t k = 0; k + 1 < lig.bonds.size(); k += 2)
		if (lig.bonds[k] < nlig && lig.bonds[k + 1] < nlig)
			bonds.push_back({lig.bonds[k], lig.bonds[k + 1]});
	std::ostringstream out;
	char buf[128];
	for (int p = 0; p < nposes; p++)
	{
		const double* c = conf + size_t(p) * conf_stride;
		const std::vector<double> x = pose_coords(f, c);
		out << name << "\n  sgpu\n\n";
		std::snprintf(buf, sizeof buf, "%3d%3d  0  0  0  0  0  0  0  0999 V2000", nlig, int(bonds.size()));
		out << buf << "\n";
		for (int i = 0; i < nlig; i++)
		{
			std::snprintf(buf, sizeof buf, "%10.4f%10.4f%10.4f %-3s 0  0  0  0  0  0  0  0  0  0  0  0",
					x[3 * size_t(i)], x[3 * size_t(i) + 1], x[3 * size_t(i) + 2], element_of(lig.type[size_t(i)]));
			out << buf << "\n";
		}
		for (const auto& b : bonds)
		{
			std::snprintf(buf, sizeof buf, "%3d%3d  1  0  0  0  0", b.first + 1, b.second + 1);
			out << buf << "\n";
		}
		out << "M  END\n";
		std::snprintf(buf, sizeof buf, "%.5f", e[p]);
		out << "> <minimizedAffinity>\n" << buf << "\n\n";
		if (rmsd && rmsd[p] >= 0)
		{
			std::snprintf(buf, sizeof buf, "%.5f", rmsd[p]);
			out << "> <minimizedRMSD>\n" << buf << "\n\n";
		}
		if (sf && rec)
			out << "> <atomic_interaction_terms>\n" << atom_terms(*sf, *rec, lig, c) << "\n\n";
		out << "$$$$\n";
	}
	return out.str();
}

} // namespace sgpu_host

using namespace sgpu_host;

namespace
{
sgpu_mol* guarded_read(sgpu_mol* (*fn)(const char*), const char* text, char* err, int err_len)
{
	try
	{
		if (!text)
			throw std::invalid_argument("no text");
		return fn(text);
	}
	catch (std::exception& e)
	{
		if (err && err_len > 0)
			std::snprintf(err, size_t(err_len), "%s", e.what());
		return nullptr;
	}
}
} // namespace

namespace
{
// the reference splits header lines on every single whitespace character
// (boost::algorithm::split with is_space, no compression): "A  B" gives
// "A", "", "B"
std::vector<std::string> split_each_space(const std::string& line)
{
	std::vector<std::string> out(1);
	for (char ch : line)
		if (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f')
			out.emplace_back();
		else
			out.back() += ch;
	return out;
}

double field(const std::vector<std::string>& t, size_t i)
{
	return i < t.size() ? std::atof(t[i].c_str()) : 0.0;
}

// setup_user_gd (src/main/main.cpp:772-810) then grid::init's value loop
// (src/lib/grid.cpp:77-108): three lines skipped; SPACING s; NELEMENTS
// nx ny nz; CENTER cx cy cz; the box spans (N + 1) s per axis around the
// center moved by s / 2, n = ceil(span / s) intervals; then n0 n1 n2 values,
// one per line, z slowest, each stored as -(value x scaling); the last
// plane of each axis stays 0; a missing line reads as 0
int read_user_grid(const std::string& text, double scaling, sgpu_grid_dims& gd, std::vector<double>& data)
{
	std::istringstream in(text);
	std::string line;
	for (int k = 0; k < 3; k++)
		std::getline(in, line);
	std::getline(in, line);
	const double gran = field(split_each_space(line), 1);
	std::getline(in, line);
	std::vector<std::string> t = split_each_space(line);
	const double size[3] = {(field(t, 1) + 1) * gran, (field(t, 2) + 1) * gran, (field(t, 3) + 1) * gran};
	std::getline(in, line);
	t = split_each_space(line);
	const double center[3] = {field(t, 1) + 0.5 * gran, field(t, 2) + 0.5 * gran, field(t, 3) + 0.5 * gran};
	if (!(gran > 0))
		throw std::invalid_argument("user grid: SPACING must be positive");
	for (int i = 0; i < 3; i++)
	{
		const double n = std::ceil(size[i] / gran);
		if (!(n >= 1 && n <= 4096))
			throw std::invalid_argument("user grid: NELEMENTS out of range");
		gd.n[i] = int(n);
		const double real_span = gran * gd.n[i];
		gd.begin[i] = center[i] - real_span / 2;
		gd.end[i] = gd.begin[i] + real_span;
	}
	const int d0 = gd.n[0] + 1, d1 = gd.n[1] + 1, d2 = gd.n[2] + 1;
	data.assign(size_t(d0) * d1 * d2, 0.0);
	for (int z = 0; z < gd.n[2]; z++)
		for (int y = 0; y < gd.n[1]; y++)
			for (int x = 0; x < gd.n[0]; x++)
			{
				if (!std::getline(in, line))
					line.clear();
				data[size_t(x) + size_t(d0) * (size_t(y) + size_t(d1) * z)] = -(std::atof(line.c_str()) * scaling);
			}
	return int(data.size());
}
} // namespace

extern "C" {

int sgpu_read_user_grid(const char* text, double scaling, sgpu_grid_dims* gd, double* data, int capacity, char* err,
		int err_len)
{
	try
	{
		if (!text || !gd)
			throw std::invalid_argument("user grid: no text or box");
		std::vector<double> v;
		const int n = read_user_grid(text, scaling, *gd, v);
		if (data)
			std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), size_t(std::max(0, capacity))), data);
		return n;
	}
	catch (std::exception& e)
	{
		if (err && err_len > 0)
			std::snprintf(err, size_t(err_len), "%s", e.what());
		return -1;
	}
}


sgpu_mol* sgpu_read_pdbqt_ligand(const char* text, char* err, int err_len)
{
	return guarded_read(read_ligand, text, err, err_len);
}

sgpu_mol* sgpu_read_pdbqt_receptor(const char* text, char* err, int err_len)
{
	return guarded_read(read_receptor, text, err, err_len);
}

sgpu_mol* sgpu_read_pdbqt_flex(const sgpu_mol* ligand, const sgpu_mol* rigid, const char* text, char* err,
		int err_len)
{
	try
	{
		if (!ligand || !rigid || !text)
			throw std::invalid_argument("ligand, rigid receptor and text are required");
		if (ligand->nmovable != ligand->natoms() || ligand->nligands < 1)
			throw std::invalid_argument("the ligand already carries side chains");
		return read_flex(*ligand, *rigid, text);
	}
	catch (std::exception& e)
	{
		if (err && err_len > 0)
			std::snprintf(err, size_t(err_len), "%s", e.what());
		return nullptr;
	}
}

int sgpu_write_pdbqt_poses(const sgpu_mol* lig, int nposes, const double* conf, int conf_stride, const double* e,
		const double* rmsd, char* out, int out_len)
{
	try
	{
		if (!lig || nposes < 0 || (nposes > 0 && (!conf || !e)))
			return -1;
		const std::string s = write_pdbqt_poses(*lig, nposes, conf, conf_stride, e, rmsd);
		if (out && out_len > 0)
			std::snprintf(out, size_t(out_len), "%s", s.c_str());
		return int(s.size());
	}
	catch (...)
	{
		return -1;
	}
}

int sgpu_mol_set_charges(sgpu_mol* m, const double* charges)
{
	if (!m)
		return -1;
	if (charges)
		m->charge.assign(charges, charges + m->natoms());
	else
		m->charge.clear();
	m->publish();
	return 0;
}

int sgpu_atom_terms(const sgpu_scoring* sf, const sgpu_mol* receptor, const sgpu_mol* ligand, const double* conf,
		char* out, int out_len)
{
	try
	{
		if (!sf || !receptor || !ligand || !conf)
			return -1;
		const std::string s = atom_terms(*sf, *receptor, *ligand, conf);
		if (out && out_len > 0)
			std::snprintf(out, size_t(out_len), "%s", s.c_str());
		return int(s.size());
	}
	catch (...)
	{
		return -1;
	}
}

int sgpu_write_sdf_poses(const sgpu_mol* lig, int nposes, const double* conf, int conf_stride, const double* e,
		const double* rmsd, const sgpu_scoring* atom_terms_sf, const sgpu_mol* receptor, char* out, int out_len)
{
	try
	{
		if (!lig || nposes < 0 || (nposes > 0 && (!conf || !e)) || (atom_terms_sf && !receptor))
			return -1;
		const std::string s = write_sdf_poses(*lig, nposes, conf, conf_stride, e, rmsd, atom_terms_sf, receptor);
		if (out && out_len > 0)
			std::snprintf(out, size_t(out_len), "%s", s.c_str());
		return int(s.size());
	}
	catch (...)
	{
		return -1;
	}
}

} // extern "C"

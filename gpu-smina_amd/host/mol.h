// Owned molecule storage behind sgpu_mol / sgpu_model_desc.
#ifndef SGPU_HOST_MOL_H
#define SGPU_HOST_MOL_H

#include <string>
#include <vector>
#include "../../include/sgpu_host.h"

struct sgpu_mol
{
	std::vector<int> type;
	std::vector<double> xyz;
	std::vector<int> bonds;
	std::vector<int> frame_parent, frame_begin, frame_end, frame_origin_atom, frame_axis_atom;
	std::vector<int> tree_frame;
	std::vector<double> charge; // partial charges (empty: all 0)
	int nmovable = 0;
	int nligands = 0;
	sgpu_model_desc desc{};
	// a parsed PDBQT file: its lines, and for an atom record the index of
	// its atom in the model (-1 for other lines), for writing poses back
	std::vector<std::string> text;
	std::vector<int> text_atom;

	int natoms() const { return int(type.size()); }
	// point desc at the vectors (after any change of them)
	void publish()
	{
		desc.natoms = natoms();
		desc.nmovable = nmovable;
		desc.type = type.data();
		desc.xyz = xyz.data();
		desc.nbonds = int(bonds.size() / 2);
		desc.bonds = bonds.data();
		desc.nframes = int(frame_parent.size());
		desc.frame_parent = frame_parent.data();
		desc.frame_begin = frame_begin.data();
		desc.frame_end = frame_end.data();
		desc.frame_origin_atom = frame_origin_atom.data();
		desc.frame_axis_atom = frame_axis_atom.data();
		desc.nligands = nligands;
		desc.ntrees = int(tree_frame.size());
		desc.tree_frame = tree_frame.data();
		desc.charge = charge.size() == type.size() ? charge.data() : nullptr;
	}
};

#endif

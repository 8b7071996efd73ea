/*
 * Seeded synthetic docking inputs.  They stand in for PDBQT files (the
 * reference reads receptors and ligands with OpenBabel / parse_pdbqt,
 * src/lib/parse_pdbqt.cpp:255-330, which this build does not carry): a
 * receptor of random-walk chains around a carved pocket, and tree-shaped
 * ligands with a chosen number of rotatable bonds, laid out in the atom and
 * frame order smina's parser produces (postprocess_branch,
 * src/lib/parse_pdbqt.cpp:318-346: a frame's atoms in listing order, each
 * followed by the first atoms of the branches rooted at it).
 */
#include <algorithm>
#include <functional>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <unordered_map>
#include <vector>

#include "../csrc/atom_types.h"
#include "../csrc/rng.h"
#include "mol.h"

namespace
{

using namespace sgpu;

struct V3
{
	double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(V3 a) { return std::sqrt(dot(a, a)); }
V3 unit(V3 a) { return (1.0 / norm(a)) * a; }

struct Gen
{
	MTState s;
	explicit Gen(uint32_t seed) { mt_seed(s, seed); }
	uint32_t next() { return mt_next(s); }
	double u(double a, double b) { return random_fl(a, b, *this); }
	int i(int a, int b) { return random_int(a, b, *this); }
	V3 unit_vec()
	{
		for (;;)
		{
			V3 v{u(-1, 1), u(-1, 1), u(-1, 1)};
			double n = norm(v);
			if (n > 1e-3 && n <= 1)
				return (1.0 / n) * v;
		}
	}
	V3 perpendicular(V3 d)
	{
		for (;;)
		{
			V3 w = cross(d, unit_vec());
			if (norm(w) > 1e-3)
				return unit(w);
		}
	}
};

// uniform-cell spatial hash for clearance checks
struct SpatialHash
{
	double cell;
	std::unordered_map<long long, std::vector<int>> m;
	explicit SpatialHash(double c) : cell(c) {}
	long long key(int a, int b, int c) const
	{
		return ((long long)(a + 100000) * 200003LL + (b + 100000)) * 200003LL + (c + 100000);
	}
	void add(int idx, V3 p)
	{
		m[key(int(std::floor(p.x / cell)), int(std::floor(p.y / cell)), int(std::floor(p.z / cell)))].push_back(idx);
	}
	bool clear_of(const std::vector<V3>& pts, V3 p, double r, int skip) const
	{
		int cx = int(std::floor(p.x / cell)), cy = int(std::floor(p.y / cell)), cz = int(std::floor(p.z / cell));
		int reach = int(std::ceil(r / cell));
		for (int a = -reach; a <= reach; a++)
			for (int b = -reach; b <= reach; b++)
				for (int c = -reach; c <= reach; c++)
				{
					auto it = m.find(key(cx + a, cy + b, cz + c));
					if (it == m.end())
						continue;
					for (int k : it->second)
						if (k != skip && norm(pts[k] - p) < r)
							return false;
				}
		return true;
	}
};

// a direction for a new bond from atom p with existing bond directions
V3 bond_direction(Gen& g, const std::vector<V3>& pos, const std::vector<std::vector<int>>& nb, int p)
{
	const auto& n = nb[p];
	if (n.empty())
		return g.unit_vec();
	if (n.size() == 1)
	{
		V3 u = unit(pos[p] - pos[n[0]]);
		V3 w = g.perpendicular(u);
		const double a = (180.0 - 109.5) * 3.14159265358979 / 180.0;
		return unit(std::cos(a) * u + std::sin(a) * w);
	}
	V3 s{0, 0, 0};
	for (int q : n)
		s = s + unit(pos[q] - pos[p]);
	V3 back = norm(s) > 1e-6 ? unit(-1.0 * s) : g.unit_vec();
	if (n.size() >= 3)
		return back;
	V3 u1 = unit(pos[n[0]] - pos[p]), u2 = unit(pos[n[1]] - pos[p]);
	V3 c = cross(u1, u2);
	V3 side = norm(c) > 1e-6 ? unit(c) : g.perpendicular(back);
	if (g.u(0, 1) < 0.5)
		side = -1.0 * side;
	const double t = 54.75 * 3.14159265358979 / 180.0;
	return unit(std::cos(t) * back + std::sin(t) * side);
}

} // namespace

extern "C" {

void sgpu_mol_free(sgpu_mol* m) { delete m; }
const sgpu_model_desc* sgpu_mol_desc(const sgpu_mol* m) { return &m->desc; }

sgpu_mol* sgpu_synth_receptor(uint32_t seed, int nheavy, double pocket_radius)
{
	Gen g(seed);
	const double density = 0.05; // heavy atoms per A^3, protein-like
	const double r_out = std::cbrt(nheavy / density * 3.0 / (4.0 * 3.14159265358979) +
			pocket_radius * pocket_radius * pocket_radius);
	std::vector<V3> pos;
	std::vector<std::vector<int>> nb;
	std::vector<int> chain_of;
	SpatialHash hash(2.0);
	auto in_shell = [&](V3 p) {
		double r = norm(p);
		return r >= pocket_radius && r <= r_out;
	};
	int chain = 0;
	int failures = 0;
	while (int(pos.size()) < nheavy && failures < 100000)
	{
		V3 start{g.u(-r_out, r_out), g.u(-r_out, r_out), g.u(-r_out, r_out)};
		if (!in_shell(start) || !hash.clear_of(pos, start, 1.3, -1))
		{
			++failures;
			continue;
		}
		int cur = int(pos.size());
		pos.push_back(start);
		nb.emplace_back();
		chain_of.push_back(chain);
		hash.add(cur, start);
		V3 dir = g.unit_vec();
		int len = 1;
		const int max_len = 40 + g.i(0, 80);
		while (len < max_len && int(pos.size()) < nheavy)
		{
			bool placed = false;
			for (int tries = 0; tries < 30 && !placed; tries++)
			{
				V3 d = unit(dir + 0.9 * g.unit_vec());
				if (!nb[cur].empty() && dot(d, unit(pos[cur] - pos[nb[cur][0]])) < -0.2)
					continue; // keep bond angles open
				V3 p = pos[cur] + 1.5 * d;
				if (!in_shell(p) || !hash.clear_of(pos, p, 1.3, cur))
					continue;
				int k = int(pos.size());
				pos.push_back(p);
				nb.emplace_back();
				chain_of.push_back(chain);
				nb[cur].push_back(k);
				nb[k].push_back(cur);
				hash.add(k, p);
				cur = k;
				dir = d;
				placed = true;
			}
			if (!placed)
				break;
			++len;
		}
		++chain;
	}
	const int nh = int(pos.size());
	// elements C 63 / N 17 / O 19 / S 1 %, then smina types with polar H
	std::vector<int> elem(nh); // 0 C, 1 N, 2 O, 3 S
	for (int i = 0; i < nh; i++)
	{
		double r = g.u(0, 1);
		elem[i] = r < 0.63 ? 0 : r < 0.80 ? 1 : r < 0.99 ? 2 : 3;
	}
	sgpu_mol* m = new sgpu_mol;
	std::vector<int> hparent;
	for (int i = 0; i < nh; i++)
	{
		int t = T_CAliph_H;
		if (elem[i] == 0)
		{
			bool hetero = false;
			for (int q : nb[i])
				hetero = hetero || elem[q] != 0;
			t = hetero ? T_CAliph_P : T_CAliph_H;
		}
		else if (elem[i] == 1)
		{
			double r = g.u(0, 1);
			if (r < 0.6)
			{
				t = T_N_D;
				hparent.push_back(i);
			}
			else
				t = r < 0.8 ? T_N_A : T_N_P;
		}
		else if (elem[i] == 2)
		{
			if (g.u(0, 1) < 0.3)
			{
				t = T_O_DA;
				hparent.push_back(i);
			}
			else
				t = T_O_A;
		}
		else
			t = T_S_P;
		m->type.push_back(t);
		m->xyz.insert(m->xyz.end(), {pos[i].x, pos[i].y, pos[i].z});
	}
	for (int i = 0; i < nh; i++)
		for (int q : nb[i])
			if (q > i)
				m->bonds.insert(m->bonds.end(), {i, q});
	for (int p : hparent)
	{
		V3 d = bond_direction(g, pos, nb, p);
		V3 h = pos[p] + 1.0 * d;
		int k = int(m->type.size());
		m->type.push_back(T_PolarHydrogen);
		m->xyz.insert(m->xyz.end(), {h.x, h.y, h.z});
		m->bonds.insert(m->bonds.end(), {p, k});
	}
	m->nmovable = 0;
	m->nligands = 0;
	m->publish();
	return m;
}

sgpu_mol* sgpu_synth_ligand(uint32_t seed, int nheavy, int ntors, const double* center)
{
	Gen g(seed * 2654435761u + 12345u);
	if (nheavy < 2)
		nheavy = 2;
	std::vector<V3> pos;
	std::vector<std::vector<int>> nb;
	std::vector<std::pair<int, int>> rot;
	for (int attempt = 0;; attempt++)
	{
		if (attempt > 2000)
			throw std::runtime_error("could not build a ligand with the requested torsions");
		pos.assign(1, V3{0, 0, 0});
		nb.assign(1, {});
		SpatialHash hash(2.5);
		hash.add(0, pos[0]);
		int guard = 0;
		while (int(pos.size()) < nheavy && guard++ < 5000)
		{
			// grow mostly from recent atoms: long, branched chains
			int n = int(pos.size());
			int p = g.u(0, 1) < 0.75 ? std::max(0, n - 1 - g.i(0, 2)) : g.i(0, n - 1);
			if (nb[p].size() >= 3)
				continue;
			V3 d = bond_direction(g, pos, nb, p);
			V3 c = pos[p] + 1.5 * d;
			if (!hash.clear_of(pos, c, 2.2, p))
				continue;
			bool ok = true;
			for (int q : nb[p])
				ok = ok && norm(pos[q] - c) > 2.0;
			if (!ok)
				continue;
			int k = n;
			pos.push_back(c);
			nb.emplace_back();
			nb[p].push_back(k);
			nb[k].push_back(p);
			hash.add(k, c);
		}
		if (int(pos.size()) < nheavy)
			continue;
		std::vector<std::pair<int, int>> cand;
		for (int i = 0; i < nheavy; i++)
			for (int q : nb[i])
				if (q > i && nb[i].size() >= 2 && nb[q].size() >= 2)
					cand.push_back({i, q});
		if (int(cand.size()) < ntors)
			continue;
		for (int k = 0; k < ntors; k++)
		{
			int j = g.i(k, int(cand.size()) - 1);
			std::swap(cand[k], cand[j]);
		}
		rot.assign(cand.begin(), cand.begin() + ntors);
		break;
	}
	// elements C 70 / N 15 / O 15 %: O on atoms with <= 2 heavy neighbours
	std::vector<int> elem(nheavy, 0);
	int n_n = int(std::lround(0.15 * nheavy)), n_o = int(std::lround(0.15 * nheavy));
	std::vector<int> order(nheavy);
	std::iota(order.begin(), order.end(), 0);
	for (int k = nheavy - 1; k > 0; k--)
		std::swap(order[k], order[g.i(0, k)]);
	for (int a : order)
	{
		if (n_o > 0 && nb[a].size() <= 2)
		{
			elem[a] = 2;
			--n_o;
		}
		else if (n_n > 0)
		{
			elem[a] = 1;
			--n_n;
		}
	}
	std::vector<int> type(nheavy);
	std::vector<int> hparent;
	for (int i = 0; i < nheavy; i++)
	{
		if (elem[i] == 0)
		{
			bool hetero = false;
			for (int q : nb[i])
				hetero = hetero || elem[q] != 0;
			type[i] = hetero ? T_CAliph_P : T_CAliph_H;
		}
		else if (elem[i] == 1)
		{
			if (nb[i].size() >= 3)
				type[i] = T_N_P;
			else if (g.u(0, 1) < 0.5)
			{
				type[i] = T_N_DA;
				hparent.push_back(i);
			}
			else
				type[i] = T_N_A;
		}
		else
		{
			if (nb[i].size() == 1 && g.u(0, 1) < 0.5)
			{
				type[i] = T_O_DA;
				hparent.push_back(i);
			}
			else
				type[i] = T_O_A;
		}
	}
	// polar hydrogens
	std::vector<V3> hpos;
	for (int p : hparent)
		hpos.push_back(pos[p] + 1.0 * bond_direction(g, pos, nb, p));
	const int nh = int(hparent.size());
	const int natoms = nheavy + nh;
	// all-atom adjacency (heavy indices, then hydrogens)
	std::vector<std::vector<int>> adj(natoms);
	for (int i = 0; i < nheavy; i++)
		adj[i] = nb[i];
	for (int k = 0; k < nh; k++)
	{
		adj[nheavy + k].push_back(hparent[k]);
		adj[hparent[k]].push_back(nheavy + k);
	}
	auto is_rot = [&](int a, int b) {
		for (auto& r : rot)
			if ((r.first == a && r.second == b) || (r.first == b && r.second == a))
				return true;
		return false;
	};
	// fragments: components without the rotatable bonds
	std::vector<int> frag(natoms, -1);
	int nfrag = 0;
	std::vector<int> fsize;
	for (int s = 0; s < natoms; s++)
	{
		if (frag[s] >= 0)
			continue;
		std::vector<int> st{s};
		frag[s] = nfrag;
		int cnt = 0;
		while (!st.empty())
		{
			int a = st.back();
			st.pop_back();
			++cnt;
			for (int q : adj[a])
				if (frag[q] < 0 && !is_rot(a, q))
				{
					frag[q] = nfrag;
					st.push_back(q);
				}
		}
		fsize.push_back(cnt);
		++nfrag;
	}
	// root: the largest fragment
	int root = int(std::max_element(fsize.begin(), fsize.end()) - fsize.begin());
	// frame tree in preorder with smina's atom listing
	struct F
	{
		int frag, entry, axis, parent;
	};
	std::vector<F> frames;
	std::vector<int> new_index(natoms, -1);
	std::vector<int> order_atoms;
	std::vector<int> f_parent, f_begin, f_end, f_origin, f_axis;
	// listing of a fragment: BFS from its entry atom inside the fragment
	auto listing = [&](int fr, int entry) {
		std::vector<int> out{entry};
		std::vector<char> seen(natoms, 0);
		seen[entry] = 1;
		for (size_t h = 0; h < out.size(); h++)
			for (int q : adj[out[h]])
				if (!seen[q] && frag[q] == fr && !is_rot(out[h], q))
				{
					seen[q] = 1;
					out.push_back(q);
				}
		return out;
	};
	// recursive placement
	std::vector<int> frag_done(nfrag, 0);
	struct Rec
	{
		static void place(int fr, int entry, int parent_frame, int axis_atom, bool is_root,
				const std::function<std::vector<int>(int, int)>& listing_fn,
				const std::vector<std::vector<int>>& adj, const std::vector<int>& frag,
				const std::function<bool(int, int)>& is_rot, std::vector<int>& frag_done,
				std::vector<int>& new_index, std::vector<int>& order_atoms, std::vector<int>& f_parent,
				std::vector<int>& f_begin, std::vector<int>& f_end, std::vector<int>& f_origin,
				std::vector<int>& f_axis)
		{
			frag_done[fr] = 1;
			int me = int(f_parent.size());
			f_parent.push_back(parent_frame);
			f_origin.push_back(entry); // old index, remapped below
			f_axis.push_back(axis_atom);
			f_begin.push_back(int(order_atoms.size()));
			f_end.push_back(-1);
			std::vector<int> lst = listing_fn(fr, entry);
			std::vector<std::pair<int, int>> children; // (axis atom, child entry)
			for (int a : lst)
			{
				if (is_root || a != entry)
				{
					new_index[a] = int(order_atoms.size());
					order_atoms.push_back(a);
				}
				for (int q : adj[a])
					if (is_rot(a, q) && !frag_done[frag[q]])
					{
						new_index[q] = int(order_atoms.size());
						order_atoms.push_back(q);
						children.push_back({a, q});
						frag_done[frag[q]] = 2; // reserved
					}
			}
			f_end[me] = int(order_atoms.size());
			for (auto& c : children)
				place(frag[c.second], c.second, me, c.first, false, listing_fn, adj, frag, is_rot, frag_done,
						new_index, order_atoms, f_parent, f_begin, f_end, f_origin, f_axis);
		}
	};
	int root_entry = -1;
	for (int a = 0; a < nheavy && root_entry < 0; a++)
		if (frag[a] == root)
			root_entry = a;
	std::function<std::vector<int>(int, int)> listing_fn = listing;
	std::function<bool(int, int)> is_rot_fn = is_rot;
	Rec::place(root, root_entry, -1, -1, true, listing_fn, adj, frag, is_rot_fn, frag_done, new_index, order_atoms,
			f_parent, f_begin, f_end, f_origin, f_axis);
	if (int(order_atoms.size()) != natoms)
		throw std::runtime_error("ligand frame layout lost atoms");
	// random pose: rotate about the centroid, move near `center`
	V3 cen{0, 0, 0};
	for (int i = 0; i < nheavy; i++)
		cen = cen + pos[i];
	cen = (1.0 / nheavy) * cen;
	V3 ax = g.unit_vec();
	double ang = g.u(-3.14159265358979, 3.14159265358979);
	V3 shift{center ? center[0] : 0.0, center ? center[1] : 0.0, center ? center[2] : 0.0};
	shift = shift + V3{g.u(-2, 2), g.u(-2, 2), g.u(-2, 2)};
	auto place_pt = [&](V3 p) {
		V3 v = p - cen;
		// Rodrigues rotation
		V3 r = std::cos(ang) * v + std::sin(ang) * cross(ax, v) + (dot(ax, v) * (1 - std::cos(ang))) * ax;
		return r + shift;
	};
	sgpu_mol* m = new sgpu_mol;
	m->type.resize(natoms);
	m->xyz.resize(3 * natoms);
	for (int k = 0; k < natoms; k++)
	{
		int a = order_atoms[k];
		int t = a < nheavy ? type[a] : T_PolarHydrogen;
		V3 p = place_pt(a < nheavy ? pos[a] : hpos[a - nheavy]);
		m->type[k] = t;
		m->xyz[3 * k] = p.x;
		m->xyz[3 * k + 1] = p.y;
		m->xyz[3 * k + 2] = p.z;
	}
	for (int a = 0; a < natoms; a++)
		for (int q : adj[a])
			if (q > a)
				m->bonds.insert(m->bonds.end(), {new_index[a], new_index[q]});
	m->frame_parent = f_parent;
	m->frame_begin = f_begin;
	m->frame_end = f_end;
	for (size_t f = 0; f < f_origin.size(); f++)
	{
		m->frame_origin_atom.push_back(new_index[f_origin[f]]);
		m->frame_axis_atom.push_back(f_axis[f] < 0 ? -1 : new_index[f_axis[f]]);
	}
	// the root frame's origin is its first listed atom
	m->frame_origin_atom[0] = f_begin[0];
	m->tree_frame = {0};
	m->nmovable = natoms;
	m->nligands = 1;
	m->publish();
	return m;
}

// Flexible side chains cut from a synthetic receptor, in the layout smina's
// flexible-residue parser builds (src/lib/parse_pdbqt.cpp:360-380,
// postprocess_residue): per residue, the root atom stays fixed ("inflex"),
// each rotatable bond below it is a branch, the first one a first_segment
// rotating about the bond from the root atom.  A side chain with c torsions
// is the last c + 2 heavy atoms of a receptor chain that ends in the pocket:
// anchor, a1 .. a(c+1); frame k (1..c) holds a(k) with its hydrogens, the
// last frame also a(c+1); frame k turns about the bond a(k-1) -> a(k)
// (a0 = anchor).  The model is the ligand's trees, then one tree per side
// chain, then the anchors; *rigid receives the receptor without the side
// chains' atoms (anchors included).
sgpu_mol* sgpu_synth_flex(const sgpu_mol* receptor, const sgpu_mol* ligand, uint32_t seed, int nflex, int min_chi,
		int max_chi, const double* center, sgpu_mol** rigid)
{
	if (!receptor || !ligand || !rigid || nflex < 0 || min_chi < 1 || max_chi < min_chi ||
			ligand->nmovable != ligand->natoms())
		return nullptr;
	*rigid = nullptr;
	Gen g(seed * 2246822519u + 77u);
	const int na = receptor->natoms();
	std::vector<std::vector<int>> heavy_nb(na), hyd(na);
	for (size_t k = 0; k + 1 < receptor->bonds.size(); k += 2)
	{
		const int a = receptor->bonds[k], b = receptor->bonds[k + 1];
		const bool ha = receptor->type[a] == T_PolarHydrogen, hb = receptor->type[b] == T_PolarHydrogen;
		if (!ha && !hb)
		{
			heavy_nb[a].push_back(b);
			heavy_nb[b].push_back(a);
		}
		else if (ha != hb)
			hyd[ha ? b : a].push_back(ha ? a : b);
	}
	const V3 c0{center ? center[0] : 0, center ? center[1] : 0, center ? center[2] : 0};
	auto at = [&](int a) { return V3{receptor->xyz[3 * a], receptor->xyz[3 * a + 1], receptor->xyz[3 * a + 2]}; };
	// candidate side chains: every chain end with its path; the chosen ones
	// are those whose farthest atom is nearest the centre
	std::vector<int> ends;
	for (int a = 0; a < na; a++)
		if (receptor->type[a] != T_PolarHydrogen && heavy_nb[a].size() == 1)
			ends.push_back(a);
	std::vector<std::vector<int>> cand; // [a(c+1), ..., a1, anchor]
	std::vector<double> reach;
	for (int e : ends)
	{
		const int c = g.i(min_chi, max_chi);
		std::vector<int> path{e};
		while (int(path.size()) < c + 2)
		{
			const int cur = path.back(), prev = path.size() > 1 ? path[path.size() - 2] : -1;
			int next = -1;
			for (int q : heavy_nb[cur])
				if (q != prev)
					next = q;
			// a movable atom bonds to no heavy atom beyond the path
			if (next < 0 || (path.size() > 1 && heavy_nb[cur].size() != 2))
				break;
			path.push_back(next);
		}
		if (int(path.size()) < c + 2)
			continue;
		double r = 0;
		for (int a : path)
			r = std::max(r, norm(at(a) - c0));
		cand.push_back(path);
		reach.push_back(r);
	}
	std::vector<int> order(cand.size());
	std::iota(order.begin(), order.end(), 0);
	std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reach[a] < reach[b]; });
	std::vector<char> taken(na, 0);
	std::vector<std::vector<int>> chains;
	for (int k : order)
	{
		if (int(chains.size()) == nflex)
			break;
		bool free = true;
		for (int a : cand[k])
			free = free && !taken[a];
		if (!free)
			continue;
		for (int a : cand[k])
			taken[a] = 1;
		chains.push_back(cand[k]);
	}
	if (int(chains.size()) < nflex)
		return nullptr;
	sgpu_mol* m = new sgpu_mol(*ligand);
	std::vector<int> new_index(na, -1);
	auto add_atom = [&](int a) {
		new_index[a] = m->natoms();
		m->type.push_back(receptor->type[a]);
		V3 p = at(a);
		m->xyz.insert(m->xyz.end(), {p.x, p.y, p.z});
	};
	struct Fr
	{
		int parent, begin, end, origin, axis;
	};
	std::vector<Fr> frames;
	for (const auto& path : chains)
	{
		const int c = int(path.size()) - 2;
		m->tree_frame.push_back(int(m->frame_parent.size()));
		for (int k = 1; k <= c; k++)
		{
			const int ak = path[c + 1 - k];
			const int first = m->natoms();
			add_atom(ak);
			for (int h : hyd[ak])
				add_atom(h);
			if (k == c)
			{
				add_atom(path[0]);
				for (int h : hyd[path[0]])
					add_atom(h);
			}
			const int f = int(m->frame_parent.size());
			m->frame_parent.push_back(k == 1 ? -1 : f - 1);
			m->frame_begin.push_back(first);
			m->frame_end.push_back(m->natoms());
			m->frame_origin_atom.push_back(first);
			m->frame_axis_atom.push_back(-1 - path[c + 2 - k]); // receptor index, resolved below
		}
	}
	m->nmovable = m->natoms();
	for (const auto& path : chains)
	{
		const int anchor = path.back();
		add_atom(anchor);
		for (int h : hyd[anchor])
			add_atom(h);
	}
	for (size_t f = ligand->frame_parent.size(); f < m->frame_axis_atom.size(); f++)
		m->frame_axis_atom[f] = new_index[-1 - m->frame_axis_atom[f]];
	for (size_t k = 0; k + 1 < receptor->bonds.size(); k += 2)
	{
		const int a = receptor->bonds[k], b = receptor->bonds[k + 1];
		if (new_index[a] >= 0 && new_index[b] >= 0)
			m->bonds.insert(m->bonds.end(), {new_index[a], new_index[b]});
	}
	m->nligands = ligand->nligands;
	m->publish();
	sgpu_mol* r = new sgpu_mol;
	std::vector<int> rigid_index(na, -1);
	for (int a = 0; a < na; a++)
		if (new_index[a] < 0)
		{
			rigid_index[a] = r->natoms();
			r->type.push_back(receptor->type[a]);
			r->xyz.insert(r->xyz.end(), {receptor->xyz[3 * a], receptor->xyz[3 * a + 1], receptor->xyz[3 * a + 2]});
		}
	for (size_t k = 0; k + 1 < receptor->bonds.size(); k += 2)
	{
		const int a = rigid_index[receptor->bonds[k]], b = rigid_index[receptor->bonds[k + 1]];
		if (a >= 0 && b >= 0)
			r->bonds.insert(r->bonds.end(), {a, b});
	}
	r->publish();
	*rigid = r;
	return m;
}

} // extern "C"

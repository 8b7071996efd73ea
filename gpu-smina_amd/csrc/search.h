/*
 * The docking search on one wavefront: model::eval_deriv, BFGS with its line
 * searches, mutation, Metropolis and the output container, for one chain
 * (or one item) per wave64.
 *
 * Work layout inside a wave:
 *   conf and change vectors: element k in lane k (k < 64);
 *   atoms: atom i in lane i mod 64, coordinates and forces in LDS;
 *   frames: one per lane within a tree level (set_conf top-down, derivative
 *     bottom-up, level-serial as the tree requires);
 *   intramolecular pairs: atom-centric; lane i walks atom i's pair list in
 *     evaluation order and recomputes each pair force, so every atom's force
 *     accumulates in exactly smina's order with no atomics and no pair
 *     buffer; the first atom of a pair accounts for its energy;
 *   receptor maps: one corner-packed cell gather per heavy atom.
 * Control (BFGS, line search, Metropolis, RNG) is wave-uniform: every lane
 * computes the same scalars.
 *
 * With T = double the search is "ordered": every sum runs in smina's CPU
 * order (energies serially over atoms and pairs in list order, dot products
 * serially over the change vector), reproducing the CPU restatement bit for
 * bit.  With T = float, energies use a fixed butterfly reduction.
 */
#ifndef SGPU_SEARCH_H
#define SGPU_SEARCH_H

#include "wave.h"
#include "rng.h"
#include "atom_types.h"
#include "../../include/sgpu.h"

namespace sgpu
{

// sizes of a wave's LDS region (batch maxima)
struct WavePlan
{
	int nmax;   // movable atoms
	int fmax;   // frames
	int cmax;   // conf
	int gmax;   // change
	int pmax;   // pairs (FP64 energy scratch, global)
	int hmax;   // packed Hessian entries
	int bytes;  // LDS bytes per wave
};

template<typename T>
struct WaveMem
{
	T* coords;
	T* forces;
	T* fo;
	T* fR;
	T* fq;
	T* fax;
	T* ff;
	T* ft;
	T* gbuf;
	T* ebuf;
	T* xstage;
	T* pstage;
	T* hess;
};

template<typename T>
__host__ __device__ inline int wave_lds_elems(const WavePlan& p)
{
	return 3 * p.nmax * 2 + p.fmax * (3 + 9 + 4 + 3 + 3 + 3) + p.gmax + p.nmax + p.cmax + p.gmax + p.hmax;
}

template<typename T>
__device__ inline WaveMem<T> carve(T* base, const WavePlan& p)
{
	WaveMem<T> w;
	T* q = base;
	w.coords = q;
	q += 3 * p.nmax;
	w.forces = q;
	q += 3 * p.nmax;
	w.fo = q;
	q += 3 * p.fmax;
	w.fR = q;
	q += 9 * p.fmax;
	w.fq = q;
	q += 4 * p.fmax;
	w.fax = q;
	q += 3 * p.fmax;
	w.ff = q;
	q += 3 * p.fmax;
	w.ft = q;
	q += 3 * p.fmax;
	w.gbuf = q;
	q += p.gmax;
	w.ebuf = q;
	q += p.nmax;
	w.xstage = q;
	q += p.cmax;
	w.pstage = q;
	q += p.gmax;
	w.hess = q;
	return w;
}

// the evaluation back end
template<typename T>
struct EvalCtx
{
	TablePtrs<T> tab;
	int backend;       // 0: receptor maps (cache), 1: direct receptor (non_cache)
	GridPtrs<T> grid;
	ReceptorPtrs<T> rec;
	T slope;
	T v[3];
	int* error;        // set to 1 on an exact atom overlap (non_cache)
};

// the model a wave works on, with its per-model offsets applied
template<typename T>
struct MV
{
	ModelHdr h;
	const int* type;
	const T* xyz0;
	const int* afr;
	const T* local;
	const int* nbr_start;
	const int* code;
	const int* fparent;
	const int* fkind;
	const int* fbegin;
	const int* fend;
	const int* fconf;
	const int* fchange;
	const int* cstart;
	const int* ccount;
	const int* clist;
	const T* forig;
	const T* faxis;
	const int* lstart;
	const int* lframes;
	const int* pa;
	const int* pb;
	const int* ptri;
	const int* heavy;
	const int* troot;
	const int* tconf;
	const int* tchange;
	const int* tlig;
	const int* tntors;
	const int* tab_;
	const int* tae;
	const int* tpb;
	const int* tpe;
	const int* ckind; // conf element -> ConfKind
	const int* cchg;  // conf element -> change index (orientation: its first)
	const int* gkind; // change element -> ChangeKind
	const int* gconf; // change element -> conf index (orientation: quaternion start)
};

template<typename T>
__device__ inline MV<T> model_view(const BatchPtrs<T>& B, int m)
{
	MV<T> v;
	v.h = B.hdr[m];
	const ModelHdr& h = v.h;
	v.type = B.atom_type + h.atom_off;
	v.xyz0 = B.atom_xyz0 + 3 * h.atom_off;
	v.afr = B.atom_frame + h.mov_off;
	v.local = B.atom_local + 3 * h.mov_off;
	v.nbr_start = B.nbr_start + h.nbr_off;
	v.code = B.nbr_code + h.code_off;
	v.fparent = B.frame_parent + h.frame_off;
	v.fkind = B.frame_kind + h.frame_off;
	v.fbegin = B.frame_begin + h.frame_off;
	v.fend = B.frame_end + h.frame_off;
	v.fconf = B.frame_conf + h.frame_off;
	v.fchange = B.frame_change + h.frame_off;
	v.cstart = B.child_start + h.frame_off;
	v.ccount = B.child_count + h.frame_off;
	v.clist = B.child_list + h.child_off;
	v.forig = B.frame_origin + 3 * h.frame_off;
	v.faxis = B.frame_axis + 3 * h.frame_off;
	v.lstart = B.level_start + h.level_off;
	v.lframes = B.level_frames + h.lvf_off;
	v.pa = B.pair_a + h.pair_off;
	v.pb = B.pair_b + h.pair_off;
	v.ptri = B.pair_tri + h.pair_off;
	v.heavy = B.heavy + h.heavy_off;
	v.troot = B.tree_root + h.tree_off;
	v.tconf = B.tree_conf + h.tree_off;
	v.tchange = B.tree_change + h.tree_off;
	v.tlig = B.tree_ligand + h.tree_off;
	v.tntors = B.tree_ntors + h.tree_off;
	v.tab_ = B.tree_atom_begin + h.tree_off;
	v.tae = B.tree_atom_end + h.tree_off;
	v.tpb = B.tree_pair_begin + h.tree_off;
	v.tpe = B.tree_pair_end + h.tree_off;
	v.ckind = B.maps + h.map_off;
	v.cchg = v.ckind + h.conf_size;
	v.gkind = v.cchg + h.conf_size;
	v.gconf = v.gkind + h.change_size;
	return v;
}

// ------------------------------------------------------------------ set_conf
// model::set (src/lib/model.cpp:837-841): frames level by level
// (rigid_body::set_conf, first_segment::set_conf, segment::set_conf,
// src/lib/tree.h:111-195), then every atom x = origin + R * local
template<typename T>
__device__ void set_conf(const MV<T>& m, WaveMem<T>& W, T x_lane)
{
	const int lane = lane_id();
	if (lane < m.h.conf_size)
		W.xstage[lane] = x_lane;
	wsync();
	for (int L = 0; L < m.h.nlevels; L++)
	{
		const int s = m.lstart[L], e = m.lstart[L + 1];
		for (int j = s + lane; j < e; j += WAVE)
		{
			const int f = m.lframes[j];
			const int kind = m.fkind[f];
			const int ci = m.fconf[f];
			Q4<T> q;
			V3<T> o;
			if (kind == SGPU_FRAME_RIGID_ROOT)
			{
				o = v3(W.xstage[ci], W.xstage[ci + 1], W.xstage[ci + 2]);
				q = {W.xstage[ci + 3], W.xstage[ci + 4], W.xstage[ci + 5], W.xstage[ci + 6]};
			}
			else if (kind == SGPU_FRAME_FIRST_SEGMENT)
			{
				o = ld3(m.forig + 3 * f);
				const V3<T> ax = ld3(m.faxis + 3 * f);
				q = axis_angle_q(ax, W.xstage[ci]);
				st3(W.fax + 3 * f, ax);
			}
			else
			{
				const int p = m.fparent[f];
				const T* pR = W.fR + 9 * p;
				const V3<T> po = ld3(W.fo + 3 * p);
				const Q4<T> pq = {W.fq[4 * p], W.fq[4 * p + 1], W.fq[4 * p + 2], W.fq[4 * p + 3]};
				o = po + mat_vec(pR, ld3(m.forig + 3 * f));
				const V3<T> ax = mat_vec(pR, ld3(m.faxis + 3 * f));
				st3(W.fax + 3 * f, ax);
				q = qmul(axis_angle_q(ax, W.xstage[ci]), pq);
				q_normalize_approx(q);
			}
			st3(W.fo + 3 * f, o);
			W.fq[4 * f] = q.a;
			W.fq[4 * f + 1] = q.b;
			W.fq[4 * f + 2] = q.c;
			W.fq[4 * f + 3] = q.d;
			q_to_mat(q, W.fR + 9 * f);
		}
		wsync();
	}
	for (int i = lane; i < m.h.natoms; i += WAVE)
	{
		const int f = m.afr[i];
		const V3<T> x = ld3(W.fo + 3 * f) + mat_vec(W.fR + 9 * f, ld3(m.local + 3 * i));
		st3(W.coords + 3 * i, x);
	}
	wsync();
}

template<typename T>
__device__ inline V3<T> atom_pos(const MV<T>& m, const WaveMem<T>& W, int i)
{
	return i < m.h.natoms ? ld3(W.coords + 3 * i) : ld3(m.xyz0 + 3 * i);
}

// ------------------------------------------------------- receptor map term
// grid::evaluate_aux (src/lib/grid.cpp:111-236) on a corner-packed cell
template<typename T>
__device__ inline T grid_term(const GridPtrs<T>& g, int slot, V3<T> loc, T v, V3<T>& deriv)
{
	T s[3] = {(loc.x - T(g.init[0])) * T(g.factor[0]), (loc.y - T(g.init[1])) * T(g.factor[1]),
			(loc.z - T(g.init[2])) * T(g.factor[2])};
	T miss[3] = {0, 0, 0};
	int region[3];
	int a[3];
	for (int i = 0; i < 3; i++)
	{
		const T dm1 = T(g.dim_fl_minus_1[i]);
		if (s[i] < 0)
		{
			miss[i] = -s[i];
			region[i] = -1;
			a[i] = 0;
			s[i] = 0;
		}
		else if (s[i] >= dm1)
		{
			miss[i] = s[i] - dm1;
			region[i] = 1;
			a[i] = g.dim[i] - 2;
			s[i] = 1;
		}
		else
		{
			region[i] = 0;
			a[i] = int(s[i]);
			s[i] -= T(a[i]);
		}
		// in range for every finite location; keeps a NaN location's gather in bounds
		a[i] = a[i] < 0 ? 0 : (a[i] > g.dim[i] - 2 ? g.dim[i] - 2 : a[i]);
	}
	const T penalty = T(g.slope) * (miss[0] * T(g.factor_inv[0]) + miss[1] * T(g.factor_inv[1]) +
			miss[2] * T(g.factor_inv[2]));
	const size_t cell = size_t(a[0]) + size_t(g.dim[0] - 1) * (size_t(a[1]) + size_t(g.dim[1] - 1) * a[2]);
	const T* c = g.cells + (size_t(slot) * g.ncells + cell) * 8;
	const T f000 = c[0], f100 = c[1], f010 = c[2], f110 = c[3];
	const T f001 = c[4], f101 = c[5], f011 = c[6], f111 = c[7];
	const T x = s[0], y = s[1], z = s[2];
	const T mx = 1 - x, my = 1 - y, mz = 1 - z;
	T f = f000 * mx * my * mz + f100 * x * my * mz + f010 * mx * y * mz + f110 * x * y * mz +
			f001 * mx * my * z + f101 * x * my * z + f011 * mx * y * z + f111 * x * y * z;
	// (-1) and 1 factors of the CPU expression are exact negations / no-ops
	const T x_g = -f000 * my * mz + f100 * my * mz + -f010 * y * mz + f110 * y * mz + -f001 * my * z +
			f101 * my * z + -f011 * y * z + f111 * y * z;
	const T y_g = (f000 * mx) * (-1) * mz + (f100 * x) * (-1) * mz + f010 * mx * mz + f110 * x * mz +
			(f001 * mx) * (-1) * z + (f101 * x) * (-1) * z + f011 * mx * z + f111 * x * z;
	const T z_g = (f000 * mx * my) * (-1) + (f100 * x * my) * (-1) + (f010 * mx * y) * (-1) +
			(f110 * x * y) * (-1) + f001 * mx * my + f101 * x * my + f011 * mx * y + f111 * x * y;
	V3<T> grad = v3(x_g, y_g, z_g);
	curl(f, grad, v);
	const T gr[3] = {grad.x, grad.y, grad.z};
	T d[3];
	for (int i = 0; i < 3; i++)
	{
		const T ge = region[i] == 0 ? gr[i] : T(0);
		d[i] = T(g.factor[i]) * ge + T(g.slope) * T(region[i]);
	}
	deriv = v3(d[0], d[1], d[2]);
	return f + penalty;
}

// precalculate_linear_element::eval_deriv, one component
// (src/lib/precalculate.h:111-151)
template<typename T>
__device__ inline void table_deriv(const TablePtrs<T>& tab, int tri, T r2, T& e, T& dor)
{
	const T r2f = T(tab.factor) * r2;
	const int i1 = int(r2f);
	const T rem = r2f - T(i1);
	const T* p = tab.smooth + (size_t(tri) * tab.n + i1) * 2;
	const T e1 = p[0], d1 = p[1], e2 = p[2], d2 = p[3];
	e = e1 + rem * (e2 - e1);
	dor = d1 + rem * (d2 - d1);
}

// non_cache::eval_deriv of one atom (src/lib/non_cache.cpp:109-191): box
// clamp with penalty, then the receptor atoms of the clamped point's cell
template<typename T>
__device__ inline T direct_term(const EvalCtx<T>& ec, int t1, V3<T> ac, V3<T>& out)
{
	const ReceptorPtrs<T>& r = ec.rec;
	T a[3] = {ac.x, ac.y, ac.z};
	T adj[3] = {ac.x, ac.y, ac.z};
	T oobd[3] = {0, 0, 0};
	T oob = 0;
	for (int j = 0; j < 3; j++)
		if (r.enabled[j])
		{
			if (a[j] < T(r.begin[j]))
			{
				adj[j] = T(r.begin[j]);
				oobd[j] = -1;
				oob += Real<T>::abs(a[j] - T(r.begin[j]));
			}
			else if (a[j] > T(r.end[j]))
			{
				adj[j] = T(r.end[j]);
				oobd[j] = 1;
				oob += Real<T>::abs(a[j] - T(r.end[j]));
			}
		}
	oob *= ec.slope;
	for (int j = 0; j < 3; j++)
		oobd[j] *= ec.slope;
	int ix[3];
	for (int j = 0; j < 3; j++)
	{
		ix[j] = int(__builtin_floor(double(adj[j]) / 3.0)) - r.cell_off[j];
		ix[j] = ix[j] < 0 ? 0 : (ix[j] >= r.cell_dim[j] ? r.cell_dim[j] - 1 : ix[j]);
	}
	const int cell = ix[0] + r.cell_dim[0] * (ix[1] + r.cell_dim[1] * ix[2]);
	T this_e = 0;
	V3<T> deriv = v3(T(0), T(0), T(0));
	const V3<T> adjv = v3(adj[0], adj[1], adj[2]);
	for (int k = r.cell_start[cell]; k < r.cell_start[cell + 1]; k++)
	{
		const int j = r.cell_atoms[k];
		const V3<T> rba = adjv - ld3(r.xyz + 3 * j);
		const T r2 = norm_sqr(rba);
		if (r2 < T(ec.tab.cutoff_sqr))
		{
			if (r2 < Real<T>::eps())
				*ec.error = 1;
			T pe, pd;
			table_deriv(ec.tab, tri_index_permissive(t1, r.type[j]), r2, pe, pd);
			this_e += pe;
			deriv.x += pd * rba.x;
			deriv.y += pd * rba.y;
			deriv.z += pd * rba.z;
		}
	}
	// no user grid: deriv += (0, 0, 0)
	deriv.x += T(0);
	deriv.y += T(0);
	deriv.z += T(0);
	curl(this_e, deriv, ec.v[1]);
	out = v3(deriv.x + oobd[0], deriv.y + oobd[1], deriv.z + oobd[2]);
	return this_e + oob;
}

// ------------------------------------------------------------------ eval_deriv
// model::eval_deriv (src/lib/model.cpp:969-983).  x: conf element in this
// lane; returns e (uniform) and the gradient element in g_lane.
template<typename T>
__device__ T eval_deriv(const MV<T>& m, const EvalCtx<T>& ec, WaveMem<T>& W, T* pbuf, T x_lane, T& g_lane)
{
	const int lane = lane_id();
	set_conf(m, W, x_lane);
	const int N = m.h.natoms;
	T e_part = 0;
	// inter: receptor maps or direct receptor, one heavy atom per lane
	for (int i = lane; i < N; i += WAVE)
	{
		const int t = m.type[i];
		V3<T> d = v3(T(0), T(0), T(0));
		T e = 0;
		if (t < NUM_ATOM_TYPES && !is_hydrogen_type(t))
		{
			const V3<T> x = ld3(W.coords + 3 * i);
			if (ec.backend == 0)
				e = grid_term(ec.grid, ec.grid.slot_of[t], x, ec.v[1], d);
			else
				e = direct_term(ec, t, x, d);
		}
		W.ebuf[i] = e;
		e_part += e;
		// intramolecular pairs of atom i, in evaluation order
		T fx = d.x, fy = d.y, fz = d.z;
		const V3<T> xi = ld3(W.coords + 3 * i);
		for (int k = m.nbr_start[i]; k < m.nbr_start[i + 1]; k++)
		{
			const int code = m.code[k];
			const int p = code >> 1;
			const bool is_b = (code & 1) != 0;
			const int a = m.pa[p], b = m.pb[p];
			const V3<T> xa = is_b ? atom_pos(m, W, a) : xi;
			const V3<T> xb = is_b ? xi : atom_pos(m, W, b);
			const V3<T> r = xb - xa;
			const T r2 = norm_sqr(r);
			T pe = 0;
			if (r2 < T(ec.tab.cutoff_sqr))
			{
				T dor;
				table_deriv(ec.tab, m.ptri[p], r2, pe, dor);
				V3<T> f = dor * r;
				curl(pe, f, p < m.h.nother ? ec.v[2] : ec.v[0]);
				if (is_b)
				{
					fx += f.x;
					fy += f.y;
					fz += f.z;
				}
				else
				{
					fx -= f.x;
					fy -= f.y;
					fz -= f.z;
				}
			}
			if (!is_b)
			{
				if (Real<T>::ordered)
					pbuf[p] = pe;
				else
					e_part += pe;
			}
		}
		st3(W.forces + 3 * i, v3(fx, fy, fz));
	}
	wsync();
	// energy
	T e;
	if (Real<T>::ordered)
	{
		e = 0;
		for (int i = 0; i < N; i++)
			e += W.ebuf[i];
		T eo = 0;
		for (int p = 0; p < m.h.nother; p++)
			eo += pbuf[p];
		e += eo;
		for (int t = 0; t < m.h.ntrees; t++)
		{
			if (!m.tlig[t])
				continue;
			T el = 0;
			for (int p = m.tpb[t]; p < m.tpe[t]; p++)
				el += pbuf[p];
			e += el;
		}
	}
	else
		e = wave_sum(e_part);
	// derivative: per-frame sums, then the tree bottom-up
	const int F = m.h.nframes;
	for (int f = lane; f < F; f += WAVE)
	{
		const V3<T> o = ld3(W.fo + 3 * f);
		V3<T> fo = v3(T(0), T(0), T(0)), to = v3(T(0), T(0), T(0));
		for (int i = m.fbegin[f]; i < m.fend[f]; i++)
		{
			const V3<T> fi = ld3(W.forces + 3 * i);
			fo = fo + fi;
			to = to + cross(ld3(W.coords + 3 * i) - o, fi);
		}
		st3(W.ff + 3 * f, fo);
		st3(W.ft + 3 * f, to);
	}
	wsync();
	for (int L = m.h.nlevels - 1; L >= 0; L--)
	{
		const int s = m.lstart[L], en = m.lstart[L + 1];
		for (int j = s + lane; j < en; j += WAVE)
		{
			const int f = m.lframes[j];
			V3<T> fo = ld3(W.ff + 3 * f), to = ld3(W.ft + 3 * f);
			const V3<T> o = ld3(W.fo + 3 * f);
			for (int k = 0; k < m.ccount[f]; k++)
			{
				const int c = m.clist[m.cstart[f] + k];
				const V3<T> cf = ld3(W.ff + 3 * c);
				fo = fo + cf;
				const V3<T> r = ld3(W.fo + 3 * c) - o;
				to = to + (cross(r, cf) + ld3(W.ft + 3 * c));
			}
			st3(W.ff + 3 * f, fo);
			st3(W.ft + 3 * f, to);
			const int gi = m.fchange[f];
			if (m.fkind[f] == SGPU_FRAME_RIGID_ROOT)
			{
				W.gbuf[gi] = fo.x;
				W.gbuf[gi + 1] = fo.y;
				W.gbuf[gi + 2] = fo.z;
				W.gbuf[gi + 3] = to.x;
				W.gbuf[gi + 4] = to.y;
				W.gbuf[gi + 5] = to.z;
			}
			else
				W.gbuf[gi] = dot(to, ld3(W.fax + 3 * f));
		}
		wsync();
	}
	g_lane = lane < m.h.change_size ? W.gbuf[lane] : T(0);
	wsync();
	return e;
}

} // namespace sgpu

#endif

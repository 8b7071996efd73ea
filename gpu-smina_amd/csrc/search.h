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
 *   intramolecular pairs: atom-centric; lane i walks atom i's neighbour
 *     list (packed codes in LDS) in evaluation order and recomputes each
 *     pair force, so every atom's force accumulates in exactly smina's order
 *     with no atomics and no pair buffer; the pair's first atom accounts for
 *     its energy;
 *   receptor maps: one corner-packed cell gather per heavy atom.
 * The model itself (types, frames, neighbour lists) is staged in LDS once
 * per workgroup; the hot loops read global memory only for the receptor map
 * cells and the pair tables.  Control (BFGS, line search, Metropolis, RNG)
 * is wave-uniform: every lane computes the same scalars.
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
#include "spline.h"
#include "gridterm.h"
#include "../../include/sgpu.h"

namespace sgpu
{

// neighbours per batch of the intramolecular pair loop
constexpr int PAIR_BATCH = 4;
// list entries per batch of the FP32 per-atom pair-force gather
constexpr int GATHER_BATCH = 4;
// the unordered search evaluates each pair once, pair-parallel, and sums
// each atom's pair forces from LDS; the ordered one walks each atom's
// neighbour list in smina's order, evaluating every pair from both ends
template<typename T>
constexpr bool pair_parallel()
{
	return !Real<T>::ordered;
}

// LDS of a workgroup: the staged model (int blob, real blob, map slots)
// then each wave's working state
struct WavePlan
{
	int nmax;   // movable atoms
	int fmax;   // frames
	int cmax;   // conf
	int gmax;   // change
	int pmax;   // pairs (FP64: energy scratch in global memory; FP32: pair forces in LDS)
	int hmax;   // Hessian entries (change size squared)
	int imax;   // model int blob
	int tmax;   // model real blob
	int amax;   // all atoms (movable + fixed)
	int bytes;  // LDS bytes per workgroup
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
	T* pforce; // pair forces, then prefix sums of the atoms' torque terms (pair-parallel search only)
};

// The search's own vectors in the search precision S: the staged conf and
// direction of a line-search trial, the Hessian and the BFGS update's
// scratch.  With S = T the conf stage and the scratch share the evaluation's
// buffers (xstage, gbuf).
template<typename S>
struct SearchMem
{
	S* xstage;
	S* pstage;
	S* hess;
	S* smhy;
};

// Per-wave LDS: the evaluation's region in T, then the search's in S
template<typename T>
__host__ __device__ inline int eval_lds_elems(const WavePlan& p)
{
	return 3 * p.amax + 3 * p.nmax + p.fmax * (3 + 9 + 4 + 3 + 3 + 3) + p.gmax + p.nmax + p.cmax +
			(pair_parallel<T>() ? 3 * (p.pmax > p.nmax ? p.pmax : p.nmax) : 0);
}
template<typename T, typename S>
__host__ __device__ inline int search_lds_elems(const WavePlan& p)
{
	return p.gmax + p.hmax + (sizeof(S) != sizeof(T) ? p.cmax + p.gmax : 0);
}
template<typename T, typename S>
__host__ __device__ inline int wave_lds_bytes(const WavePlan& p)
{
	const int eb = (eval_lds_elems<T>(p) * int(sizeof(T)) + 15) / 16 * 16;
	return eb + (search_lds_elems<T, S>(p) * int(sizeof(S)) + 15) / 16 * 16;
}
// model region in bytes, rounded to 16
template<typename T>
__host__ __device__ inline int model_lds_bytes(const WavePlan& p)
{
	const int ints = (p.imax + p.amax + 3) / 4 * 4;
	return ints * 4 + (p.tmax * int(sizeof(T)) + 15) / 16 * 16;
}
template<typename T, typename S = T>
__host__ __device__ inline int block_lds_bytes(const WavePlan& p, int waves)
{
	return model_lds_bytes<T>(p) + waves * wave_lds_bytes<T, S>(p);
}

// one wave's LDS (16-byte aligned) as its evaluation and search buffers
template<typename T, typename S>
__device__ inline void carve(unsigned char* base, const WavePlan& p, WaveMem<T>& w, SearchMem<S>& sm)
{
	T* q = reinterpret_cast<T*>(base);
	w.coords = q; // movable atoms, then the fixed atoms (copied once)
	q += 3 * p.amax;
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
	w.pforce = q;
	S* r = reinterpret_cast<S*>(base + (eval_lds_elems<T>(p) * int(sizeof(T)) + 15) / 16 * 16);
	sm.pstage = r;
	r += p.gmax;
	sm.hess = r;
	r += p.hmax;
	if constexpr (sizeof(S) == sizeof(T))
	{
		sm.xstage = reinterpret_cast<S*>(w.xstage);
		sm.smhy = reinterpret_cast<S*>(w.gbuf);
	}
	else
	{
		sm.xstage = r;
		r += p.cmax;
		sm.smhy = r;
	}
}

// receptor back ends: precomputed maps (smina's cache) or the receptor
// atoms themselves (non_cache)
enum Backend
{
	BK_MAPS = 0,
	BK_DIRECT = 1
};

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
	// the user grid (--user_grid), one corner-packed map: added by the
	// direct back end at each heavy atom with its slope, and by the final
	// rescoring (the maps carry it from their build)
	GridPtrs<T> ug;
	int has_ug;
};

// the model a wave works on: counts and arrays in LDS; the pair arrays
// (used outside the hot loop) in global memory
template<typename T>
struct MV
{
	ModelHdr h;
	const int* type;
	const int* slot; // map slot of each movable atom's type
	const T* xyz0;
	const int* afr;
	const T* local;
	const int* nbr_start;
	const int* code;
	const int2* prec; // the pairs as two-word records (pair-parallel loop)
	const int* fparent;
	const int* fkind;
	const int* fbegin;
	const int* fend;
	const int* fconf;
	const int* fchange;
	const int* fsend; // end of the frame's subtree atoms
	const int* fdepth; // frames on the root-to-frame path
	const int* fpath;  // that path, frame indices as bytes, MAX_PATH / 4 words per frame
	const int* cstart;
	const int* ccount;
	const int* clist;
	const T* forig;
	const T* faxis;
	const T* charge; // partial charges, movable then fixed atoms
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

// copy model m into the workgroup's LDS (every thread of the block takes
// part; the caller synchronises the block afterwards) and view it there
template<typename T>
__device__ inline MV<T> stage_model(const BatchPtrs<T>& B, int m, unsigned char* lds, const WavePlan& plan,
		const int* slot_of)
{
	const ModelHdr h = B.hdr[m];
	int* li = reinterpret_cast<int*>(lds);
	int* ls = li + plan.imax;
	T* lt = reinterpret_cast<T*>(lds + ((plan.imax + plan.amax + 3) / 4 * 4) * 4);
	const int tid = int(threadIdx.x), nt = int(blockDim.x);
	for (int k = tid; k < h.iblob_len; k += nt)
		li[k] = B.iblob[h.iblob_off + k];
	for (int k = tid; k < h.tblob_len; k += nt)
		lt[k] = B.tblob[h.tblob_off + k];
	if (slot_of)
		for (int k = tid; k < h.natoms; k += nt)
		{
			const int t = B.iblob[h.iblob_off + h.o_type + k];
			ls[k] = is_hydrogen_type(t) ? 0 : slot_of[t];
		}
	MV<T> v;
	v.h = h;
	v.type = li + h.o_type;
	v.slot = ls;
	v.xyz0 = lt + h.o_xyz0;
	v.afr = li + h.o_afr;
	v.local = lt + h.o_local;
	v.nbr_start = li + h.o_nbr;
	v.code = li + h.o_code;
	v.prec = reinterpret_cast<const int2*>(li + h.o_prec);
	v.fparent = li + h.o_fparent;
	v.fkind = li + h.o_fkind;
	v.fbegin = li + h.o_fbegin;
	v.fend = li + h.o_fend;
	v.fconf = li + h.o_fconf;
	v.fchange = li + h.o_fchange;
	v.fsend = li + h.o_fsend;
	v.fdepth = li + h.o_fdepth;
	v.fpath = li + h.o_fpath;
	v.cstart = li + h.o_cstart;
	v.ccount = li + h.o_ccount;
	v.clist = li + h.o_clist;
	v.forig = lt + h.o_forig;
	v.faxis = lt + h.o_faxis;
	v.charge = lt + h.o_charge;
	v.lstart = li + h.o_lstart;
	v.lframes = li + h.o_lframes;
	v.pa = B.pair_a + h.pair_off;
	v.pb = B.pair_b + h.pair_off;
	v.ptri = B.pair_tri + h.pair_off;
	v.heavy = li + h.o_heavy;
	v.troot = li + h.o_troot;
	v.tconf = li + h.o_tconf;
	v.tchange = li + h.o_tchange;
	v.tlig = li + h.o_tlig;
	v.tntors = li + h.o_tntors;
	v.tab_ = li + h.o_tab;
	v.tae = li + h.o_tae;
	v.tpb = li + h.o_tpb;
	v.tpe = li + h.o_tpe;
	v.ckind = li + h.o_maps;
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
__device__ __forceinline__ void set_conf(const MV<T>& m, WaveMem<T>& W, T x_lane)
{
	const int lane = lane_id();
	if (lane < m.h.conf_size)
		W.xstage[lane] = x_lane;
	lsync();
	if (pair_parallel<T>() && m.h.path_ok)
	{
		// Ancestor walk, one lane per frame (at most 64).  A segment's
		// rotation q(R_p l, theta) * q_p equals q_p * q(l, theta), l its axis
		// in the input coordinates (every frame starts unrotated), so each
		// frame's own factor is made once by its lane and the walk composes
		// the path's factors and rotates its relative origins.
		const int F = m.h.nframes;
		const int f = lane;
		if (f < F)
		{
			const int ci = m.fconf[f];
			Q4<T> lq;
			if (m.fkind[f] == SGPU_FRAME_RIGID_ROOT)
				lq = {W.xstage[ci + 3], W.xstage[ci + 4], W.xstage[ci + 5], W.xstage[ci + 6]};
			else
				lq = axis_angle_q(ld3(m.faxis + 3 * f), W.xstage[ci]);
			W.fq[4 * f] = lq.a;
			W.fq[4 * f + 1] = lq.b;
			W.fq[4 * f + 2] = lq.c;
			W.fq[4 * f + 3] = lq.d;
		}
		lsync();
		V3<T> o = v3(T(0), T(0), T(0)), ax = o;
		Q4<T> q = {T(1), T(0), T(0), T(0)};
		if (f < F)
		{
			const int d = m.fdepth[f];
			const int* path = m.fpath + f * (MAX_PATH / 4);
			const int r = path[0] & 255;
			const int rci = m.fconf[r];
			o = m.fkind[r] == SGPU_FRAME_RIGID_ROOT ? v3(W.xstage[rci], W.xstage[rci + 1], W.xstage[rci + 2])
			                                        : ld3(m.forig + 3 * r);
			q = {W.fq[4 * r], W.fq[4 * r + 1], W.fq[4 * r + 2], W.fq[4 * r + 3]};
			Q4<T> qp = {T(1), T(0), T(0), T(0)};
			for (int k = 1; k < d; k++)
			{
				const int c = (path[k >> 2] >> (8 * (k & 3))) & 255;
				o = o + q_rotate(q, ld3(m.forig + 3 * c));
				qp = q;
				q = qmul(q, Q4<T>{W.fq[4 * c], W.fq[4 * c + 1], W.fq[4 * c + 2], W.fq[4 * c + 3]});
			}
			// the world axis: the frame's own axis under its parent's rotation
			// (a first segment's parent rotation is the identity)
			if (m.fkind[f] != SGPU_FRAME_RIGID_ROOT)
				ax = q_rotate(qp, ld3(m.faxis + 3 * f));
		}
		lsync(); // every lane has read the path's factors
		if (f < F)
		{
			st3(W.fo + 3 * f, o);
			W.fq[4 * f] = q.a;
			W.fq[4 * f + 1] = q.b;
			W.fq[4 * f + 2] = q.c;
			W.fq[4 * f + 3] = q.d;
			q_to_mat(q, W.fR + 9 * f);
			if (m.fkind[f] != SGPU_FRAME_RIGID_ROOT)
				st3(W.fax + 3 * f, ax);
		}
		lsync();
	}
	else
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
		lsync();
	}
	for (int i = lane; i < m.h.natoms; i += WAVE)
	{
		const int f = m.afr[i];
		const V3<T> x = ld3(W.fo + 3 * f) + mat_vec(W.fR + 9 * f, ld3(m.local + 3 * i));
		st3(W.coords + 3 * i, x);
	}
	lsync();
}

// the fixed atoms (flexible-residue atoms that do not move) after the
// movable ones in the wave's coordinate array, once per kernel
template<typename T>
__device__ inline void init_fixed_atoms(const MV<T>& m, WaveMem<T>& W)
{
	for (int k = lane_id(); k < 3 * m.h.ninflex; k += WAVE)
		W.coords[3 * m.h.natoms + k] = m.xyz0[3 * m.h.natoms + k];
	lsync();
}

template<typename T>
__device__ inline V3<T> atom_pos(const MV<T>& m, const WaveMem<T>& W, int i)
{
	return i < m.h.natoms ? ld3(W.coords + 3 * i) : ld3(m.xyz0 + 3 * i);
}

// precalculate_splines::eval_deriv, one component, no slow terms
// (src/lib/precalculate.h:418-450) with Spline::eval_deriv
// (src/lib/splines.h:103-121): the spline at r = sqrt(r2), its derivative
// over r
// over r, in FP64 (the FP32 search rounds the results)
template<typename T>
__device__ inline void spline_deriv(const TablePtrs<T>& tab, int tri, T r2, T& e, T& dor)
{
	const double r = __builtin_sqrt(double(r2));
	double val, der;
	spline_eval(tab.spl, tab.spl_cut, tab.fraction, tab.nseg, tri, r, val, der);
	e = T(val);
	dor = T(der / r);
}

// precalculate_linear_element::eval_deriv, one component
// (src/lib/precalculate.h:111-151), or the spline's
template<typename T>
__device__ inline void table_deriv(const TablePtrs<T>& tab, int tri, T r2, T& e, T& dor)
{
	if (tab.kind == 1)
	{
		spline_deriv(tab, tri, r2, e, dor);
		return;
	}
	const T r2f = T(tab.factor) * r2;
	const int i1 = int(r2f);
	const T rem = r2f - T(i1);
	const T* p = tab.smooth + (size_t(tri) * tab.n + i1) * 2;
	const T e1 = p[0], d1 = p[1], e2 = p[2], d2 = p[3];
	e = e1 + rem * (e2 - e1);
	dor = d1 + rem * (d2 - d1);
}

// the four-component knots of a pair with charges qa, qb in table order
// (the atom of the smaller type first): result_components::eval
// (src/lib/result_components.h:73-79) of each knot's value and derivative,
// c0 + |qa| c1 + |qb| c2 + qa qb c3 summed in that order.  p: the pair's
// (e, dor) row at knot i1 in smooth, q: its row in qsmooth.
template<typename T>
__device__ inline void charged_knots(const T* p, const T* q, T qa, T qb, T& e1, T& d1, T& e2, T& d2)
{
	const T aa = Real<T>::abs(qa), ab = Real<T>::abs(qb), qq = qa * qb;
	e1 = p[0] + aa * q[0] + ab * q[2] + qq * q[4];
	d1 = p[1] + aa * q[1] + ab * q[3] + qq * q[5];
	e2 = p[2] + aa * q[6] + ab * q[8] + qq * q[10];
	d2 = p[3] + aa * q[7] + ab * q[9] + qq * q[11];
}

// precalculate_linear::eval_deriv (src/lib/precalculate.h:245-293) of
// atoms of types t1, t2 and charges q1, q2 with four-component tables: the
// element of (min type, max type) with the atoms in that order
template<typename T>
__device__ inline void table_deriv_charged(const TablePtrs<T>& tab, int t1, int t2, T q1, T q2, T r2, T& e, T& dor)
{
	const T r2f = T(tab.factor) * r2;
	const int i1 = int(r2f);
	const T rem = r2f - T(i1);
	const size_t k = size_t(tri_index_permissive(t1, t2)) * tab.n + i1;
	const bool in_order = t1 <= t2;
	T e1, d1, e2, d2;
	charged_knots(tab.smooth + 2 * k, tab.qsmooth + 6 * k, in_order ? q1 : q2, in_order ? q2 : q1, e1, d1, e2, d2);
	e = e1 + rem * (e2 - e1);
	dor = d1 + rem * (d2 - d1);
}

// non_cache::eval_deriv of one atom (src/lib/non_cache.cpp:109-191): box
// clamp with penalty, then the receptor atoms of the clamped point's cell.
// q1: the atom's partial charge (read with four-component tables only).
template<typename T>
__device__ inline T direct_term(const EvalCtx<T>& ec, int t1, T q1, V3<T> ac, V3<T>& out)
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
				atomicOr(ec.error, 1); // bit 0: exact overlap (other bits belong to the caller)
			T pe, pd;
			if (ec.tab.ncomp == 4)
				table_deriv_charged(ec.tab, t1, r.type[j], q1, r.charge[j], r2, pe, pd);
			else
				table_deriv(ec.tab, tri_index_permissive(t1, r.type[j]), r2, pe, pd);
			this_e += pe;
			deriv.x += pd * rba.x;
			deriv.y += pd * rba.y;
			deriv.z += pd * rba.z;
		}
	}
	// the user grid at the atom's own position, with the back end's slope
	V3<T> ug_deriv = v3(T(0), T(0), T(0));
	if (ec.has_ug)
		this_e += grid_finish(ec.ug, grid_fetch(ec.ug, 0, ac), T(1000), ec.slope, ug_deriv);
	deriv.x += ug_deriv.x;
	deriv.y += ug_deriv.y;
	deriv.z += ug_deriv.z;
	curl(this_e, deriv, ec.v[1]);
	out = v3(deriv.x + oobd[0], deriv.y + oobd[1], deriv.z + oobd[2]);
	return this_e + oob;
}

// ------------------------------------------------------------------ eval_deriv
// model::eval_deriv (src/lib/model.cpp:969-983).  x: conf element in this
// lane; returns e (uniform) and the gradient element in g_lane.  pbuf: the
// wave's pair-energy scratch in global memory (ordered mode only).  BK: the
// receptor back end (BK_MAPS or BK_DIRECT), fixed per kernel.  SPL: the
// pair loops can read the spline tables (off in the search kernel of
// smina's own search, which keeps its registers without them).
template<typename T, int BK, bool SPL = true>
__device__ __forceinline__ T eval_deriv(const MV<T>& m, const EvalCtx<T>& ec, WaveMem<T>& W, T* pbuf, T x_lane,
		T& g_lane)
{
	const int lane = lane_id();
	set_conf(m, W, x_lane);
	const int N = m.h.natoms;
	const T cutoff_sqr = T(ec.tab.cutoff_sqr);
	T e_part = 0;
	if constexpr (!pair_parallel<T>())
	{
		// atom-centric: every force accumulates in smina's order
		for (int i = lane; i < N; i += WAVE)
		{
			const int t = m.type[i];
			const V3<T> xi = ld3(W.coords + 3 * i);
			V3<T> d = v3(T(0), T(0), T(0));
			T e = 0;
			if (!is_hydrogen_type(t))
			{
				if constexpr (BK == BK_MAPS)
				{
					e = grid_term(ec.grid, m.slot[i], xi, ec.v[1], d);
					if (SPL && ec.grid.qcells && m.charge[i] != T(0))
						grid_charge_term(ec.grid, m.slot[i], xi, ec.v[1], m.charge[i], e, d);
				}
				else
					e = direct_term(ec, t, m.charge[i], xi, d);
			}
			if constexpr (Real<T>::ordered)
				W.ebuf[i] = e;
			else
				e_part += e;
			// intramolecular pairs of atom i, in evaluation order, PB at a time:
			// the batch's table rows are fetched together, then accumulated in
			// list order
			T fx = d.x, fy = d.y, fz = d.z;
			const int k0 = m.nbr_start[i], k1 = m.nbr_start[i + 1];
			for (int k = k0; k < k1; k += PAIR_BATCH)
			{
				int cd[PAIR_BATCH];
				T rx[PAIR_BATCH], ry[PAIR_BATCH], rz[PAIR_BATCH], r2[PAIR_BATCH], rem[PAIR_BATCH];
				T e1[PAIR_BATCH], d1[PAIR_BATCH], e2[PAIR_BATCH], d2[PAIR_BATCH];
#pragma unroll
				for (int u = 0; u < PAIR_BATCH; u++)
				{
					const int kk = k + u < k1 ? k + u : k0;
					const int code = m.code[kk];
					cd[u] = code;
					const int j = code_partner(code);
					const bool is_b = code_is_b(code);
					const V3<T> xj = atom_pos(m, W, j);
					// r = x_b - x_a
					rx[u] = is_b ? xi.x - xj.x : xj.x - xi.x;
					ry[u] = is_b ? xi.y - xj.y : xj.y - xi.y;
					rz[u] = is_b ? xi.z - xj.z : xj.z - xi.z;
					r2[u] = sqr(rx[u]) + sqr(ry[u]) + sqr(rz[u]);
					// precalculate_linear_element::eval_deriv (src/lib/precalculate.h:111-151)
					if (SPL && ec.tab.kind == 1)
					{
						// the spline's value and derivative over r, carried
						// through the interpolation below unchanged (rem 0)
						if (r2[u] < cutoff_sqr)
							spline_deriv(ec.tab, code_tri(code), r2[u], e1[u], d1[u]);
						else
							e1[u] = d1[u] = 0;
						rem[u] = 0;
						e2[u] = e1[u];
						d2[u] = d1[u];
						continue;
					}
					const T r2f = T(ec.tab.factor) * r2[u];
					const int i1 = r2[u] < cutoff_sqr ? int(r2f) : 0;
					rem[u] = r2f - T(i1);
					const size_t kn = size_t(code_tri(code)) * ec.tab.n + i1;
					const T* tp = ec.tab.smooth + kn * 2;
					if (SPL && ec.tab.ncomp == 4)
					{
						// four components: the pair's charges in table order
						const int ia = is_b ? j : i, ib = is_b ? i : j;
						const bool in_order = m.type[ia] <= m.type[ib];
						const T qa = m.charge[in_order ? ia : ib], qb = m.charge[in_order ? ib : ia];
						charged_knots(tp, ec.tab.qsmooth + kn * 6, qa, qb, e1[u], d1[u], e2[u], d2[u]);
						continue;
					}
					e1[u] = tp[0];
					d1[u] = tp[1];
					e2[u] = tp[2];
					d2[u] = tp[3];
				}
#pragma unroll
				for (int u = 0; u < PAIR_BATCH; u++)
				{
					if (k + u >= k1)
						break;
					const int code = cd[u];
					const bool is_b = code_is_b(code);
					T pe = 0;
					if (r2[u] < cutoff_sqr)
					{
						pe = e1[u] + rem[u] * (e2[u] - e1[u]);
						const T dor = d1[u] + rem[u] * (d2[u] - d1[u]);
						V3<T> f = v3(dor * rx[u], dor * ry[u], dor * rz[u]);
						curl(pe, f, code_other(code) ? ec.v[2] : ec.v[0]);
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
						if constexpr (Real<T>::ordered)
							pbuf[code_pair(code)] = pe;
						else
							e_part += pe;
					}
				}
			}
			st3(W.forces + 3 * i, v3(fx, fy, fz));
		}
	}
	else
	{
		// intramolecular pairs, pair-parallel (pair t * 64 + lane), each
		// pair's force into its own LDS slot
		const int P = m.h.npairs;
		// the curl caps of the two pair kinds; a cap below epsilon zeroes the
		// term, as v < epsilon does in curl
		const T v_lig = ec.v[0] < Real<T>::eps() ? T(0) : ec.v[0];
		const T v_oth = ec.v[2] < Real<T>::eps() ? T(0) : ec.v[2];
		const T factor = T(ec.tab.factor);
		const int last_knot = ec.tab.n - 2;
		const int row_stride = 2 * ec.tab.n;
		for (int t0 = 0; t0 * WAVE < P; t0 += PAIR_BATCH)
		{
			// phases across the batch: records, coordinates, table rows, then
			// the arithmetic, so each phase's loads are in flight together
			int2 rc[PAIR_BATCH];
#pragma unroll
			for (int u = 0; u < PAIR_BATCH; u++)
			{
				const int p = (t0 + u) * WAVE + lane;
				rc[u] = m.prec[p < P ? p : 0];
			}
			V3<T> xa[PAIR_BATCH], xb[PAIR_BATCH];
#pragma unroll
			for (int u = 0; u < PAIR_BATCH; u++)
			{
				xa[u] = ld3(W.coords + prec_xa(rc[u].x));
				xb[u] = ld3(W.coords + prec_xb(rc[u].x));
			}
			T rx[PAIR_BATCH], ry[PAIR_BATCH], rz[PAIR_BATCH], r2[PAIR_BATCH], rem[PAIR_BATCH];
			T e1[PAIR_BATCH], d1[PAIR_BATCH], e2[PAIR_BATCH], d2[PAIR_BATCH];
#pragma unroll
			for (int u = 0; u < PAIR_BATCH; u++)
			{
				rx[u] = xb[u].x - xa[u].x;
				ry[u] = xb[u].y - xa[u].y;
				rz[u] = xb[u].z - xa[u].z;
				r2[u] = sqr(rx[u]) + sqr(ry[u]) + sqr(rz[u]);
				// precalculate_linear_element::eval_deriv (src/lib/precalculate.h:111-151);
				// the knot is clamped so that the row read stays in the table.
				// With the spline tables, the spline's value and derivative
				// over r pass through the interpolation below unchanged.
				if (SPL && ec.tab.kind == 1)
				{
					if (r2[u] < cutoff_sqr)
						spline_deriv(ec.tab, prec_tri(rc[u].y), r2[u], e1[u], d1[u]);
					else
						e1[u] = d1[u] = T(0);
					rem[u] = T(0);
					e2[u] = e1[u];
					d2[u] = d1[u];
					continue;
				}
				const T r2f = factor * r2[u];
				const int i1 = min(int(r2f), last_knot);
				rem[u] = r2f - T(i1);
				const T* tp = ec.tab.smooth + int(__umul24(unsigned(prec_tri(rc[u].y)), unsigned(row_stride)) + 2 * i1);
				if (SPL && ec.tab.ncomp == 4)
				{
					// four components: the pair's charges in table order
					const int ia = prec_xa(rc[u].x) / 3, ib = prec_xb(rc[u].x) / 3;
					const bool in_order = m.type[ia] <= m.type[ib];
					const T qa = m.charge[in_order ? ia : ib], qb = m.charge[in_order ? ib : ia];
					charged_knots(tp, ec.tab.qsmooth + 3 * (tp - ec.tab.smooth), qa, qb, e1[u], d1[u], e2[u], d2[u]);
					continue;
				}
				e1[u] = tp[0];
				d1[u] = tp[1];
				e2[u] = tp[2];
				d2[u] = tp[3];
			}
#pragma unroll
			for (int u = 0; u < PAIR_BATCH; u++)
			{
				const int p = (t0 + u) * WAVE + lane;
				const bool ok = p < P && r2[u] < cutoff_sqr;
				// interpolation and curl (src/lib/curl.h:28-43) without branches
				T pe = e1[u] + rem[u] * (e2[u] - e1[u]);
				const T dor = d1[u] + rem[u] * (d2[u] - d1[u]);
				const T v = prec_other(rc[u].y) ? v_oth : v_lig;
				const T k = pe > 0 ? Real<T>::fdiv(v, v + pe) : T(1);
				const T fk = ok ? dor * (k * k) : T(0);
				e_part += ok ? pe * k : T(0);
				if (p < P)
					st3(W.pforce + 3 * p, v3(fk * rx[u], fk * ry[u], fk * rz[u]));
			}
		}
		lsync();
		// per movable atom: receptor term, then its pair forces in list order;
		// the map gather is issued first and consumed last, so its latency
		// overlaps the list walk
		for (int i = lane; i < N; i += WAVE)
		{
			const int t = m.type[i];
			const bool heavy = !is_hydrogen_type(t);
			const V3<T> xi = ld3(W.coords + 3 * i);
			GridSample<T> gs;
			if constexpr (BK == BK_MAPS)
				gs = grid_fetch(ec.grid, heavy ? m.slot[i] : 0, xi);
			V3<T> fp = v3(T(0), T(0), T(0));
			const int k0 = m.nbr_start[i], k1 = m.nbr_start[i + 1];
			for (int k = k0; k < k1; k += GATHER_BATCH)
			{
				int cd[GATHER_BATCH];
#pragma unroll
				for (int u = 0; u < GATHER_BATCH; u++)
					cd[u] = m.code[k + u < k1 ? k + u : k0];
				V3<T> f[GATHER_BATCH];
#pragma unroll
				for (int u = 0; u < GATHER_BATCH; u++)
					f[u] = ld3(W.pforce + 3 * code_pair(cd[u]));
#pragma unroll
				for (int u = 0; u < GATHER_BATCH; u++)
				{
					const T sg = k + u >= k1 ? T(0) : (code_is_b(cd[u]) ? T(1) : T(-1));
					fp = v3(fp.x + sg * f[u].x, fp.y + sg * f[u].y, fp.z + sg * f[u].z);
				}
			}
			V3<T> d = v3(T(0), T(0), T(0));
			if (heavy)
			{
				if constexpr (BK == BK_MAPS)
				{
					T ea = grid_finish(ec.grid, gs, ec.v[1], d);
					if (SPL && ec.grid.qcells && m.charge[i] != T(0))
						grid_charge_term(ec.grid, m.slot[i], xi, ec.v[1], m.charge[i], ea, d);
					e_part += ea;
				}
				else
					e_part += direct_term(ec, t, m.charge[i], xi, d);
			}
			st3(W.forces + 3 * i, d + fp);
		}
		lsync();
		// Inclusive prefix sums over the atoms of the force (in place) and of
		// (x - o0) x force (o0: frame 0's origin; into the pair-force
		// buffer, free now): every frame's subtree is a contiguous atom range,
		// so the derivative reads its sums as two differences
		const V3<T> o0 = ld3(W.fo);
		V3<T> carry_f = v3(T(0), T(0), T(0)), carry_t = carry_f;
		for (int base = 0; base < N; base += WAVE)
		{
			const int i = base + lane;
			V3<T> force = v3(T(0), T(0), T(0)), tq = force;
			if (i < N)
			{
				force = ld3(W.forces + 3 * i);
				tq = cross(ld3(W.coords + 3 * i) - o0, force);
			}
			const V3<T> sf = v3(wave_inclusive_scan(force.x), wave_inclusive_scan(force.y),
							  wave_inclusive_scan(force.z)) + carry_f;
			const V3<T> st = v3(wave_inclusive_scan(tq.x), wave_inclusive_scan(tq.y), wave_inclusive_scan(tq.z)) +
					carry_t;
			if (i < N)
			{
				st3(W.forces + 3 * i, sf);
				st3(W.pforce + 3 * i, st);
			}
			carry_f = v3(readlane(sf.x, WAVE - 1), readlane(sf.y, WAVE - 1), readlane(sf.z, WAVE - 1));
			carry_t = v3(readlane(st.x, WAVE - 1), readlane(st.y, WAVE - 1), readlane(st.z, WAVE - 1));
		}
	}
	T e;
	if (Real<T>::ordered)
	{
		gsync(); // the pair energies went through global memory
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
	{
		lsync();
		e = wave_sum(e_part);
	}
	if constexpr (pair_parallel<T>())
	{
		// derivative, direct: each frame's force and torque over its
		// subtree's atoms (a contiguous range) from the prefix sums, equal to
		// smina's bottom-up accumulation (src/lib/tree.h:88-97, :214-289) up
		// to rounding: torque about the frame's origin o is
		// sum (x - o0) x f - (o - o0) x sum f
		const int F = m.h.nframes;
		const int f = lane;
		if (f < F)
		{
			const int b = m.fbegin[f], en = m.fsend[f];
			V3<T> fs = ld3(W.forces + 3 * (en - 1)), ts = ld3(W.pforce + 3 * (en - 1));
			if (b > 0)
			{
				fs = fs - ld3(W.forces + 3 * (b - 1));
				ts = ts - ld3(W.pforce + 3 * (b - 1));
			}
			const V3<T> to = ts - cross(ld3(W.fo + 3 * f) - ld3(W.fo), fs);
			const int gi = m.fchange[f];
			if (m.fkind[f] == SGPU_FRAME_RIGID_ROOT)
			{
				W.gbuf[gi] = fs.x;
				W.gbuf[gi + 1] = fs.y;
				W.gbuf[gi + 2] = fs.z;
				W.gbuf[gi + 3] = to.x;
				W.gbuf[gi + 4] = to.y;
				W.gbuf[gi + 5] = to.z;
			}
			else
				W.gbuf[gi] = dot(to, ld3(W.fax + 3 * f));
		}
		lsync();
	}
	else
	{
		// derivative: per-frame sums, then the tree bottom-up
		const int F = m.h.nframes;
		for (int f = lane; f < F; f += WAVE)
		{
			const V3<T> o = ld3(W.fo + 3 * f);
			V3<T> fo = v3(T(0), T(0), T(0)), to = v3(T(0), T(0), T(0));
			const int b = m.fbegin[f], en = m.fend[f];
	#pragma unroll 4
			for (int i = b; i < en; i++)
			{
				const V3<T> fi = ld3(W.forces + 3 * i);
				fo = fo + fi;
				to = to + cross(ld3(W.coords + 3 * i) - o, fi);
			}
			st3(W.ff + 3 * f, fo);
			st3(W.ft + 3 * f, to);
		}
		lsync();
		for (int L = m.h.nlevels - 1; L >= 0; L--)
		{
			const int s = m.lstart[L], en = m.lstart[L + 1];
			for (int j = s + lane; j < en; j += WAVE)
			{
				const int f = m.lframes[j];
				V3<T> fo = ld3(W.ff + 3 * f), to = ld3(W.ft + 3 * f);
				const V3<T> o = ld3(W.fo + 3 * f);
				const int cs = m.cstart[f], cc = m.ccount[f];
				for (int k = 0; k < cc; k++)
				{
					const int c = m.clist[cs + k];
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
			lsync();
		}
	}
	g_lane = lane < m.h.change_size ? W.gbuf[lane] : T(0);
	lsync();
	return e;
}

} // namespace sgpu

#endif

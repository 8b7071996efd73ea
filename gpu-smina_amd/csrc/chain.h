/*
 * BFGS, line searches, mutation, Metropolis, the output container and the
 * Monte Carlo chain on one wavefront (src/lib/bfgs.h, src/lib/mutate.cpp,
 * src/lib/monte_carlo.cpp, src/lib/coords.cpp; the reference's device
 * statement of the same semantics is src/gpu/gpu_mc.cuh).
 *
 * Vectors of the change and conf spaces live one element per lane; the
 * Hessian lives in LDS as the full symmetric matrix.  Scalars are
 * wave-uniform.  With T = double every dot product is a serial sum in
 * index order, as on the CPU.
 *
 * Two real types: T for the model evaluation (coordinates, maps, pair
 * tables, forces) and S for the search state (conf, gradient and direction
 * vectors, Hessian, line search, energies the chain compares).  The served
 * search uses T = S = float, the oracle-parity search T = S = double, and
 * the mixed search T = float with S = double.
 */
#ifndef SGPU_CHAIN_H
#define SGPU_CHAIN_H

#include "search.h"

namespace sgpu
{

// scalar_product (src/lib/bfgs.h:44-50)
template<typename T>
__device__ inline T dot_n(T a, T b, int n)
{
	const T prod = lane_id() < n ? a * b : T(0);
	if (Real<T>::ordered)
	{
		T s = 0;
		for (int k = 0; k < n; k++)
			s += readlane(prod, k);
		return s;
	}
	return wave_sum(prod);
}

// The Hessian is held in LDS as the full symmetric n x n matrix, h(i, j) at
// i * n + j, both halves written with the same value, so reads by column
// are conflict-free.  An n x n sweep gives each lane one entry per pass:
// HessMap holds the lane's (row offset, column) and the rows per pass.
struct HessMap
{
	int rows; // rows per pass (64 / n)
	int q;    // this lane's row within a pass
	int c;    // this lane's column
};
__device__ inline HessMap hess_map(int n)
{
	const int lane = lane_id();
	HessMap hm;
	hm.rows = WAVE / n;
	hm.q = lane / n;
	hm.c = lane - hm.q * n;
	return hm;
}

// minus_mat_vec_product (src/lib/bfgs.h:31-42)
template<typename T>
__device__ inline T minus_mat_vec(const T* h, T in, int n)
{
	const int lane = lane_id();
	const T* col = h + (lane < n ? lane : 0);
	T sum = 0;
#pragma unroll 4
	for (int j = 0; j < n; j++)
	{
		const T inj = readlane(in, j);
		sum += col[j * n] * inj;
	}
	return lane < n ? -sum : -T(0);
}

// bfgs_update (src/lib/bfgs.h:52-72): entry (i, j) of the upper triangle
// and its mirror take the same value, each computed by its own lane with
// the CPU's expression on (min, max) of the two indices
template<typename T>
__device__ inline void bfgs_update(T* h, T* sp, T* smhy, T p, T y, T alpha, int n, const HessMap& hm)
{
	const T yp = dot_n(y, p, n);
	if (alpha * yp < Real<T>::eps())
		return;
	const T mhy = minus_mat_vec(h, y, n);
	const T yhy = -dot_n(y, mhy, n);
	const T r = Real<T>::rcp(alpha * yp);
	const int lane = lane_id();
	if (lane < n)
	{
		sp[lane] = p;
		smhy[lane] = mhy;
	}
	lsync();
	// this lane's column values, then one row value pair per pass
	const bool col_ok = hm.q < hm.rows && hm.c < n;
	const T pc = col_ok ? sp[hm.c] : T(0), mc = col_ok ? smhy[hm.c] : T(0);
#pragma unroll 2
	for (int r0 = 0; r0 < n; r0 += hm.rows)
	{
		const int row = r0 + hm.q;
		if (col_ok && row < n)
		{
			const T pr = sp[row], mr = smhy[row];
			// (i, j) = (min, max) of (row, column)
			const bool lower = row < hm.c;
			const T pi = lower ? pr : pc, pj = lower ? pc : pr, mi = lower ? mr : mc, mj = lower ? mc : mr;
			h[row * n + hm.c] += alpha * r * (mi * pj + mj * pi) + +alpha * alpha * (r * r * yhy + r) * pi * pj;
		}
	}
	lsync();
}

// conf::increment (src/lib/conf.h:316-321) of the staged conf by
// alpha * staged change; returns this lane's new conf element
template<typename T, typename S>
__device__ inline S conf_increment(const MV<T>& m, const SearchMem<S>& W, S alpha)
{
	const int k = lane_id();
	if (k >= m.h.conf_size)
		return S(0);
	const int kind = m.ckind[k];
	const int gi = m.cchg[k];
	const S x = W.xstage[k];
	if (kind == CK_POSITION)
	{
		const S d = alpha * W.pstage[gi];
		return x + d;
	}
	if (kind == CK_TORSION)
	{
		S t = x + normalized_angle(alpha * W.pstage[gi]);
		normalize_angle(t);
		return t;
	}
	const int comp = kind - CK_QUAT0;
	const int q0 = k - comp;
	const V3<S> rot = v3(alpha * W.pstage[gi], alpha * W.pstage[gi + 1], alpha * W.pstage[gi + 2]);
	const Q4<S> q = {W.xstage[q0], W.xstage[q0 + 1], W.xstage[q0 + 2], W.xstage[q0 + 3]};
	const Q4<S> r = q_increment(q, rot);
	return comp == 0 ? r.a : comp == 1 ? r.b : comp == 2 ? r.c : r.d;
}

// conf::operator()(index) (src/lib/conf.h:376-397) for the staged conf
template<typename T, typename S>
__device__ inline S conf_value(const MV<T>& m, const SearchMem<S>& W)
{
	const int i = lane_id();
	if (i >= m.h.change_size)
		return S(0);
	const int kind = m.gkind[i];
	const int ci = m.gconf[i];
	if (kind == GK_POSITION || kind == GK_TORSION)
		return W.xstage[ci];
	const Q4<S> q = {W.xstage[ci], W.xstage[ci + 1], W.xstage[ci + 2], W.xstage[ci + 3]};
	const V3<S> a = q_to_angle(q);
	const int comp = kind - GK_ORIENT0;
	return comp == 0 ? a.x : comp == 1 ? a.y : a.z;
}

template<typename S>
__device__ inline void stage(SearchMem<S>& W, S x, S p, int c, int n)
{
	const int lane = lane_id();
	if (lane < c)
		W.xstage[lane] = x;
	if (lane < n)
		W.pstage[lane] = p;
	lsync();
}

// QuickVina 2's history of visited points (not smina's search; the
// GPU-style --fast search, src/gpu/gpu_mc.cuh:146-264; Alhossary et al.,
// Bioinformatics 2015): a ring of the last 5n points a chain's local
// optimizations visited (n = change size <= 64), each with its energy and
// the masks of its zero and positive gradient components.  Points are stored
// coordinate-major, x[i * stride + e]: the lanes of a wave read 64 entries of
// one coordinate in one coalesced load.  The ring's state (size, next slot,
// full) is uniform and lives in registers during a launch.
template<typename S>
struct Visited
{
	S* x;            // null: no history
	S* f;
	uint64_t* bits;  // [2e] zero mask, [2e + 1] positive mask
	int stride;      // 5 x the largest change size the storage holds
	int size, next, full;

	__device__ __forceinline__ bool on(int n) const { return x != nullptr && 5 * n <= stride; }

	// visited::interesting: until the ring is full every point is;
	// otherwise the nearest stored point (by squared distance, then slot)
	// that passes the first-order consistency check must rank below 4n.  A
	// point passes unless a coordinate where both gradients have the same
	// nonzero sign breaks (g > 0) == (now > x_e) xor (f > f_e).  now, g:
	// this lane's change-indexed component.
	__device__ __forceinline__ bool interesting(int n, S now, S fv, S g) const
	{
		if (!full)
			return true;
		const int cap = 5 * n, lane = lane_id();
		S d[5];
		S best = Real<S>::max();
		int besti = cap;
#pragma unroll
		for (int k = 0; k < 5; k++)
		{
			const int e = lane + WAVE * k;
			const bool valid = e < cap;
			const bool newY = valid && fv - f[e] > 0;
			const uint64_t zb = valid ? bits[2 * e] : 0, pb = valid ? bits[2 * e + 1] : 0;
			S d2 = 0;
			bool pass = true;
			for (int i = 0; i < n; i++)
			{
				const S nowi = readlane(now, i), gi = readlane(g, i);
				const S xe = valid ? x[size_t(i) * stride + e] : S(0);
				const S dx = xe - nowi;
				d2 += dx * dx;
				if (((zb >> i) & 1) || gi == 0)
					continue;
				const bool nowPos = gi > 0;
				if (nowPos != bool((pb >> i) & 1))
					continue;
				const bool newX = nowi - xe > 0;
				if (nowPos ^ newX ^ newY)
					pass = false;
			}
			d[k] = d2;
			if (valid && pass && d2 < best)
			{
				best = d2;
				besti = e;
			}
		}
		for (int off = WAVE / 2; off >= 1; off >>= 1)
		{
			const S ob = __shfl_xor(best, off);
			const int oi = __shfl_xor(besti, off);
			if (ob < best || (ob == best && oi < besti))
			{
				best = ob;
				besti = oi;
			}
		}
		if (besti >= cap)
			return false;
		int ahead = 0;
#pragma unroll
		for (int k = 0; k < 5; k++)
		{
			const int e = lane + WAVE * k;
			ahead += __popcll(__ballot(e < cap && (d[k] < best || (d[k] == best && e < besti))));
		}
		return ahead < 4 * n;
	}

	// visited_add: the point takes the next slot of the ring
	__device__ __forceinline__ void add(int n, S xv, S fv, S g)
	{
		const int cap = 5 * n, lane = lane_id();
		int slot;
		if (!full)
		{
			slot = size++;
			if (size >= cap)
			{
				full = 1;
				next = 0;
			}
		}
		else
		{
			slot = next;
			next = next + 1 == cap ? 0 : next + 1;
		}
		const uint64_t zero = __ballot(lane < n && g == 0), pos = __ballot(lane < n && g > 0);
		if (lane == 0)
		{
			f[slot] = fv;
			bits[2 * slot] = zero;
			bits[2 * slot + 1] = pos;
		}
		if (lane < n)
			x[size_t(lane) * stride + slot] = xv;
		gsync();
	}
};

// Speculative trials of the fast line search.  When a launch has fewer
// chains than the device has wave slots, a workgroup gives one chain
// `width` waves: wave 0 runs the chain, and at each round of its fast line
// search the other waves evaluate the trials the round would reach next.
// Wave 0 posts the staged conf x, the direction p, the step alpha and the
// curl caps (barrier A); wave w evaluates x + alpha 2^-w p while wave 0
// evaluates x + alpha p (the halvings are exact, so these are the confs
// the sequential search evaluates); after barrier B wave 0 takes the
// trials in order, as the sequential search would, and the first that
// passes the Armijo test ends the line search.  The chain's confs,
// energies and evaluation counts are the sequential search's.  The helpers
// touch no chain state.
constexpr int SPEC_MAX = 8;
constexpr int SPEC_REQ = MAX_CONF + MAX_CHANGE + 4;  // x, p, alpha, caps
constexpr int SPEC_RS = 1 + MAX_CONF + MAX_CHANGE;   // per wave: f, conf, gradient
template<typename S>
struct SpecLink
{
	int width;   // waves per chain (0 or 1: no speculation)
	int* cmd;    // 0: a request is posted, 1: the chain is done
	S* req;      // [SPEC_REQ]
	S* res;      // [SPEC_MAX][SPEC_RS]
};

// bfgs (src/lib/bfgs.h:201-282) with fast_line_search (:76-96) or
// accurate_line_search (:101-190), written as one loop around a single
// model evaluation: the kernels inline the evaluation once.  Each pass of
// the loop evaluates the conf the state machine asks for next: the start
// point, then the line search's trials; a finished line search runs the
// BFGS step body and starts the next line search.  The order of every
// evaluation and every arithmetic operation is the CPU's.
template<typename T, typename S, int BK, bool VIS = false, bool SPL = true, bool SPEC = false>
struct Searcher
{
	const MV<T>& m;
	const EvalCtx<T>& ec;
	WaveMem<T>& W;
	SearchMem<S>& SM;
	T* pbuf;
	int64_t evals;
	Visited<S> vis; // QuickVina 2's history (vis.x null: none)
	SpecLink<S> spec{}; // speculative trials (wave 0 of a chain's workgroup)

	// post a round of trials at alpha (barrier A); the helpers start on it
	__device__ __forceinline__ void spec_post(S x, S p, S alpha)
	{
		const int lane = lane_id();
		if (lane < m.h.conf_size)
			spec.req[lane] = x;
		if (lane < m.h.change_size)
			spec.req[MAX_CONF + lane] = p;
		if (lane == 0)
		{
			spec.req[MAX_CONF + MAX_CHANGE] = alpha;
			for (int d = 0; d < 3; d++)
				spec.req[MAX_CONF + MAX_CHANGE + 1 + d] = S(ec.v[d]);
			*spec.cmd = 0;
		}
		__syncthreads();
	}

	// this lane's change-indexed component of the conf x
	__device__ __forceinline__ S change_value(S x)
	{
		if (lane_id() < m.h.conf_size)
			SM.xstage[lane_id()] = x;
		lsync();
		const S v = conf_value(m, SM);
		lsync();
		return v;
	}

	// the staged conf x plus alpha * p, this lane's element
	__device__ __forceinline__ S trial_conf(S x, S p, S alpha)
	{
		stage(SM, x, p, m.h.conf_size, m.h.change_size);
		const S xn = conf_increment(m, SM, alpha);
		lsync();
		return xn;
	}

	// x, g: this lane's elements; returns the final energy.  With a history
	// (src/gpu/gpu_mc.cuh:440-515): an uninteresting start returns its
	// energy unoptimized; the start and every completed step join the history.
	__device__ __forceinline__ S bfgs(S& x, S& g, int maxiters, bool accurate, bool early_term)
	{
		const int n = m.h.change_size, c = m.h.conf_size;
		const int lane = lane_id();
		const bool history = VIS && vis.on(n);
		const bool speculate = SPEC && spec.width > 1 && !accurate;
		// identity
		const HessMap hm = hess_map(n > 0 ? n : 1);
		for (int r0 = 0; r0 < n; r0 += hm.rows)
		{
			const int row = r0 + hm.q;
			if (hm.q < hm.rows && row < n)
				SM.hess[row * n + hm.c] = row == hm.c ? S(1) : S(0);
		}
		lsync();
		S f0 = 0, f_orig = 0, g_orig = 0, x_orig = x;
		S p = 0, x_new = x, g_new = 0, f1 = 0;
		S alpha = 1, slope = 0;   // slope: p.g (fast) or g.p (accurate)
		S alamin = 0, alpha2 = 0, f2 = 0;
		int trial = 0, step = 0;
		bool initial = true;
		S xe = x; // the conf evaluated next
		// speculative rounds: whether the last trial taken in the sequential
		// order was a helper's (then this wave's coordinates are not that
		// trial's), and that trial's conf
		bool helper_last = false;
		S x_last = 0;
		for (;;)
		{
			++evals;
			T ge_t;
			const S fe = S(eval_deriv<T, BK, SPL>(m, ec, W, pbuf, T(xe), ge_t));
			const S ge = S(ge_t);
			helper_last = false;
			bool begin = false;
			if (initial)
			{
				initial = false;
				f0 = fe;
				g = ge;
				f_orig = f0;
				g_orig = g;
				x_orig = x;
				x_new = x;
				g_new = g;
				if (history)
				{
					const S xv = change_value(x);
					if (!vis.interesting(n, xv, f0, g))
						break;
					vis.add(n, xv, f0, g);
				}
				if (maxiters <= 0)
					break;
				begin = true;
			}
			else
			{
				// a line-search trial at alpha
				f1 = fe;
				g_new = ge;
				x_new = xe;
				bool done = false;
				if constexpr (SPEC)
					if (speculate)
				{
					// barrier B: the helpers' trials of this round are in;
					// take the round's trials in order
					__syncthreads();
					for (int w = 0; w < spec.width && !done; w++)
					{
						if (w > 0)
						{
							const S* r = spec.res + w * SPEC_RS;
							++evals;
							f1 = r[0];
							x_new = lane < c ? r[1 + lane] : S(0);
							g_new = lane < n ? r[1 + MAX_CONF + lane] : S(0);
							helper_last = true;
							x_last = x_new;
						}
						if (f1 - f0 < S(0.0001) * alpha * slope)
							done = true;
						else
						{
							alpha *= S(0.5);
							done = ++trial >= 10;
						}
					}
				}
				if (speculate)
				{
					// taken above
				}
				else if (!accurate)
				{
					if (f1 - f0 < S(0.0001) * alpha * slope)
						done = true;
					else
					{
						alpha *= S(0.5);
						done = ++trial >= 10;
					}
				}
				else
				{
					const S ALF = S(1.0e-4), FIRST = S(1.0);
					if (alpha < alamin)
					{
						x_new = x;
						g_new = 0;
						alpha = 0;
						done = true;
					}
					else if (f1 <= f0 + ALF * alpha * slope)
						done = true;
					else
					{
						S tmplam;
						if (alpha == FIRST)
							tmplam = -slope / (S(2.0) * (f1 - f0 - slope));
						else
						{
							const S rhs1 = f1 - f0 - alpha * slope;
							const S rhs2 = f2 - f0 - alpha2 * slope;
							const S a = (rhs1 / (alpha * alpha) - rhs2 / (alpha2 * alpha2)) / (alpha - alpha2);
							const S b = (-alpha2 * rhs1 / (alpha * alpha) + alpha * rhs2 / (alpha2 * alpha2)) /
									(alpha - alpha2);
							if (a == S(0.0))
								tmplam = -slope / (S(2.0) * b);
							else
							{
								const S disc = b * b - S(3.0) * a * slope;
								if (disc < 0)
									tmplam = S(0.5) * alpha;
								else if (b <= 0)
									tmplam = (-b + Real<S>::sqrt(disc)) / (S(3.0) * a);
								else
									tmplam = -slope / (b + Real<S>::sqrt(disc));
							}
							if (tmplam > S(.5) * alpha)
								tmplam = S(.5) * alpha;
						}
						alpha2 = alpha;
						f2 = f1;
						alpha = fmax(tmplam, S(0.1) * alpha);
						// every backtrack at least halves alpha, so a finite
						// search ends within ~1100 trials; the cap only stops
						// a NaN search
						if (++trial >= 2048)
						{
							x_new = x;
							g_new = 0;
							alpha = 0;
							done = true;
						}
						else
							done = false;
					}
				}
				if (!done)
				{
					if constexpr (SPEC)
						if (speculate)
							spec_post(x, p, alpha);
					xe = trial_conf(x, p, alpha);
					continue;
				}
				// the step body after the line search
				if (alpha == 0)
					break;
				const S y = g_new - g;
				const S prevf0 = f0;
				f0 = f1;
				x = x_new;
				if (early_term)
				{
					const S diff = prevf0 - f0;
					if (Real<S>::abs(diff) < S(1e-5))
						break;
				}
				g = g_new;
				const S gradnormsq = dot_n(g, g, n);
				if (!(gradnormsq >= S(1e-4)))
					break;
				if (step == 0)
				{
					const S yy = dot_n(y, y, n);
					if (Real<S>::abs(yy) > Real<S>::eps())
					{
						const S diag = alpha * dot_n(y, p, n) / yy;
						if (lane < n)
							SM.hess[lane * n + lane] = diag;
						lsync();
					}
				}
				bfgs_update(SM.hess, SM.pstage, SM.smhy, p, y, alpha, n, hm);
				if (history)
					vis.add(n, change_value(x), f0, g);
				begin = ++step < maxiters;
				if (!begin)
					break;
			}
			if (begin)
			{
				// the next step's direction and the start of its line search
				p = minus_mat_vec(SM.hess, g, n);
				trial = 0;
				alpha = 1;
				if (!accurate)
					slope = dot_n(p, g, n);
				else
				{
					(void)dot_n(p, p, n); // the unused norm of p
					slope = dot_n(g, p, n);
					if (slope >= 0)
						break; // alpha 0
					stage(SM, x, p, c, n);
					const S xv = conf_value(m, SM);
					const S temp = lane < n ? Real<S>::abs(p) / fmax(Real<S>::abs(xv), S(1.0)) : S(0);
					S test = 0;
					for (int i = 0; i < n; i++)
					{
						const S ti = readlane(temp, i);
						if (ti > test)
							test = ti;
					}
					alamin = Real<S>::eps() / test;
					alpha2 = 0;
					f2 = 0;
				}
				if constexpr (SPEC)
					if (speculate)
						spec_post(x, p, alpha);
				xe = trial_conf(x, p, alpha);
			}
		}
		// the model's coordinates are the last evaluated conf's, as after
		// the sequential search (smina's mutation reads them next)
		if constexpr (SPEC)
			if (helper_last)
				set_conf(m, W, T(x_last));
		if (!(f0 <= f_orig))
		{
			f0 = f_orig;
			x = x_orig;
			g = g_orig;
		}
		return f0;
	}
};

// A helper wave of a speculative chain (rank 1..width-1 of the workgroup):
// until wave 0 reports the chain done, evaluate at each posted round the
// trial x + alpha 2^-rank p on the receptor maps, as Searcher::trial_conf and
// the evaluation at the top of Searcher::bfgs do, and leave its energy, conf
// and gradient for wave 0.  Two workgroup barriers per round (A, B), as wave
// 0 takes them.
template<typename T, typename S, bool SPL>
__device__ inline void spec_helper(const MV<T>& m, const EvalCtx<T>& ec, WaveMem<T>& W, SearchMem<S>& SM,
		const SpecLink<S>& L, int rank, T* pbuf)
{
	EvalCtx<T> ev = ec;
	const int lane = lane_id();
	const int c = m.h.conf_size, n = m.h.change_size;
	for (;;)
	{
		__syncthreads(); // A: a round (or the end) is posted
		if (*L.cmd)
			break;
		S alpha = L.req[MAX_CONF + MAX_CHANGE];
		for (int k = 0; k < rank; k++)
			alpha *= S(0.5);
		for (int d = 0; d < 3; d++)
			ev.v[d] = T(L.req[MAX_CONF + MAX_CHANGE + 1 + d]);
		const S x = lane < c ? L.req[lane] : S(0);
		const S p = lane < n ? L.req[MAX_CONF + lane] : S(0);
		stage(SM, x, p, c, n);
		const S xe = conf_increment(m, SM, alpha);
		lsync();
		T g;
		const S f = S(eval_deriv<T, BK_MAPS, SPL>(m, ev, W, pbuf, T(xe), g));
		S* r = L.res + rank * SPEC_RS;
		if (lane == 0)
			r[0] = f;
		if (lane < c)
			r[1 + lane] = xe;
		if (lane < n)
			r[1 + MAX_CONF + lane] = S(g);
		__syncthreads(); // B: this round's trials are in
	}
}

// ------------------------------------------------------------------ RNG
// MT19937's regeneration of the 624 words in place, lane-parallel.  Out of
// line: it runs once per 624 draws and would otherwise be inlined at every
// draw site.  new[k] uses old k, k+1 and (k+397 mod 624), which is already
// new for k >= 227: three dependent ranges, chunks of 64 in order.
static __device__ __noinline__ void mt_regenerate_wave(uint32_t* x)
{
	const int lane = lane_id();
	for (int base = 0; base < 227; base += WAVE)
	{
		const int k = base + lane;
		uint32_t v = 0;
		if (k < 227)
			v = mt_mix(x[k], x[k + 1], x[k + MT_M]);
		gsync();
		if (k < 227)
			x[k] = v;
		gsync();
	}
	for (int base = 227; base < 454; base += WAVE)
	{
		const int k = base + lane;
		uint32_t v = 0;
		if (k < 454)
			v = mt_mix(x[k], x[k + 1], x[k - 227]);
		gsync();
		if (k < 454)
			x[k] = v;
		gsync();
	}
	for (int base = 454; base < MT_N; base += WAVE)
	{
		const int k = base + lane;
		uint32_t v = 0;
		if (k < MT_N)
			v = mt_mix(x[k], x[(k + 1) % MT_N], x[k - 227]);
		gsync();
		if (k < MT_N)
			x[k] = v;
		gsync();
	}
}

// the chain's MT19937 in global memory; every lane draws the same numbers
struct WaveMT
{
	uint32_t* x; // 624 words
	int i;

	__device__ __forceinline__ uint32_t next()
	{
		if (i >= MT_N)
		{
			mt_regenerate_wave(x);
			i = 0;
		}
		const uint32_t v = x[i];
		++i;
		return mt_temper(v);
	}
};

// random_inside_sphere (src/lib/random.cpp:66-76)
template<typename G>
__device__ inline V3<double> random_inside_sphere(G& g)
{
	for (;;)
	{
		const double r1 = random_fl(-1.0, 1.0, g);
		const double r2 = random_fl(-1.0, 1.0, g);
		const double r3 = random_fl(-1.0, 1.0, g);
		if (r1 * r1 + r2 * r2 + r3 * r3 < 1)
			return v3(r1, r2, r3);
	}
}

// per-chain state in global memory: confs and energies in the search
// precision S, coordinates in the evaluation precision T
template<typename T, typename S>
struct ChainState
{
	uint32_t* mt;     // 624 words
	int* ints;        // [0] mt position, [1] pose count, [2] started, [3..3+MAX_POSES) order
	S* tmp;           // current conf (conf stride)
	S* scal;          // [0] current e, [1] best e
	S* out_e;         // [MAX_POSES]
	S* out_conf;      // [MAX_POSES][conf stride]
	T* out_coords;    // [MAX_POSES][coord stride]
	T* model_coords;  // the model's coordinates (3 * movable) between launches
	T* root_origin;   // [MAX_TREES * 3]
	T* pbuf;          // pair-energy scratch (ordered mode)
	int64_t* evals;
	S* vis_x;         // QuickVina 2's history (null: none)
	S* vis_f;
	uint64_t* vis_bits;
	int* vis_state;   // size, next slot, full
};

struct ChainLayout
{
	int conf_stride;
	int coord_stride;
	int atom_stride;   // 3 * max movable atoms
	int pair_stride;
	int max_poses;
	size_t bytes;      // per chain
	size_t off_mt, off_ints, off_tmp, off_scal, off_oute, off_outc, off_outx, off_mc, off_ro, off_pbuf, off_ev;
	size_t off_hpbuf;  // FP64: the ordered pair loop's scratch of a speculative chain's helper waves
	int vis_n;         // QuickVina 2's history for change sizes up to this (0: none)
	size_t off_vx, off_vf, off_vb, off_vs;
};

template<typename T, typename S>
__device__ inline ChainState<T, S> chain_state(char* base, const ChainLayout& L, int c)
{
	char* p = base + size_t(c) * L.bytes;
	ChainState<T, S> s;
	s.mt = reinterpret_cast<uint32_t*>(p + L.off_mt);
	s.ints = reinterpret_cast<int*>(p + L.off_ints);
	s.tmp = reinterpret_cast<S*>(p + L.off_tmp);
	s.scal = reinterpret_cast<S*>(p + L.off_scal);
	s.out_e = reinterpret_cast<S*>(p + L.off_oute);
	s.out_conf = reinterpret_cast<S*>(p + L.off_outc);
	s.out_coords = reinterpret_cast<T*>(p + L.off_outx);
	s.model_coords = reinterpret_cast<T*>(p + L.off_mc);
	s.root_origin = reinterpret_cast<T*>(p + L.off_ro);
	s.pbuf = reinterpret_cast<T*>(p + L.off_pbuf);
	s.evals = reinterpret_cast<int64_t*>(p + L.off_ev);
	s.vis_x = L.vis_n > 0 ? reinterpret_cast<S*>(p + L.off_vx) : nullptr;
	s.vis_f = reinterpret_cast<S*>(p + L.off_vf);
	s.vis_bits = reinterpret_cast<uint64_t*>(p + L.off_vb);
	s.vis_state = reinterpret_cast<int*>(p + L.off_vs);
	return s;
}

struct McParams
{
	double temperature;
	double amplitude;
	double min_rmsd;
	int num_saved;
	double hunt[3];
	double authentic[3];
	int accurate;
	int early_term;
	int visited;  // QuickVina 2's history check
};

// model::gyration_radius (src/lib/model.cpp:868-883) on the model's current
// coordinates and root origin
template<typename T>
__device__ inline double gyration_radius(const MV<T>& m, const WaveMem<T>& W, int tree)
{
	const int root = m.troot[tree];
	const V3<T> o = ld3(W.fo + 3 * root);
	T acc = 0;
	unsigned counter = 0;
	if constexpr (Real<T>::ordered)
	{
		for (int i = m.tab_[tree]; i < m.tae[tree]; i++)
			if (!is_hydrogen_type(m.type[i]))
			{
				const V3<T> x = ld3(W.coords + 3 * i);
				acc += sqr(x.x - o.x) + sqr(x.y - o.y) + sqr(x.z - o.z);
				++counter;
			}
	}
	else
	{
		T part = 0, cnt = 0;
		for (int i = m.tab_[tree] + lane_id(); i < m.tae[tree]; i += WAVE)
			if (!is_hydrogen_type(m.type[i]))
			{
				const V3<T> x = ld3(W.coords + 3 * i);
				part += sqr(x.x - o.x) + sqr(x.y - o.y) + sqr(x.z - o.z);
				cnt += 1;
			}
		acc = wave_sum(part);
		counter = unsigned(wave_sum(cnt));
	}
	return counter > 0 ? double(Real<T>::sqrt(acc / T(counter))) : 0.0;
}

// mutate_conf (src/lib/mutate.cpp:35-63): one of position, orientation or a
// torsion; returns this lane's conf element
template<typename T, typename S, typename G>
__device__ inline S mutate(const MV<T>& m, const WaveMem<T>& W, S x, double amplitude, G& g)
{
	const int lane = lane_id();
	int count = 0;
	for (int t = 0; t < m.h.ntrees; t++)
		count += m.tlig[t] ? 2 + m.tntors[t] : m.tntors[t];
	if (count == 0)
		return x;
	int which = random_int(0, count - 1, g);
	for (int t = 0; t < m.h.ntrees; t++)
	{
		if (!m.tlig[t])
			continue;
		const int cc = m.tconf[t];
		if (which == 0)
		{
			const V3<double> s = random_inside_sphere(g);
			const V3<double> d = amplitude * s;
			if (lane == cc)
				x += S(d.x);
			if (lane == cc + 1)
				x += S(d.y);
			if (lane == cc + 2)
				x += S(d.z);
			return x;
		}
		--which;
		if (which == 0)
		{
			const double gr = gyration_radius(m, W, t);
			if (gr > DBL_EPSILON)
			{
				const V3<double> s = random_inside_sphere(g);
				const double k = amplitude / gr;
				const V3<S> rot = v3(S(k * s.x), S(k * s.y), S(k * s.z));
				const Q4<S> q = {readlane(x, cc + 3), readlane(x, cc + 4), readlane(x, cc + 5), readlane(x, cc + 6)};
				const Q4<S> r = q_increment(q, rot);
				if (lane == cc + 3)
					x = r.a;
				if (lane == cc + 4)
					x = r.b;
				if (lane == cc + 5)
					x = r.c;
				if (lane == cc + 6)
					x = r.d;
			}
			return x;
		}
		--which;
		if (which < m.tntors[t])
		{
			const double v = random_fl(-3.1415926535897931, 3.1415926535897931, g);
			if (lane == cc + 7 + which)
				x = S(v);
			return x;
		}
		which -= m.tntors[t];
	}
	for (int t = 0; t < m.h.ntrees; t++)
	{
		if (m.tlig[t])
			continue;
		if (which < m.tntors[t])
		{
			const double v = random_fl(-3.1415926535897931, 3.1415926535897931, g);
			if (lane == m.tconf[t] + which)
				x = S(v);
			return x;
		}
		which -= m.tntors[t];
	}
	return x;
}

// metropolis_accept (src/lib/monte_carlo.cpp:37-41)
template<typename G>
__device__ inline bool metropolis_accept(double old_f, double new_f, double temperature, G& g)
{
	if (new_f < old_f)
		return true;
	const double p = pmath::exp((old_f - new_f) / temperature);
	return random_fl(0.0, 1.0, g) < p;
}

// add_to_output_container (src/lib/coords.cpp:34-58) with find_closest's
// rmsd_upper_bound over the heavy movable atoms, then a stable sort
template<typename T, typename S>
__device__ inline void add_to_output(const MV<T>& m, const WaveMem<T>& W, ChainState<T, S>& s, const ChainLayout& L,
		S x, S e, double min_rmsd, int max_size, int& count)
{
	const int lane = lane_id();
	const int H = m.h.nheavy;
	int* order = s.ints + 3;
	T rmsd = 0;
	if (lane < count)
	{
		const int slot = order[lane];
		const T* oc = s.out_coords + size_t(slot) * L.coord_stride;
		T acc = 0;
		for (int h = 0; h < H; h++)
		{
			const V3<T> a = ld3(W.coords + 3 * m.heavy[h]);
			acc += sqr(a.x - oc[3 * h]) + sqr(a.y - oc[3 * h + 1]) + sqr(a.z - oc[3 * h + 2]);
		}
		rmsd = H > 0 ? Real<T>::sqrt(acc / T(H)) : T(0);
	}
	int best_p = count;
	T best = Real<T>::max();
	for (int p = 0; p < count; p++)
	{
		const T r = readlane(rmsd, p);
		if (p == 0 || r < best)
		{
			best_p = p;
			best = r;
		}
	}
	int slot = -1;
	if (best_p < count && best < T(min_rmsd))
	{
		if (e < s.out_e[order[best_p]])
			slot = order[best_p];
	}
	else
	{
		if (count < max_size)
		{
			slot = count;
			if (lane == 0)
				order[count] = count;
			++count;
		}
		else if (count > 0 && e < s.out_e[order[count - 1]])
			slot = order[count - 1];
	}
	if (slot >= 0)
	{
		if (lane < m.h.conf_size)
			s.out_conf[size_t(slot) * L.conf_stride + lane] = x;
		for (int h = lane; h < H; h += WAVE)
		{
			const V3<T> a = ld3(W.coords + 3 * m.heavy[h]);
			st3(s.out_coords + size_t(slot) * L.coord_stride + 3 * h, a);
		}
		if (lane == 0)
			s.out_e[slot] = e;
		gsync();
		if (lane == 0)
		{ // stable insertion sort of the order by energy
			for (int a = 1; a < count; a++)
			{
				const int v = order[a];
				const S ev = s.out_e[v];
				int b = a - 1;
				while (b >= 0 && s.out_e[order[b]] > ev)
				{
					order[b + 1] = order[b];
					--b;
				}
				order[b + 1] = v;
			}
			s.ints[1] = count;
		}
		gsync();
	}
}

// within the box (non_cache::within, src/lib/non_cache.cpp:93-107)
template<typename T>
__device__ inline bool within_box(const MV<T>& m, const WaveMem<T>& W, const ReceptorPtrs<T>& r)
{
	bool ok = true;
	for (int i = lane_id(); i < m.h.natoms; i += WAVE)
	{
		if (is_hydrogen_type(m.type[i]))
			continue;
		const V3<T> x = ld3(W.coords + 3 * i);
		const T c[3] = {x.x, x.y, x.z};
		for (int j = 0; j < 3; j++)
			if (r.enabled[j])
				if (c[j] < T(r.begin[j] - 0.0001) || c[j] > T(r.end[j] + 0.0001))
					ok = false;
	}
	return wave_all(ok);
}

} // namespace sgpu

#endif

/*
 * BFGS, line searches, mutation, Metropolis, the output container and the
 * Monte Carlo chain on one wavefront (src/lib/bfgs.h, src/lib/mutate.cpp,
 * src/lib/monte_carlo.cpp, src/lib/coords.cpp; the reference's device
 * statement of the same semantics is src/gpu/gpu_mc.cuh).
 *
 * Vectors of the change and conf spaces live one element per lane; the
 * packed Hessian (h(i, j) at i + j(j+1)/2, i <= j) lives in LDS.  Scalars
 * are wave-uniform.  With T = double every dot product is a serial sum in
 * index order, as on the CPU.
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

__device__ inline int hidx(int i, int j) { return i < j ? i + j * (j + 1) / 2 : j + i * (i + 1) / 2; }

// minus_mat_vec_product (src/lib/bfgs.h:31-42)
template<typename T>
__device__ inline T minus_mat_vec(const T* h, T in, int n)
{
	const int lane = lane_id();
	T sum = 0;
	for (int j = 0; j < n; j++)
	{
		const T inj = readlane(in, j);
		if (lane < n)
			sum += h[hidx(lane, j)] * inj;
	}
	return -sum;
}

// bfgs_update (src/lib/bfgs.h:52-72)
template<typename T>
__device__ inline void bfgs_update(T* h, T p, T y, T alpha, int n)
{
	const T yp = dot_n(y, p, n);
	if (alpha * yp < Real<T>::eps())
		return;
	const T mhy = minus_mat_vec(h, y, n);
	const T yhy = -dot_n(y, mhy, n);
	const T r = 1 / (alpha * yp);
	const int lane = lane_id();
	for (int j = 0; j < n; j++)
	{
		const T pj = readlane(p, j);
		const T mhyj = readlane(mhy, j);
		if (lane <= j)
		{
			T& hij = h[lane + j * (j + 1) / 2];
			hij += alpha * r * (mhy * pj + mhyj * p) + +alpha * alpha * (r * r * yhy + r) * p * pj;
		}
	}
	wsync();
}

// conf::increment (src/lib/conf.h:316-321) of the staged conf by
// alpha * staged change; returns this lane's new conf element
template<typename T>
__device__ inline T conf_increment(const MV<T>& m, const WaveMem<T>& W, T alpha)
{
	const int k = lane_id();
	if (k >= m.h.conf_size)
		return T(0);
	const int kind = m.ckind[k];
	const int gi = m.cchg[k];
	const T x = W.xstage[k];
	if (kind == CK_POSITION)
	{
		const T d = alpha * W.pstage[gi];
		return x + d;
	}
	if (kind == CK_TORSION)
	{
		T t = x + normalized_angle(alpha * W.pstage[gi]);
		normalize_angle(t);
		return t;
	}
	const int comp = kind - CK_QUAT0;
	const int q0 = k - comp;
	const V3<T> rot = v3(alpha * W.pstage[gi], alpha * W.pstage[gi + 1], alpha * W.pstage[gi + 2]);
	const Q4<T> q = {W.xstage[q0], W.xstage[q0 + 1], W.xstage[q0 + 2], W.xstage[q0 + 3]};
	const Q4<T> r = q_increment(q, rot);
	return comp == 0 ? r.a : comp == 1 ? r.b : comp == 2 ? r.c : r.d;
}

// conf::operator()(index) (src/lib/conf.h:376-397) for the staged conf
template<typename T>
__device__ inline T conf_value(const MV<T>& m, const WaveMem<T>& W)
{
	const int i = lane_id();
	if (i >= m.h.change_size)
		return T(0);
	const int kind = m.gkind[i];
	const int ci = m.gconf[i];
	if (kind == GK_POSITION || kind == GK_TORSION)
		return W.xstage[ci];
	const Q4<T> q = {W.xstage[ci], W.xstage[ci + 1], W.xstage[ci + 2], W.xstage[ci + 3]};
	const V3<T> a = q_to_angle(q);
	const int comp = kind - GK_ORIENT0;
	return comp == 0 ? a.x : comp == 1 ? a.y : a.z;
}

template<typename T>
__device__ inline void stage(WaveMem<T>& W, T x, T p, int c, int n)
{
	const int lane = lane_id();
	if (lane < c)
		W.xstage[lane] = x;
	if (lane < n)
		W.pstage[lane] = p;
	wsync();
}

template<typename T>
struct Searcher
{
	const MV<T>& m;
	const EvalCtx<T>& ec;
	WaveMem<T>& W;
	T* pbuf;
	int64_t evals;

	__device__ T f(T x, T& g)
	{
		++evals;
		return eval_deriv(m, ec, W, pbuf, x, g);
	}

	// fast_line_search (src/lib/bfgs.h:76-96)
	__device__ T fast_line_search(T x, T g, T f0, T p, T& x_new, T& g_new, T& f1)
	{
		const int n = m.h.change_size, c = m.h.conf_size;
		const T c0 = T(0.0001);
		const T multiplier = T(0.5);
		T alpha = 1;
		const T pg = dot_n(p, g, n);
		for (int trial = 0; trial < 10; trial++)
		{
			stage(W, x, p, c, n);
			x_new = conf_increment(m, W, alpha);
			wsync();
			f1 = f(x_new, g_new);
			if (f1 - f0 < c0 * alpha * pg)
				break;
			alpha *= multiplier;
		}
		return alpha;
	}

	// accurate_line_search (src/lib/bfgs.h:101-190)
	__device__ T accurate_line_search(T x, T g, T f0, T p, T& x_new, T& g_new, T& f1)
	{
		const int n = m.h.change_size, c = m.h.conf_size;
		const T ALF = T(1.0e-4);
		const T FIRST = T(1.0);
		T a, alpha, alpha2 = 0, alamin, b, disc, f2 = 0;
		T rhs1, rhs2, slope, test, tmplam = 0;
		T sum = dot_n(p, p, n);
		sum = Real<T>::sqrt(sum);
		(void)sum;
		slope = dot_n(g, p, n);
		if (slope >= 0)
		{
			x_new = x;
			g_new = 0;
			return 0;
		}
		stage(W, x, p, c, n);
		const T xv = conf_value(m, W);
		T temp = lane_id() < n ? Real<T>::abs(p) / fmax(Real<T>::abs(xv), T(1.0)) : T(0);
		test = 0;
		for (int i = 0; i < n; i++)
		{
			const T ti = readlane(temp, i);
			if (ti > test)
				test = ti;
		}
		alamin = Real<T>::eps() / test;
		alpha = FIRST;
		// every backtrack at least halves alpha, so a finite search ends
		// within ~1100 trials; the cap only stops a NaN search
		for (int trial = 0; trial < 2048; trial++)
		{
			stage(W, x, p, c, n);
			x_new = conf_increment(m, W, alpha);
			wsync();
			f1 = f(x_new, g_new);
			if (alpha < alamin)
			{
				x_new = x;
				g_new = 0;
				return 0;
			}
			else if (f1 <= f0 + ALF * alpha * slope)
				return alpha;
			else
			{
				if (alpha == FIRST)
					tmplam = -slope / (T(2.0) * (f1 - f0 - slope));
				else
				{
					rhs1 = f1 - f0 - alpha * slope;
					rhs2 = f2 - f0 - alpha2 * slope;
					a = (rhs1 / (alpha * alpha) - rhs2 / (alpha2 * alpha2)) / (alpha - alpha2);
					b = (-alpha2 * rhs1 / (alpha * alpha) + alpha * rhs2 / (alpha2 * alpha2)) / (alpha - alpha2);
					if (a == T(0.0))
						tmplam = -slope / (T(2.0) * b);
					else
					{
						disc = b * b - T(3.0) * a * slope;
						if (disc < 0)
							tmplam = T(0.5) * alpha;
						else if (b <= 0)
							tmplam = (-b + Real<T>::sqrt(disc)) / (T(3.0) * a);
						else
							tmplam = -slope / (b + Real<T>::sqrt(disc));
					}
					if (tmplam > T(.5) * alpha)
						tmplam = T(.5) * alpha;
				}
			}
			alpha2 = alpha;
			f2 = f1;
			alpha = fmax(tmplam, T(0.1) * alpha);
		}
		x_new = x;
		g_new = 0;
		return 0;
	}

	// bfgs (src/lib/bfgs.h:201-282); x, g: this lane's elements
	__device__ T bfgs(T& x, T& g, int maxiters, bool accurate, bool early_term)
	{
		const int n = m.h.change_size;
		const int lane = lane_id();
		const int hn = n * (n + 1) / 2;
		for (int k = lane; k < hn; k += WAVE)
			W.hess[k] = 0;
		wsync();
		if (lane < n)
			W.hess[lane + lane * (lane + 1) / 2] = 1;
		wsync();
		T f0 = f(x, g);
		const T f_orig = f0;
		const T g_orig = g, x_orig = x;
		T x_new = x, g_new = g, p = g;
		for (int step = 0; step < maxiters; step++)
		{
			p = minus_mat_vec(W.hess, g, n);
			T f1 = 0;
			T alpha;
			if (accurate)
				alpha = accurate_line_search(x, g, f0, p, x_new, g_new, f1);
			else
				alpha = fast_line_search(x, g, f0, p, x_new, g_new, f1);
			if (alpha == 0)
				break;
			const T y = g_new - g;
			const T prevf0 = f0;
			f0 = f1;
			x = x_new;
			if (early_term)
			{
				const T diff = prevf0 - f0;
				if (Real<T>::abs(diff) < T(1e-5))
					break;
			}
			g = g_new;
			const T gradnormsq = dot_n(g, g, n);
			if (!(gradnormsq >= T(1e-4)))
				break;
			if (step == 0)
			{
				const T yy = dot_n(y, y, n);
				if (Real<T>::abs(yy) > Real<T>::eps())
				{
					const T diag = alpha * dot_n(y, p, n) / yy;
					if (lane < n)
						W.hess[lane + lane * (lane + 1) / 2] = diag;
					wsync();
				}
			}
			bfgs_update(W.hess, p, y, alpha, n);
		}
		if (!(f0 <= f_orig))
		{
			f0 = f_orig;
			x = x_orig;
			g = g_orig;
		}
		return f0;
	}
};

// ------------------------------------------------------------------ RNG
// the chain's MT19937 in global memory; every lane draws the same numbers
struct WaveMT
{
	uint32_t* x; // 624 words
	int i;

	__device__ void regenerate()
	{
		const int lane = lane_id();
		// new[k] uses old k, k+1 and (k+397 mod 624), which is already new for
		// k >= 227: three dependent ranges, chunks of 64 in order
		for (int base = 0; base < 227; base += WAVE)
		{
			const int k = base + lane;
			uint32_t v = 0;
			if (k < 227)
				v = mt_mix(x[k], x[k + 1], x[k + MT_M]);
			wsync();
			if (k < 227)
				x[k] = v;
			wsync();
		}
		for (int base = 227; base < 454; base += WAVE)
		{
			const int k = base + lane;
			uint32_t v = 0;
			if (k < 454)
				v = mt_mix(x[k], x[k + 1], x[k - 227]);
			wsync();
			if (k < 454)
				x[k] = v;
			wsync();
		}
		for (int base = 454; base < MT_N; base += WAVE)
		{
			const int k = base + lane;
			uint32_t v = 0;
			if (k < MT_N)
				v = mt_mix(x[k], x[(k + 1) % MT_N], x[k - 227]);
			wsync();
			if (k < MT_N)
				x[k] = v;
			wsync();
		}
		i = 0;
	}
	__device__ uint32_t next()
	{
		if (i >= MT_N)
			regenerate();
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

// per-chain state in global memory
template<typename T>
struct ChainState
{
	uint32_t* mt;     // 624 words
	int* ints;        // [0] mt position, [1] pose count, [2] started, [3..3+MAX_POSES) order
	T* tmp;           // current conf (conf stride)
	T* scal;          // [0] current e, [1] best e
	T* out_e;         // [MAX_POSES]
	T* out_conf;      // [MAX_POSES][conf stride]
	T* out_coords;    // [MAX_POSES][coord stride]
	T* model_coords;  // the model's coordinates (3 * movable) between launches
	T* root_origin;   // [MAX_TREES * 3]
	T* pbuf;          // pair-energy scratch (ordered mode)
	int64_t* evals;
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
};

template<typename T>
__device__ inline ChainState<T> chain_state(char* base, const ChainLayout& L, int c)
{
	char* p = base + size_t(c) * L.bytes;
	ChainState<T> s;
	s.mt = reinterpret_cast<uint32_t*>(p + L.off_mt);
	s.ints = reinterpret_cast<int*>(p + L.off_ints);
	s.tmp = reinterpret_cast<T*>(p + L.off_tmp);
	s.scal = reinterpret_cast<T*>(p + L.off_scal);
	s.out_e = reinterpret_cast<T*>(p + L.off_oute);
	s.out_conf = reinterpret_cast<T*>(p + L.off_outc);
	s.out_coords = reinterpret_cast<T*>(p + L.off_outx);
	s.model_coords = reinterpret_cast<T*>(p + L.off_mc);
	s.root_origin = reinterpret_cast<T*>(p + L.off_ro);
	s.pbuf = reinterpret_cast<T*>(p + L.off_pbuf);
	s.evals = reinterpret_cast<int64_t*>(p + L.off_ev);
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
	for (int i = m.tab_[tree]; i < m.tae[tree]; i++)
		if (!is_hydrogen_type(m.type[i]))
		{
			const V3<T> x = ld3(W.coords + 3 * i);
			acc += sqr(x.x - o.x) + sqr(x.y - o.y) + sqr(x.z - o.z);
			++counter;
		}
	return counter > 0 ? double(Real<T>::sqrt(acc / T(counter))) : 0.0;
}

// mutate_conf (src/lib/mutate.cpp:35-63): one of position, orientation or a
// torsion; returns this lane's conf element
template<typename T, typename G>
__device__ inline T mutate(const MV<T>& m, const WaveMem<T>& W, T x, double amplitude, G& g)
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
				x += T(d.x);
			if (lane == cc + 1)
				x += T(d.y);
			if (lane == cc + 2)
				x += T(d.z);
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
				const V3<T> rot = v3(T(k * s.x), T(k * s.y), T(k * s.z));
				const Q4<T> q = {readlane(x, cc + 3), readlane(x, cc + 4), readlane(x, cc + 5), readlane(x, cc + 6)};
				const Q4<T> r = q_increment(q, rot);
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
				x = T(v);
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
				x = T(v);
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
template<typename T>
__device__ inline void add_to_output(const MV<T>& m, const WaveMem<T>& W, ChainState<T>& s, const ChainLayout& L,
		T x, T e, double min_rmsd, int max_size, int& count)
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
		wsync();
		if (lane == 0)
		{ // stable insertion sort of the order by energy
			for (int a = 1; a < count; a++)
			{
				const int v = order[a];
				const T ev = s.out_e[v];
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
		wsync();
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

// TEST INFRASTRUCTURE ONLY: CPU FP64 restatement of smina's docking search.
// See oracle.h for what it is, what pins it and who may use it.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <array>
#include <map>
#include <sstream>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <atomic>
#include <vector>

#include "../gpu-smina_amd/csrc/atom_types.h"
#include "../gpu-smina_amd/csrc/pmath.h"

namespace orc
{

using sgpu::atom_type_info;
using sgpu::is_hydrogen_type;

typedef double fl;
const fl pi = fl(3.1415926535897931);
const fl epsilon_fl = std::numeric_limits<fl>::epsilon();
const fl max_fl = std::numeric_limits<fl>::max();
const int NUM_TYPES = sgpu::NUM_ATOM_TYPES;

inline fl sqr(fl x) { return x * x; }
inline bool not_max(fl x) { return x < 0.1 * max_fl; }

// ---------------------------------------------------------------- vectors
// vec, mat (column-major) with smina's operand order (src/lib/common.h)
struct vec
{
	fl d[3];
	vec() : d{0, 0, 0} {}
	vec(fl x, fl y, fl z) : d{x, y, z} {}
	fl& operator[](int i) { return d[i]; }
	fl operator[](int i) const { return d[i]; }
	fl norm_sqr() const { return sqr(d[0]) + sqr(d[1]) + sqr(d[2]); }
	fl norm() const { return std::sqrt(norm_sqr()); }
	fl operator*(const vec& v) const { return d[0] * v[0] + d[1] * v[1] + d[2] * v[2]; }
	vec& operator+=(const vec& v) { d[0] += v[0]; d[1] += v[1]; d[2] += v[2]; return *this; }
	vec& operator-=(const vec& v) { d[0] -= v[0]; d[1] -= v[1]; d[2] -= v[2]; return *this; }
	vec& operator*=(fl s) { d[0] *= s; d[1] *= s; d[2] *= s; return *this; }
	vec operator+(const vec& v) const { return vec(d[0] + v[0], d[1] + v[1], d[2] + v[2]); }
	vec operator-(const vec& v) const { return vec(d[0] - v[0], d[1] - v[1], d[2] - v[2]); }
};
inline vec operator*(fl s, const vec& v) { return vec(s * v[0], s * v[1], s * v[2]); }
inline vec cross_product(const vec& a, const vec& b)
{
	return vec(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline fl vec_distance_sqr(const vec& a, const vec& b)
{
	return sqr(a[0] - b[0]) + sqr(a[1] - b[1]) + sqr(a[2] - b[2]);
}
struct mat
{
	fl d[9];
	vec operator*(const vec& v) const
	{
		return vec(d[0] * v[0] + d[3] * v[1] + d[6] * v[2], d[1] * v[0] + d[4] * v[1] + d[7] * v[2],
				d[2] * v[0] + d[5] * v[1] + d[8] * v[2]);
	}
};

// common.h:245-270
inline void normalize_angle(fl& x)
{
	if (x > 3 * pi)
	{
		fl n = (x - pi) / (2 * pi);
		x -= 2 * pi * std::ceil(n);
		normalize_angle(x);
	}
	else if (x < -3 * pi)
	{
		fl n = (-x - pi) / (2 * pi);
		x += 2 * pi * std::ceil(n);
		normalize_angle(x);
	}
	else if (x > pi)
		x -= 2 * pi;
	else if (x < -pi)
		x += 2 * pi;
}
inline fl normalized_angle(fl x)
{
	normalize_angle(x);
	return x;
}

// curl.h:28-43 (soft curl)
inline void curl(fl& e, vec& deriv, fl v)
{
	if (e > 0 && not_max(v))
	{
		fl tmp = (v < epsilon_fl) ? 0 : (v / (v + e));
		e *= tmp;
		deriv *= sqr(tmp);
	}
}
inline void curl(fl& e, fl v)
{
	if (e > 0 && not_max(v))
	{
		fl tmp = (v < epsilon_fl) ? 0 : (v / (v + e));
		e *= tmp;
	}
}

// ---------------------------------------------------------------- quaternions
// boost::math::quaternion<double> as smina uses it (src/lib/quaternion.cpp)
struct qt
{
	fl a, b, c, d;
};
const qt qt_identity = {1, 0, 0, 0};
inline qt operator*(const qt& l, const qt& r) // Hamilton product, Boost's operand order
{
	return {+l.a * r.a - l.b * r.b - l.c * r.c - l.d * r.d, +l.a * r.b + l.b * r.a + l.c * r.d - l.d * r.c,
			+l.a * r.c - l.b * r.d + l.c * r.a + l.d * r.b, +l.a * r.d + l.b * r.c - l.c * r.b + l.d * r.a};
}
inline fl quaternion_norm_sqr(const qt& q) { return sqr(q.a) + sqr(q.b) + sqr(q.c) + sqr(q.d); }
// quaternion.h:69-80
inline void quaternion_normalize_approx(qt& q)
{
	const fl s = quaternion_norm_sqr(q);
	if (std::abs(s - 1) < 1e-6)
		;
	else
	{
		const fl a = std::sqrt(s);
		const fl inv = 1 / a;
		q.a *= inv;
		q.b *= inv;
		q.c *= inv;
		q.d *= inv;
	}
}
// quaternion.cpp:42-50
inline qt angle_to_quaternion(const vec& axis, fl angle)
{
	normalize_angle(angle);
	fl c = pmath::cos(angle / 2);
	fl s = pmath::sin(angle / 2);
	return {c, s * axis[0], s * axis[1], s * axis[2]};
}
// quaternion.cpp:52-62
inline qt angle_to_quaternion(const vec& rotation)
{
	fl angle = rotation.norm();
	if (angle > epsilon_fl)
	{
		vec axis = (1 / angle) * rotation;
		return angle_to_quaternion(axis, angle);
	}
	return qt_identity;
}
// quaternion.cpp:63-79
inline vec quaternion_to_angle(const qt& q)
{
	const fl c = q.a;
	if (c > -1 && c < 1)
	{
		fl angle = 2 * pmath::acos(c);
		if (angle > pi)
			angle -= 2 * pi;
		vec axis(q.b, q.c, q.d);
		fl s = pmath::sin(angle / 2);
		if (std::abs(s) < epsilon_fl)
			return vec(0, 0, 0);
		axis *= (angle / s);
		return axis;
	}
	return vec(0, 0, 0);
}
// quaternion.cpp:81-107
inline mat quaternion_to_r3(const qt& q)
{
	const fl a = q.a, b = q.b, c = q.c, d = q.d;
	const fl aa = a * a, ab = a * b, ac = a * c, ad = a * d;
	const fl bb = b * b, bc = b * c, bd = b * d;
	const fl cc = c * c, cd = c * d, dd = d * d;
	mat m;
	m.d[0] = (aa + bb - cc - dd);
	m.d[3] = 2 * (-ad + bc);
	m.d[6] = 2 * (ac + bd);
	m.d[1] = 2 * (ad + bc);
	m.d[4] = (aa - bb + cc - dd);
	m.d[7] = 2 * (-ab + cd);
	m.d[2] = 2 * (-ac + bd);
	m.d[5] = 2 * (ab + cd);
	m.d[8] = (aa - bb - cc + dd);
	return m;
}
// quaternion.cpp:123-127
inline void quaternion_increment(qt& q, const vec& rotation)
{
	q = angle_to_quaternion(rotation) * q;
	quaternion_normalize_approx(q);
}

// ---------------------------------------------------------------- random
// boost::mt19937 is std::mt19937's algorithm; the distributions follow
// Boost's uniform_real / uniform_int (src/lib/random.cpp:28-50)
struct rng
{
	std::mt19937 g;
	explicit rng(uint32_t seed) : g(seed) {}
	uint32_t next() { return uint32_t(g()); }
};
fl random_fl(fl a, fl b, rng& gen)
{
	for (;;)
	{
		fl numerator = fl(gen.next());
		fl divisor = fl(4294967295.0) + 1;
		fl result = numerator / divisor * (b - a) + a;
		if (result < b)
			return result;
	}
}
int random_int(int a, int b, rng& gen)
{
	uint32_t range = uint32_t(b) - uint32_t(a);
	if (range == 0)
		return a;
	uint32_t brange = 0xffffffffu;
	uint32_t bucket = brange / (range + 1);
	if (brange % (range + 1) == range)
		++bucket;
	for (;;)
	{
		uint32_t r = gen.next() / bucket;
		if (r <= range)
			return int(r + uint32_t(a));
	}
}
// The normal draws of random_orientation (src/lib/quaternion.cpp:129-141).
// Parity unpinned: Boost's normal_distribution is not available; this is
// Marsaglia's polar method, one fresh pair per call (smina builds a new
// distribution object per draw, so nothing is cached between calls).
fl random_normal(fl mean, fl sigma, rng& gen)
{
	for (;;)
	{
		fl u = random_fl(-1, 1, gen);
		fl v = random_fl(-1, 1, gen);
		fl s = u * u + v * v;
		if (s < 1 && s > 0)
			return mean + sigma * (u * std::sqrt(-2 * std::log(s) / s));
	}
}
// random.cpp:66-76
vec random_inside_sphere(rng& gen)
{
	for (;;)
	{
		fl r1 = random_fl(-1, 1, gen);
		fl r2 = random_fl(-1, 1, gen);
		fl r3 = random_fl(-1, 1, gen);
		vec tmp(r1, r2, r3);
		if (tmp.norm_sqr() < 1)
			return tmp;
	}
}
vec random_in_box(const vec& c1, const vec& c2, rng& gen)
{
	vec tmp;
	for (int i = 0; i < 3; i++)
		tmp[i] = random_fl(c1[i], c2[i], gen);
	return tmp;
}
// quaternion.cpp:129-141 with boost::math::abs (scaled 2-norm)
qt random_orientation(rng& gen)
{
	for (;;)
	{
		fl x0 = random_normal(0, 1, gen);
		fl x1 = random_normal(0, 1, gen);
		fl x2 = random_normal(0, 1, gen);
		fl x3 = random_normal(0, 1, gen);
		fl m = std::max(std::max(std::abs(x0), std::abs(x1)), std::max(std::abs(x2), std::abs(x3)));
		fl nrm = 0;
		if (m != 0)
		{
			fl mixam = 1 / m;
			fl t0 = x0 * mixam, t1 = x1 * mixam, t2 = x2 * mixam, t3 = x3 * mixam;
			nrm = m * std::sqrt(t0 * t0 + t1 * t1 + t2 * t2 + t3 * t3);
		}
		if (nrm > epsilon_fl)
			return {x0 / nrm, x1 / nrm, x2 / nrm, x3 / nrm};
	}
}

// ---------------------------------------------------------------- scoring
// terms of src/lib/everything.h:156-541, weighted sum of
// weighted_terms::eval_fast (src/lib/weighted_terms.cpp:52-62)
// atom properties from the scoring function's parameter set: smina's
// default table, or vinardo's, which builtin_scoring installs in its place
// (src/lib/builtinscoring.cpp:83)
inline fl optimal_distance(int t1, int t2, int ap)
{
	return atom_type_info(t1, ap).xs_radius + atom_type_info(t2, ap).xs_radius;
}
inline fl slope_step(fl x_bad, fl x_good, fl x)
{
	if (x_bad < x_good)
	{
		if (x <= x_bad)
			return 0;
		if (x >= x_good)
			return 1;
	}
	else
	{
		if (x >= x_bad)
			return 0;
		if (x <= x_good)
			return 1;
	}
	return (x - x_bad) / (x_good - x_bad);
}
inline bool xs_h_bond_possible(int t1, int t2, int ap)
{
	return (atom_type_info(t1, ap).donor && atom_type_info(t2, ap).acceptor) ||
			(atom_type_info(t2, ap).donor && atom_type_info(t1, ap).acceptor);
}
inline bool xs_anti_h_bond(int t1, int t2, int ap)
{
	const auto a = atom_type_info(t1, ap), b = atom_type_info(t2, ap);
	if (a.donor && !a.acceptor)
		return b.donor && !b.acceptor;
	if (!a.donor && a.acceptor)
		return !b.donor && b.acceptor;
	return false;
}
inline fl int_pow(fl x, int n)
{
	fl r = 1;
	for (int i = 0; i < n; i++)
		r *= x;
	return r;
}
fl term_eval(const sgpu_term& t, int t1, int t2, fl r, int ap)
{
	const auto a = atom_type_info(t1, ap), b = atom_type_info(t2, ap);
	switch (t.kind)
	{
	case SGPU_TERM_GAUSS: // gauss(o, w): exp(-((r - (d0 + o)) / w)^2) with std::exp
	{
		fl x = r - (optimal_distance(t1, t2, ap) + t.p[0]);
		return std::exp(-sqr(x / t.p[1]));
	}
	case SGPU_TERM_REPULSION:
	{
		fl d = r - (optimal_distance(t1, t2, ap) + t.p[0]);
		if (d > 0)
			return 0;
		return d * d;
	}
	case SGPU_TERM_HYDROPHOBIC:
		if (a.hydrophobe && b.hydrophobe)
			return slope_step(t.p[1], t.p[0], r - optimal_distance(t1, t2, ap));
		return 0;
	case SGPU_TERM_NON_HYDROPHOBIC:
		if (!a.hydrophobe && !b.hydrophobe)
			return slope_step(t.p[1], t.p[0], r - optimal_distance(t1, t2, ap));
		return 0;
	case SGPU_TERM_NON_DIR_H_BOND:
		if (xs_h_bond_possible(t1, t2, ap))
			return slope_step(t.p[1], t.p[0], r - optimal_distance(t1, t2, ap));
		return 0;
	case SGPU_TERM_NON_DIR_ANTI_H_BOND_QUADRATIC:
	case SGPU_TERM_DONOR_DONOR_QUADRATIC:
	case SGPU_TERM_ACCEPTOR_ACCEPTOR_QUADRATIC:
	{
		bool on = t.kind == SGPU_TERM_NON_DIR_ANTI_H_BOND_QUADRATIC ? xs_anti_h_bond(t1, t2, ap)
				: t.kind == SGPU_TERM_DONOR_DONOR_QUADRATIC ? (a.donor && b.donor)
				: (a.acceptor && b.acceptor);
		if (!on)
			return 0;
		fl d = r - (optimal_distance(t1, t2, ap) + t.p[0]);
		if (d > 0)
			return 0;
		return d * d;
	}
	case SGPU_TERM_VDW:
	{
		int i = int(t.p[0]), j = int(t.p[1]);
		fl smoothing = t.p[2], cap = t.p[3];
		fl d0 = optimal_distance(t1, t2, ap);
		fl depth = 1;
		fl c_i = int_pow(d0, i) * depth * j / (fl(i) - fl(j));
		fl c_j = int_pow(d0, j) * depth * i / (fl(j) - fl(i));
		if (r > d0 + smoothing)
			r -= smoothing;
		else if (r < d0 - smoothing)
			r += smoothing;
		else
			r = d0;
		fl r_i = int_pow(r, i), r_j = int_pow(r, j);
		if (r_i > epsilon_fl && r_j > epsilon_fl)
			return std::min(cap, c_i / r_i + c_j / r_j);
		return cap;
	}
	case SGPU_TERM_NON_DIR_H_BOND_LJ:
	{
		if (!xs_h_bond_possible(t1, t2, ap))
			return 0;
		fl cap = t.p[1];
		fl d0 = optimal_distance(t1, t2, ap) + t.p[0];
		fl depth = 5;
		fl c_i = int_pow(d0, 10) * depth * 12 / (fl(10) - fl(12));
		fl c_j = int_pow(d0, 12) * depth * 10 / (fl(12) - fl(10));
		fl r_i = int_pow(r, 10), r_j = int_pow(r, 12);
		if (r_i > epsilon_fl && r_j > epsilon_fl)
			return std::min(cap, c_i / r_i + c_j / r_j);
		return cap;
	}
	}
	throw std::runtime_error("unknown term");
}

// smooth_div (src/lib/everything.cpp)
inline fl smooth_div(fl x, fl y)
{
	if (std::abs(x) < epsilon_fl)
		return 0;
	if (std::abs(y) < epsilon_fl)
		return (x * y > 0) ? max_fl : -max_fl;
	return x / y;
}

inline bool charge_dependent(const sgpu_term& t)
{
	return t.kind == SGPU_TERM_ELECTROSTATIC || t.kind == SGPU_TERM_AD4_SOLVATION;
}

// charge_dependent::eval_components (src/lib/everything.h:64-156): the four
// result components (type only, |q_a|, |q_b|, q_a q_b dependent)
void term_components(const sgpu_term& t, int t1, int t2, fl r, fl* out)
{
	out[0] = out[1] = out[2] = out[3] = 0;
	if (t.kind == SGPU_TERM_ELECTROSTATIC)
	{
		const fl tmp = int_pow(r, int(t.p[0]));
		const fl cap = t.p[1];
		if (tmp < epsilon_fl)
			out[3] = cap;
		else
			out[3] = std::min(cap, 1 / tmp);
		return;
	}
	// ad4_solvation
	const fl solv1 = sgpu::ad_solvation(t1), solv2 = sgpu::ad_solvation(t2);
	const fl volume1 = sgpu::ad_volume(t1), volume2 = sgpu::ad_volume(t2);
	const fl mysolv = t.p[1];
	const fl distfactor = std::exp(-sqr(r / (2 * t.p[0])));
	out[0] += solv1 * volume2 * distfactor;
	out[1] += mysolv * volume2 * distfactor;
	out[0] += solv2 * volume1 * distfactor;
	out[2] += mysolv * volume1 * distfactor;
}

struct Scoring
{
	std::vector<sgpu_term> terms;
	std::vector<sgpu_ci_term> ci;
	int atom_params = 0;
	fl cutoff = 8;
	bool charged() const
	{
		for (const auto& t : terms)
			if (charge_dependent(t))
				return true;
		return false;
	}
	// weighted_terms::eval_fast (src/lib/weighted_terms.cpp:52-62): the
	// charge-independent terms into the first component, then each
	// charge-dependent term's components times its weight
	void eval_fast4(int t1, int t2, fl r, fl* acc) const
	{
		acc[0] = acc[1] = acc[2] = acc[3] = 0;
		for (const auto& t : terms)
			if (!charge_dependent(t))
				acc[0] += t.weight * term_eval(t, t1, t2, r, atom_params);
		for (const auto& t : terms)
			if (charge_dependent(t))
			{
				fl c[4];
				term_components(t, t1, t2, r, c);
				for (int k = 0; k < 4; k++)
					acc[k] += c[k] * t.weight;
			}
	}
	fl eval_fast(int t1, int t2, fl r) const
	{
		fl acc[4];
		eval_fast4(t1, t2, r, acc);
		return acc[0];
	}
	// weighted_terms::conf_independent (src/lib/weighted_terms.cpp:74-80):
	// each conf-independent term's eval in order (src/lib/everything.h:730-856,
	// :980-1002)
	fl conf_independent(fl num_tors, fl x) const
	{
		for (const sgpu_ci_term& c : ci)
		{
			const fl w = c.weight;
			if (c.kind == SGPU_CI_NUM_TORS_DIV)
				x = smooth_div(x, 1 + 0.1 * (w + 1) * num_tors / 5.0);
			else if (c.kind == SGPU_CI_NUM_TORS_ADD)
				x = x + w * num_tors;
			else if (c.kind == SGPU_CI_NUM_TORS_SQR)
				x = x + 0.1 * w * sqr(num_tors) / 5;
			else if (c.kind == SGPU_CI_NUM_TORS_SQRT)
				x = x + 0.1 * w * std::sqrt(num_tors) / std::sqrt(5.0);
			else if (c.kind == SGPU_CI_NUM_TORS_DIV_SIMPLE)
				x = smooth_div(x, 1 + w * num_tors);
			else if (c.kind == SGPU_CI_CONSTANT)
				x = x + w;
			else
				throw std::runtime_error("unknown conf-independent term");
		}
		return x;
	}
};

// precalculate_linear (src/lib/precalculate.h:92-308), one component
struct Tables
{
	int n = 0;
	fl factor = 32;
	fl cutoff_sqr = 64;
	std::vector<fl> rs;
	std::vector<fl> e, dor, fast; // [tri][n], the first (charge-independent) component
	// with charge-dependent terms: 4 components, the other three here
	// ([tri][n][3]: |q_a|, |q_b|, q_a q_b dependent)
	int ncomp = 1;
	std::vector<fl> eq, dorq, fastq;
	void build(const Scoring& sf, fl factor_)
	{
		factor = factor_;
		cutoff_sqr = sqr(sf.cutoff);
		n = int(size_t(factor * cutoff_sqr) + 3);
		rs.assign(n + 2, 0);
		for (int i = 0; i < n + 2; i++)
			rs[i] = std::sqrt(i / factor);
		ncomp = sf.charged() ? 4 : 1;
		const int ntri = NUM_TYPES * (NUM_TYPES + 1) / 2;
		e.assign(size_t(ntri) * n, 0);
		dor.assign(size_t(ntri) * n, 0);
		fast.assign(size_t(ntri) * n, 0);
		eq.assign(ncomp == 4 ? size_t(ntri) * n * 3 : 0, 0);
		dorq.assign(eq.size(), 0);
		fastq.assign(eq.size(), 0);
		std::vector<fl> comp(4 * size_t(n));
		for (int t1 = 0; t1 < NUM_TYPES; t1++)
			for (int t2 = t1; t2 < NUM_TYPES; t2++)
			{
				const size_t base = size_t(sgpu::tri_index(t1, t2)) * n;
				for (int i = 0; i < n; i++)
				{
					sf.eval_fast4(t1, t2, rs[i], &comp[4 * size_t(i)]);
					e[base + i] = comp[4 * size_t(i)];
				}
				// init_from_smooth_fst, component by component
				for (int c = 0; c < ncomp; c++)
				{
					auto val = [&](int i) { return comp[4 * size_t(i) + c]; };
					for (int i = 0; i < n; i++)
					{
						fl d = 0;
						if (!(i == 0 || i == n - 1))
						{
							fl delta = rs[i + 1] - rs[i - 1];
							fl r = rs[i];
							d = (val(i + 1) - val(i - 1)) / (delta * r);
						}
						fl f1 = val(i);
						fl f2 = (i + 1 >= n) ? 0 : val(i + 1);
						const fl fa = (f2 + f1) / 2;
						if (c == 0)
						{
							dor[base + i] = d;
							fast[base + i] = fa;
						}
						else
						{
							eq[(base + i) * 3 + c - 1] = val(i);
							dorq[(base + i) * 3 + c - 1] = d;
							fastq[(base + i) * 3 + c - 1] = fa;
						}
					}
				}
			}
	}
	// precalculate_splines (src/lib/precalculate.h:313-455) with Spline
	// (src/lib/splines.h:37-121), one component.  Per type pair: the terms
	// at n = factor x cutoff distances i x cutoff / n plus (cutoff, 0), and
	// the cubic spline through them with zero first derivative at both
	// ends.  Spline::initialize inverts the dense (n+1) x (n+1) matrix
	// (Eigen); the system is tridiagonal and is solved here by elimination
	// (Thomas algorithm), equal up to rounding.  A pair whose terms vanish at
	// every point keeps an uninitialized spline (cutoff 0: value 0).
	int kind = 0;      // 0: precalculate_linear, 1: precalculate_splines
	int nseg = 0;      // spline segments per type pair
	fl fraction = 0;   // segment length in A
	std::vector<fl> spl;     // [tri][nseg][5]: x, a, b, c, d
	std::vector<fl> spl_cut; // [tri]: the spline's cutoff, 0 for a zero spline
	void build_splines(const Scoring& sf, fl factor_)
	{
		kind = 1;
		const fl cutoff = sf.cutoff;
		cutoff_sqr = sqr(cutoff);
		const unsigned np = unsigned(factor_ * cutoff);
		nseg = int(np);
		fraction = cutoff / fl(np);
		const int ntri = NUM_TYPES * (NUM_TYPES + 1) / 2;
		spl.assign(size_t(ntri) * nseg * 5, 0);
		spl_cut.assign(size_t(ntri), 0);
		const int e = int(np); // points 0..e
		std::vector<fl> xs(e + 1), ys(e + 1), ddy(e + 1), cp(e + 1), rp(e + 1);
		for (int t1 = 0; t1 < NUM_TYPES; t1++)
			for (int t2 = t1; t2 < NUM_TYPES; t2++)
			{
				const int tri = sgpu::tri_index(t1, t2);
				bool nonzero = false;
				for (int i = 0; i < e; i++)
				{
					xs[i] = i * fraction;
					ys[i] = sf.eval_fast(t1, t2, xs[i]);
					if (ys[i] != 0)
						nonzero = true;
				}
				xs[e] = cutoff;
				ys[e] = 0;
				if (!nonzero)
					continue;
				const fl h = xs[1] - xs[0];
				const fl hlast = xs[e] - xs[e - 1];
				// the equations: row 0: 2h y0'' + h y1'' = 6 (y1 - y0) / h;
				// row i: hi y''(i-1) + 2 (h + hi) y''i + hi y''(i+1) =
				// 6 ((y(i+1) - yi) / hi - (yi - y(i-1)) / h), hi = hlast for
				// the last interior row; row e: hlast y''(e-1) + 2 hlast y''e
				// = -6 (ye - y(e-1)) / hlast
				auto sub = [&](int i) { return i == e ? hlast : (i == e - 1 ? hlast : h); };
				auto diag = [&](int i) {
					if (i == 0)
						return 2 * h;
					if (i == e)
						return 2 * hlast;
					const fl hi = i == e - 1 ? hlast : h;
					return 2 * (h + hi);
				};
				auto sup = [&](int i) { return i == 0 ? h : (i == e - 1 ? hlast : h); };
				auto rhs = [&](int i) {
					if (i == 0)
						return 6 * ((ys[1] - ys[0]) / h);
					if (i == e)
						return 6 * (-(ys[e] - ys[e - 1]) / hlast);
					const fl hi = i == e - 1 ? hlast : h;
					return 6 * ((ys[i + 1] - ys[i]) / hi - (ys[i] - ys[i - 1]) / h);
				};
				cp[0] = sup(0) / diag(0);
				rp[0] = rhs(0) / diag(0);
				for (int i = 1; i <= e; i++)
				{
					const fl m = diag(i) - sub(i) * cp[i - 1];
					cp[i] = i < e ? sup(i) / m : 0;
					rp[i] = (rhs(i) - sub(i) * rp[i - 1]) / m;
				}
				ddy[e] = rp[e];
				for (int i = e - 1; i >= 0; i--)
					ddy[i] = rp[i] - cp[i] * ddy[i + 1];
				for (int i = 0; i < e; i++)
				{
					const fl hi = i == e - 1 ? hlast : h;
					fl* s = &spl[(size_t(tri) * nseg + i) * 5];
					s[0] = xs[i];
					s[1] = (ddy[i + 1] - ddy[i]) / (6 * hi);
					s[2] = ddy[i] / 2;
					s[3] = (ys[i + 1] - ys[i]) / hi - ddy[i + 1] * hi / 6 - ddy[i] * hi / 3;
					s[4] = ys[i];
				}
				spl_cut[size_t(tri)] = cutoff;
			}
	}
	// Spline::eval_deriv (src/lib/splines.h:103-121) at distance r
	void spline_eval(int tri, fl r, fl& val, fl& der) const
	{
		if (r >= spl_cut[size_t(tri)])
		{
			val = 0;
			der = 0;
			return;
		}
		const unsigned index = unsigned(r / fraction);
		const fl* s = &spl[(size_t(tri) * nseg + index) * 5];
		const fl lx = r - s[0];
		val = ((s[1] * lx + s[2]) * lx + s[3]) * lx + s[4];
		der = (3 * s[1] * lx + 2 * s[2]) * lx + s[3];
	}
	// precalculate_linear_element::eval_deriv with one component, or
	// precalculate_splines::eval_deriv (src/lib/precalculate.h:418-450, no
	// slow terms): the spline's derivative over r
	// precalculate_linear::eval_deriv (src/lib/precalculate.h:245-293) of
	// atoms of types t1, t2 and charges q1, q2: the element of (min type,
	// max type) with the atoms in that order; with four components each
	// knot's value is result_components::eval (c0 + |qa| c1 + |qb| c2 +
	// qa qb c3, src/lib/result_components.h:73-79) before interpolation
	void eval_deriv(int t1, int t2, fl r2, fl& eo, fl& doro, fl q1 = 0, fl q2 = 0) const
	{
		if (kind == 1)
		{
			const fl r = std::sqrt(r2);
			fl val, der;
			spline_eval(sgpu::tri_index_permissive(t1, t2), r, val, der);
			eo = val;
			doro = der / r;
			return;
		}
		const size_t base = size_t(sgpu::tri_index_permissive(t1, t2)) * n;
		fl r2_factored = factor * r2;
		size_t i1 = size_t(r2_factored);
		size_t i2 = i1 + 1;
		fl rem = r2_factored - i1;
		fl e1, e2, d1, d2;
		if (ncomp == 1)
		{
			e1 = e[base + i1];
			e2 = e[base + i2];
			d1 = dor[base + i1];
			d2 = dor[base + i2];
		}
		else
		{
			const fl qa = t1 <= t2 ? q1 : q2, qb = t1 <= t2 ? q2 : q1;
			auto comb = [&](const std::vector<fl>& c0, const std::vector<fl>& cq, size_t i) {
				const fl* q = &cq[(base + i) * 3];
				return c0[base + i] + std::abs(qa) * q[0] + std::abs(qb) * q[1] + qa * qb * q[2];
			};
			e1 = comb(e, eq, i1);
			e2 = comb(e, eq, i2);
			d1 = comb(dor, dorq, i1);
			d2 = comb(dor, dorq, i2);
		}
		eo = e1 + rem * (e2 - e1);
		doro = d1 + rem * (d2 - d1);
	}
	// precalculate::eval (eval_fast(...).eval(a, b) + eval_slow, no slow
	// terms): the fast components of (min type, max type), swapped
	// (swapOrder) when t1 > t2, then result_components::eval with the atoms
	// in the given order
	fl eval(int t1, int t2, fl r2, fl q1 = 0, fl q2 = 0) const
	{
		fl c[4] = {fast_value(t1, t2, r2), 0, 0, 0};
		if (ncomp == 4)
		{
			const size_t base = size_t(sgpu::tri_index_permissive(t1, t2)) * n;
			const fl* q = &fastq[(base + size_t(factor * r2)) * 3];
			c[1] = t1 <= t2 ? q[0] : q[1];
			c[2] = t1 <= t2 ? q[1] : q[0];
			c[3] = q[2];
		}
		fl ret = c[0] + std::abs(q1) * c[1] + std::abs(q2) * c[2] + q1 * q2 * c[3];
		return ret + 0.0;
	}
	// the four fast components of receptor type tr and ligand type tl in
	// that order (eval_fast_data, swapped when tr > tl), for the maps
	void fast_components(int tr, int tl, fl r2, fl* c) const
	{
		c[0] = fast_value(tr, tl, r2);
		c[1] = c[2] = c[3] = 0;
		if (ncomp == 4)
		{
			const size_t base = size_t(sgpu::tri_index_permissive(tr, tl)) * n;
			const fl* q = &fastq[(base + size_t(factor * r2)) * 3];
			c[1] = tr <= tl ? q[0] : q[1];
			c[2] = tr <= tl ? q[1] : q[0];
			c[3] = q[2];
		}
	}
	fl fast_value(int t1, int t2, fl r2) const
	{
		if (kind == 1)
		{
			// precalculate_splines::eval_fast: the spline's value
			fl val, der;
			spline_eval(sgpu::tri_index_permissive(t1, t2), std::sqrt(r2), val, der);
			return val;
		}
		const size_t base = size_t(sgpu::tri_index_permissive(t1, t2)) * n;
		return fast[base + size_t(factor * r2)];
	}
};

// ---------------------------------------------------------------- grid
struct GridDims
{
	fl begin[3], end[3];
	int n[3];
};
// grid::init (src/lib/grid.cpp:60-75) + evaluate_aux (:111-236)
struct Grid
{
	int dim[3] = {0, 0, 0};
	vec init, factor, dim_fl_minus_1, factor_inv;
	std::vector<fl> data; // x fastest
	void setup(const GridDims& gd)
	{
		for (int i = 0; i < 3; i++)
		{
			dim[i] = gd.n[i] + 1;
			init[i] = gd.begin[i];
			fl range = gd.end[i] - gd.begin[i];
			dim_fl_minus_1[i] = dim[i] - 1.0;
			factor[i] = dim_fl_minus_1[i] / range;
			factor_inv[i] = 1 / factor[i];
		}
	}
	fl& at(int x, int y, int z) { return data[size_t(x) + size_t(dim[0]) * (size_t(y) + size_t(dim[1]) * z)]; }
	fl at(int x, int y, int z) const { return data[size_t(x) + size_t(dim[0]) * (size_t(y) + size_t(dim[1]) * z)]; }
	vec index_to_argument(int x, int y, int z) const
	{
		return vec(init[0] + factor_inv[0] * x, init[1] + factor_inv[1] * y, init[2] + factor_inv[2] * z);
	}
	fl evaluate(const vec& location, fl slope, fl v, vec* deriv) const
	{
		vec s = location - init;
		s = vec(s[0] * factor[0], s[1] * factor[1], s[2] * factor[2]);
		vec miss(0, 0, 0);
		int region[3];
		int a[3];
		for (int i = 0; i < 3; i++)
		{
			if (s[i] < 0)
			{
				miss[i] = -s[i];
				region[i] = -1;
				a[i] = 0;
				s[i] = 0;
			}
			else if (s[i] >= dim_fl_minus_1[i])
			{
				miss[i] = s[i] - dim_fl_minus_1[i];
				region[i] = 1;
				a[i] = dim[i] - 2;
				s[i] = 1;
			}
			else
			{
				region[i] = 0;
				a[i] = int(size_t(s[i]));
				s[i] -= size_t(a[i]);
			}
		}
		const fl penalty = slope * (miss * factor_inv);
		const int x0 = a[0], y0 = a[1], z0 = a[2];
		const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
		const fl f000 = at(x0, y0, z0), f100 = at(x1, y0, z0), f010 = at(x0, y1, z0), f110 = at(x1, y1, z0);
		const fl f001 = at(x0, y0, z1), f101 = at(x1, y0, z1), f011 = at(x0, y1, z1), f111 = at(x1, y1, z1);
		const fl x = s[0], y = s[1], z = s[2];
		const fl mx = 1 - x, my = 1 - y, mz = 1 - z;
		fl f = f000 * mx * my * mz + f100 * x * my * mz + f010 * mx * y * mz + f110 * x * y * mz +
				f001 * mx * my * z + f101 * x * my * z + f011 * mx * y * z + f111 * x * y * z;
		if (deriv)
		{
			const fl x_g = f000 * (-1) * my * mz + f100 * 1 * my * mz + f010 * (-1) * y * mz + f110 * 1 * y * mz +
					f001 * (-1) * my * z + f101 * 1 * my * z + f011 * (-1) * y * z + f111 * 1 * y * z;
			const fl y_g = f000 * mx * (-1) * mz + f100 * x * (-1) * mz + f010 * mx * 1 * mz + f110 * x * 1 * mz +
					f001 * mx * (-1) * z + f101 * x * (-1) * z + f011 * mx * 1 * z + f111 * x * 1 * z;
			const fl z_g = f000 * mx * my * (-1) + f100 * x * my * (-1) + f010 * mx * y * (-1) +
					f110 * x * y * (-1) + f001 * mx * my * 1 + f101 * x * my * 1 + f011 * mx * y * 1 +
					f111 * x * y * 1;
			vec gradient(x_g, y_g, z_g);
			curl(f, gradient, v);
			for (int i = 0; i < 3; i++)
			{
				fl ge = (region[i] == 0) ? gradient[i] : 0;
				(*deriv)[i] = factor[i] * ge + slope * region[i];
			}
			return f + penalty;
		}
		curl(f, v);
		return f + penalty;
	}
};

inline fl closest_between(fl begin, fl end, fl x)
{
	if (x <= begin)
		return begin;
	else if (x >= end)
		return end;
	return x;
}
inline fl brick_distance_sqr(const vec& begin, const vec& end, const vec& v)
{
	vec c(closest_between(begin[0], end[0], v[0]), closest_between(begin[1], end[1], v[1]),
			closest_between(begin[2], end[2], v[2]));
	return vec_distance_sqr(c, v);
}

struct RecAtom
{
	vec coords;
	int type;
	fl charge;
};

// ---------------------------------------------------------------- model
struct Frame
{
	int kind;   // SGPU_FRAME_*
	int parent; // -1 root
	int begin, end;
	int tree;
	std::vector<int> children;
	vec origin;
	qt q = qt_identity;
	mat R;
	vec axis;
	vec rel_origin, rel_axis;
	void set_orientation(const qt& nq)
	{
		q = nq;
		R = quaternion_to_r3(q);
	}
};

struct Tree
{
	bool ligand;
	int root;
	int conf_off, change_off;
	int ntors;
	int atom_begin, atom_end;
};

struct Pair
{
	int a, b;
};

struct Model
{
	int N = 0;       // movable atoms
	int natoms = 0;  // movable + inflex
	int nligands = 0;
	std::vector<int> type;
	std::vector<fl> charge;  // partial charges, every atom
	std::vector<vec> local;  // movable: relative to frame origin at identity; inflex: absolute
	std::vector<std::vector<int>> bonds;
	std::vector<Frame> frames;
	std::vector<Tree> trees;
	std::vector<Pair> other_pairs;
	std::vector<std::vector<Pair>> lig_pairs;
	int conf_size = 0, change_size = 0;
	fl num_tors = 0;
	// state
	std::vector<vec> coords, minus_forces;

	int find_ligand(int a) const
	{
		for (int i = 0; i < nligands; i++)
			if (a >= trees[i].atom_begin && a < trees[i].atom_end)
				return i;
		return nligands;
	}
	// model::bonded_to (src/lib/model.cpp:576-589), depth-first with the
	// "already listed" check
	void bonded_to(int a, int n, std::vector<int>& out) const
	{
		if (std::find(out.begin(), out.end(), a) == out.end())
		{
			out.push_back(a);
			if (n > 0)
				for (int b : bonds[a])
					bonded_to(b, n - 1, out);
		}
	}
	void build(const sgpu_model_desc& d)
	{
		N = d.nmovable;
		natoms = d.natoms;
		nligands = d.nligands;
		type.assign(d.type, d.type + natoms);
		charge.assign(natoms, 0);
		if (d.charge)
			charge.assign(d.charge, d.charge + natoms);
		std::vector<vec> xyz(natoms);
		for (int i = 0; i < natoms; i++)
			xyz[i] = vec(d.xyz[3 * i], d.xyz[3 * i + 1], d.xyz[3 * i + 2]);
		bonds.assign(natoms, {});
		for (int k = 0; k < d.nbonds; k++)
		{
			int a = d.bonds[2 * k], b = d.bonds[2 * k + 1];
			bonds[a].push_back(b);
			bonds[b].push_back(a);
		}
		frames.resize(d.nframes);
		trees.resize(d.ntrees);
		std::vector<int> atom_frame(N, -1);
		for (int t = 0; t < d.ntrees; t++)
		{
			trees[t].ligand = t < d.nligands;
			trees[t].root = d.tree_frame[t];
		}
		for (int f = 0; f < d.nframes; f++)
		{
			Frame& fr = frames[f];
			fr.parent = d.frame_parent[f];
			fr.begin = d.frame_begin[f];
			fr.end = d.frame_end[f];
			fr.tree = -1;
			for (int t = d.ntrees - 1; t >= 0; t--)
				if (f >= d.tree_frame[t])
				{
					fr.tree = t;
					break;
				}
			for (int i = fr.begin; i < fr.end; i++)
				atom_frame[i] = f;
			fr.origin = xyz[d.frame_origin_atom[f]];
			if (fr.parent < 0)
				fr.kind = trees[fr.tree].ligand ? SGPU_FRAME_RIGID_ROOT : SGPU_FRAME_FIRST_SEGMENT;
			else
			{
				fr.kind = SGPU_FRAME_SEGMENT;
				frames[fr.parent].children.push_back(f);
			}
			if (fr.kind != SGPU_FRAME_RIGID_ROOT)
			{ // axis_frame (src/lib/tree.h:116-125)
				vec diff = fr.origin - xyz[d.frame_axis_atom[f]];
				fl nrm = diff.norm();
				if (!(nrm >= epsilon_fl))
					throw std::runtime_error("zero-length rotatable bond");
				fr.axis = (1 / nrm) * diff;
			}
			if (fr.kind == SGPU_FRAME_SEGMENT)
			{ // segment (src/lib/tree.h:139-144)
				fr.rel_axis = fr.axis;
				fr.rel_origin = fr.origin - frames[fr.parent].origin;
			}
		}
		local.resize(natoms);
		for (int i = 0; i < natoms; i++)
			local[i] = i < N ? xyz[i] - frames[atom_frame[i]].origin : xyz[i];
		// conf / change layout
		conf_size = change_size = 0;
		for (int t = 0; t < d.ntrees; t++)
		{
			Tree& tr = trees[t];
			int last = (t + 1 < d.ntrees) ? d.tree_frame[t + 1] : d.nframes;
			tr.ntors = (last - tr.root) - (tr.ligand ? 1 : 0);
			tr.conf_off = conf_size;
			tr.change_off = change_size;
			conf_size += tr.ligand ? 7 + tr.ntors : tr.ntors;
			change_size += tr.ligand ? 6 + tr.ntors : tr.ntors;
			tr.atom_begin = frames[tr.root].begin;
			tr.atom_end = frames[last - 1].end;
			for (int f = tr.root; f < last; f++)
				tr.atom_end = std::max(tr.atom_end, frames[f].end);
		}
		// mobility: fixed within a frame and between a segment's axis atoms
		// and its atoms (src/lib/parse_pdbqt.cpp:279-336); else variable
		auto fixed = [&](int i, int j) -> bool {
			if (i >= N && j >= N)
				return true;
			if (i >= N || j >= N)
			{
				int m = i < N ? i : j, o = i < N ? j : i;
				int f = atom_frame[m];
				return frames[f].kind == SGPU_FRAME_FIRST_SEGMENT && d.frame_axis_atom[f] == o;
			}
			int fi = atom_frame[i], fj = atom_frame[j];
			if (fi == fj)
				return true;
			if (frames[fj].kind != SGPU_FRAME_RIGID_ROOT &&
					(d.frame_axis_atom[fj] == i || d.frame_origin_atom[fj] == i))
				return true;
			if (frames[fi].kind != SGPU_FRAME_RIGID_ROOT &&
					(d.frame_axis_atom[fi] == j || d.frame_origin_atom[fi] == j))
				return true;
			return false;
		};
		// model::initialize_pairs (src/lib/model.cpp:598-624)
		lig_pairs.assign(nligands, {});
		other_pairs.clear();
		for (int i = 0; i < natoms; i++)
		{
			int i_lig = find_ligand(i);
			std::vector<int> bonded;
			bonded_to(i, 3, bonded);
			for (int j = i + 1; j < natoms; j++)
			{
				if (i >= N && j >= N)
					continue;
				if (!fixed(i, j) && std::find(bonded.begin(), bonded.end(), j) == bonded.end())
				{
					int t1 = type[i], t2 = type[j];
					if (t1 < NUM_TYPES && t2 < NUM_TYPES && !is_hydrogen_type(t1) && !is_hydrogen_type(t2))
					{
						if (i_lig < nligands && find_ligand(j) == i_lig)
							lig_pairs[i_lig].push_back({i, j});
						else
							other_pairs.push_back({i, j});
					}
				}
			}
		}
		// conf_independent_inputs::num_tors (src/lib/terms.cpp:39-110) with
		// fixed rotatable hydrogens: rotatable = a bond across two frames
		num_tors = 0;
		auto heavy_nbrs = [&](int a) {
			int k = 0;
			for (int b : bonds[a])
				if (!is_hydrogen_type(type[b]))
					++k;
			return k;
		};
		for (int l = 0; l < nligands; l++)
			for (int i = trees[l].atom_begin; i < trees[l].atom_end; i++)
			{
				if (is_hydrogen_type(type[i]))
					continue;
				unsigned ar = 0;
				for (int b : bonds[i])
				{
					bool rot = b < N && atom_frame[b] != atom_frame[i] && !fixed(i, b);
					// a bond is rotatable when it is a tree bond between frames
					for (int f = 0; f < int(frames.size()) && !rot; f++)
						if (frames[f].kind == SGPU_FRAME_SEGMENT &&
								((d.frame_axis_atom[f] == i && d.frame_origin_atom[f] == b) ||
										(d.frame_axis_atom[f] == b && d.frame_origin_atom[f] == i)))
							rot = true;
					if (rot && !is_hydrogen_type(type[b]) && heavy_nbrs(b) > 1 && heavy_nbrs(i) > 1)
						++ar;
				}
				num_tors += 0.5 * ar;
			}
		coords.assign(xyz.begin(), xyz.end());
		minus_forces.assign(natoms, vec(0, 0, 0));
		init_coords = coords;
		init_frames = frames;
	}
	// the state of a freshly parsed model: input coordinates, frames at
	// their input origins with identity orientation
	std::vector<vec> init_coords;
	std::vector<Frame> init_frames;
	void reset()
	{
		coords = init_coords;
		frames = init_frames;
		minus_forces.assign(natoms, vec(0, 0, 0));
	}

	// ---- model::set (src/lib/model.cpp:837-841) through tree.h
	void set_coords_of(const Frame& fr)
	{
		for (int i = fr.begin; i < fr.end; i++)
			coords[i] = fr.origin + fr.R * local[i]; // atom_frame::set_coords
	}
	void set_branch(int f, const fl*& tors)
	{
		Frame& fr = frames[f];
		const Frame& par = frames[fr.parent];
		const fl torsion = *tors++;
		fr.origin = par.origin + par.R * fr.rel_origin;
		fr.axis = par.R * fr.rel_axis;
		qt tmp = angle_to_quaternion(fr.axis, torsion) * par.q;
		quaternion_normalize_approx(tmp);
		fr.set_orientation(tmp);
		set_coords_of(fr);
		for (int c : fr.children)
			set_branch(c, tors);
	}
	void set(const fl* c)
	{
		for (const Tree& tr : trees)
		{
			Frame& root = frames[tr.root];
			const fl* p = c + tr.conf_off;
			if (tr.ligand)
			{ // rigid_body::set_conf
				root.origin = vec(p[0], p[1], p[2]);
				root.set_orientation({p[3], p[4], p[5], p[6]});
				set_coords_of(root);
				const fl* tors = p + 7;
				for (int ch : root.children)
					set_branch(ch, tors);
			}
			else
			{ // first_segment::set_conf
				const fl* tors = p;
				root.set_orientation(angle_to_quaternion(root.axis, *tors++));
				set_coords_of(root);
				for (int ch : root.children)
					set_branch(ch, tors);
			}
		}
	}
	// ---- derivative (src/lib/tree.h:88-97, 214-289)
	void sum_force_and_torque(const Frame& fr, vec& f, vec& t) const
	{
		f = vec(0, 0, 0);
		t = vec(0, 0, 0);
		for (int i = fr.begin; i < fr.end; i++)
		{
			f += minus_forces[i];
			t += cross_product(coords[i] - fr.origin, minus_forces[i]);
		}
	}
	void branch_derivative(int f, vec& fo, vec& to, fl*& d) const
	{
		const Frame& fr = frames[f];
		sum_force_and_torque(fr, fo, to);
		fl* mine = d++;
		for (int ch : fr.children)
		{
			vec cf, ct;
			branch_derivative(ch, cf, ct, d);
			fo += cf;
			vec r = frames[ch].origin - fr.origin;
			to += cross_product(r, cf) + ct;
		}
		*mine = to * fr.axis;
	}
	void derivative(fl* g) const
	{
		for (const Tree& tr : trees)
		{
			const Frame& root = frames[tr.root];
			fl* p = g + tr.change_off;
			vec fo, to;
			sum_force_and_torque(root, fo, to);
			if (tr.ligand)
			{
				fl* d = p + 6;
				for (int ch : root.children)
				{
					vec cf, ct;
					branch_derivative(ch, cf, ct, d);
					fo += cf;
					vec r = frames[ch].origin - root.origin;
					to += cross_product(r, cf) + ct;
				}
				p[0] = fo[0];
				p[1] = fo[1];
				p[2] = fo[2];
				p[3] = to[0];
				p[4] = to[1];
				p[5] = to[2];
			}
			else
			{
				fl* mine = p;
				fl* d = p + 1;
				for (int ch : root.children)
				{
					vec cf, ct;
					branch_derivative(ch, cf, ct, d);
					fo += cf;
					vec r = frames[ch].origin - root.origin;
					to += cross_product(r, cf) + ct;
				}
				*mine = to * root.axis;
			}
		}
	}
	// heavy movable coords (model::get_heavy_atom_movable_coords)
	std::vector<vec> heavy_coords() const
	{
		std::vector<vec> out;
		for (int i = 0; i < N; i++)
			if (!is_hydrogen_type(type[i]))
				out.push_back(coords[i]);
		return out;
	}
	int heavy_count() const
	{
		int k = 0;
		for (int i = 0; i < N; i++)
			if (!is_hydrogen_type(type[i]))
				++k;
		return k;
	}
	// model::gyration_radius (src/lib/model.cpp:868-883)
	fl gyration_radius(int l) const
	{
		const Tree& tr = trees[l];
		fl acc = 0;
		unsigned counter = 0;
		for (int i = tr.atom_begin; i < tr.atom_end; i++)
			if (!is_hydrogen_type(type[i]))
			{
				acc += vec_distance_sqr(coords[i], frames[tr.root].origin);
				++counter;
			}
		return counter > 0 ? std::sqrt(acc / counter) : 0;
	}
};

// ---------------------------------------------------------------- conf ops
struct ConfOps
{
	const Model* m;
	int change_size() const { return m->change_size; }
	int conf_size() const { return m->conf_size; }
	// conf::increment (src/lib/conf.h:316-321, 168-172, 73-77)
	void increment(fl* x, const fl* c, fl factor) const
	{
		for (const Tree& tr : m->trees)
		{
			fl* p = x + tr.conf_off;
			const fl* q = c + tr.change_off;
			int nt = tr.ntors;
			fl* tors = p;
			const fl* dt = q;
			if (tr.ligand)
			{
				vec dp = factor * vec(q[0], q[1], q[2]);
				p[0] += dp[0];
				p[1] += dp[1];
				p[2] += dp[2];
				vec rotation = factor * vec(q[3], q[4], q[5]);
				qt o = {p[3], p[4], p[5], p[6]};
				quaternion_increment(o, rotation);
				p[3] = o.a;
				p[4] = o.b;
				p[5] = o.c;
				p[6] = o.d;
				tors = p + 7;
				dt = q + 6;
			}
			for (int i = 0; i < nt; i++)
			{
				tors[i] += normalized_angle(factor * dt[i]);
				normalize_angle(tors[i]);
			}
		}
	}
	// conf::operator()(index) (src/lib/conf.h:376-397)
	fl value(const fl* x, int index) const
	{
		for (const Tree& tr : m->trees)
		{
			const fl* p = x + tr.conf_off;
			if (tr.ligand)
			{
				if (index < 3)
					return p[index];
				index -= 3;
				if (index < 3)
				{
					vec ang = quaternion_to_angle({p[3], p[4], p[5], p[6]});
					return ang[index];
				}
				index -= 3;
				if (index < tr.ntors)
					return p[7 + index];
				index -= tr.ntors;
			}
			else
			{
				if (index < tr.ntors)
					return p[index];
				index -= tr.ntors;
			}
		}
		return 0;
	}
};

// ---------------------------------------------------------------- context
struct Ctx
{
	Scoring sf;
	Tables tab;
	GridDims gd;
	Grid geom;
	std::vector<RecAtom> rec; // all receptor atoms (incl. H)
	std::vector<Grid> grids;  // per type
	std::vector<Grid> cgrids; // per type: the charge map (four-component tables)
	std::vector<bool> have_grid;
	// --user_grid (grid::init from a map file, src/lib/grid.cpp:77-108)
	Grid ug;
	bool has_ug = false;
	fl ug_map_slope = 1e6;
	// grid::evaluate_user (src/lib/grid.cpp:50-53): evaluate_aux with cap 1000
	fl user_grid(const vec& location, fl slope, vec* deriv) const { return ug.evaluate(location, slope, 1000, deriv); }
	std::vector<std::unique_ptr<Model>> models;
	// non_cache cells (full 3 A bricks over the relevant atoms)
	int cell_off[3], cell_dim[3];
	std::vector<std::vector<int>> cells;
	std::string err;

	void setup_noncache()
	{
		const fl cutoff_sqr = tab.cutoff_sqr;
		const fl gran = 3.0;
		std::vector<int> relevant;
		vec start(gd.begin[0], gd.begin[1], gd.begin[2]), end(gd.end[0], gd.end[1], gd.end[2]);
		for (size_t i = 0; i < rec.size(); i++)
			if (!is_hydrogen_type(rec[i].type) && brick_distance_sqr(start, end, rec[i].coords) < cutoff_sqr)
				relevant.push_back(int(i));
		for (int i = 0; i < 3; i++)
		{
			cell_off[i] = int(std::floor(gd.begin[i] / gran));
			cell_dim[i] = int(std::ceil(gd.end[i] / gran)) - cell_off[i] + 1;
		}
		cells.assign(size_t(cell_dim[0]) * cell_dim[1] * cell_dim[2], {});
		for (int z = 0; z < cell_dim[2]; z++)
			for (int y = 0; y < cell_dim[1]; y++)
				for (int x = 0; x < cell_dim[0]; x++)
				{
					vec lo((cell_off[0] + x) * gran, (cell_off[1] + y) * gran, (cell_off[2] + z) * gran);
					vec hi(lo[0] + gran, lo[1] + gran, lo[2] + gran);
					auto& c = cells[size_t(x) + size_t(cell_dim[0]) * (y + size_t(cell_dim[1]) * z)];
					for (int i : relevant)
						if (brick_distance_sqr(lo, hi, rec[i].coords) < cutoff_sqr)
							c.push_back(i);
				}
	}
	const std::vector<int>& cell_of(const vec& p) const
	{
		int ix[3];
		for (int i = 0; i < 3; i++)
		{
			ix[i] = int(std::floor(p[i] / 3.0)) - cell_off[i];
			ix[i] = std::max(0, std::min(cell_dim[i] - 1, ix[i]));
		}
		return cells[size_t(ix[0]) + size_t(cell_dim[0]) * (ix[1] + size_t(cell_dim[1]) * ix[2])];
	}

	// cache::populate (src/lib/cache.cpp:116-210) with the szv_grid_cache
	// semantics of src/lib/szv_grid.h:56-165: a cell's atom list is made by
	// the first probe that falls into it
	void populate(const std::vector<int>& types)
	{
		std::vector<int> needed;
		for (int t : types)
			if (!have_grid[t])
			{
				needed.push_back(t);
				grids[t].setup(gd);
				grids[t].data.assign(size_t(grids[t].dim[0]) * grids[t].dim[1] * grids[t].dim[2], 0);
				if (tab.ncomp == 4)
				{
					cgrids[t].setup(gd);
					cgrids[t].data.assign(grids[t].data.size(), 0);
				}
				have_grid[t] = true;
			}
		if (needed.empty())
			return;
		const fl cutoff_sqr = tab.cutoff_sqr;
		const fl gran = 3.0;
		std::vector<int> relevant;
		vec start(gd.begin[0], gd.begin[1], gd.begin[2]), end(gd.end[0], gd.end[1], gd.end[2]);
		for (size_t i = 0; i < rec.size(); i++)
			if (!is_hydrogen_type(rec[i].type) && brick_distance_sqr(start, end, rec[i].coords) < cutoff_sqr)
				relevant.push_back(int(i));
		std::map<std::array<int, 3>, std::vector<int>> cache;
		Grid& g = grids[needed.front()];
		std::vector<fl> aff(needed.size()), caff(needed.size());
		const bool charged = tab.ncomp == 4;
		for (int x = 0; x < g.dim[0]; x++)
			for (int y = 0; y < g.dim[1]; y++)
				for (int z = 0; z < g.dim[2]; z++)
				{
					std::fill(aff.begin(), aff.end(), 0);
					std::fill(caff.begin(), caff.end(), 0);
					vec probe = g.index_to_argument(x, y, z);
					std::array<int, 3> key;
					for (int i = 0; i < 3; i++)
						key[i] = int(std::floor(probe[i] / gran));
					auto it = cache.find(key);
					if (it == cache.end())
					{
						vec lower, upper;
						for (int i = 0; i < 3; i++)
						{
							lower[i] = std::floor(probe[i] / gran) * gran;
							upper[i] = std::ceil(probe[i] / gran) * gran;
						}
						std::vector<int> atoms;
						for (int i : relevant)
							if (brick_distance_sqr(lower, upper, rec[i].coords) < cutoff_sqr)
								atoms.push_back(i);
						it = cache.emplace(key, std::move(atoms)).first;
					}
					for (int i : it->second)
					{
						const RecAtom& a = rec[i];
						const fl r2 = vec_distance_sqr(a.coords, probe);
						if (r2 <= cutoff_sqr)
							for (size_t j = 0; j < needed.size(); j++)
							{
								if (!charged)
								{
									aff[j] += tab.fast_value(a.type, needed[j], r2);
									continue;
								}
								// src/lib/cache.cpp:172-187: the receptor atom's
								// charge in the map, the ligand atom's applied
								// to the charge map at evaluation
								fl val[4];
								tab.fast_components(a.type, needed[j], r2, val);
								aff[j] += val[0] + val[1] * std::abs(a.charge);
								caff[j] += val[2] + val[3] * a.charge;
							}
					}
					for (size_t j = 0; j < needed.size(); j++)
					{
						grids[needed[j]].at(x, y, z) = aff[j];
						if (charged)
							cgrids[needed[j]].at(x, y, z) = caff[j];
						// src/lib/cache.cpp:204-205: the user grid at the
						// point's integer indices taken as a location
						if (has_ug)
							grids[needed[j]].at(x, y, z) += user_grid(vec(x, y, z), ug_map_slope, nullptr);
					}
				}
	}

	// ---- back ends of model::eval_deriv
	// cache::eval_deriv (src/lib/cache.cpp:69-90)
	fl cache_eval_deriv(Model& m, fl slope, fl v) const
	{
		fl e = 0;
		for (int i = 0; i < m.N; i++)
		{
			int t = m.type[i];
			if (t >= NUM_TYPES || is_hydrogen_type(t))
			{
				m.minus_forces[i] = vec(0, 0, 0);
				continue;
			}
			if (!have_grid[t])
				throw std::runtime_error("grid for a ligand type was not built");
			// grid::evaluate (src/lib/grid.cpp:28-48): the charge map scaled by
			// the atom's charge, value and derivative, when it has one
			vec deriv;
			fl ge = grids[t].evaluate(m.coords[i], slope, v, &deriv);
			if (m.charge[i] != 0 && tab.ncomp == 4)
			{
				vec cderiv(0, 0, 0);
				ge += m.charge[i] * cgrids[t].evaluate(m.coords[i], slope, v, &cderiv);
				deriv += m.charge[i] * cderiv;
			}
			e += ge;
			m.minus_forces[i] = deriv;
		}
		return e;
	}
	// non_cache::eval_deriv (src/lib/non_cache.cpp:109-191)
	fl noncache_eval_deriv(Model& m, fl slope, fl v) const
	{
		fl e = 0;
		const fl cutoff_sqr = tab.cutoff_sqr;
		for (int i = 0; i < m.N; i++)
		{
			fl this_e = 0;
			vec deriv(0, 0, 0);
			vec ug_deriv(0, 0, 0);
			vec oob_deriv(0, 0, 0);
			fl oob_penalty = 0;
			int t1 = m.type[i];
			if (t1 >= NUM_TYPES || is_hydrogen_type(t1))
			{
				m.minus_forces[i] = vec(0, 0, 0);
				continue;
			}
			const vec& ac = m.coords[i];
			vec adj = ac;
			for (int j = 0; j < 3; j++)
				if (gd.n[j] > 0)
				{
					if (ac[j] < gd.begin[j])
					{
						adj[j] = gd.begin[j];
						oob_deriv[j] = -1;
						oob_penalty += std::abs(ac[j] - gd.begin[j]);
					}
					else if (ac[j] > gd.end[j])
					{
						adj[j] = gd.end[j];
						oob_deriv[j] = 1;
						oob_penalty += std::abs(ac[j] - gd.end[j]);
					}
				}
			oob_penalty *= slope;
			oob_deriv *= slope;
			for (int j : cell_of(adj))
			{
				const RecAtom& b = rec[j];
				vec r_ba = adj - b.coords;
				fl r2 = r_ba.norm_sqr();
				if (r2 < cutoff_sqr)
				{
					if (r2 < epsilon_fl)
						throw std::runtime_error("Ligand atom exactly overlaps receptor atom.  I can't deal with this.");
					fl pe, pd;
					tab.eval_deriv(t1, b.type, r2, pe, pd, m.charge[i], b.charge);
					this_e += pe;
					deriv += pd * r_ba;
				}
			}
			if (has_ug)
				this_e += user_grid(ac, slope, &ug_deriv);
			deriv += ug_deriv;
			curl(this_e, deriv, v);
			m.minus_forces[i] = deriv + oob_deriv;
			e += this_e + oob_penalty;
		}
		return e;
	}
	// non_cache::eval (src/lib/non_cache.cpp:24-79)
	fl noncache_eval(const Model& m, fl slope, fl v) const
	{
		fl e = 0;
		const fl cutoff_sqr = tab.cutoff_sqr;
		for (int i = 0; i < m.N; i++)
		{
			fl this_e = 0;
			fl oob_penalty = 0;
			int t1 = m.type[i];
			if (t1 >= NUM_TYPES || is_hydrogen_type(t1))
				continue;
			const vec& ac = m.coords[i];
			vec adj = ac;
			for (int j = 0; j < 3; j++)
				if (gd.n[j] > 0)
				{
					if (ac[j] < gd.begin[j])
					{
						adj[j] = gd.begin[j];
						oob_penalty += std::abs(ac[j] - gd.begin[j]);
					}
					else if (ac[j] > gd.end[j])
					{
						adj[j] = gd.end[j];
						oob_penalty += std::abs(ac[j] - gd.end[j]);
					}
				}
			oob_penalty *= slope;
			for (int j : cell_of(adj))
			{
				const RecAtom& b = rec[j];
				vec r_ba = adj - b.coords;
				fl r2 = r_ba.norm_sqr();
				if (r2 < cutoff_sqr)
					this_e += tab.eval(t1, b.type, r2, m.charge[i], b.charge);
			}
			curl(this_e, v);
			e += this_e + oob_penalty;
		}
		return e;
	}
	// model::eval_interacting_pairs_deriv (src/lib/model.cpp:904-928)
	fl pairs_deriv(Model& m, fl v, const std::vector<Pair>& pairs) const
	{
		const fl cutoff_sqr = tab.cutoff_sqr;
		fl e = 0;
		for (const Pair& ip : pairs)
		{
			vec r = m.coords[ip.b] - m.coords[ip.a];
			fl r2 = r.norm_sqr();
			if (r2 < cutoff_sqr)
			{
				fl pe, pd;
				tab.eval_deriv(m.type[ip.a], m.type[ip.b], r2, pe, pd, m.charge[ip.a], m.charge[ip.b]);
				vec force = pd * r;
				curl(pe, force, v);
				e += pe;
				m.minus_forces[ip.a] -= force;
				m.minus_forces[ip.b] += force;
			}
		}
		return e;
	}
	// model::eval_interacting_pairs (src/lib/model.cpp:885-902)
	fl pairs_eval(const Model& m, fl v, const std::vector<Pair>& pairs) const
	{
		const fl cutoff_sqr = tab.cutoff_sqr;
		fl e = 0;
		for (const Pair& ip : pairs)
		{
			fl r2 = vec_distance_sqr(m.coords[ip.a], m.coords[ip.b]);
			if (r2 < cutoff_sqr)
			{
				fl tmp = tab.eval(m.type[ip.a], m.type[ip.b], r2, m.charge[ip.a], m.charge[ip.b]);
				curl(tmp, v);
				e += tmp;
			}
		}
		return e;
	}
	// model::eval_deriv (src/lib/model.cpp:969-983)
	fl eval_deriv(Model& m, int backend, fl slope, const fl* v, const fl* c, fl* g) const
	{
		m.set(c);
		fl e = backend == 0 ? cache_eval_deriv(m, slope, v[1]) : noncache_eval_deriv(m, slope, v[1]);
		e += pairs_deriv(m, v[2], m.other_pairs);
		for (int l = 0; l < m.nligands; l++)
			e += pairs_deriv(m, v[0], m.lig_pairs[l]);
		m.derivative(g);
		return e;
	}
	// model::eval (src/lib/model.cpp:931-953) with the non_cache back end
	fl eval(Model& m, fl slope, const fl* v, const fl* c) const
	{
		m.set(c);
		fl e = noncache_eval(m, slope, v[1]);
		e += pairs_eval(m, v[2], m.other_pairs);
		for (int l = 0; l < m.nligands; l++)
			e += pairs_eval(m, v[0], m.lig_pairs[l]);
		// src/lib/model.cpp:953-964: the user grid at every atom of the
		// single ligand (get_ligand_coords), slope 1000
		if (has_ug)
			for (int i = 0; i < m.N; i++)
				if (m.find_ligand(i) == 0)
					e += user_grid(m.coords[i], 1000, nullptr);
		return e;
	}
	// model::eval_intramolecular (src/lib/model.cpp:1028-1083)
	fl eval_intramolecular(Model& m, const fl* v, const fl* c) const
	{
		m.set(c);
		fl e = 0;
		for (int l = 0; l < m.nligands; l++)
			e += pairs_eval(m, v[0], m.lig_pairs[l]);
		const fl cutoff_sqr = tab.cutoff_sqr;
		for (int i = 0; i < m.N; i++)
		{
			if (m.find_ligand(i) < m.nligands)
				continue;
			int t1 = m.type[i];
			if (t1 >= NUM_TYPES)
				continue;
			for (const RecAtom& b : rec)
			{
				if (b.type >= NUM_TYPES)
					continue;
				fl r2 = vec_distance_sqr(m.coords[i], b.coords);
				if (r2 < cutoff_sqr)
				{
					fl this_e = tab.eval(t1, b.type, r2, m.charge[i], b.charge);
					curl(this_e, v[1]);
					e += this_e;
				}
			}
		}
		for (const Pair& p : m.other_pairs)
		{
			if (m.find_ligand(p.a) < m.nligands || m.find_ligand(p.b) < m.nligands)
				continue;
			fl r2 = vec_distance_sqr(m.coords[p.a], m.coords[p.b]);
			if (r2 < cutoff_sqr)
			{
				fl this_e = tab.eval(m.type[p.a], m.type[p.b], r2, m.charge[p.a], m.charge[p.b]);
				curl(this_e, v[2]);
				e += this_e;
			}
		}
		return e;
	}
};

// ---------------------------------------------------------------- BFGS
struct Objective
{
	const Ctx* ctx;
	Model* m;
	int backend;
	fl slope;
	fl v[3];
	int64_t* evals;
	fl operator()(const fl* x, fl* g) const
	{
		if (evals)
			++*evals;
		return ctx->eval_deriv(*m, backend, slope, v, x, g);
	}
};

inline fl scalar_product(const fl* a, const fl* b, int n)
{
	fl tmp = 0;
	for (int i = 0; i < n; i++)
		tmp += a[i] * b[i];
	return tmp;
}
// triangular matrix h(i, j), i <= j at i + j(j+1)/2 (src/lib/matrix.h)
struct Hess
{
	int n;
	std::vector<fl> d;
	fl& at(int i, int j) { return d[size_t(i + j * (j + 1) / 2)]; }
	fl perm(int i, int j) const { return i < j ? d[size_t(i + j * (j + 1) / 2)] : d[size_t(j + i * (i + 1) / 2)]; }
};
// bfgs.h:31-42
void minus_mat_vec_product(const Hess& h, const fl* in, fl* out)
{
	for (int i = 0; i < h.n; i++)
	{
		fl sum = 0;
		for (int j = 0; j < h.n; j++)
			sum += h.perm(i, j) * in[j];
		out[i] = -sum;
	}
}
// bfgs.h:52-72
bool bfgs_update(Hess& h, const fl* p, const fl* y, fl alpha)
{
	const int n = h.n;
	const fl yp = scalar_product(y, p, n);
	if (alpha * yp < epsilon_fl)
		return false;
	std::vector<fl> minus_hy(n);
	minus_mat_vec_product(h, y, minus_hy.data());
	const fl yhy = -scalar_product(y, minus_hy.data(), n);
	const fl r = 1 / (alpha * yp);
	for (int i = 0; i < n; i++)
		for (int j = i; j < n; j++)
			h.at(i, j) += alpha * r * (minus_hy[i] * p[j] + minus_hy[j] * p[i]) +
					+alpha * alpha * (r * r * yhy + r) * p[i] * p[j];
	return true;
}
// bfgs.h:76-96
template<typename F, typename Ops>
fl fast_line_search(const F& f, const Ops& ops, int n, const fl* x, const fl* g, fl f0,
		const fl* p, fl* x_new, fl* g_new, fl& f1, int conf_size)
{
	const fl c0 = 0.0001;
	const unsigned max_trials = 10;
	const fl multiplier = 0.5;
	fl alpha = 1;
	const fl pg = scalar_product(p, g, n);
	for (unsigned trial = 0; trial < max_trials; trial++)
	{
		std::copy(x, x + conf_size, x_new);
		ops.increment(x_new, p, alpha);
		f1 = f(x_new, g_new);
		if (f1 - f0 < c0 * alpha * pg)
			break;
		alpha *= multiplier;
	}
	return alpha;
}
// bfgs.h:101-190
template<typename F, typename Ops>
fl accurate_line_search(const F& f, const Ops& ops, int n, const fl* x, const fl* g, fl f0,
		const fl* p, fl* x_new, fl* g_new, fl& f1, int conf_size)
{
	fl a, alpha, alpha2 = 0, alamin, b, disc, f2 = 0;
	fl rhs1, rhs2, slope = 0, sum = 0, temp, test, tmplam = 0;
	const fl ALF = 1.0e-4;
	const fl FIRST = 1.0;
	sum = scalar_product(p, p, n);
	sum = std::sqrt(sum);
	slope = scalar_product(g, p, n);
	if (slope >= 0)
	{
		std::copy(x, x + conf_size, x_new);
		std::fill(g_new, g_new + n, 0.0);
		return 0;
	}
	test = 0;
	for (int i = 0; i < n; i++)
	{
		temp = std::fabs(p[i]) / std::max(std::fabs(ops.value(x, i)), 1.0);
		if (temp > test)
			test = temp;
	}
	alamin = std::numeric_limits<fl>::epsilon() / test;
	alpha = FIRST;
	for (;;)
	{
		std::copy(x, x + conf_size, x_new);
		ops.increment(x_new, p, alpha);
		f1 = f(x_new, g_new);
		if (alpha < alamin)
		{
			std::copy(x, x + conf_size, x_new);
			std::fill(g_new, g_new + n, 0.0);
			return 0;
		}
		else if (f1 <= f0 + ALF * alpha * slope)
			return alpha;
		else
		{
			if (alpha == FIRST)
				tmplam = -slope / (2.0 * (f1 - f0 - slope));
			else
			{
				rhs1 = f1 - f0 - alpha * slope;
				rhs2 = f2 - f0 - alpha2 * slope;
				a = (rhs1 / (alpha * alpha) - rhs2 / (alpha2 * alpha2)) / (alpha - alpha2);
				b = (-alpha2 * rhs1 / (alpha * alpha) + alpha * rhs2 / (alpha2 * alpha2)) / (alpha - alpha2);
				if (a == 0.0)
					tmplam = -slope / (2.0 * b);
				else
				{
					disc = b * b - 3.0 * a * slope;
					if (disc < 0)
						tmplam = 0.5 * alpha;
					else if (b <= 0)
						tmplam = (-b + std::sqrt(disc)) / (3.0 * a);
					else
						tmplam = -slope / (b + std::sqrt(disc));
				}
				if (tmplam > .5 * alpha)
					tmplam = .5 * alpha;
			}
		}
		alpha2 = alpha;
		f2 = f1;
		alpha = std::max(tmplam, (fl)0.1 * alpha);
	}
}
// QuickVina 2's history of visited points (Alhossary et al., Bioinformatics
// 2015, github.com/QVina/qvina branch qvina2: visited.h, visited.cpp; as the
// GPU-style search restates it, src/gpu/gpu_mc.cuh:146-264).  Not smina's
// search.  A chain keeps the last 5n points its local optimizations visited
// (n degrees of freedom, conf in change indexing), each with its energy and
// the signs of its gradient, in a ring of 5n slots filled in order.
struct Visited
{
	int n, cap, size = 0, next = 0;
	bool full = false;
	std::vector<fl> x, f;          // x: [cap][n]
	std::vector<signed char> sign; // [cap][n]: -1, 0 (zero gradient) or +1; NaN: -1
	explicit Visited(int n_) : n(n_), cap(5 * n_), x(size_t(cap) * n_), f(cap), sign(size_t(cap) * n_) {}
	void add(const fl* xv, fl fv, const fl* g)
	{
		int slot;
		if (!full)
		{
			slot = size++;
			if (size >= cap)
			{
				full = true;
				next = 0;
			}
		}
		else
		{
			slot = next;
			next = (next + 1) % cap;
		}
		f[size_t(slot)] = fv;
		for (int i = 0; i < n; i++)
		{
			x[size_t(slot) * n + i] = xv[i];
			sign[size_t(slot) * n + i] = g[i] == 0 ? 0 : g[i] > 0 ? 1 : -1;
		}
	}
	// visited::interesting in QuickVina 2's own form: until the ring is full
	// every point is; otherwise the 4n stored points nearest to xv, in order
	// of (squared distance, slot), are checked and the first that passes the
	// first-order consistency check makes xv interesting.  The check: in
	// every coordinate where both gradients have the same nonzero sign,
	// (gradient > 0) must equal (xv > x_e) xor (fv > f_e).
	bool interesting(const fl* xv, fl fv, const fl* g) const
	{
		if (!full)
			return true;
		std::vector<std::pair<fl, int>> order;
		for (int e = 0; e < cap; e++)
		{
			fl d2 = 0;
			for (int i = 0; i < n; i++)
			{
				const fl dx = x[size_t(e) * n + i] - xv[i];
				d2 += dx * dx;
			}
			order.push_back({std::isnan(d2) ? max_fl : d2, e});
		}
		std::sort(order.begin(), order.end());
		for (int r = 0; r < 4 * n; r++)
		{
			const int e = order[size_t(r)].second;
			const bool newY = fv - f[size_t(e)] > 0;
			bool pass = true;
			for (int i = 0; i < n; i++)
			{
				const int se = sign[size_t(e) * n + i];
				if (se == 0 || g[i] == 0)
					continue;
				const bool nowPos = g[i] > 0;
				if (nowPos != (se > 0))
					continue;
				const bool newX = xv[i] - x[size_t(e) * n + i] > 0;
				if (nowPos ^ newX ^ newY)
					pass = false;
			}
			if (pass)
				return true;
		}
		return false;
	}
};

// bfgs.h:201-282; with a history (QuickVina 2, src/gpu/gpu_mc.cuh:440-515):
// an uninteresting start returns its energy unoptimized, and the start and
// every completed step join the history
template<typename F, typename Ops>
fl bfgs(const F& f, const Ops& ops, fl* x, fl* g, int maxiters, bool accurate, bool early_term,
		Visited* vis = nullptr)
{
	const int n = ops.change_size();
	const int cs = ops.conf_size();
	Hess h{n, std::vector<fl>(size_t(n) * (n + 1) / 2, 0)};
	for (int i = 0; i < n; i++)
		h.at(i, i) = 1;
	std::vector<fl> g_new(g, g + n), x_new(x, x + cs);
	fl f0 = f(x, g);
	std::vector<fl> xv(vis ? n : 0);
	auto values = [&]() {
		for (int i = 0; i < n; i++)
			xv[size_t(i)] = ops.value(x, i);
		return xv.data();
	};
	if (vis)
	{
		if (!vis->interesting(values(), f0, g))
			return f0;
		vis->add(xv.data(), f0, g);
	}
	fl f_orig = f0;
	std::vector<fl> g_orig(g, g + n), x_orig(x, x + cs);
	std::vector<fl> p(g, g + n), y(n);
	for (int step = 0; step < maxiters; step++)
	{
		minus_mat_vec_product(h, g, p.data());
		fl f1 = 0;
		fl alpha;
		if (accurate)
			alpha = accurate_line_search(f, ops, n, x, g, f0, p.data(), x_new.data(), g_new.data(), f1, cs);
		else
			alpha = fast_line_search(f, ops, n, x, g, f0, p.data(), x_new.data(), g_new.data(), f1, cs);
		if (alpha == 0)
			break;
		for (int i = 0; i < n; i++)
			y[i] = g_new[i] - g[i];
		fl prevf0 = f0;
		f0 = f1;
		std::copy(x_new.begin(), x_new.end(), x);
		if (early_term)
		{
			fl diff = prevf0 - f0;
			if (std::fabs(diff) < 1e-5)
				break;
		}
		std::copy(g_new.begin(), g_new.end(), g);
		fl gradnormsq = scalar_product(g, g, n);
		if (!(gradnormsq >= 1e-4))
			break;
		if (step == 0)
		{
			const fl yy = scalar_product(y.data(), y.data(), n);
			if (std::abs(yy) > epsilon_fl)
			{
				const fl diag = alpha * scalar_product(y.data(), p.data(), n) / yy;
				for (int i = 0; i < n; i++)
					h.at(i, i) = diag;
			}
		}
		bfgs_update(h, p.data(), y.data(), alpha);
		if (vis)
			vis->add(values(), f0, g);
	}
	if (!(f0 <= f_orig))
	{
		f0 = f_orig;
		std::copy(x_orig.begin(), x_orig.end(), x);
		std::copy(g_orig.begin(), g_orig.end(), g);
	}
	return f0;
}

// a plain vector space: x += alpha * p (for pinning BFGS on a quadratic)
struct FlatOps
{
	int n;
	int change_size() const { return n; }
	int conf_size() const { return n; }
	void increment(fl* x, const fl* c, fl factor) const
	{
		for (int i = 0; i < n; i++)
		{
			const fl d = factor * c[i];
			x[i] += d;
		}
	}
	fl value(const fl* x, int i) const { return x[i]; }
};
// f(x) = 1/2 x^T A x - b^T x, g = A x - b
struct Quadratic
{
	int n;
	const fl* A;
	const fl* b;
	int64_t* evals;
	fl operator()(const fl* x, fl* g) const
	{
		if (evals)
			++*evals;
		fl f = 0;
		for (int i = 0; i < n; i++)
		{
			fl ax = 0;
			for (int j = 0; j < n; j++)
				ax += A[i * n + j] * x[j];
			g[i] = ax - b[i];
			f += 0.5 * x[i] * ax - b[i] * x[i];
		}
		return f;
	}
};

// ---------------------------------------------------------------- MC
struct Output
{
	std::vector<fl> c;
	fl e;
	std::vector<vec> coords;
};
// coords.cpp:25-32
fl rmsd_upper_bound(const std::vector<vec>& a, const std::vector<vec>& b)
{
	fl acc = 0;
	for (size_t i = 0; i < a.size(); i++)
		acc += vec_distance_sqr(a[i], b[i]);
	return (a.size() > 0) ? std::sqrt(acc / a.size()) : 0;
}
void sort_container(std::vector<Output>& out)
{
	std::stable_sort(out.begin(), out.end(), [](const Output& a, const Output& b) { return a.e < b.e; });
}
// coords.cpp:34-58
void add_to_output_container(std::vector<Output>& out, const Output& t, fl min_rmsd, size_t max_size)
{
	size_t best_i = out.size();
	fl best = max_fl;
	for (size_t i = 0; i < out.size(); i++)
	{
		fl res = rmsd_upper_bound(t.coords, out[i].coords);
		if (i == 0 || res < best)
		{
			best_i = i;
			best = res;
		}
	}
	if (best_i < out.size() && best < min_rmsd)
	{
		if (t.e < out[best_i].e)
			out[best_i] = t;
	}
	else
	{
		if (out.size() < max_size)
			out.push_back(t);
		else if (!out.empty() && t.e < out.back().e)
			out.back() = t;
	}
	sort_container(out);
}
// mutate.cpp:35-63
void mutate_conf(fl* c, const Model& m, fl amplitude, rng& gen)
{
	int count = 0;
	for (const Tree& tr : m.trees)
		count += tr.ligand ? 2 + tr.ntors : tr.ntors;
	if (count == 0)
		return;
	int which = random_int(0, count - 1, gen);
	for (int l = 0; l < int(m.trees.size()); l++)
	{
		const Tree& tr = m.trees[l];
		fl* p = c + tr.conf_off;
		if (!tr.ligand)
			continue;
		if (which == 0)
		{
			vec d = amplitude * random_inside_sphere(gen);
			p[0] += d[0];
			p[1] += d[1];
			p[2] += d[2];
			return;
		}
		--which;
		if (which == 0)
		{
			fl gr = m.gyration_radius(l);
			if (gr > epsilon_fl)
			{
				vec rotation = amplitude / gr * random_inside_sphere(gen);
				qt o = {p[3], p[4], p[5], p[6]};
				quaternion_increment(o, rotation);
				p[3] = o.a;
				p[4] = o.b;
				p[5] = o.c;
				p[6] = o.d;
			}
			return;
		}
		--which;
		if (which < tr.ntors)
		{
			p[7 + which] = random_fl(-pi, pi, gen);
			return;
		}
		which -= tr.ntors;
	}
	for (const Tree& tr : m.trees)
	{
		if (tr.ligand)
			continue;
		if (which < tr.ntors)
		{
			c[tr.conf_off + which] = random_fl(-pi, pi, gen);
			return;
		}
		which -= tr.ntors;
	}
}
// monte_carlo.cpp:37-41
bool metropolis_accept(fl old_f, fl new_f, fl temperature, rng& gen)
{
	if (new_f < old_f)
		return true;
	const fl acceptance_probability = pmath::exp((old_f - new_f) / temperature);
	return random_fl(0, 1, gen) < acceptance_probability;
}
// conf::randomize (src/lib/conf.h:362-367)
void randomize_conf(const Model& m, fl* c, const vec& c1, const vec& c2, rng& gen)
{
	for (const Tree& tr : m.trees)
	{
		fl* p = c + tr.conf_off;
		if (tr.ligand)
		{
			vec pos = random_in_box(c1, c2, gen);
			qt o = random_orientation(gen);
			p[0] = pos[0];
			p[1] = pos[1];
			p[2] = pos[2];
			p[3] = o.a;
			p[4] = o.b;
			p[5] = o.c;
			p[6] = o.d;
			for (int i = 0; i < tr.ntors; i++)
				p[7 + i] = random_fl(-pi, pi, gen);
		}
	}
	for (const Tree& tr : m.trees)
		if (!tr.ligand)
			for (int i = 0; i < tr.ntors; i++)
				c[tr.conf_off + i] = random_fl(-pi, pi, gen);
}

// monte_carlo::operator() (src/lib/monte_carlo.cpp:95-130) from a start conf
void mc_chain(const Ctx& ctx, Model& m, const sgpu_mc_params& p, int num_steps, int maxiters,
		std::vector<fl> start, rng& gen, std::vector<Output>& out, int64_t* evals)
{
	const fl authentic_v[3] = {p.authentic_v[0], p.authentic_v[1], p.authentic_v[2]};
	const int n = m.change_size;
	ConfOps ops{&m};
	std::vector<fl> g(n, 0);
	Output tmp{start, 0, {}};
	fl best_e = max_fl;
	// QuickVina 2's history: models of at most 64 degrees of freedom
	Visited history(n);
	Visited* vis = p.visited && n <= 64 ? &history : nullptr;
	Objective hunt{&ctx, &m, 0, p.cache_slope, {p.hunt_cap[0], p.hunt_cap[1], p.hunt_cap[2]}, evals};
	Objective auth{&ctx, &m, 0, p.cache_slope, {authentic_v[0], authentic_v[1], authentic_v[2]}, evals};
	for (int step = 0; step < num_steps; step++)
	{
		Output candidate = tmp;
		mutate_conf(candidate.c.data(), m, p.mutation_amplitude, gen);
		candidate.e = bfgs(hunt, ops, candidate.c.data(), g.data(), maxiters, p.accurate_line_search,
				p.early_term, vis);
		if (step == 0 || metropolis_accept(tmp.e, candidate.e, p.temperature, gen))
		{
			tmp = candidate;
			m.set(tmp.c.data());
			if (tmp.e < best_e || out.size() < size_t(p.num_saved_mins))
			{
				tmp.e = bfgs(auth, ops, tmp.c.data(), g.data(), maxiters, p.accurate_line_search, p.early_term,
						vis);
				m.set(tmp.c.data());
				tmp.coords = m.heavy_coords();
				add_to_output_container(out, tmp, p.min_rmsd, size_t(p.num_saved_mins));
				if (tmp.e < best_e)
					best_e = tmp.e;
			}
		}
	}
}

// main.cpp:165-191 with non_cache::within (src/lib/non_cache.cpp:93-107)
bool within(const Ctx& ctx, const Model& m, fl margin = 0.0001)
{
	for (int i = 0; i < m.N; i++)
	{
		if (is_hydrogen_type(m.type[i]))
			continue;
		for (int j = 0; j < 3; j++)
			if (ctx.gd.n[j] > 0)
				if (m.coords[i][j] < ctx.gd.begin[j] - margin || m.coords[i][j] > ctx.gd.end[j] + margin)
					return false;
	}
	return true;
}
fl refine_structure(const Ctx& ctx, Model& m, Output& out, const fl* cap, int maxiters, bool accurate,
		bool early_term, int64_t* evals)
{
	ConfOps ops{&m};
	std::vector<fl> g(m.change_size, 0);
	fl slope = 10;
	for (int p = 0; p < 5; p++)
	{
		Objective f{&ctx, &m, 1, slope, {cap[0], cap[1], cap[2]}, evals};
		out.e = bfgs(f, ops, out.c.data(), g.data(), maxiters, accurate, early_term);
		m.set(out.c.data());
		if (within(ctx, m))
			break;
		slope *= 10;
	}
	out.coords = m.heavy_coords();
	if (!within(ctx, m))
		out.e = max_fl;
	return out.e;
}

} // namespace orc

using namespace orc;

struct orc_ctx
{
	Ctx c;
};

extern "C" {

orc_ctx* orc_create(const sgpu_scoring* sf, const sgpu_receptor* rec, const sgpu_grid_dims* gd)
{
	orc_ctx* o = new orc_ctx;
	Ctx& c = o->c;
	c.sf.terms.assign(sf->terms, sf->terms + sf->nterms);
	c.sf.ci.assign(sf->ci, sf->ci + sf->nci);
	c.sf.atom_params = sf->atom_params;
	c.tab.build(c.sf, 32);
	for (int i = 0; i < 3; i++)
	{
		c.gd.begin[i] = gd->begin[i];
		c.gd.end[i] = gd->end[i];
		c.gd.n[i] = gd->n[i];
	}
	c.geom.setup(c.gd);
	c.rec.resize(rec->natoms);
	for (int i = 0; i < rec->natoms; i++)
		c.rec[i] = {vec(rec->xyz[3 * i], rec->xyz[3 * i + 1], rec->xyz[3 * i + 2]), rec->type[i],
				rec->charge ? rec->charge[i] : 0.0};
	c.grids.resize(NUM_TYPES);
	c.cgrids.resize(NUM_TYPES);
	c.have_grid.assign(NUM_TYPES, false);
	c.setup_noncache();
	return o;
}
// --user_grid: values as grid::init stores them ((n+1)^3, x fastest); the
// maps built afterwards carry it
void orc_set_user_grid(orc_ctx* c, const sgpu_grid_dims* gd, const double* data, double map_slope)
{
	Ctx& x = c->c;
	if (!gd)
	{
		x.has_ug = false;
		return;
	}
	GridDims g;
	for (int i = 0; i < 3; i++)
	{
		g.begin[i] = gd->begin[i];
		g.end[i] = gd->end[i];
		g.n[i] = gd->n[i];
	}
	x.ug.setup(g);
	x.ug.data.assign(data, data + size_t(x.ug.dim[0]) * x.ug.dim[1] * x.ug.dim[2]);
	x.ug_map_slope = map_slope;
	x.has_ug = true;
}
void orc_destroy(orc_ctx* c) { delete c; }
const char* orc_error(orc_ctx* c) { return c->c.err.c_str(); }
int orc_table_n(orc_ctx* c) { return c->c.tab.n; }
int orc_use_splines(orc_ctx* c, double factor, double* coef, double* cut)
{
	if (!(factor > 0))
	{
		c->c.err = "spline factor must be positive";
		return -1;
	}
	Tables& t = c->c.tab;
	if (c->c.sf.charged())
	{
		c->c.err = "the spline tables carry one component: no charge-dependent terms";
		return -1;
	}
	t.build_splines(c->c.sf, factor);
	if (coef)
		std::copy(t.spl.begin(), t.spl.end(), coef);
	if (cut)
		std::copy(t.spl_cut.begin(), t.spl_cut.end(), cut);
	// the maps built so far used the linear tables
	std::fill(c->c.have_grid.begin(), c->c.have_grid.end(), false);
	return t.nseg;
}
void orc_tables(orc_ctx* c, double* smooth, double* fast)
{
	const Tables& t = c->c.tab;
	const size_t total = t.e.size();
	for (size_t k = 0; k < total; k++)
	{
		if (smooth)
		{
			smooth[2 * k] = t.e[k];
			smooth[2 * k + 1] = t.dor[k];
		}
		if (fast)
			fast[k] = t.fast[k];
	}
}
double orc_eval_terms(orc_ctx* c, int t1, int t2, double r) { return c->c.sf.eval_fast(t1, t2, r); }
void orc_eval_components(orc_ctx* c, int t1, int t2, double r, double* out4) { c->c.sf.eval_fast4(t1, t2, r, out4); }
double orc_conf_independent(orc_ctx* c, double num_tors, double e) { return c->c.sf.conf_independent(num_tors, e); }
int orc_build_grids(orc_ctx* c, const int* types, int ntypes)
{
	try
	{
		c->c.populate(std::vector<int>(types, types + ntypes));
	}
	catch (std::exception& e)
	{
		c->c.err = e.what();
		return 1;
	}
	return 0;
}
int orc_table_ncomp(orc_ctx* c) { return c->c.tab.ncomp; }
// the three charge-dependent components: qsmooth [tri][n][3][2], qfast [tri][n][3]
void orc_qtables(orc_ctx* c, double* qsmooth, double* qfast)
{
	const Tables& t = c->c.tab;
	for (size_t k = 0; k < t.eq.size(); k++)
	{
		if (qsmooth)
		{
			qsmooth[2 * k] = t.eq[k];
			qsmooth[2 * k + 1] = t.dorq[k];
		}
		if (qfast)
			qfast[k] = t.fastq[k];
	}
}
int orc_charge_grid_values(orc_ctx* c, int type, double* out)
{
	if (!c->c.have_grid[type] || c->c.tab.ncomp != 4)
		return 1;
	std::copy(c->c.cgrids[type].data.begin(), c->c.cgrids[type].data.end(), out);
	return 0;
}
int orc_grid_values(orc_ctx* c, int type, double* out)
{
	if (!c->c.have_grid[type])
		return 1;
	std::copy(c->c.grids[type].data.begin(), c->c.grids[type].data.end(), out);
	return 0;
}
int orc_add_model(orc_ctx* c, const sgpu_model_desc* d)
{
	try
	{
		std::unique_ptr<Model> m(new Model);
		m->build(*d);
		c->c.models.push_back(std::move(m));
		return int(c->c.models.size()) - 1;
	}
	catch (std::exception& e)
	{
		c->c.err = e.what();
		return -1;
	}
}
void orc_model_info(orc_ctx* c, int model, int* info)
{
	const Model& m = *c->c.models[model];
	size_t np = m.other_pairs.size();
	for (auto& lp : m.lig_pairs)
		np += lp.size();
	info[0] = m.N;
	info[1] = m.natoms;
	info[2] = int(m.frames.size());
	info[3] = int(np);
	info[4] = int(m.other_pairs.size());
	info[5] = m.conf_size;
	info[6] = m.change_size;
	info[7] = m.heavy_count();
	info[8] = m.nligands;
	info[9] = int(m.trees.size());
}
void orc_initial_conf(orc_ctx* c, int model, double* conf)
{ // model::get_initial_conf (src/lib/model.cpp:650-659)
	const Model& m = *c->c.models[model];
	std::fill(conf, conf + m.conf_size, 0.0);
	for (const Tree& tr : m.trees)
		if (tr.ligand)
		{
			double* p = conf + tr.conf_off;
			const vec& o = m.frames[tr.root].origin;
			p[0] = o[0];
			p[1] = o[1];
			p[2] = o[2];
			p[3] = 1;
		}
}
void orc_model_pairs(orc_ctx* c, int model, int* a, int* b, int* list)
{
	const Model& m = *c->c.models[model];
	int k = 0;
	for (const Pair& p : m.other_pairs)
	{
		a[k] = p.a;
		b[k] = p.b;
		list[k++] = 2;
	}
	for (auto& lp : m.lig_pairs)
		for (const Pair& p : lp)
		{
			a[k] = p.a;
			b[k] = p.b;
			list[k++] = 0;
		}
}
double orc_num_tors(orc_ctx* c, int model) { return c->c.models[model]->num_tors; }
double orc_eval_deriv(orc_ctx* c, int model, int backend, double slope, const double* v, const double* conf,
		double* g, double* coords)
{
	Model& m = *c->c.models[model];
	try
	{
		double e = c->c.eval_deriv(m, backend, slope, v, conf, g);
		if (coords)
			for (int i = 0; i < m.N; i++)
				for (int k = 0; k < 3; k++)
					coords[3 * i + k] = m.coords[i][k];
		return e;
	}
	catch (std::exception& e)
	{
		c->c.err = e.what();
		return std::numeric_limits<double>::quiet_NaN();
	}
}
double orc_eval(orc_ctx* c, int model, double slope, const double* v, const double* conf)
{
	return c->c.eval(*c->c.models[model], slope, v, conf);
}
double orc_eval_intramolecular(orc_ctx* c, int model, const double* v, const double* conf)
{
	return c->c.eval_intramolecular(*c->c.models[model], v, conf);
}
double orc_bfgs(orc_ctx* c, int model, int backend, double slope, const double* v, int maxiters, int accurate,
		int early_term, double* conf, int64_t* evals)
{
	Model& m = *c->c.models[model];
	ConfOps ops{&m};
	Objective f{&c->c, &m, backend, slope, {v[0], v[1], v[2]}, evals};
	std::vector<double> g(m.change_size, 0);
	return bfgs(f, ops, conf, g.data(), maxiters, accurate != 0, early_term != 0);
}
void orc_randomize(orc_ctx* c, int model, uint32_t seed, double* conf, uint32_t* rng_state)
{
	const Model& m = *c->c.models[model];
	rng gen(seed);
	std::fill(conf, conf + m.conf_size, 0.0);
	vec c1(c->c.gd.begin[0], c->c.gd.begin[1], c->c.gd.begin[2]);
	vec c2(c->c.gd.end[0], c->c.gd.end[1], c->c.gd.end[2]);
	randomize_conf(m, conf, c1, c2, gen);
	if (rng_state)
	{
		// libstdc++ writes the 624 words in order followed by the position
		std::ostringstream os;
		os << gen.g;
		std::istringstream is(os.str());
		for (int k = 0; k < 625; k++)
			is >> rng_state[k];
	}
}
} // extern "C"

namespace
{
rng rng_from_state(const uint32_t* s)
{
	rng gen(0);
	std::ostringstream os;
	for (int k = 0; k < 625; k++)
		os << s[k] << (k + 1 < 625 ? " " : "");
	std::istringstream is(os.str());
	is >> gen.g;
	return gen;
}
void write_out(const Model& m, const std::vector<Output>& out, int max_out, double* e, double* conf,
		double* coords)
{
	int k = 0;
	for (const Output& o : out)
	{
		if (k >= max_out)
			break;
		if (e)
			e[k] = o.e;
		if (conf)
			std::copy(o.c.begin(), o.c.end(), conf + size_t(k) * m.conf_size);
		if (coords)
			for (size_t a = 0; a < o.coords.size(); a++)
				for (int d = 0; d < 3; d++)
					coords[size_t(k) * 3 * m.heavy_count() + 3 * a + d] = o.coords[a][d];
		++k;
	}
}
} // namespace

extern "C" {

int orc_mc_chain(orc_ctx* c, int model, const sgpu_mc_params* p, int num_steps, int maxiters,
		const double* conf0, const uint32_t* rng_state, int max_out, double* e, double* conf, double* coords,
		int64_t* evals)
{
	Model& m = *c->c.models[model];
	rng gen = rng_from_state(rng_state);
	std::vector<Output> out;
	try
	{
		// each task starts from the parsed model's state (parallel_mc_task copies it)
		m.reset();
		mc_chain(c->c, m, *p, num_steps, maxiters, std::vector<double>(conf0, conf0 + m.conf_size), gen, out,
				evals);
	}
	catch (std::exception& ex)
	{
		c->c.err = ex.what();
		return -1;
	}
	write_out(m, out, max_out, e, conf, coords);
	return int(out.size());
}

// add_to_output_container on a caller-held container of n poses (energies
// e[n], coordinates coords[n][3 * natoms], capacity max_size): adds the pose
// (te, tcoords) and returns the new count; conf slots carry each pose's
// original index (tag) so the caller sees which pose went where
int orc_container_add(int n, int natoms, int max_size, double min_rmsd, double* e, double* coords, int* tag,
		double te, const double* tcoords, int ttag)
{
	std::vector<Output> out;
	for (int i = 0; i < n; i++)
	{
		Output o{{double(tag[i])}, e[i], {}};
		for (int a = 0; a < natoms; a++)
			o.coords.push_back(vec{coords[(size_t(i) * natoms + a) * 3], coords[(size_t(i) * natoms + a) * 3 + 1],
					coords[(size_t(i) * natoms + a) * 3 + 2]});
		out.push_back(o);
	}
	Output t{{double(ttag)}, te, {}};
	for (int a = 0; a < natoms; a++)
		t.coords.push_back(vec{tcoords[3 * a], tcoords[3 * a + 1], tcoords[3 * a + 2]});
	add_to_output_container(out, t, min_rmsd, size_t(max_size));
	for (size_t i = 0; i < out.size(); i++)
	{
		e[i] = out[i].e;
		tag[i] = int(out[i].c[0]);
		for (int a = 0; a < natoms; a++)
			for (int d = 0; d < 3; d++)
				coords[(i * natoms + a) * 3 + d] = out[i].coords[size_t(a)][d];
	}
	return int(out.size());
}
}

extern "C" {

double orc_bfgs_quadratic(int n, const double* A, const double* b, double* x, int maxiters, int accurate,
		int64_t* evals)
{
	Quadratic q{n, A, b, evals};
	FlatOps ops{n};
	std::vector<double> g(n, 0);
	return bfgs(q, ops, x, g.data(), maxiters, accurate != 0, false);
}

int orc_visited_check(int n, int count, const double* xs, const double* fs, const double* gs, const double* xv,
		double fv, const double* g)
{
	Visited v(n);
	for (int k = 0; k < count; k++)
		v.add(xs + size_t(k) * n, fs[k], gs + size_t(k) * n);
	return v.interesting(xv, fv, g) ? 1 : 0;
}

double orc_refine(orc_ctx* c, int model, const double* v, int maxiters, double* conf, double* coords)
{
	return orc_refine_opts(c, model, v, maxiters, 0, conf, coords);
}

double orc_refine_opts(orc_ctx* c, int model, const double* v, int maxiters, int accurate, double* conf,
		double* coords)
{
	Model& m = *c->c.models[model];
	Output o{std::vector<double>(conf, conf + m.conf_size), 0, {}};
	try
	{
		refine_structure(c->c, m, o, v, maxiters, accurate != 0, false, nullptr);
	}
	catch (std::exception& ex)
	{
		c->c.err = ex.what();
		return std::numeric_limits<double>::quiet_NaN();
	}
	std::copy(o.c.begin(), o.c.end(), conf);
	if (coords)
		for (size_t a = 0; a < o.coords.size(); a++)
			for (int d = 0; d < 3; d++)
				coords[3 * a + d] = o.coords[a][d];
	return o.e;
}

int orc_dock(orc_ctx* c, int model, const sgpu_mc_params* p, int seed, int exhaustiveness, int num_modes,
		double energy_range, int threads, double step_fraction, int split, int max_out, double* e, double* conf,
		int64_t* evals, double force_cap, double out_min_rmsd, int minimize_iters)
{
	const double cap[3] = {force_cap > 0 ? force_cap : p->authentic_v[0], force_cap > 0 ? force_cap : p->authentic_v[1],
			force_cap > 0 ? force_cap : p->authentic_v[2]};
	if (split < 1)
	{
		c->c.err = "split must be at least 1";
		return -1;
	}
	const Model& proto = *c->c.models[model];
	// setup_search (src/main/main.cpp:513-529)
	const int N = proto.N;
	int dof = 6 * proto.nligands;
	for (const Tree& tr : proto.trees)
		dof += tr.ntors;
	const int heuristic = N + 10 * dof;
	const int full_steps = int(unsigned(70 * 3 * (50 + heuristic) / 2));
	// a shortened or split search (the GPU-style options; add_chains and
	// chain_steps, src/gpu/gpu_docking.cpp:200-220): exhaustiveness x split
	// chains seeded in order, each with max(1, ceil(steps x fraction /
	// split)) steps
	const int num_steps = step_fraction >= 1 && split == 1
	                              ? full_steps
	                              : std::max(1, int(std::ceil(double(full_steps) * step_fraction / double(split))));
	const int maxiters = minimize_iters > 0 ? minimize_iters : (25 + N) / 3;
	// parallel_mc::operator() (src/lib/parallel_mc.cpp:80-98)
	const int tasks = exhaustiveness * split;
	rng generator{uint32_t(seed)};
	std::vector<int> seeds(tasks);
	for (int i = 0; i < tasks; i++)
		seeds[i] = random_int(0, 1000000, generator);
	std::vector<std::vector<Output>> task_out(tasks);
	std::vector<int64_t> task_evals(tasks, 0);
	std::vector<std::string> errs(tasks);
	vec c1(c->c.gd.begin[0], c->c.gd.begin[1], c->c.gd.begin[2]);
	vec c2(c->c.gd.end[0], c->c.gd.end[1], c->c.gd.end[2]);
	int steps_now = num_steps;
	auto run_task = [&](int t) {
		try
		{
			Model m = proto; // parallel_mc_task copies the model
			m.reset();
			rng gen(uint32_t(seeds[t]));
			std::vector<double> start(m.conf_size, 0.0);
			randomize_conf(m, start.data(), c1, c2, gen);
			task_out[t].clear();
			task_evals[t] = 0;
			mc_chain(c->c, m, *p, steps_now, maxiters, start, gen, task_out[t], &task_evals[t]);
		}
		catch (std::exception& ex)
		{
			errs[t] = ex.what();
		}
	};
	auto run_all = [&](int steps) {
		steps_now = steps;
		if (threads <= 1)
			for (int t = 0; t < tasks; t++)
				run_task(t);
		else
		{
			std::vector<std::thread> pool;
			std::atomic<int> next(0);
			for (int w = 0; w < threads; w++)
				pool.emplace_back([&]() {
					for (int t = next++; t < tasks; t = next++)
						run_task(t);
				});
			for (auto& th : pool)
				th.join();
		}
		for (auto& s : errs)
			if (!s.empty())
				throw std::runtime_error(s);
	};
	try
	{
		// early stopping (the GPU-style --gpu_early_stop; EarlyStop and
		// stop_agreeing, src/gpu/gpu_kernels_impl.cuh:607-657, :742-808):
		// at checkpoint k of n every chain has run ceil(steps x k / n) steps
		// (a chain cut there is the prefix of the longer one); the ligand
		// stops at the first checkpoint from `first` on, before the last,
		// where at least max(2, ceil(fraction x units)) of its units
		// (`split` consecutive chains each) have best energies within eps of
		// the lowest
		int stop_at = 0;
		const int n = p->early_stop_checkpoints;
		if (n > 0)
		{
			const int first = std::min(n, std::max(1, p->early_stop_first));
			for (int k = first; k < n && !stop_at; k++)
			{
				run_all(int((int64_t(num_steps) * k + n - 1) / n));
				std::vector<fl> ub(size_t(exhaustiveness), max_fl);
				for (int t = 0; t < tasks; t++)
					for (const Output& o : task_out[t])
						ub[size_t(t / split)] = std::min(ub[size_t(t / split)], o.e);
				if (ub.size() < 2)
					continue;
				const fl best = *std::min_element(ub.begin(), ub.end());
				int agree = 0;
				for (fl u : ub)
					agree += (u - best <= p->early_stop_eps);
				if (agree >= std::max(2, int(std::ceil(p->early_stop_fraction * double(ub.size()) - 1e-9))))
					stop_at = k;
			}
		}
		if (!stop_at)
			run_all(num_steps);
	}
	catch (std::exception& ex)
	{
		c->c.err = ex.what();
		return -1;
	}
	// merge_output_containers with min_rmsd 2 (src/lib/parallel_mc.cpp:71-78)
	std::vector<Output> out;
	for (int t = 0; t < tasks; t++)
		for (const Output& o : task_out[t])
			add_to_output_container(out, o, 2, size_t(p->num_saved_mins));
	sort_container(out);
	int64_t total = 0;
	for (int64_t k : task_evals)
		total += k;
	Model m = proto;
	m.reset();
	try
	{
		// refine_structure of every pose (do_search, src/main/main.cpp:478-482)
		for (Output& o : out)
			refine_structure(c->c, m, o, cap, maxiters, p->accurate_line_search != 0,
					p->early_term != 0, &total);
		// finish_docking (src/main/main.cpp:256-324)
		const double slope = p->cache_slope;
		if (!out.empty())
		{
			sort_container(out);
			const double intra = c->c.eval_intramolecular(m, cap, out[0].c.data());
			for (Output& o : out)
				if (not_max(o.e))
				{
					double ev = c->c.eval(m, slope, cap, o.c.data());
					o.e = c->c.sf.conf_independent(m.num_tors, ev - intra);
				}
			sort_container(out);
		}
		// remove_redundant (src/main/main.cpp:204-210), out_min_rmsd 1
		std::vector<Output> tmp;
		for (const Output& o : out)
			add_to_output_container(tmp, o, out_min_rmsd > 0 ? out_min_rmsd : 1.0, out.size());
		out.swap(tmp);
	}
	catch (std::exception& ex)
	{
		c->c.err = ex.what();
		return -1;
	}
	int k = 0;
	for (size_t i = 0; i < out.size() && k < max_out; i++)
	{
		if (k >= num_modes || !not_max(out[i].e) || out[i].e > out[0].e + energy_range)
			break;
		e[k] = out[i].e;
		if (conf)
			std::copy(out[i].c.begin(), out[i].c.end(), conf + size_t(k) * m.conf_size);
		++k;
	}
	if (evals)
		*evals = total;
	return k;
}
}

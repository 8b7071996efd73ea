/*
 * Host side of the docking path: scoring tables, model flattening, chain
 * seeding, the GpuDocking driver and its C wrappers (include/sgpu_host.h).
 * Compiled with -ffp-contract=off: table values, local coordinates and
 * initial confs must be the exact doubles smina computes.
 */
#include "gpu_docking.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>

#include "../csrc/atom_types.h"
#include "../csrc/rng.h"
#include "mol.h"

namespace sgpu_host
{

using namespace sgpu;

namespace
{
const double kPi = 3.1415926535897931;
const double kEps = std::numeric_limits<double>::epsilon();
const double kMax = std::numeric_limits<double>::max();

inline double sq(double x) { return x * x; }

// ---- terms (src/lib/everything.h:156-541), charge independent, on the
// scoring function's atom parameter set
double optimal_distance(int t1, int t2, int ap)
{
	return atom_type_info(t1, ap).xs_radius + atom_type_info(t2, ap).xs_radius;
}
double slope_step(double bad, double good, double x)
{
	if (bad < good)
	{
		if (x <= bad)
			return 0;
		if (x >= good)
			return 1;
	}
	else
	{
		if (x >= bad)
			return 0;
		if (x <= good)
			return 1;
	}
	return (x - bad) / (good - bad);
}
bool h_bond_possible(int t1, int t2, int ap)
{
	const AtomTypeInfo a = atom_type_info(t1, ap), b = atom_type_info(t2, ap);
	return (a.donor && b.acceptor) || (b.donor && a.acceptor);
}
bool anti_h_bond(int t1, int t2, int ap)
{
	const AtomTypeInfo a = atom_type_info(t1, ap), b = atom_type_info(t2, ap);
	if (a.donor && !a.acceptor)
		return b.donor && !b.acceptor;
	if (!a.donor && a.acceptor)
		return !b.donor && b.acceptor;
	return false;
}
double ipow(double x, int n)
{
	double r = 1;
	for (int i = 0; i < n; i++)
		r *= x;
	return r;
}
double quadratic_below(double r, double d0)
{
	const double d = r - d0;
	return d > 0 ? 0.0 : d * d;
}
double lj(double r, double d0, double depth, int i, int j, double cap)
{
	const double ci = ipow(d0, i) * depth * j / (double(i) - double(j));
	const double cj = ipow(d0, j) * depth * i / (double(j) - double(i));
	const double ri = ipow(r, i), rj = ipow(r, j);
	if (ri > kEps && rj > kEps)
		return std::min(cap, ci / ri + cj / rj);
	return cap;
}
double term(const sgpu_term& t, int t1, int t2, double r, int ap)
{
	const AtomTypeInfo a = atom_type_info(t1, ap), b = atom_type_info(t2, ap);
	switch (t.kind)
	{
	case SGPU_TERM_GAUSS:
	{
		const double x = r - (optimal_distance(t1, t2, ap) + t.p[0]);
		return std::exp(-sq(x / t.p[1]));
	}
	case SGPU_TERM_REPULSION:
		return quadratic_below(r, optimal_distance(t1, t2, ap) + t.p[0]);
	case SGPU_TERM_HYDROPHOBIC:
		return (a.hydrophobe && b.hydrophobe) ? slope_step(t.p[1], t.p[0], r - optimal_distance(t1, t2, ap)) : 0.0;
	case SGPU_TERM_NON_HYDROPHOBIC:
		return (!a.hydrophobe && !b.hydrophobe) ? slope_step(t.p[1], t.p[0], r - optimal_distance(t1, t2, ap)) : 0.0;
	case SGPU_TERM_NON_DIR_H_BOND:
		return h_bond_possible(t1, t2, ap) ? slope_step(t.p[1], t.p[0], r - optimal_distance(t1, t2, ap)) : 0.0;
	case SGPU_TERM_NON_DIR_ANTI_H_BOND_QUADRATIC:
		return anti_h_bond(t1, t2, ap) ? quadratic_below(r, optimal_distance(t1, t2, ap) + t.p[0]) : 0.0;
	case SGPU_TERM_DONOR_DONOR_QUADRATIC:
		return (a.donor && b.donor) ? quadratic_below(r, optimal_distance(t1, t2, ap) + t.p[0]) : 0.0;
	case SGPU_TERM_ACCEPTOR_ACCEPTOR_QUADRATIC:
		return (a.acceptor && b.acceptor) ? quadratic_below(r, optimal_distance(t1, t2, ap) + t.p[0]) : 0.0;
	case SGPU_TERM_VDW:
	{
		const double d0 = optimal_distance(t1, t2, ap), s = t.p[2];
		double rr = r;
		if (rr > d0 + s)
			rr -= s;
		else if (rr < d0 - s)
			rr += s;
		else
			rr = d0;
		return lj(rr, d0, 1, int(t.p[0]), int(t.p[1]), t.p[3]);
	}
	case SGPU_TERM_NON_DIR_H_BOND_LJ:
		return h_bond_possible(t1, t2, ap) ? lj(r, optimal_distance(t1, t2, ap) + t.p[0], 5, 10, 12, t.p[1]) : 0.0;
	}
	throw std::runtime_error("unknown scoring term");
}

} // namespace

bool charge_dependent(const sgpu_term& t)
{
	return t.kind == SGPU_TERM_ELECTROSTATIC || t.kind == SGPU_TERM_AD4_SOLVATION;
}

namespace
{

// the four result components of a charge-dependent term (type only, |q_a|,
// |q_b| and q_a q_b dependent): electrostatic and ad4_solvation's
// eval_components (src/lib/everything.h:64-156)
void components(const sgpu_term& t, int t1, int t2, double r, double* out)
{
	out[0] = out[1] = out[2] = out[3] = 0;
	if (t.kind == SGPU_TERM_ELECTROSTATIC)
	{
		const double p = ipow(r, int(t.p[0]));
		out[3] = p < kEps ? t.p[1] : std::min(t.p[1], 1 / p);
		return;
	}
	if (t.kind != SGPU_TERM_AD4_SOLVATION)
		throw std::runtime_error("not a charge-dependent term");
	const double df = std::exp(-sq(r / (2 * t.p[0])));
	out[0] += ad_solvation(t1) * ad_volume(t2) * df;
	out[1] += t.p[1] * ad_volume(t2) * df;
	out[0] += ad_solvation(t2) * ad_volume(t1) * df;
	out[2] += t.p[1] * ad_volume(t1) * df;
}

// smooth_div (src/lib/everything.cpp)
double smooth_div(double x, double y)
{
	if (std::abs(x) < kEps)
		return 0;
	if (std::abs(y) < kEps)
		return (x * y > 0) ? kMax : -kMax;
	return x / y;
}

// weighted_terms::conf_independent (src/lib/weighted_terms.cpp:74-80): the
// conf-independent terms in order (src/lib/everything.h:730-856, :980-1002)
double conf_independent(const sgpu_scoring& sf, double num_tors, double x)
{
	for (int k = 0; k < sf.nci; k++)
	{
		const double w = sf.ci[k].weight;
		switch (sf.ci[k].kind)
		{
		case SGPU_CI_NUM_TORS_DIV:
			x = smooth_div(x, 1 + 0.1 * (w + 1) * num_tors / 5.0);
			break;
		case SGPU_CI_NUM_TORS_ADD:
			x = x + w * num_tors;
			break;
		case SGPU_CI_NUM_TORS_SQR:
			x = x + 0.1 * w * sq(num_tors) / 5;
			break;
		case SGPU_CI_NUM_TORS_SQRT:
			x = x + 0.1 * w * std::sqrt(num_tors) / std::sqrt(5.0);
			break;
		case SGPU_CI_NUM_TORS_DIV_SIMPLE:
			x = smooth_div(x, 1 + w * num_tors);
			break;
		case SGPU_CI_CONSTANT:
			x = x + w;
			break;
		default:
			throw std::runtime_error("unknown conf-independent term");
		}
	}
	return x;
}

// ---- chain initialisation randoms, on the product's MT19937
struct Gen
{
	MTState s;
	explicit Gen(uint32_t seed) { mt_seed(s, seed); }
	uint32_t next() { return mt_next(s); }
};
// random_normal: Marsaglia's polar method, one fresh pair per call
double normal(Gen& g)
{
	for (;;)
	{
		const double u = random_fl(-1.0, 1.0, g);
		const double v = random_fl(-1.0, 1.0, g);
		const double s = u * u + v * v;
		if (s < 1 && s > 0)
			return 0 + 1 * (u * std::sqrt(-2 * std::log(s) / s));
	}
}
} // namespace

double term_value(const sgpu_term& t, int t1, int t2, double r, int atom_params, double q1, double q2)
{
	if (!charge_dependent(t))
		return term(t, t1, t2, r, atom_params);
	// charge_dependent::eval (src/lib/terms.h:90-94): the components of
	// (t1, t2) as given, then result_components::eval
	double c[4];
	components(t, t1, t2, r, c);
	return c[0] + std::abs(q1) * c[1] + std::abs(q2) * c[2] + q1 * q2 * c[3];
}

// each term's cutoff: the last parameter its name carries (c)
double term_cutoff(const sgpu_term& t)
{
	switch (t.kind)
	{
	case SGPU_TERM_REPULSION:
	case SGPU_TERM_NON_DIR_ANTI_H_BOND_QUADRATIC:
	case SGPU_TERM_DONOR_DONOR_QUADRATIC:
	case SGPU_TERM_ACCEPTOR_ACCEPTOR_QUADRATIC:
		return t.p[1];
	case SGPU_TERM_VDW:
		return t.p[4];
	default:
		return t.p[2];
	}
}

// the term's name as smina prints it (src/lib/everything.h: parameters with
// the stream's default formatting, vdw's powers as integers)
std::string term_name(const sgpu_term& t)
{
	auto f = [](double x) {
		std::ostringstream o;
		o << x;
		return o.str();
	};
	switch (t.kind)
	{
	case SGPU_TERM_GAUSS: return "gauss(o=" + f(t.p[0]) + ",_w=" + f(t.p[1]) + ",_c=" + f(t.p[2]) + ")";
	case SGPU_TERM_REPULSION: return "repulsion(o=" + f(t.p[0]) + ",_c=" + f(t.p[1]) + ")";
	case SGPU_TERM_HYDROPHOBIC: return "hydrophobic(g=" + f(t.p[0]) + ",_b=" + f(t.p[1]) + ",_c=" + f(t.p[2]) + ")";
	case SGPU_TERM_NON_HYDROPHOBIC:
		return "non_hydrophobic(g=" + f(t.p[0]) + ",_b=" + f(t.p[1]) + ",_c=" + f(t.p[2]) + ")";
	case SGPU_TERM_NON_DIR_H_BOND:
		return "non_dir_h_bond(g=" + f(t.p[0]) + ",_b=" + f(t.p[1]) + ",_c=" + f(t.p[2]) + ")";
	case SGPU_TERM_NON_DIR_ANTI_H_BOND_QUADRATIC:
		return "non_dir_anti_h_bond_quadratic(o=" + f(t.p[0]) + ",_c=" + f(t.p[1]) + ")";
	case SGPU_TERM_DONOR_DONOR_QUADRATIC: return "donor_donor_quadratic(o=" + f(t.p[0]) + ",_c=" + f(t.p[1]) + ")";
	case SGPU_TERM_ACCEPTOR_ACCEPTOR_QUADRATIC:
		return "acceptor_acceptor_quadratic(o=" + f(t.p[0]) + ",_c=" + f(t.p[1]) + ")";
	case SGPU_TERM_VDW:
		return "vdw(i=" + f(double(int(t.p[0]))) + ",_j=" + f(double(int(t.p[1]))) + ",_s=" + f(t.p[2]) +
				",_^=" + f(t.p[3]) + ",_c=" + f(t.p[4]) + ")";
	case SGPU_TERM_NON_DIR_H_BOND_LJ:
		return "non_dir_h_bond_lj(o=" + f(t.p[0]) + ",_^=" + f(t.p[1]) + ",_c=" + f(t.p[2]) + ")";
	case SGPU_TERM_ELECTROSTATIC:
		return "electrostatic(i=" + f(double(int(t.p[0]))) + ",_^=" + f(t.p[1]) + ",_c=" + f(t.p[2]) + ")";
	case SGPU_TERM_AD4_SOLVATION:
		return "ad4_solvation(d-sigma=" + f(t.p[0]) + ",_s/q=" + f(t.p[1]) + ",_c=" + f(t.p[2]) + ")";
	}
	throw std::runtime_error("unknown scoring term");
}

Tables build_tables(const sgpu_scoring& sf, double factor)
{
	Tables t;
	t.factor = factor;
	const double cutoff = 8;
	t.cutoff_sqr = cutoff * cutoff;
	t.n = int(size_t(factor * t.cutoff_sqr) + 3);
	const int n = t.n;
	std::vector<double> rs(n + 2);
	for (int i = 0; i < n + 2; i++)
		rs[i] = std::sqrt(i / factor);
	bool charged = false;
	for (int k = 0; k < sf.nterms; k++)
		charged = charged || charge_dependent(sf.terms[k]);
	t.ncomp = charged ? 4 : 1;
	t.smooth.assign(size_t(SGPU_NUM_TYPE_PAIRS) * n * 2, 0);
	t.fast.assign(size_t(SGPU_NUM_TYPE_PAIRS) * n, 0);
	if (charged)
	{
		t.qsmooth.assign(size_t(SGPU_NUM_TYPE_PAIRS) * n * 6, 0);
		t.qfast.assign(size_t(SGPU_NUM_TYPE_PAIRS) * n * 3, 0);
	}
	// per knot the four components: the charge-independent terms summed
	// into the first, then each charge-dependent term's components times
	// its weight (weighted_terms::eval_fast, src/lib/weighted_terms.cpp:52-62)
	std::vector<double> e(size_t(n) * 4);
	for (int t1 = 0; t1 < NUM_ATOM_TYPES; t1++)
		for (int t2 = t1; t2 < NUM_ATOM_TYPES; t2++)
		{
			for (int i = 0; i < n; i++)
			{
				double* acc = &e[4 * size_t(i)];
				acc[0] = acc[1] = acc[2] = acc[3] = 0;
				for (int k = 0; k < sf.nterms; k++)
					if (!charge_dependent(sf.terms[k]))
						acc[0] += sf.terms[k].weight * term(sf.terms[k], t1, t2, rs[i], sf.atom_params);
				for (int k = 0; k < sf.nterms; k++)
					if (charge_dependent(sf.terms[k]))
					{
						double c[4];
						components(sf.terms[k], t1, t2, rs[i], c);
						for (int j = 0; j < 4; j++)
							acc[j] += c[j] * sf.terms[k].weight;
					}
			}
			const size_t base = size_t(tri_index(t1, t2)) * n;
			for (int c = 0; c < t.ncomp; c++)
			{
				auto val = [&](int i) { return e[4 * size_t(i) + c]; };
				for (int i = 0; i < n; i++)
				{
					double dor = 0;
					if (!(i == 0 || i == n - 1))
					{
						const double delta = rs[i + 1] - rs[i - 1];
						const double r = rs[i];
						dor = (val(i + 1) - val(i - 1)) / (delta * r);
					}
					const double f1 = val(i);
					const double f2 = (i + 1 >= n) ? 0 : val(i + 1);
					if (c == 0)
					{
						t.smooth[2 * (base + i)] = val(i);
						t.smooth[2 * (base + i) + 1] = dor;
						t.fast[base + i] = (f2 + f1) / 2;
					}
					else
					{
						t.qsmooth[2 * ((base + i) * 3 + c - 1)] = val(i);
						t.qsmooth[2 * ((base + i) * 3 + c - 1) + 1] = dor;
						t.qfast[(base + i) * 3 + c - 1] = (f2 + f1) / 2;
					}
				}
			}
		}
	return t;
}

// The terms at nseg = factor x cutoff distances i x cutoff / nseg plus a
// last point (cutoff, 0); per type pair the cubic spline through them with
// zero slope at both ends (Spline::initialize, src/lib/splines.h:37-99).
// Its second derivatives solve a tridiagonal system: row 0 is
// 2h m0 + h m1 = 6 (y1 - y0) / h; interior row i is
// hi m(i-1) + 2 (h + hi) mi + hi m(i+1) = 6 ((y(i+1) - yi) / hi - (yi - y(i-1)) / h)
// with hi the last segment's length on the last interior row; the last row
// is hl m(e-1) + 2 hl me = -6 (ye - y(e-1)) / hl.  The reference inverts the
// dense matrix; elimination gives the same solution up to rounding.  Pairs
// whose terms vanish at every point get no spline (cutoff 0).
Splines build_splines(const sgpu_scoring& sf, double factor)
{
	if (!(factor > 0))
		throw std::invalid_argument("spline factor must be positive");
	Splines s;
	const double cutoff = 8;
	s.cutoff_sqr = cutoff * cutoff;
	const unsigned np = unsigned(factor * cutoff);
	if (np < 2)
		throw std::invalid_argument("spline factor too small");
	const int e = int(np);
	s.nseg = e;
	s.fraction = cutoff / double(np);
	s.coef.assign(size_t(SGPU_NUM_TYPE_PAIRS) * e * 5, 0);
	s.cut.assign(size_t(SGPU_NUM_TYPE_PAIRS), 0);
	std::vector<double> x(e + 1), y(e + 1), lo(e + 1), di(e + 1), up(e + 1), rhs(e + 1), cp(e + 1), rp(e + 1),
			m(e + 1);
	for (int t1 = 0; t1 < NUM_ATOM_TYPES; t1++)
		for (int t2 = t1; t2 < NUM_ATOM_TYPES; t2++)
		{
			bool nonzero = false;
			for (int i = 0; i < e; i++)
			{
				x[i] = i * s.fraction;
				double acc = 0;
				for (int k = 0; k < sf.nterms; k++)
					acc += sf.terms[k].weight * term(sf.terms[k], t1, t2, x[i], sf.atom_params);
				y[i] = acc;
				nonzero = nonzero || acc != 0;
			}
			x[e] = cutoff;
			y[e] = 0;
			if (!nonzero)
				continue;
			const double h = x[1] - x[0], hl = x[e] - x[e - 1];
			di[0] = 2 * h;
			up[0] = h;
			rhs[0] = 6 * ((y[1] - y[0]) / h);
			for (int i = 1; i < e; i++)
			{
				const double hi = i == e - 1 ? hl : h;
				lo[i] = hi;
				di[i] = 2 * (h + hi);
				up[i] = hi;
				rhs[i] = 6 * ((y[i + 1] - y[i]) / hi - (y[i] - y[i - 1]) / h);
			}
			lo[e] = hl;
			di[e] = 2 * hl;
			rhs[e] = 6 * (-(y[e] - y[e - 1]) / hl);
			cp[0] = up[0] / di[0];
			rp[0] = rhs[0] / di[0];
			for (int i = 1; i <= e; i++)
			{
				const double piv = di[i] - lo[i] * cp[i - 1];
				cp[i] = i < e ? up[i] / piv : 0;
				rp[i] = (rhs[i] - lo[i] * rp[i - 1]) / piv;
			}
			m[e] = rp[e];
			for (int i = e - 1; i >= 0; i--)
				m[i] = rp[i] - cp[i] * m[i + 1];
			const size_t tri = size_t(tri_index(t1, t2));
			for (int i = 0; i < e; i++)
			{
				const double hi = i == e - 1 ? hl : h;
				double* c = &s.coef[(tri * e + i) * 5];
				c[0] = x[i];
				c[1] = (m[i + 1] - m[i]) / (6 * hi);
				c[2] = m[i] / 2;
				c[3] = (y[i + 1] - y[i]) / hi - m[i + 1] * hi / 6 - m[i] * hi / 3;
				c[4] = y[i];
			}
			s.cut[tri] = cutoff;
		}
	return s;
}

void FlatModel::publish()
{
	view.atom_type = atom_type.data();
	view.atom_frame = atom_frame.data();
	view.atom_local = atom_local.data();
	view.atom_init = atom_init.data();
	view.frame_parent = frame_parent.data();
	view.frame_kind = frame_kind.data();
	view.frame_begin = frame_begin.data();
	view.frame_end = frame_end.data();
	view.frame_tree = frame_tree.data();
	view.frame_conf = frame_conf.data();
	view.frame_change = frame_change.data();
	view.frame_origin = frame_origin.data();
	view.frame_axis = frame_axis.data();
	view.pair_a = pair_a.data();
	view.pair_b = pair_b.data();
	view.pair_tri = pair_tri.data();
	view.atom_charge = atom_charge.empty() ? nullptr : atom_charge.data();
}

FlatModel flatten(const sgpu_model_desc& d)
{
	if (d.nmovable <= 0 || d.nmovable > d.natoms || d.nframes <= 0 || d.ntrees <= 0)
		throw std::runtime_error("invalid model description");
	FlatModel m;
	const int N = d.nmovable, NA = d.natoms;
	m.atom_type.assign(d.type, d.type + NA);
	for (int t : m.atom_type)
		if (t < 0 || t >= NUM_ATOM_TYPES)
			throw std::runtime_error("atom type out of range");
	m.atom_init.assign(d.xyz, d.xyz + 3 * size_t(NA));
	if (d.charge)
		m.atom_charge.assign(d.charge, d.charge + NA);
	auto X = [&](int a, int k) { return d.xyz[3 * a + k]; };
	std::vector<int> tree_of(d.nframes, -1);
	for (int t = 0; t < d.ntrees; t++)
	{
		const int last = t + 1 < d.ntrees ? d.tree_frame[t + 1] : d.nframes;
		for (int f = d.tree_frame[t]; f < last; f++)
			tree_of[f] = t;
	}
	m.atom_frame.assign(N, -1);
	m.frame_origin.assign(3 * size_t(d.nframes), 0);
	m.frame_axis.assign(3 * size_t(d.nframes), 0);
	std::vector<double> abs_origin(3 * size_t(d.nframes));
	int conf_off = 0, change_off = 0;
	std::vector<int> torsion_count(d.ntrees, 0);
	for (int t = 0; t < d.ntrees; t++)
	{
		const bool lig = t < d.nligands;
		const int first = d.tree_frame[t];
		const int last = t + 1 < d.ntrees ? d.tree_frame[t + 1] : d.nframes;
		const int ntors = (last - first) - (lig ? 1 : 0);
		m.tree_ligand.push_back(lig ? 1 : 0);
		m.tree_ntors.push_back(ntors);
		m.tree_conf.push_back(conf_off);
		int k = 0;
		for (int f = first; f < last; f++)
		{
			const int parent = d.frame_parent[f];
			int kind = parent < 0 ? (lig ? SGPU_FRAME_RIGID_ROOT : SGPU_FRAME_FIRST_SEGMENT) : SGPU_FRAME_SEGMENT;
			m.frame_parent.push_back(parent);
			m.frame_kind.push_back(kind);
			m.frame_begin.push_back(d.frame_begin[f]);
			m.frame_end.push_back(d.frame_end[f]);
			m.frame_tree.push_back(t);
			for (int a = d.frame_begin[f]; a < d.frame_end[f]; a++)
				m.atom_frame[a] = f;
			const int oa = d.frame_origin_atom[f];
			for (int c = 0; c < 3; c++)
				abs_origin[3 * f + c] = X(oa, c);
			if (kind == SGPU_FRAME_RIGID_ROOT)
			{
				m.frame_conf.push_back(conf_off);
				m.frame_change.push_back(change_off);
				for (int c = 0; c < 3; c++)
					m.frame_origin[3 * f + c] = X(oa, c);
			}
			else
			{
				const int idx = lig ? k - 1 : k;
				m.frame_conf.push_back(conf_off + (lig ? 7 : 0) + idx);
				m.frame_change.push_back(change_off + (lig ? 6 : 0) + idx);
				// axis_frame (src/lib/tree.h:116-125): unit vector from the axis
				// atom to the origin
				const int aa = d.frame_axis_atom[f];
				if (aa < 0 || aa >= NA)
					throw std::runtime_error("rotatable frame without an axis atom");
				const double diff[3] = {X(oa, 0) - X(aa, 0), X(oa, 1) - X(aa, 1), X(oa, 2) - X(aa, 2)};
				const double nrm = std::sqrt(sq(diff[0]) + sq(diff[1]) + sq(diff[2]));
				if (!(nrm >= kEps))
					throw std::runtime_error("zero-length rotatable bond");
				const double inv = 1 / nrm;
				for (int c = 0; c < 3; c++)
					m.frame_axis[3 * f + c] = inv * diff[c];
				if (kind == SGPU_FRAME_SEGMENT)
				{ // segment (src/lib/tree.h:139-144): relative to the parent's origin
					for (int c = 0; c < 3; c++)
						m.frame_origin[3 * f + c] = X(oa, c) - abs_origin[3 * parent + c];
				}
				else
					for (int c = 0; c < 3; c++)
						m.frame_origin[3 * f + c] = X(oa, c);
			}
			++k;
		}
		conf_off += lig ? 7 + ntors : ntors;
		change_off += lig ? 6 + ntors : ntors;
	}
	for (int a = 0; a < N; a++)
		if (m.atom_frame[a] < 0)
			throw std::runtime_error("movable atom outside every frame");
	m.atom_local.resize(3 * size_t(N));
	for (int a = 0; a < N; a++)
		for (int c = 0; c < 3; c++)
			m.atom_local[3 * a + c] = X(a, c) - abs_origin[3 * m.atom_frame[a] + c];
	// bonds and mobility (src/lib/parse_pdbqt.cpp:279-336)
	std::vector<std::vector<int>> bonds(NA);
	for (int k = 0; k < d.nbonds; k++)
	{
		const int a = d.bonds[2 * k], b = d.bonds[2 * k + 1];
		if (a < 0 || b < 0 || a >= NA || b >= NA)
			throw std::runtime_error("bond atom out of range");
		bonds[a].push_back(b);
		bonds[b].push_back(a);
	}
	auto fixed = [&](int i, int j) -> bool {
		if (i >= N && j >= N)
			return true;
		if (i >= N || j >= N)
		{
			const int mv = i < N ? i : j, o = i < N ? j : i;
			const int f = m.atom_frame[mv];
			return m.frame_kind[f] == SGPU_FRAME_FIRST_SEGMENT && d.frame_axis_atom[f] == o;
		}
		const int fi = m.atom_frame[i], fj = m.atom_frame[j];
		if (fi == fj)
			return true;
		if (m.frame_kind[fj] != SGPU_FRAME_RIGID_ROOT && (d.frame_axis_atom[fj] == i || d.frame_origin_atom[fj] == i))
			return true;
		if (m.frame_kind[fi] != SGPU_FRAME_RIGID_ROOT && (d.frame_axis_atom[fi] == j || d.frame_origin_atom[fi] == j))
			return true;
		return false;
	};
	std::vector<int> lig_begin(d.nligands, N), lig_end(d.nligands, 0);
	for (int f = 0; f < d.nframes; f++)
	{
		const int t = tree_of[f];
		if (t < d.nligands)
		{
			lig_begin[t] = std::min(lig_begin[t], d.frame_begin[f]);
			lig_end[t] = std::max(lig_end[t], d.frame_end[f]);
		}
	}
	auto find_ligand = [&](int a) {
		for (int l = 0; l < d.nligands; l++)
			if (a >= lig_begin[l] && a < lig_end[l])
				return l;
		return d.nligands;
	};
	// bonded_to (src/lib/model.cpp:576-589): depth first, an atom already
	// listed is not expanded again
	std::function<void(int, int, std::vector<int>&)> bonded_to = [&](int a, int n, std::vector<int>& out) {
		if (std::find(out.begin(), out.end(), a) != out.end())
			return;
		out.push_back(a);
		if (n > 0)
			for (int b : bonds[a])
				bonded_to(b, n - 1, out);
	};
	// initialize_pairs (src/lib/model.cpp:598-624)
	std::vector<std::pair<int, int>> other;
	std::vector<std::vector<std::pair<int, int>>> ligp(d.nligands);
	for (int i = 0; i < NA; i++)
	{
		const int il = find_ligand(i);
		std::vector<int> near;
		bonded_to(i, 3, near);
		for (int j = i + 1; j < NA; j++)
		{
			if (i >= N && j >= N)
				continue;
			if (fixed(i, j) || std::find(near.begin(), near.end(), j) != near.end())
				continue;
			const int t1 = m.atom_type[i], t2 = m.atom_type[j];
			if (is_hydrogen_type(t1) || is_hydrogen_type(t2))
				continue;
			if (il < d.nligands && find_ligand(j) == il)
				ligp[il].push_back({i, j});
			else
				other.push_back({i, j});
		}
	}
	auto push_pair = [&](std::pair<int, int> p) {
		m.pair_a.push_back(p.first);
		m.pair_b.push_back(p.second);
		m.pair_tri.push_back(tri_index_permissive(m.atom_type[p.first], m.atom_type[p.second]));
	};
	for (auto& p : other)
		push_pair(p);
	for (auto& lp : ligp)
		for (auto& p : lp)
			push_pair(p);
	// num_tors (src/lib/terms.cpp:39-110), rotatable hydrogens fixed
	auto heavy_nbrs = [&](int a) {
		int k = 0;
		for (int b : bonds[a])
			if (!is_hydrogen_type(m.atom_type[b]))
				++k;
		return k;
	};
	auto rotatable = [&](int a, int b) {
		for (int f = 0; f < d.nframes; f++)
			if (m.frame_kind[f] == SGPU_FRAME_SEGMENT &&
					((d.frame_axis_atom[f] == a && d.frame_origin_atom[f] == b) ||
							(d.frame_axis_atom[f] == b && d.frame_origin_atom[f] == a)))
				return true;
		return false;
	};
	m.num_tors = 0;
	for (int l = 0; l < d.nligands; l++)
		for (int i = lig_begin[l]; i < lig_end[l]; i++)
		{
			if (is_hydrogen_type(m.atom_type[i]))
				continue;
			unsigned ar = 0;
			for (int b : bonds[i])
				if (rotatable(i, b) && !is_hydrogen_type(m.atom_type[b]) && heavy_nbrs(b) > 1 && heavy_nbrs(i) > 1)
					++ar;
			m.num_tors += 0.5 * ar;
		}
	m.nheavy = 0;
	for (int a = 0; a < N; a++)
		if (!is_hydrogen_type(m.atom_type[a]))
			++m.nheavy;
	sgpu_flat_model& v = m.view;
	v.natoms = N;
	v.ninflex = NA - N;
	v.nframes = d.nframes;
	v.nligands = d.nligands;
	v.ntrees = d.ntrees;
	v.npairs = int(m.pair_a.size());
	v.nother = int(other.size());
	v.conf_size = conf_off;
	v.change_size = change_off;
	v.num_tors = int(m.num_tors);
	m.publish();
	return m;
}

std::vector<double> initial_conf(const FlatModel& m)
{ // model::get_initial_conf (src/lib/model.cpp:650-659)
	std::vector<double> c(m.view.conf_size, 0.0);
	for (int f = 0; f < m.view.nframes; f++)
		if (m.frame_kind[f] == SGPU_FRAME_RIGID_ROOT)
		{
			const int ci = m.frame_conf[f];
			for (int k = 0; k < 3; k++)
				c[ci + k] = m.frame_origin[3 * f + k];
			c[ci + 3] = 1;
		}
	return c;
}

void make_chain(const FlatModel& m, uint32_t seed, const sgpu_grid_dims& gd, double* conf, uint32_t* rng_state)
{
	Gen g(seed);
	std::fill(conf, conf + m.view.conf_size, 0.0);
	// conf::randomize (src/lib/conf.h:362-367): ligands, then side chains
	for (size_t t = 0; t < m.tree_ligand.size(); t++)
	{
		if (!m.tree_ligand[t])
			continue;
		double* p = conf + m.tree_conf[t];
		for (int k = 0; k < 3; k++)
			p[k] = random_fl(gd.begin[k], gd.end[k], g);
		// random_orientation (src/lib/quaternion.cpp:129-141) with
		// boost::math::abs's scaled norm
		for (;;)
		{
			const double x0 = normal(g), x1 = normal(g), x2 = normal(g), x3 = normal(g);
			const double mx = std::max(std::max(std::abs(x0), std::abs(x1)), std::max(std::abs(x2), std::abs(x3)));
			double nrm = 0;
			if (mx != 0)
			{
				const double k = 1 / mx;
				const double a = x0 * k, b = x1 * k, c = x2 * k, d = x3 * k;
				nrm = mx * std::sqrt(a * a + b * b + c * c + d * d);
			}
			if (nrm > kEps)
			{
				p[3] = x0 / nrm;
				p[4] = x1 / nrm;
				p[5] = x2 / nrm;
				p[6] = x3 / nrm;
				break;
			}
		}
		for (int k = 0; k < m.tree_ntors[t]; k++)
			p[7 + k] = random_fl(-kPi, kPi, g);
	}
	for (size_t t = 0; t < m.tree_ligand.size(); t++)
		if (!m.tree_ligand[t])
			for (int k = 0; k < m.tree_ntors[t]; k++)
				conf[m.tree_conf[t] + k] = random_fl(-kPi, kPi, g);
	if (rng_state)
	{
		std::memcpy(rng_state, g.s.x, sizeof(g.s.x));
		rng_state[624] = g.s.i;
	}
}

SearchParams setup_search(const FlatModel& m, int exhaustiveness, int num_modes, int minimize_iters)
{
	SearchParams p;
	p.num_saved_mins = std::max(20, num_modes);
	int dof = 0;
	for (size_t t = 0; t < m.tree_ligand.size(); t++)
		dof += m.tree_ntors[t] + (m.tree_ligand[t] ? 6 : 0);
	const int heuristic = m.view.natoms + 10 * dof;
	p.num_steps = int(unsigned(70 * 3 * (50 + heuristic) / 2));
	p.maxiters = minimize_iters > 0 ? minimize_iters : (25 + m.view.natoms) / 3;
	p.num_tasks = exhaustiveness;
	return p;
}

// add_to_output_container (src/lib/coords.cpp:34-58); the container stays
// sorted by energy, ties in insertion order
void add_to_output_container(std::vector<Pose>& out, const Pose& t, double min_rmsd, size_t max_size)
{
	size_t best_i = out.size();
	double best = kMax;
	for (size_t i = 0; i < out.size(); i++)
	{
		double acc = 0;
		const size_t na = t.coords.size() / 3;
		for (size_t a = 0; a < na; a++)
			acc += sq(t.coords[3 * a] - out[i].coords[3 * a]) + sq(t.coords[3 * a + 1] - out[i].coords[3 * a + 1]) +
					sq(t.coords[3 * a + 2] - out[i].coords[3 * a + 2]);
		const double res = na > 0 ? std::sqrt(acc / na) : 0;
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
	std::stable_sort(out.begin(), out.end(), [](const Pose& a, const Pose& b) { return a.e < b.e; });
}

namespace
{
void check(sgpu_ctx* c, sgpu_status s)
{
	if (s != SGPU_OK)
		throw std::runtime_error(std::string("sgpu: ") + sgpu_last_error(c));
}
} // namespace

GpuDocking::GpuDocking(const GpuSettings& s, const sgpu_scoring& sf, const sgpu_receptor& rec,
		const sgpu_grid_dims& gd) :
		s_(s), gd_(gd)
{
	terms_.assign(sf.terms, sf.terms + sf.nterms);
	ci_.assign(sf.ci, sf.ci + sf.nci);
	sf_ = sf;
	sf_.terms = terms_.data();
	sf_.ci = ci_.data();
	if (s.approximation != GpuSettings::LINEAR && s.approximation != GpuSettings::SPLINE)
		throw std::invalid_argument("approximation must be LINEAR or SPLINE");
	if (s.approximation == GpuSettings::SPLINE)
		for (const sgpu_term& t : terms_)
			if (charge_dependent(t))
				throw std::invalid_argument("charge-dependent terms need the linear tables (approximation LINEAR)");
	sgpu_status st = sgpu_create(s.device, &ctx_);
	if (st != SGPU_OK)
		throw std::runtime_error("sgpu_create failed: no HIP device");
	if (s.approximation == GpuSettings::SPLINE)
	{
		Splines sp = build_splines(sf_, s.approx_factor);
		sgpu_splines ss{sp.nseg, sp.fraction, sp.cutoff_sqr, sp.coef.data(), sp.cut.data()};
		check(ctx_, sgpu_upload_splines(ctx_, &ss));
	}
	else
	{
		Tables t = build_tables(sf_, s.approx_factor);
		sgpu_tables tt{t.n, t.factor, t.cutoff_sqr, t.smooth.data(), t.fast.data(), t.ncomp,
				t.ncomp == 4 ? t.qsmooth.data() : nullptr, t.ncomp == 4 ? t.qfast.data() : nullptr};
		check(ctx_, sgpu_upload_tables(ctx_, &tt));
	}
	check(ctx_, sgpu_upload_receptor(ctx_, &rec, &receptor_));
	if (s.user_grid)
		check(ctx_, sgpu_set_user_grid(ctx_, s.user_grid));
	s_.user_grid = nullptr; // the device holds its copy
	check(ctx_, sgpu_set_progress(ctx_, s.progress_chunks, s.progress, s.progress_user));
	check(ctx_, sgpu_mc_capacity(ctx_, s.precision, &capacity_));
}

GpuDocking::GpuDocking(const GpuSettings& s, const sgpu_scoring& sf, const sgpu_receptor& rec,
		const sgpu_grid_dims& gd, sgpu_ctx* shared) :
		s_(s), gd_(gd)
{
	if (!shared)
		throw std::invalid_argument("no shared context");
	ctx_ = shared;
	owns_ctx_ = false;
	terms_.assign(sf.terms, sf.terms + sf.nterms);
	ci_.assign(sf.ci, sf.ci + sf.nci);
	sf_ = sf;
	sf_.terms = terms_.data();
	sf_.ci = ci_.data();
	if (s.approximation != GpuSettings::LINEAR)
		throw std::invalid_argument("a shared context carries the linear tables");
	Tables t = build_tables(sf_, s.approx_factor);
	sgpu_tables tt{t.n, t.factor, t.cutoff_sqr, t.smooth.data(), t.fast.data(), t.ncomp,
			t.ncomp == 4 ? t.qsmooth.data() : nullptr, t.ncomp == 4 ? t.qfast.data() : nullptr};
	check(ctx_, sgpu_upload_tables(ctx_, &tt));
	check(ctx_, sgpu_upload_receptor(ctx_, &rec, &receptor_));
	check(ctx_, sgpu_mc_capacity(ctx_, s.precision, &capacity_));
}

GpuDocking::~GpuDocking()
{
	if (owns_ctx_)
		sgpu_destroy(ctx_);
	else if (grid_ >= 0)
		sgpu_release_grid(ctx_, grid_);
}

size_t GpuDocking::map_bytes() const
{
	if (grid_ < 0)
		return 0;
	size_t cells = 1;
	for (int d = 0; d < 3; d++)
		cells *= size_t(gd_.n[d]);
	const size_t points = size_t(gd_.n[0] + 1) * (gd_.n[1] + 1) * (gd_.n[2] + 1);
	return grid_types_.size() * (cells * 8 * (s_.precision == SGPU_FP64 ? 8 : 4) + points * 8);
}

std::string GpuDocking::description() const
{
	char name[256] = {0};
	int cus = 0;
	sgpu_device_info(ctx_, name, sizeof(name), &cus);
	return std::string(name) + ", " + std::to_string(cus) + " CUs, " + (s_.precision == SGPU_FP64 ? "FP64" : s_.precision == SGPU_MIXED ? "FP32 evaluation, FP64 search" : "FP32");
}

void GpuDocking::adopt_grid(sgpu_ctx* store, int store_grid, const std::vector<int>& types)
{
	if (grid_ >= 0)
		sgpu_release_grid(ctx_, grid_);
	grid_ = -1;
	check(ctx_, sgpu_share_grid(ctx_, store, store_grid, &grid_));
	grid_types_ = types;
	std::sort(grid_types_.begin(), grid_types_.end());
	shared_grid_ = true;
}

void GpuDocking::ensure_grids(const std::vector<int>& types)
{
	bool have = grid_ >= 0;
	for (int t : types)
		have = have && std::find(grid_types_.begin(), grid_types_.end(), t) != grid_types_.end();
	if (have)
		return;
	if (shared_grid_)
		throw std::invalid_argument("a ligand atom type is not among the served receptor's map types");
	std::vector<int> all = grid_types_;
	for (int t : types)
		if (std::find(all.begin(), all.end(), t) == all.end())
			all.push_back(t);
	std::sort(all.begin(), all.end());
	if (grid_ >= 0)
		sgpu_release_grid(ctx_, grid_);
	check(ctx_, sgpu_build_grids(ctx_, receptor_, &gd_, all.data(), int(all.size()), s_.precision == SGPU_FP64 ? SGPU_FP64 : SGPU_FP32,
						&grid_));
	grid_types_ = all;
}

void GpuDocking::prepare(const std::vector<const sgpu_model_desc*>& models)
{
	std::vector<int> types;
	for (const sgpu_model_desc* d : models)
		for (int a = 0; a < d->nmovable; a++)
			if (!is_hydrogen_type(d->type[a]) && std::find(types.begin(), types.end(), d->type[a]) == types.end())
				types.push_back(d->type[a]);
	ensure_grids(types);
}

void apply_fast(GpuSettings& s)
{
	s.chain_split = 0;
	s.step_fraction = 0;
	s.skip_visited = true;
	if (s.early_stop == GpuSettings::EARLY_STOP_OFF)
		s.early_stop = GpuSettings::EARLY_STOP_FULL;
	s.fast_top = 0.1;
}

// auto_split (src/gpu/gpu_docking.cpp:87-102): the largest power of two from
// 4 to 32 that keeps tasks x split within 4 times the chains the device holds
// with fraction_for and auto_fraction (src/gpu/gpu_docking.cpp:104-112,
// :167-171): a first-pass budget, else the set fraction, else half the steps
// for a split in [4, 32) and all of them otherwise; and stops / stop_for
// (:172-189): no early stopping in a first pass, EARLY_STOP_FULL only below
// split 32, checkpoints after every tenth of smina's steps the chains run
// (at least first + 1), the rule from the second on
SearchPlan plan_search(const GpuSettings& s, size_t tasks, int capacity, double budget)
{
	SearchPlan p;
	if (s.chain_split > 0)
		p.split = s.chain_split;
	else
	{
		p.split = 32;
		while (p.split > 4 && tasks * size_t(p.split) > 4 * size_t(std::max(capacity, 1)))
			p.split /= 2;
	}
	p.fraction = budget > 0 ? budget : s.step_fraction > 0 ? s.step_fraction
	                                                        : (p.split >= 4 && p.split < 32 ? 0.5 : 1.0);
	const bool stops = !(budget > 0) && (s.early_stop == GpuSettings::EARLY_STOP_ALWAYS ||
	                                     (s.early_stop == GpuSettings::EARLY_STOP_FULL && p.split < 32));
	if (stops)
	{
		p.first = 2;
		p.checkpoints = std::max(p.first + 1, int(std::lround(10 * p.fraction)));
	}
	return p;
}

void GpuDocking::dock_pass(std::vector<DockJob*>& jobs, double budget)
{
	const size_t per = s_.ligands_per_launch > 0 ? size_t(s_.ligands_per_launch) : jobs.size();
	for (size_t b = 0; b < jobs.size(); b += per)
		dock_sub(jobs.data() + b, std::min(per, jobs.size() - b), budget);
}

// src/main/main.cpp:1578-1586: the score of the history's best fast_top
// (the ceil(fast_top x n)-th lowest) plus the margin
double full_search_cutoff(std::vector<double> history, double fast_top, double margin)
{
	if (history.empty())
		throw std::invalid_argument("no first-pass scores");
	const size_t q = std::max<size_t>(1, size_t(std::ceil(fast_top * double(history.size())))) - 1;
	std::nth_element(history.begin(), history.begin() + ptrdiff_t(q), history.end());
	return history[q] + margin;
}

// the batch flush of src/main/main.cpp:1551-1597: one pass, or with fast_top
// in (0, 1) and at least 20 ligands seen, a first pass at fast_first of the
// steps whose top scores join the history, then the full search for the
// ligands within fast_top_margin of the history's best fast_top
void GpuDocking::tune(std::vector<DockJob*>& jobs, int probe_steps, sgpu_mc_tune_report& report)
{
	if (probe_steps < 1)
		throw std::invalid_argument("probe_steps must be positive");
	tune_steps_ = probe_steps;
	tune_report_ = &report;
	struct Reset
	{
		GpuDocking* d;
		~Reset() { d->tune_report_ = nullptr; }
	} reset{this};
	dock_pass(jobs, 0);
}

void GpuDocking::dock_batch(std::vector<DockJob*>& jobs)
{
	full_searches_ = jobs.size();
	if (!(s_.fast_top > 0 && s_.fast_top < 1 && top_history_.size() + jobs.size() >= 20))
	{
		dock_pass(jobs, 0);
		return;
	}
	if (!(s_.fast_first > 0))
		throw std::invalid_argument("fast_first must be positive");
	dock_pass(jobs, s_.fast_first);
	std::vector<double> score(jobs.size());
	for (size_t k = 0; k < jobs.size(); k++)
		score[k] = jobs[k]->out.empty() ? kMax : jobs[k]->out[0].e;
	top_history_.insert(top_history_.end(), score.begin(), score.end());
	const double cutoff = full_search_cutoff(top_history_, s_.fast_top, s_.fast_top_margin);
	std::vector<DockJob*> again;
	std::vector<int64_t> first_evals;
	for (size_t k = 0; k < jobs.size(); k++)
		if (score[k] <= cutoff)
		{
			jobs[k]->out.clear();
			again.push_back(jobs[k]);
			first_evals.push_back(jobs[k]->evals);
		}
	full_searches_ = again.size();
	dock_pass(again, 0);
	for (size_t k = 0; k < again.size(); k++)
		again[k]->evals += first_evals[k];
}

sgpu_grid_dims movable_atoms_box(const FlatModel& m, double add, double granularity)
{
	return movable_atoms_box_at(m.atom_init.data(), m.view.natoms, add, granularity);
}

sgpu_grid_dims movable_atoms_box_at(const double* xyz, int n, double add, double granularity)
{
	double c1[3] = {0, 0, 0}, c2[3] = {0, 0, 0};
	for (int i = 0; i < n; i++)
		for (int j = 0; j < 3; j++)
		{
			const double v = xyz[size_t(3 * i + j)];
			if (i == 0 || v < c1[j])
				c1[j] = v;
			if (i == 0 || v > c2[j])
				c2[j] = v;
		}
	sgpu_grid_dims gd{};
	for (int j = 0; j < 3; j++)
	{
		c1[j] -= add / 2;
		c2[j] += add / 2;
	}
	for (int j = 0; j < 3; j++)
	{
		const double center = 0.5 * (c2[j] + c1[j]);
		gd.n[j] = int(std::ceil((c2[j] - c1[j]) / granularity));
		const double span = granularity * gd.n[j];
		gd.begin[j] = center - span / 2;
		gd.end[j] = gd.begin[j] + span;
	}
	return gd;
}

void GpuDocking::minimize_batch(std::vector<MinimizeJob*>& jobs, double cap)
{
	if (jobs.empty())
		return;
	if (!(cap > 0))
		throw std::invalid_argument("minimize: the force cap must be positive");
	std::vector<FlatModel> fms;
	for (MinimizeJob* j : jobs)
		fms.push_back(flatten(*j->m));
	std::vector<sgpu_flat_model> views;
	for (auto& f : fms)
		views.push_back(f.view);
	int batch = -1;
	check(ctx_, sgpu_upload_models(ctx_, views.data(), int(views.size()), SGPU_FP64, &batch));
	int cstride = 1, hstride = 1;
	for (auto& f : fms)
	{
		cstride = std::max(cstride, f.view.conf_size);
		hstride = std::max(hstride, 3 * f.nheavy);
	}
	// one refine request per line-search kind; cells over the union of boxes
	for (int acc = 0; acc < 2; acc++)
	{
		std::vector<size_t> idx;
		for (size_t k = 0; k < jobs.size(); k++)
			if (jobs[k]->accurate == (acc == 1))
				idx.push_back(k);
		if (idx.empty())
			continue;
		const int n = int(idx.size());
		std::vector<int> model(n), iters(n);
		std::vector<double> conf(size_t(n) * cstride, 0.0), boxes(size_t(n) * 6), e(n), x(size_t(n) * hstride);
		std::vector<std::vector<double>> orig(n);
		sgpu_grid_dims un{};
		for (int i = 0; i < n; i++)
		{
			MinimizeJob& j = *jobs[idx[size_t(i)]];
			const FlatModel& f = fms[idx[size_t(i)]];
			if (j.gd.n[0] == 0)
				j.gd = movable_atoms_box(f, 4, 0.375);
			if (j.conf.empty())
				j.conf = initial_conf(f);
			if (int(j.conf.size()) != f.view.conf_size)
				throw std::invalid_argument("minimize: conf size does not match the model");
			model[size_t(i)] = int(idx[size_t(i)]);
			iters[size_t(i)] = j.maxiters > 0 ? j.maxiters : 10000;
			std::copy(j.conf.begin(), j.conf.end(), conf.begin() + size_t(i) * cstride);
			for (int d = 0; d < 3; d++)
			{
				boxes[6 * size_t(i) + d] = j.gd.begin[d];
				boxes[6 * size_t(i) + 3 + d] = j.gd.end[d];
				un.begin[d] = i == 0 ? j.gd.begin[d] : std::min(un.begin[d], j.gd.begin[d]);
				un.end[d] = i == 0 ? j.gd.end[d] : std::max(un.end[d], j.gd.end[d]);
			}
			for (int a = 0; a < f.view.natoms; a++)
				if (!is_hydrogen_type(f.atom_type[size_t(a)]))
					for (int d = 0; d < 3; d++)
						orig[size_t(i)].push_back(f.atom_init[size_t(3 * a + d)]);
		}
		for (int d = 0; d < 3; d++)
			un.n[d] = std::max(1, int(std::ceil((un.end[d] - un.begin[d]) / 0.375)));
		sgpu_refine_request rq{};
		rq.batch_id = batch;
		rq.receptor_id = receptor_;
		rq.gd = &un;
		for (int k = 0; k < 3; k++)
			rq.v[k] = cap;
		rq.accurate_line_search = acc;
		rq.nposes = n;
		rq.model = model.data();
		rq.conf = conf.data();
		rq.conf_stride = cstride;
		rq.maxiters = iters.data();
		rq.e = e.data();
		rq.coords = x.data();
		rq.coord_stride = hstride;
		rq.boxes = boxes.data();
		const sgpu_status st = sgpu_run_refine(ctx_, &rq);
		if (st != SGPU_OK)
		{
			const std::string msg = sgpu_last_error(ctx_);
			(void)sgpu_release_models(ctx_, batch);
			throw std::runtime_error(msg);
		}
		for (int i = 0; i < n; i++)
		{
			MinimizeJob& j = *jobs[idx[size_t(i)]];
			const FlatModel& f = fms[idx[size_t(i)]];
			j.conf.assign(conf.begin() + size_t(i) * cstride, conf.begin() + size_t(i) * cstride + f.view.conf_size);
			j.e = e[size_t(i)];
			j.coords.assign(x.begin() + size_t(i) * hstride, x.begin() + size_t(i) * hstride + 3 * f.nheavy);
			double acc2 = 0;
			for (int a = 0; a < f.nheavy; a++)
			{
				double d2 = 0;
				for (int d = 0; d < 3; d++)
					d2 += (j.coords[size_t(3 * a + d)] - orig[size_t(i)][size_t(3 * a + d)]) *
							(j.coords[size_t(3 * a + d)] - orig[size_t(i)][size_t(3 * a + d)]);
				acc2 += d2;
			}
			j.rmsd = f.nheavy > 0 ? std::sqrt(acc2 / f.nheavy) : 0;
		}
	}
	check(ctx_, sgpu_release_models(ctx_, batch));
}

// the FP64 error of an FP32 or mixed search per job: its top pose's
// refined energy against the FP64 value of the refinement's own function at
// the same conf and box slope (sgpu_score mode 2)
void GpuDocking::fp64_errors(DockJob* const* jobs, size_t nj, int batch, int cstride)
{
	for (size_t j = 0; j < nj; j++)
		jobs[j]->fp64_error = -1;
	if (s_.precision == SGPU_FP64)
		return;
	for (int tries = 0; tries < 5; tries++)
	{
		std::vector<int> model;
		std::vector<double> conf;
		std::vector<size_t> which;
		for (size_t j = 0; j < nj; j++)
		{
			const auto& out = jobs[j]->out;
			if (out.empty() || out[0].tries != tries)
				continue;
			model.push_back(int(j));
			std::vector<double> c(size_t(cstride), 0.0);
			std::copy(out[0].c.begin(), out[0].c.end(), c.begin());
			conf.insert(conf.end(), c.begin(), c.end());
			which.push_back(j);
		}
		if (model.empty())
			continue;
		std::vector<double> e(model.size());
		sgpu_score_request sq{};
		sq.batch_id = batch;
		sq.receptor_id = receptor_;
		sq.gd = &gd_;
		sq.slope = 10 * std::pow(10.0, tries);
		for (int k = 0; k < 3; k++)
			sq.v[k] = s_.force_cap;
		sq.mode = 2;
		sq.nconfs = int(model.size());
		sq.model = model.data();
		sq.conf = conf.data();
		sq.conf_stride = cstride;
		sq.e = e.data();
		check(ctx_, sgpu_score(ctx_, &sq));
		for (size_t q = 0; q < which.size(); q++)
		{
			DockJob* job = jobs[which[q]];
			job->e_search = job->out[0].e_search;
			job->e_fp64 = e[q];
			job->fp64_error = std::fabs(job->e_search - job->e_fp64);
		}
	}
}

void GpuDocking::dock_sub(DockJob* const* jobs, size_t nj, double budget)
{
	if (nj == 0)
		return;
	const auto t0 = std::chrono::steady_clock::now();
	const int prec = s_.precision;
	std::vector<FlatModel> fms;
	std::vector<int> types;
	for (size_t j = 0; j < nj; j++)
	{
		fms.push_back(flatten(*jobs[j]->m));
		for (int a = 0; a < fms.back().view.natoms; a++)
		{
			const int t = fms.back().atom_type[a];
			if (!is_hydrogen_type(t) && std::find(types.begin(), types.end(), t) == types.end())
				types.push_back(t);
		}
	}
	ensure_grids(types);
	std::vector<sgpu_flat_model> views;
	for (auto& f : fms)
		views.push_back(f.view);
	int batch = -1;
	check(ctx_, sgpu_upload_models(ctx_, views.data(), int(views.size()), prec, &batch));
	// chains: per ligand rng(seed), tasks seeded in order (parallel_mc)
	int cstride = 1, hstride = 1;
	for (auto& f : fms)
	{
		cstride = std::max(cstride, f.view.conf_size);
		hstride = std::max(hstride, 3 * f.nheavy);
	}
	// with split n, n chains per exhaustiveness unit, each with 1/n of the
	// steps (add_chains and chain_steps, src/gpu/gpu_docking.cpp:200-220)
	std::vector<int> cmodel, csteps, citers, cunit;
	std::vector<double> cconf;
	std::vector<uint32_t> crng;
	std::vector<size_t> first(nj + 1, 0);
	std::vector<SearchParams> sp;
	size_t tasks = 0;
	for (size_t j = 0; j < nj; j++)
		tasks += size_t(jobs[j]->exhaustiveness > 0 ? jobs[j]->exhaustiveness : s_.exhaustiveness);
	const SearchPlan plan = plan_search(s_, tasks, capacity_, budget);
	const int split = plan.split;
	const double fraction = plan.fraction;
	for (size_t j = 0; j < nj; j++)
	{
		const int exh = jobs[j]->exhaustiveness > 0 ? jobs[j]->exhaustiveness : s_.exhaustiveness;
		sp.push_back(setup_search(fms[j], exh, s_.num_modes, s_.minimize_iters));
		Gen g(uint32_t(s_.seed));
		first[j] = cmodel.size();
		const int steps = std::max(1, int(std::ceil(double(sp[j].num_steps) * fraction / double(split))));
		for (int t = 0; t < exh * split; t++)
		{
			const int seed = random_int(0, 1000000, g);
			std::vector<double> conf(cstride, 0.0);
			std::vector<uint32_t> st(625);
			make_chain(fms[j], uint32_t(seed), gd_, conf.data(), st.data());
			cmodel.push_back(int(j));
			csteps.push_back(steps);
			citers.push_back(sp[j].maxiters);
			cunit.push_back(t / split);
			cconf.insert(cconf.end(), conf.begin(), conf.end());
			crng.insert(crng.end(), st.begin(), st.end());
		}
	}
	first[nj] = cmodel.size();
	const int nchains = int(cmodel.size());
	sgpu_chains ch{nchains, cmodel.data(), cconf.data(), cstride, crng.data(), csteps.data(), citers.data(),
			cunit.data()};
	sgpu_mc_params mp{};
	mp.temperature = 1.2;
	mp.mutation_amplitude = 2;
	mp.min_rmsd = 1.0;
	// setup_search (src/main/main.cpp:513-529); the chains' second pass
	// keeps monte_carlo's own cap of 1000 (src/lib/monte_carlo.cpp:96),
	// whatever --force_cap says
	mp.num_saved_mins = std::max(20, s_.num_modes);
	for (int k = 0; k < 3; k++)
	{
		mp.hunt_cap[k] = 10;
		mp.authentic_v[k] = 1000;
	}
	mp.cache_slope = 1e6;
	mp.accurate_line_search = s_.accurate_line ? 1 : 0;
	mp.early_term = s_.early_term ? 1 : 0;
	mp.visited = s_.skip_visited ? 1 : 0;
	if (plan.checkpoints > 0)
	{
		mp.early_stop_first = plan.first;
		mp.early_stop_checkpoints = plan.checkpoints;
		mp.early_stop_fraction = 0.5;
		mp.early_stop_eps = 0.05;
	}
	const int maxp = mp.num_saved_mins;
	std::vector<int> cnt(nchains);
	std::vector<double> oe(size_t(nchains) * maxp), oc(size_t(nchains) * maxp * cstride),
			ox(size_t(nchains) * maxp * hstride);
	std::vector<int64_t> evals(nchains, 0);
	sgpu_chain_results res{maxp, cstride, hstride, cnt.data(), oe.data(), oc.data(), ox.data(), evals.data()};
	if (tune_report_)
	{
		// a tuning call: probe the launch plans on these chains, search nothing
		check(ctx_, sgpu_tune_mc(ctx_, batch, grid_, &mp, &ch, tune_steps_, tune_report_));
		sgpu_release_models(ctx_, batch);
		return;
	}
	check(ctx_, sgpu_run_mc(ctx_, batch, grid_, &mp, &ch, &res));
	// merge_output_containers with min_rmsd 2 (src/gpu/gpu_docking.cpp:576-598)
	std::vector<int> pose_model, pose_iters;
	std::vector<double> pose_conf;
	std::vector<std::pair<size_t, size_t>> pose_ref;
	for (size_t j = 0; j < nj; j++)
	{
		auto& out = jobs[j]->out;
		out.clear();
		jobs[j]->evals = 0;
		for (size_t c = first[j]; c < first[j + 1]; c++)
		{
			jobs[j]->evals += evals[c];
			for (int k = 0; k < cnt[c]; k++)
			{
				Pose p;
				const size_t slot = c * maxp + k;
				p.e = oe[slot];
				p.c.assign(oc.begin() + slot * cstride, oc.begin() + slot * cstride + fms[j].view.conf_size);
				p.coords.assign(ox.begin() + slot * hstride, ox.begin() + slot * hstride + 3 * fms[j].nheavy);
				add_to_output_container(out, p, 2, size_t(sp[j].num_saved_mins));
			}
		}
		for (size_t k = 0; k < out.size(); k++)
		{
			pose_model.push_back(int(j));
			pose_iters.push_back(sp[j].maxiters);
			std::vector<double> c(cstride, 0.0);
			std::copy(out[k].c.begin(), out[k].c.end(), c.begin());
			pose_conf.insert(pose_conf.end(), c.begin(), c.end());
			pose_ref.push_back({j, k});
		}
	}
	// refine_structure of every pose (do_search, src/main/main.cpp:478-482)
	const int np = int(pose_model.size());
	if (np > 0)
	{
		std::vector<double> pe(np), px(size_t(np) * hstride);
		std::vector<int> ptries(np, 0);
		sgpu_refine_request rq{};
		rq.batch_id = batch;
		rq.receptor_id = receptor_;
		rq.gd = &gd_;
		for (int k = 0; k < 3; k++)
			rq.v[k] = s_.force_cap;
		rq.accurate_line_search = s_.accurate_line ? 1 : 0;
		rq.early_term = s_.early_term ? 1 : 0;
		rq.nposes = np;
		rq.model = pose_model.data();
		rq.conf = pose_conf.data();
		rq.conf_stride = cstride;
		rq.maxiters = pose_iters.data();
		rq.e = pe.data();
		rq.coords = px.data();
		rq.coord_stride = hstride;
		rq.tries = ptries.data();
		check(ctx_, sgpu_run_refine(ctx_, &rq));
		for (int i = 0; i < np; i++)
		{
			Pose& p = jobs[pose_ref[i].first]->out[pose_ref[i].second];
			const FlatModel& f = fms[pose_ref[i].first];
			p.c.assign(pose_conf.begin() + size_t(i) * cstride, pose_conf.begin() + size_t(i) * cstride + f.view.conf_size);
			p.e = pe[i];
			p.e_search = pe[i];
			p.tries = ptries[i];
			p.coords.assign(px.begin() + size_t(i) * hstride, px.begin() + size_t(i) * hstride + 3 * f.nheavy);
		}
	}
	// finish_docking (src/main/main.cpp:256-324): FP64 rescoring on the device
	std::vector<int> best_model;
	std::vector<double> best_conf;
	std::vector<size_t> with_poses;
	for (size_t j = 0; j < nj; j++)
	{
		auto& out = jobs[j]->out;
		std::stable_sort(out.begin(), out.end(), [](const Pose& a, const Pose& b) { return a.e < b.e; });
		if (out.empty())
			continue;
		with_poses.push_back(j);
		best_model.push_back(int(j));
		std::vector<double> c(cstride, 0.0);
		std::copy(out[0].c.begin(), out[0].c.end(), c.begin());
		best_conf.insert(best_conf.end(), c.begin(), c.end());
	}
	if (!with_poses.empty())
	{
		const double v[3] = {s_.force_cap, s_.force_cap, s_.force_cap};
		std::vector<double> intra(with_poses.size());
		sgpu_score_request sq{};
		sq.batch_id = batch;
		sq.receptor_id = receptor_;
		sq.gd = &gd_;
		sq.slope = 1e6;
		for (int k = 0; k < 3; k++)
			sq.v[k] = v[k];
		sq.mode = 1;
		sq.nconfs = int(with_poses.size());
		sq.model = best_model.data();
		sq.conf = best_conf.data();
		sq.conf_stride = cstride;
		sq.e = intra.data();
		check(ctx_, sgpu_score(ctx_, &sq));
		std::vector<int> smodel;
		std::vector<double> sconf;
		std::vector<std::pair<size_t, size_t>> sref;
		for (size_t j : with_poses)
			for (size_t k = 0; k < jobs[j]->out.size(); k++)
				if (jobs[j]->out[k].e < 0.1 * kMax)
				{
					smodel.push_back(int(j));
					std::vector<double> c(cstride, 0.0);
					std::copy(jobs[j]->out[k].c.begin(), jobs[j]->out[k].c.end(), c.begin());
					sconf.insert(sconf.end(), c.begin(), c.end());
					sref.push_back({j, k});
				}
		if (!smodel.empty())
		{
			std::vector<double> se(smodel.size());
			sq.mode = 0;
			sq.nconfs = int(smodel.size());
			sq.model = smodel.data();
			sq.conf = sconf.data();
			sq.e = se.data();
			check(ctx_, sgpu_score(ctx_, &sq));
			std::vector<double> intra_of(nj, 0);
			for (size_t q = 0; q < with_poses.size(); q++)
				intra_of[with_poses[q]] = intra[q];
			for (size_t i = 0; i < smodel.size(); i++)
			{
				const size_t j = sref[i].first;
				jobs[j]->out[sref[i].second].e = conf_independent(sf_, fms[j].num_tors, se[i] - intra_of[j]);
			}
		}
	}
	for (size_t j = 0; j < nj; j++)
	{
		auto& out = jobs[j]->out;
		std::stable_sort(out.begin(), out.end(), [](const Pose& a, const Pose& b) { return a.e < b.e; });
		// remove_redundant with --min_rmsd_filter (src/main/main.cpp:204-210, :276)
		std::vector<Pose> tmp;
		for (const Pose& p : out)
			add_to_output_container(tmp, p, s_.out_min_rmsd, out.size());
		// the reported modes: num_modes within energy_range of the best
		std::vector<Pose> fin;
		for (size_t i = 0; i < tmp.size(); i++)
		{
			if (int(i) >= s_.num_modes || !(tmp[i].e < 0.1 * kMax) || tmp[i].e > tmp[0].e + s_.energy_range)
				break;
			fin.push_back(tmp[i]);
		}
		out.swap(fin);
	}
	fp64_errors(jobs, nj, batch, cstride);
	sgpu_release_models(ctx_, batch);
	const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	for (size_t j = 0; j < nj; j++)
		jobs[j]->seconds = secs / double(nj);
}

} // namespace sgpu_host

// ---------------------------------------------------------------- C wrappers
using namespace sgpu_host;

struct sgpu_flat
{
	FlatModel m;
};

struct sgpu_docking
{
	std::unique_ptr<GpuDocking> d;
	std::string err;
	std::vector<sgpu_term> terms;
	std::vector<sgpu_ci_term> ci;
};

// custom_terms::add (src/lib/custom_terms.cpp:17-37) over the creators of
// src/lib/everything.h:1006-1037: a term's name is its parameterized form,
// "name(k1=v1,_k2=v2,..)" with the keys in a fixed order, or a bare name
// for the conf-independent terms.
namespace sgpu_host
{
namespace
{
struct TermForm
{
	const char* name;
	int kind;
	std::vector<std::string> keys;
	bool charge_dependent;
};

const std::vector<TermForm>& term_forms()
{
	static const std::vector<TermForm> f = {
			{"electrostatic", SGPU_TERM_ELECTROSTATIC, {"i", "^", "c"}, true},
			{"ad4_solvation", SGPU_TERM_AD4_SOLVATION, {"d-sigma", "s/q", "c"}, true},
			{"gauss", SGPU_TERM_GAUSS, {"o", "w", "c"}, false},
			{"repulsion", SGPU_TERM_REPULSION, {"o", "c"}, false},
			{"hydrophobic", SGPU_TERM_HYDROPHOBIC, {"g", "b", "c"}, false},
			{"non_hydrophobic", SGPU_TERM_NON_HYDROPHOBIC, {"g", "b", "c"}, false},
			{"vdw", SGPU_TERM_VDW, {"i", "j", "s", "^", "c"}, false},
			{"non_dir_h_bond_lj", SGPU_TERM_NON_DIR_H_BOND_LJ, {"o", "^", "c"}, false},
			{"non_dir_anti_h_bond_quadratic", SGPU_TERM_NON_DIR_ANTI_H_BOND_QUADRATIC, {"o", "c"}, false},
			{"non_dir_h_bond", SGPU_TERM_NON_DIR_H_BOND, {"g", "b", "c"}, false},
			{"acceptor_acceptor_quadratic", SGPU_TERM_ACCEPTOR_ACCEPTOR_QUADRATIC, {"o", "c"}, false},
			{"donor_donor_quadratic", SGPU_TERM_DONOR_DONOR_QUADRATIC, {"o", "c"}, false},
	};
	return f;
}

// boost::lexical_cast<fl>: the whole string is one number
double whole_number(const std::string& v, const std::string& term)
{
	char* end = nullptr;
	const double x = v.empty() ? 0 : std::strtod(v.c_str(), &end);
	if (v.empty() || end != v.c_str() + v.size())
		throw std::invalid_argument(term + ": could not convert parameters");
	return x;
}
} // namespace

void parse_custom_scoring(const std::string& text, double scale, std::vector<sgpu_term>& ind,
		std::vector<sgpu_term>& dep, std::vector<sgpu_ci_term>& ci)
{
	static const std::vector<std::pair<std::string, int>> ci_forms = {{"num_tors_add", SGPU_CI_NUM_TORS_ADD},
			{"num_tors_sqr", SGPU_CI_NUM_TORS_SQR}, {"num_tors_sqrt", SGPU_CI_NUM_TORS_SQRT},
			{"num_tors_div", SGPU_CI_NUM_TORS_DIV}, {"num_tors_div_simple", SGPU_CI_NUM_TORS_DIV_SIMPLE},
			{"constant_term", SGPU_CI_CONSTANT}};
	static const char* refused[] = {"atom_type_gaussian", "atom_type_linear", "atom_type_quadratic",
			"atom_type_inverse_power", "ligand_length", "num_ligands", "num_heavy_atoms_div", "num_heavy_atoms",
			"num_hydrophobic_atoms"};
	std::istringstream in(text);
	std::string line;
	int lineno = 0;
	while (std::getline(in, line))
	{
		++lineno;
		if (line.empty() || line[0] == '#')
			continue;
		std::istringstream ls(line);
		double w = 0;
		std::string name;
		if (!(ls >> w >> name) || name.empty())
			continue;
		const std::string where = "line " + std::to_string(lineno) + ": " + name;
		bool done = false;
		for (const auto& c : ci_forms)
			if (name == c.first)
			{
				ci.push_back({c.second, w * scale});
				done = true;
				break;
			}
		if (done)
			continue;
		const size_t open = name.find('(');
		const std::string base = name.substr(0, open);
		for (const char* r : refused)
			if (base == r)
				throw std::invalid_argument(where + ": a term this build does not evaluate");
		const TermForm* form = nullptr;
		for (const TermForm& f : term_forms())
			if (base == f.name)
				form = &f;
		if (!form || open == std::string::npos || name.back() != ')')
			throw std::invalid_argument(where + ": unknown term");
		// "k1=v1,_k2=v2,..": each value runs to the next ",_<key>=" or the end
		const std::string inner = name.substr(open + 1, name.size() - open - 2);
		sgpu_term t{form->kind, {0, 0, 0, 0, 0}, w * scale};
		size_t pos = 0;
		for (size_t k = 0; k < form->keys.size(); k++)
		{
			const std::string lead = (k == 0 ? "" : ",_") + form->keys[k] + "=";
			if (inner.compare(pos, lead.size(), lead) != 0)
				throw std::invalid_argument(where + ": unknown term");
			pos += lead.size();
			size_t end = inner.size();
			if (k + 1 < form->keys.size())
			{
				end = inner.find(",_" + form->keys[k + 1] + "=", pos);
				if (end == std::string::npos)
					throw std::invalid_argument(where + ": unknown term");
			}
			t.p[k] = whole_number(inner.substr(pos, end - pos), where);
			pos = end;
		}
		if (form->kind == SGPU_TERM_ELECTROSTATIC && t.p[0] != 1 && t.p[0] != 2)
			throw std::invalid_argument(where + ": invalid exponent (1 or 2 only)");
		if (form->kind == SGPU_TERM_VDW && !((t.p[0] == 4 && t.p[1] == 8) || (t.p[0] == 6 && t.p[1] == 12)))
			throw std::invalid_argument(where + ": unsupported LJ exponents (4,8 or 6,12)");
		(form->charge_dependent ? dep : ind).push_back(t);
	}
}
} // namespace sgpu_host

extern "C" {

int sgpu_builtin_scoring(const char* name, double hydrophobic_scale, sgpu_term* t, sgpu_ci_term* ci,
		sgpu_scoring* out)
{ // builtin_scoring (src/lib/builtinscoring.cpp:65-113), terms in its order
	const std::string n = name ? name : "vina";
	int k = 0, c = 0;
	out->atom_params = SGPU_ATOM_PARAMS_DEFAULT;
	if (n == "vina" || n == "default")
	{ // src/main/main.cpp:1347-1353
		t[k++] = {SGPU_TERM_GAUSS, {0, 0.5, 8, 0, 0}, -0.035579};
		t[k++] = {SGPU_TERM_GAUSS, {3, 2, 8, 0, 0}, -0.005156};
		t[k++] = {SGPU_TERM_REPULSION, {0, 8, 0, 0, 0}, 0.840245};
		t[k++] = {SGPU_TERM_HYDROPHOBIC, {0.5, 1.5, 8, 0, 0}, -0.035069 * hydrophobic_scale};
		t[k++] = {SGPU_TERM_NON_DIR_H_BOND, {-0.7, 0, 8, 0, 0}, -0.587439};
		ci[c++] = {SGPU_CI_NUM_TORS_DIV, 5 * 0.05846 / 0.1 - 1};
	}
	else if (n == "vinardo")
	{
		t[k++] = {SGPU_TERM_GAUSS, {0, 0.8, 8, 0, 0}, -0.045};
		t[k++] = {SGPU_TERM_REPULSION, {0, 8, 0, 0, 0}, 0.80};
		t[k++] = {SGPU_TERM_HYDROPHOBIC, {0.0, 2.5, 8, 0, 0}, -0.035 * hydrophobic_scale};
		t[k++] = {SGPU_TERM_NON_DIR_H_BOND, {-0.6, 0, 8, 0, 0}, -0.60};
		ci[c++] = {SGPU_CI_NUM_TORS_DIV, 5 * 0.02 / 0.1 - 1};
		out->atom_params = SGPU_ATOM_PARAMS_VINARDO;
	}
	else if (n == "dkoes_fast")
	{
		t[k++] = {SGPU_TERM_VDW, {4, 8, 0, 100, 8}, 0.008962};
		t[k++] = {SGPU_TERM_NON_DIR_H_BOND, {-0.7, 0, 8, 0, 0}, 0.387739};
		ci[c++] = {SGPU_CI_NUM_TORS_SQR, .285035};
		ci[c++] = {SGPU_CI_CONSTANT, -2.467357};
	}
	else if (n == "dkoes_scoring_old")
	{
		t[k++] = {SGPU_TERM_VDW, {4, 8, 0, 100, 8}, 0.010607};
		t[k++] = {SGPU_TERM_NON_DIR_H_BOND, {-0.7, 0, 8, 0, 0}, 0.197201};
		ci[c++] = {SGPU_CI_NUM_TORS_SQR, .285035};
		ci[c++] = {SGPU_CI_CONSTANT, -2.585651};
	}
	else if (n == "dkoes_scoring")
	{ // with charge-dependent desolvation: four-component tables
		t[k++] = {SGPU_TERM_VDW, {4, 8, 0, 100, 8}, 0.009900};
		t[k++] = {SGPU_TERM_NON_DIR_H_BOND, {-0.7, 0, 8, 0, 0}, -0.153055};
		t[k++] = {SGPU_TERM_AD4_SOLVATION, {3.6, 0.01097, 8, 0, 0}, 0.048934};
		ci[c++] = {SGPU_CI_NUM_TORS_SQR, 0.317267};
		ci[c++] = {SGPU_CI_CONSTANT, -2.469020};
	}
	else if (n == "ad4_scoring")
	{
		t[k++] = {SGPU_TERM_VDW, {6, 12, 0, 100, 8}, 0.1560};
		t[k++] = {SGPU_TERM_NON_DIR_H_BOND_LJ, {-0.7, 100, 8, 0, 0}, -0.0974};
		t[k++] = {SGPU_TERM_AD4_SOLVATION, {3.5, 0.01097, 8, 0, 0}, 0.1159};
		t[k++] = {SGPU_TERM_ELECTROSTATIC, {1, 100, 8, 0, 0}, 0.1465};
		ci[c++] = {SGPU_CI_NUM_TORS_ADD, 0.2744};
	}
	else
		return -1;
	out->nterms = k;
	out->terms = t;
	out->nci = c;
	out->ci = ci;
	return 0;
}

int sgpu_parse_custom_scoring(const char* text, double scale, sgpu_term* terms, int max_terms, sgpu_ci_term* ci,
		int max_ci, sgpu_scoring* out, char* err, int err_len)
{
	try
	{
		if (!text || !terms || !ci || !out)
			throw std::invalid_argument("custom scoring: missing input");
		std::vector<sgpu_term> ind, dep;
		std::vector<sgpu_ci_term> cis;
		parse_custom_scoring(text, scale, ind, dep, cis);
		if (int(ind.size() + dep.size()) > max_terms || int(cis.size()) > max_ci)
			throw std::invalid_argument("custom scoring: more terms than the arrays hold");
		int k = 0;
		for (const sgpu_term& t : ind)
			terms[k++] = t;
		for (const sgpu_term& t : dep)
			terms[k++] = t;
		for (size_t i = 0; i < cis.size(); i++)
			ci[i] = cis[i];
		out->nterms = k;
		out->terms = terms;
		out->nci = int(cis.size());
		out->ci = ci;
		out->atom_params = SGPU_ATOM_PARAMS_DEFAULT;
		return 0;
	}
	catch (std::exception& e)
	{
		if (err && err_len > 0)
			std::snprintf(err, size_t(err_len), "%s", e.what());
		return -1;
	}
}

double sgpu_conf_independent(const sgpu_scoring* sf, double num_tors, double e)
{
	return conf_independent(*sf, num_tors, e);
}

void sgpu_vina_scoring(sgpu_term* t, sgpu_ci_term* ci, sgpu_scoring* out, double hydrophobic_scale)
{
	sgpu_builtin_scoring("vina", hydrophobic_scale, t, ci, out);
}

int sgpu_build_splines(const sgpu_scoring* sf, double factor, double* coef, double* cut)
{
	try
	{
		Splines s = build_splines(*sf, factor);
		if (coef)
			std::copy(s.coef.begin(), s.coef.end(), coef);
		if (cut)
			std::copy(s.cut.begin(), s.cut.end(), cut);
		return s.nseg;
	}
	catch (...)
	{
		return -1;
	}
}

int sgpu_build_qtables(const sgpu_scoring* sf, double factor, double* qsmooth, double* qfast)
{
	try
	{
		Tables t = build_tables(*sf, factor);
		if (qsmooth)
			std::copy(t.qsmooth.begin(), t.qsmooth.end(), qsmooth);
		if (qfast)
			std::copy(t.qfast.begin(), t.qfast.end(), qfast);
		return t.ncomp;
	}
	catch (...)
	{
		return -1;
	}
}

int sgpu_build_tables(const sgpu_scoring* sf, double factor, double* smooth, double* fast)
{
	try
	{
		Tables t = build_tables(*sf, factor);
		if (smooth)
			std::copy(t.smooth.begin(), t.smooth.end(), smooth);
		if (fast)
			std::copy(t.fast.begin(), t.fast.end(), fast);
		return t.n;
	}
	catch (...)
	{
		return -1;
	}
}

sgpu_flat* sgpu_flatten(const sgpu_model_desc* d)
{
	try
	{
		sgpu_flat* f = new sgpu_flat;
		f->m = flatten(*d);
		f->m.publish();
		return f;
	}
	catch (...)
	{
		return nullptr;
	}
}
const sgpu_flat_model* sgpu_flat_get(const sgpu_flat* f) { return &f->m.view; }
void sgpu_flat_free(sgpu_flat* f) { delete f; }
void sgpu_initial_conf(const sgpu_flat* f, double* conf)
{
	std::vector<double> c = initial_conf(f->m);
	std::copy(c.begin(), c.end(), conf);
}
void sgpu_make_chain(const sgpu_flat* f, uint32_t seed, const sgpu_grid_dims* gd, double* conf, uint32_t* rng_state)
{
	make_chain(f->m, seed, *gd, conf, rng_state);
}

sgpu_docking* sgpu_docking_create(const sgpu_dock_settings* s, const sgpu_mol* receptor, const sgpu_grid_dims* gd)
{
	sgpu_docking* d = new sgpu_docking;
	try
	{
		GpuSettings gs;
		gs.device = s->device;
		if (s->precision != SGPU_FP32 && s->precision != SGPU_FP64 && s->precision != SGPU_MIXED)
			throw std::invalid_argument("precision must be SGPU_FP32, SGPU_FP64 or SGPU_MIXED");
		gs.precision = s->precision;
		gs.exhaustiveness = s->exhaustiveness;
		gs.seed = s->seed;
		if (s->num_modes < 1 || s->num_modes > 64)
			throw std::invalid_argument("num_modes must be in 1..64");
		gs.num_modes = s->num_modes;
		gs.energy_range = s->energy_range;
		if (s->force_cap < 0)
			throw std::invalid_argument("force_cap must be positive (0: 1000)");
		gs.force_cap = s->force_cap > 0 ? s->force_cap : 1000;
		gs.accurate_line = s->accurate_line != 0;
		if (s->min_rmsd_filter < 0)
			throw std::invalid_argument("min_rmsd_filter must be positive (0: 1)");
		gs.out_min_rmsd = s->min_rmsd_filter > 0 ? s->min_rmsd_filter : 1;
		if (s->minimize_iters < 0)
			throw std::invalid_argument("minimize_iters must be positive (0: default)");
		gs.minimize_iters = s->minimize_iters;
		gs.early_term = s->minimize_early_term != 0;
		gs.ligands_per_launch = s->ligands_per_launch;
		if (!(s->step_fraction > 0) || s->step_fraction > 1)
			throw std::invalid_argument("step_fraction must be in (0, 1]");
		gs.step_fraction = s->step_fraction;
		if (s->chain_split < 0)
			throw std::invalid_argument("chain_split must be at least 1 (0: default)");
		gs.chain_split = s->chain_split > 0 ? s->chain_split : 1;
		gs.early_stop = s->early_stop ? GpuSettings::EARLY_STOP_ALWAYS : GpuSettings::EARLY_STOP_OFF;
		gs.skip_visited = s->skip_visited != 0;
		if (s->approximation != 0 && s->approximation != 1)
			throw std::invalid_argument("approximation must be 0 (linear) or 1 (spline)");
		gs.approximation = s->approximation == 1 ? GpuSettings::SPLINE : GpuSettings::LINEAR;
		if (s->approx_factor < 0)
			throw std::invalid_argument("approx_factor must be positive (0: default)");
		gs.approx_factor = s->approx_factor > 0 ? s->approx_factor : (s->approximation == 1 ? 10 : 32);
		if (s->fast)
		{
			// explicit chain_split / step_fraction take precedence
			apply_fast(gs);
			if (s->chain_split > 0)
				gs.chain_split = s->chain_split;
			if (s->step_fraction < 1)
				gs.step_fraction = s->step_fraction;
		}
		gs.progress_chunks = s->progress_chunks > 0 ? s->progress_chunks : 1;
		gs.progress = s->progress;
		gs.progress_user = s->progress_user;
		gs.user_grid = s->user_grid;
		sgpu_scoring sf;
		if (s->custom_scoring)
		{
			std::vector<sgpu_term> ind, dep;
			std::vector<sgpu_ci_term> cis;
			parse_custom_scoring(s->custom_scoring, 1.0, ind, dep, cis);
			d->terms = ind;
			d->terms.insert(d->terms.end(), dep.begin(), dep.end());
			d->ci = cis;
			if (d->terms.empty())
				throw std::invalid_argument("custom scoring: no pairwise terms");
			sf.nterms = int(d->terms.size());
			sf.terms = d->terms.data();
			sf.nci = int(d->ci.size());
			sf.ci = d->ci.data();
			sf.atom_params = SGPU_ATOM_PARAMS_DEFAULT;
		}
		else
		{
			d->terms.resize(SGPU_MAX_TERMS);
			d->ci.resize(SGPU_MAX_CI);
			const int rc = sgpu_builtin_scoring(s->scoring, s->hydrophobic_scale, d->terms.data(), d->ci.data(), &sf);
			if (rc != 0)
				throw std::invalid_argument("unknown scoring function");
		}
		const sgpu_model_desc* rd = sgpu_mol_desc(receptor);
		sgpu_receptor rec{rd->natoms, rd->xyz, rd->type, rd->charge};
		d->d.reset(new GpuDocking(gs, sf, rec, *gd));
	}
	catch (std::exception& e)
	{
		d->err = e.what();
	}
	return d;
}
void sgpu_docking_free(sgpu_docking* d) { delete d; }
int sgpu_docking_full_searches(const sgpu_docking* d) { return d->d ? int(d->d->full_searches()) : 0; }

double sgpu_full_search_cutoff(const double* history, int n, double fast_top, double margin)
{
	if (!history || n <= 0 || !(fast_top > 0) || fast_top > 1)
		return std::numeric_limits<double>::quiet_NaN();
	return full_search_cutoff(std::vector<double>(history, history + n), fast_top, margin);
}

int sgpu_search_plan(int chain_split, double step_fraction, int early_stop, int tasks, int capacity, double budget,
		int* split, double* fraction, int* checkpoints)
{
	if (chain_split < 0 || step_fraction < 0 || early_stop < 0 || early_stop > 2 || tasks < 0 || !split ||
			!fraction || !checkpoints)
		return -1;
	GpuSettings s;
	s.chain_split = chain_split;
	s.step_fraction = step_fraction;
	s.early_stop = early_stop;
	const SearchPlan p = plan_search(s, size_t(tasks), capacity, budget);
	*split = p.split;
	*fraction = p.fraction;
	*checkpoints = p.checkpoints;
	return 0;
}
const char* sgpu_docking_error(const sgpu_docking* d) { return d->err.c_str(); }
sgpu_ctx* sgpu_docking_ctx(sgpu_docking* d) { return d->d ? d->d->ctx() : nullptr; }
int sgpu_docking_receptor_id(sgpu_docking* d) { return d->d ? d->d->receptor_id() : -1; }

int sgpu_docking_prepare(sgpu_docking* d, const sgpu_mol* const* ligands, int n)
{
	if (!d->d)
		return -1;
	try
	{
		std::vector<const sgpu_model_desc*> ms;
		for (int i = 0; i < n; i++)
			ms.push_back(sgpu_mol_desc(ligands[i]));
		d->d->prepare(ms);
		return 0;
	}
	catch (std::exception& e)
	{
		d->err = e.what();
		return -1;
	}
}

int sgpu_docking_tune(sgpu_docking* d, const sgpu_mol* const* ligands, int n, int probe_steps,
		sgpu_mc_tune_report* report)
{
	if (!d->d || !report)
		return -1;
	try
	{
		std::vector<GpuDocking::DockJob> jobs(n);
		std::vector<GpuDocking::DockJob*> jp;
		for (int i = 0; i < n; i++)
		{
			jobs[i].m = sgpu_mol_desc(ligands[i]);
			jp.push_back(&jobs[i]);
		}
		d->d->tune(jp, probe_steps, *report);
		return 0;
	}
	catch (std::exception& e)
	{
		d->err = e.what();
		return -1;
	}
}

int sgpu_dock(sgpu_docking* d, const sgpu_mol* const* ligands, int n, sgpu_dock_result* out)
{
	if (!d->d)
		return -1;
	try
	{
		std::vector<GpuDocking::DockJob> jobs(n);
		std::vector<GpuDocking::DockJob*> jp;
		for (int i = 0; i < n; i++)
		{
			jobs[i].m = sgpu_mol_desc(ligands[i]);
			jp.push_back(&jobs[i]);
		}
		d->d->dock_batch(jp);
		for (int i = 0; i < n; i++)
		{
			out[i].nposes = int(std::min<size_t>(32, jobs[i].out.size()));
			for (int k = 0; k < out[i].nposes; k++)
				out[i].e[k] = jobs[i].out[k].e;
			out[i].evals = jobs[i].evals;
			out[i].seconds = jobs[i].seconds;
			out[i].fp64_error = jobs[i].fp64_error;
			out[i].e_search = jobs[i].e_search;
			out[i].e_fp64 = jobs[i].e_fp64;
			out[i].conf_size = 0;
			if (out[i].confs && !jobs[i].out.empty())
			{
				const size_t cs = jobs[i].out[0].c.size();
				out[i].conf_size = int(cs);
				for (int k = 0; k < out[i].nposes; k++)
					if (size_t(k + 1) * cs <= size_t(std::max(0, out[i].confs_capacity)))
						std::copy(jobs[i].out[size_t(k)].c.begin(), jobs[i].out[size_t(k)].c.end(),
								out[i].confs + size_t(k) * cs);
			}
			out[i].ncoords = 0;
			if (out[i].coords && !jobs[i].out.empty())
			{
				const std::vector<double>& x = jobs[i].out[0].coords;
				out[i].ncoords = int(std::min<size_t>(x.size(), size_t(std::max(0, out[i].coords_capacity))));
				std::copy(x.begin(), x.begin() + out[i].ncoords, out[i].coords);
			}
		}
		return 0;
	}
	catch (std::exception& e)
	{
		d->err = e.what();
		return -1;
	}
}

int sgpu_autobox(const sgpu_mol* const* mols, int n, double add, double granularity, const double* size,
		sgpu_grid_dims* out)
{
	if (!mols || n < 1 || !out || !(granularity > 0) || !(add >= 0))
		return -1;
	double lo[3], hi[3];
	bool any = false;
	for (int k = 0; k < n; k++)
	{
		const sgpu_model_desc* m = mols[k] ? sgpu_mol_desc(mols[k]) : nullptr;
		if (!m)
			return -1;
		for (int a = 0; a < m->natoms; a++)
			for (int d = 0; d < 3; d++)
			{
				const double x = m->xyz[3 * a + d];
				lo[d] = any ? std::min(lo[d], x) : x;
				hi[d] = any ? std::max(hi[d], x) : x;
				if (d == 2)
					any = true;
			}
	}
	if (!any)
		return -1;
	for (int d = 0; d < 3; d++)
	{
		const double center = (hi[d] + lo[d]) / 2.0;
		const double span = size && size[d] > 0 ? size[d] : (hi[d] + add) - (lo[d] - add);
		out->n[d] = int(std::ceil(span / granularity));
		const double real_span = granularity * out->n[d];
		out->begin[d] = center - real_span / 2;
		out->end[d] = out->begin[d] + real_span;
	}
	return 0;
}

int sgpu_minimize(sgpu_docking* d, const sgpu_mol* const* ligands, int n, int maxiters, int accurate,
		double force_cap, sgpu_minimize_result* out)
{
	if (!d->d)
		return -1;
	try
	{
		std::vector<GpuDocking::MinimizeJob> jobs(n);
		std::vector<GpuDocking::MinimizeJob*> jp;
		for (int i = 0; i < n; i++)
		{
			jobs[i].m = sgpu_mol_desc(ligands[i]);
			jobs[i].maxiters = maxiters;
			jobs[i].accurate = accurate != 0;
			jp.push_back(&jobs[i]);
		}
		d->d->minimize_batch(jp, force_cap);
		for (int i = 0; i < n; i++)
		{
			out[i].e = jobs[i].e;
			out[i].rmsd = jobs[i].rmsd;
			for (int k = 0; k < 3; k++)
			{
				out[i].box_begin[k] = jobs[i].gd.begin[k];
				out[i].box_end[k] = jobs[i].gd.end[k];
			}
		}
		return 0;
	}
	catch (std::exception& e)
	{
		d->err = e.what();
		return -1;
	}
}

} // extern "C"

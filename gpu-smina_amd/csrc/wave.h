/*
 * Wave64 helpers and the smina geometry kernels in device form, templated on
 * the arithmetic type T (float for the served search, double for the
 * oracle-parity search).  Expressions keep smina's operand order
 * (src/lib/common.h vec / mat, src/lib/quaternion.cpp, src/lib/curl.h) so the
 * double instantiation, built with -ffp-contract=off, reproduces the CPU
 * bit for bit.
 */
#ifndef SGPU_WAVE_H
#define SGPU_WAVE_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include "pmath.h"
#include "layout.h"

namespace sgpu
{

__device__ inline int lane_id() { return int(threadIdx.x) & (WAVE - 1); }
__device__ inline int wave_in_block() { return int(threadIdx.x) / WAVE; }

// Syncs inside one wavefront; waves never wait for each other.
// lsync: LDS written by some lanes is read by others next.  A wavefront's
// LDS instructions execute in order, so this only has to stop the compiler
// from reordering them (a wavefront-scope fence emits no wait).
__device__ inline void lsync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// gsync: global memory written by some lanes is read by others next: wait
// for the stores (workgroup scope: the same CU's L1 serves the reads)
__device__ inline void gsync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ inline float readlane(float v, int l)
{
	return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ inline double readlane(double v, int l)
{
	const uint64_t u = __builtin_bit_cast(uint64_t, v);
	const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u)), l));
	const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u >> 32)), l));
	return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
__device__ inline int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// DPP lane exchange within rows of 16 (no LDS traffic)
template<int CTRL>
__device__ inline float dpp(float v)
{
	return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
constexpr int DPP_QUAD_1032 = 0xB1;    // quad_perm [1, 0, 3, 2]
constexpr int DPP_QUAD_2301 = 0x4E;    // quad_perm [2, 3, 0, 1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_ROW_MIRROR = 0x140;

// wave-wide sum in a fixed order (deterministic), every lane gets it: DPP
// within each row of 16 lanes, then the four row sums.  The whole wave must
// be active.
__device__ inline float wave_sum(float v)
{
	v += dpp<DPP_QUAD_1032>(v);
	v += dpp<DPP_QUAD_2301>(v);
	v += dpp<DPP_ROW_HALF_MIRROR>(v);
	v += dpp<DPP_ROW_MIRROR>(v);
	return (__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0)) +
				   __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16))) +
			(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32)) +
					__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48)));
}
// DPP with row and bank write masks; lanes masked off, and sources outside
// the row when bound_ctrl is set, give 0
template<int CTRL, int ROW_MASK, int BANK_MASK, bool BOUND_ZERO>
__device__ inline float dpp_masked(float v)
{
	return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK,
											 BANK_MASK, BOUND_ZERO));
}
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR3 = 0x113, DPP_ROW_SHR4 = 0x114,
			  DPP_ROW_SHR8 = 0x118, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

// inclusive prefix sum over the 64 lanes: sums of 4 within each row from
// the input, widened to 8 and 16 with bank masks, then the row totals
// carried into the next rows (row_bcast).  The whole wave must be active.
__device__ inline float wave_inclusive_scan(float v)
{
	float s = v + dpp_masked<DPP_ROW_SHR1, 0xF, 0xF, true>(v);
	s += dpp_masked<DPP_ROW_SHR2, 0xF, 0xF, true>(v);
	s += dpp_masked<DPP_ROW_SHR3, 0xF, 0xF, true>(v);
	s += dpp_masked<DPP_ROW_SHR4, 0xF, 0xE, false>(s);
	s += dpp_masked<DPP_ROW_SHR8, 0xF, 0xC, false>(s);
	s += dpp_masked<DPP_ROW_BCAST15, 0xA, 0xF, false>(s);
	s += dpp_masked<DPP_ROW_BCAST31, 0xC, 0xF, false>(s);
	return s;
}

// double: butterfly through lane permutes (the ordered search never uses it)
__device__ inline double wave_sum(double v)
{
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off, WAVE);
	return v;
}
__device__ inline bool wave_any(bool p) { return __ballot(p) != 0; }
__device__ inline bool wave_all(bool p) { return __ballot(!p) == 0; }

template<typename T> struct Real;
template<> struct Real<float>
{
	static constexpr bool ordered = false;
	// hardware sine/cosine: every argument the search passes is a normalised
	// half angle, |x| <= pi / 2
	__device__ static float sin(float x) { return __sinf(x); }
	__device__ static float cos(float x) { return __cosf(x); }
	__device__ static float acos(float x) { return ::acosf(x); }
	__device__ static float sqrt(float x) { return ::sqrtf(x); }
	__device__ static float abs(float x) { return ::fabsf(x); }
	__device__ static float ceil(float x) { return ::ceilf(x); }
	__device__ static float eps() { return FLT_EPSILON; }
	__device__ static float max() { return FLT_MAX; }
	// a / b, 1 / x and 1 / sqrt(x) through the hardware reciprocal and
	// reciprocal square root (1 ulp), for the served search
	__device__ static float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
	__device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
	__device__ static float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
};
template<> struct Real<double>
{
	static constexpr bool ordered = true;
	__device__ static double sin(double x) { return pmath::sin(x); }
	__device__ static double cos(double x) { return pmath::cos(x); }
	__device__ static double acos(double x) { return pmath::acos(x); }
	__device__ static double sqrt(double x) { return __builtin_sqrt(x); }
	__device__ static double abs(double x) { return __builtin_fabs(x); }
	__device__ static double ceil(double x) { return __builtin_ceil(x); }
	__device__ static double eps() { return DBL_EPSILON; }
	__device__ static double max() { return DBL_MAX; }
	__device__ static double fdiv(double a, double b) { return a / b; }
	__device__ static double rcp(double x) { return 1 / x; }
	__device__ static double rsqrt(double x) { return 1 / __builtin_sqrt(x); }
};

template<typename T>
struct V3
{
	T x, y, z;
};
template<typename T> __device__ inline V3<T> v3(T x, T y, T z) { return {x, y, z}; }
template<typename T> __device__ inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template<typename T> __device__ inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template<typename T> __device__ inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template<typename T> __device__ inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template<typename T> __device__ inline T sqr(T x) { return x * x; }
template<typename T> __device__ inline T norm_sqr(V3<T> a) { return sqr(a.x) + sqr(a.y) + sqr(a.z); }
template<typename T> __device__ inline V3<T> cross(V3<T> a, V3<T> b)
{
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template<typename T> __device__ inline V3<T> ld3(const T* p) { return {p[0], p[1], p[2]}; }
template<typename T> __device__ inline void st3(T* p, V3<T> v)
{
	p[0] = v.x;
	p[1] = v.y;
	p[2] = v.z;
}

// column-major 3x3 (m[i + 3j] = M(i, j)); M * v in smina's order
template<typename T> __device__ inline V3<T> mat_vec(const T* m, V3<T> v)
{
	return {m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
			m[2] * v.x + m[5] * v.y + m[8] * v.z};
}

template<typename T>
struct Q4
{
	T a, b, c, d;
};
// Hamilton product in boost::math::quaternion's operand order
template<typename T> __device__ inline Q4<T> qmul(Q4<T> l, Q4<T> r)
{
	return {l.a * r.a - l.b * r.b - l.c * r.c - l.d * r.d, l.a * r.b + l.b * r.a + l.c * r.d - l.d * r.c,
			l.a * r.c - l.b * r.d + l.c * r.a + l.d * r.b, l.a * r.d + l.b * r.c - l.c * r.b + l.d * r.a};
}

// normalize_angle (src/lib/common.h:245-270)
template<typename T> __device__ inline void normalize_angle(T& x)
{
	const T pi = T(3.1415926535897931);
	for (;;)
	{
		if (x > 3 * pi)
		{
			T n = (x - pi) / (2 * pi);
			x -= 2 * pi * Real<T>::ceil(n);
		}
		else if (x < -3 * pi)
		{
			T n = (-x - pi) / (2 * pi);
			x += 2 * pi * Real<T>::ceil(n);
		}
		else
		{
			if (x > pi)
				x -= 2 * pi;
			else if (x < -pi)
				x += 2 * pi;
			return;
		}
	}
}
template<typename T> __device__ inline T normalized_angle(T x)
{
	normalize_angle(x);
	return x;
}

// angle_to_quaternion (src/lib/quaternion.cpp:42-62)
template<typename T> __device__ inline Q4<T> axis_angle_q(V3<T> axis, T angle)
{
	normalize_angle(angle);
	const T c = Real<T>::cos(angle / 2);
	const T s = Real<T>::sin(angle / 2);
	return {c, s * axis.x, s * axis.y, s * axis.z};
}
template<typename T> __device__ inline Q4<T> rotation_q(V3<T> rot)
{
	const T angle = Real<T>::sqrt(norm_sqr(rot));
	if (angle > Real<T>::eps())
	{
		V3<T> axis = Real<T>::rcp(angle) * rot;
		return axis_angle_q(axis, angle);
	}
	return {T(1), T(0), T(0), T(0)};
}
// v rotated by the unit quaternion q (the matrix of q_to_mat times v):
// v + 2a (u x v) + 2 u x (u x v), u the vector part
template<typename T> __device__ inline V3<T> q_rotate(Q4<T> q, V3<T> v)
{
	const T tx = 2 * (q.c * v.z - q.d * v.y), ty = 2 * (q.d * v.x - q.b * v.z), tz = 2 * (q.b * v.y - q.c * v.x);
	return {v.x + q.a * tx + (q.c * tz - q.d * ty), v.y + q.a * ty + (q.d * tx - q.b * tz),
			v.z + q.a * tz + (q.b * ty - q.c * tx)};
}
// quaternion_normalize_approx (src/lib/quaternion.h:69-80)
template<typename T> __device__ inline void q_normalize_approx(Q4<T>& q)
{
	const T s = sqr(q.a) + sqr(q.b) + sqr(q.c) + sqr(q.d);
	if (Real<T>::abs(s - 1) < T(1e-6))
		return;
	const T inv = Real<T>::rsqrt(s);
	q.a *= inv;
	q.b *= inv;
	q.c *= inv;
	q.d *= inv;
}
// quaternion_increment (src/lib/quaternion.cpp:123-127)
template<typename T> __device__ inline Q4<T> q_increment(Q4<T> q, V3<T> rot)
{
	Q4<T> r = qmul(rotation_q(rot), q);
	q_normalize_approx(r);
	return r;
}
// quaternion_to_angle (src/lib/quaternion.cpp:63-79)
template<typename T> __device__ inline V3<T> q_to_angle(Q4<T> q)
{
	const T c = q.a;
	const T pi = T(3.1415926535897931);
	if (c > -1 && c < 1)
	{
		T angle = 2 * Real<T>::acos(c);
		if (angle > pi)
			angle -= 2 * pi;
		V3<T> axis = v3(q.b, q.c, q.d);
		const T s = Real<T>::sin(angle / 2);
		if (Real<T>::abs(s) < Real<T>::eps())
			return v3(T(0), T(0), T(0));
		const T k = angle / s;
		return v3(axis.x * k, axis.y * k, axis.z * k);
	}
	return v3(T(0), T(0), T(0));
}
// quaternion_to_r3 (src/lib/quaternion.cpp:81-107)
template<typename T> __device__ inline void q_to_mat(Q4<T> q, T* m)
{
	const T a = q.a, b = q.b, c = q.c, d = q.d;
	const T aa = a * a, ab = a * b, ac = a * c, ad = a * d;
	const T bb = b * b, bc = b * c, bd = b * d;
	const T cc = c * c, cd = c * d, dd = d * d;
	m[0] = (aa + bb - cc - dd);
	m[3] = 2 * (-ad + bc);
	m[6] = 2 * (ac + bd);
	m[1] = 2 * (ad + bc);
	m[4] = (aa - bb + cc - dd);
	m[7] = 2 * (-ab + cd);
	m[2] = 2 * (-ac + bd);
	m[5] = 2 * (ab + cd);
	m[8] = (aa - bb - cc + dd);
}

// curl (src/lib/curl.h:28-43); every cap the search passes is below max/10
template<typename T> __device__ inline void curl(T& e, V3<T>& d, T v)
{
	if (e > 0)
	{
		T tmp = (v < Real<T>::eps()) ? T(0) : Real<T>::fdiv(v, v + e);
		e *= tmp;
		const T t2 = sqr(tmp);
		d.x *= t2;
		d.y *= t2;
		d.z *= t2;
	}
}
template<typename T> __device__ inline void curl(T& e, T v)
{
	if (e > 0)
	{
		T tmp = (v < Real<T>::eps()) ? T(0) : (v / (v + e));
		e *= tmp;
	}
}

} // namespace sgpu

#endif

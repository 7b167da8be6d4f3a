/*
 * se3_math.h -- deterministic SE(3) state arithmetic shared by host C/C++ and
 * gfx950 device code.
 *
 * Every function here is written so that the x86-64 host build (gcc or clang,
 * -ffp-contract=off) and the gfx950 device build produce bit-identical
 * results: only IEEE-754 binary64 +, -, *, /, sqrt, fabs, floor and
 * comparisons are used, with an explicit association order and FMA
 * contraction disabled.  Division and sqrt are correctly rounded on both
 * targets.  That makes every interpolated segment state computed on the GPU
 * equal, bit for bit, to the same state computed on the CPU.
 *
 * The transcendental functions (sin, cos, acos, atan2) that Eigen calls
 * inside Quaternion::slerp, angularDistance and AngleAxis (lib/osr/
 * osr_state.cc:39, :135, :121-128) are this library's correctly rounded ones
 * (se3_crmath.h): the same bits as the C library of the reference's
 * platform, which rounds them correctly, on the CPU and on the GPU.
 *
 * State layout (lib/osr/osr_state.h:44-50, lib/osr/unit_world.h:44-48):
 *   q = [tx, ty, tz, qw, qx, qy, qz] in binary64.
 *
 * Association orders follow Eigen 3.3 as built for x86-64 with SSE2 (the
 * reference's Release build): 4-element quaternion reductions pair
 * (x,z)+(y,w); 3-element vector reductions run left to right.
 */
#ifndef SE3_MATH_H
#define SE3_MATH_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define SE3_HD __host__ __device__
#else
#define SE3_HD
#endif

#ifdef __clang__
#pragma clang fp contract(off)
#endif

#include "se3_crmath.h"

/* ---------------------------------------------------------------------- */
/* Constants (binary64, hi+lo splits where a second word is needed).      */
/* ---------------------------------------------------------------------- */
#define SE3_PI_HI 0x1.921fb54442d18p+1
#define SE3_PI_LO 0x1.1a62633145c07p-53
#define SE3_PIO2_HI 0x1.921fb54442d18p+0
#define SE3_PIO2_LO 0x1.1a62633145c07p-54
#define SE3_PIO4_HI 0x1.921fb54442d18p-1
#define SE3_PIO4_LO 0x1.1a62633145c07p-55
#define SE3_TWO_OVER_PI 0x1.45f306dc9c883p-1
/* pi/2 = P1 + P2 + P3; P1 and P2 carry 33 significant bits so k*P1 and k*P2
 * are exact for |k| < 2^20. */
#define SE3_RED_P1 0x1.921fb54400000p+0
#define SE3_RED_P2 0x1.0b4611a600000p-34
#define SE3_RED_P3 0x1.3198a2e037073p-69
#define SE3_TAN_PI_8 0x1.a827999fcef32p-2
/* Beyond |x| = 2^20 the three-part reduction loses accuracy (k*P1 is no
 * longer exact); it stays deterministic, which is what the state arithmetic
 * needs, and slerp/apply never produce such arguments. */
#define SE3_RED_MAX 0x1p20

/* sin(r) on |r| <= pi/4 (+ a little): r + r^3 * P(r^2), Taylor to r^21. */
SE3_HD static inline double se3_sin_kernel(double r)
{
	double r2 = r * r;
	double p = 0x1.71b8ef6dcf572p-66;
	p = -0x1.2f49b46814157p-57 + r2 * p;
	p = 0x1.952c77030ad4ap-49 + r2 * p;
	p = -0x1.ae7f3e733b81fp-41 + r2 * p;
	p = 0x1.6124613a86d09p-33 + r2 * p;
	p = -0x1.ae64567f544e4p-26 + r2 * p;
	p = 0x1.71de3a556c734p-19 + r2 * p;
	p = -0x1.a01a01a01a01ap-13 + r2 * p;
	p = 0x1.1111111111111p-7 + r2 * p;
	p = -0x1.5555555555555p-3 + r2 * p;
	return r + (r2 * r) * p;
}

/* cos(r) on |r| <= pi/4 (+ a little): 1 - r^2/2 + r^4 * Q(r^2), Taylor to r^20. */
SE3_HD static inline double se3_cos_kernel(double r)
{
	double r2 = r * r;
	double p = 0x1.e542ba4020225p-62;
	p = -0x1.6827863b97d97p-53 + r2 * p;
	p = 0x1.ae7f3e733b81fp-45 + r2 * p;
	p = -0x1.93974a8c07c9dp-37 + r2 * p;
	p = 0x1.1eed8eff8d898p-29 + r2 * p;
	p = -0x1.27e4fb7789f5cp-22 + r2 * p;
	p = 0x1.a01a01a01a01ap-16 + r2 * p;
	p = -0x1.6c16c16c16c17p-10 + r2 * p;
	p = 0x1.5555555555555p-5 + r2 * p;
	return 1.0 + r2 * (-0.5 + r2 * p);
}

/* Reduce x to r in [-pi/4, pi/4] (approximately) and a quadrant q = k mod 4;
 * |x| <= pi/4 is left as is (k = 0). */
SE3_HD static inline double se3_reduce_pio2(double x, int* quadrant)
{
	double k = floor(x * SE3_TWO_OVER_PI + 0.5);
	if (fabs(x) <= SE3_PIO4_HI)
		k = 0.0;
	double r = ((x - k * SE3_RED_P1) - k * SE3_RED_P2) - k * SE3_RED_P3;
	*quadrant = (int)(k - 4.0 * floor(k * 0.25)); /* k mod 4, also for huge k */
	return r;
}

/* Both kernels are evaluated and the quadrant selects: straight-line code on
 * the GPU (no divergent branches), the same values as a switch. */
SE3_HD static inline double se3_sin_legacy(double x)
{
	if (x != x || fabs(x) == INFINITY)
		return x - x; /* NaN */
	int q;
	const double r = se3_reduce_pio2(x, &q);
	const double sv = se3_sin_kernel(r), cv = se3_cos_kernel(r);
	const double v = (q & 1) ? cv : sv;
	return (q & 2) ? -v : v;
}

SE3_HD static inline double se3_cos_legacy(double x)
{
	if (x != x || fabs(x) == INFINITY)
		return x - x; /* NaN */
	int q;
	const double r = se3_reduce_pio2(x, &q);
	const double sv = se3_sin_kernel(r), cv = se3_cos_kernel(r);
	const double v = (q & 1) ? sv : cv;
	return ((q + 1) & 2) ? -v : v;
}

/* The correctly rounded functions (se3_crmath.h); sin and cos beyond
 * |x| = 2^20 use the reduction above (deterministic on both targets, not
 * correctly rounded: slerp and apply never produce such arguments).
 * SE3_CR_INLINE forces their fast paths inline on the device (the slow
 * paths stay calls); by default the compiler decides. */
#if defined(__HIP_DEVICE_COMPILE__) && defined(SE3_CR_INLINE) && SE3_CR_INLINE
#define SE3_CR_FAST __attribute__((always_inline))
#else
#define SE3_CR_FAST
#endif
SE3_HD static SE3_CR_FAST inline double se3_sin(double x)
{
	double r;
	if (se3_cr_sincos(x, 0, &r))
		return r;
	return se3_sin_legacy(x);
}

SE3_HD static SE3_CR_FAST inline double se3_cos(double x)
{
	double r;
	if (se3_cr_sincos(x, 1, &r))
		return r;
	return se3_cos_legacy(x);
}

SE3_HD static SE3_CR_FAST inline double se3_acos(double x) { return se3_cr_acos(x); }

SE3_HD static SE3_CR_FAST inline double se3_atan2(double y, double x) { return se3_cr_atan2(y, x); }

/* ---------------------------------------------------------------------- */
/* Quaternion and state arithmetic (Eigen 3.3 semantics, see file header). */
/* Quaternions are passed as (w, x, y, z).                                */
/* ---------------------------------------------------------------------- */

/* Eigen QuaternionBase::normalize: coeffs /= sqrt(squaredNorm) if > 0. */
SE3_HD static inline void se3_quat_normalize(double* q)
{
	double w = q[0], x = q[1], y = q[2], z = q[3];
	double n2 = (x * x + z * z) + (y * y + w * w);
	if (n2 > 0.0) {
		double n = sqrt(n2);
		q[0] = w / n;
		q[1] = x / n;
		q[2] = y / n;
		q[3] = z / n;
	}
}

/* Eigen quaternion dot: coeffs (x,y,z,w) reduced as (x,z)+(y,w). */
SE3_HD static inline double se3_quat_dot(const double* a, const double* b)
{
	return (a[1] * b[1] + a[3] * b[3]) + (a[2] * b[2] + a[0] * b[0]);
}

/* Hamilton product a*b with Eigen's SSE2 binary64 grouping. */
SE3_HD static inline void se3_quat_mul(const double* a, const double* b, double* r)
{
	double wa = a[0], xa = a[1], ya = a[2], za = a[3];
	double wb = b[0], xb = b[1], yb = b[2], zb = b[3];
	double rx = (wa * xb + ya * zb) - (za * yb - xa * wb);
	double ry = (wa * yb + ya * wb) + (za * xb - xa * zb);
	double rz = (wa * zb - ya * xb) + (za * wb + xa * yb);
	double rw = (wa * wb - ya * yb) - (za * zb + xa * xb);
	r[0] = rw;
	r[1] = rx;
	r[2] = ry;
	r[3] = rz;
}

/*
 * Eigen QuaternionBase::toRotationMatrix on the quaternion exactly as given
 * (not normalised), row-major 3x3.  translate_state_to_transform
 * (lib/osr/osr_state.cc:16-25) feeds the raw state quaternion to it.
 */
SE3_HD static inline void se3_quat_to_rotmat(double w, double x, double y, double z, double* R)
{
	double tx = 2.0 * x, ty = 2.0 * y, tz = 2.0 * z;
	double twx = tx * w, twy = ty * w, twz = tz * w;
	double txx = tx * x, txy = ty * x, txz = tz * x;
	double tyy = ty * y, tyz = tz * y, tzz = tz * z;
	R[0] = 1.0 - (tyy + tzz);
	R[1] = txy - twz;
	R[2] = txz + twy;
	R[3] = txy + twz;
	R[4] = 1.0 - (txx + tzz);
	R[5] = tyz - twx;
	R[6] = txz - twy;
	R[7] = tyz + twx;
	R[8] = 1.0 - (txx + tyy);
}

/*
 * Box-box separating-axis test of FCL's box shapes (FCL boxBox2, itself ODE's
 * dBoxBox; external), boolean part: box i has rotation Ri (row-major, column
 * k = axis k), centre pi and half extents hi.  Returns 1 when one of the 15
 * axes separates the boxes (|projected centre distance| exceeds the summed
 * projected radii).  The nine edge-edge axes use |R| + 1e-6, the guard those
 * implementations add against near-parallel edges.  Association orders follow
 * the macro expansions of dBoxBox.
 */
SE3_HD static inline int se3_boxes_disjoint(const double* R1, const double* p1, const double* A, const double* R2,
                                            const double* p2, const double* B)
{
	double p[3], pp[3], R[9], Q[9];
	for (int k = 0; k < 3; k++)
		p[k] = p2[k] - p1[k];
	for (int i = 0; i < 3; i++) {
		pp[i] = (R1[i] * p[0] + R1[3 + i] * p[1]) + R1[6 + i] * p[2];
		for (int j = 0; j < 3; j++) {
			R[3 * i + j] = (R1[i] * R2[j] + R1[3 + i] * R2[3 + j]) + R1[6 + i] * R2[6 + j];
			Q[3 * i + j] = fabs(R[3 * i + j]);
		}
	}
	/* axes of box 1 */
	for (int i = 0; i < 3; i++)
		if (fabs(pp[i]) - (((A[i] + B[0] * Q[3 * i]) + B[1] * Q[3 * i + 1]) + B[2] * Q[3 * i + 2]) > 0.0)
			return 1;
	/* axes of box 2 */
	for (int j = 0; j < 3; j++) {
		const double d = (R2[j] * p[0] + R2[3 + j] * p[1]) + R2[6 + j] * p[2];
		if (fabs(d) - (((A[0] * Q[j] + A[1] * Q[3 + j]) + A[2] * Q[6 + j]) + B[j]) > 0.0)
			return 1;
	}
	for (int k = 0; k < 9; k++)
		Q[k] = Q[k] + 1e-6;
	/* u_i x v_j: (i1, i2) are the other two axes of box 1 and (j1, j2) of
	 * box 2, each pair in increasing order, which is the term order of the
	 * dBoxBox listing (A[i1] before A[i2], B[j1] before B[j2]); the sign of
	 * e follows the cyclic order of (i, i+1, i+2) */
	for (int i = 0; i < 3; i++) {
		const int c1 = (i + 1) % 3, c2 = (i + 2) % 3;
		const int i1 = c1 < c2 ? c1 : c2, i2 = c1 < c2 ? c2 : c1;
		for (int j = 0; j < 3; j++) {
			const int k1 = (j + 1) % 3, k2 = (j + 2) % 3;
			const int j1 = k1 < k2 ? k1 : k2, j2 = k1 < k2 ? k2 : k1;
			const double e = pp[c2] * R[3 * c1 + j] - pp[c1] * R[3 * c2 + j];
			const double r = ((A[i1] * Q[3 * i2 + j] + A[i2] * Q[3 * i1 + j]) + B[j1] * Q[3 * i + j2]) +
			                 B[j2] * Q[3 * i + j1];
			if (fabs(e) - r > 0.0)
				return 1;
		}
	}
	return 0;
}

/*
 * Intersection segment of two triangles, after Moller, "A Fast
 * Triangle-Triangle Intersection Test" (JGT 1997), interval variant with the
 * segment end points, as lib/tritri wraps it for UnitWorld::intersectingSegments
 * (lib/osr/unit_world.cc:959-1012): plane distances below 1e-6 count as
 * zero; the triangles' intervals on the line where their planes meet are
 * compared, and the overlap's end points are returned in p1, p2.  Returns 1
 * when the triangles meet; coplanar triangles return 1 with *coplanar set and
 * p1 = p2 = 0 (the segment is undefined there).
 */
SE3_HD static inline void se3_mol_isect2(const double* v0, const double* v1, const double* v2, double vv0, double vv1,
                                         double vv2, double d0, double d1, double d2, double* i0, double* i1,
                                         double* p0, double* p1)
{
	double t = d0 / (d0 - d1);
	*i0 = vv0 + (vv1 - vv0) * t;
	for (int k = 0; k < 3; k++)
		p0[k] = (v1[k] - v0[k]) * t + v0[k];
	t = d0 / (d0 - d2);
	*i1 = vv0 + (vv2 - vv0) * t;
	for (int k = 0; k < 3; k++)
		p1[k] = v0[k] + (v2[k] - v0[k]) * t;
}

/* Interval of triangle (v0, v1, v2) on the common line; returns 1 if the
 * triangle lies in the other's plane. */
SE3_HD static inline int se3_mol_interval(const double* v0, const double* v1, const double* v2, double vv0, double vv1,
                                          double vv2, double d0, double d1, double d2, double d0d1, double d0d2,
                                          double* i0, double* i1, double* p0, double* p1)
{
	if (d0d1 > 0.0)
		se3_mol_isect2(v2, v0, v1, vv2, vv0, vv1, d2, d0, d1, i0, i1, p0, p1);
	else if (d0d2 > 0.0)
		se3_mol_isect2(v1, v0, v2, vv1, vv0, vv2, d1, d0, d2, i0, i1, p0, p1);
	else if (d1 * d2 > 0.0 || d0 != 0.0)
		se3_mol_isect2(v0, v1, v2, vv0, vv1, vv2, d0, d1, d2, i0, i1, p0, p1);
	else if (d1 != 0.0)
		se3_mol_isect2(v1, v0, v2, vv1, vv0, vv2, d1, d0, d2, i0, i1, p0, p1);
	else if (d2 != 0.0)
		se3_mol_isect2(v2, v0, v1, vv2, vv0, vv1, d2, d0, d1, i0, i1, p0, p1);
	else
		return 1;
	return 0;
}

SE3_HD static inline double se3_mol_dot(const double* a, const double* b)
{
	return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

SE3_HD static inline void se3_mol_cross(const double* a, const double* b, double* r)
{
	r[0] = a[1] * b[2] - a[2] * b[1];
	r[1] = a[2] * b[0] - a[0] * b[2];
	r[2] = a[0] * b[1] - a[1] * b[0];
}

SE3_HD static inline int se3_tri_isect_segment(const double* V0, const double* V1, const double* V2, const double* U0,
                                               const double* U1, const double* U2, int* coplanar, double* p1,
                                               double* p2)
{
	const double eps = 0.000001;
	double e1[3], e2[3], n1[3], n2[3], dd[3];
	*coplanar = 0;
	for (int k = 0; k < 3; k++) {
		p1[k] = 0.0;
		p2[k] = 0.0;
		e1[k] = V1[k] - V0[k];
		e2[k] = V2[k] - V0[k];
	}
	se3_mol_cross(e1, e2, n1);
	const double d1 = -se3_mol_dot(n1, V0);
	double du0 = se3_mol_dot(n1, U0) + d1, du1 = se3_mol_dot(n1, U1) + d1, du2 = se3_mol_dot(n1, U2) + d1;
	if (fabs(du0) < eps) du0 = 0.0;
	if (fabs(du1) < eps) du1 = 0.0;
	if (fabs(du2) < eps) du2 = 0.0;
	const double du0du1 = du0 * du1, du0du2 = du0 * du2;
	if (du0du1 > 0.0 && du0du2 > 0.0)
		return 0;
	for (int k = 0; k < 3; k++) {
		e1[k] = U1[k] - U0[k];
		e2[k] = U2[k] - U0[k];
	}
	se3_mol_cross(e1, e2, n2);
	const double d2 = -se3_mol_dot(n2, U0);
	double dv0 = se3_mol_dot(n2, V0) + d2, dv1 = se3_mol_dot(n2, V1) + d2, dv2 = se3_mol_dot(n2, V2) + d2;
	if (fabs(dv0) < eps) dv0 = 0.0;
	if (fabs(dv1) < eps) dv1 = 0.0;
	if (fabs(dv2) < eps) dv2 = 0.0;
	const double dv0dv1 = dv0 * dv1, dv0dv2 = dv0 * dv2;
	if (dv0dv1 > 0.0 && dv0dv2 > 0.0)
		return 0;
	se3_mol_cross(n1, n2, dd);
	double mx = fabs(dd[0]);
	int idx = 0;
	if (fabs(dd[1]) > mx) {
		mx = fabs(dd[1]);
		idx = 1;
	}
	if (fabs(dd[2]) > mx)
		idx = 2;
	double i1a, i1b, i2a, i2b, a1[3], a2[3], b1[3], b2[3];
	if (se3_mol_interval(V0, V1, V2, V0[idx], V1[idx], V2[idx], dv0, dv1, dv2, dv0dv1, dv0dv2, &i1a, &i1b, a1, a2)) {
		*coplanar = 1;
		return 1;
	}
	if (se3_mol_interval(U0, U1, U2, U0[idx], U1[idx], U2[idx], du0, du1, du2, du0du1, du0du2, &i2a, &i2b, b1, b2)) {
		/* U lies in V's plane within the snap while V does not lie in U's:
		 * Moller's code leaves U's interval unset; report the pair as
		 * coplanar */
		*coplanar = 1;
		return 1;
	}
	int s1 = 0, s2 = 0;
	if (i1a > i1b) {
		const double t = i1a; i1a = i1b; i1b = t;
		s1 = 1;
	}
	if (i2a > i2b) {
		const double t = i2a; i2a = i2b; i2b = t;
		s2 = 1;
	}
	if (i1b < i2a || i2b < i1a)
		return 0;
	const double* q1;
	const double* q2;
	if (i2a < i1a) {
		q1 = s1 == 0 ? a1 : a2;
		if (i2b < i1b)
			q2 = s2 == 0 ? b2 : b1;
		else
			q2 = s1 == 0 ? a2 : a1;
	} else {
		q1 = s2 == 0 ? b1 : b2;
		if (i2b > i1b)
			q2 = s1 == 0 ? a2 : a1;
		else
			q2 = s2 == 0 ? b2 : b1;
	}
	for (int k = 0; k < 3; k++) {
		p1[k] = q1[k];
		p2[k] = q2[k];
	}
	return 1;
}

/* Centre of a model's box in the world: the model transform (rotation R,
 * translation t) applied to the box's local centre c. */
SE3_HD static inline void se3_place(const double* R, const double* t, const double* c, double* out)
{
	for (int i = 0; i < 3; i++)
		out[i] = ((R[3 * i] * c[0] + R[3 * i + 1] * c[1]) + R[3 * i + 2] * c[2]) + t[i];
}

/* Eigen's Quaternion * Vector3 (_transformVector) for a unit quaternion
 * q = (w, x, y, z): uv = 2 (vec x v); v + w uv + vec x uv. */
SE3_HD static inline void se3_quat_apply(const double* q, const double* v, double* out)
{
	const double u[3] = {q[1], q[2], q[3]};
	double uv[3], c[3];
	uv[0] = u[1] * v[2] - u[2] * v[1];
	uv[1] = u[2] * v[0] - u[0] * v[2];
	uv[2] = u[0] * v[1] - u[1] * v[0];
	for (int k = 0; k < 3; k++)
		uv[k] = uv[k] + uv[k];
	c[0] = u[1] * uv[2] - u[2] * uv[1];
	c[1] = u[2] * uv[0] - u[0] * uv[2];
	c[2] = u[0] * uv[1] - u[1] * uv[0];
	for (int k = 0; k < 3; k++)
		out[k] = (v[k] + q[0] * uv[k]) + c[k];
}

/* translateToUnitState (lib/osr/unit_world.cc:526-539) then applyPertubation
 * (713-724) of a w-first state: the translation through the calibration
 * matrix M (row-major 4x4, the binary32 matrix widened), then the
 * perturbation pert = (t, w, x, y, z) applied on the left. */
SE3_HD static inline void se3_to_unit(const double* M, const double* pert, const double* q, double* out)
{
	const double t4[4] = {q[0], q[1], q[2], 1.0};
	double nt[3];
	for (int r = 0; r < 3; r++)
		nt[r] = ((M[4 * r] * t4[0] + M[4 * r + 1] * t4[1]) + M[4 * r + 2] * t4[2]) + M[4 * r + 3] * t4[3];
	double rot[4] = {q[3], q[4], q[5], q[6]};
	se3_quat_normalize(rot);
	double prot[4] = {pert[3], pert[4], pert[5], pert[6]};
	se3_quat_normalize(prot);
	double qret[4];
	se3_quat_mul(prot, rot, qret);
	double PR[9];
	se3_quat_to_rotmat(prot[0], prot[1], prot[2], prot[3], PR);
	for (int r = 0; r < 3; r++)
		out[r] = pert[r] + ((PR[3 * r] * nt[0] + PR[3 * r + 1] * nt[1]) + PR[3 * r + 2] * nt[2]);
	se3_quat_normalize(qret);
	for (int i = 0; i < 4; i++)
		out[3 + i] = qret[i];
}

/* translateOMPLStateToUnitState (lib/osr/unit_world.cc:689-711) of one OMPL
 * state (x, y, z, qx, qy, qz, qw): reorder to w first, translate to the unit
 * world, then move the translation by -R delta, where delta = (OMPL centre -
 * robot centre) * scene scale is non-zero after enforceRobotCenter. */
SE3_HD static inline void se3_ompl_to_unit(const double* M, const double* pert, const double* delta, const double* q,
                                           double* out)
{
	const double wf[7] = {q[0], q[1], q[2], q[6], q[3], q[4], q[5]};
	double u[7];
	se3_to_unit(M, pert, wf, u);
	double rot[4] = {u[3], u[4], u[5], u[6]};
	se3_quat_normalize(rot); /* decompose */
	double rd[3];
	se3_quat_apply(rot, delta, rd);
	for (int k = 0; k < 3; k++)
		out[k] = u[k] - rd[k];
	se3_quat_normalize(rot); /* compose */
	for (int k = 0; k < 4; k++)
		out[3 + k] = rot[k];
}

/* ---------------------------------------------------------------------- */
/* OMPL's SE(3) state space and discrete motion validator (OMPL 1.4.1, the */
/* version the reference's ompl.app fork builds on; restated from OMPL's    */
/* published sources): SE3StateSpace is the compound of a 3-D real vector   */
/* space and SO3StateSpace with weights 1.  States are in OMPL's order      */
/* (x, y, z, qx, qy, qz, qw) with unit quaternions.                         */
/* ---------------------------------------------------------------------- */
#define SE3_OMPL_QUAT_EPS 1e-9 /* MAX_QUATERNION_NORM_ERROR */
#define SE3_DBL_EPS 0x1p-52

/* qs1->x * qs2->x + qs1->y * qs2->y + qs1->z * qs2->z + qs1->w * qs2->w */
SE3_HD static inline double se3_ompl_qdot(const double* a, const double* b)
{
	return ((a[3] * b[3] + a[4] * b[4]) + a[5] * b[5]) + a[6] * b[6];
}

/* SO3StateSpace arcLength: 0 when |dot| > 1 - 1e-9, else acos(|dot|). */
SE3_HD static inline double se3_ompl_arc(const double* a, const double* b)
{
	const double dq = fabs(se3_ompl_qdot(a, b));
	if (dq > 1.0 - SE3_OMPL_QUAT_EPS)
		return 0.0;
	return se3_acos(dq);
}

/* CompoundStateSpace::distance: RealVectorStateSpace's Euclidean distance
 * (squares summed in index order) plus the SO(3) arc length. */
SE3_HD static inline double se3_ompl_distance(const double* a, const double* b)
{
	double e = 0.0;
	for (int i = 0; i < 3; i++) {
		const double d = a[i] - b[i];
		e += d * d;
	}
	return sqrt(e) + se3_ompl_arc(a, b);
}

/* getMaximumExtent() * resolution, the longest valid segment after
 * setStateValidityCheckingResolution(resolution): the bounding box diagonal
 * (RealVectorStateSpace) plus pi / 2 (SO3StateSpace). */
SE3_HD static inline double se3_ompl_longest_segment(const double* lo, const double* hi, double resolution)
{
	double e = 0.0;
	for (int i = 0; i < 3; i++) {
		const double d = hi[i] - lo[i];
		e += d * d;
	}
	double ext = 0.0;
	ext += 1.0 * sqrt(e);
	ext += 1.0 * (0.5 * 3.141592653589793);
	return ext * resolution;
}

/* One motion of DiscreteMotionValidator::checkMotion: nd =
 * validSegmentCount = ceil(distance / longest valid segment); the validator
 * tests the end state and, when nd >= 2, interpolate(from, to, k / nd) for
 * k = 1 .. nd-1.  The start state is not tested. */
typedef struct se3_ompl_motion {
	double a[7], b[7];
	double theta, inv_sin; /* slerp angle and 1 / sin(theta) */
	int slerp;             /* theta > DBL_EPSILON */
	int neg;               /* signed dot < 0: the long-arc case */
	int64_t nd;            /* -1: not a finite count below 2^31 */
} se3_ompl_motion;

SE3_HD static inline void se3_ompl_motion_init(se3_ompl_motion* m, const double* from, const double* to, double lvs)
{
	for (int k = 0; k < 7; k++) {
		m->a[k] = from[k];
		m->b[k] = to[k];
	}
	const double c = ceil(se3_ompl_distance(from, to) / lvs);
	m->nd = (c >= 0.0 && c < 2147483648.0) ? (int64_t)c : -1;
	m->theta = se3_ompl_arc(from, to);
	m->slerp = m->theta > SE3_DBL_EPS;
	m->inv_sin = m->slerp ? 1.0 / se3_sin(m->theta) : 0.0;
	m->neg = se3_ompl_qdot(from, to) < 0.0;
}

/* CompoundStateSpace::interpolate: from + (to - from) t for the
 * translation; SO3StateSpace's slerp (q1 sin((1-t) theta) + q2 sin(t theta))
 * / sin(theta), with the second weight negated on the long arc, or the
 * `from` rotation when theta <= DBL_EPSILON. */
SE3_HD static inline void se3_ompl_interpolate(const se3_ompl_motion* m, double t, double* out)
{
	for (int i = 0; i < 3; i++)
		out[i] = m->a[i] + (m->b[i] - m->a[i]) * t;
	if (m->slerp) {
		const double s0 = se3_sin((1.0 - t) * m->theta);
		double s1 = se3_sin(t * m->theta);
		if (m->neg)
			s1 = -s1;
		for (int k = 3; k < 7; k++)
			out[k] = (m->a[k] * s0 + m->b[k] * s1) * m->inv_sin;
	} else {
		for (int k = 3; k < 7; k++)
			out[k] = m->a[k];
	}
}

/* se3_ompl_interpolate with se3_sin_legacy for the weights: the states the
 * GPU's motion walk tests (see se3_segment_at_fast). */
SE3_HD static inline void se3_ompl_interpolate_fast(const se3_ompl_motion* m, double t, double* out)
{
	for (int i = 0; i < 3; i++)
		out[i] = m->a[i] + (m->b[i] - m->a[i]) * t;
	if (m->slerp) {
		const double s0 = se3_sin_legacy((1.0 - t) * m->theta);
		double s1 = se3_sin_legacy(t * m->theta);
		if (m->neg)
			s1 = -s1;
		for (int k = 3; k < 7; k++)
			out[k] = (m->a[k] * s0 + m->b[k] * s1) * m->inv_sin;
	} else {
		for (int k = 3; k < 7; k++)
			out[k] = m->a[k];
	}
}

/* The k-th state the validator tests (1 <= k <= max(nd, 1)): the end state
 * itself for the last, interpolate(k / nd) for the others (fast = 1: with
 * se3_ompl_interpolate_fast). */
SE3_HD static inline void se3_ompl_motion_sample(const se3_ompl_motion* m, int64_t k, double* out, int fast)
{
	const int64_t last = m->nd > 1 ? m->nd : 1;
	if (k >= last) {
		for (int i = 0; i < 7; i++)
			out[i] = m->b[i];
	} else if (fast) {
		se3_ompl_interpolate_fast(m, (double)k / (double)m->nd, out);
	} else {
		se3_ompl_interpolate(m, (double)k / (double)m->nd, out);
	}
}

/* distance() of lib/osr/osr_state.cc:130-154: |t0-t1| + angularDistance. */
SE3_HD static inline double se3_distance(const double* a, const double* b)
{
	double qa[4] = {a[3], a[4], a[5], a[6]};
	double qb[4] = {b[3], b[4], b[5], b[6]};
	se3_quat_normalize(qa);
	se3_quat_normalize(qb);
	double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
	double trdist = sqrt((dx * dx + dy * dy) + dz * dz);
	/* angularDistance: d = qa * conj(qb); 2 atan2(|d.vec|, |d.w|). */
	double cb[4] = {qb[0], -qb[1], -qb[2], -qb[3]};
	double d[4];
	se3_quat_mul(qa, cb, d);
	double vn = sqrt((d[1] * d[1] + d[2] * d[2]) + d[3] * d[3]);
	double rotdist = 2.0 * se3_atan2(vn, fabs(d[0]));
	return trdist + rotdist;
}

/*
 * Segment interpolation, interpolate() of lib/osr/osr_state.cc:27-44, split
 * into a per-segment part (normalised endpoint quaternions and the slerp
 * angle, which Eigen's slerp recomputes identically on every call) and a
 * per-tau part.
 */
typedef struct se3_segment {
	double p0[3], p1[3];
	double q0[4], q1[4]; /* normalised, (w,x,y,z) */
	double theta, sin_theta;
	int lerp;  /* |dot| >= 1 - eps: linear weights */
	int neg;   /* dot < 0: negate the second weight */
} se3_segment;

SE3_HD static inline void se3_segment_init(se3_segment* s, const double* from, const double* to)
{
	for (int i = 0; i < 3; i++) {
		s->p0[i] = from[i];
		s->p1[i] = to[i];
	}
	for (int i = 0; i < 4; i++) {
		s->q0[i] = from[3 + i];
		s->q1[i] = to[3 + i];
	}
	se3_quat_normalize(s->q0);
	se3_quat_normalize(s->q1);
	const double one = 1.0 - 0x1p-52; /* 1 - NumTraits<double>::epsilon() */
	double d = se3_quat_dot(s->q0, s->q1);
	double absd = fabs(d);
	s->neg = d < 0.0;
	if (absd >= one) {
		s->lerp = 1;
		s->theta = 0.0;
		s->sin_theta = 0.0;
	} else {
		s->lerp = 0;
		s->theta = se3_acos(absd);
		s->sin_theta = se3_sin(s->theta);
	}
}

/*
 * The pose a segment walk tests at tau, as the GPU kernels evaluate it inside
 * their traversals: se3_segment_at with the two sines of the slerp weights
 * from se3_sin_legacy (within 2 ulp of the correctly rounded value).  The
 * correctly rounded evaluation needs more registers than the traversal
 * loops can give without spilling; a tested pose differs from the
 * reference's by at most a few ulp per coordinate, which can change a
 * collision answer only for a pose that close to touching.  The returned
 * states (last free and first colliding) use se3_segment_at.
 */
SE3_HD static inline void se3_segment_at_fast(const se3_segment* s, double tau, double* out)
{
	double omt = 1.0 - tau;
	for (int i = 0; i < 3; i++)
		out[i] = s->p0[i] * omt + s->p1[i] * tau;
	double scale0, scale1;
	if (s->lerp) {
		scale0 = 1.0 - tau;
		scale1 = tau;
	} else {
		scale0 = se3_sin_legacy((1.0 - tau) * s->theta) / s->sin_theta;
		scale1 = se3_sin_legacy(tau * s->theta) / s->sin_theta;
	}
	if (s->neg)
		scale1 = -scale1;
	for (int i = 0; i < 4; i++)
		out[3 + i] = scale0 * s->q0[i] + scale1 * s->q1[i];
}

SE3_HD static inline void se3_segment_at(const se3_segment* s, double tau, double* out)
{
	double omt = 1.0 - tau;
	for (int i = 0; i < 3; i++)
		out[i] = s->p0[i] * omt + s->p1[i] * tau;
	double scale0, scale1;
	if (s->lerp) {
		scale0 = 1.0 - tau;
		scale1 = tau;
	} else {
		scale0 = se3_sin((1.0 - tau) * s->theta) / s->sin_theta;
		scale1 = se3_sin(tau * s->theta) / s->sin_theta;
	}
	if (s->neg)
		scale1 = -scale1;
	for (int i = 0; i < 4; i++)
		out[3 + i] = scale0 * s->q0[i] + scale1 * s->q1[i];
}

#endif

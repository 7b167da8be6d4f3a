/*
 * cull.h -- the conservative binary32 culling test of the BVH traversals
 * (host and device): the culling transform of one query, a robot box placed
 * in the environment's frame, and the 15-axis separating test.  The kernels
 * include it; a host build lets the tests check, in exact rational
 * arithmetic, that every box pair it culls is disjoint.
 */
#ifndef SE3CQ_CULL_H
#define SE3CQ_CULL_H

#include <math.h>
#include <string.h>

#include "engine.h"
#include "se3_math.h"

namespace se3cq {

/* rsqrt: the hardware's approximation on the device (within an ulp or two,
 * which the norm deviation term below absorbs), 1 / sqrt on the host. */
SE3_HD __attribute__((always_inline)) inline float se3_rsqrtf(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
	return rsqrtf(x);
#else
	return 1.0f / sqrtf(x);
#endif
}

SE3_HD __attribute__((always_inline)) inline int se3_f2i(float f)
{
	int i;
	memcpy(&i, &f, sizeof(i));
	return i;
}

/*
 * Culling transform of one query (binary32).  O orients robot boxes in the
 * env frame and T places them.  The exact map the leaf test applies is
 * L q + T' with L = Re^T R(q) and T' = Re^T (t - te) in binary64, where R(q)
 * is Eigen's toRotationMatrix of the raw (possibly non-unit) quaternion.
 * Since R(q) = n^2 R(q/n) + (1 - n^2) I (n = |q|), ||L - O||_F is at most
 * ||Re||_F * 2.83 * |1 - n^2| + sqrt(3) ||Re - R(pert/|pert|)||_F plus the
 * binary32 rounding of O (below 3e-5); robot boxes grown by that bound times
 * their reach (|centre| + |extents|) contain the exact images of their
 * triangles.  eta absorbs the binary32 rounding of translations.
 */
struct QCull {
	float O[9];
	float T[3];
	float infl;
	float eta;
};

SE3_HD __attribute__((always_inline)) inline void make_qcull(const DevScene& s, const double* q, QCull& c)
{
	const float w = (float)q[3], x = (float)q[4], y = (float)q[5], z = (float)q[6];
	const float n2 = fmaf(w, w, fmaf(x, x, fmaf(y, y, z * z)));
	float Rq[9];
	if (n2 > 0.0f) {
		const float inv = se3_rsqrtf(n2);
		const float hw = w * inv, hx = x * inv, hy = y * inv, hz = z * inv;
		const float tx = 2.0f * hx, ty = 2.0f * hy, tz = 2.0f * hz;
		Rq[0] = 1.0f - fmaf(ty, hy, tz * hz);
		Rq[1] = fmaf(ty, hx, -tz * hw);
		Rq[2] = fmaf(tz, hx, ty * hw);
		Rq[3] = fmaf(ty, hx, tz * hw);
		Rq[4] = 1.0f - fmaf(tx, hx, tz * hz);
		Rq[5] = fmaf(tz, hy, -tx * hw);
		Rq[6] = fmaf(tz, hx, -ty * hw);
		Rq[7] = fmaf(tz, hy, tx * hw);
		Rq[8] = 1.0f - fmaf(tx, hx, ty * hy);
	} else {
		/* R(0) is exactly the identity */
#pragma unroll
		for (int i = 0; i < 9; i++)
			Rq[i] = (i % 4 == 0) ? 1.0f : 0.0f;
	}
	float dl1 = 0.0f;
	float fro_Re = 1.7320509f, fro_dRe = 0.0f; /* sqrt(3) rounded up, 0 */
	if (s.pert_identity) {
#pragma unroll
		for (int i = 0; i < 9; i++)
			c.O[i] = Rq[i];
#pragma unroll
		for (int i = 0; i < 3; i++)
			c.T[i] = (float)q[i];
	} else {
		const PertData* P = s.pert;
		fro_Re = P->fro_Re;
		fro_dRe = P->fro_dRe;
		float d[3];
#pragma unroll
		for (int i = 0; i < 3; i++) {
			d[i] = (float)(q[i] - P->te[i]);
			dl1 += fabsf(d[i]);
		}
#pragma unroll
		for (int i = 0; i < 3; i++) {
#pragma unroll
			for (int j = 0; j < 3; j++)
				c.O[3 * i + j] = fmaf(P->PeT[3 * i], Rq[j], fmaf(P->PeT[3 * i + 1], Rq[3 + j], P->PeT[3 * i + 2] * Rq[6 + j]));
			c.T[i] = fmaf(P->PeT[3 * i], d[0], fmaf(P->PeT[3 * i + 1], d[1], P->PeT[3 * i + 2] * d[2]));
		}
	}
	const float dev = fabsf(1.0f - n2) + 1e-6f * n2;
	const float delta = fmaf(fro_Re * 2.83f, dev, fro_dRe + 3e-5f);
	c.infl = delta * s.rob_reach;
	const float tm = fmaxf(fabsf(c.T[0]), fmaxf(fabsf(c.T[1]), fabsf(c.T[2])));
	c.eta = fmaf(1e-5f, tm + s.scene_mag, fro_dRe * dl1);
}

/* Robot box in the env frame, cached across consecutive tests that share the
 * robot node. */
struct RobBox {
	float W[9]; /* column j = robot axis j in the env frame */
	float c[3];
	float e[3];
	int child;
};

SE3_HD __attribute__((always_inline)) inline void load_rob_box(const GNode& B, const QCull& t, RobBox& r)
{
	const float b[3][3] = {{B.a0_ey.x, B.a0_ey.y, B.a0_ey.z},
	                       {B.a1_ez.x, B.a1_ez.y, B.a1_ez.z},
	                       {B.a2_child.x, B.a2_child.y, B.a2_child.z}};
#pragma unroll
	for (int i = 0; i < 3; i++) {
#pragma unroll
		for (int j = 0; j < 3; j++)
			r.W[3 * i + j] = fmaf(t.O[3 * i], b[j][0], fmaf(t.O[3 * i + 1], b[j][1], t.O[3 * i + 2] * b[j][2]));
		r.c[i] = fmaf(t.O[3 * i], B.c_ex.x, fmaf(t.O[3 * i + 1], B.c_ex.y, fmaf(t.O[3 * i + 2], B.c_ex.z, t.T[i])));
	}
	r.e[0] = B.c_ex.w + t.infl;
	r.e[1] = B.a0_ey.w + t.infl;
	r.e[2] = B.a1_ez.w + t.infl;
	r.child = se3_f2i(B.a2_child.w);
}

/*
 * Gottschalk's 15-axis separating test in the env box frame, binary32.
 * Disjoint only when a separation exceeds the summed radii by a relative
 * 3e-5 plus eta = 1e-5 * (|T| + scene magnitude): binary32 rounding of
 * these ~10-operation expressions is below 2e-6 relative, so a binary32
 * "disjoint" implies the exact boxes are disjoint.  |R| gets FCL's 1e-6.
 * The three env-axis tests run first, row by row, so most disjoint pairs
 * exit before the full 3x3 relative rotation is formed.
 */
#ifndef SE3CQ_OBB_EDGE_AXES
#define SE3CQ_OBB_EDGE_AXES 1
#endif
SE3_HD __attribute__((always_inline)) inline bool obb_disjoint(const GNode& A, const RobBox& B, float eta)
{
	const float K = 1.00003f;
	const float d[3] = {B.c[0] - A.c_ex.x, B.c[1] - A.c_ex.y, B.c[2] - A.c_ex.z};
	const float ea[3] = {A.c_ex.w, A.a0_ey.w, A.a1_ez.w};
	const float a[3][3] = {{A.a0_ey.x, A.a0_ey.y, A.a0_ey.z},
	                       {A.a1_ez.x, A.a1_ez.y, A.a1_ez.z},
	                       {A.a2_child.x, A.a2_child.y, A.a2_child.z}};
	float t[3], Rr[3][3], Bf[3][3];
#pragma unroll
	for (int i = 0; i < 3; i++) {
		t[i] = fmaf(a[i][0], d[0], fmaf(a[i][1], d[1], a[i][2] * d[2]));
#pragma unroll
		for (int j = 0; j < 3; j++) {
			Rr[i][j] = fmaf(a[i][0], B.W[j], fmaf(a[i][1], B.W[3 + j], a[i][2] * B.W[6 + j]));
			Bf[i][j] = fabsf(Rr[i][j]) + 1e-6f;
		}
		const float r = ea[i] + fmaf(B.e[0], Bf[i][0], fmaf(B.e[1], Bf[i][1], B.e[2] * Bf[i][2]));
		if (fabsf(t[i]) > fmaf(r, K, eta))
			return true;
	}
#pragma unroll
	for (int j = 0; j < 3; j++) {
		const float sj = fmaf(t[0], Rr[0][j], fmaf(t[1], Rr[1][j], t[2] * Rr[2][j]));
		const float r = B.e[j] + fmaf(ea[0], Bf[0][j], fmaf(ea[1], Bf[1][j], ea[2] * Bf[2][j]));
		if (fabsf(sj) > fmaf(r, K, eta))
			return true;
	}
#if SE3CQ_OBB_EDGE_AXES
#pragma unroll
	for (int i = 0; i < 3; i++) {
		const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
#pragma unroll
		for (int j = 0; j < 3; j++) {
			const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
			const float sx = fmaf(t[i2], Rr[i1][j], -t[i1] * Rr[i2][j]);
			const float r = fmaf(ea[i1], Bf[i2][j], fmaf(ea[i2], Bf[i1][j], fmaf(B.e[j1], Bf[i][j2], B.e[j2] * Bf[i][j1])));
			if (fabsf(sx) > fmaf(r, K, eta))
				return true;
		}
	}
#endif
	return false;
}

/* The smallest binary32 value not below x (host). */
inline float round_up_f(double x)
{
	float f = (float)x;
	if ((double)f < x)
		f = nextafterf(f, INFINITY);
	return f;
}

/* Perturbation data for the device (see PertData; host). */
inline PertData make_pert(const double* p)
{
	PertData d;
	memset(&d, 0, sizeof(d));
	se3_quat_to_rotmat(p[3], p[4], p[5], p[6], d.Re);
	for (int i = 0; i < 3; i++)
		d.te[i] = p[i];
	double qn[4] = {p[3], p[4], p[5], p[6]};
	se3_quat_normalize(qn);
	double Rh[9];
	se3_quat_to_rotmat(qn[0], qn[1], qn[2], qn[3], Rh);
	double fr = 0.0, fd = 0.0;
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) {
			d.PeT[3 * i + j] = (float)Rh[3 * j + i];
			fr += d.Re[3 * i + j] * d.Re[3 * i + j];
			const double e = d.Re[3 * i + j] - Rh[3 * i + j];
			fd += e * e;
		}
	d.fro_Re = round_up_f(sqrt(fr) * (1.0 + 1e-9));
	d.fro_dRe = round_up_f(sqrt(3.0 * fd) * (1.0 + 1e-9) + (fd > 0 ? 1e-15 : 0.0));
	return d;
}

} // namespace se3cq

#endif

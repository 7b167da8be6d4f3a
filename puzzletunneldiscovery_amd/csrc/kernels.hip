/*
 * kernels.hip -- gfx950 kernels of the SE(3) collision-query engine.
 *
 * Compiled with -ffp-contract=off: the binary64 leaf test and the state
 * arithmetic must round exactly like the CPU restatement of FCL.  The
 * binary32 box culling uses explicit fmaf and is conservative by
 * construction (margins below), so it never changes a boolean.
 *
 * Kernels
 *   k_valid_states   one thread per state: BVH x BVH depth-first traversal
 *                    (UnitWorld::isValid, lib/osr/unit_world.cc:215-247 ->
 *                    fcl::collide via lib/osr/cdmodel.cc:162-178).
 *   k_segments       one wavefront per segment: the 64 lanes evaluate 64
 *                    consecutive discrete steps at once, a ballot finds the
 *                    first colliding step (transitStateToWithContact,
 *                    unit_world.cc:405-474; isValidTransition 477-486), and
 *                    the touch search refines contact with an exact
 *                    5-levels-per-round parallel bisection.
 *   k_visibility     one wavefront per state pair (calculateVisibility*,
 *                    unit_world.cc:740-840).
 */
#include <hip/hip_runtime.h>

#include "engine.h"
#include "se3_math.h"
#include "launch.h"

namespace se3cq {

enum { MODE_TRANSIT = 0, MODE_TRANSITION = 1, MODE_TOUCH = 2 };
enum { VIS_PAIR = 0, VIS_MATRIX = 1, VIS_MATRIX2 = 2 };

/* Per-query transforms.  R, T: the exact map applied to robot vertices in
 * the leaf test (FCL relative transform Re^T Rr, Re^T (tr - te)).  O: the
 * same rotation from normalised quaternions, rounded to binary32, used to
 * orient robot boxes; infl bounds ||R - O||_F * |box reach| so robot boxes
 * grown by it contain the exact images of their triangles. */
struct QTf {
	double R[9], T[3];
	float O[9], Tf[3];
	float infl, eta;
};

__device__ __forceinline__ void make_qtf(const DevScene& s, const double* q, QTf& t)
{
	double Rr[9];
	se3_quat_to_rotmat(q[3], q[4], q[5], q[6], Rr);
	double d[3] = {q[0] - s.te[0], q[1] - s.te[1], q[2] - s.te[2]};
#pragma unroll
	for (int i = 0; i < 3; i++) {
#pragma unroll
		for (int j = 0; j < 3; j++)
			t.R[3 * i + j] = (s.Re[i] * Rr[j] + s.Re[3 + i] * Rr[3 + j]) + s.Re[6 + i] * Rr[6 + j];
		t.T[i] = (s.Re[i] * d[0] + s.Re[3 + i] * d[1]) + s.Re[6 + i] * d[2];
	}
	double qn[4] = {q[3], q[4], q[5], q[6]};
	se3_quat_normalize(qn);
	double Rn[9];
	se3_quat_to_rotmat(qn[0], qn[1], qn[2], qn[3], Rn);
	double s2 = 0.0;
#pragma unroll
	for (int i = 0; i < 3; i++)
#pragma unroll
		for (int j = 0; j < 3; j++) {
			double o = (s.ReHat[i] * Rn[j] + s.ReHat[3 + i] * Rn[3 + j]) + s.ReHat[6 + i] * Rn[6 + j];
			float of = (float)o;
			t.O[3 * i + j] = of;
			double dd = t.R[3 * i + j] - (double)of;
			s2 += dd * dd;
		}
	/* (1 + 2^-20) absorbs the binary32 rounding of the product */
	t.infl = (float)(sqrt(s2) * (double)s.rob_reach * (1.0 + 0x1p-20));
	float tm = 0.0f;
#pragma unroll
	for (int i = 0; i < 3; i++) {
		t.Tf[i] = (float)t.T[i];
		tm = fmaxf(tm, fabsf(t.Tf[i]));
	}
	t.eta = 1e-5f * (tm + s.scene_mag);
}

/* Transformed robot box (binary32), cached across consecutive tests that
 * share the robot node. */
struct RobBox {
	float W[9]; /* column j = robot axis j in the env frame */
	float c[3];
	float e[3];
	float size;
	int child;
};

__device__ __forceinline__ void load_rob_box(const GNode& B, const QTf& t, RobBox& r)
{
	const float b[3][3] = {{B.a0_ey.x, B.a0_ey.y, B.a0_ey.z},
	                       {B.a1_ez.x, B.a1_ez.y, B.a1_ez.z},
	                       {B.a2_child.x, B.a2_child.y, B.a2_child.z}};
#pragma unroll
	for (int i = 0; i < 3; i++) {
#pragma unroll
		for (int j = 0; j < 3; j++)
			r.W[3 * i + j] = fmaf(t.O[3 * i], b[j][0], fmaf(t.O[3 * i + 1], b[j][1], t.O[3 * i + 2] * b[j][2]));
		r.c[i] = fmaf(t.O[3 * i], B.c_ex.x, fmaf(t.O[3 * i + 1], B.c_ex.y, fmaf(t.O[3 * i + 2], B.c_ex.z, t.Tf[i])));
	}
	r.e[0] = B.c_ex.w + t.infl;
	r.e[1] = B.a0_ey.w + t.infl;
	r.e[2] = B.a1_ez.w + t.infl;
	r.size = fmaf(B.c_ex.w, B.c_ex.w, fmaf(B.a0_ey.w, B.a0_ey.w, B.a1_ez.w * B.a1_ez.w));
	r.child = __float_as_int(B.a2_child.w);
}

/*
 * Gottschalk's 15-axis separating test in the env box frame, binary32.
 * Disjoint only when a separation exceeds the summed radii by a relative
 * 3e-5 plus eta = 1e-5 * (|T| + scene magnitude): binary32 rounding of
 * these ~10-operation expressions is below 2e-6 relative, so a binary32
 * "disjoint" implies the exact boxes are disjoint.  |R| gets FCL's 1e-6.
 */
__device__ __forceinline__ bool obb_disjoint(const GNode& A, const RobBox& B, float eta)
{
	const float a[3][3] = {{A.a0_ey.x, A.a0_ey.y, A.a0_ey.z},
	                       {A.a1_ez.x, A.a1_ez.y, A.a1_ez.z},
	                       {A.a2_child.x, A.a2_child.y, A.a2_child.z}};
	const float ea[3] = {A.c_ex.w, A.a0_ey.w, A.a1_ez.w};
	const float d[3] = {B.c[0] - A.c_ex.x, B.c[1] - A.c_ex.y, B.c[2] - A.c_ex.z};
	const float K = 1.00003f;
	float t[3], Rr[3][3], Bf[3][3];
#pragma unroll
	for (int i = 0; i < 3; i++) {
		t[i] = fmaf(a[i][0], d[0], fmaf(a[i][1], d[1], a[i][2] * d[2]));
#pragma unroll
		for (int j = 0; j < 3; j++) {
			Rr[i][j] = fmaf(a[i][0], B.W[j], fmaf(a[i][1], B.W[3 + j], a[i][2] * B.W[6 + j]));
			Bf[i][j] = fabsf(Rr[i][j]) + 1e-6f;
		}
	}
#pragma unroll
	for (int i = 0; i < 3; i++) {
		float r = ea[i] + fmaf(B.e[0], Bf[i][0], fmaf(B.e[1], Bf[i][1], B.e[2] * Bf[i][2]));
		if (fabsf(t[i]) > fmaf(r, K, eta))
			return true;
	}
#pragma unroll
	for (int j = 0; j < 3; j++) {
		float s = fmaf(t[0], Rr[0][j], fmaf(t[1], Rr[1][j], t[2] * Rr[2][j]));
		float r = B.e[j] + fmaf(ea[0], Bf[0][j], fmaf(ea[1], Bf[1][j], ea[2] * Bf[2][j]));
		if (fabsf(s) > fmaf(r, K, eta))
			return true;
	}
#pragma unroll
	for (int i = 0; i < 3; i++) {
		const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
#pragma unroll
		for (int j = 0; j < 3; j++) {
			const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
			float s = fmaf(t[i2], Rr[i1][j], -t[i1] * Rr[i2][j]);
			float r = fmaf(ea[i1], Bf[i2][j], fmaf(ea[i2], Bf[i1][j], fmaf(B.e[j1], Bf[i][j2], B.e[j2] * Bf[i][j1])));
			if (fabsf(s) > fmaf(r, K, eta))
				return true;
		}
	}
	return false;
}

/* ---- exact binary64 leaf test (FCL/PQP 17-axis SAT, contraction off) ---- */
__device__ __forceinline__ double dot3d(const double* a, const double* b)
{
	return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
__device__ __forceinline__ void cross3d(const double* a, const double* b, double* r)
{
	r[0] = a[1] * b[2] - a[2] * b[1];
	r[1] = a[2] * b[0] - a[0] * b[2];
	r[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ bool overlap6(const double* ax, const double* p1, const double* p2,
                                         const double* p3, const double* q1, const double* q2,
                                         const double* q3)
{
	double P1 = dot3d(ax, p1), P2 = dot3d(ax, p2), P3 = dot3d(ax, p3);
	double Q1 = dot3d(ax, q1), Q2 = dot3d(ax, q2), Q3 = dot3d(ax, q3);
	/* std::max / std::min semantics (also for NaN operands) */
	double t = (P2 < P3) ? P3 : P2;
	const double mx1 = (P1 < t) ? t : P1;
	t = (P3 < P2) ? P3 : P2;
	const double mn1 = (t < P1) ? t : P1;
	t = (Q2 < Q3) ? Q3 : Q2;
	const double mx2 = (Q1 < t) ? t : Q1;
	t = (Q3 < Q2) ? Q3 : Q2;
	const double mn2 = (t < Q1) ? t : Q1;
	return !(mn1 > mx2) && !(mn2 > mx1);
}

__device__ bool tri_intersect(const double* __restrict__ P, const double* __restrict__ Q)
{
	double p1[3], p2[3], p3[3], q1[3], q2[3], q3[3];
#pragma unroll
	for (int k = 0; k < 3; k++) {
		p1[k] = P[k] - P[k];
		p2[k] = P[3 + k] - P[k];
		p3[k] = P[6 + k] - P[k];
		q1[k] = Q[k] - P[k];
		q2[k] = Q[3 + k] - P[k];
		q3[k] = Q[6 + k] - P[k];
	}
	double e[3][3], f[3][3], n1[3], m1[3], ax[3];
#pragma unroll
	for (int k = 0; k < 3; k++) {
		e[0][k] = p2[k] - p1[k];
		e[1][k] = p3[k] - p2[k];
		e[2][k] = p1[k] - p3[k];
		f[0][k] = q2[k] - q1[k];
		f[1][k] = q3[k] - q2[k];
		f[2][k] = q1[k] - q3[k];
	}
	cross3d(e[0], e[1], n1);
	if (!overlap6(n1, p1, p2, p3, q1, q2, q3)) return false;
	cross3d(f[0], f[1], m1);
	if (!overlap6(m1, p1, p2, p3, q1, q2, q3)) return false;
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) {
			cross3d(e[i], f[j], ax);
			if (!overlap6(ax, p1, p2, p3, q1, q2, q3)) return false;
		}
	for (int i = 0; i < 3; i++) {
		cross3d(e[i], n1, ax);
		if (!overlap6(ax, p1, p2, p3, q1, q2, q3)) return false;
	}
	for (int i = 0; i < 3; i++) {
		cross3d(f[i], m1, ax);
		if (!overlap6(ax, p1, p2, p3, q1, q2, q3)) return false;
	}
	return true;
}

__device__ __forceinline__ bool leaf_test(const DevScene& s, const QTf& t, int te, int tr)
{
	const double* P = s.env.tris + 9 * (int64_t)te;
	const double* q = s.rob.tris + 9 * (int64_t)tr;
	double Pl[9], Q[9];
#pragma unroll
	for (int k = 0; k < 9; k++)
		Pl[k] = P[k];
#pragma unroll
	for (int k = 0; k < 3; k++) {
		double q0 = q[3 * k], q1 = q[3 * k + 1], q2 = q[3 * k + 2];
#pragma unroll
		for (int r = 0; r < 3; r++)
			Q[3 * k + r] = ((t.R[3 * r] * q0 + t.R[3 * r + 1] * q1) + t.R[3 * r + 2] * q2) + t.T[r];
	}
	return tri_intersect(Pl, Q);
}

/* ---- depth-first BVH x BVH traversal, one thread per query ---- */
/* Stack entries live in LDS, strided by the block size so that lanes at the
 * same depth touch consecutive banks.  Narrow scenes pack (env, rob) node ids
 * into one 32-bit word; wide scenes use two words. */
template <bool WIDE>
struct PairStack {
	uint32_t* base;
	int stride;
	__device__ __forceinline__ void put(int k, int a, int b) const
	{
		if (WIDE) {
			base[(2 * k) * stride] = (uint32_t)a;
			base[(2 * k + 1) * stride] = (uint32_t)b;
		} else {
			base[k * stride] = ((uint32_t)a << 16) | (uint32_t)b;
		}
	}
	__device__ __forceinline__ void get(int k, int& a, int& b) const
	{
		if (WIDE) {
			a = (int)base[(2 * k) * stride];
			b = (int)base[(2 * k + 1) * stride];
		} else {
			uint32_t v = base[k * stride];
			a = (int)(v >> 16);
			b = (int)(v & 0xffffu);
		}
	}
};

__device__ __forceinline__ GNode load_node(const GNode* __restrict__ nodes, int i)
{
	const float4* p = reinterpret_cast<const float4*>(nodes + i);
	GNode n;
	n.c_ex = p[0];
	n.a0_ey = p[1];
	n.a1_ez = p[2];
	n.a2_child = p[3];
	return n;
}

template <bool WIDE>
__device__ bool collide_dfs(const DevScene& s, const QTf& t, const PairStack<WIDE>& stk)
{
	int sp = 0;
	stk.put(sp++, 0, 0);
	int cached_b = -1;
	RobBox B;
	const int cap = s.stack_depth;
	while (sp > 0) {
		int a, b;
		stk.get(--sp, a, b);
		GNode A = load_node(s.env.nodes, a);
		if (b != cached_b) {
			GNode Bn = load_node(s.rob.nodes, b);
			load_rob_box(Bn, t, B);
			cached_b = b;
		}
		if (obb_disjoint(A, B, t.eta))
			continue;
		const int ca = __float_as_int(A.a2_child.w);
		const bool la = ca < 0, lb = B.child < 0;
		if (la && lb) {
			if (leaf_test(s, t, ~ca, ~B.child))
				return true;
			continue;
		}
		if (sp + 2 > cap) {
			atomicOr(s.error_flag, 1u);
			return true;
		}
		const float sa = fmaf(A.c_ex.w, A.c_ex.w, fmaf(A.a0_ey.w, A.a0_ey.w, A.a1_ez.w * A.a1_ez.w));
		if (!la && (lb || sa > B.size)) {
			stk.put(sp++, ca + 1, b);
			stk.put(sp++, ca, b);
		} else {
			stk.put(sp++, a, B.child + 1);
			stk.put(sp++, a, B.child);
		}
	}
	return false;
}

template <bool WIDE>
__device__ __forceinline__ bool state_collides(const DevScene& s, const double* q, uint32_t* smem)
{
	QTf t;
	make_qtf(s, q, t);
	PairStack<WIDE> stk{smem + threadIdx.x, (int)blockDim.x};
	return collide_dfs<WIDE>(s, t, stk);
}

template <bool WIDE>
__global__ void __launch_bounds__(256) k_valid_states(DevScene s, const double* __restrict__ qs, int64_t n,
                                                      uint8_t* __restrict__ out)
{
	extern __shared__ uint32_t smem[];
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	double q[7];
#pragma unroll
	for (int k = 0; k < 7; k++)
		q[k] = qs[7 * i + k];
	out[i] = state_collides<WIDE>(s, q, smem) ? 0 : 1;
}

/* ---- segments: one wavefront per segment ---- */
__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }

/* Exact parallel form of `levels` bisection halvings: each round evaluates
 * the 31 midpoints of a 5-level probe tree (every midpoint computed by the
 * same 0.5*(lo+hi) sequence the sequential search would use), then walks the
 * tree with the ballot of results. */
template <bool WIDE>
__device__ void bisect_wave(const DevScene& s, const se3_segment& sg, double& lo, double& hi, int levels,
                            uint32_t* smem)
{
	const int lane = threadIdx.x & 63;
	for (int done = 0; done < levels;) {
		const int k = (levels - done < 5) ? (levels - done) : 5;
		const int nodes = (1 << k) - 1;
		const bool act = lane < nodes;
		double mid = 0.0;
		bool ok = true;
		if (act) {
			const int h = lane + 1;
			const int dh = 31 - __clz(h);
			double l0 = lo, h0 = hi;
			for (int bbit = dh - 1; bbit >= 0; bbit--) {
				double m = 0.5 * (l0 + h0);
				if ((h >> bbit) & 1)
					l0 = m;
				else
					h0 = m;
			}
			mid = 0.5 * (l0 + h0);
			double st[7];
			se3_segment_at(&sg, mid, st);
			ok = !state_collides<WIDE>(s, st, smem);
		}
		const uint64_t vm = __ballot(act && ok);
		int h = 1;
		for (int d = 0; d < k; d++) {
			const double m = shfl_d(mid, h - 1);
			const int bit = (int)((vm >> (h - 1)) & 1ull);
			if (bit)
				lo = m;
			else
				hi = m;
			h = 2 * h + bit;
		}
		done += k;
	}
}

/* Wave-uniform outcome of one segment walk. */
struct SegResult {
	int done;        /* 1 = collision free (reference tuple element 2) */
	int early;       /* verify_delta >= distance: nothing evaluated */
	int lo_is_from;  /* last free state is `from` itself (no free step) */
	int32_t steps;   /* isValid calls under the reference's semantics (+ probes) */
	double last_tau, tau, lo, hi;
};

template <bool WIDE, int MODE>
__device__ SegResult segment_wave(const DevScene& s, const double* __restrict__ from, const double* __restrict__ to,
                                  double vd_in, int levels, se3_segment& sg, uint32_t* smem)
{
	const int lane = threadIdx.x & 63;
	SegResult r;
	r.early = 0;
	r.lo_is_from = 0;
	const double dist = se3_distance(from, to);
	double vd = vd_in;
	if (MODE == MODE_TRANSITION) {
		const double half = dist / 2.0;
		vd = (vd_in < half) ? vd_in : half; /* std::min(d / 2, initial_verify_delta) */
	}
	if (vd >= dist) {
		r.done = 0;
		r.early = 1;
		r.lo_is_from = 1;
		r.steps = 0;
		r.last_tau = 0.0;
		r.tau = 1.0;
		r.lo = 0.0;
		r.hi = 1.0;
		return r;
	}
	const double inv_dist = 1.0 / dist;
	se3_segment_init(&sg, from, to);
	double dbase = vd;
	double prev_tau = 0.0;
	int64_t base = 0;
	for (;;) {
		double d = dbase;
		for (int k = 0; k < lane; k++) {
			if (d < dist) {
				d += vd;
				d = (d < dist) ? d : dist;
			} else {
				d += vd;
			}
		}
		const bool active = d <= dist;
		const double tau = d * inv_dist;
		bool ok = true;
		if (active) {
			double st[7];
			se3_segment_at(&sg, tau, st);
			ok = !state_collides<WIDE>(s, st, smem);
		}
		const uint64_t cm = __ballot(active && !ok);
		const uint64_t am = __ballot(active);
		if (cm) {
			const int f = __ffsll((long long)cm) - 1;
			const double thi = shfl_d(tau, f);
			const double tprev = shfl_d(tau, f > 0 ? f - 1 : 0);
			r.done = 0;
			r.last_tau = (f > 0) ? tprev : prev_tau;
			r.tau = thi;
			r.lo_is_from = (f == 0) && (base == 0);
			r.steps = (int32_t)(base + f + 1);
			r.lo = r.last_tau;
			r.hi = thi;
			if (MODE == MODE_TOUCH) {
				bisect_wave<WIDE>(s, sg, r.lo, r.hi, levels, smem);
				r.steps += levels;
			}
			return r;
		}
		const int na = __popcll(am);
		if (na < 64) {
			const double tl = shfl_d(tau, na > 0 ? na - 1 : 0);
			r.done = 1;
			r.last_tau = (na > 0) ? tl : prev_tau;
			r.tau = 0.0; /* the reference returns its shadowed outer tau */
			r.steps = (int32_t)(base + na);
			r.lo = r.last_tau;
			r.hi = 1.0;
			return r;
		}
		const double d63 = shfl_d(d, 63);
		double nd;
		if (d63 < dist) {
			nd = d63 + vd;
			nd = (nd < dist) ? nd : dist;
		} else {
			nd = d63 + vd;
		}
		dbase = nd;
		prev_tau = shfl_d(tau, 63);
		base += 64;
	}
}

template <bool WIDE, int MODE>
__global__ void __launch_bounds__(256) k_segments(DevScene s, const double* __restrict__ from_all,
                                                  const double* __restrict__ to_all, int64_t n, double vd,
                                                  int levels, SegOut o)
{
	extern __shared__ uint32_t smem[];
	const int64_t seg = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
	if (seg >= n)
		return;
	const int lane = threadIdx.x & 63;
	double from[7], to[7];
#pragma unroll
	for (int k = 0; k < 7; k++) {
		from[k] = from_all[7 * seg + k];
		to[k] = to_all[7 * seg + k];
	}
	se3_segment sg;
	const SegResult r = segment_wave<WIDE, MODE>(s, from, to, vd, levels, sg, smem);
	if (lane == 0) {
		o.done[seg] = (uint8_t)r.done;
		if (o.last_tau) o.last_tau[seg] = r.last_tau;
		if (o.tau) o.tau[seg] = r.tau;
		if (o.lo) o.lo[seg] = r.lo;
		if (o.hi) o.hi[seg] = r.hi;
		if (o.steps) o.steps[seg] = r.steps;
	}
	if (o.last_free && lane < 7) {
		/* reconstruct the returned states; interpolation is deterministic so
		 * these equal the states the walk evaluated */
		double a[7], b[7];
		if (r.lo_is_from) {
			for (int k = 0; k < 7; k++) a[k] = from[k];
		} else {
			se3_segment_at(&sg, r.last_tau, a);
		}
		if (r.early) {
			for (int k = 0; k < 7; k++) b[k] = to[k];
		} else if (r.done) {
			for (int k = 0; k < 7; k++) b[k] = a[k];
		} else {
			se3_segment_at(&sg, r.tau, b);
		}
		double va = a[0], vb = b[0];
		for (int k = 1; k < 7; k++)
			if (lane == k) {
				va = a[k];
				vb = b[k];
			}
		o.last_free[7 * seg + lane] = va;
		o.state[7 * seg + lane] = vb;
	}
}

/* Row of the upper-triangle pair p in an n x n matrix (pairs fi < ti in
 * row-major order). */
__device__ __forceinline__ void tri_pair(int64_t p, int64_t n, int64_t& fi, int64_t& ti)
{
	double b = 2.0 * (double)n - 1.0;
	int64_t i = (int64_t)floor((b - sqrt(b * b - 8.0 * (double)p)) * 0.5);
	if (i < 0) i = 0;
	if (i > n - 2) i = n - 2;
	while (i > 0 && i * (2 * n - i - 1) / 2 > p) i--;
	while ((i + 1) * (2 * n - i - 2) / 2 <= p) i++;
	fi = i;
	ti = p - i * (2 * n - i - 1) / 2 + i + 1;
}

/* Visibility: one wavefront per pair.  out8: int8 matrix (or pair vector),
 * bits: bit-packed row-major n x n (upper and lower triangles), either may be
 * null. */
template <bool WIDE, int VIS>
__global__ void __launch_bounds__(256) k_visibility(DevScene s, const double* __restrict__ qs0,
                                                    const double* __restrict__ qs1, int64_t m, int64_t n,
                                                    int64_t npairs, double ivd, int8_t* __restrict__ out8,
                                                    uint32_t* __restrict__ bits, int32_t* __restrict__ steps)
{
	extern __shared__ uint32_t smem[];
	const int64_t p = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
	if (p >= npairs)
		return;
	int64_t fi, ti;
	if (VIS == VIS_MATRIX) {
		tri_pair(p, n, fi, ti);
	} else if (VIS == VIS_MATRIX2) {
		fi = p / n;
		ti = p % n;
	} else {
		fi = p;
		ti = p;
	}
	double from[7], to[7];
#pragma unroll
	for (int k = 0; k < 7; k++) {
		from[k] = qs0[7 * fi + k];
		to[k] = qs1[7 * ti + k];
	}
	se3_segment sg;
	const SegResult r = segment_wave<WIDE, MODE_TRANSITION>(s, from, to, ivd, 0, sg, smem);
	const uint8_t done = (uint8_t)r.done;
	const int32_t st = r.steps;
	const int lane = threadIdx.x & 63;
	if (lane == 0) {
		if (VIS == VIS_MATRIX) {
			if (out8) {
				out8[fi * n + ti] = (int8_t)done;
				out8[ti * n + fi] = (int8_t)done;
			}
			if (bits && done) {
				int64_t b1 = fi * n + ti, b2 = ti * n + fi;
				atomicOr(bits + (b1 >> 5), 1u << (b1 & 31));
				atomicOr(bits + (b2 >> 5), 1u << (b2 & 31));
			}
		} else if (VIS == VIS_MATRIX2) {
			if (out8) out8[fi * n + ti] = (int8_t)done;
			if (bits && done) {
				int64_t b1 = fi * n + ti;
				atomicOr(bits + (b1 >> 5), 1u << (b1 & 31));
			}
		} else {
			if (out8) out8[p] = (int8_t)done;
		}
		if (steps) steps[p] = st;
	}
}

/* ---- launch wrappers (host) ---- */
static inline size_t stack_bytes_per_thread(const DevScene& s)
{
	return (size_t)s.stack_depth * (s.wide ? 2 : 1) * sizeof(uint32_t);
}

/* Largest block (multiple of 64, at most 256) whose traversal stacks fit in
 * 48 KiB of LDS. */
static inline int pick_block(const DevScene& s)
{
	int block = 256;
	while (block > 64 && (size_t)block * stack_bytes_per_thread(s) > 48 * 1024)
		block >>= 1;
	return block;
}

static inline size_t stack_bytes(const DevScene& s, int block) { return stack_bytes_per_thread(s) * (size_t)block; }

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
	return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
	return hipGetLastError();
}

hipError_t launch_valid_states(const DevScene& s, const double* q, int64_t n, uint8_t* out, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	const int block = pick_block(s);
	const int64_t grid = (n + block - 1) / block;
	const size_t lds = stack_bytes(s, block);
	if (s.wide)
		hipLaunchKernelGGL(k_valid_states<true>, dim3((unsigned)grid), dim3(block), lds, st, s, q, n, out);
	else
		hipLaunchKernelGGL(k_valid_states<false>, dim3((unsigned)grid), dim3(block), lds, st, s, q, n, out);
	return hipGetLastError();
}

template <int MODE>
static hipError_t launch_seg_mode(const DevScene& s, const double* from, const double* to, int64_t n, double vd,
                                  int levels, const SegOut& o, hipStream_t st)
{
	const int block = pick_block(s);
	const int per_block = block / 64;
	const int64_t grid = (n + per_block - 1) / per_block;
	const size_t lds = stack_bytes(s, block);
	if (s.wide)
		hipLaunchKernelGGL((k_segments<true, MODE>), dim3((unsigned)grid), dim3(block), lds, st, s, from, to, n, vd,
		                   levels, o);
	else
		hipLaunchKernelGGL((k_segments<false, MODE>), dim3((unsigned)grid), dim3(block), lds, st, s, from, to, n, vd,
		                   levels, o);
	return hipGetLastError();
}

hipError_t launch_segments(const DevScene& s, int mode, const double* from, const double* to, int64_t n, double vd,
                           int levels, const SegOut& o, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	switch (mode) {
	case MODE_TRANSIT: return launch_seg_mode<MODE_TRANSIT>(s, from, to, n, vd, levels, o, st);
	case MODE_TRANSITION: return launch_seg_mode<MODE_TRANSITION>(s, from, to, n, vd, levels, o, st);
	case MODE_TOUCH: return launch_seg_mode<MODE_TOUCH>(s, from, to, n, vd, levels, o, st);
	default: return hipErrorInvalidValue;
	}
}

template <int VIS>
static hipError_t launch_vis_mode(const DevScene& s, const double* qs0, const double* qs1, int64_t m, int64_t n,
                                  int64_t npairs, double ivd, int8_t* out8, uint32_t* bits, int32_t* steps,
                                  hipStream_t st)
{
	const int block = pick_block(s);
	const int per_block = block / 64;
	const int64_t grid = (npairs + per_block - 1) / per_block;
	const size_t lds = stack_bytes(s, block);
	if (s.wide)
		hipLaunchKernelGGL((k_visibility<true, VIS>), dim3((unsigned)grid), dim3(block), lds, st, s, qs0, qs1, m, n,
		                   npairs, ivd, out8, bits, steps);
	else
		hipLaunchKernelGGL((k_visibility<false, VIS>), dim3((unsigned)grid), dim3(block), lds, st, s, qs0, qs1, m, n,
		                   npairs, ivd, out8, bits, steps);
	return hipGetLastError();
}

hipError_t launch_visibility(const DevScene& s, int vis, const double* qs0, const double* qs1, int64_t m, int64_t n,
                             double ivd, int8_t* out8, uint32_t* bits, int32_t* steps, hipStream_t st)
{
	int64_t npairs = 0;
	if (vis == VIS_MATRIX)
		npairs = n * (n - 1) / 2;
	else if (vis == VIS_MATRIX2)
		npairs = m * n;
	else
		npairs = n;
	if (npairs <= 0)
		return hipSuccess;
	switch (vis) {
	case VIS_PAIR: return launch_vis_mode<VIS_PAIR>(s, qs0, qs1, m, n, npairs, ivd, out8, bits, steps, st);
	case VIS_MATRIX: return launch_vis_mode<VIS_MATRIX>(s, qs0, qs1, m, n, npairs, ivd, out8, bits, steps, st);
	case VIS_MATRIX2: return launch_vis_mode<VIS_MATRIX2>(s, qs0, qs1, m, n, npairs, ivd, out8, bits, steps, st);
	default: return hipErrorInvalidValue;
	}
}

} // namespace se3cq

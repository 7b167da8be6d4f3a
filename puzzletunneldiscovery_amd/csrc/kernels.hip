/*
 * kernels.hip -- gfx950 kernels of the SE(3) collision-query engine.
 *
 * Compiled with -ffp-contract=off: the binary64 leaf test and the state
 * arithmetic must round exactly like the CPU restatement of FCL.  The
 * binary32 box culling uses explicit fmaf and is conservative by
 * construction (margins below), so it never changes a boolean.
 *
 * Kernels (the query kernels up to k_batch are persistent: the grid holds
 * the resident wavefronts, which take work from a per-launch counter; the
 * last three are one thread per state)
 *   k_valid_states   one state per lane, lanes refilled as they finish:
 *                    BVH x BVH depth-first traversal (UnitWorld::isValid,
 *                    lib/osr/unit_world.cc:215-247 -> fcl::collide via
 *                    lib/osr/cdmodel.cc:162-178); states that run long are
 *                    finished by the whole wavefront.
 *   k_segments       one wavefront per segment at a time: the 64 lanes
 *                    evaluate 64 consecutive discrete steps in one packet
 *                    traversal, which finds the first colliding step
 *                    (transitStateToWithContact, unit_world.cc:405-474;
 *                    isValidTransition 477-486), and the touch search
 *                    refines contact with an exact 5-levels-per-round
 *                    parallel bisection.
 *   k_visibility     one wavefront per state pair (calculateVisibility*,
 *                    unit_world.cc:740-840).
 *   k_ompl_states,   OMPL-convention states and motions
 *   k_ompl_motions   (lib/pyse3ompl/ompldriver.cc:566-620).
 *   k_batch          uneven segments and single states in one launch.
 *   k_contacts,      every intersecting triangle pair of a state and its
 *   k_contact_segments  intersection segment (unit_world.cc:959-1012).
 *   k_disentangled   the two models' boxes (unit_world.cc:249-264).
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine.h"
#include "launch.h"
#include "se3_math.h"
#include "cull.h"

/* Minimum resident wavefronts per SIMD each traversal kernel is compiled
 * for (caps its VGPR budget; 3 waves = 168 VGPRs, 4 = 128, 6 = 80, 7 = 72).
 * The segment walks wait on dependent node loads, and for the plain segment
 * and visibility walks more resident waves hide that better than the
 * registers they give up, even with the spills that follow.  The touch
 * search, the OMPL motions and the mixed batch are fastest at 4, and the
 * per-lane state kernels at 3, where they do not spill. */
#ifndef SE3CQ_MIN_WAVES
#define SE3CQ_MIN_WAVES 4
#endif
#ifndef SE3CQ_STATES_WAVES
#define SE3CQ_STATES_WAVES 3
#endif
#ifndef SE3CQ_SEG_WAVES
#define SE3CQ_SEG_WAVES 6
#endif
#ifndef SE3CQ_VIS_WAVES
#define SE3CQ_VIS_WAVES 6
#endif
#ifndef SE3CQ_TOUCH_WAVES
#define SE3CQ_TOUCH_WAVES 4
#endif
#ifndef SE3CQ_OMPL_WAVES
#define SE3CQ_OMPL_WAVES 4
#endif

namespace se3cq {

enum { MODE_TRANSIT = 0, MODE_TRANSITION = 1, MODE_TOUCH = 2 };
enum { VIS_PAIR = 0, VIS_MATRIX = 1, VIS_MATRIX2 = 2 };

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
/* project6: the two triangles' intervals on axis ax overlap (closed). */
__device__ __forceinline__ bool overlap6(const double* ax, const double* p1, const double* p2,
                                         const double* p3, const double* q1, const double* q2,
                                         const double* q3)
{
	const double P1 = dot3d(ax, p1), P2 = dot3d(ax, p2), P3 = dot3d(ax, p3);
	const double Q1 = dot3d(ax, q1), Q2 = dot3d(ax, q2), Q3 = dot3d(ax, q3);
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

/*
 * 17-axis SAT on triangles already translated by -P1 (p1 is the zero vector,
 * as FCL's P1 - P1).  Edges are formed on the fly (e1 = p2 - p1,
 * e2 = p3 - p2, e3 = p1 - p3, f likewise) so only the six points and the two
 * face normals stay live.  Axis order and arithmetic follow FCL's
 * intersect_Triangle; the boolean does not depend on the order.
 */
__device__ __forceinline__ bool sat17(const double* p1, const double* p2, const double* p3, const double* q1,
                                      const double* q2, const double* q3)
{
	double e1[3], e2[3], n1[3], m1[3], ax[3], u[3], v[3];
#pragma unroll
	for (int k = 0; k < 3; k++) {
		e1[k] = p2[k] - p1[k];
		e2[k] = p3[k] - p2[k];
	}
	cross3d(e1, e2, n1);
	if (!overlap6(n1, p1, p2, p3, q1, q2, q3)) return false;
#pragma unroll
	for (int k = 0; k < 3; k++) {
		u[k] = q2[k] - q1[k];
		v[k] = q3[k] - q2[k];
	}
	cross3d(u, v, m1);
	if (!overlap6(m1, p1, p2, p3, q1, q2, q3)) return false;
#pragma unroll 1
	for (int i = 0; i < 3; i++) {
		const double* a0 = (i == 0) ? p2 : (i == 1) ? p3 : p1;
		const double* a1 = (i == 0) ? p1 : (i == 1) ? p2 : p3;
#pragma unroll
		for (int k = 0; k < 3; k++)
			u[k] = a0[k] - a1[k];
#pragma unroll 1
		for (int j = 0; j < 3; j++) {
			const double* b0 = (j == 0) ? q2 : (j == 1) ? q3 : q1;
			const double* b1 = (j == 0) ? q1 : (j == 1) ? q2 : q3;
#pragma unroll
			for (int k = 0; k < 3; k++)
				v[k] = b0[k] - b1[k];
			cross3d(u, v, ax);
			if (!overlap6(ax, p1, p2, p3, q1, q2, q3)) return false;
		}
	}
#pragma unroll 1
	for (int i = 0; i < 3; i++) {
		const double* a0 = (i == 0) ? p2 : (i == 1) ? p3 : p1;
		const double* a1 = (i == 0) ? p1 : (i == 1) ? p2 : p3;
#pragma unroll
		for (int k = 0; k < 3; k++)
			u[k] = a0[k] - a1[k];
		cross3d(u, n1, ax);
		if (!overlap6(ax, p1, p2, p3, q1, q2, q3)) return false;
	}
#pragma unroll 1
	for (int j = 0; j < 3; j++) {
		const double* b0 = (j == 0) ? q2 : (j == 1) ? q3 : q1;
		const double* b1 = (j == 0) ? q1 : (j == 1) ? q2 : q3;
#pragma unroll
		for (int k = 0; k < 3; k++)
			v[k] = b0[k] - b1[k];
		cross3d(v, m1, ax);
		if (!overlap6(ax, p1, p2, p3, q1, q2, q3)) return false;
	}
	return true;
}

/* Exact binary64 map of robot vertices into the env frame for state q: the
 * FCL relative transform R = Re^T R(q), T = Re^T (t - te). */
struct ExactTf {
	double R[9], T[3];
};

__device__ __forceinline__ void make_exact(const DevScene& s, const double* q, ExactTf& x)
{
	double Rr[9];
	se3_quat_to_rotmat(q[3], q[4], q[5], q[6], Rr);
	if (s.pert_identity) {
		/* Re = I, te = 0: Re^T Rr and Re^T t equal Rr and t up to the sign
		 * of zero entries, which no comparison in the SAT can observe. */
#pragma unroll
		for (int i = 0; i < 9; i++)
			x.R[i] = Rr[i];
#pragma unroll
		for (int i = 0; i < 3; i++)
			x.T[i] = q[i];
		return;
	}
	const PertData* P = s.pert;
	const double d[3] = {q[0] - P->te[0], q[1] - P->te[1], q[2] - P->te[2]};
#pragma unroll
	for (int i = 0; i < 3; i++) {
#pragma unroll
		for (int j = 0; j < 3; j++)
			x.R[3 * i + j] = (P->Re[i] * Rr[j] + P->Re[3 + i] * Rr[3 + j]) + P->Re[6 + i] * Rr[6 + j];
		x.T[i] = (P->Re[i] * d[0] + P->Re[3 + i] * d[1]) + P->Re[6 + i] * d[2];
	}
}

/* Exact test of two leaves (defined with the packet traversal below). */
__device__ __forceinline__ bool leaf_pair_x(const DevScene& s, const ExactTf& x, int le, int lr);

/* The leaves (codes le, lr) at state q. */
__device__ __forceinline__ bool leaf_pair(const DevScene& s, const double* q, int le, int lr)
{
	ExactTf x;
	make_exact(s, q, x);
	return leaf_pair_x(s, x, le, lr);
}

/* Leaf code (the first triangle and the triangle count, see GNode) of leaf
 * node i. */
__device__ __forceinline__ int leaf_code(const GNode* __restrict__ nodes, int i)
{
	return ~__float_as_int(nodes[i].a2_child.w);
}

/* ---- depth-first BVH x BVH traversal, one thread per query ---- */
/* Stack entries and the pending-leaf buffer live in LDS, strided by the
 * block size so that lanes at the same depth touch consecutive banks.
 * Narrow scenes pack (env, rob) ids into one 32-bit word; wide scenes use two
 * words. */
enum { LEAF_BUF = 4 };

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
			const uint32_t v = base[k * stride];
			a = (int)(v >> 16);
			b = (int)(v & 0xffffu);
		}
	}
};

/* Where a traversal finds its exact binary64 state when a leaf test needs it. */
struct StateFromGlobal {
	const double* q;
	__device__ __forceinline__ void get(double* o) const
	{
#pragma unroll
		for (int k = 0; k < 7; k++)
			o[k] = q[k];
	}
};

struct StateFromOmpl {
	const double* q; /* OMPL order (x, y, z, qx, qy, qz, qw) */
	const OmplXf* x;
	__device__ __forceinline__ void get(double* o) const { se3_ompl_to_unit(x->M, x->pert, x->delta, q, o); }
};

struct StateFromSegment {
	const se3_segment* sg; /* LDS */
	double tau;
	__device__ __forceinline__ void get(double* o) const { se3_segment_at_fast(sg, tau, o); }
};

/*
 * Box pairs are tested in binary32; leaf-leaf pairs that survive are queued
 * (up to LEAF_BUF per thread) and tested exactly in a separate phase, so the
 * traversal's registers are free while the binary64 SAT runs and the lanes
 * of a wave run their SATs together.  Stops at the first intersecting pair.
 */
template <bool WIDE, class Src>
__device__ bool collide_dfs(const DevScene& s, const Src& src, const PairStack<WIDE>& stk,
                            const PairStack<WIDE>& leaves)
{
	int sp = 0;
	stk.put(sp++, 0, 0);
	const int cap = s.stack_depth;
	for (;;) {
		int nleaf = 0;
		{
			/* the culling transform lives only through the traversal phase */
			QCull c;
			{
				double q[7];
				src.get(q);
				make_qcull(s, q, c);
			}
			int cached_b = -1;
			RobBox B;
			while (sp > 0 && nleaf < LEAF_BUF) {
				int a, b;
				stk.get(--sp, a, b);
				const GNode A = load_node(s.env.nodes, a);
				if (b != cached_b) {
					const GNode Bn = load_node(s.rob.nodes, b);
					load_rob_box(Bn, c, B);
					cached_b = b;
				}
				if (obb_disjoint(A, B, c.eta))
					continue;
				const int ca = __float_as_int(A.a2_child.w);
				const bool la = ca < 0, lb = B.child < 0;
				if (la && lb) {
					leaves.put(nleaf++, a, b);
					continue;
				}
				if (sp + 2 > cap) {
					atomicOr(s.error_flag, 1u);
					return true;
				}
				/* FCL's rule: split the env node unless it is a leaf or the
				 * robot box is the larger one (squared extent norms). */
				const float sa = fmaf(A.c_ex.w, A.c_ex.w, fmaf(A.a0_ey.w, A.a0_ey.w, A.a1_ez.w * A.a1_ez.w));
				const float eb0 = B.e[0] - c.infl, eb1 = B.e[1] - c.infl, eb2 = B.e[2] - c.infl;
				const float sb = fmaf(eb0, eb0, fmaf(eb1, eb1, eb2 * eb2));
				if (!la && (lb || sa > sb)) {
					stk.put(sp++, ca + 1, b);
					stk.put(sp++, ca, b);
				} else {
					stk.put(sp++, a, B.child + 1);
					stk.put(sp++, a, B.child);
				}
			}
		}
		for (int i = 0; i < nleaf; i++) {
			int a, b;
			leaves.get(i, a, b);
			double q[7];
			src.get(q);
			if (leaf_pair(s, q, leaf_code(s.env.nodes, a), leaf_code(s.rob.nodes, b)))
				return true;
		}
		if (sp == 0)
			return false;
	}
}

/* LDS words per thread: traversal stack plus the pending-leaf buffer. */
__host__ __device__ __forceinline__ int lds_words_per_thread(const DevScene& s)
{
	return (s.stack_depth + LEAF_BUF) * (s.wide ? 2 : 1);
}

template <bool WIDE>
__device__ __forceinline__ void thread_stacks(const DevScene& s, uint32_t* smem, PairStack<WIDE>& stk,
                                              PairStack<WIDE>& leaves)
{
	const int stride = (int)blockDim.x;
	stk = PairStack<WIDE>{smem + threadIdx.x, stride};
	leaves = PairStack<WIDE>{smem + threadIdx.x + (size_t)s.stack_depth * (WIDE ? 2 : 1) * stride, stride};
}

/* Where a state query finds its binary64 unit-world state i. */
struct StatesPlain {
	const double* qs;
	__device__ __forceinline__ void get(int64_t i, double* o) const
	{
#pragma unroll
		for (int k = 0; k < 7; k++)
			o[k] = qs[7 * i + k];
	}
};

/* States picked by an index list (a prescreen's survivors): item i is
 * state idx[i]; an index outside [0, n) reads state 0 and is never stored. */
struct StatesSubset {
	const double* qs;
	const int32_t* idx;
	int64_t n;
	__device__ __forceinline__ void get(int64_t i, double* o) const
	{
		int64_t j = idx[i];
		if (j < 0 || j >= n)
			j = 0;
#pragma unroll
		for (int k = 0; k < 7; k++)
			o[k] = qs[7 * j + k];
	}
};

struct StatesOmpl {
	const double* qs; /* OMPL order (x, y, z, qx, qy, qz, qw) */
	const OmplXf* x;
	__device__ __forceinline__ void get(int64_t i, double* o) const
	{
		se3_ompl_to_unit(x->M, x->pert, x->delta, qs + 7 * i, o);
	}
};

#ifndef SE3CQ_REFILL_MIN
#define SE3CQ_REFILL_MIN 16
#endif
#ifndef SE3CQ_LEAF_PATIENCE
#define SE3CQ_LEAF_PATIENCE 16
#endif
#ifndef SE3CQ_REFILL_CHUNK
#define SE3CQ_REFILL_CHUNK 64
#endif
enum { REFILL_MIN = SE3CQ_REFILL_MIN, LEAF_PATIENCE = SE3CQ_LEAF_PATIENCE, REFILL_CHUNK = SE3CQ_REFILL_CHUNK };

__device__ __forceinline__ int lane_rank(uint64_t m);

/*
 * State validity with lane refill.  Independent states traverse very
 * differently: most are culled at the root within a step, a few near the
 * surface walk deep, and a wavefront holding one state per lane runs until
 * its slowest lane ends (7 % of VALU lanes were doing work).  Here every lane
 * keeps one state in flight: once REFILL_MIN lanes are idle they take the
 * next states from the wavefront's pool, which takes REFILL_CHUNK states at
 * a time from the launch's counter (one atomic per chunk), and each
 * pass of the loop is, wave-uniformly, either a traversal pass (one node
 * pair per traversing lane) or a leaf pass (one exact binary64 SAT per lane
 * whose leaf buffer is full or whose traversal is over), so the expensive
 * leaf test never diverges against traversal.  A leaf pass runs once
 * REFILL_MIN lanes wait for one, when no lane traverses, or after
 * LEAF_PATIENCE traversal passes, so no lane waits without bound.  The
 * answer, "some triangle pair intersects", does not depend on the order
 * pairs are visited in.  set(i, hit) stores state i's answer.
 */
/* defined with the packet traversal below */
__device__ __forceinline__ bool leaf_pair_x(const DevScene& s, const ExactTf& x, int te, int tr);
__device__ __forceinline__ void wave_lds_sync();
__device__ __forceinline__ int rdl(int v, int k);
__device__ __forceinline__ int wrl(int val, int k, int old);

/* Where a one-lane traversal finds state i of a state loader. */
template <class Loader>
struct StateFromLoader {
	const Loader* ld;
	int64_t i;
	__device__ __forceinline__ void get(double* o) const { ld->get(i, o); }
};

/* Entry e of a pair stack shared by the whole wavefront: slot e / 64 of lane
 * e % 64's traversal stack and leaf buffer (own = the calling lane's). */
template <bool WIDE>
__device__ __forceinline__ PairStack<WIDE> wave_entry(const PairStack<WIDE>& own, int e)
{
	return PairStack<WIDE>{own.base - (threadIdx.x & 63) + (e & 63), own.stride};
}

/*
 * One state's traversal by the whole wavefront, for states whose one-lane
 * traversal ran long.  Each step pops up to 64 pairs off a stack shared by
 * the wavefront (its lanes' traversal stacks, free by then), one per lane,
 * tests them with the state's culling transform, tests the leaf pairs among
 * them exactly, and pushes the children of the overlapping inner pairs.
 * The answer is the one-lane traversal's: whether some triangle pair
 * intersects.  Host replays of the heaviest states of a million random
 * torus states (up to 1,597 pair tests one lane at a time) take at most 29
 * steps and 304 stack entries.  Should the shared stack fill up, lane 0
 * walks the state alone.
 */
template <bool WIDE, class Loader>
__device__ bool coop_collide(const DevScene& s, const Loader& ld, int64_t i, const PairStack<WIDE>& stk,
                             const PairStack<WIDE>& leaves)
{
	const int lane = threadIdx.x & 63;
	const int cap = 64 * (s.stack_depth + LEAF_BUF);
	QCull c;
	{
		double q[7];
		ld.get(i, q);
		make_qcull(s, q, c);
	}
	if (lane == 0)
		wave_entry(stk, 0).put(0, 0, 0);
	int sp = 1;
	wave_lds_sync();
	for (;;) {
		const int nproc = sp < 64 ? sp : 64;
		const bool act = lane < nproc;
		int a = 0, b = 0;
		if (act) {
			const int e = sp - 1 - lane;
			wave_entry(stk, e).get(e >> 6, a, b);
		}
		sp -= nproc;
		wave_lds_sync();
		bool leaf = false, inner = false;
		int a1 = 0, b1 = 0, a2 = 0, b2 = 0;
		if (act) {
			const GNode A = load_node(s.env.nodes, a);
			RobBox B;
			load_rob_box(load_node(s.rob.nodes, b), c, B);
			if (!obb_disjoint(A, B, c.eta)) {
				const int ca = __float_as_int(A.a2_child.w);
				const bool la = ca < 0, lb = B.child < 0;
				if (la && lb) {
					leaf = true;
					a1 = ~ca;
					b1 = ~B.child;
				} else {
					inner = true;
					/* FCL's rule, as in the one-lane traversal */
					const float sa = fmaf(A.c_ex.w, A.c_ex.w, fmaf(A.a0_ey.w, A.a0_ey.w, A.a1_ez.w * A.a1_ez.w));
					const float eb0 = B.e[0] - c.infl, eb1 = B.e[1] - c.infl, eb2 = B.e[2] - c.infl;
					const float sb = fmaf(eb0, eb0, fmaf(eb1, eb1, eb2 * eb2));
					if (!la && (lb || sa > sb)) {
						a1 = ca + 1; b1 = b; a2 = ca; b2 = b;
					} else {
						a1 = a; b1 = B.child + 1; a2 = a; b2 = B.child;
					}
				}
			}
		}
		if (__ballot(leaf)) {
			ExactTf x;
			{
				double q[7];
				ld.get(i, q);
				make_exact(s, q, x);
			}
			const bool h = leaf && leaf_pair_x(s, x, a1, b1);
			if (__ballot(h))
				return true;
		}
		const uint64_t mi = __ballot(inner);
		const int ni = __popcll(mi);
		if (sp + 2 * ni > cap) {
			/* the shared stack is full: lane 0 starts over alone */
			bool r = false;
			if (lane == 0)
				r = collide_dfs<WIDE>(s, StateFromLoader<Loader>{&ld, i}, stk, leaves);
			return __ballot(r) != 0;
		}
		if (inner) {
			const int e = sp + 2 * lane_rank(mi);
			wave_entry(stk, e).put(e >> 6, a1, b1);
			wave_entry(stk, e + 1).put((e + 1) >> 6, a2, b2);
		}
		sp += 2 * ni;
		wave_lds_sync();
		if (sp == 0)
			return false;
	}
}

/* A lane whose state has taken more pair tests than SE3CQ_HEAVY_PER_LEVEL
 * times the scene's stack depth (the two trees' depths plus 2) sets it aside
 * (at most SE3CQ_HEAVY_MAX per wavefront) and takes a new one; the
 * wavefront finishes the states set aside with coop_collide once its last
 * lane is done. */
#ifndef SE3CQ_HEAVY_PER_LEVEL
#define SE3CQ_HEAVY_PER_LEVEL 14
#endif
#ifndef SE3CQ_HEAVY_MAX
#define SE3CQ_HEAVY_MAX 12
#endif
/* Once the launch's counter is spent and the wavefront's pool is empty, no
 * lane takes a new state, and a lane still walking one alone holds the
 * wavefront: states past SE3CQ_HEAVY_TAIL times the stack depth (if that is
 * the lower limit) are set aside then too, up to SE3CQ_HEAVY_TAIL_MAX per
 * wavefront, and all states set aside are walked together by
 * coop_collide_many. */
#ifndef SE3CQ_HEAVY_TAIL
#define SE3CQ_HEAVY_TAIL 8
#endif
#ifndef SE3CQ_HEAVY_TAIL_MAX
#define SE3CQ_HEAVY_TAIL_MAX 64
#endif
/* In that tail, once at most SE3CQ_TAIL_LANES lanes still walk a state, all
 * of those states are set aside as well: the joint walk takes 64 node
 * pairs per pass, where a lone lane's walk is one dependent chain. */
#ifndef SE3CQ_TAIL_LANES
#define SE3CQ_TAIL_LANES 8
#endif
static_assert(SE3CQ_HEAVY_TAIL_MAX <= 64 && SE3CQ_HEAVY_MAX <= SE3CQ_HEAVY_TAIL_MAX, "set-aside states live one per lane");

/* OR of a 64-bit value over the wavefront's lanes. */
__device__ __forceinline__ uint64_t wave_or64(uint64_t v)
{
	uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) {
		lo |= (uint32_t)__shfl_xor((int)lo, off, 64);
		hi |= (uint32_t)__shfl_xor((int)hi, off, 64);
	}
	return ((uint64_t)hi << 32) | lo;
}

/*
 * Several states, each "some triangle pair intersects", decided together
 * by the whole wavefront: slot j (j < nh) is the state whose index lane j
 * of (hv_hi, hv_lo) holds.  A LIFO stack of (environment node, slot and
 * robot node) entries, two words each, fills the lanes' traversal stacks
 * (free by then); every pass pops up to 64 entries of any slots, one per
 * lane, reads the entry's culling transform from its slot's lane (computed
 * once per slot), tests the pair, tests it exactly if both are leaves, and
 * pushes the children of an overlapping inner pair.  Entries of a slot
 * already found colliding are dropped.  Returns the mask of colliding
 * slots; *ok is false when the stack would overflow (the mask then holds
 * only the collisions found).  Robot node ids must stay below 2^26.
 */
template <bool WIDE, class Loader>
__device__ uint64_t coop_collide_many(const DevScene& s, const Loader& ld, int hv_lo, int hv_hi, int nh,
                                      const PairStack<WIDE>& stk, bool* ok)
{
	const int lane = threadIdx.x & 63;
	uint32_t* col0 = stk.base - lane; /* lane 0's column of this wavefront */
	const int stride = stk.stride;
	const int rows = (s.stack_depth + LEAF_BUF) * (WIDE ? 2 : 1);
	const int cap = 32 * rows; /* entries of two words */
	const uint64_t all = nh >= 64 ? ~0ull : ((1ull << nh) - 1ull);
	auto word = [&](int w) -> uint32_t* { return col0 + (w & 63) + (size_t)(w >> 6) * stride; };
	/* lane j keeps slot j's culling transform; a pass reads each entry's
	 * from its slot's lane */
	QCull mc{};
	if (lane < nh) {
		const int64_t i = (int64_t)(((uint64_t)(uint32_t)hv_hi << 32) | (uint64_t)(uint32_t)hv_lo);
		double q[7];
		ld.get(i, q);
		make_qcull(s, q, mc);
		*word(2 * lane) = 0u;
		*word(2 * lane + 1) = (uint32_t)lane << 26;
	}
	int sp = nh;
	uint64_t hit = 0;
	wave_lds_sync();
	for (;;) {
		const int nproc = sp < 64 ? sp : 64;
		int a = 0, b = 0, slot = 0;
		if (lane < nproc) {
			const int e = sp - 1 - lane;
			a = (int)*word(2 * e);
			const uint32_t w1 = *word(2 * e + 1);
			slot = (int)(w1 >> 26);
			b = (int)(w1 & ((1u << 26) - 1u));
		}
		sp -= nproc;
		wave_lds_sync();
		/* every lane takes part in the shuffles */
		QCull c;
#pragma unroll
		for (int k = 0; k < 9; k++)
			c.O[k] = __shfl(mc.O[k], slot, 64);
#pragma unroll
		for (int k = 0; k < 3; k++)
			c.T[k] = __shfl(mc.T[k], slot, 64);
		c.infl = __shfl(mc.infl, slot, 64);
		c.eta = __shfl(mc.eta, slot, 64);
		const int ilo = __shfl(hv_lo, slot, 64), ihi = __shfl(hv_hi, slot, 64);
		const bool live = lane < nproc && !((hit >> slot) & 1ull);
		bool inner = false, h = false;
		int a1 = 0, b1 = 0, a2 = 0, b2 = 0;
		if (live) {
			const GNode A = load_node(s.env.nodes, a);
			RobBox B;
			load_rob_box(load_node(s.rob.nodes, b), c, B);
			if (!obb_disjoint(A, B, c.eta)) {
				const int ca = __float_as_int(A.a2_child.w);
				const bool la = ca < 0, lb = B.child < 0;
				if (la && lb) {
					const int64_t i = (int64_t)(((uint64_t)(uint32_t)ihi << 32) | (uint64_t)(uint32_t)ilo);
					double q[7];
					ld.get(i, q);
					h = leaf_pair(s, q, ~ca, ~B.child);
				} else {
					inner = true;
					/* FCL's rule, as in the one-lane traversal */
					const float sa = fmaf(A.c_ex.w, A.c_ex.w, fmaf(A.a0_ey.w, A.a0_ey.w, A.a1_ez.w * A.a1_ez.w));
					const float eb0 = B.e[0] - c.infl, eb1 = B.e[1] - c.infl, eb2 = B.e[2] - c.infl;
					const float sb = fmaf(eb0, eb0, fmaf(eb1, eb1, eb2 * eb2));
					if (!la && (lb || sa > sb)) {
						a1 = ca + 1; b1 = b; a2 = ca; b2 = b;
					} else {
						a1 = a; b1 = B.child + 1; a2 = a; b2 = B.child;
					}
				}
			}
		}
		if (__ballot(h)) {
			hit |= wave_or64(h ? (1ull << slot) : 0ull);
			if ((hit & all) == all) {
				*ok = true;
				return hit;
			}
		}
		const uint64_t mi = __ballot(inner);
		const int ni = __popcll(mi);
		if (sp + 2 * ni > cap) {
			*ok = false;
			return hit;
		}
		if (inner) {
			const int e = sp + 2 * lane_rank(mi);
			const uint32_t tag = (uint32_t)slot << 26;
			*word(2 * e) = (uint32_t)a1;
			*word(2 * e + 1) = tag | (uint32_t)b1;
			*word(2 * e + 2) = (uint32_t)a2;
			*word(2 * e + 3) = tag | (uint32_t)b2;
		}
		sp += 2 * ni;
		wave_lds_sync();
		if (sp == 0) {
			*ok = true;
			return hit;
		}
	}
}


template <bool WIDE, class Loader, class Store>
__device__ void states_refill(const DevScene& s, const Loader& ld, int64_t n, const Store& set,
                              const PairStack<WIDE>& stk, const PairStack<WIDE>& leaves)
{
	const int lane = threadIdx.x & 63;
	const int cap = s.stack_depth;
	int64_t idx = -1;
	int sp = 0, nleaf = 0, cached_b = -1;
	QCull c;
	RobBox B;
	/* wave-uniform: the counter is spent, and the pool [pool_base,
	 * pool_base + pool_left) of states taken but not yet started */
	bool exhausted = false;
	int64_t pool_base = 0;
	int pool_left = 0;
	int since_leaf = 0;
	int tests = 0;
	/* states set aside for coop_collide: lane j of (hv_hi, hv_lo) holds the
	 * j-th index */
	int hv_lo = 0, hv_hi = 0, nheavy = 0;
	const int heavy_tests = SE3CQ_HEAVY_PER_LEVEL * s.stack_depth;
	const int tail_tests = SE3CQ_HEAVY_TAIL * s.stack_depth;
	/* the multi-state walk tags entries with 26-bit robot node ids */
	const bool many = s.rob.nnodes < (1 << 26);
	/* states per counter add (at least 1; the launch sets it) */
	const int chunk = s.pool_chunk > 0 ? s.pool_chunk : REFILL_CHUNK;
	for (;;) {
		{
			const bool tail = many && exhausted && pool_left == 0;
			const uint64_t busy = __ballot(idx >= 0);
			/* in the tail, once few lanes still walk, all of them stop and
			 * the wavefront walks their states together */
			const bool few = tail && __popcll(busy) <= SE3CQ_TAIL_LANES;
			const int lim = few ? -1 : (tail && tail_tests < heavy_tests ? tail_tests : heavy_tests);
			const int hmax = tail ? SE3CQ_HEAVY_TAIL_MAX : SE3CQ_HEAVY_MAX;
			uint64_t dm = __ballot(idx >= 0 && tests > lim);
			uint64_t taken = 0;
			while (dm && nheavy < hmax) {
				const int L = __ffsll((long long)dm) - 1;
				dm &= dm - 1;
				hv_lo = wrl(rdl((int)(uint32_t)idx, L), nheavy, hv_lo);
				hv_hi = wrl(rdl((int)(uint32_t)((uint64_t)idx >> 32), L), nheavy, hv_hi);
				nheavy++;
				taken |= 1ull << L;
			}
			if ((taken >> lane) & 1ull)
				idx = -1;
		}
		const uint64_t idle = __ballot(idx < 0);
		const int nidle = __popcll(idle);
		if ((nidle >= REFILL_MIN || nidle == 64) && (pool_left > 0 || !exhausted)) {
			int64_t base2 = 0;
			int left2 = 0;
			if (pool_left < nidle && !exhausted) {
				unsigned long long got = 0;
				if (lane == 0)
					got = atomicAdd(s.queue, (unsigned long long)chunk);
				base2 = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(got >> 32)) << 32) |
				                  (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)got));
				if (base2 + chunk >= n)
					exhausted = true;
				left2 = base2 < n ? (int)((n - base2) < chunk ? (n - base2) : chunk) : 0;
			}
			const int take = nidle < pool_left + left2 ? nidle : pool_left + left2;
			if (idx < 0) {
				const int r = lane_rank(idle);
				int64_t i = -1;
				if (r < pool_left)
					i = pool_base + r;
				else if (r - pool_left < left2)
					i = base2 + (r - pool_left);
				if (i >= 0) {
					idx = i;
					double q[7];
					ld.get(i, q);
					make_qcull(s, q, c);
					sp = 0;
					stk.put(sp++, 0, 0);
					nleaf = 0;
					cached_b = -1;
					tests = 0;
				}
			}
			if (take <= pool_left) {
				pool_base += take;
				pool_left -= take;
			} else {
				pool_base = base2 + (take - pool_left);
				pool_left = left2 - (take - pool_left);
			}
		}
		const bool act = idx >= 0;
		const uint64_t am = __ballot(act);
		if (am == 0) {
			if (exhausted && pool_left == 0)
				break;
			continue;
		}

		const bool want_leaf = act && (nleaf == LEAF_BUF || (sp == 0 && nleaf > 0));
		const uint64_t wl = __ballot(want_leaf);
		if (wl != 0 && (__popcll(wl) >= REFILL_MIN || wl == am || since_leaf >= LEAF_PATIENCE)) {
			since_leaf = 0;
			if (want_leaf) {
				int a, b;
				leaves.get(--nleaf, a, b);
				double q[7];
				ld.get(idx, q);
				if (leaf_pair(s, q, leaf_code(s.env.nodes, a), leaf_code(s.rob.nodes, b))) {
					set(idx, true);
					idx = -1;
				} else if (sp == 0 && nleaf == 0) {
					set(idx, false);
					idx = -1;
				}
			}
			continue;
		}
		since_leaf++;
		if (!act || want_leaf)
			continue;
		int a, b;
		stk.get(--sp, a, b);
		tests++;
		const GNode A = load_node(s.env.nodes, a);
		if (b != cached_b) {
			const GNode Bn = load_node(s.rob.nodes, b);
			load_rob_box(Bn, c, B);
			cached_b = b;
		}
		if (!obb_disjoint(A, B, c.eta)) {
			const int ca = __float_as_int(A.a2_child.w);
			const bool la = ca < 0, lb = B.child < 0;
			if (la && lb) {
				leaves.put(nleaf++, a, b);
			} else if (sp + 2 > cap) {
				atomicOr(s.error_flag, 1u);
				set(idx, true);
				idx = -1;
				continue;
			} else {
				/* FCL's rule: split the env node unless it is a leaf or the
				 * robot box is the larger one (squared extent norms). */
				const float sa = fmaf(A.c_ex.w, A.c_ex.w, fmaf(A.a0_ey.w, A.a0_ey.w, A.a1_ez.w * A.a1_ez.w));
				const float eb0 = B.e[0] - c.infl, eb1 = B.e[1] - c.infl, eb2 = B.e[2] - c.infl;
				const float sb = fmaf(eb0, eb0, fmaf(eb1, eb1, eb2 * eb2));
				if (!la && (lb || sa > sb)) {
					stk.put(sp++, ca + 1, b);
					stk.put(sp++, ca, b);
				} else {
					stk.put(sp++, a, B.child + 1);
					stk.put(sp++, a, B.child);
				}
			}
		}
		if (sp == 0 && nleaf == 0) {
			set(idx, false);
			idx = -1;
		}
	}
	/* every lane is idle and the states have run out: the whole wavefront
	 * now walks the states set aside, together; should their stack
	 * overflow, each undecided one alone */
	if (nheavy == 0)
		return;
	/* decided: slots whose answer is known; hit: those found colliding */
	uint64_t decided = 0, hit = 0;
	if (many) {
		bool ok = false;
		hit = coop_collide_many<WIDE>(s, ld, hv_lo, hv_hi, nheavy, stk, &ok);
		decided = ok ? (nheavy >= 64 ? ~0ull : ((1ull << nheavy) - 1ull)) : hit;
		wave_lds_sync();
	}
	if (lane < nheavy && ((decided >> lane) & 1ull)) {
		const int64_t i = (int64_t)(((uint64_t)(uint32_t)hv_hi << 32) | (uint64_t)(uint32_t)hv_lo);
		set(i, (hit >> lane) & 1ull);
	}
	/* should the shared stack have overflowed: each undecided state alone */
	for (int j = 0; j < nheavy; j++) {
		if ((decided >> j) & 1ull)
			continue;
		const int64_t i = (int64_t)(((uint64_t)(uint32_t)rdl(hv_hi, j) << 32) | (uint64_t)(uint32_t)rdl(hv_lo, j));
		const bool h = coop_collide<WIDE>(s, ld, i, stk, leaves);
		if (lane == 0)
			set(i, h);
	}
}

struct StoreValidU8 {
	uint8_t* out;
	__device__ __forceinline__ void operator()(int64_t i, bool hit) const { out[i] = hit ? 0 : 1; }
};

struct StoreValidSubset {
	uint8_t* out;
	const int32_t* idx;
	int64_t n;
	__device__ __forceinline__ void operator()(int64_t i, bool hit) const
	{
		const int64_t j = idx[i];
		if (j >= 0 && j < n)
			out[j] = hit ? 0 : 1;
	}
};

struct StoreValidI32 {
	int32_t* out;
	__device__ __forceinline__ void operator()(int64_t i, bool hit) const { out[i] = hit ? 0 : 1; }
};

template <bool WIDE>
__global__ void __launch_bounds__(256, SE3CQ_STATES_WAVES) k_valid_states(DevScene s, const double* __restrict__ qs, int64_t n,
                                                      uint8_t* __restrict__ out)
{
	extern __shared__ uint32_t smem[];
	PairStack<WIDE> stk, leaves;
	thread_stacks<WIDE>(s, smem, stk, leaves);
	states_refill<WIDE>(s, StatesPlain{qs}, n, StoreValidU8{out}, stk, leaves);
}

/* The states a prescreen kept: items [0, min(*count, n)) of idx, each an
 * index into qs; out[idx[i]] is written, other entries are left alone. */
template <bool WIDE>
__global__ void __launch_bounds__(256, SE3CQ_STATES_WAVES) k_valid_states_subset(DevScene s, const double* __restrict__ qs,
                                                      int64_t n, const int32_t* __restrict__ idx,
                                                      const uint32_t* __restrict__ count, uint8_t* __restrict__ out)
{
	extern __shared__ uint32_t smem[];
	PairStack<WIDE> stk, leaves;
	thread_stacks<WIDE>(s, smem, stk, leaves);
	const int64_t c = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)*count);
	const int64_t items = c < n ? c : n;
	states_refill<WIDE>(s, StatesSubset{qs, idx, n}, items, StoreValidSubset{out, idx, n}, stk, leaves);
}

/* OMPL-convention states (OmplDriver::validateStates,
 * lib/pyse3ompl/ompldriver.cc:566-589): converted to the unit world on load. */
template <bool WIDE>
__global__ void __launch_bounds__(256, SE3CQ_STATES_WAVES) k_ompl_states(DevScene s, OmplXf x, const double* __restrict__ qs,
                                                     int64_t n, int32_t* __restrict__ out)
{
	extern __shared__ uint32_t smem[];
	PairStack<WIDE> stk, leaves;
	thread_stacks<WIDE>(s, smem, stk, leaves);
	states_refill<WIDE>(s, StatesOmpl{qs, &x}, n, StoreValidI32{out}, stk, leaves);
}

/* ---- packet traversal: one wavefront, one pose per lane ---- */
/*
 * Segment steps, bisection probes and visibility pairs hand a wavefront 64
 * nearly identical poses.  The wave walks a single BVH x BVH traversal; each
 * stack entry carries the 64-bit mask of lanes whose boxes still overlap, a
 * pair is expanded if any lane overlaps it, and each lane tests only its own
 * pose.  Control flow and node addresses are wave-uniform, so nodes and
 * triangles are read with scalar loads (constant address space) and the
 * stack lives in VGPR lanes (entry k in lane k, writelane/readlane).
 */
__device__ __forceinline__ GNode load_node_uniform(const GNode* nodes, int i)
{
#if defined(__HIP_DEVICE_COMPILE__)
	const __attribute__((address_space(4))) float4* p =
	    (const __attribute__((address_space(4))) float4*)(nodes + i);
	GNode n;
	n.c_ex = p[0];
	n.a0_ey = p[1];
	n.a1_ez = p[2];
	n.a2_child = p[3];
	return n;
#else
	return nodes[i];
#endif
}

__device__ __forceinline__ int rdl(int v, int k) { return __builtin_amdgcn_readlane(v, k); }
/* llvm.amdgcn.writelane: put a wave-uniform value into lane k of a VGPR. */
__device__ int se3cq_amdgcn_writelane(int val, int lane, int old) __asm("llvm.amdgcn.writelane.i32");
__device__ __forceinline__ int wrl(int val, int k, int old) { return se3cq_amdgcn_writelane(val, k, old); }

/*
 * One (env triangle, robot triangle) pair under the exact transform x:
 * robot vertices R q + T, both triangles translated by -P1, then the SAT.
 * P1 - P1 is the zero vector for the finite coordinates the host admits, so
 * p1 is the literal 0.  Triangles are numbered in leaf order.
 */
__device__ __forceinline__ bool tri_pair_x(const DevScene& s, const ExactTf& x, int te, int tr)
{
	const double* P = s.env.tris + 9 * (int64_t)te;
	const double* qv = s.rob.tris + 9 * (int64_t)tr;
	const double P1[3] = {P[0], P[1], P[2]};
	const double p1[3] = {0.0, 0.0, 0.0};
	double p2[3], p3[3], q1[3], q2[3], q3[3];
#pragma unroll
	for (int k = 0; k < 3; k++) {
		p2[k] = P[3 + k] - P1[k];
		p3[k] = P[6 + k] - P1[k];
	}
	double* qq[3] = {q1, q2, q3};
#pragma unroll
	for (int v = 0; v < 3; v++) {
		const double a = qv[3 * v], b = qv[3 * v + 1], c = qv[3 * v + 2];
#pragma unroll
		for (int r = 0; r < 3; r++)
			qq[v][r] = (((x.R[3 * r] * a + x.R[3 * r + 1] * b) + x.R[3 * r + 2] * c) + x.T[r]) - P1[r];
	}
	return sat17(p1, p2, p3, q1, q2, q3);
}

/* Every triangle pair of two leaves (codes le, lr); true at the first that
 * intersects. */
__device__ __forceinline__ bool leaf_pair_x(const DevScene& s, const ExactTf& x, int le, int lr)
{
	const int e0 = le >> 3, ne = (le & 7) + 1, r0 = lr >> 3, nr = (lr & 7) + 1;
	for (int i = 0; i < ne; i++)
		for (int j = 0; j < nr; j++)
			if (tri_pair_x(s, x, e0 + i, r0 + j))
				return true;
	return false;
}

/*
 * Returns the mask of lanes in `live` whose pose (the exact binary64 state
 * from `pose`) puts the robot in collision.  With FIRST, only the lowest
 * colliding lane is asked for: once a hit is known, lanes above it are
 * dropped, and the mask returned is exact up to and including its lowest set
 * bit.  Traversal phases cull in binary32 and queue the (env triangle, robot
 * triangle, lane mask) leaf candidates in VGPR lanes; leaf phases run the
 * exact binary64 tests on the queue.  The culling transform is rebuilt at
 * the start of every traversal phase so it is dead while the SATs run.
 *
 * Here a stack entry is a node pair to test for the lanes of its mask; a
 * pop tests it and pushes both children of an overlapping inner pair.  In
 * packet_collide_eager a stack entry is a pair known to overlap for its
 * lanes (the root pair is tested first); a pop splits it by FCL's rule and
 * tests both children in the same pass, then pushes the overlapping inner
 * children and queues the overlapping leaf pairs: the two children's node
 * loads overlap, a child that overlaps for no lane never costs a pop (about
 * half of all tested pairs on the touch workload), and the pop's
 * bookkeeping is paid once per two tests.  Measured on MI355X, the eager
 * walk takes the visibility matrix from 2.78 to 2.50 s and OMPL motions
 * from 112 to 108 ms, leaves the touch search and uneven segments within
 * 1 %, and makes the plain transition walk of the shell scene slower (exact
 * segment checks 595 -> 615 ms, prescreen re-verification 157 -> 182 ms),
 * so the plain segment kernels keep this pair-at-a-time walk.
 *
 * With FIRST, pairs are taken in "focus" order: the topmost stack entry
 * holding the lowest lane that still has pending pairs.  Lanes are then
 * finished one after another from the bottom, so the pair that collides for
 * the lowest colliding lane is reached early and its hit drops every lane
 * above it.  Near contact, where the 64 consecutive poses straddle the first
 * colliding one, this takes about a third fewer pair tests than plain depth
 * first order, and half as many on the heaviest windows.  Focus order lets
 * entries of the lanes above pile up, so once the stack holds more than
 * 62 - stack_depth entries, pairs are popped from the top: a depth-first
 * descent from there adds at most stack_depth - 2 entries, which keeps the
 * stack within its 64 lanes.
 */
#ifndef SE3CQ_PACKET_LEAVES
#define SE3CQ_PACKET_LEAVES 8
#endif
enum { PACKET_LEAVES = SE3CQ_PACKET_LEAVES };  /* at most 64: one lane per entry */
static_assert(PACKET_LEAVES >= 2 && PACKET_LEAVES <= 64, "leaf entries live one per lane");
/* the touch search walks with packet_collide_eager (1) or packet_collide (0) */
#ifndef SE3CQ_TOUCH_EAGER
#define SE3CQ_TOUCH_EAGER 1
#endif

/*
 * Where the traversal gets its pose.  The pose is read at the two places
 * that need it (the culling transform of each traversal phase and the exact
 * transform of each leaf phase) instead of staying live in 14 VGPRs through
 * the traversal, which cuts the spills of the packet kernels by two thirds.
 *
 * LdsPose: the wavefront's 64 poses, evaluated once per packet and staged in
 * LDS as 7 rows of 64 binary64 values (lane l's pose is column l, so a row
 * read touches consecutive banks).  A read costs seven LDS loads instead of
 * the interpolation (two sines, two divisions) per phase.  The empty asm
 * makes the lane index opaque, so the compiler cannot hoist the loads out of
 * the traversal loop.
 */
enum { POSE_LDS_DOUBLES = 7 * 64, POSE_LDS_BYTES = POSE_LDS_DOUBLES * 8 };

struct LdsPose {
	const double* pq; /* LDS, POSE_LDS_DOUBLES */
	__device__ __forceinline__ void get(double* o) const
	{
		int l = threadIdx.x & 63;
		asm volatile("" : "+v"(l));
#pragma unroll
		for (int k = 0; k < 7; k++)
			o[k] = pq[64 * k + l];
	}
};

/* Stores each lane's pose q into the wavefront's pose rows (every lane of
 * the wavefront calls it). */
__device__ __forceinline__ void wave_lds_sync();
__device__ __forceinline__ void stage_pose(double* pq, const double* q)
{
	const int lane = threadIdx.x & 63;
#pragma unroll
	for (int k = 0; k < 7; k++)
		pq[64 * k + lane] = q[k];
	wave_lds_sync();
}

template <bool FIRST, class Pose>
__device__ uint64_t packet_collide(const DevScene& s, const Pose& pose, uint64_t live)
{
	const int lane = threadIdx.x & 63;
	int va = 0, vb = 0, vml = 0, vmh = 0;
	va = wrl(0, 0, va);
	vb = wrl(0, 0, vb);
	vml = wrl((int)(uint32_t)live, 0, vml);
	vmh = wrl((int)(uint32_t)(live >> 32), 0, vmh);
	int sp = 1;
	uint64_t hit = 0;
	int focus = __ffsll((long long)live) - 1;
	const int focus_cap = 62 - s.stack_depth;
	for (;;) {
		int lte = 0, ltr = 0, lml = 0, lmh = 0;
		int nl = 0;
		{
			QCull c;
			{
				double q[7];
				pose.get(q);
				make_qcull(s, q, c);
			}
			/* the robot box of the last robot node, reused while the env
			 * side is split */
			int cached_b = -1;
			RobBox B;
			while (sp > 0 && nl < PACKET_LEAVES) {
				int p = sp - 1;
				if (FIRST && sp <= focus_cap) {
					/* lanes below `focus` have no pending pairs; lanes at
					 * or above the lowest hit need none */
					uint64_t hb = 0;
					const int stop = hit ? __ffsll((long long)hit) - 1 : 64;
					while (focus < stop) {
						const int word = (focus < 32) ? vml : vmh;
						hb = __ballot(lane < sp && ((word >> (focus & 31)) & 1));
						if (hb)
							break;
						focus++;
					}
					if (!hb) {
						sp = 0;
						break;
					}
					p = 63 - __clzll((long long)hb);
				}
				const int a = rdl(va, p), b = rdl(vb, p);
				uint64_t m = ((uint64_t)(uint32_t)rdl(vmh, p) << 32) | (uint64_t)(uint32_t)rdl(vml, p);
				sp--;
				if (p != sp) {
					/* the top entry fills the hole */
					va = wrl(rdl(va, sp), p, va);
					vb = wrl(rdl(vb, sp), p, vb);
					vml = wrl(rdl(vml, sp), p, vml);
					vmh = wrl(rdl(vmh, sp), p, vmh);
				}
				m &= ~hit;
				if (FIRST && hit)
					m &= (hit & (~hit + 1ull)) - 1ull;
				if (m == 0)
					continue;
				const GNode A = load_node_uniform(s.env.nodes, a);
				if (b != cached_b) {
					load_rob_box(load_node_uniform(s.rob.nodes, b), c, B);
					cached_b = b;
				}
				const bool ov = ((m >> lane) & 1ull) && !obb_disjoint(A, B, c.eta);
				const uint64_t mo = __ballot(ov);
				if (mo == 0)
					continue;
				const int ca = __float_as_int(A.a2_child.w);
				const int cb = B.child;
				const bool la = ca < 0, lb = cb < 0;
				const int ml = (int)(uint32_t)mo, mh = (int)(uint32_t)(mo >> 32);
				if (la && lb) {
					lte = wrl(~ca, nl, lte);
					ltr = wrl(~cb, nl, ltr);
					lml = wrl(ml, nl, lml);
					lmh = wrl(mh, nl, lmh);
					nl++;
					continue;
				}
				if (sp + 2 > 64 || (!FIRST && sp + 2 > s.stack_depth)) {
					if (lane == 0)
						atomicOr(s.error_flag, 1u);
					return live;
				}
				/* FCL's rule: split the env node unless it is a leaf or the
				 * robot box is the larger one (squared extent norms). */
				const float sa = fmaf(A.c_ex.w, A.c_ex.w, fmaf(A.a0_ey.w, A.a0_ey.w, A.a1_ez.w * A.a1_ez.w));
				const float eb0 = B.e[0] - c.infl, eb1 = B.e[1] - c.infl, eb2 = B.e[2] - c.infl;
				const float sb = fmaf(eb0, eb0, fmaf(eb1, eb1, eb2 * eb2));
				int a1, b1, a2, b2;
				if (!la && (lb || sa > sb)) {
					a1 = ca + 1; b1 = b; a2 = ca; b2 = b;
				} else {
					a1 = a; b1 = cb + 1; a2 = a; b2 = cb;
				}
				va = wrl(a1, sp, va); vb = wrl(b1, sp, vb); vml = wrl(ml, sp, vml); vmh = wrl(mh, sp, vmh);
				sp++;
				va = wrl(a2, sp, va); vb = wrl(b2, sp, vb); vml = wrl(ml, sp, vml); vmh = wrl(mh, sp, vmh);
				sp++;
			}
		}
		if (nl > 0) {
			ExactTf x;
			{
				double q[7];
				pose.get(q);
				make_exact(s, q, x);
			}
			for (int i = 0; i < nl; i++) {
				uint64_t m = ((uint64_t)(uint32_t)rdl(lmh, i) << 32) | (uint64_t)(uint32_t)rdl(lml, i);
				m &= ~hit;
				if (FIRST && hit)
					m &= (hit & (~hit + 1ull)) - 1ull;
				if (m == 0)
					continue;
				const int te = rdl(lte, i), tr = rdl(ltr, i);
				const bool h = ((m >> lane) & 1ull) && leaf_pair_x(s, x, te, tr);
				hit |= __ballot(h);
			}
			if ((hit & live) == live || (FIRST && (hit & live & (~live + 1ull))))
				return hit & live;
		}
		if (sp == 0)
			return hit & live;
	}
}

/* The same answer as packet_collide, walking pairs known to overlap and
 * testing both children of each (see above). */
template <bool FIRST, class Pose>
__device__ uint64_t packet_collide_eager(const DevScene& s, const Pose& pose, uint64_t live)
{
	const int lane = threadIdx.x & 63;
	int va = 0, vb = 0, vml = 0, vmh = 0;
	va = wrl(0, 0, va);
	vb = wrl(0, 0, vb);
	vml = wrl((int)(uint32_t)live, 0, vml);
	vmh = wrl((int)(uint32_t)(live >> 32), 0, vmh);
	int sp = 1;
	uint64_t hit = 0;
	int focus = __ffsll((long long)live) - 1;
	const int focus_cap = 62 - s.stack_depth;
	bool root = true;
	for (;;) {
		int lte = 0, ltr = 0, lml = 0, lmh = 0;
		int nl = 0;
		{
			QCull c;
			{
				double q[7];
				pose.get(q);
				make_qcull(s, q, c);
			}
			/* the robot box of the last robot node, reused while the env
			 * side is split */
			int cached_b = -1;
			RobBox B;
			if (root) {
				/* stack entries are pairs known to overlap (for their lane
				 * mask): the root pair is tested first */
				root = false;
				const GNode A = load_node_uniform(s.env.nodes, 0);
				load_rob_box(load_node_uniform(s.rob.nodes, 0), c, B);
				cached_b = 0;
				const bool ov = ((live >> lane) & 1ull) && !obb_disjoint(A, B, c.eta);
				const uint64_t mo = __ballot(ov);
				sp = 0;
				if (mo) {
					const int ca = __float_as_int(A.a2_child.w);
					const int ml = (int)(uint32_t)mo, mh = (int)(uint32_t)(mo >> 32);
					if (ca < 0 && B.child < 0) {
						lte = wrl(~ca, 0, lte);
						ltr = wrl(~B.child, 0, ltr);
						lml = wrl(ml, 0, lml);
						lmh = wrl(mh, 0, lmh);
						nl = 1;
					} else {
						va = wrl(0, 0, va); vb = wrl(0, 0, vb); vml = wrl(ml, 0, vml); vmh = wrl(mh, 0, vmh);
						sp = 1;
					}
				}
			}
			while (sp > 0 && nl < PACKET_LEAVES - 1) {
				int p = sp - 1;
				if (FIRST && sp <= focus_cap) {
					/* lanes below `focus` have no pending pairs; lanes at
					 * or above the lowest hit need none */
					uint64_t hb = 0;
					const int stop = hit ? __ffsll((long long)hit) - 1 : 64;
					while (focus < stop) {
						const int word = (focus < 32) ? vml : vmh;
						hb = __ballot(lane < sp && ((word >> (focus & 31)) & 1));
						if (hb)
							break;
						focus++;
					}
					if (!hb) {
						sp = 0;
						break;
					}
					p = 63 - __clzll((long long)hb);
				}
				const int a = rdl(va, p), b = rdl(vb, p);
				uint64_t m = ((uint64_t)(uint32_t)rdl(vmh, p) << 32) | (uint64_t)(uint32_t)rdl(vml, p);
				sp--;
				if (p != sp) {
					/* the top entry fills the hole */
					va = wrl(rdl(va, sp), p, va);
					vb = wrl(rdl(vb, sp), p, vb);
					vml = wrl(rdl(vml, sp), p, vml);
					vmh = wrl(rdl(vmh, sp), p, vmh);
				}
				m &= ~hit;
				if (FIRST && hit)
					m &= (hit & (~hit + 1ull)) - 1ull;
				if (m == 0)
					continue;
				/* (a, b) overlaps for the lanes of m and is not a pair of
				 * leaves: test both of its children */
				const GNode A = load_node_uniform(s.env.nodes, a);
				const GNode Bn = load_node_uniform(s.rob.nodes, b);
				const int ca = __float_as_int(A.a2_child.w);
				const int cb = __float_as_int(Bn.a2_child.w);
				const bool la = ca < 0, lb = cb < 0;
				if (sp + 2 > 64 || (!FIRST && sp + 2 > s.stack_depth)) {
					if (lane == 0)
						atomicOr(s.error_flag, 1u);
					return live;
				}
				const bool in_m = (m >> lane) & 1ull;
				/* FCL's rule: split the env node unless it is a leaf or the
				 * robot box is the larger one (squared extent norms). */
				const float sa = fmaf(A.c_ex.w, A.c_ex.w, fmaf(A.a0_ey.w, A.a0_ey.w, A.a1_ez.w * A.a1_ez.w));
				const float sb = fmaf(Bn.c_ex.w, Bn.c_ex.w, fmaf(Bn.a0_ey.w, Bn.a0_ey.w, Bn.a1_ez.w * Bn.a1_ez.w));
				int a1, b1, a2, b2, c1a, c1b, c2a, c2b;
				bool ov1, ov2;
				if (!la && (lb || sa > sb)) {
					a1 = ca + 1; b1 = b; a2 = ca; b2 = b;
					if (b != cached_b) {
						load_rob_box(Bn, c, B);
						cached_b = b;
					}
					const GNode A1 = load_node_uniform(s.env.nodes, a1);
					const GNode A2 = load_node_uniform(s.env.nodes, a2);
					ov1 = in_m && !obb_disjoint(A1, B, c.eta);
					ov2 = in_m && !obb_disjoint(A2, B, c.eta);
					c1a = __float_as_int(A1.a2_child.w);
					c2a = __float_as_int(A2.a2_child.w);
					c1b = cb;
					c2b = cb;
				} else {
					a1 = a; b1 = cb + 1; a2 = a; b2 = cb;
					RobBox B1;
					load_rob_box(load_node_uniform(s.rob.nodes, b1), c, B1);
					ov1 = in_m && !obb_disjoint(A, B1, c.eta);
					c1b = B1.child;
					load_rob_box(load_node_uniform(s.rob.nodes, b2), c, B);
					cached_b = b2;
					ov2 = in_m && !obb_disjoint(A, B, c.eta);
					c2b = B.child;
					c1a = ca;
					c2a = ca;
				}
				const uint64_t mo1 = __ballot(ov1), mo2 = __ballot(ov2);
				/* child 1 first, so that child 2 is on top */
				if (mo1) {
					const int ml = (int)(uint32_t)mo1, mh = (int)(uint32_t)(mo1 >> 32);
					if (c1a < 0 && c1b < 0) {
						lte = wrl(~c1a, nl, lte); ltr = wrl(~c1b, nl, ltr); lml = wrl(ml, nl, lml); lmh = wrl(mh, nl, lmh);
						nl++;
					} else {
						va = wrl(a1, sp, va); vb = wrl(b1, sp, vb); vml = wrl(ml, sp, vml); vmh = wrl(mh, sp, vmh);
						sp++;
					}
				}
				if (mo2) {
					const int ml = (int)(uint32_t)mo2, mh = (int)(uint32_t)(mo2 >> 32);
					if (c2a < 0 && c2b < 0) {
						lte = wrl(~c2a, nl, lte); ltr = wrl(~c2b, nl, ltr); lml = wrl(ml, nl, lml); lmh = wrl(mh, nl, lmh);
						nl++;
					} else {
						va = wrl(a2, sp, va); vb = wrl(b2, sp, vb); vml = wrl(ml, sp, vml); vmh = wrl(mh, sp, vmh);
						sp++;
					}
				}
			}
		}
		if (nl > 0) {
			ExactTf x;
			{
				double q[7];
				pose.get(q);
				make_exact(s, q, x);
			}
			for (int i = 0; i < nl; i++) {
				uint64_t m = ((uint64_t)(uint32_t)rdl(lmh, i) << 32) | (uint64_t)(uint32_t)rdl(lml, i);
				m &= ~hit;
				if (FIRST && hit)
					m &= (hit & (~hit + 1ull)) - 1ull;
				if (m == 0)
					continue;
				const int te = rdl(lte, i), tr = rdl(ltr, i);
				const bool h = ((m >> lane) & 1ull) && leaf_pair_x(s, x, te, tr);
				hit |= __ballot(h);
			}
			if ((hit & live) == live || (FIRST && (hit & live & (~live + 1ull))))
				return hit & live;
		}
		if (sp == 0)
			return hit & live;
	}
}

/* ---- segments: one wavefront per segment ---- */
__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }

/* Per-wave segment record in LDS. */
struct SegShared {
	se3_segment sg;
	double dist, inv_dist, vd;
	int exact; /* every partial sum k*vd of the step sequence is exact */
	int early;
	int bad;   /* not an early return, yet 2^31 or more steps (or none) */
};

/* Wave-level check of one pose per lane (lanes outside `live` are ignored):
 * returns the mask of colliding lanes.  pq: the wavefront's pose rows. */
template <bool FIRST, bool EAGER>
__device__ __forceinline__ uint64_t seg_poses_collide(const DevScene& s, const SegShared* sh, double* pq, double tau,
                                                      uint64_t live)
{
	{
		double q[7];
		se3_segment_at_fast(&sh->sg, tau, q);
		stage_pose(pq, q);
	}
	return EAGER ? packet_collide_eager<FIRST>(s, LdsPose{pq}, live) : packet_collide<FIRST>(s, LdsPose{pq}, live);
}

/*
 * True when the sequential sums vd, vd+vd, ... of transitStateToWithContact's
 * stepping loop are all exact up to dist + vd, so the k-th step sits at
 * exactly k*vd: vd's lowest set bit is 2^e and every sum stays below
 * 2^(e+53).
 */
__device__ __forceinline__ int steps_exact(double vd, double dist)
{
	if (!(vd > 0.0) || !(dist < 1e300))
		return 0;
	const unsigned long long bits = (unsigned long long)__double_as_longlong(vd);
	const int bexp = (int)((bits >> 52) & 0x7ff);
	if (bexp == 0 || bexp == 0x7ff)
		return 0;
	const unsigned long long mant = (bits & 0xfffffffffffffULL) | (1ULL << 52);
	const int e = bexp - 1075 + __ffsll((long long)mant) - 1;
	if (e + 53 > 1000)
		return 0;
	return (dist + 2.0 * vd) < ldexp(1.0, e + 53);
}

/* Exact parallel form of `levels` bisection halvings: each round evaluates
 * the 31 midpoints of a 5-level probe tree (every midpoint computed by the
 * same 0.5*(lo+hi) sequence the sequential search would use), then walks the
 * tree with the ballot of results. */
template <bool WIDE>
__device__ void bisect_wave(const DevScene& s, const SegShared* sh, double* pq, double& lo, double& hi, int levels)
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
				const double m = 0.5 * (l0 + h0);
				if ((h >> bbit) & 1)
					l0 = m;
				else
					h0 = m;
			}
			mid = 0.5 * (l0 + h0);
		}
		const uint64_t live = __ballot(act);
		const uint64_t hitm = seg_poses_collide<false, SE3CQ_TOUCH_EAGER != 0>(s, sh, pq, act ? mid : lo, live);
		ok = !((hitm >> lane) & 1ull);
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

/* ---- touch search: bisection over a swept candidate list ---- */
/*
 * Every probe of a contact bisection lies on the segment between the bracket
 * ends [lo, hi].  One traversal with the robot boxes grown by the largest
 * displacement any robot point undergoes over that interval lists the leaf
 * pairs (env leaf, robot leaf) that can touch at any probe; each round then
 * tests only those pairs: a wave-uniform loop of binary32 leaf-box tests with
 * one probe per lane, and the exact binary64 triangle tests packed 64
 * (probe, pair) combinations per pass.  The outcome per probe is the same
 * "some triangle pair intersects" the full traversal decides.
 */
enum { SWEEP_QUEUE = 256, SWEEP_CANDS = 128 };

/* Per-wavefront LDS of a swept candidate search (3,328 B). */
struct SweepShared {
	uint32_t qa[SWEEP_QUEUE], qb[SWEEP_QUEUE];   /* breadth-first queue of node pairs (ring) */
	uint32_t ca[SWEEP_CANDS], cb[SWEEP_CANDS];   /* candidate leaf node pairs */
	uint32_t combo[64];                          /* (candidate << 8) | probe, one pass */
};

/* Per-wavefront LDS of a touch search: the sweep lists and the exact
 * transforms of one bisection round's probes. */
struct BisectShared {
	double xf[32][12];                           /* exact transform per probe: R row-major, T */
	SweepShared sw;
};
static_assert(sizeof(BisectShared) >= POSE_LDS_BYTES, "a touch search's pose rows live in its BisectShared");

/* Make this wave's LDS stores visible to its own later LDS loads. */
__device__ __forceinline__ void wave_lds_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Persistent wavefronts: the grid holds only as many wavefronts as stay
 * resident, and each takes its next work item from the launch's counter
 * (lane 0 adds, the wave reads the value in scalar registers) until the
 * items run out.  A long item then holds one wavefront, not a whole block,
 * and a finished wavefront starts its next item at once. */
__device__ __forceinline__ int64_t next_item(unsigned long long* queue, int take = 1)
{
	unsigned long long got = 0;
	if ((threadIdx.x & 63) == 0)
		got = atomicAdd(queue, (unsigned long long)take);
	return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(got >> 32)) << 32) |
	                 (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)got));
}

/* Number of set bits of m below this lane. */
__device__ __forceinline__ int lane_rank(uint64_t m)
{
	return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/*
 * Candidate leaf pairs for all poses se3_segment_at_fast(sg, t), t in [t0, t1]:
 * breadth first, one node pair per lane.  Poses are placed at the middle
 * parameter tm; a slerp moves the rotation by at most 2 * theta * |t - tm|
 * radians and the translation by |p1 - p0| * |t - tm|, so growing the robot
 * boxes by that rotation times the robot's reach plus that translation (with
 * a relative 1e-3 to spare) keeps every probe's boxes inside.  Returns the
 * number of candidates, or -1 if the queue or the list would overflow.
 */
__device__ int sweep_candidates(const DevScene& s, const se3_segment* sg, double t0, double t1, SweepShared* bs)
{
	const int lane = threadIdx.x & 63;
	QCull c;
	{
		double q[7];
		se3_segment_at_fast(sg, 0.5 * (t0 + t1), q);
		make_qcull(s, q, c);
	}
	const double half = 0.5 * (t1 - t0);
	/* normalised lerp between quaternions closer than acos(1 - 2^-52) */
	const double th = sg->lerp ? 4.5e-8 : sg->theta;
	double d2 = 0.0;
#pragma unroll
	for (int k = 0; k < 3; k++) {
		const double d = sg->p1[k] - sg->p0[k];
		d2 += d * d;
	}
	const double grow = (2.0 * th * half * (double)s.rob_reach + sqrt(d2) * half) * 1.001 + 1e-12;
	c.infl += (float)grow * 1.0001f;
	int head = 0, tail = 1, nc = 0;
	if (lane == 0) {
		bs->qa[0] = 0u;
		bs->qb[0] = 0u;
	}
	wave_lds_sync();
	while (head < tail) {
		const int nproc = (tail - head < 64) ? (tail - head) : 64;
		const bool act = lane < nproc;
		int a = 0, b = 0;
		if (act) {
			const int slot = (head + lane) & (SWEEP_QUEUE - 1);
			a = (int)bs->qa[slot];
			b = (int)bs->qb[slot];
		}
		head += nproc;
		bool leaf = false, inner = false;
		int a1 = 0, b1 = 0, a2 = 0, b2 = 0;
		if (act) {
			const GNode A = load_node(s.env.nodes, a);
			const GNode Bn = load_node(s.rob.nodes, b);
			bool ov;
			{
				RobBox B;
				load_rob_box(Bn, c, B);
				ov = !obb_disjoint(A, B, c.eta);
			}
			const int ca = __float_as_int(A.a2_child.w);
			const int cb = __float_as_int(Bn.a2_child.w);
			const bool la = ca < 0, lb = cb < 0;
			leaf = ov && la && lb;
			inner = ov && !(la && lb);
			const float sa = fmaf(A.c_ex.w, A.c_ex.w, fmaf(A.a0_ey.w, A.a0_ey.w, A.a1_ez.w * A.a1_ez.w));
			const float sb = fmaf(Bn.c_ex.w, Bn.c_ex.w, fmaf(Bn.a0_ey.w, Bn.a0_ey.w, Bn.a1_ez.w * Bn.a1_ez.w));
			if (!la && (lb || sa > sb)) {
				a1 = ca; b1 = b; a2 = ca + 1; b2 = b;
			} else {
				a1 = a; b1 = cb; a2 = a; b2 = cb + 1;
			}
		}
		const uint64_t ml = __ballot(leaf), mi = __ballot(inner);
		const int nl = __popcll(ml), ni = __popcll(mi);
		if (nc + nl > SWEEP_CANDS || (tail - head) + 2 * ni > SWEEP_QUEUE)
			return -1;
		if (leaf) {
			const int r = nc + lane_rank(ml);
			bs->ca[r] = (uint32_t)a;
			bs->cb[r] = (uint32_t)b;
		}
		if (inner) {
			const int r = tail + 2 * lane_rank(mi);
			bs->qa[r & (SWEEP_QUEUE - 1)] = (uint32_t)a1;
			bs->qb[r & (SWEEP_QUEUE - 1)] = (uint32_t)b1;
			bs->qa[(r + 1) & (SWEEP_QUEUE - 1)] = (uint32_t)a2;
			bs->qb[(r + 1) & (SWEEP_QUEUE - 1)] = (uint32_t)b2;
		}
		nc += nl;
		tail += 2 * ni;
		wave_lds_sync();
	}
	return nc;
}

/* Exact tests of the first nb (candidate, probe) combinations in bs->sw.combo,
 * one per lane; returns the mask of probes found colliding. */
__device__ uint64_t sat_pass(const DevScene& s, const BisectShared* bs, int nb)
{
	const int lane = threadIdx.x & 63;
	bool h = false;
	int p = 0;
	if (lane < nb) {
		const uint32_t cmb = bs->sw.combo[lane];
		const int i = (int)(cmb >> 8);
		p = (int)(cmb & 0xffu);
		ExactTf x;
#pragma unroll
		for (int k = 0; k < 9; k++)
			x.R[k] = bs->xf[p][k];
#pragma unroll
		for (int k = 0; k < 3; k++)
			x.T[k] = bs->xf[p][9 + k];
		const int te = ~__float_as_int(s.env.nodes[bs->sw.ca[i]].a2_child.w);
		const int tr = ~__float_as_int(s.rob.nodes[bs->sw.cb[i]].a2_child.w);
		h = leaf_pair_x(s, x, te, tr);
	}
	uint64_t hb = __ballot(h);
	uint64_t out = 0;
	while (hb) {
		const int j = __ffsll((long long)hb) - 1;
		hb &= hb - 1;
		out |= 1ull << rdl(p, j);
	}
	return out;
}

/*
 * Mask of the lanes in `live` (all below lane 32) whose pose
 * se3_segment_at_fast(sg, tau, .) collides, decided on the nc candidate pairs in
 * bs, which must cover every such pose.  Each pass tests two candidates
 * against all probes: lane l takes probe l % 32 (the pose of lane l % 32)
 * and candidate i + l / 32, so no lane idles while a round has at most 32
 * probes.
 */
__device__ uint64_t cands_collide(const DevScene& s, const se3_segment* sg, BisectShared* bs, int nc, double tau,
                                  uint64_t live)
{
	const int lane = threadIdx.x & 63;
	const int p = lane & 31, h = lane >> 5;
	const bool act = (live >> p) & 1ull;
	QCull c;
	{
		double q[7];
		se3_segment_at_fast(sg, shfl_d(tau, p), q);
		make_qcull(s, q, c);
		if (h == 0 && act) {
			ExactTf x;
			make_exact(s, q, x);
#pragma unroll
			for (int t = 0; t < 9; t++)
				bs->xf[lane][t] = x.R[t];
#pragma unroll
			for (int t = 0; t < 3; t++)
				bs->xf[lane][9 + t] = x.T[t];
		}
	}
	wave_lds_sync();
	uint64_t hitm = 0;
	int nb = 0;
	for (int i = 0; i < nc; i += 2) {
		const int ci = i + h;
		bool ov = false;
		if (act && ci < nc && !((hitm >> p) & 1ull)) {
			const GNode A = load_node(s.env.nodes, (int)bs->sw.ca[ci]);
			RobBox B;
			load_rob_box(load_node(s.rob.nodes, (int)bs->sw.cb[ci]), c, B);
			ov = !obb_disjoint(A, B, c.eta);
		}
		const uint64_t m = __ballot(ov);
		const int cnt = __popcll(m);
		if (cnt == 0)
			continue;
		if (nb + cnt > 64) {
			hitm |= sat_pass(s, bs, nb);
			nb = 0;
			wave_lds_sync();
		}
		if (ov)
			bs->sw.combo[nb + lane_rank(m)] = ((uint32_t)ci << 8) | (uint32_t)p;
		nb += cnt;
		wave_lds_sync();
	}
	if (nb > 0)
		hitm |= sat_pass(s, bs, nb);
	wave_lds_sync();
	return hitm & live;
}

/* `levels` exact bisection halvings (the probe-tree rounds of bisect_wave)
 * evaluated on the nc candidate pairs in bs. */
__device__ void bisect_cands(const DevScene& s, const SegShared* sh, double& lo, double& hi, int levels,
                             BisectShared* bs, int nc)
{
	const int lane = threadIdx.x & 63;
	for (int done = 0; done < levels;) {
		const int k = (levels - done < 5) ? (levels - done) : 5;
		const int nodes = (1 << k) - 1;
		const bool act = lane < nodes;
		double mid = 0.0;
		if (act) {
			const int h = lane + 1;
			const int dh = 31 - __clz(h);
			double l0 = lo, h0 = hi;
			for (int bbit = dh - 1; bbit >= 0; bbit--) {
				const double m = 0.5 * (l0 + h0);
				if ((h >> bbit) & 1)
					l0 = m;
				else
					h0 = m;
			}
			mid = 0.5 * (l0 + h0);
		}
		const uint64_t hitm = cands_collide(s, &sh->sg, bs, nc, act ? mid : lo, __ballot(act));
		const bool ok = !((hitm >> lane) & 1ull);
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

/* Contact refinement of a touch search: candidate-list bisection, or the
 * packet traversal per probe round when the candidate list overflows. */
template <bool WIDE>
__device__ void bisect_touch(const DevScene& s, const SegShared* sh, double& lo, double& hi, int levels,
                             BisectShared* bs)
{
	if (levels <= 0)
		return;
	const int nc = sweep_candidates(s, &sh->sg, lo, hi, &bs->sw);
	if (nc >= 0)
		bisect_cands(s, sh, lo, hi, levels, bs, nc);
	else
		bisect_wave<WIDE>(s, sh, reinterpret_cast<double*>(bs), lo, hi, levels);
}

/* ---- walk chunks: the same candidate search over 64 consecutive steps ---- */
/*
 * A walk chunk tests the 64 poses se3_segment_at_fast(sg, tau_l), one per
 * lane, whose parameters increase with the lane.  sweep_candidates over
 * [tau of the first live lane, tau of the last] lists, 64 node pairs per
 * breadth-first pass, every leaf pair whose boxes can overlap at any of
 * them; each candidate is then tested against all 64 poses at once.  The
 * packet traversal instead tests every inner node pair pose by pose, one
 * pair per pass.  When the lists overflow the chunk falls back to the
 * packet traversal.  Measured on MI355X: the visibility matrix of 2,048
 * shell milestones runs 4 % faster this way; the touch search and the
 * uneven shell segments run 6 % and 16 % slower (their chunks sit closer to
 * contact, with longer lists), so only the visibility kernel sweeps.
 */
#ifndef SE3CQ_WALK_SWEEP
#define SE3CQ_WALK_SWEEP 1
#endif

/* Exact tests of the first nb (candidate, probe) combinations of sw->combo,
 * one per lane; the probe is a lane of the chunk, whose exact transform is
 * rebuilt from its pose (the same se3_segment_at_fast and make_exact the
 * packet traversal evaluates).  Returns the mask of probes found colliding. */
__device__ uint64_t walk_sat_pass(const DevScene& s, const se3_segment* sg, const SweepShared* sw, int nb, double tau)
{
	const int lane = threadIdx.x & 63;
	const uint32_t cmb = lane < nb ? sw->combo[lane] : 0u;
	const int p = (int)(cmb & 0xffu);
	const double tp = shfl_d(tau, p);
	bool h = false;
	if (lane < nb) {
		const int i = (int)(cmb >> 8);
		double q[7];
		se3_segment_at_fast(sg, tp, q);
		ExactTf x;
		make_exact(s, q, x);
		const int te = ~__float_as_int(s.env.nodes[sw->ca[i]].a2_child.w);
		const int tr = ~__float_as_int(s.rob.nodes[sw->cb[i]].a2_child.w);
		h = leaf_pair_x(s, x, te, tr);
	}
	uint64_t hb = __ballot(h);
	uint64_t out = 0;
	while (hb) {
		const int j = __ffsll((long long)hb) - 1;
		hb &= hb - 1;
		out |= 1ull << rdl(p, j);
	}
	return out;
}

/* Colliding lanes of `live` among the chunk's poses, decided on the nc
 * candidates in sw (which cover every live pose).  Only the lowest
 * colliding lane is exact: lanes above a hit found so far stop testing. */
__device__ uint64_t walk_cands_collide(const DevScene& s, const se3_segment* sg, SweepShared* sw, int nc, double tau,
                                       uint64_t live)
{
	const int lane = threadIdx.x & 63;
	QCull c;
	{
		double q[7];
		se3_segment_at_fast(sg, tau, q);
		make_qcull(s, q, c);
	}
	uint64_t hitm = 0;
	int nb = 0;
	for (int i = 0; i < nc; i++) {
		const uint64_t below = hitm ? (hitm & (~hitm + 1ull)) - 1ull : ~0ull;
		const uint64_t want = live & below;
		const int ca = __builtin_amdgcn_readfirstlane((int)sw->ca[i]);
		const int cb = __builtin_amdgcn_readfirstlane((int)sw->cb[i]);
		bool ov = false;
		if ((want >> lane) & 1ull) {
			const GNode A = load_node_uniform(s.env.nodes, ca);
			RobBox B;
			load_rob_box(load_node_uniform(s.rob.nodes, cb), c, B);
			ov = !obb_disjoint(A, B, c.eta);
		}
		const uint64_t m = __ballot(ov);
		const int cnt = __popcll(m);
		if (cnt == 0)
			continue;
		if (nb + cnt > 64) {
			hitm |= walk_sat_pass(s, sg, sw, nb, tau);
			nb = 0;
			wave_lds_sync();
		}
		if (ov)
			sw->combo[nb + lane_rank(m)] = ((uint32_t)i << 8) | (uint32_t)lane;
		nb += cnt;
		wave_lds_sync();
	}
	if (nb > 0)
		hitm |= walk_sat_pass(s, sg, sw, nb, tau);
	wave_lds_sync();
	return hitm & live;
}

/* Colliding lanes of a walk chunk (lowest one exact): with SWEEP by
 * candidate lists when they fit, else by the packet traversal. */
template <bool SWEEP, bool EAGER>
__device__ __forceinline__ uint64_t walk_collide(const DevScene& s, const SegShared* sh, SweepShared* sw, double* pq,
                                                 double tau, uint64_t live)
{
	if (SWEEP) {
		const int l0 = __ffsll((long long)live) - 1;
		const int l1 = 63 - __clzll((long long)live);
		const double t0 = shfl_d(tau, l0), t1 = shfl_d(tau, l1);
		const int nc = sweep_candidates(s, &sh->sg, t0, t1, sw);
		if (nc >= 0)
			return walk_cands_collide(s, &sh->sg, sw, nc, tau, live);
	}
	return seg_poses_collide<true, EAGER>(s, sh, pq, tau, live);
}

/* Wave-uniform outcome of one segment walk. */
struct SegResult {
	int done;        /* 1 = collision free (reference tuple element 2) */
	int early;       /* verify_delta >= distance: nothing evaluated */
	int lo_is_from;  /* last free state is `from` itself (no free step) */
	int32_t steps;   /* isValid calls under the reference's semantics (+ probes) */
	double last_tau, tau, lo, hi;
};

/* Fill the wave's segment record.  Lane 0 alone computes it: the correctly
 * rounded atan2, acos and sin it calls are out-of-line functions whose
 * register saves go to scratch memory, so the other lanes sit out (their
 * saves would be 63 times the traffic for the same record). */
template <int MODE>
__device__ __forceinline__ void seg_record(SegShared* sh, const double* __restrict__ from,
                                           const double* __restrict__ to, double vd_in)
{
	double f[7], t[7];
#pragma unroll
	for (int k = 0; k < 7; k++) {
		f[k] = from[k];
		t[k] = to[k];
	}
	const double dist = se3_distance(f, t);
	double vd = vd_in;
	if (MODE == MODE_TRANSITION) {
		const double half = dist / 2.0;
		vd = (vd_in < half) ? vd_in : half; /* std::min(d / 2, initial_verify_delta) */
	}
	se3_segment sg;
	se3_segment_init(&sg, f, t);
	{
		sh->sg = sg;
		sh->dist = dist;
		sh->inv_dist = 1.0 / dist;
		sh->vd = vd;
		sh->early = vd >= dist;
		sh->exact = steps_exact(vd, dist);
		/* a NaN distance never enters the stepping loop (the reference
		 * returns "free"); an infinite one, a step that is not positive,
		 * or 2^31 or more steps would never end */
		sh->bad = (dist == dist) && !(vd >= dist) && !(dist / vd < 2147483648.0 && vd > 0.0);
	}
}

/* Lane 0 computes the wavefront's segment record. */
template <int MODE>
__device__ __forceinline__ void seg_setup(SegShared* sh, const double* __restrict__ from,
                                          const double* __restrict__ to, double vd_in)
{
	if ((threadIdx.x & 63) == 0)
		seg_record<MODE>(sh, from, to, vd_in);
}

/*
 * Every segment's record before the walk, one segment per lane.  In the
 * walk a record is one dependent chain of correctly rounded functions on a
 * single lane (about 54 us per segment on a wavefront of the touch search:
 * 0.86 ms of wavefront time for 16 segments each); here 64 segments share
 * each instruction.  The walk copies its record into LDS.  (Its register
 * bound is the segment kernels': the out-of-line functions it shares with
 * them are compiled once for all their callers.)
 */
template <int MODE>
__global__ void __launch_bounds__(256, SE3CQ_SEG_WAVES) k_seg_records(const double* __restrict__ from, const double* __restrict__ to,
                                                     int64_t n, double vd, SegShared* __restrict__ rec)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		seg_record<MODE>(rec + i, from + 7 * i, to + 7 * i, vd);
}
enum { SEG_REC_WORDS = sizeof(SegShared) / 8 };
static_assert(sizeof(SegShared) % 8 == 0 && SEG_REC_WORDS <= 64, "a record is copied one word per lane");

template <bool WIDE, int MODE, bool SWEEP = false, bool EAGER = (MODE == MODE_TOUCH && SE3CQ_TOUCH_EAGER)>
__device__ SegResult segment_walk(const DevScene& s, const SegShared* sh, double* pq, int levels, BisectShared* bs,
                                  SweepShared* sw = nullptr)
{
	const int lane = threadIdx.x & 63;
	SegResult r;
	r.early = 0;
	r.lo_is_from = 0;
	if (sh->bad) {
		/* the walk would not end: flag it and report a collision at the
		 * start, which the host turns into an error */
		if (lane == 0)
			atomicOr(s.error_flag, 2u);
		r.done = 0;
		r.lo_is_from = 1;
		r.steps = 0;
		r.last_tau = 0.0;
		r.tau = 0.0;
		r.lo = 0.0;
		r.hi = 0.0;
		return r;
	}
	if (sh->early) {
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
	const double dist = sh->dist, inv_dist = sh->inv_dist, vd = sh->vd;
	const int exact = sh->exact;
	double dbase = vd;
	double prev_tau = 0.0;
	int64_t base = 0;
	for (;;) {
		double d;
		bool active;
		if (exact) {
			/* k-th step of the chunk at exactly dbase + k*vd, clamped to dist */
			d = dbase + (double)lane * vd;
			const double dprev = dbase + (double)(lane - 1) * vd;
			active = (lane == 0) ? (dbase <= dist) : (dprev < dist);
			d = (d < dist) ? d : dist;
		} else {
			d = dbase;
			for (int k = 0; k < lane; k++) {
				if (d < dist) {
					d += vd;
					d = (d < dist) ? d : dist;
				} else {
					d += vd;
				}
			}
			active = d <= dist;
		}
		const double tau = d * inv_dist;
		const uint64_t am = __ballot(active);
		const uint64_t cm = am ? walk_collide<SWEEP, EAGER>(s, sh, sw, pq, active ? tau : sh->vd * inv_dist, am) : 0ull;
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
				bisect_touch<WIDE>(s, sh, r.lo, r.hi, levels, bs);
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

/* One segment's outputs: lane 0 stores the scalars, lanes 0-6 the two
 * returned states, rebuilt from the segment record (the interpolation is
 * deterministic, so they equal the states the walk evaluated). */
__device__ __forceinline__ void seg_store(const SegOut& o, int64_t seg, const SegResult& r, const SegShared* sh,
                                          const double* from, const double* to)
{
	const int lane = threadIdx.x & 63;
	if (lane == 0) {
		o.done[seg] = (uint8_t)r.done;
		if (o.last_tau) o.last_tau[seg] = r.last_tau;
		if (o.tau) o.tau[seg] = r.tau;
		if (o.lo) o.lo[seg] = r.lo;
		if (o.hi) o.hi[seg] = r.hi;
		if (o.steps) o.steps[seg] = r.steps;
	}
	/* lane 0 alone again: se3_segment_at calls the out-of-line correctly
	 * rounded sin and acos */
	if (o.last_free && lane == 0) {
		double a[7], b[7];
		if (r.lo_is_from) {
			for (int k = 0; k < 7; k++) a[k] = from[k];
		} else {
			se3_segment_at(&sh->sg, r.last_tau, a);
		}
		if (r.early) {
			for (int k = 0; k < 7; k++) b[k] = to[k];
		} else if (r.done) {
			for (int k = 0; k < 7; k++) b[k] = a[k];
		} else {
			se3_segment_at(&sh->sg, r.tau, b);
		}
		for (int k = 0; k < 7; k++) {
			o.last_free[7 * seg + k] = a[k];
			o.state[7 * seg + k] = b[k];
		}
	}
}

template <bool WIDE, int MODE>
__global__ void __launch_bounds__(256, MODE == MODE_TOUCH ? SE3CQ_TOUCH_WAVES : SE3CQ_SEG_WAVES) k_segments(DevScene s, const double* __restrict__ from_all,
                                                  const double* __restrict__ to_all, int64_t n, double vd,
                                                  int levels, SegOut o)
{
	extern __shared__ double seg_lds[];
	__shared__ SegShared segsh[4];
	const int wave = threadIdx.x >> 6;
	SegShared* sh = &segsh[wave];
	BisectShared* bs = (MODE == MODE_TOUCH) ? reinterpret_cast<BisectShared*>(seg_lds) + wave : nullptr;
	/* the pose rows: a touch search's walk and its candidate bisection take
	 * turns on the wavefront's BisectShared */
	double* pq = (MODE == MODE_TOUCH) ? reinterpret_cast<double*>(bs) : seg_lds + wave * POSE_LDS_DOUBLES;
	int64_t items = n;
	if (s.n_items) {
		const int64_t c = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)*s.n_items);
		items = c < n ? c : n;
	}
	/* items [got, got + left) are this wavefront's (queue_depth per add) */
	const int depth = s.queue_depth > 0 ? s.queue_depth : 1;
	int64_t got = 0;
	int left = 0;
	for (;;) {
		if (left == 0) {
			got = next_item(s.queue, depth);
			left = depth;
		}
		const int64_t item = got++;
		left--;
		if (item >= items)
			break;
		const int64_t seg = s.order ? (int64_t)s.order[item] : item;
		if (seg < 0 || seg >= n)
			continue; /* an index list from the caller may hold anything */
		if (s.seg_rec) {
			const int lane = threadIdx.x & 63;
			if (lane < SEG_REC_WORDS)
				reinterpret_cast<uint64_t*>(sh)[lane] =
				    reinterpret_cast<const uint64_t*>(static_cast<const SegShared*>(s.seg_rec) + seg)[lane];
		} else {
			seg_setup<MODE>(sh, from_all + 7 * seg, to_all + 7 * seg, vd);
		}
		wave_lds_sync();
		const SegResult r = segment_walk<WIDE, MODE>(s, sh, pq, levels, bs);
		seg_store(o, seg, r, sh, from_all + 7 * seg, to_all + 7 * seg);
		wave_lds_sync();
	}
}

/* ---- mixed batches: persistent wavefronts over a work queue ---- */
/*
 * One launch answers a batch of uneven segments (transitStateToWithContact,
 * lib/osr/unit_world.cc:405-474, at a fixed verify step) and of single
 * states (isValid, 215-247).  The grid holds only as many wavefronts as the
 * device keeps resident; each takes work items from a 64-bit counter
 * (zeroed by a fill kernel on the same stream, so the pair of launches can
 * be captured in a HIP graph) until the queue is empty.  Items
 * [0, n_seg) are segments, one per wavefront, walked as k_segments walks
 * them; the rest are chunks of 64 * states_per_lane states, one per lane
 * with the per-thread traversal of k_valid_states.  A long segment occupies
 * one wavefront while the others drain the queue, so uneven lengths do not
 * idle whole blocks.  Only a batch with states (STATES) reserves the
 * per-thread traversal stacks in LDS, which limit how many blocks a CU
 * holds; a batch of segments alone runs at full occupancy.
 */
template <bool WIDE, bool STATES>
__global__ void __launch_bounds__(256, SE3CQ_MIN_WAVES) k_batch(DevScene s, BatchArgs a)
{
	extern __shared__ uint32_t smem[];
	__shared__ SegShared segsh[4];
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;
	SegShared* sh = &segsh[wave];
	PairStack<WIDE> stk, leaves;
	if (STATES)
		thread_stacks<WIDE>(s, smem, stk, leaves);
	/* the pose rows follow the traversal stacks */
	double* pq = reinterpret_cast<double*>(smem + (STATES ? (size_t)lds_words_per_thread(s) * blockDim.x : 0)) +
	             wave * POSE_LDS_DOUBLES;
	const int64_t per_item = 64 * (int64_t)a.states_per_lane;
	const int64_t total = a.n_seg + (a.n_states + per_item - 1) / per_item;
	const int depth = s.queue_depth > 0 ? s.queue_depth : 1;
	int64_t got = 0;
	int left = 0;
	for (;;) {
		if (left == 0) {
			got = next_item(s.queue, depth);
			left = depth;
		}
		const int64_t item = got++;
		left--;
		if (item >= total)
			break;
		if (item < a.n_seg) {
			const double* from = a.seg_from + 7 * item;
			const double* to = a.seg_to + 7 * item;
			seg_setup<MODE_TRANSIT>(sh, from, to, a.vd);
			wave_lds_sync();
			const SegResult r = segment_walk<WIDE, MODE_TRANSIT>(s, sh, pq, 0, nullptr);
			seg_store(a.seg, item, r, sh, from, to);
			/* every lane is done with the record before the next item
			 * overwrites it */
			wave_lds_sync();
		} else if (STATES) {
			const int64_t base = (item - a.n_seg) * per_item;
			for (int j = 0; j < a.states_per_lane; j++) {
				const int64_t i = base + 64 * j + lane;
				if (i < a.n_states)
					a.valid[i] = collide_dfs<WIDE>(s, StateFromGlobal{a.states + 7 * i}, stk, leaves) ? 0 : 1;
			}
		}
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
/* A visibility wavefront's LDS: its pose rows or its swept-search lists. */
enum { VIS_WAVE_DOUBLES = (sizeof(SweepShared) > POSE_LDS_BYTES ? sizeof(SweepShared) : POSE_LDS_BYTES) / 8 };
static_assert(sizeof(SweepShared) % 8 == 0, "the regions are whole binary64 values");

template <bool WIDE, int VIS>
__global__ void __launch_bounds__(256, SE3CQ_VIS_WAVES) k_visibility(DevScene s, const double* __restrict__ qs0,
                                                    const double* __restrict__ qs1, int64_t m, int64_t n,
                                                    int64_t npairs, double ivd, int8_t* __restrict__ out8,
                                                    uint32_t* __restrict__ bits, int32_t* __restrict__ steps)
{
	extern __shared__ double vis_lds[];
	__shared__ SegShared segsh[4];
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;
	SegShared* sh = &segsh[wave];
	/* the pose rows of the packet traversal and the candidate lists of the
	 * swept search take turns on the wavefront's region */
	double* pq = vis_lds + wave * VIS_WAVE_DOUBLES;
	SweepShared* sw = reinterpret_cast<SweepShared*>(pq);
	for (;;) {
		const int64_t p = next_item(s.queue);
		if (p >= npairs)
			break;
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
		seg_setup<MODE_TRANSITION>(sh, qs0 + 7 * fi, qs1 + 7 * ti, ivd);
		wave_lds_sync();
		const SegResult r = segment_walk<WIDE, MODE_TRANSITION, SE3CQ_WALK_SWEEP != 0, true>(s, sh, pq, 0, nullptr, sw);
		const uint8_t done = (uint8_t)r.done;
		if (lane == 0) {
			if (VIS == VIS_MATRIX) {
				if (out8) {
					out8[fi * n + ti] = (int8_t)done;
					out8[ti * n + fi] = (int8_t)done;
				}
				if (bits && done) {
					const int64_t b1 = fi * n + ti, b2 = ti * n + fi;
					atomicOr(bits + (b1 >> 5), 1u << (b1 & 31));
					atomicOr(bits + (b2 >> 5), 1u << (b2 & 31));
				}
			} else if (VIS == VIS_MATRIX2) {
				if (out8) out8[fi * n + ti] = (int8_t)done;
				if (bits && done) {
					const int64_t b1 = fi * n + ti;
					atomicOr(bits + (b1 >> 5), 1u << (b1 & 31));
				}
			} else {
				if (out8) out8[p] = (int8_t)done;
			}
			if (steps) steps[p] = r.steps;
		}
		wave_lds_sync();
	}
}

/* ---- OMPL discrete motions: one wavefront per motion ---- */
/*
 * OmplDriver::validateMotionPairs (lib/pyse3ompl/ompldriver.cc:592-620) runs
 * DiscreteMotionValidator::checkMotion on each pair: the motion is valid when
 * the end state and every interior sample interpolate(k / nd), 0 < k < nd,
 * are valid.  The answer does not depend on the order the samples are
 * tested in.  The wave first tests 64 samples spread over the whole motion
 * (1, 1 + C, 1 + 2C, ... with C = ceil(K / 64)), the way the validator's
 * midpoint-first order starts, so most colliding motions end after one
 * packet traversal; then it tests all K samples 64 neighbours at a time,
 * whose nearly equal poses keep the packet traversal narrow.  (Measured on
 * the alpha puzzle's key screening: 2.5x faster than a strided order for
 * every pass, 1.1x faster than neighbours only.)
 * out: 1 valid, 0 invalid, -1 the sample count is not a finite number below
 * 2^31.  samples (optional): max(nd, 1), the states a valid motion costs.
 */
template <bool WIDE>
__global__ void __launch_bounds__(256, SE3CQ_OMPL_WAVES) k_ompl_motions(DevScene s, OmplXf x,
                                                      const double* __restrict__ q0,
                                                      const double* __restrict__ q1, int64_t n, double lvs,
                                                      int32_t* __restrict__ out, int32_t* __restrict__ samples)
{
	/* the motion record and the transform sit in LDS (broadcast reads), so
	 * the packet traversal keeps its registers; so do the pose rows */
	__shared__ se3_ompl_motion msh[4];
	__shared__ OmplXf xsh;
	extern __shared__ double ompl_lds[];
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;
	if (threadIdx.x < 16)
		xsh.M[threadIdx.x] = x.M[threadIdx.x];
	else if (threadIdx.x < 23)
		xsh.pert[threadIdx.x - 16] = x.pert[threadIdx.x - 16];
	else if (threadIdx.x < 26)
		xsh.delta[threadIdx.x - 23] = x.delta[threadIdx.x - 23];
	__syncthreads();
	se3_ompl_motion* m = &msh[wave];
	double* pq = ompl_lds + wave * POSE_LDS_DOUBLES;
	/* sample k of the motion, in the unit world, staged as the pose rows */
	auto stage = [&](int64_t k) {
		double st[7], o[7];
		se3_ompl_motion_sample(m, k, st, 1);
		se3_ompl_to_unit(xsh.M, xsh.pert, xsh.delta, st, o);
		stage_pose(pq, o);
	};
	for (;;) {
		const int64_t p = next_item(s.queue);
		if (p >= n)
			break;
		if (lane == 0) {
			double qa[7], qb[7];
			for (int k = 0; k < 7; k++) {
				qa[k] = q0[7 * p + k];
				qb[k] = q1[7 * p + k];
			}
			se3_ompl_motion_init(m, qa, qb, lvs);
		}
		wave_lds_sync();
		const int64_t nd = m->nd;
		int valid = 1;
		int64_t K = 0;
		if (nd < 0) {
			valid = -1;
		} else {
			K = nd > 1 ? nd : 1;
			const int64_t C = (K + 63) / 64;
			if (K > 64) {
				/* coarse pass: samples 1, 1 + C, 1 + 2C, ... span the motion */
				const int64_t k = 1 + C * lane;
				const bool active = k <= K;
				const uint64_t am = __ballot(active);
				stage(active ? k : K);
				if (packet_collide_eager<true>(s, LdsPose{pq}, am))
					valid = 0;
			}
			/* then every sample, 64 neighbours per pass */
			for (int64_t c = 0; valid == 1 && c < C; c++) {
				const int64_t k = 1 + 64 * c + lane;
				const bool active = k <= K;
				const uint64_t am = __ballot(active);
				stage(active ? k : K);
				if (packet_collide_eager<true>(s, LdsPose{pq}, am))
					valid = 0;
			}
		}
		if (lane == 0) {
			out[p] = valid;
			if (samples)
				samples[p] = (int32_t)K;
		}
		wave_lds_sync();
	}
}

/* ---- contact enumeration: one thread per state ---- */
/*
 * Every (env triangle, robot triangle) pair whose exact test intersects at
 * the state (CDModel::collideForDetails, lib/osr/cdmodel.cc:201-229, asks FCL
 * for all contacts), found by the depth-first traversal of k_valid_states
 * without its early exit.  The first `cap` pairs found are written; count
 * holds how many there are.
 */
template <bool WIDE>
__global__ void __launch_bounds__(256, SE3CQ_MIN_WAVES) k_contacts(DevScene s, const double* __restrict__ qs, int64_t n,
                                                      int cap, int32_t* __restrict__ pairs,
                                                      int32_t* __restrict__ count)
{
	extern __shared__ uint32_t smem[];
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const double* qp = qs + 7 * i;
	PairStack<WIDE> stk, leaves;
	thread_stacks<WIDE>(s, smem, stk, leaves);
	int found = 0;
	int sp = 0;
	stk.put(sp++, 0, 0);
	const int depth_cap = s.stack_depth;
	while (sp > 0) {
		int nleaf = 0;
		{
			QCull c;
			{
				double q[7];
				for (int k = 0; k < 7; k++)
					q[k] = qp[k];
				make_qcull(s, q, c);
			}
			while (sp > 0 && nleaf < LEAF_BUF) {
				int a, b;
				stk.get(--sp, a, b);
				const GNode A = load_node(s.env.nodes, a);
				const GNode Bn = load_node(s.rob.nodes, b);
				RobBox B;
				load_rob_box(Bn, c, B);
				if (obb_disjoint(A, B, c.eta))
					continue;
				const int ca = __float_as_int(A.a2_child.w);
				const bool la = ca < 0, lb = B.child < 0;
				if (la && lb) {
					leaves.put(nleaf++, a, b);
					continue;
				}
				if (sp + 2 > depth_cap) {
					atomicOr(s.error_flag, 1u);
					count[i] = -1;
					return;
				}
				const float sa = fmaf(A.c_ex.w, A.c_ex.w, fmaf(A.a0_ey.w, A.a0_ey.w, A.a1_ez.w * A.a1_ez.w));
				const float sb = fmaf(Bn.c_ex.w, Bn.c_ex.w, fmaf(Bn.a0_ey.w, Bn.a0_ey.w, Bn.a1_ez.w * Bn.a1_ez.w));
				if (!la && (lb || sa > sb)) {
					stk.put(sp++, ca + 1, b);
					stk.put(sp++, ca, b);
				} else {
					stk.put(sp++, a, B.child + 1);
					stk.put(sp++, a, B.child);
				}
			}
		}
		for (int j = 0; j < nleaf; j++) {
			int a, b;
			leaves.get(j, a, b);
			const int le = leaf_code(s.env.nodes, a), lr = leaf_code(s.rob.nodes, b);
			double q[7];
			for (int k = 0; k < 7; k++)
				q[k] = qp[k];
			ExactTf x;
			make_exact(s, q, x);
			for (int u = 0; u < (le & 7) + 1; u++)
				for (int v = 0; v < (lr & 7) + 1; v++) {
					const int te = (le >> 3) + u, tr = (lr >> 3) + v;
					if (tri_pair_x(s, x, te, tr)) {
						if (found < cap) {
							pairs[(i * cap + found) * 2] = s.env.tri_id[te];
							pairs[(i * cap + found) * 2 + 1] = s.rob.tri_id[tr];
						}
						found++;
					}
				}
		}
	}
	count[i] = found;
}

/*
 * Intersection segment of each listed pair, in world coordinates as
 * UnitWorld::intersectingSegments computes it (unit_world.cc:959-1012): the
 * environment triangle moved by the perturbation transform and the robot
 * triangle by the state's transform.  One thread per (state, slot).
 */
__global__ void k_contact_segments(DevScene s, const double* __restrict__ qs, int64_t n, int cap,
                                   const int32_t* __restrict__ pairs, const int32_t* __restrict__ count,
                                   double* __restrict__ segs, int8_t* __restrict__ flags)
{
	const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n * (int64_t)cap)
		return;
	const int64_t i = t / cap;
	const int slot = (int)(t % cap);
	if (slot >= count[i])
		return;
	const double* q = qs + 7 * i;
	double Rr[9];
	se3_quat_to_rotmat(q[3], q[4], q[5], q[6], Rr);
	const double* E = s.env.tris + 9 * (int64_t)s.env.tri_pos[pairs[2 * t]];
	const double* Q = s.rob.tris + 9 * (int64_t)s.rob.tri_pos[pairs[2 * t + 1]];
	double V[9], U[9];
	for (int v = 0; v < 3; v++) {
		se3_place(s.pert->Re, s.pert->te, E + 3 * v, V + 3 * v);
		se3_place(Rr, q, Q + 3 * v, U + 3 * v);
	}
	int cop = 0;
	double p1[3], p2[3];
	const int hit = se3_tri_isect_segment(V, V + 3, V + 6, U, U + 3, U + 6, &cop, p1, p2);
	for (int k = 0; k < 3; k++) {
		segs[6 * t + k] = p1[k];
		segs[6 * t + 3 + k] = p2[k];
	}
	flags[t] = (int8_t)(hit | (cop << 1));
}

/* ---- isDisentangled: one thread per state ---- */
/* The robot's and the environment's AABB boxes, placed by their model
 * transforms (the robot's from the raw state, the environment's from the
 * perturbation), tested with FCL's box-box test (UnitWorld::isDisentangled,
 * lib/osr/unit_world.cc:249-264 -> CDModel::collideBB, cdmodel.cc:180-198).
 * out = 1 when the boxes are disjoint. */
__global__ void k_disentangled(BoxPair bp, const PertData* __restrict__ pert, const double* __restrict__ qs,
                               int64_t n, uint8_t* __restrict__ out)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const double* q = qs + 7 * i;
	double Rr[9], pe[3], pr[3];
	se3_quat_to_rotmat(q[3], q[4], q[5], q[6], Rr);
	const double tr[3] = {q[0], q[1], q[2]};
	se3_place(pert->Re, pert->te, bp.ce, pe);
	se3_place(Rr, tr, bp.cr, pr);
	out[i] = (uint8_t)se3_boxes_disjoint(Rr, pr, bp.hr, pert->Re, pe, bp.he);
}

/* ---- launch wrappers (host) ---- */
static inline size_t stack_bytes_per_thread(const DevScene& s)
{
	return (size_t)lds_words_per_thread(s) * sizeof(uint32_t);
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

/* ---- longest-first order of a segment launch ----
 *
 * Persistent wavefronts take segments in item order, so a long segment taken
 * late in the launch walks alone while the other wavefronts idle.  A
 * segment launch can instead walk its segments longest first: a counting
 * sort on the estimated step count (distance / verify step), 16 bins per
 * octave, builds order[] from one pass over the inputs.  Segments within a
 * bin come in no fixed order.  The order changes which wavefront walks a
 * segment and when, never what it computes. */
constexpr int ORDER_BINS = 256;

/* Bin of an estimated step count: 0 for the longest (2^16 steps and more),
 * 255 for fewer than one step (the early return) or a NaN distance. */
__device__ __forceinline__ int order_bin(double steps)
{
	if (!(steps >= 1.0))
		return ORDER_BINS - 1;
	if (steps >= 65536.0)
		return 0;
	const uint32_t b = __float_as_uint((float)steps);
	const int octave = (int)(b >> 23) - 127;  /* 0 .. 15 */
	const int sixteenth = (int)((b >> 19) & 15u);
	return ORDER_BINS - 1 - (octave * 16 + sixteenth);
}

/* The SE(3) distance of lib/osr/osr_state.cc:130-154 to a few digits, for
 * the order's bins only: translation length plus twice the relative
 * rotation's half-angle, with the binary32 atan2 of the device library. */
__device__ __forceinline__ double order_distance(const double* a, const double* b)
{
	const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
	const double na = sqrt(a[3] * a[3] + a[4] * a[4] + a[5] * a[5] + a[6] * a[6]);
	const double nb = sqrt(b[3] * b[3] + b[4] * b[4] + b[5] * b[5] + b[6] * b[6]);
	const double dot = fabs(a[3] * b[3] + a[4] * b[4] + a[5] * b[5] + a[6] * b[6]) / (na * nb);
	const double vn = sqrt(fmax(0.0, 1.0 - dot * dot));
	return sqrt(dx * dx + dy * dy + dz * dz) + 2.0 * (double)atan2f((float)vn, (float)dot);
}

template <int MODE>
__global__ void __launch_bounds__(256) k_order_keys(const double* __restrict__ from, const double* __restrict__ to,
                                                    int64_t n, double vd, uint8_t* __restrict__ key,
                                                    uint32_t* __restrict__ hist)
{
	__shared__ uint32_t lh[ORDER_BINS];
	lh[threadIdx.x] = 0;
	__syncthreads();
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) {
		const double d = order_distance(from + 7 * i, to + 7 * i);
		double step = vd;
		if (MODE == MODE_TRANSITION)
			step = (vd < d / 2.0) ? vd : d / 2.0;
		const int b = order_bin(d / step);
		key[i] = (uint8_t)b;
		atomicAdd(&lh[b], 1u);
	}
	__syncthreads();
	if (lh[threadIdx.x])
		atomicAdd(&hist[threadIdx.x], lh[threadIdx.x]);
}

/* Exclusive prefix sum of the 256 bin counts into the bins' first slots. */
__global__ void __launch_bounds__(256) k_order_scan(const uint32_t* __restrict__ hist, uint32_t* __restrict__ cursor)
{
	__shared__ uint32_t v[ORDER_BINS];
	const int t = threadIdx.x;
	v[t] = hist[t];
	__syncthreads();
	for (int off = 1; off < ORDER_BINS; off <<= 1) {
		const uint32_t add = t >= off ? v[t - off] : 0u;
		__syncthreads();
		v[t] += add;
		__syncthreads();
	}
	cursor[t] = v[t] - hist[t];
}

/* Each block reserves one run of slots per bin, then its threads fill it. */
__global__ void __launch_bounds__(256) k_order_scatter(const uint8_t* __restrict__ key, int64_t n,
                                                       uint32_t* __restrict__ cursor, int32_t* __restrict__ order)
{
	__shared__ uint32_t cnt[ORDER_BINS], base[ORDER_BINS];
	cnt[threadIdx.x] = 0;
	__syncthreads();
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	int b = 0;
	uint32_t r = 0;
	if (i < n) {
		b = key[i];
		r = atomicAdd(&cnt[b], 1u);
	}
	__syncthreads();
	if (cnt[threadIdx.x])
		base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], cnt[threadIdx.x]);
	__syncthreads();
	if (i < n)
		order[base[b] + r] = (int32_t)i;
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v);

/* Zero the launch's work counter with a kernel (not hipMemsetAsync, so the
 * launch may be captured in a HIP graph). */
static hipError_t zero_queue(const DevScene& s, hipStream_t st)
{
	hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t*>(s.queue), (int64_t)2, 0u);
	return hipGetLastError();
}

/* Blocks of a persistent launch: as many as stay resident on every CU, but
 * no more than the items need. */
template <class K>
static unsigned persistent_grid(const DevScene& s, K kernel, int block, size_t lds, int64_t items)
{
	int nb = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, lds) != hipSuccess || nb < 1)
		nb = 1;
	const int64_t full = (int64_t)(s.num_cu > 0 ? s.num_cu : 1) * nb;
	const int64_t waves = block / 64;
	const int64_t need = (items + waves - 1) / waves;
	return (unsigned)(need < full ? need : full);
}


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

hipError_t launch_contacts(const DevScene& s, const double* q, int64_t n, int cap, int32_t* pairs, int32_t* count,
                           double* segs, int8_t* flags, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	const int block = pick_block(s);
	const int64_t grid = (n + block - 1) / block;
	const size_t lds = stack_bytes(s, block);
	if (s.wide)
		hipLaunchKernelGGL(k_contacts<true>, dim3((unsigned)grid), dim3(block), lds, st, s, q, n, cap, pairs, count);
	else
		hipLaunchKernelGGL(k_contacts<false>, dim3((unsigned)grid), dim3(block), lds, st, s, q, n, cap, pairs, count);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return e;
	const int64_t slots = n * (int64_t)cap;
	hipLaunchKernelGGL(k_contact_segments, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, st, s, q, n, cap,
	                   pairs, count, segs, flags);
	return hipGetLastError();
}

hipError_t launch_disentangled(const BoxPair& b, const PertData* pert, const double* q, int64_t n, uint8_t* out,
                               hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_disentangled, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, b, pert, q, n, out);
	return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
	return hipGetLastError();
}

__global__ void k_take_u32(uint32_t* flag, uint32_t* out)
{
	if (threadIdx.x == 0)
		*out = atomicExch(flag, 0u);
}

hipError_t launch_take_u32(uint32_t* flag, uint32_t* out, hipStream_t st)
{
	hipLaunchKernelGGL(k_take_u32, dim3(1), dim3(64), 0, st, flag, out);
	return hipGetLastError();
}

/* *flag |= 1 if bytes [0, n) of a and b differ (grid-stride, any grid). */
__global__ void k_differ(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b, int64_t n, uint32_t* flag)
{
	bool d = false;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
		d |= a[i] != b[i];
	if (__ballot(d) && (threadIdx.x & 63) == 0)
		atomicOr(flag, 1u);
}

hipError_t launch_differ(const void* a, const void* b, int64_t n, uint32_t* flag, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	const int64_t blocks = (n + 255) / 256;
	hipLaunchKernelGGL(k_differ, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, st,
	                   static_cast<const uint8_t*>(a), static_cast<const uint8_t*>(b), n, flag);
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
	const size_t lds = stack_bytes(s, block);
	hipError_t e = zero_queue(s, st);
	if (e != hipSuccess)
		return e;
	if (s.wide)
		hipLaunchKernelGGL(k_valid_states<true>, dim3(persistent_grid(s, k_valid_states<true>, block, lds, (n + 63) / 64)),
		                   dim3(block), lds, st, s, q, n, out);
	else
		hipLaunchKernelGGL(k_valid_states<false>,
		                   dim3(persistent_grid(s, k_valid_states<false>, block, lds, (n + 63) / 64)), dim3(block), lds,
		                   st, s, q, n, out);
	return hipGetLastError();
}

hipError_t launch_valid_states_subset(const DevScene& s, const double* q, int64_t n, const int32_t* idx,
                                      const uint32_t* count, uint8_t* out, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	const int block = pick_block(s);
	const size_t lds = stack_bytes(s, block);
	hipError_t e = zero_queue(s, st);
	if (e != hipSuccess)
		return e;
	if (s.wide)
		hipLaunchKernelGGL(k_valid_states_subset<true>,
		                   dim3(persistent_grid(s, k_valid_states_subset<true>, block, lds, (n + 63) / 64)), dim3(block),
		                   lds, st, s, q, n, idx, count, out);
	else
		hipLaunchKernelGGL(k_valid_states_subset<false>,
		                   dim3(persistent_grid(s, k_valid_states_subset<false>, block, lds, (n + 63) / 64)),
		                   dim3(block), lds, st, s, q, n, idx, count, out);
	return hipGetLastError();
}

hipError_t launch_ompl_states(const DevScene& s, const OmplXf& x, const double* q, int64_t n, int32_t* out,
                              hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	const int block = pick_block(s);
	const size_t lds = stack_bytes(s, block);
	hipError_t e = zero_queue(s, st);
	if (e != hipSuccess)
		return e;
	if (s.wide)
		hipLaunchKernelGGL(k_ompl_states<true>, dim3(persistent_grid(s, k_ompl_states<true>, block, lds, (n + 63) / 64)),
		                   dim3(block), lds, st, s, x, q, n, out);
	else
		hipLaunchKernelGGL(k_ompl_states<false>,
		                   dim3(persistent_grid(s, k_ompl_states<false>, block, lds, (n + 63) / 64)), dim3(block), lds,
		                   st, s, x, q, n, out);
	return hipGetLastError();
}

hipError_t launch_ompl_motions(const DevScene& s, const OmplXf& x, const double* q0, const double* q1, int64_t n,
                               double lvs, int32_t* out, int32_t* samples, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	const int block = 256;
	hipError_t e = zero_queue(s, st);
	if (e != hipSuccess)
		return e;
	const size_t lds = (block / 64) * (size_t)POSE_LDS_BYTES;
	if (s.wide)
		hipLaunchKernelGGL(k_ompl_motions<true>, dim3(persistent_grid(s, k_ompl_motions<true>, block, lds, n)),
		                   dim3(block), lds, st, s, x, q0, q1, n, lvs, out, samples);
	else
		hipLaunchKernelGGL(k_ompl_motions<false>, dim3(persistent_grid(s, k_ompl_motions<false>, block, lds, n)),
		                   dim3(block), lds, st, s, x, q0, q1, n, lvs, out, samples);
	return hipGetLastError();
}

/* k_batch's LDS: the traversal stacks of a batch with states, then one
 * wavefront's pose rows per wavefront. */
static inline size_t batch_lds_bytes(const DevScene& s, int block, bool states)
{
	return (states ? stack_bytes(s, block) : 0) + (size_t)(block / 64) * POSE_LDS_BYTES;
}

hipError_t launch_batch(const DevScene& s, const BatchArgs& a, int block, int grid, hipStream_t st)
{
	if (a.n_seg <= 0 && a.n_states <= 0)
		return hipSuccess;
	if (block < 64 || block > 256 || block % 64 != 0 || grid <= 0 || a.states_per_lane < 1)
		return hipErrorInvalidValue;
	hipError_t e = zero_queue(s, st);
	if (e != hipSuccess)
		return e;
	const bool states = a.n_states > 0;
	const size_t lds = batch_lds_bytes(s, block, states);
	if (states && s.wide)
		hipLaunchKernelGGL((k_batch<true, true>), dim3((unsigned)grid), dim3(block), lds, st, s, a);
	else if (states)
		hipLaunchKernelGGL((k_batch<false, true>), dim3((unsigned)grid), dim3(block), lds, st, s, a);
	else if (s.wide)
		hipLaunchKernelGGL((k_batch<true, false>), dim3((unsigned)grid), dim3(block), lds, st, s, a);
	else
		hipLaunchKernelGGL((k_batch<false, false>), dim3((unsigned)grid), dim3(block), lds, st, s, a);
	return hipGetLastError();
}

int batch_default_block(const DevScene& s, bool states) { return states ? pick_block(s) : 256; }

int batch_blocks_per_cu(const DevScene& s, int block, bool states)
{
	int nb = 0;
	const size_t lds = batch_lds_bytes(s, block, states);
	hipError_t e;
	if (states)
		e = s.wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<true, true>, block, lds)
		           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<false, true>, block, lds);
	else
		e = s.wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<true, false>, block, lds)
		           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<false, false>, block, lds);
	return (e == hipSuccess && nb > 0) ? nb : 1;
}

static constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static constexpr size_t order_bytes_for(int64_t n)
{
	return 2 * ORDER_BINS * sizeof(uint32_t) + align256((size_t)n) + (size_t)n * sizeof(int32_t);
}

size_t segment_record_bytes(int64_t n)
{
	if (n <= 0 || n >= ((int64_t)1 << 31))
		return 0;
	return align256((size_t)n * sizeof(SegShared));
}

size_t segment_order_bytes(int64_t n)
{
	if (n <= 0 || n >= ((int64_t)1 << 31))
		return 0;
	return order_bytes_for(n);
}

/* ---- scratch of the allocation-free touch search ----
 * One array per device, in the code object: the launch's work counter
 * (first 256 bytes), then the longest-first order of up to
 * NOALLOC_MAX_SEGMENTS segments.  Host code never passes the array's name to
 * a kernel: the host-side symbol is not a device address.  The device
 * address comes from hipGetSymbolAddress, once per device. */
static_assert(NOALLOC_MAX_SEGMENTS == SE3CQ_NOALLOC_MAX_SEGMENTS, "one limit");
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_COUNTER_BYTES + order_bytes_for(NOALLOC_MAX_SEGMENTS);
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch(void** out)
{
	static std::mutex mu;
	static void* ptr[64] = {};
	int dev = -1;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess)
		return e;
	if (dev < 0 || dev >= 64)
		return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> lk(mu);
	if (!ptr[dev]) {
		void* p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess)
			return e;
		ptr[dev] = p;
	}
	*out = ptr[dev];
	return hipSuccess;
}

/* The longest-first order of n segments into `scratch` (segment_order_bytes(n)
 * bytes: bin counts, bin cursors, one key byte per segment, the order). */
template <int MODE>
static hipError_t build_order(const double* from, const double* to, int64_t n, double vd, void* scratch,
                              const int32_t** order, hipStream_t st)
{
	uint32_t* hist = static_cast<uint32_t*>(scratch);
	uint32_t* cursor = hist + ORDER_BINS;
	uint8_t* key = reinterpret_cast<uint8_t*>(cursor + ORDER_BINS);
	int32_t* ord = reinterpret_cast<int32_t*>(key + align256((size_t)n));
	const unsigned blocks = (unsigned)((n + 255) / 256);
	hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(ORDER_BINS), 0, st, hist, (int64_t)ORDER_BINS, 0u);
	hipLaunchKernelGGL(k_order_keys<MODE>, dim3(blocks), dim3(256), 0, st, from, to, n, vd, key, hist);
	hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(ORDER_BINS), 0, st, hist, cursor);
	hipLaunchKernelGGL(k_order_scatter, dim3(blocks), dim3(256), 0, st, key, n, cursor, ord);
	*order = ord;
	return hipGetLastError();
}

template <int MODE>
static hipError_t launch_seg_mode(const DevScene& s_in, const double* from, const double* to, int64_t n, double vd,
                                  int levels, const SegOut& o, hipStream_t st, void* order_scratch, void* rec_scratch)
{
	/* one wavefront per segment, four per block; each wavefront keeps its
	 * pose rows in LDS, within a BisectShared record for the touch search */
	const int block = 256;
	const int per_block = block / 64;
	const size_t lds = per_block * ((MODE == MODE_TOUCH) ? sizeof(BisectShared) : (size_t)POSE_LDS_BYTES);
	DevScene s = s_in;
	const unsigned grid = s.wide ? persistent_grid(s, k_segments<true, MODE>, block, lds, n)
	                             : persistent_grid(s, k_segments<false, MODE>, block, lds, n);
	/* with no more segments than resident wavefronts every segment starts
	 * at once, and the order would change nothing */
	if (order_scratch && !s.order && n > (int64_t)grid * per_block && segment_order_bytes(n)) {
		hipError_t e = build_order<MODE>(from, to, n, vd, order_scratch, &s.order, st);
		if (e != hipSuccess)
			return e;
		/* many segments per wavefront: their records first, 64 per
		 * wavefront instruction (uneven shell segments 300 -> 290 ms; the
		 * touch search, whose set-up chains hide behind other wavefronts'
		 * walks, measured no gain and computes them in the walk) */
		if (MODE != MODE_TOUCH && rec_scratch && segment_record_bytes(n)) {
			hipLaunchKernelGGL(k_seg_records<MODE>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, from, to, n, vd,
			                   static_cast<SegShared*>(rec_scratch));
			e = hipGetLastError();
			if (e != hipSuccess)
				return e;
			s.seg_rec = rec_scratch;
		}
	}
	hipError_t e = zero_queue(s, st);
	if (e != hipSuccess)
		return e;
	if (s.wide)
		hipLaunchKernelGGL((k_segments<true, MODE>), dim3(grid), dim3(block), lds, st, s, from, to, n, vd, levels, o);
	else
		hipLaunchKernelGGL((k_segments<false, MODE>), dim3(grid), dim3(block), lds, st, s, from, to, n, vd, levels, o);
	return hipGetLastError();
}

hipError_t launch_segments(const DevScene& s, int mode, const double* from, const double* to, int64_t n, double vd,
                           int levels, const SegOut& o, hipStream_t st, void* order_scratch, void* rec_scratch)
{
	if (n <= 0)
		return hipSuccess;
	switch (mode) {
	case MODE_TRANSIT:
		return launch_seg_mode<MODE_TRANSIT>(s, from, to, n, vd, levels, o, st, order_scratch, rec_scratch);
	case MODE_TRANSITION:
		return launch_seg_mode<MODE_TRANSITION>(s, from, to, n, vd, levels, o, st, order_scratch, rec_scratch);
	case MODE_TOUCH:
		return launch_seg_mode<MODE_TOUCH>(s, from, to, n, vd, levels, o, st, order_scratch, rec_scratch);
	default: return hipErrorInvalidValue;
	}
}

template <int VIS>
static hipError_t launch_vis_mode(const DevScene& s, const double* qs0, const double* qs1, int64_t m, int64_t n,
                                  int64_t npairs, double ivd, int8_t* out8, uint32_t* bits, int32_t* steps,
                                  hipStream_t st)
{
	const int block = 256;
	const size_t lds = (block / 64) * (size_t)VIS_WAVE_DOUBLES * sizeof(double);
	hipError_t e = zero_queue(s, st);
	if (e != hipSuccess)
		return e;
	if (s.wide)
		hipLaunchKernelGGL((k_visibility<true, VIS>),
		                   dim3(persistent_grid(s, k_visibility<true, VIS>, block, lds, npairs)), dim3(block), lds, st,
		                   s, qs0, qs1, m, n, npairs, ivd, out8, bits, steps);
	else
		hipLaunchKernelGGL((k_visibility<false, VIS>),
		                   dim3(persistent_grid(s, k_visibility<false, VIS>, block, lds, npairs)), dim3(block), lds,
		                   st, s, qs0, qs1, m, n, npairs, ivd, out8, bits, steps);
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

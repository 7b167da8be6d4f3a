/*
 * engine.h -- internal data layout shared by the host-side world
 * (world.cpp) and the gfx950 kernels (kernels.hip).
 *
 * HBM layout per model (environment or robot), built once by angle_model:
 *   nodes : GNode[2*nf-1], 64 B each, children of an internal node are
 *           adjacent (first, first+1), nodes in breadth-first order so the
 *           top levels shared by every query sit in a few cache lines.
 *   tris  : binary64 xyz of the three vertices of every triangle, in leaf
 *           order (the triangles of a leaf are consecutive; 72 B per
 *           triangle); the exact leaf test reads these.
 *   tri_id / tri_pos : the mesh's face index of each triangle in leaf order,
 *           and its inverse (contact enumeration reports face indices).
 * Node boxes are binary32 and conservative: every vertex of every triangle
 * under a node lies inside the box defined by the stored binary32 centre,
 * axes and extents, checked in binary64 at build time.
 */
#ifndef SE3CQ_ENGINE_H
#define SE3CQ_ENGINE_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#else
struct alignas(16) float4 {
	float x, y, z, w;
};
#endif

namespace se3cq {

/* c_ex = (cx, cy, cz, ex), a0_ey = (axis0, ey), a1_ez = (axis1, ez),
 * a2_child = (axis2, child) with child >= 0 an internal node whose children
 * are child and child+1, and child < 0 a leaf whose code ~child is
 * 8 * first + count - 1: triangles first .. first + count - 1 of the leaf
 * order, 1 <= count <= 8 (the world's leaf size bounds it). */
struct alignas(16) GNode {
	float4 c_ex;
	float4 a0_ey;
	float4 a1_ez;
	float4 a2_child;
};

struct DevModel {
	const GNode* nodes;
	const double* tris;
	const int32_t* tri_id;  /* leaf order -> mesh face index */
	const int32_t* tri_pos; /* mesh face index -> leaf order */
	int32_t nnodes;
	int32_t ntris;
	int32_t depth;
	int32_t pad;
};

/* Perturbation (environment transform) data, read only when the
 * perturbation is not the identity. */
struct PertData {
	double Re[9];  /* rotation of the raw perturbation quaternion, row-major */
	double te[3];  /* perturbation translation */
	float PeT[9];  /* binary32 transpose of the normalised perturbation rotation */
	float fro_Re;  /* ||Re||_F, rounded up */
	float fro_dRe; /* sqrt(3) * ||Re - R(normalised)||_F, rounded up */
};

/* OMPL state -> unit-world state (se3_ompl_to_unit), passed by value. */
struct OmplXf {
	double M[16];    /* calibration matrix, row-major */
	double pert[7];  /* perturbation (t, w, x, y, z) */
	double delta[3]; /* (OMPL centre - robot centre) * scene scale */
};

/* Everything a kernel needs about the scene, passed by value. */
struct DevScene {
	DevModel env;
	DevModel rob;
	const PertData* pert;  /* device copy */
	float rob_reach;       /* max over robot nodes of |centre| + |extents|, rounded up */
	float scene_mag;       /* max |centre|_inf + max |extent| over both models */
	int32_t pert_identity; /* perturbation is exactly [0,0,0,1,0,0,0] */
	int32_t stack_depth;   /* per-thread traversal stack entries */
	int32_t wide;          /* 1 if node ids need more than 16 bits */
	uint32_t* error_flag;  /* device word: bit 0 traversal overflow, bit 1 a segment with 2^31+ steps */
	unsigned long long* queue; /* this launch's work counter (persistent kernels zero it first) */
	const int32_t* order;  /* segment kernels: item k walks segment order[k] (null: segment k) */
	const void* seg_rec;   /* segment kernels: segment k's precomputed record (null: computed in the walk) */
	const uint32_t* n_items; /* segment kernels: device count of items (null: one per segment) */
	int32_t num_cu;        /* compute units of the device (persistent grid size) */
	int32_t pool_chunk;    /* state kernels: states a wavefront takes per counter add */
	int32_t queue_depth;   /* segment and batch kernels: items a wavefront takes per counter add */
};

} // namespace se3cq

#endif

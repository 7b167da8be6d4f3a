/*
 * world.cpp -- host side of the SE(3) collision-query engine: geometry
 * ingest, unit-world normalisation, BVH construction and upload, and the
 * extern "C" entry points declared in include/se3cq.h.
 *
 * Geometry follows the reference bit for bit where it defines the collision
 * vertices:
 *   - Scene::load keeps binary32 positions and takes the binary32 mean of the
 *     vertices as the centre; the bounding box is seeded at the origin
 *     (lib/osr/scene.cc:40-50,135-145, lib/osr/bbox.h:22-51).
 *   - scaleToUnit: scale = float(1.0 / max(span_env, span_rob))
 *     (lib/osr/unit_world.cc:123-131).
 *   - angleModel(0,0): env matrix diag(s); robot matrix scale(s)*translate(-c)
 *     built in binary32 (glm), then CDModel::addVF applies it in binary64:
 *     (m0*v0 + m1*v1) + (m2*v2 + m3) (lib/osr/cdmodel.cc:120-160).
 * The BVH is this engine's own (PCA-fitted boxes, median split along the
 * box axis that minimises the children's summed surface area, binary32
 * storage made conservative against the binary64 vertices); only the exact
 * leaf test
 * decides a boolean, so the tree shape never changes an answer.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/se3cq.h"
#include "engine.h"
#include "launch.h"
#include "se3_math.h"
#include "cull.h"

using namespace se3cq;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
	g_err = msg;
	return code;
}

struct Mesh {
	std::vector<float> V;
	std::vector<uint32_t> F;
	float center[3] = {0, 0, 0};
	float ompl_center[3] = {0, 0, 0};
	float bbox[6] = {0, 0, 0, 0, 0, 0}; /* l r b t k f */
	bool loaded = false;
};

struct HostModel {
	std::vector<double> Vcd;      /* unit-world vertices [nv*3] */
	std::vector<GNode> nodes;     /* BFS order */
	std::vector<double> tris;     /* [nf*9], leaf order */
	std::vector<int32_t> tri_id;  /* leaf order -> mesh face index */
	std::vector<int32_t> tri_pos; /* mesh face index -> leaf order */
	int depth = 0;
	float reach = 0.0f;           /* max |c| + |e| over nodes */
	float mag = 0.0f;             /* max |c|_inf + max extent */
	double bc[3] = {0, 0, 0};     /* AABB of Vcd: centre (min+max)*0.5 */
	double bh[3] = {0, 0, 0};     /* and half extents (max-min)*0.5 */
};

struct DevBuf {
	void* p = nullptr;
	size_t n = 0;
	void release()
	{
		if (p)
			(void)hipFree(p);
		p = nullptr;
		n = 0;
	}
	bool reserve(size_t bytes)
	{
		if (bytes <= n)
			return true;
		release();
		if (hipMalloc(&p, bytes) != hipSuccess) {
			p = nullptr;
			return false;
		}
		n = bytes;
		return true;
	}
};

} // namespace

/* Work counters of the persistent kernels per world: QUEUE_SLOTS words 64 B
 * apart for eager launches, GRAPH_WORDS more for launches captured into HIP
 * graphs (one each, never handed out again, since a graph may replay at
 * any later time). */
constexpr int QUEUE_SLOTS = 64;
constexpr int GRAPH_WORDS = 4096;

namespace {

/* One eager work counter and what the world knows of its last launch: the
 * stream and an event recorded after it.  A launch on another stream that
 * takes the slot waits for that event first, so two launches never share a
 * counter while both run.  `order` is the slot's segment-order scratch. */
struct QueueSlot {
	hipEvent_t ev = nullptr;
	hipStream_t last = nullptr;
	bool used = false;
	DevBuf order;
};

/* What one blocking (_host) call works with: its own stream, staging
 * buffers and error word, so calls from several threads run side by side. */
struct HostCtx {
	hipStream_t st = nullptr;
	DevBuf ws[12];
	DevBuf err;
	~HostCtx()
	{
		for (auto& b : ws)
			b.release();
		err.release();
		if (st)
			(void)hipStreamDestroy(st);
	}
};

/* A perturbation as the kernels read it.  Blocks are never overwritten:
 * set_perturbation selects (or adds) the block of the new value, so a
 * launch in flight, or a captured graph, keeps reading the value it was
 * launched with. */
struct PertBlock {
	double q[7];
	bool identity;
	DevBuf buf;
};

} // namespace

struct se3cq_world {
	Mesh mesh[2];
	float scale = 1.0f;
	float xf[2][16];       /* binary32 column-major (glm) */
	bool has_model[2] = {false, false};
	HostModel hm[2];
	double pert[7] = {0, 0, 0, 1, 0, 0, 0};
	double cres = 0.0;
	/* device */
	int device = -1;
	bool uploaded = false;
	hipStream_t stream = nullptr; /* the world's own stream (tuning, error checks) */
	DevBuf d_nodes[2], d_tris[2], d_tid[2], d_tpos[2];
	int leaf_size = 1;     /* triangles per BVH leaf, at most */
	int pool_chunk = 64;   /* states a wavefront takes per counter add (state kernels) */
	int queue_depth = 1;   /* work items a wavefront takes per counter add (segments, batches) */
	/* [0]: error bits of _device launches, kept until se3cq_check takes
	 * them; [1]: where se3cq_check's kernel puts them */
	DevBuf d_err;
	DevBuf d_queue;
	QueueSlot qslot[QUEUE_SLOTS];
	uint32_t q_rr = 0;
	DevBuf d_graph;
	int graph_used = 0;
	std::mutex qmu; /* slots, graph words and the current perturbation */
	std::vector<std::unique_ptr<PertBlock>> perts;
	const PertBlock* pert_cur = nullptr;
	std::vector<std::unique_ptr<HostCtx>> hctx;
	std::vector<HostCtx*> hfree;
	std::mutex hmu;
	int num_cu = 0;
	se3cq_batch_config batch_cfg = {0, 0, 1, 1};
	std::mutex mu; /* set-up: geometry, upload, perturbation, configuration */

	se3cq_world()
	{
		for (int g = 0; g < 2; g++)
			for (int i = 0; i < 16; i++)
				xf[g][i] = (i % 5 == 0) ? 1.0f : 0.0f;
	}
	~se3cq_world() { release_device(); }
	void release_device()
	{
		if (uploaded)
			(void)hipDeviceSynchronize();
		for (int g = 0; g < 2; g++) {
			d_nodes[g].release();
			d_tris[g].release();
			d_tid[g].release();
			d_tpos[g].release();
		}
		d_err.release();
		d_queue.release();
		d_graph.release();
		graph_used = 0;
		for (auto& q : qslot) {
			if (q.ev)
				(void)hipEventDestroy(q.ev);
			q.ev = nullptr;
			q.last = nullptr;
			q.used = false;
			q.order.release();
		}
		perts.clear();
		pert_cur = nullptr;
		hfree.clear();
		hctx.clear();
		if (stream)
			(void)hipStreamDestroy(stream);
		stream = nullptr;
		uploaded = false;
	}
	bool ready() const { return has_model[0] && has_model[1]; }
};

namespace {

/* ---------------------------------------------------------------------- */
/* Mesh ingest                                                            */
/* ---------------------------------------------------------------------- */
void mesh_finish(Mesh& m)
{
	float l = 0, r = 0, b = 0, t = 0, k = 0, f = 0;
	float sx = 0.0f, sy = 0.0f, sz = 0.0f;
	const size_t nv = m.V.size() / 3;
	for (size_t i = 0; i < nv; i++) {
		const float x = m.V[3 * i], y = m.V[3 * i + 1], z = m.V[3 * i + 2];
		r = std::max(x, r);
		l = std::min(x, l);
		t = std::max(y, t);
		b = std::min(y, b);
		f = std::max(z, f);
		k = std::min(z, k);
		sx += x;
		sy += y;
		sz += z;
	}
	const float bb[6] = {l, r, b, t, k, f};
	std::memcpy(m.bbox, bb, sizeof(bb));
	const float n = (float)nv;
	m.ompl_center[0] = sx / n;
	m.ompl_center[1] = sy / n;
	m.ompl_center[2] = sz / n;
	std::memcpy(m.center, m.ompl_center, sizeof(m.center));
	m.loaded = true;
}

int mesh_from_arrays(Mesh& m, const float* V, int64_t nv, const uint32_t* F, int64_t nf)
{
	if (!V || !F || nv <= 0 || nf <= 0)
		return fail(SE3CQ_EINVAL, "load_mesh: empty mesh");
	for (int64_t i = 0; i < 3 * nf; i++)
		if ((int64_t)F[i] >= nv)
			return fail(SE3CQ_EINVAL, "load_mesh: face index out of range");
	/* The exact leaf test relies on P1 - P1 being the zero vector and on
	 * products staying finite, so coordinates must be finite and moderate. */
	for (int64_t i = 0; i < 3 * nv; i++)
		if (!(std::fabs(V[i]) <= 1e30f))
			return fail(SE3CQ_EINVAL, "load_mesh: non-finite or out-of-range vertex coordinate");
	m = Mesh();
	m.V.assign(V, V + 3 * nv);
	m.F.assign(F, F + 3 * nf);
	mesh_finish(m);
	return SE3CQ_OK;
}

/* Wavefront OBJ as assimp (Triangulate | JoinIdenticalVertices) presents it
 * to Mesh::Mesh (lib/osr/mesh.cc:24-63): polygons become triangle fans, and
 * each face corner, in file order, joins the first already-kept vertex at the
 * same position whose normal and uv are each within a squared binary32
 * distance of (1e-5f)^2 of its own (assimp's JoinVerticesProcess, external);
 * a corner that joins none becomes a new vertex. */
namespace {
struct ObjCorner {
	float a[8]; // pos3, uv2, nrm3
};
struct PosKey {
	float p[3];
	bool operator==(const PosKey& o) const { return p[0] == o.p[0] && p[1] == o.p[1] && p[2] == o.p[2]; }
};
struct PosKeyHash {
	size_t operator()(const PosKey& k) const
	{
		float q[3] = {k.p[0] + 0.0f, k.p[1] + 0.0f, k.p[2] + 0.0f}; // -0 == +0
		uint64_t h = 1469598103934665603ULL;
		const unsigned char* b = reinterpret_cast<const unsigned char*>(q);
		for (size_t i = 0; i < sizeof(q); i++)
			h = (h ^ b[i]) * 1099511628211ULL;
		return (size_t)h;
	}
};
float sq_gap(const float* a, const float* b, int n)
{
	float s = 0.0f;
	for (int k = 0; k < n; k++) {
		const float d = a[k] - b[k];
		s = s + d * d;
	}
	return s;
}
bool corners_join(const ObjCorner& kept, const ObjCorner& c)
{
	const float eps = 1e-5f;
	const float eps2 = eps * eps;
	return !(sq_gap(kept.a + 5, c.a + 5, 3) > eps2) && !(sq_gap(kept.a + 3, c.a + 3, 2) > eps2);
}
} // namespace

int mesh_from_obj(Mesh& m, const char* path)
{
	FILE* fp = std::fopen(path, "r");
	if (!fp)
		return fail(SE3CQ_EIO, std::string("cannot open ") + path);
	std::vector<float> pv, pt, pn;
	std::unordered_map<PosKey, std::vector<uint32_t>, PosKeyHash> at_pos;
	std::vector<ObjCorner> kept;
	std::vector<float> V;
	std::vector<uint32_t> F;
	char line[8192];
	while (std::fgets(line, sizeof(line), fp)) {
		char* s = line;
		while (*s == ' ' || *s == '\t')
			s++;
		if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
			char* e = s + 1;
			for (int k = 0; k < 3; k++)
				pv.push_back(std::strtof(e, &e));
		} else if (s[0] == 'v' && s[1] == 't') {
			char* e = s + 2;
			for (int k = 0; k < 2; k++)
				pt.push_back(std::strtof(e, &e));
		} else if (s[0] == 'v' && s[1] == 'n') {
			char* e = s + 2;
			for (int k = 0; k < 3; k++)
				pn.push_back(std::strtof(e, &e));
		} else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
			std::vector<uint32_t> ids;
			char* e = s + 1;
			for (;;) {
				while (*e == ' ' || *e == '\t')
					e++;
				if (*e == '\0' || *e == '\n' || *e == '\r')
					break;
				long vi = std::strtol(e, &e, 10), ti = 0, ni = 0;
				if (*e == '/') {
					e++;
					if (*e != '/')
						ti = std::strtol(e, &e, 10);
					if (*e == '/') {
						e++;
						ni = std::strtol(e, &e, 10);
					}
				}
				while (*e && *e != ' ' && *e != '\t' && *e != '\n' && *e != '\r')
					e++;
				const long nvv = (long)pv.size() / 3, ntt = (long)pt.size() / 2, nnn = (long)pn.size() / 3;
				if (vi < 0) vi = nvv + 1 + vi;
				if (ti < 0) ti = ntt + 1 + ti;
				if (ni < 0) ni = nnn + 1 + ni;
				ObjCorner c;
				std::memset(c.a, 0, sizeof(c.a));
				if (vi >= 1 && vi <= nvv)
					std::memcpy(c.a, &pv[3 * (vi - 1)], 3 * sizeof(float));
				if (ti >= 1 && ti <= ntt)
					std::memcpy(c.a + 3, &pt[2 * (ti - 1)], 2 * sizeof(float));
				if (ni >= 1 && ni <= nnn)
					std::memcpy(c.a + 5, &pn[3 * (ni - 1)], 3 * sizeof(float));
				std::vector<uint32_t>& here = at_pos[PosKey{{c.a[0], c.a[1], c.a[2]}}];
				uint32_t id = UINT32_MAX;
				for (uint32_t u : here)
					if (corners_join(kept[u], c)) {
						id = u;
						break;
					}
				if (id == UINT32_MAX) {
					id = (uint32_t)kept.size();
					kept.push_back(c);
					here.push_back(id);
					V.insert(V.end(), c.a, c.a + 3);
				}
				ids.push_back(id);
			}
			for (size_t k = 1; k + 1 < ids.size(); k++) {
				F.push_back(ids[0]);
				F.push_back(ids[k]);
				F.push_back(ids[k + 1]);
			}
		}
	}
	std::fclose(fp);
	if (V.empty() || F.empty())
		return fail(SE3CQ_EIO, std::string("no triangles in ") + path);
	return mesh_from_arrays(m, V.data(), (int64_t)V.size() / 3, F.data(), (int64_t)F.size() / 3);
}

float bbox_span(const float* b)
{
	return std::max(std::max(b[1] - b[0], b[3] - b[2]), b[5] - b[4]);
}

/* ---------------------------------------------------------------------- */
/* BVH                                                                    */
/* ---------------------------------------------------------------------- */
void sym_eigen3(double A[3][3], double E[3][3])
{
	/* cyclic Jacobi; E columns are eigenvectors */
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++)
			E[i][j] = (i == j) ? 1.0 : 0.0;
	for (int sweep = 0; sweep < 50; sweep++) {
		const double off = std::fabs(A[0][1]) + std::fabs(A[0][2]) + std::fabs(A[1][2]);
		const double dia = std::fabs(A[0][0]) + std::fabs(A[1][1]) + std::fabs(A[2][2]);
		if (off == 0.0 || off <= 1e-24 * dia)
			break;
		for (int p = 0; p < 2; p++)
			for (int q = p + 1; q < 3; q++) {
				if (A[p][q] == 0.0)
					continue;
				const double th = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
				const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
				const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
				for (int k = 0; k < 3; k++) {
					const double kp = A[k][p], kq = A[k][q];
					A[k][p] = c * kp - s * kq;
					A[k][q] = s * kp + c * kq;
				}
				for (int k = 0; k < 3; k++) {
					const double pk = A[p][k], qk = A[q][k];
					A[p][k] = c * pk - s * qk;
					A[q][k] = s * pk + c * qk;
				}
				for (int k = 0; k < 3; k++) {
					const double kp = E[k][p], kq = E[k][q];
					E[k][p] = c * kp - s * kq;
					E[k][q] = s * kp + c * kq;
				}
			}
	}
}


/* Fit a box to the vertices of triangles perm[lo..hi) and store it in
 * binary32 such that every vertex v satisfies |a_i . (v - c)| <= e_i for the
 * stored a_i, c, e_i evaluated exactly. */
void fit_node(const HostModel& m, const uint32_t* F, const std::vector<int>& perm, int lo, int hi, GNode& g,
              double* c_out, double* e_out)
{
	double mean[3] = {0, 0, 0};
	const int cnt = 3 * (hi - lo);
	for (int i = lo; i < hi; i++)
		for (int k = 0; k < 3; k++) {
			const double* v = &m.Vcd[3 * F[3 * perm[i] + k]];
			for (int d = 0; d < 3; d++)
				mean[d] += v[d];
		}
	for (int d = 0; d < 3; d++)
		mean[d] /= cnt;
	double C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
	for (int i = lo; i < hi; i++)
		for (int k = 0; k < 3; k++) {
			const double* v = &m.Vcd[3 * F[3 * perm[i] + k]];
			const double d[3] = {v[0] - mean[0], v[1] - mean[1], v[2] - mean[2]};
			for (int r = 0; r < 3; r++)
				for (int c = 0; c < 3; c++)
					C[r][c] += d[r] * d[c];
		}
	double E[3][3];
	sym_eigen3(C, E);
	int ord[3] = {0, 1, 2};
	std::sort(ord, ord + 3, [&](int a, int b) { return C[a][a] > C[b][b]; });
	double ax[3][3];
	for (int k = 0; k < 2; k++) {
		double n = 0;
		for (int r = 0; r < 3; r++)
			n += E[r][ord[k]] * E[r][ord[k]];
		n = std::sqrt(n);
		for (int r = 0; r < 3; r++)
			ax[k][r] = E[r][ord[k]] / n;
	}
	ax[2][0] = ax[0][1] * ax[1][2] - ax[0][2] * ax[1][1];
	ax[2][1] = ax[0][2] * ax[1][0] - ax[0][0] * ax[1][2];
	ax[2][2] = ax[0][0] * ax[1][1] - ax[0][1] * ax[1][0];
	{
		const double n = std::sqrt(ax[2][0] * ax[2][0] + ax[2][1] * ax[2][1] + ax[2][2] * ax[2][2]);
		for (int r = 0; r < 3; r++)
			ax[2][r] /= n;
	}
	float af[3][3];
	for (int i = 0; i < 3; i++)
		for (int r = 0; r < 3; r++)
			af[i][r] = (float)ax[i][r];
	double plo[3] = {INFINITY, INFINITY, INFINITY}, phi[3] = {-INFINITY, -INFINITY, -INFINITY};
	for (int i = lo; i < hi; i++)
		for (int k = 0; k < 3; k++) {
			const double* v = &m.Vcd[3 * F[3 * perm[i] + k]];
			for (int a = 0; a < 3; a++) {
				const double p = (double)af[a][0] * v[0] + (double)af[a][1] * v[1] + (double)af[a][2] * v[2];
				plo[a] = std::min(plo[a], p);
				phi[a] = std::max(phi[a], p);
			}
		}
	double c[3] = {0, 0, 0};
	for (int a = 0; a < 3; a++) {
		const double mid = 0.5 * (plo[a] + phi[a]);
		for (int r = 0; r < 3; r++)
			c[r] += mid * (double)af[a][r];
	}
	float cf[3] = {(float)c[0], (float)c[1], (float)c[2]};
	double e[3] = {0, 0, 0};
	for (int i = lo; i < hi; i++)
		for (int k = 0; k < 3; k++) {
			const double* v = &m.Vcd[3 * F[3 * perm[i] + k]];
			const double d[3] = {v[0] - (double)cf[0], v[1] - (double)cf[1], v[2] - (double)cf[2]};
			for (int a = 0; a < 3; a++) {
				const double p = (double)af[a][0] * d[0] + (double)af[a][1] * d[1] + (double)af[a][2] * d[2];
				e[a] = std::max(e[a], std::fabs(p));
			}
		}
	float ef[3];
	for (int a = 0; a < 3; a++)
		ef[a] = round_up_f(e[a] * (1.0 + 0x1p-40) + 0x1p-60);
	g.c_ex = {cf[0], cf[1], cf[2], ef[0]};
	g.a0_ey = {af[0][0], af[0][1], af[0][2], ef[1]};
	g.a1_ez = {af[1][0], af[1][1], af[1][2], ef[2]};
	g.a2_child = {af[2][0], af[2][1], af[2][2], 0.0f};
	for (int r = 0; r < 3; r++) {
		c_out[r] = cf[r];
		e_out[r] = ef[r];
	}
}

void set_child(GNode& g, int32_t child)
{
	float f;
	std::memcpy(&f, &child, sizeof(f));
	g.a2_child.w = f;
}

/* Unit-world geometry and BVH of one model; leaves hold at most leaf_size
 * (1-8) triangles. */
void build_model(HostModel& m, const Mesh& me, const float* M, int leaf_size)
{
	const size_t nv = me.V.size() / 3, nf = me.F.size() / 3;
	m.Vcd.resize(3 * nv);
	for (size_t i = 0; i < nv; i++) {
		const float* v = &me.V[3 * i];
		for (int r = 0; r < 3; r++) {
			const double a = (double)M[0 * 4 + r] * (double)v[0] + (double)M[1 * 4 + r] * (double)v[1];
			const double b = (double)M[2 * 4 + r] * (double)v[2] + (double)M[3 * 4 + r] * 1.0;
			m.Vcd[3 * i + r] = a + b;
		}
	}
	/* FCL's BVHModel::computeLocalAABB over the collision vertices, wrapped
	 * in an fcl::Box for isDisentangled (cdmodel.cc:90-100) */
	{
		double lo[3], hi[3];
		for (int k = 0; k < 3; k++)
			lo[k] = hi[k] = nv ? m.Vcd[k] : 0.0;
		for (size_t i = 1; i < nv; i++)
			for (int k = 0; k < 3; k++) {
				const double v = m.Vcd[3 * i + k];
				if (v < lo[k])
					lo[k] = v;
				if (v > hi[k])
					hi[k] = v;
			}
		for (int k = 0; k < 3; k++) {
			m.bc[k] = (lo[k] + hi[k]) * 0.5;
			m.bh[k] = (hi[k] - lo[k]) * 0.5;
		}
	}
	m.tris.resize(9 * nf);
	for (size_t t = 0; t < nf; t++)
		for (int k = 0; k < 3; k++)
			for (int r = 0; r < 3; r++)
				m.tris[9 * t + 3 * k + r] = m.Vcd[3 * me.F[3 * t + k] + r];
	std::vector<double> cen(3 * nf);
	for (size_t t = 0; t < nf; t++)
		for (int r = 0; r < 3; r++)
			cen[3 * t + r] = (m.tris[9 * t + r] + m.tris[9 * t + 3 + r] + m.tris[9 * t + 6 + r]) / 3.0;
	std::vector<int> perm(nf);
	for (size_t t = 0; t < nf; t++)
		perm[t] = (int)t;
	m.nodes.assign(2 * nf - 1, GNode());
	struct Item {
		int idx, lo, hi, depth;
	};
	std::deque<Item> q;
	q.push_back({0, 0, (int)nf, 0});
	int next = 1;
	m.depth = 0;
	double reach = 0.0, cmax = 0.0, emax = 0.0;
	while (!q.empty()) {
		Item it = q.front();
		q.pop_front();
		m.depth = std::max(m.depth, it.depth);
		GNode& g = m.nodes[it.idx];
		double c[3], e[3];
		fit_node(m, me.F.data(), perm, it.lo, it.hi, g, c, e);
		reach = std::max(reach, std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) +
		                            std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]));
		cmax = std::max(cmax, std::max(std::fabs(c[0]), std::max(std::fabs(c[1]), std::fabs(c[2]))));
		emax = std::max(emax, std::max(e[0], std::max(e[1], e[2])));
		if (it.hi - it.lo <= leaf_size) {
			/* leaf code 8 * first + count - 1 (GNode) */
			set_child(g, ~(8 * it.lo + (it.hi - it.lo - 1)));
			continue;
		}
		/* Median split along whichever of the node box's three axes gives
		 * the smallest summed surface area of the two child boxes: the
		 * expected number of box tests in a BVH x BVH traversal grows with
		 * child box size. */
		const int mid = (it.lo + it.hi) / 2;
		const float* axes[3] = {&g.a0_ey.x, &g.a1_ez.x, &g.a2_child.x};
		double best = INFINITY;
		std::vector<int> best_perm;
		for (int ax = 0; ax < 3; ax++) {
			const double a0 = axes[ax][0], a1 = axes[ax][1], a2 = axes[ax][2];
			auto proj = [&](int t) { return a0 * cen[3 * t] + a1 * cen[3 * t + 1] + a2 * cen[3 * t + 2]; };
			std::nth_element(perm.begin() + it.lo, perm.begin() + mid, perm.begin() + it.hi,
			                 [&](int a, int b) { return proj(a) < proj(b); });
			GNode gl, gr;
			double cl[3], el[3], cr[3], er[3];
			fit_node(m, me.F.data(), perm, it.lo, mid, gl, cl, el);
			fit_node(m, me.F.data(), perm, mid, it.hi, gr, cr, er);
			const double cost = (el[0] * el[1] + el[1] * el[2] + el[0] * el[2]) +
			                    (er[0] * er[1] + er[1] * er[2] + er[0] * er[2]);
			if (cost < best) {
				best = cost;
				best_perm.assign(perm.begin() + it.lo, perm.begin() + it.hi);
			}
		}
		std::copy(best_perm.begin(), best_perm.end(), perm.begin() + it.lo);
		const int first = next;
		next += 2;
		set_child(g, first);
		q.push_back({first, it.lo, mid, it.depth + 1});
		q.push_back({first + 1, mid, it.hi, it.depth + 1});
	}
	m.nodes.resize(next);
	/* each box's axes stored thinnest first: the culling test tries its
	 * face axes in storage order and a wavefront leaves it once every lane
	 * has found a separating axis, which the thinnest axis is the likeliest
	 * to be (the separating-axis test itself does not depend on the order
	 * or the handedness of the axes) */
#ifndef SE3CQ_AXES_THIN_FIRST
#define SE3CQ_AXES_THIN_FIRST 1
#endif
	if (SE3CQ_AXES_THIN_FIRST)
		for (GNode& g : m.nodes) {
			float ax[3][4] = {{g.a0_ey.x, g.a0_ey.y, g.a0_ey.z, g.c_ex.w},
			                  {g.a1_ez.x, g.a1_ez.y, g.a1_ez.z, g.a0_ey.w},
			                  {g.a2_child.x, g.a2_child.y, g.a2_child.z, g.a1_ez.w}};
			int o[3] = {0, 1, 2};
			std::stable_sort(o, o + 3, [&](int a, int b) { return ax[a][3] < ax[b][3]; });
			g.a0_ey = {ax[o[0]][0], ax[o[0]][1], ax[o[0]][2], ax[o[1]][3]};
			g.a1_ez = {ax[o[1]][0], ax[o[1]][1], ax[o[1]][2], ax[o[2]][3]};
			g.c_ex.w = ax[o[0]][3];
			g.a2_child.x = ax[o[2]][0];
			g.a2_child.y = ax[o[2]][1];
			g.a2_child.z = ax[o[2]][2];
		}
	/* triangles in leaf order */
	std::vector<double> mesh_order;
	mesh_order.swap(m.tris);
	m.tris.resize(9 * nf);
	m.tri_id.resize(nf);
	m.tri_pos.resize(nf);
	for (size_t i = 0; i < nf; i++) {
		const int t = perm[i];
		std::copy(&mesh_order[9 * (size_t)t], &mesh_order[9 * (size_t)t] + 9, &m.tris[9 * i]);
		m.tri_id[i] = t;
		m.tri_pos[t] = (int32_t)i;
	}
	m.reach = round_up_f(reach * (1.0 + 1e-6));
	m.mag = round_up_f(cmax + emax);
}

/* ---------------------------------------------------------------------- */
/* Device scene                                                           */
/* ---------------------------------------------------------------------- */
bool pert_is_identity(const double* p)
{
	return p[0] == 0.0 && p[1] == 0.0 && p[2] == 0.0 && p[3] == 1.0 && p[4] == 0.0 && p[5] == 0.0 && p[6] == 0.0;
}

/* Everything of the scene but the launch's own work counter and error word
 * (call with qmu held, so the perturbation is read once). */
DevScene base_scene(const se3cq_world* w)
{
	DevScene s;
	std::memset(&s, 0, sizeof(s));
	for (int g = 0; g < 2; g++) {
		DevModel& dm = g ? s.rob : s.env;
		dm.nodes = (const GNode*)w->d_nodes[g].p;
		dm.tris = (const double*)w->d_tris[g].p;
		dm.tri_id = (const int32_t*)w->d_tid[g].p;
		dm.tri_pos = (const int32_t*)w->d_tpos[g].p;
		dm.nnodes = (int32_t)w->hm[g].nodes.size();
		dm.ntris = (int32_t)(w->hm[g].tris.size() / 9);
		dm.depth = w->hm[g].depth;
	}
	s.pert = (const PertData*)w->pert_cur->buf.p;
	s.pert_identity = w->pert_cur->identity ? 1 : 0;
	s.rob_reach = w->hm[1].reach;
	s.scene_mag = w->hm[0].mag + w->hm[1].mag;
	s.stack_depth = w->hm[0].depth + w->hm[1].depth + 2;
	s.wide = (w->hm[0].nodes.size() > 65536 || w->hm[1].nodes.size() > 65536) ? 1 : 0;
	s.num_cu = w->num_cu;
	s.pool_chunk = w->pool_chunk;
	s.queue_depth = w->queue_depth;
	return s;
}


/* Select the device block of w->pert, adding one if this value has none
 * yet (call with mu held, after the upload). */
int select_pert(se3cq_world* w)
{
	for (const auto& b : w->perts)
		if (std::memcmp(b->q, w->pert, sizeof(b->q)) == 0) {
			std::lock_guard<std::mutex> lk(w->qmu);
			w->pert_cur = b.get();
			return SE3CQ_OK;
		}
	std::unique_ptr<PertBlock> b(new (std::nothrow) PertBlock());
	if (!b || !b->buf.reserve(sizeof(PertData)))
		return fail(SE3CQ_ENOMEM, "device allocation failed");
	std::memcpy(b->q, w->pert, sizeof(b->q));
	b->identity = pert_is_identity(w->pert);
	const PertData d = make_pert(w->pert);
	if (hipMemcpy(b->buf.p, &d, sizeof(d), hipMemcpyHostToDevice) != hipSuccess)
		return fail(SE3CQ_EDEVICE, "device upload failed");
	std::lock_guard<std::mutex> lk(w->qmu);
	w->perts.push_back(std::move(b));
	w->pert_cur = w->perts.back().get();
	return SE3CQ_OK;
}

/* Copy both trees to the device.  The node buffer is sized for one
 * triangle per leaf (2 nf - 1 nodes), the most any leaf size needs, so a
 * later change of leaf size rewrites the same buffers. */
int upload_trees(se3cq_world* w)
{
	for (int g = 0; g < 2; g++) {
		const HostModel& m = w->hm[g];
		const size_t nf = m.tri_id.size();
		const size_t nb = m.nodes.size() * sizeof(GNode), tb = m.tris.size() * sizeof(double);
		const size_t ib = nf * sizeof(int32_t);
		if (!w->d_nodes[g].reserve((2 * nf - 1) * sizeof(GNode)) || !w->d_tris[g].reserve(tb) ||
		    !w->d_tid[g].reserve(ib) || !w->d_tpos[g].reserve(ib))
			return fail(SE3CQ_ENOMEM, "device allocation failed");
		if (hipMemcpy(w->d_nodes[g].p, m.nodes.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
		    hipMemcpy(w->d_tris[g].p, m.tris.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
		    hipMemcpy(w->d_tid[g].p, m.tri_id.data(), ib, hipMemcpyHostToDevice) != hipSuccess ||
		    hipMemcpy(w->d_tpos[g].p, m.tri_pos.data(), ib, hipMemcpyHostToDevice) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "device upload failed");
	}
	return SE3CQ_OK;
}

int upload(se3cq_world* w)
{
	if (hipGetDevice(&w->device) != hipSuccess)
		return fail(SE3CQ_EDEVICE, "no HIP device");
	if (!w->stream && hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess)
		return fail(SE3CQ_EDEVICE, "hipStreamCreate failed");
	{
		const int rc = upload_trees(w);
		if (rc)
			return rc;
	}
	if (!w->d_err.reserve(2 * sizeof(uint32_t)) || hipMemset(w->d_err.p, 0, 2 * sizeof(uint32_t)) != hipSuccess)
		return fail(SE3CQ_EDEVICE, "device allocation failed");
	if (!w->d_queue.reserve(QUEUE_SLOTS * 64) || !w->d_graph.reserve((size_t)GRAPH_WORDS * 64))
		return fail(SE3CQ_EDEVICE, "device allocation failed");
	for (auto& q : w->qslot)
		if (!q.ev && hipEventCreateWithFlags(&q.ev, hipEventDisableTiming) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "hipEventCreate failed");
	if (hipDeviceGetAttribute(&w->num_cu, hipDeviceAttributeMultiprocessorCount, w->device) != hipSuccess ||
	    w->num_cu <= 0)
		return fail(SE3CQ_EDEVICE, "cannot read the device's CU count");
	{
		const int rc = select_pert(w);
		if (rc)
			return rc;
	}
	const int need = (int)(w->hm[0].depth + w->hm[1].depth + 2) * 2 * 4 * 64;
	if (need > 48 * 1024)
		return fail(SE3CQ_EINVAL, "BVH too deep for the traversal stack");
	w->uploaded = true;
	return SE3CQ_OK;
}

/* Upload the models on first use (angle_model itself needs no device, so the
 * host-side geometry can be inspected on machines without one).  Queries
 * without a device fail: there is no CPU fallback. */
int ensure_device(se3cq_world* w)
{
	if (w->uploaded)
		return SE3CQ_OK;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
		return fail(SE3CQ_EDEVICE, "no HIP device available (the engine has no CPU fallback)");
	return upload(w);
}

int setup_device(se3cq_world* w)
{
	std::lock_guard<std::mutex> lk(w->mu);
	return ensure_device(w);
}

int error_bits(uint32_t flag)
{
	if (flag & 1u)
		return fail(SE3CQ_EDEVICE, "traversal stack overflow on the device");
	if (flag & 2u)
		return fail(SE3CQ_EDEVICE, "a segment needs 2^31 or more verify steps (distance / verify step)");
	if (flag)
		return fail(SE3CQ_EDEVICE, "device error flag set");
	return SE3CQ_OK;
}

/* Take (read and clear, atomically on the device) the error bits of the
 * world's _device launches. */
int take_device_errors(se3cq_world* w)
{
	uint32_t* e = (uint32_t*)w->d_err.p;
	uint32_t flag = 0;
	if (launch_take_u32(e, e + 1, w->stream) != hipSuccess ||
	    hipMemcpyAsync(&flag, e + 1, sizeof(flag), hipMemcpyDeviceToHost, w->stream) != hipSuccess ||
	    hipStreamSynchronize(w->stream) != hipSuccess)
		return fail(SE3CQ_EDEVICE, "device error check failed");
	return error_bits(flag);
}

hipStream_t pick_stream(se3cq_world* w, void* s) { return s ? (hipStream_t)s : w->stream; }

/* A verify step that is not positive and finite never advances the walk
 * (the reference loops forever on it); such calls are rejected. */
bool check_step(double vd) { return vd > 0.0 && std::isfinite(vd); }

/* The eager slot for a launch on st: the one st used last (stream order
 * already separates its launches), else one never used or whose last
 * launch has finished, else the next in turn, which st then waits for. */
int pick_slot(se3cq_world* w, hipStream_t st)
{
	for (int k = 0; k < QUEUE_SLOTS; k++)
		if (w->qslot[k].used && w->qslot[k].last == st)
			return k;
	for (int k = 0; k < QUEUE_SLOTS; k++)
		if (!w->qslot[k].used)
			return k;
	for (int k = 0; k < QUEUE_SLOTS; k++)
		if (hipEventQuery(w->qslot[k].ev) == hipSuccess)
			return k;
	return (int)(w->q_rr++ % QUEUE_SLOTS);
}

/*
 * Enqueue one launch on st: fn(scene, order_scratch) with the scene's work
 * counter and error word filled in.  An eager launch takes a slot (waiting
 * on the slot's last launch if that ran on another stream) and, when
 * order_bytes > 0, that slot's segment-order scratch.  A launch captured
 * into a HIP graph gets a counter word of its own for the graph's lifetime
 * and walks in input order (no scratch is allocated during a capture).
 */
template <class F>
int with_launch(se3cq_world* w, hipStream_t st, uint32_t* err, size_t order_bytes, F&& fn)
{
	hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
	if (hipStreamIsCapturing(st, &cs) != hipSuccess)
		return fail(SE3CQ_EDEVICE, "hipStreamIsCapturing failed");
	std::lock_guard<std::mutex> lk(w->qmu);
	DevScene s = base_scene(w);
	s.error_flag = err;
	if (cs != hipStreamCaptureStatusNone) {
		if (w->graph_used >= GRAPH_WORDS)
			return fail(SE3CQ_ENOMEM, "too many launches captured on this world (" + std::to_string(GRAPH_WORDS) +
			                              " work counters)");
		s.queue = (unsigned long long*)((char*)w->d_graph.p + 64 * (size_t)w->graph_used++);
		const hipError_t e = fn(s, (void*)nullptr);
		if (e != hipSuccess)
			return fail(SE3CQ_EDEVICE, std::string("launch failed: ") + hipGetErrorString(e));
		return SE3CQ_OK;
	}
	const int k = pick_slot(w, st);
	QueueSlot& q = w->qslot[k];
	if (q.used && q.last != st && hipStreamWaitEvent(st, q.ev, 0) != hipSuccess)
		return fail(SE3CQ_EDEVICE, "hipStreamWaitEvent failed");
	void* ord = nullptr;
	if (order_bytes > 0 && q.order.reserve(order_bytes))
		ord = q.order.p;
	s.queue = (unsigned long long*)((char*)w->d_queue.p + 64 * (size_t)k);
	const hipError_t e = fn(s, ord);
	if (e != hipSuccess)
		return fail(SE3CQ_EDEVICE, std::string("launch failed: ") + hipGetErrorString(e));
	if (hipEventRecord(q.ev, st) != hipSuccess)
		return fail(SE3CQ_EDEVICE, "hipEventRecord failed");
	q.last = st;
	q.used = true;
	return SE3CQ_OK;
}

/* A launch of a _device entry point: errors go to the world's word. */
template <class F>
int device_launch(se3cq_world* w, void* stream, size_t order_bytes, F&& fn)
{
	return with_launch(w, pick_stream(w, stream), (uint32_t*)w->d_err.p, order_bytes, std::forward<F>(fn));
}

/* A blocking call's context, taken from the world's pool for the call. */
class HostCall {
  public:
	explicit HostCall(se3cq_world* w) : w_(w)
	{
		std::lock_guard<std::mutex> lk(w->hmu);
		if (!w->hfree.empty()) {
			c_ = w->hfree.back();
			w->hfree.pop_back();
			return;
		}
		std::unique_ptr<HostCtx> c(new (std::nothrow) HostCtx());
		if (!c || hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
		    !c->err.reserve(sizeof(uint32_t)))
			return;
		c_ = c.get();
		w->hctx.push_back(std::move(c));
	}
	~HostCall()
	{
		if (!c_)
			return;
		std::lock_guard<std::mutex> lk(w_->hmu);
		w_->hfree.push_back(c_);
	}
	bool ok() const { return c_ != nullptr; }
	hipStream_t stream() const { return c_->st; }
	uint32_t* err() const { return (uint32_t*)c_->err.p; }
	template <typename T>
	T* buf(int slot, size_t count)
	{
		if (!c_->ws[slot].reserve(std::max<size_t>(count, 1) * sizeof(T)))
			return nullptr;
		return (T*)c_->ws[slot].p;
	}
	/* zero this call's error word (before its launches) */
	int begin()
	{
		if (hipMemsetAsync(c_->err.p, 0, sizeof(uint32_t), c_->st) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "hipMemsetAsync failed");
		return SE3CQ_OK;
	}
	template <class F>
	int launch(size_t order_bytes, F&& fn)
	{
		return with_launch(w_, c_->st, (uint32_t*)c_->err.p, order_bytes, std::forward<F>(fn));
	}
	/* wait for the call's work and report its device-side errors */
	int finish()
	{
		uint32_t flag = 0;
		const hipError_t e = hipStreamSynchronize(c_->st);
		if (e != hipSuccess)
			return fail(SE3CQ_EDEVICE, std::string("device error: ") + hipGetErrorString(e));
		if (hipMemcpy(&flag, c_->err.p, sizeof(flag), hipMemcpyDeviceToHost) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "device error check failed");
		return error_bits(flag);
	}
	int h2d(void* dst, const void* src, size_t bytes)
	{
		if (bytes && hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c_->st) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "hipMemcpyAsync failed");
		return SE3CQ_OK;
	}
	int d2h(void* dst, const void* src, size_t bytes)
	{
		if (bytes && hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c_->st) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "hipMemcpyAsync failed");
		return SE3CQ_OK;
	}
	int sync()
	{
		if (hipStreamSynchronize(c_->st) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "hipStreamSynchronize failed");
		return SE3CQ_OK;
	}

  private:
	se3cq_world* w_;
	HostCtx* c_ = nullptr;
};

#define CHECK_WORLD(w)                                                   \
	do {                                                                 \
		if (!(w))                                                        \
			return fail(SE3CQ_EINVAL, "null world");                     \
	} while (0)

#define HIPCHK(x)                                                        \
	do {                                                                 \
		hipError_t e_ = (x);                                             \
		if (e_ != hipSuccess)                                            \
			return fail(SE3CQ_EDEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
	} while (0)

void calib_rowmajor(const se3cq_world* w, double* M)
{
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			M[4 * i + j] = (double)w->xf[0][4 * j + i];
}

void unit_row(const se3cq_world* w, const double* q, double* out)
{
	/* translateToUnitState (unit_world.cc:526-539) + applyPertubation (713-724) */
	double M[16];
	calib_rowmajor(w, M);
	se3_to_unit(M, w->pert, q, out);
}

std::vector<double> unit_rows(const se3cq_world* w, const double* qs, int64_t n, int is_unit)
{
	std::vector<double> u((size_t)std::max<int64_t>(n, 1) * 7);
	for (int64_t i = 0; i < n; i++) {
		if (is_unit)
			std::memcpy(&u[7 * i], qs + 7 * i, 7 * sizeof(double));
		else
			unit_row(w, qs + 7 * i, &u[7 * i]);
	}
	return u;
}

} // namespace

/* ======================================================================= */
/* extern "C" entry points                                                 */
/* ======================================================================= */
extern "C" {

const char* se3cq_last_error(void) { return g_err.c_str(); }
const char* se3cq_version(void) { return "se3cq 0.1 gfx950"; }

int se3cq_world_create(se3cq_world** out)
{
	if (!out)
		return fail(SE3CQ_EINVAL, "null output");
	*out = new (std::nothrow) se3cq_world();
	return *out ? SE3CQ_OK : fail(SE3CQ_ENOMEM, "out of memory");
}

int se3cq_world_destroy(se3cq_world* w)
{
	delete w;
	return SE3CQ_OK;
}

int se3cq_world_copy_from(se3cq_world* dst, const se3cq_world* src)
{
	CHECK_WORLD(dst);
	CHECK_WORLD(src);
	std::lock_guard<std::mutex> lk(dst->mu);
	for (int g = 0; g < 2; g++) {
		dst->mesh[g] = src->mesh[g];
		dst->hm[g] = src->hm[g];
		dst->has_model[g] = src->has_model[g];
	}
	std::memcpy(dst->xf, src->xf, sizeof(dst->xf));
	dst->leaf_size = src->leaf_size;
	dst->pool_chunk = src->pool_chunk;
	dst->queue_depth = src->queue_depth;
	dst->scale = src->scale;
	std::memcpy(dst->pert, src->pert, sizeof(dst->pert));
	dst->cres = src->cres;
	if (dst->ready())
		return upload(dst);
	return SE3CQ_OK;
}

int se3cq_load_file(se3cq_world* w, int geo, const char* path)
{
	CHECK_WORLD(w);
	if (geo < 0 || geo > 1 || !path)
		return fail(SE3CQ_EINVAL, "bad geo or path");
	std::lock_guard<std::mutex> lk(w->mu);
	Mesh m;
	const int rc = mesh_from_obj(m, path);
	if (rc)
		return rc;
	w->mesh[geo] = std::move(m);
	w->has_model[geo] = false;
	return SE3CQ_OK;
}

int se3cq_load_mesh(se3cq_world* w, int geo, const float* V, int64_t nv, const uint32_t* F, int64_t nf)
{
	CHECK_WORLD(w);
	if (geo < 0 || geo > 1)
		return fail(SE3CQ_EINVAL, "bad geo");
	std::lock_guard<std::mutex> lk(w->mu);
	Mesh m;
	const int rc = mesh_from_arrays(m, V, nv, F, nf);
	if (rc)
		return rc;
	w->mesh[geo] = std::move(m);
	w->has_model[geo] = false;
	return SE3CQ_OK;
}

int se3cq_get_center(const se3cq_world* w, int geo, float* c3)
{
	CHECK_WORLD(w);
	if (geo < 0 || geo > 1 || !w->mesh[geo].loaded)
		return fail(SE3CQ_ESTATE, "mesh not loaded");
	std::memcpy(c3, w->mesh[geo].center, 3 * sizeof(float));
	return SE3CQ_OK;
}

int se3cq_get_ompl_center(const se3cq_world* w, int geo, float* c3)
{
	CHECK_WORLD(w);
	if (geo < 0 || geo > 1 || !w->mesh[geo].loaded)
		return fail(SE3CQ_ESTATE, "mesh not loaded");
	std::memcpy(c3, w->mesh[geo].ompl_center, 3 * sizeof(float));
	return SE3CQ_OK;
}

int se3cq_enforce_robot_center(se3cq_world* w, const double* c3)
{
	CHECK_WORLD(w);
	if (!w->mesh[1].loaded)
		return fail(SE3CQ_ESTATE, "robot not loaded");
	for (int i = 0; i < 3; i++)
		w->mesh[1].center[i] = (float)c3[i];
	return SE3CQ_OK;
}

int se3cq_scale_to_unit(se3cq_world* w)
{
	CHECK_WORLD(w);
	if (!w->mesh[0].loaded)
		return fail(SE3CQ_ESTATE, "environment not loaded");
	const float se = bbox_span(w->mesh[0].bbox);
	const float sr = w->mesh[1].loaded ? bbox_span(w->mesh[1].bbox) : 1.0f;
	w->scale = (float)(1.0 / (double)std::max(se, sr));
	return SE3CQ_OK;
}

int se3cq_angle_model(se3cq_world* w, float latitude, float longitude)
{
	CHECK_WORLD(w);
	if (latitude != 0.0f || longitude != 0.0f)
		return fail(SE3CQ_EINVAL,
		            "The arguments of UnitWorld::angleModel are reserved for compatibility reason. Both values must be 0.");
	if (!w->mesh[0].loaded)
		return fail(SE3CQ_ESTATE, "environment not loaded");
	std::lock_guard<std::mutex> lk(w->mu);
	const float s = w->scale;
	for (int g = 0; g < 2; g++) {
		float* M = w->xf[g];
		std::memset(M, 0, 16 * sizeof(float));
		M[0] = M[5] = M[10] = s;
		M[15] = 1.0f;
		if (g == 1) {
			const float* c = w->mesh[1].center;
			M[12] = s * (-c[0]);
			M[13] = s * (-c[1]);
			M[14] = s * (-c[2]);
		}
	}
	for (int g = 0; g < 2; g++) {
		w->has_model[g] = false;
		if (!w->mesh[g].loaded)
			continue;
		w->hm[g] = HostModel();
		build_model(w->hm[g], w->mesh[g], w->xf[g], w->leaf_size);
		w->has_model[g] = true;
	}
	w->release_device();
	return SE3CQ_OK;
}

int se3cq_get_scene_scale(const se3cq_world* w, float* scale)
{
	CHECK_WORLD(w);
	*scale = w->scale;
	return SE3CQ_OK;
}

int se3cq_get_matrix(const se3cq_world* w, int geo, double* m16)
{
	CHECK_WORLD(w);
	if (geo < 0 || geo > 1)
		return fail(SE3CQ_EINVAL, "bad geo");
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			m16[4 * i + j] = (double)w->xf[geo][4 * j + i];
	return SE3CQ_OK;
}

int se3cq_get_mesh_size(const se3cq_world* w, int geo, int64_t* nv, int64_t* nf)
{
	CHECK_WORLD(w);
	if (geo < 0 || geo > 1)
		return fail(SE3CQ_EINVAL, "bad geo");
	*nv = (int64_t)w->mesh[geo].V.size() / 3;
	*nf = (int64_t)w->mesh[geo].F.size() / 3;
	return SE3CQ_OK;
}

int se3cq_get_mesh(const se3cq_world* w, int geo, float* V, uint32_t* F)
{
	CHECK_WORLD(w);
	if (geo < 0 || geo > 1)
		return fail(SE3CQ_EINVAL, "bad geo");
	std::memcpy(V, w->mesh[geo].V.data(), w->mesh[geo].V.size() * sizeof(float));
	std::memcpy(F, w->mesh[geo].F.data(), w->mesh[geo].F.size() * sizeof(uint32_t));
	return SE3CQ_OK;
}

int se3cq_get_cd_vertices(const se3cq_world* w, int geo, double* V)
{
	CHECK_WORLD(w);
	if (geo < 0 || geo > 1 || !w->has_model[geo])
		return fail(SE3CQ_ESTATE, "model not built (call angle_model)");
	std::memcpy(V, w->hm[geo].Vcd.data(), w->hm[geo].Vcd.size() * sizeof(double));
	return SE3CQ_OK;
}

int se3cq_get_bvh_info(const se3cq_world* w, int geo, int32_t* depth, int64_t* nnodes)
{
	CHECK_WORLD(w);
	if (geo < 0 || geo > 1 || !w->has_model[geo])
		return fail(SE3CQ_ESTATE, "model not built (call angle_model)");
	*depth = w->hm[geo].depth;
	*nnodes = (int64_t)w->hm[geo].nodes.size();
	return SE3CQ_OK;
}

int se3cq_set_perturbation(se3cq_world* w, const double* q7)
{
	CHECK_WORLD(w);
	std::lock_guard<std::mutex> lk(w->mu);
	std::memcpy(w->pert, q7, 7 * sizeof(double));
	/* launches already issued, and graphs already captured, keep reading the
	 * block of the value they were launched with */
	if (w->uploaded)
		return select_pert(w);
	return SE3CQ_OK;
}

int se3cq_get_tuning(const se3cq_world* w, se3cq_tuning* t)
{
	CHECK_WORLD(w);
	if (!t)
		return fail(SE3CQ_EINVAL, "null output");
	t->leaf_size = w->leaf_size;
	t->pool_chunk = w->pool_chunk;
	t->queue_depth = w->queue_depth;
	return SE3CQ_OK;
}

/* Rebuild both trees with leaves of at most leaf_size triangles (call with
 * mu held, no launch of this world in flight). */
static int rebuild_trees(se3cq_world* w, int leaf_size)
{
	w->leaf_size = leaf_size;
	for (int g = 0; g < 2; g++) {
		if (!w->has_model[g])
			continue;
		w->hm[g] = HostModel();
		build_model(w->hm[g], w->mesh[g], w->xf[g], leaf_size);
	}
	if (!w->uploaded)
		return SE3CQ_OK;
	const int need = (int)(w->hm[0].depth + w->hm[1].depth + 2) * 2 * 4 * 64;
	if (need > 48 * 1024)
		return fail(SE3CQ_EINVAL, "BVH too deep for the traversal stack");
	return upload_trees(w);
}

int se3cq_set_tuning(se3cq_world* w, const se3cq_tuning* t)
{
	CHECK_WORLD(w);
	if (!t || t->leaf_size < 1 || t->leaf_size > 8 || t->pool_chunk < 16 || t->pool_chunk > 1024 ||
	    t->queue_depth < 1 || t->queue_depth > 16)
		return fail(SE3CQ_EINVAL, "bad tuning (leaf_size 1-8, pool_chunk 16-1024, queue_depth 1-16)");
	std::lock_guard<std::mutex> lk(w->mu);
	if (t->leaf_size != w->leaf_size) {
		/* launches already issued read the trees being replaced */
		if (w->uploaded && hipDeviceSynchronize() != hipSuccess)
			return fail(SE3CQ_EDEVICE, "hipDeviceSynchronize failed");
		const int rc = rebuild_trees(w, t->leaf_size);
		if (rc)
			return rc;
	}
	w->pool_chunk = t->pool_chunk;
	w->queue_depth = t->queue_depth;
	return SE3CQ_OK;
}

int se3cq_get_perturbation(const se3cq_world* w, double* q7)
{
	CHECK_WORLD(w);
	std::memcpy(q7, w->pert, 7 * sizeof(double));
	return SE3CQ_OK;
}

int se3cq_set_recommended_cres(se3cq_world* w, double cres)
{
	CHECK_WORLD(w);
	w->cres = cres;
	return SE3CQ_OK;
}

int se3cq_get_recommended_cres(const se3cq_world* w, double* cres)
{
	CHECK_WORLD(w);
	*cres = w->cres;
	return SE3CQ_OK;
}

int se3cq_check(se3cq_world* w)
{
	CHECK_WORLD(w);
	if (!w->ready() || !w->uploaded)
		return SE3CQ_OK;
	return take_device_errors(w);
}

/* ---- hot path ---- */
int se3cq_valid_states_device(se3cq_world* w, const double* d_q, int64_t n, uint8_t* d_out, void* stream)
{
	CHECK_WORLD(w);
	if (n < 0)
		return fail(SE3CQ_EINVAL, "negative count");
	if (!w->ready()) {
		if (n > 0)
			HIPCHK(launch_fill_u8(d_out, n, 1, pick_stream(w, stream)));
		return SE3CQ_OK;
	}
	const int rc0 = setup_device(w);
	if (rc0)
		return rc0;
	return device_launch(w, stream, 0, [&](const DevScene& s, void*) {
		return launch_valid_states(s, d_q, n, d_out, pick_stream(w, stream));
	});
}

int se3cq_valid_states_subset_device(se3cq_world* w, const double* d_states, int64_t n, const int32_t* d_idx,
                                     const uint32_t* d_count, uint8_t* d_out, void* stream)
{
	CHECK_WORLD(w);
	if (n < 0 || n >= ((int64_t)1 << 31))
		return fail(SE3CQ_EINVAL, "n must lie in [0, 2^31)");
	if (n > 0 && (!d_states || !d_idx || !d_count || !d_out))
		return fail(SE3CQ_EINVAL, "bad arguments");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	const int rc0 = setup_device(w);
	if (rc0 || n == 0)
		return rc0;
	return device_launch(w, stream, 0, [&](const DevScene& s, void*) {
		return launch_valid_states_subset(s, d_states, n, d_idx, d_count, d_out, pick_stream(w, stream));
	});
}

int se3cq_valid_states_host(se3cq_world* w, const double* q, int64_t n, uint8_t* out)
{
	CHECK_WORLD(w);
	if (n < 0 || (n > 0 && (!q || !out)))
		return fail(SE3CQ_EINVAL, "bad arguments");
	if (n == 0)
		return SE3CQ_OK;
	if (!w->ready()) {
		std::memset(out, 1, (size_t)n);
		return SE3CQ_OK;
	}
	const int rc0 = setup_device(w);
	if (rc0)
		return rc0;
	HostCall hc(w);
	if (!hc.ok())
		return fail(SE3CQ_EDEVICE, "cannot create a stream for the call");
	double* dq = hc.buf<double>(0, 7 * n);
	uint8_t* dout = hc.buf<uint8_t>(1, n);
	if (!dq || !dout)
		return fail(SE3CQ_ENOMEM, "device allocation failed");
	int rc = hc.begin();
	if (!rc)
		rc = hc.h2d(dq, q, 7 * n * sizeof(double));
	if (!rc)
		rc = hc.launch(0, [&](const DevScene& s, void*) { return launch_valid_states(s, dq, n, dout, hc.stream()); });
	if (!rc)
		rc = hc.finish();
	if (!rc)
		rc = hc.d2h(out, dout, (size_t)n);
	return rc ? rc : hc.sync();
}

int se3cq_contacts_host(se3cq_world* w, const double* q, int64_t n, int32_t cap, int32_t* pairs, double* segs,
                        int8_t* flags, int32_t* counts)
{
	CHECK_WORLD(w);
	if (n < 0 || cap < 0 || (n > 0 && (!q || !counts)) || (n > 0 && cap > 0 && (!pairs || !segs || !flags)))
		return fail(SE3CQ_EINVAL, "bad arguments");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	if (n == 0)
		return SE3CQ_OK;
	const int rc0 = setup_device(w);
	if (rc0)
		return rc0;
	HostCall hc(w);
	if (!hc.ok())
		return fail(SE3CQ_EDEVICE, "cannot create a stream for the call");
	const size_t slots = (size_t)n * (size_t)(cap > 0 ? cap : 1);
	double* dq = hc.buf<double>(0, 7 * n);
	int32_t* dpairs = hc.buf<int32_t>(2, 2 * slots);
	int32_t* dcount = hc.buf<int32_t>(3, (size_t)n);
	double* dsegs = hc.buf<double>(4, 6 * slots);
	int8_t* dflags = hc.buf<int8_t>(5, slots);
	if (!dq || !dpairs || !dcount || !dsegs || !dflags)
		return fail(SE3CQ_ENOMEM, "device allocation failed");
	int rc = hc.begin();
	if (!rc)
		rc = hc.h2d(dq, q, 7 * n * sizeof(double));
	if (!rc)
		rc = hc.launch(0, [&](const DevScene& s, void*) {
			return launch_contacts(s, dq, n, cap, dpairs, dcount, dsegs, dflags, hc.stream());
		});
	if (!rc)
		rc = hc.finish();
	if (rc)
		return rc;
	std::vector<int32_t> hp(cap > 0 ? 2 * slots : 0);
	std::vector<double> hs(cap > 0 ? 6 * slots : 0);
	std::vector<int8_t> hf(cap > 0 ? slots : 0);
	rc = hc.d2h(counts, dcount, (size_t)n * sizeof(int32_t));
	if (!rc && cap > 0) {
		rc = hc.d2h(hp.data(), dpairs, hp.size() * sizeof(int32_t));
		if (!rc)
			rc = hc.d2h(hs.data(), dsegs, hs.size() * sizeof(double));
		if (!rc)
			rc = hc.d2h(hf.data(), dflags, hf.size());
	}
	if (!rc)
		rc = hc.sync();
	if (rc || cap == 0)
		return rc;
	/* report the listed pairs of each state in (env, robot) order */
	std::vector<int> ord;
	for (int64_t i = 0; i < n; i++) {
		const int m = counts[i] < cap ? counts[i] : cap;
		ord.resize((size_t)m);
		for (int k = 0; k < m; k++)
			ord[(size_t)k] = k;
		const int32_t* P = &hp[(size_t)i * cap * 2];
		std::sort(ord.begin(), ord.end(), [P](int x, int y) {
			return P[2 * x] != P[2 * y] ? P[2 * x] < P[2 * y] : P[2 * x + 1] < P[2 * y + 1];
		});
		for (int k = 0; k < m; k++) {
			const size_t src = (size_t)i * cap + (size_t)ord[(size_t)k], dst = (size_t)i * cap + (size_t)k;
			pairs[2 * dst] = hp[2 * src];
			pairs[2 * dst + 1] = hp[2 * src + 1];
			for (int c = 0; c < 6; c++)
				segs[6 * dst + c] = hs[6 * src + c];
			flags[dst] = hf[src];
		}
	}
	return SE3CQ_OK;
}

static BoxPair box_pair(const se3cq_world* w)
{
	BoxPair b;
	for (int k = 0; k < 3; k++) {
		b.ce[k] = w->hm[0].bc[k];
		b.he[k] = w->hm[0].bh[k];
		b.cr[k] = w->hm[1].bc[k];
		b.hr[k] = w->hm[1].bh[k];
	}
	return b;
}

int se3cq_disentangled_device(se3cq_world* w, const double* d_q, int64_t n, uint8_t* d_out, void* stream)
{
	CHECK_WORLD(w);
	if (n < 0)
		return fail(SE3CQ_EINVAL, "negative count");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	const int rc0 = setup_device(w);
	if (rc0)
		return rc0;
	return device_launch(w, stream, 0, [&](const DevScene& s, void*) {
		return launch_disentangled(box_pair(w), s.pert, d_q, n, d_out, pick_stream(w, stream));
	});
}

int se3cq_disentangled_host(se3cq_world* w, const double* q, int64_t n, uint8_t* out)
{
	CHECK_WORLD(w);
	if (n < 0 || (n > 0 && (!q || !out)))
		return fail(SE3CQ_EINVAL, "bad arguments");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	if (n == 0)
		return SE3CQ_OK;
	const int rc0 = setup_device(w);
	if (rc0)
		return rc0;
	HostCall hc(w);
	if (!hc.ok())
		return fail(SE3CQ_EDEVICE, "cannot create a stream for the call");
	double* dq = hc.buf<double>(0, 7 * n);
	uint8_t* dout = hc.buf<uint8_t>(1, n);
	if (!dq || !dout)
		return fail(SE3CQ_ENOMEM, "device allocation failed");
	int rc = hc.begin();
	if (!rc)
		rc = hc.h2d(dq, q, 7 * n * sizeof(double));
	if (!rc)
		rc = hc.launch(0, [&](const DevScene& s, void*) {
			return launch_disentangled(box_pair(w), s.pert, dq, n, dout, hc.stream());
		});
	if (!rc)
		rc = hc.finish();
	if (!rc)
		rc = hc.d2h(out, dout, (size_t)n);
	return rc ? rc : hc.sync();
}

static int seg_device(se3cq_world* w, int mode, const double* d_from, const double* d_to, int64_t n, double vd,
                      int levels, const SegOut& o, void* stream)
{
	hipStream_t st = pick_stream(w, stream);
	return device_launch(w, stream, segment_scratch_bytes(n), [&](const DevScene& s, void* ord) {
		return launch_segments(s, mode, d_from, d_to, n, vd, levels, o, st, ord, segment_records_in(ord, n));
	});
}

/* A blocking segment call: inputs in, one segment launch, outputs out. */
struct SegHostOut {
	double* last_free;
	double* state;
	uint8_t* done;
	double* last_tau;
	double* tau;
	double* lo;
	double* hi;
	int32_t* steps;
};

static int seg_host(se3cq_world* w, int mode, const double* from, const double* to, int64_t n, double vd, int levels,
                    const SegHostOut& h)
{
	HostCall hc(w);
	if (!hc.ok())
		return fail(SE3CQ_EDEVICE, "cannot create a stream for the call");
	const size_t un = (size_t)n;
	double* df = hc.buf<double>(0, 7 * un);
	double* dt = hc.buf<double>(2, 7 * un);
	double* dlf = h.last_free ? hc.buf<double>(3, 7 * un) : nullptr;
	double* dst = h.last_free ? hc.buf<double>(4, 7 * un) : nullptr;
	uint8_t* dd = hc.buf<uint8_t>(1, un);
	double* dlt = h.last_tau ? hc.buf<double>(5, un) : nullptr;
	double* dta = h.tau ? hc.buf<double>(6, un) : nullptr;
	int32_t* dsp = h.steps ? hc.buf<int32_t>(7, un) : nullptr;
	double* dlo = mode == 2 ? hc.buf<double>(8, un) : nullptr;
	double* dhi = mode == 2 ? hc.buf<double>(9, un) : nullptr;
	if (!df || !dt || !dd || (h.last_free && (!dlf || !dst)) || (h.last_tau && !dlt) || (h.tau && !dta) ||
	    (h.steps && !dsp) || (mode == 2 && (!dlo || !dhi)))
		return fail(SE3CQ_ENOMEM, "device allocation failed");
	const SegOut o{dd, dlt, dta, dlf, dst, dlo, dhi, dsp};
	int rc = hc.begin();
	if (!rc)
		rc = hc.h2d(df, from, 7 * un * sizeof(double));
	if (!rc)
		rc = hc.h2d(dt, to, 7 * un * sizeof(double));
	if (!rc)
		rc = hc.launch(segment_scratch_bytes(n), [&](const DevScene& s, void* ord) {
			return launch_segments(s, mode, df, dt, n, vd, levels, o, hc.stream(), ord, segment_records_in(ord, n));
		});
	if (!rc)
		rc = hc.finish();
	if (!rc && h.done)
		rc = hc.d2h(h.done, dd, un);
	if (!rc && h.last_tau)
		rc = hc.d2h(h.last_tau, dlt, un * sizeof(double));
	if (!rc && h.tau)
		rc = hc.d2h(h.tau, dta, un * sizeof(double));
	if (!rc && h.lo)
		rc = hc.d2h(h.lo, dlo, un * sizeof(double));
	if (!rc && h.hi)
		rc = hc.d2h(h.hi, dhi, un * sizeof(double));
	if (!rc && h.steps)
		rc = hc.d2h(h.steps, dsp, un * sizeof(int32_t));
	if (!rc && h.last_free)
		rc = hc.d2h(h.last_free, dlf, 7 * un * sizeof(double));
	if (!rc && h.state)
		rc = hc.d2h(h.state, dst, 7 * un * sizeof(double));
	return rc ? rc : hc.sync();
}

int se3cq_transit_device(se3cq_world* w, const double* d_from, const double* d_to, int64_t n, double verify_delta,
                         double* d_last_free, double* d_state, uint8_t* d_done, double* d_last_tau, double* d_tau,
                         int32_t* d_steps, void* stream)
{
	CHECK_WORLD(w);
	if (!check_step(verify_delta))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	if ((d_last_free == nullptr) != (d_state == nullptr))
		return fail(SE3CQ_EINVAL, "last_free and state must both be given or both be null");
	if (n < 0)
		return fail(SE3CQ_EINVAL, "negative count");
	const int rc0 = setup_device(w);
	if (rc0)
		return rc0;
	SegOut o{d_done, d_last_tau, d_tau, d_last_free, d_state, nullptr, nullptr, d_steps};
	return seg_device(w, 0, d_from, d_to, n, verify_delta, 0, o, stream);
}

int se3cq_transit_host(se3cq_world* w, const double* from, const double* to, int64_t n, double verify_delta,
                       double* last_free, double* state, uint8_t* done, double* last_tau, double* tau,
                       int32_t* steps)
{
	CHECK_WORLD(w);
	if (!check_step(verify_delta))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	{
		std::lock_guard<std::mutex> lk0(w->mu);
		const int rc0 = ensure_device(w);
		if (rc0)
			return rc0;
	}
	if (n <= 0)
		return n == 0 ? SE3CQ_OK : fail(SE3CQ_EINVAL, "negative count");
	if ((last_free == nullptr) != (state == nullptr))
		return fail(SE3CQ_EINVAL, "last_free and state must both be given or both be null");
	return seg_host(w, 0, from, to, n, verify_delta, 0,
	                SegHostOut{last_free, state, done, last_tau, tau, nullptr, nullptr, steps});
}

int se3cq_valid_transitions_host(se3cq_world* w, const double* from, const double* to, int64_t n,
                                 double initial_verify_delta, uint8_t* out)
{
	CHECK_WORLD(w);
	if (!check_step(initial_verify_delta))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	{
		std::lock_guard<std::mutex> lk0(w->mu);
		const int rc0 = ensure_device(w);
		if (rc0)
			return rc0;
	}
	if (n <= 0)
		return n == 0 ? SE3CQ_OK : fail(SE3CQ_EINVAL, "negative count");
	return seg_host(w, 1, from, to, n, initial_verify_delta, 0,
	                SegHostOut{nullptr, nullptr, out, nullptr, nullptr, nullptr, nullptr, nullptr});
}

int se3cq_valid_transitions_device(se3cq_world* w, const double* d_from, const double* d_to, int64_t n,
                                   double initial_verify_delta, uint8_t* d_out, int32_t* d_steps, void* stream)
{
	CHECK_WORLD(w);
	if (!check_step(initial_verify_delta))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	if (n < 0)
		return fail(SE3CQ_EINVAL, "negative count");
	const int rc0 = setup_device(w);
	if (rc0)
		return rc0;
	SegOut o{d_out, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, d_steps};
	return seg_device(w, 1, d_from, d_to, n, initial_verify_delta, 0, o, stream);
}

int se3cq_valid_transitions_subset_device(se3cq_world* w, const double* d_from, const double* d_to, int64_t n,
                                          const int32_t* d_idx, const uint32_t* d_count, double initial_verify_delta,
                                          uint8_t* d_out, int32_t* d_steps, void* stream)
{
	CHECK_WORLD(w);
	if (!check_step(initial_verify_delta))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (n < 0 || n >= ((int64_t)1 << 31))
		return fail(SE3CQ_EINVAL, "n must lie in [0, 2^31)");
	if (n > 0 && (!d_from || !d_to || !d_idx || !d_count || !d_out))
		return fail(SE3CQ_EINVAL, "bad arguments");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	const int rc0 = setup_device(w);
	if (rc0 || n == 0)
		return rc0;
	hipStream_t st = pick_stream(w, stream);
	const SegOut o{d_out, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, d_steps};
	return device_launch(w, stream, 0, [&](const DevScene& s0, void*) {
		DevScene s = s0;
		s.order = d_idx;
		s.n_items = d_count;
		return launch_segments(s, 1, d_from, d_to, n, initial_verify_delta, 0, o, st, nullptr);
	});
}

int se3cq_touch_search_device(se3cq_world* w, const double* d_from, const double* d_to, int64_t n,
                              double verify_delta, int levels, uint8_t* d_done, double* d_last_tau, double* d_tau,
                              double* d_lo, double* d_hi, int32_t* d_steps, void* stream)
{
	CHECK_WORLD(w);
	if (!check_step(verify_delta))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	if (levels < 0 || levels > 60)
		return fail(SE3CQ_EINVAL, "levels must be in [0, 60]");
	if (!d_lo || !d_hi)
		return fail(SE3CQ_EINVAL, "lo and hi outputs are required");
	if (n < 0)
		return fail(SE3CQ_EINVAL, "negative count");
	const int rc0 = setup_device(w);
	if (rc0)
		return rc0;
	SegOut o{d_done, d_last_tau, d_tau, nullptr, nullptr, d_lo, d_hi, d_steps};
	return seg_device(w, 2, d_from, d_to, n, verify_delta, levels, o, stream);
}

int se3cq_touch_search_device_noalloc(se3cq_world* w, const double* d_from, const double* d_to, int64_t n,
                                      double verify_delta, int levels, uint8_t* d_done, double* d_last_tau,
                                      double* d_tau, double* d_lo, double* d_hi, int32_t* d_steps, void* stream)
{
	/* Nothing here allocates: the work counter and the segment order live
	 * in the library's static per-device scratch array, and the world's
	 * buffers (and its error word) exist since its upload. */
	CHECK_WORLD(w);
	if (!check_step(verify_delta))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (levels < 0 || levels > 60)
		return fail(SE3CQ_EINVAL, "levels must be in [0, 60]");
	if (!d_lo || !d_hi || !d_done)
		return fail(SE3CQ_EINVAL, "done, lo and hi outputs are required");
	if (n < 0)
		return fail(SE3CQ_EINVAL, "negative count");
	if (n > SE3CQ_NOALLOC_MAX_SEGMENTS)
		return fail(SE3CQ_EINVAL, "the allocation-free touch search takes at most " +
		                              std::to_string(SE3CQ_NOALLOC_MAX_SEGMENTS) + " segments per call (got " +
		                              std::to_string(n) + ")");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	if (!w->uploaded)
		return fail(SE3CQ_ESTATE, "the world is not on the device yet; any other query on it uploads it "
		                          "(the allocation-free call does not)");
	int dev = -1;
	if (hipGetDevice(&dev) != hipSuccess || dev != w->device)
		return fail(SE3CQ_EINVAL, "the current device is not the world's device");
	if (n == 0)
		return SE3CQ_OK;
	void* scr = nullptr;
	const hipError_t se = noalloc_scratch(&scr);
	if (se != hipSuccess)
		return fail(SE3CQ_EDEVICE, std::string("scratch lookup failed: ") + hipGetErrorString(se));
	DevScene s;
	{
		std::lock_guard<std::mutex> lk(w->qmu);
		s = base_scene(w);
	}
	s.error_flag = (uint32_t*)w->d_err.p;
	s.queue = (unsigned long long*)scr;
	const SegOut o{d_done, d_last_tau, d_tau, nullptr, nullptr, d_lo, d_hi, d_steps};
	const hipError_t e = launch_segments(s, 2, d_from, d_to, n, verify_delta, levels, o, pick_stream(w, stream),
	                                     (char*)scr + NOALLOC_COUNTER_BYTES);
	if (e != hipSuccess)
		return fail(SE3CQ_EDEVICE, std::string("launch failed: ") + hipGetErrorString(e));
	return SE3CQ_OK;
}

int se3cq_touch_search_host(se3cq_world* w, const double* from, const double* to, int64_t n, double verify_delta,
                            int levels, uint8_t* done, double* last_tau, double* tau, double* lo, double* hi,
                            int32_t* steps)
{
	CHECK_WORLD(w);
	if (!check_step(verify_delta))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	{
		std::lock_guard<std::mutex> lk0(w->mu);
		const int rc0 = ensure_device(w);
		if (rc0)
			return rc0;
	}
	if (levels < 0 || levels > 60)
		return fail(SE3CQ_EINVAL, "levels must be in [0, 60]");
	if (n <= 0)
		return n == 0 ? SE3CQ_OK : fail(SE3CQ_EINVAL, "negative count");
	return seg_host(w, 2, from, to, n, verify_delta, levels,
	                SegHostOut{nullptr, nullptr, done, last_tau, tau, lo, hi, steps});
}

int se3cq_bisect_contact_host(se3cq_world* w, const double* from, const double* to, int64_t n, const double* lo_in,
                              const double* hi_in, int levels, double* lo, double* hi)
{
	/* Bisection on an explicit bracket: evaluate every probe as a batch of
	 * states per level (n states per launch, `levels` launches). */
	CHECK_WORLD(w);
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	{
		std::lock_guard<std::mutex> lk0(w->mu);
		const int rc0 = ensure_device(w);
		if (rc0)
			return rc0;
	}
	if (n <= 0)
		return n == 0 ? SE3CQ_OK : fail(SE3CQ_EINVAL, "negative count");
	std::vector<double> l(lo_in, lo_in + n), h(hi_in, hi_in + n), mids(7 * n);
	std::vector<uint8_t> ok(n);
	for (int lv = 0; lv < levels; lv++) {
		for (int64_t i = 0; i < n; i++) {
			se3_segment sg;
			se3_segment_init(&sg, from + 7 * i, to + 7 * i);
			se3_segment_at(&sg, 0.5 * (l[i] + h[i]), &mids[7 * i]);
		}
		int rc = se3cq_valid_states_host(w, mids.data(), n, ok.data());
		if (rc)
			return rc;
		for (int64_t i = 0; i < n; i++) {
			const double m = 0.5 * (l[i] + h[i]);
			if (ok[i])
				l[i] = m;
			else
				h[i] = m;
		}
	}
	std::memcpy(lo, l.data(), n * sizeof(double));
	std::memcpy(hi, h.data(), n * sizeof(double));
	return SE3CQ_OK;
}

static int visibility_host(se3cq_world* w, int vis, const std::vector<double>& u0, const std::vector<double>& u1,
                           int64_t m, int64_t n, double vd, int8_t* out, size_t out_n)
{
	HostCall hc(w);
	if (!hc.ok())
		return fail(SE3CQ_EDEVICE, "cannot create a stream for the call");
	double* d0 = hc.buf<double>(0, u0.size());
	double* d1 = hc.buf<double>(2, u1.size());
	int8_t* dout = hc.buf<int8_t>(1, out_n);
	if (!d0 || !d1 || !dout)
		return fail(SE3CQ_ENOMEM, "device allocation failed");
	int rc = hc.begin();
	if (!rc)
		rc = hc.h2d(d0, u0.data(), u0.size() * sizeof(double));
	if (!rc)
		rc = hc.h2d(d1, u1.data(), u1.size() * sizeof(double));
	if (!rc && launch_fill_u8((uint8_t*)dout, (int64_t)out_n, 0, hc.stream()) != hipSuccess)
		rc = fail(SE3CQ_EDEVICE, "fill failed");
	if (!rc)
		rc = hc.launch(0, [&](const DevScene& s, void*) {
			return launch_visibility(s, vis, d0, d1, m, n, vd, dout, nullptr, nullptr, hc.stream());
		});
	if (!rc)
		rc = hc.finish();
	if (!rc)
		rc = hc.d2h(out, dout, out_n);
	return rc ? rc : hc.sync();
}

int se3cq_visibility_matrix_host(se3cq_world* w, const double* qs, int64_t n, int qs_are_unit,
                                 double verify_magnitude, int8_t* out)
{
	CHECK_WORLD(w);
	if (!check_step(verify_magnitude))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	{
		std::lock_guard<std::mutex> lk0(w->mu);
		const int rc0 = ensure_device(w);
		if (rc0)
			return rc0;
	}
	if (n <= 0)
		return n == 0 ? SE3CQ_OK : fail(SE3CQ_EINVAL, "negative count");
	std::vector<double> u = unit_rows(w, qs, n, qs_are_unit);
	return visibility_host(w, 1, u, u, n, n, verify_magnitude, out, (size_t)(n * n));
}

int se3cq_visibility_matrix2_host(se3cq_world* w, const double* qs0, int64_t m, int qs0_are_unit, const double* qs1,
                                  int64_t n, int qs1_are_unit, double verify_magnitude, int8_t* out)
{
	CHECK_WORLD(w);
	if (!check_step(verify_magnitude))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	{
		std::lock_guard<std::mutex> lk0(w->mu);
		const int rc0 = ensure_device(w);
		if (rc0)
			return rc0;
	}
	if (m < 0 || n < 0)
		return fail(SE3CQ_EINVAL, "negative count");
	if (m == 0 || n == 0)
		return SE3CQ_OK;
	std::vector<double> u0 = unit_rows(w, qs0, m, qs0_are_unit);
	std::vector<double> u1 = unit_rows(w, qs1, n, qs1_are_unit);
	return visibility_host(w, 2, u0, u1, m, n, verify_magnitude, out, (size_t)(m * n));
}

int se3cq_visibility_pair_host(se3cq_world* w, const double* qs0, int qs0_are_unit, const double* qs1,
                               int qs1_are_unit, int64_t n, double verify_magnitude, int8_t* out)
{
	CHECK_WORLD(w);
	if (!check_step(verify_magnitude))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	{
		std::lock_guard<std::mutex> lk0(w->mu);
		const int rc0 = ensure_device(w);
		if (rc0)
			return rc0;
	}
	if (n <= 0)
		return n == 0 ? SE3CQ_OK : fail(SE3CQ_EINVAL, "negative count");
	std::vector<double> u0 = unit_rows(w, qs0, n, qs0_are_unit);
	std::vector<double> u1 = unit_rows(w, qs1, n, qs1_are_unit);
	return visibility_host(w, 0, u0, u1, n, n, verify_magnitude, out, (size_t)n);
}

int se3cq_visibility_bits_device(se3cq_world* w, const double* d_qs, int64_t n, double verify_magnitude,
                                 uint32_t* d_bits, int32_t* d_steps, void* stream)
{
	CHECK_WORLD(w);
	if (!check_step(verify_magnitude))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	{
		std::lock_guard<std::mutex> lk0(w->mu);
		const int rc0 = ensure_device(w);
		if (rc0)
			return rc0;
	}
	if (n < 0)
		return fail(SE3CQ_EINVAL, "negative count");
	hipStream_t st = pick_stream(w, stream);
	const size_t words = (size_t)((n * n + 31) / 32);
	if (words)
		HIPCHK(launch_fill_u32(d_bits, (int64_t)words, 0u, st));
	return device_launch(w, stream, 0, [&](const DevScene& s, void*) {
		return launch_visibility(s, 1, d_qs, d_qs, n, n, verify_magnitude, nullptr, d_bits, d_steps, st);
	});
}

/* ---- state arithmetic ---- */
int se3cq_translate_to_unit(const se3cq_world* w, const double* q, int64_t n, double* out)
{
	CHECK_WORLD(w);
	for (int64_t i = 0; i < n; i++)
		unit_row(w, q + 7 * i, out + 7 * i);
	return SE3CQ_OK;
}

int se3cq_translate_from_unit(const se3cq_world* w, const double* q, int64_t n, double* out)
{
	/* unapplyPertubation (unit_world.cc:726-738) then the inverse calibration
	 * (diag(1/s)) on the translation (541-555). */
	CHECK_WORLD(w);
	for (int64_t i = 0; i < n; i++) {
		const double* s = q + 7 * i;
		double rot[4] = {s[3], s[4], s[5], s[6]};
		se3_quat_normalize(rot);
		double prot[4] = {w->pert[3], w->pert[4], w->pert[5], w->pert[6]};
		se3_quat_normalize(prot);
		double pinv[4] = {prot[0], -prot[1], -prot[2], -prot[3]};
		double R[9];
		se3_quat_to_rotmat(pinv[0], pinv[1], pinv[2], pinv[3], R);
		const double d[3] = {-w->pert[0] + s[0], -w->pert[1] + s[1], -w->pert[2] + s[2]};
		double t[3];
		for (int r = 0; r < 3; r++)
			t[r] = (R[3 * r] * d[0] + R[3 * r + 1] * d[1]) + R[3 * r + 2] * d[2];
		double qr[4];
		se3_quat_mul(pinv, rot, qr);
		se3_quat_normalize(qr);
		const double inv = 1.0 / (double)w->scale;
		double* o = out + 7 * i;
		for (int r = 0; r < 3; r++)
			o[r] = t[r] * inv;
		for (int k = 0; k < 4; k++)
			o[3 + k] = qr[k];
	}
	return SE3CQ_OK;
}

namespace {
/* (OMPL centre - centre) * scene scale: the robot-frame offset between the
 * OMPL convention and the unit world when enforceRobotCenter moved the
 * centre (unit_world.cc:564-565, 702-703); glm float difference, then double. */
void ompl_delta(const se3cq_world* w, double* d)
{
	for (int k = 0; k < 3; k++)
		d[k] = (double)(w->mesh[1].ompl_center[k] - w->mesh[1].center[k]) * (double)w->scale;
}
} // namespace

int se3cq_translate_ompl_to_unit(const se3cq_world* w, const double* q, int64_t n, double* out)
{
	/* translateOMPLStateToUnitState (unit_world.cc:689-711) */
	CHECK_WORLD(w);
	if (!w->mesh[1].loaded)
		return fail(SE3CQ_ESTATE, "robot not loaded");
	double d[3], M[16];
	ompl_delta(w, d);
	calib_rowmajor(w, M);
	for (int64_t i = 0; i < n; i++)
		se3_ompl_to_unit(M, w->pert, d, q + 7 * i, out + 7 * i);
	return SE3CQ_OK;
}

int se3cq_translate_unit_to_ompl(const se3cq_world* w, const double* q, int64_t n, int to_angle_axis, double* out)
{
	/* translateUnitStateToOMPLState (unit_world.cc:558-605) */
	CHECK_WORLD(w);
	if (!w->mesh[1].loaded)
		return fail(SE3CQ_ESTATE, "robot not loaded");
	double d[3];
	ompl_delta(w, d);
	for (int64_t i = 0; i < n; i++) {
		const double* s = q + 7 * i;
		double rot[4] = {s[3], s[4], s[5], s[6]}, rd[3], c[7], f[7];
		se3_quat_normalize(rot);
		se3_quat_apply(rot, d, rd);
		for (int k = 0; k < 3; k++)
			c[k] = s[k] + rd[k];
		se3_quat_normalize(rot);
		for (int k = 0; k < 4; k++)
			c[3 + k] = rot[k];
		se3cq_translate_from_unit(w, c, 1, f);
		double* o = out + 7 * i;
		for (int k = 0; k < 3; k++)
			o[k] = f[k];
		if (to_angle_axis) {
			/* Eigen (3.3) AngleAxis from a unit quaternion: angle
			 * 2 atan2(|v|, |w|), axis v / |v| with the sign of w */
			const double vx = f[4], vy = f[5], vz = f[6];
			double nv = std::sqrt((vx * vx + vy * vy) + vz * vz);
			if (nv != 0.0) {
				o[3] = 2.0 * std::atan2(nv, std::fabs(f[3]));
				if (f[3] < 0.0)
					nv = -nv;
				o[4] = vx / nv;
				o[5] = vy / nv;
				o[6] = vz / nv;
			} else {
				o[3] = 0.0;
				o[4] = 1.0;
				o[5] = 0.0;
				o[6] = 0.0;
			}
		} else {
			/* w first -> w last */
			o[3] = f[4];
			o[4] = f[5];
			o[5] = f[6];
			o[6] = f[3];
		}
	}
	return SE3CQ_OK;
}

static int vanilla_ompl(const se3cq_world* w, const double* q, int64_t n, double sign, double* out)
{
	if (!w->mesh[1].loaded)
		return fail(SE3CQ_ESTATE, "robot not loaded");
	const double oc[3] = {w->mesh[1].ompl_center[0], w->mesh[1].ompl_center[1], w->mesh[1].ompl_center[2]};
	for (int64_t i = 0; i < n; i++) {
		const double* s = q + 7 * i;
		double rot[4] = {s[6], s[3], s[4], s[5]}, roo[3];
		se3_quat_normalize(rot);
		se3_quat_apply(rot, oc, roo);
		double* o = out + 7 * i;
		for (int k = 0; k < 3; k++)
			o[k] = sign > 0 ? s[k] + roo[k] : s[k] - roo[k];
		for (int k = 3; k < 7; k++)
			o[k] = s[k];
	}
	return SE3CQ_OK;
}

int se3cq_translate_vanilla_to_ompl(const se3cq_world* w, const double* q, int64_t n, double* out)
{
	/* translateVanillaStateToOMPLState (unit_world.cc:606-629) */
	CHECK_WORLD(w);
	return vanilla_ompl(w, q, n, 1.0, out);
}

int se3cq_translate_ompl_to_vanilla(const se3cq_world* w, const double* q, int64_t n, double* out)
{
	/* translateOMPLStateToVanillaState (unit_world.cc:639-660) */
	CHECK_WORLD(w);
	return vanilla_ompl(w, q, n, -1.0, out);
}

int se3cq_translate_vanilla_to_unit(const se3cq_world* w, const double* q, int64_t n, double* out)
{
	/* translateVanillaStateToUnitState (unit_world.cc:662-666) */
	CHECK_WORLD(w);
	std::vector<double> o((size_t)std::max<int64_t>(n, 1) * 7);
	const int rc = vanilla_ompl(w, q, n, 1.0, o.data());
	if (rc)
		return rc;
	return se3cq_translate_ompl_to_unit(w, o.data(), n, out);
}

int se3cq_translate_vanilla_pts_to_unit(const se3cq_world* w, int geo, const double* pts, int64_t n, double* out)
{
	/* translateVanillaPointsToUnitPoints (unit_world.cc:668-687): robot points
	 * relative to the OMPL centre, then the calibration matrix */
	CHECK_WORLD(w);
	if (geo < 0 || geo > 1 || !w->mesh[geo].loaded)
		return fail(SE3CQ_ESTATE, "mesh not loaded");
	double M[16];
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			M[4 * i + j] = (double)w->xf[0][4 * j + i];
	for (int64_t i = 0; i < n; i++) {
		double p[3];
		for (int k = 0; k < 3; k++)
			p[k] = geo == 0 ? pts[3 * i + k] : pts[3 * i + k] - (double)w->mesh[1].ompl_center[k];
		for (int r = 0; r < 3; r++)
			out[3 * i + r] = ((M[4 * r] * p[0] + M[4 * r + 1] * p[1]) + M[4 * r + 2] * p[2]) + M[4 * r + 3];
	}
	return SE3CQ_OK;
}

/* ---- mixed batches (persistent kernel) ---- */
static int batch_shape(se3cq_world* w, const DevScene& s, const se3cq_batch_config& c, bool states, int* block,
                       int* grid)
{
	const int bmax = batch_default_block(s, states);
	int b = c.block_waves > 0 ? 64 * c.block_waves : bmax;
	if (b > bmax)
		b = bmax;
	const int per_cu_max = batch_blocks_per_cu(s, b, states);
	const int per_cu = (c.blocks_per_cu > 0 && c.blocks_per_cu < per_cu_max) ? c.blocks_per_cu : per_cu_max;
	*block = b;
	*grid = w->num_cu * per_cu;
	return SE3CQ_OK;
}

static int batch_check(se3cq_world* w, int64_t n_seg, double vd, const double* d_last_free, const double* d_state,
                       const uint8_t* d_done, int64_t n_states, const uint8_t* d_valid)
{
	if (n_seg < 0 || n_states < 0)
		return fail(SE3CQ_EINVAL, "negative count");
	if (n_seg > 0 && !check_step(vd))
		return fail(SE3CQ_EINVAL, "the verify step must be positive and finite");
	if ((d_last_free == nullptr) != (d_state == nullptr))
		return fail(SE3CQ_EINVAL, "last_free and state must both be given or both be null");
	if ((n_seg > 0 && !d_done) || (n_states > 0 && !d_valid))
		return fail(SE3CQ_EINVAL, "missing output");
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	return setup_device(w);
}

static BatchArgs batch_args(se3cq_world* w, const double* d_seg_from, const double* d_seg_to, int64_t n_seg,
                            double vd, double* d_last_free, double* d_state, uint8_t* d_done, double* d_last_tau,
                            double* d_tau, int32_t* d_steps, const double* d_states, int64_t n_states,
                            uint8_t* d_valid, int states_per_lane)
{
	BatchArgs a;
	a.seg_from = d_seg_from;
	a.seg_to = d_seg_to;
	a.n_seg = n_seg;
	a.vd = vd;
	a.seg = SegOut{d_done, d_last_tau, d_tau, d_last_free, d_state, nullptr, nullptr, d_steps};
	a.states = d_states;
	a.n_states = n_states;
	a.valid = d_valid;
	a.states_per_lane = states_per_lane > 0 ? states_per_lane : 1;
	return a;
}

/* One batch on st with error word err (the world's for _device calls). */
static int batch_launch(se3cq_world* w, const BatchArgs& a, const se3cq_batch_config& c, hipStream_t st,
                        uint32_t* err)
{
	const bool split = a.n_states == 0 || c.split;
	return with_launch(w, st, err, split ? segment_scratch_bytes(a.n_seg) : 0, [&](const DevScene& s, void* ord) {
		if (split) {
			/* the persistent segment kernel (no traversal stacks in LDS),
			 * then the state kernel */
			hipError_t e = launch_segments(s, 0, a.seg_from, a.seg_to, a.n_seg, a.vd, 0, a.seg, st, ord,
			                               segment_records_in(ord, a.n_seg));
			if (e == hipSuccess && a.n_states > 0)
				e = launch_valid_states(s, a.states, a.n_states, a.valid, st);
			return e;
		}
		int block = 0, grid = 0;
		batch_shape(w, s, c, a.n_states > 0, &block, &grid);
		return launch_batch(s, a, block, grid, st);
	});
}

int se3cq_batch_device(se3cq_world* w, const double* d_seg_from, const double* d_seg_to, int64_t n_seg,
                       double verify_delta, double* d_last_free, double* d_state, uint8_t* d_done, double* d_last_tau,
                       double* d_tau, int32_t* d_steps, const double* d_states, int64_t n_states, uint8_t* d_valid,
                       void* stream)
{
	CHECK_WORLD(w);
	const int rc = batch_check(w, n_seg, verify_delta, d_last_free, d_state, d_done, n_states, d_valid);
	if (rc)
		return rc;
	const BatchArgs a = batch_args(w, d_seg_from, d_seg_to, n_seg, verify_delta, d_last_free, d_state, d_done,
	                               d_last_tau, d_tau, d_steps, d_states, n_states, d_valid, w->batch_cfg.states_per_lane);
	return batch_launch(w, a, w->batch_cfg, pick_stream(w, stream), (uint32_t*)w->d_err.p);
}

int se3cq_batch_host(se3cq_world* w, const double* seg_from, const double* seg_to, int64_t n_seg, double verify_delta,
                     double* last_free, double* state, uint8_t* done, double* last_tau, double* tau, int32_t* steps,
                     const double* states, int64_t n_states, uint8_t* valid)
{
	CHECK_WORLD(w);
	if ((n_seg > 0 && (!seg_from || !seg_to)) || (n_states > 0 && !states))
		return fail(SE3CQ_EINVAL, "missing input");
	const int rc0 = batch_check(w, n_seg, verify_delta, last_free, state, done, n_states, valid);
	if (rc0 || (n_seg == 0 && n_states == 0))
		return rc0;
	HostCall hc(w);
	if (!hc.ok())
		return fail(SE3CQ_EDEVICE, "cannot create a stream for the call");
	const size_t ns = (size_t)n_seg, nq = (size_t)n_states;
	double* df = hc.buf<double>(0, 7 * ns);
	double* dt = hc.buf<double>(2, 7 * ns);
	double* dlf = last_free ? hc.buf<double>(3, 7 * ns) : nullptr;
	double* dst = last_free ? hc.buf<double>(4, 7 * ns) : nullptr;
	uint8_t* dd = hc.buf<uint8_t>(1, ns);
	double* dlt = hc.buf<double>(5, ns);
	double* dta = hc.buf<double>(6, ns);
	int32_t* dsp = hc.buf<int32_t>(7, ns);
	double* dq = hc.buf<double>(8, 7 * nq);
	uint8_t* dv = hc.buf<uint8_t>(9, nq);
	if (!df || !dt || (last_free && (!dlf || !dst)) || !dd || !dlt || !dta || !dsp || !dq || !dv)
		return fail(SE3CQ_ENOMEM, "device allocation failed");
	int rc = hc.begin();
	if (!rc)
		rc = hc.h2d(df, seg_from, 7 * ns * sizeof(double));
	if (!rc)
		rc = hc.h2d(dt, seg_to, 7 * ns * sizeof(double));
	if (!rc)
		rc = hc.h2d(dq, states, 7 * nq * sizeof(double));
	se3cq_batch_config cfg;
	{
		std::lock_guard<std::mutex> lk(w->mu);
		cfg = w->batch_cfg;
	}
	const BatchArgs a = batch_args(w, df, dt, n_seg, verify_delta, dlf, dst, dd, dlt, dta, dsp, dq, n_states, dv,
	                               cfg.states_per_lane);
	if (!rc)
		rc = batch_launch(w, a, cfg, hc.stream(), hc.err());
	if (!rc)
		rc = hc.finish();
	if (!rc && ns) {
		rc = hc.d2h(done, dd, ns);
		if (!rc && last_tau)
			rc = hc.d2h(last_tau, dlt, ns * sizeof(double));
		if (!rc && tau)
			rc = hc.d2h(tau, dta, ns * sizeof(double));
		if (!rc && steps)
			rc = hc.d2h(steps, dsp, ns * sizeof(int32_t));
		if (!rc && last_free)
			rc = hc.d2h(last_free, dlf, 7 * ns * sizeof(double));
		if (!rc && last_free)
			rc = hc.d2h(state, dst, 7 * ns * sizeof(double));
	}
	if (!rc && nq)
		rc = hc.d2h(valid, dv, nq);
	return rc ? rc : hc.sync();
}

int se3cq_batch_get_config(const se3cq_world* w, se3cq_batch_config* cfg)
{
	CHECK_WORLD(w);
	if (!cfg)
		return fail(SE3CQ_EINVAL, "bad arguments");
	*cfg = w->batch_cfg;
	return SE3CQ_OK;
}

int se3cq_batch_set_config(se3cq_world* w, const se3cq_batch_config* cfg)
{
	CHECK_WORLD(w);
	if (!cfg || cfg->block_waves < 0 || cfg->block_waves > 4 || cfg->blocks_per_cu < 0 || cfg->states_per_lane < 1 ||
	    cfg->states_per_lane > 64 || cfg->split < 0 || cfg->split > 1)
		return fail(SE3CQ_EINVAL, "bad batch configuration");
	w->batch_cfg = *cfg;
	return SE3CQ_OK;
}

int se3cq_batch_tune(se3cq_world* w, const double* d_seg_from, const double* d_seg_to, int64_t n_seg,
                     double verify_delta, double* d_last_free, double* d_state, uint8_t* d_done, double* d_last_tau,
                     double* d_tau, int32_t* d_steps, const double* d_states, int64_t n_states, uint8_t* d_valid,
                     int32_t repeats, se3cq_batch_config* chosen, float* ms)
{
	/* A coordinate search: leaf size, then launch shape, then pool chunk,
	 * then queue depth, each at the best values found so far.  Every
	 * candidate's outputs are compared with the first run's on the device. */
	CHECK_WORLD(w);
	const int rc0 = batch_check(w, n_seg, verify_delta, d_last_free, d_state, d_done, n_states, d_valid);
	if (rc0)
		return rc0;
	if (repeats < 1 || repeats > 100)
		return fail(SE3CQ_EINVAL, "repeats must be in [1, 100]");
	if (n_states == 0)
		return fail(SE3CQ_EINVAL, "a batch without states has one launch shape (the segment kernel's); nothing to tune");
	const size_t ns = (size_t)n_seg;
	struct Out {
		void* p;
		size_t n;
	};
	std::vector<Out> outs;
	outs.push_back({d_done, ns});
	if (d_last_tau)
		outs.push_back({d_last_tau, ns * sizeof(double)});
	if (d_tau)
		outs.push_back({d_tau, ns * sizeof(double)});
	if (d_steps)
		outs.push_back({d_steps, ns * sizeof(int32_t)});
	if (d_last_free)
		outs.push_back({d_last_free, 7 * ns * sizeof(double)});
	if (d_state)
		outs.push_back({d_state, 7 * ns * sizeof(double)});
	outs.push_back({d_valid, (size_t)n_states});
	std::vector<DevBuf> ref(outs.size());
	DevBuf flag;
	if (!flag.reserve(sizeof(uint32_t)))
		return fail(SE3CQ_ENOMEM, "device allocation failed");
	for (size_t k = 0; k < outs.size(); k++)
		if (outs[k].n && !ref[k].reserve(outs[k].n))
			return fail(SE3CQ_ENOMEM, "device allocation failed");
	hipStream_t st = w->stream;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
		if (e0)
			(void)hipEventDestroy(e0);
		return fail(SE3CQ_EDEVICE, "hipEventCreate failed");
	}
	/* one launch untimed, then `reps` timed */
	auto run = [&](const se3cq_batch_config& c, int reps, float* t) -> int {
		const BatchArgs a = batch_args(w, d_seg_from, d_seg_to, n_seg, verify_delta, d_last_free, d_state, d_done,
		                               d_last_tau, d_tau, d_steps, d_states, n_states, d_valid, c.states_per_lane);
		int rc = batch_launch(w, a, c, st, (uint32_t*)w->d_err.p);
		if (rc)
			return rc;
		if (hipEventRecord(e0, st) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "hipEventRecord failed");
		for (int r = 0; r < reps && !rc; r++)
			rc = batch_launch(w, a, c, st, (uint32_t*)w->d_err.p);
		if (rc)
			return rc;
		if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
		    hipEventElapsedTime(t, e0, e1) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "timing failed");
		*t /= (float)reps;
		return SE3CQ_OK;
	};
	auto same = [&]() -> int {
		uint32_t h = 0;
		if (launch_fill_u32((uint32_t*)flag.p, 1, 0u, st) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "fill failed");
		for (size_t k = 0; k < outs.size(); k++)
			if (launch_differ(outs[k].p, ref[k].p, (int64_t)outs[k].n, (uint32_t*)flag.p, st) != hipSuccess)
				return fail(SE3CQ_EDEVICE, "compare failed");
		if (hipMemcpyAsync(&h, flag.p, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess ||
		    hipStreamSynchronize(st) != hipSuccess)
			return fail(SE3CQ_EDEVICE, "compare failed");
		return h ? fail(SE3CQ_EDEVICE, "two tuning candidates computed different results") : SE3CQ_OK;
	};
	se3cq_batch_config best_cfg = w->batch_cfg;
	se3cq_tuning best_t;
	se3cq_get_tuning(w, &best_t);
	float best = 0.0f;
	int rc = run(best_cfg, repeats, &best);
	for (size_t k = 0; k < outs.size() && !rc; k++)
		if (outs[k].n && hipMemcpyAsync(ref[k].p, outs[k].p, outs[k].n, hipMemcpyDeviceToDevice, st) != hipSuccess)
			rc = fail(SE3CQ_EDEVICE, "copy failed");
	if (!rc && hipStreamSynchronize(st) != hipSuccess)
		rc = fail(SE3CQ_EDEVICE, "copy failed");
	/* try one candidate: the current best with c and t; keep it if faster */
	auto trial = [&](const se3cq_batch_config& c, const se3cq_tuning& t) -> int {
		int r = se3cq_set_tuning(w, &t);
		float ti = 0.0f;
		if (!r)
			r = run(c, repeats, &ti);
		if (!r)
			r = same();
		if (!r && ti < best) {
			best = ti;
			best_cfg = c;
			best_t = t;
		}
		return r;
	};
	static const int leaves[3] = {1, 2, 4}, chunks[3] = {64, 128, 256}, depths[3] = {1, 2, 4};
	for (int i = 0; i < 3 && !rc; i++)
		if (leaves[i] != best_t.leaf_size) {
			se3cq_tuning t = best_t;
			t.leaf_size = leaves[i];
			rc = trial(best_cfg, t);
		}
	if (!rc)
		rc = se3cq_set_tuning(w, &best_t);
	if (!rc) {
		DevScene s;
		{
			std::lock_guard<std::mutex> lk(w->qmu);
			s = base_scene(w);
		}
		std::vector<se3cq_batch_config> cands;
		cands.push_back(se3cq_batch_config{0, 0, 1, 1});
		const int bmax = batch_default_block(s, true) / 64;
		for (int bw = 1; bw <= bmax; bw *= 2) {
			const int full = batch_blocks_per_cu(s, 64 * bw, true);
			const int halves[2] = {full, (full + 1) / 2};
			for (int h = 0; h < (full > 1 ? 2 : 1); h++)
				for (int spl = 1; spl <= 4; spl *= 2)
					cands.push_back(se3cq_batch_config{bw, halves[h], spl, 0});
		}
		const se3cq_tuning t = best_t;
		for (size_t i = 0; i < cands.size() && !rc; i++) {
			const se3cq_batch_config& c = cands[i];
			if (c.block_waves == best_cfg.block_waves && c.blocks_per_cu == best_cfg.blocks_per_cu &&
			    c.states_per_lane == best_cfg.states_per_lane && c.split == best_cfg.split)
				continue;
			rc = trial(c, t);
		}
	}
	for (int i = 0; i < 3 && !rc; i++)
		if (chunks[i] != best_t.pool_chunk) {
			se3cq_tuning t = best_t;
			t.pool_chunk = chunks[i];
			rc = trial(best_cfg, t);
		}
	for (int i = 0; i < 3 && !rc; i++)
		if (depths[i] != best_t.queue_depth) {
			se3cq_tuning t = best_t;
			t.queue_depth = depths[i];
			rc = trial(best_cfg, t);
		}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (!rc)
		rc = se3cq_set_tuning(w, &best_t);
	if (rc)
		return rc;
	rc = take_device_errors(w);
	if (rc)
		return rc;
	w->batch_cfg = best_cfg;
	if (chosen)
		*chosen = best_cfg;
	if (ms)
		*ms = best;
	return SE3CQ_OK;
}

/* ---- OMPL-side validity (pyse3ompl) ---- */
static void ompl_xf(const se3cq_world* w, OmplXf* x)
{
	calib_rowmajor(w, x->M);
	for (int k = 0; k < 7; k++)
		x->pert[k] = w->pert[k];
	ompl_delta(w, x->delta);
}

int se3cq_ompl_longest_valid_segment(const double* lo, const double* hi, double resolution, double* out)
{
	/* SpaceInformation::setStateValidityCheckingResolution and
	 * StateSpace::setup (OMPL 1.4.1), as OmplDriver configures them
	 * (lib/pyse3ompl/ompldriver.cc:676-685) */
	if (!lo || !hi || !out)
		return fail(SE3CQ_EINVAL, "bad arguments");
	const double eps = 0x1p-52;
	if (!(resolution >= eps && resolution <= 1.0 - eps))
		return fail(SE3CQ_EINVAL, "The fraction of the extent must be larger than 0 and less than 1");
	for (int i = 0; i < 3; i++)
		if (!(hi[i] >= lo[i]))
			return fail(SE3CQ_EINVAL, "Lower bound cannot exceed upper bound");
	const double lvs = se3_ompl_longest_segment(lo, hi, resolution);
	if (!(lvs >= eps) || !(lvs < 1e300))
		return fail(SE3CQ_EINVAL, "The longest valid segment must be positive");
	*out = lvs;
	return SE3CQ_OK;
}

static int ompl_prepare(se3cq_world* w)
{
	if (!w->ready())
		return fail(SE3CQ_ESTATE, "models not loaded");
	return setup_device(w);
}

/* The OMPL-to-unit transform with the perturbation of the launch's scene
 * (the block the scene reads), so both come from one value. */
static OmplXf ompl_xf_of(const se3cq_world* w, const DevScene& s)
{
	OmplXf x;
	ompl_xf(w, &x);
	for (const auto& b : w->perts)
		if ((const PertData*)b->buf.p == s.pert)
			for (int k = 0; k < 7; k++)
				x.pert[k] = b->q[k];
	return x;
}

int se3cq_ompl_validate_states_device(se3cq_world* w, const double* d_q, int64_t n, int32_t* d_out, void* stream)
{
	CHECK_WORLD(w);
	if (n < 0)
		return fail(SE3CQ_EINVAL, "negative count");
	const int rc0 = ompl_prepare(w);
	if (rc0)
		return rc0;
	return device_launch(w, stream, 0, [&](const DevScene& s, void*) {
		return launch_ompl_states(s, ompl_xf_of(w, s), d_q, n, d_out, pick_stream(w, stream));
	});
}

int se3cq_ompl_validate_states_host(se3cq_world* w, const double* q, int64_t n, int32_t* out)
{
	CHECK_WORLD(w);
	if (n < 0 || (n > 0 && (!q || !out)))
		return fail(SE3CQ_EINVAL, "bad arguments");
	const int rc0 = ompl_prepare(w);
	if (rc0 || n == 0)
		return rc0;
	HostCall hc(w);
	if (!hc.ok())
		return fail(SE3CQ_EDEVICE, "cannot create a stream for the call");
	double* dq = hc.buf<double>(0, 7 * n);
	int32_t* dout = hc.buf<int32_t>(3, n);
	if (!dq || !dout)
		return fail(SE3CQ_ENOMEM, "device allocation failed");
	int rc = hc.begin();
	if (!rc)
		rc = hc.h2d(dq, q, 7 * n * sizeof(double));
	if (!rc)
		rc = hc.launch(0, [&](const DevScene& s, void*) {
			return launch_ompl_states(s, ompl_xf_of(w, s), dq, n, dout, hc.stream());
		});
	if (!rc)
		rc = hc.finish();
	if (!rc)
		rc = hc.d2h(out, dout, (size_t)n * sizeof(int32_t));
	return rc ? rc : hc.sync();
}

int se3cq_ompl_validate_motions_device(se3cq_world* w, const double* d_q0, const double* d_q1, int64_t n, double lvs,
                                       int32_t* d_out, int32_t* d_samples, void* stream)
{
	CHECK_WORLD(w);
	if (n < 0 || !(lvs > 0.0))
		return fail(SE3CQ_EINVAL, "bad arguments");
	const int rc0 = ompl_prepare(w);
	if (rc0)
		return rc0;
	return device_launch(w, stream, 0, [&](const DevScene& s, void*) {
		return launch_ompl_motions(s, ompl_xf_of(w, s), d_q0, d_q1, n, lvs, d_out, d_samples, pick_stream(w, stream));
	});
}

int se3cq_ompl_validate_motions_host(se3cq_world* w, const double* q0, const double* q1, int64_t n, double lvs,
                                     int32_t* out, int32_t* samples)
{
	CHECK_WORLD(w);
	if (n < 0 || !(lvs > 0.0) || (n > 0 && (!q0 || !q1 || !out)))
		return fail(SE3CQ_EINVAL, "bad arguments");
	const int rc0 = ompl_prepare(w);
	if (rc0 || n == 0)
		return rc0;
	HostCall hc(w);
	if (!hc.ok())
		return fail(SE3CQ_EDEVICE, "cannot create a stream for the call");
	double* d0 = hc.buf<double>(0, 7 * n);
	double* d1 = hc.buf<double>(6, 7 * n);
	int32_t* dout = hc.buf<int32_t>(3, n);
	int32_t* dsam = hc.buf<int32_t>(2, n);
	if (!d0 || !d1 || !dout || !dsam)
		return fail(SE3CQ_ENOMEM, "device allocation failed");
	int rc = hc.begin();
	if (!rc)
		rc = hc.h2d(d0, q0, 7 * n * sizeof(double));
	if (!rc)
		rc = hc.h2d(d1, q1, 7 * n * sizeof(double));
	if (!rc)
		rc = hc.launch(0, [&](const DevScene& s, void*) {
			return launch_ompl_motions(s, ompl_xf_of(w, s), d0, d1, n, lvs, dout, dsam, hc.stream());
		});
	if (!rc)
		rc = hc.finish();
	if (!rc)
		rc = hc.d2h(out, dout, (size_t)n * sizeof(int32_t));
	if (!rc && samples)
		rc = hc.d2h(samples, dsam, (size_t)n * sizeof(int32_t));
	if (!rc)
		rc = hc.sync();
	if (rc)
		return rc;
	for (int64_t i = 0; i < n; i++)
		if (out[i] < 0)
			return fail(SE3CQ_EINVAL, "motion " + std::to_string(i) +
			                              " needs 2^31 or more segments (or has a non-finite state)");
	return SE3CQ_OK;
}

int se3cq_distance(const double* a7, const double* b7, double* out)
{
	*out = se3_distance(a7, b7);
	return SE3CQ_OK;
}

int se3cq_multi_distance(const double* origin7, const double* targets, int64_t n, double* out)
{
	for (int64_t i = 0; i < n; i++)
		out[i] = se3_distance(origin7, targets + 7 * i);
	return SE3CQ_OK;
}

int se3cq_interpolate(const double* a7, const double* b7, double tau, double* out7)
{
	se3_segment sg;
	se3_segment_init(&sg, a7, b7);
	se3_segment_at(&sg, tau, out7);
	return SE3CQ_OK;
}

int se3cq_apply(const double* from7, const double* tr3, const double* aa3, double* out7)
{
	double rot[4] = {from7[3], from7[4], from7[5], from7[6]};
	se3_quat_normalize(rot);
	const double angle = std::sqrt((aa3[0] * aa3[0] + aa3[1] * aa3[1]) + aa3[2] * aa3[2]);
	double axis[3] = {1.0, 0.0, 0.0};
	if (angle != 0.0)
		for (int k = 0; k < 3; k++)
			axis[k] = aa3[k] / angle;
	const double ha = 0.5 * angle;
	const double sh = se3_sin(ha);
	const double qa[4] = {se3_cos(ha), sh * axis[0], sh * axis[1], sh * axis[2]};
	double r[4];
	se3_quat_mul(qa, rot, r);
	se3_quat_normalize(r);
	for (int k = 0; k < 3; k++)
		out7[k] = from7[k] + tr3[k];
	for (int k = 0; k < 4; k++)
		out7[3 + k] = r[k];
	return SE3CQ_OK;
}

int se3cq_apply_batch(const double* from, const double* tr, const double* aa, int64_t n, double* out)
{
	if (n < 0 || (n > 0 && (!from || !tr || !aa || !out)))
		return fail(SE3CQ_EINVAL, "bad arguments");
	for (int64_t i = 0; i < n; i++)
		se3cq_apply(from + 7 * i, tr + 3 * i, aa + 3 * i, out + 7 * i);
	return SE3CQ_OK;
}

int se3cq_decompose_2(const double* q7, double* t3, double* R9)
{
	double q[4] = {q7[3], q7[4], q7[5], q7[6]};
	se3_quat_normalize(q);
	se3_quat_to_rotmat(q[0], q[1], q[2], q[3], R9);
	for (int k = 0; k < 3; k++)
		t3[k] = q7[k];
	return SE3CQ_OK;
}

int se3cq_decompose_3(const double* q7, double* t3, double* angle, double* axis3)
{
	/* decompose (osr_state.cc:85-92) normalises the quaternion; Eigen 3.3's
	 * AngleAxis from a quaternion: n = |vec| (the scaled norm when n is below
	 * the binary64 epsilon), angle = 2 atan2(n, |w|), axis = vec / n with
	 * n negated when w < 0; the identity gives angle 0 and axis (1, 0, 0). */
	if (!q7 || !t3 || !angle || !axis3)
		return fail(SE3CQ_EINVAL, "bad arguments");
	double q[4] = {q7[3], q7[4], q7[5], q7[6]};
	se3_quat_normalize(q);
	for (int k = 0; k < 3; k++)
		t3[k] = q7[k];
	double n = std::sqrt((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]);
	if (n < 0x1p-52) {
		const double sc = std::max(std::fabs(q[1]), std::max(std::fabs(q[2]), std::fabs(q[3])));
		if (sc > 0.0) {
			const double a = q[1] / sc, b = q[2] / sc, c = q[3] / sc;
			n = sc * std::sqrt((a * a + b * b) + c * c);
		}
	}
	if (n != 0.0) {
		*angle = 2.0 * se3_atan2(n, std::fabs(q[0]));
		if (q[0] < 0.0)
			n = -n;
		for (int k = 0; k < 3; k++)
			axis3[k] = q[1 + k] / n;
	} else {
		*angle = 0.0;
		axis3[0] = 1.0;
		axis3[1] = 0.0;
		axis3[2] = 0.0;
	}
	return SE3CQ_OK;
}

int se3cq_compose_from_angleaxis(const double* t3, double angle, const double* axis3, double* out7)
{
	const double ha = 0.5 * angle;
	const double sh = se3_sin(ha);
	double q[4] = {se3_cos(ha), sh * axis3[0], sh * axis3[1], sh * axis3[2]};
	se3_quat_normalize(q);
	for (int k = 0; k < 3; k++)
		out7[k] = t3[k];
	for (int k = 0; k < 4; k++)
		out7[3 + k] = q[k];
	return SE3CQ_OK;
}

} /* extern "C" */

/*
 * mlp.hip -- the learned prescreen: a 12 -> 256 -> 256 -> 256 -> 256 -> 2
 * multilayer perceptron over SE(3) states, trained and evaluated with
 * gfx950 BF16 matrix cores (v_mfma_f32_32x32x16_bf16) and FP32 accumulation.
 *
 * Input features per state: the rotation matrix of the normalised quaternion
 * (row-major, decompose_2 of lib/osr/osr_state.cc:94-99) and the translation
 * minus 0.5, padded with zeros to 32.  Output: two logits, z_free and z_coll;
 * P(free) = sigmoid(z_free - z_coll) and the decision margin is
 * z_free - z_coll.  The surrogate only ranks and prescreens: every state it
 * accepts is re-verified by the exact engine.
 *
 * Every product is one "NT" GEMM, C[m][n] = sum_k A[m][k] * Bt[n][k], with the
 * reduction index contiguous in both operands: the forward pass uses W
 * (N x K), the input gradient a transposed bf16 copy of W, and the weight
 * gradient transposed copies of the activations and output gradients, which
 * the producing kernels write alongside the row-major ones.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/se3mlp.h"

namespace se3mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

enum { IN = 32, HID = 256, OUT = 32, LAYERS = 5, OUT_USED = 2 };

__host__ __device__ inline uint16_t f2bf(float x)
{
	uint32_t u;
	std::memcpy(&u, &x, 4);
	if ((u & 0x7f800000u) == 0x7f800000u)
		return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u)); /* inf stays inf, NaN stays NaN */
	u += 0x7fffu + ((u >> 16) & 1u);                                /* round to nearest even */
	return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h)
{
	const uint32_t u = (uint32_t)h << 16;
	float x;
	std::memcpy(&x, &u, 4);
	return x;
}

/*
 * Position of element (row, col) of a [rows][K] bf16 matrix stored
 * fragment-major (rows a multiple of 32, K of 16): the 32 x 16 blocks in
 * row-block-major order, each as the 64 lanes x 8 values that
 * v_mfma_f32_32x32x16_bf16 takes for one operand (lane r + 32h holds row r,
 * columns 8h .. 8h + 7 of the block).  A wavefront's load of one fragment
 * is then 1 KB of consecutive bytes, where a row-major load touches 32
 * cache lines for 32 bytes each.
 */
__host__ __device__ inline int64_t frag_at(int64_t row, int64_t col, int64_t K)
{
	return (((row >> 5) * (K >> 4) + (col >> 4)) * 64 + (row & 31) + 32 * ((col >> 3) & 1)) * 8 + (col & 7);
}

/*
 * The training kernel's activations and gradients ([features][rows], K =
 * rows) use the same blocks with the 32 feature rows of each block in the
 * order that its copy-out writes whole 128-byte runs: feature f sits in
 * fragment row feat_pos(f), and fragment row p holds feature pos_feat(p)
 * of the block.  The weight-gradient kernel maps its rows back.
 */
__host__ __device__ inline int feat_pos(int f) { return 8 * (f & 3) + ((f & 31) >> 2); }
__host__ __device__ inline int pos_feat(int p) { return 4 * (p & 7) + (p >> 3); }
__host__ __device__ inline int64_t act_at(int64_t f, int64_t s, int64_t rows)
{
	return (((f >> 5) * (rows >> 4) + (s >> 4)) * 64 + feat_pos((int)f) + 32 * ((s >> 3) & 1)) * 8 + (s & 7);
}

/* ---- GEMM ---------------------------------------------------------------- */
struct Gemm {
	const uint16_t* A;  /* A[m][k] = A[m * lda + k] */
	const uint16_t* Bt; /* B[k][n] = Bt[n * ldb + k] */
	int64_t lda, ldb;
	int M, N, K;        /* M, N multiples of 32; K a multiple of 16 */
	const float* bias;  /* [N] or null */
	int relu;           /* max(0, .) after the bias */
	const uint16_t* mask; int64_t ldm; /* if set: zero where mask[m][n] <= 0 (ReLU backward) */
	float* C; int64_t ldc;             /* f32 out, optional */
	uint16_t* Cb; int64_t ldcb;        /* bf16 out, optional */
	uint16_t* Ct; int64_t ldct;        /* bf16 transposed out Ct[n][m], optional */
};

/* Makes this wavefront's LDS stores visible to its own later LDS loads. */
__device__ __forceinline__ void wave_lds_fence()
{
	__builtin_amdgcn_s_waitcnt(0xc07f); /* lgkmcnt(0) */
	__builtin_amdgcn_wave_barrier();
}

/*
 * acc += A[m0 + r][k0, k1) . Bt[n0 + r][k0, k1) for one 32 x 32 tile (lane
 * l holds rows m0 + l%32 and n0 + l%32 at k + 8*(l/32)): the fragments of
 * KCH k-steps are loaded at once and then multiplied in k order, so a
 * wavefront has KCH loads per operand in flight instead of one.  The
 * products are summed in the same order as one k-step at a time.
 */
#ifndef SE3MLP_KCH
#define SE3MLP_KCH 4 /* 8 took 0.78 s against 0.68 s for two epochs (200 VGPRs, half the waves) */
#endif
enum { KCH = SE3MLP_KCH };
__device__ __forceinline__ void tile_k(const uint16_t* pa, const uint16_t* pb, int k0, int k1, f32x16& acc)
{
	for (int k = k0; k < k1; k += 16 * KCH) {
		uint4 va[KCH], vb[KCH];
#pragma unroll
		for (int j = 0; j < KCH; j++) {
			va[j] = make_uint4(0, 0, 0, 0);
			vb[j] = make_uint4(0, 0, 0, 0);
			if (k + 16 * j < k1) {
				va[j] = *reinterpret_cast<const uint4*>(pa + k + 16 * j);
				vb[j] = *reinterpret_cast<const uint4*>(pb + k + 16 * j);
			}
		}
#pragma unroll
		for (int j = 0; j < KCH; j++) {
			if (k + 16 * j < k1) {
				bf16x8 a, b;
				std::memcpy(&a, &va[j], 16);
				std::memcpy(&b, &vb[j], 16);
				acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
			}
		}
	}
}

/*
 * One wavefront per 32 x 32 output tile, four per block; accumulator element
 * i of lane l is C[m0 + 8*(i/4) + 4*(l/32) + i%4][n0 + l%32].  The bf16
 * outputs go through a per-wavefront LDS tile (rows padded to 40 elements,
 * 16-byte aligned), so the row-major Cb and the transposed Ct are both
 * written 16 bytes per lane store instead of 2.  (Splitting each tile's
 * k-steps over the four wavefronts of a block, with an LDS reduction, was
 * slower: 0.87 against 0.68 s for two epochs of training.)
 */
enum { TLD = 40 };
__global__ void __launch_bounds__(256) k_gemm_nt(Gemm g)
{
	__shared__ __align__(16) uint16_t tile[4][2][32 * TLD];
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;
	const int m0 = (blockIdx.y * 2 + (wave >> 1)) * 32;
	const int n0 = (blockIdx.x * 2 + (wave & 1)) * 32;
	if (m0 >= g.M || n0 >= g.N)
		return;
	const int r = lane & 31, kq = (lane >> 5) * 8;
	f32x16 acc = {};
	tile_k(g.A + (int64_t)(m0 + r) * g.lda + kq, g.Bt + (int64_t)(n0 + r) * g.ldb + kq, 0, g.K, acc);
	const int n = n0 + r;
	const float bn = g.bias ? g.bias[n] : 0.0f;
	uint16_t* T = tile[wave][0];  /* T[m][n] */
	uint16_t* TT = tile[wave][1]; /* TT[n][m] */
#pragma unroll
	for (int i = 0; i < 16; i++) {
		const int ml = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
		const int m = m0 + ml;
		float v = acc[i] + bn;
		if (g.relu)
			v = v > 0.0f ? v : 0.0f;
		if (g.mask && !(bf2f(g.mask[(int64_t)m * g.ldm + n]) > 0.0f))
			v = 0.0f;
		if (g.C)
			g.C[(int64_t)m * g.ldc + n] = v;
		const uint16_t h = f2bf(v);
		T[ml * TLD + r] = h;
		TT[r * TLD + ml] = h;
	}
	wave_lds_fence();
	/* lane l writes 16 values of row l/2 (Cb) and of column l/2 (Ct) */
	const int row = lane >> 1, c0 = 16 * (lane & 1);
	if (g.Cb) {
		const uint4* src = reinterpret_cast<const uint4*>(T + row * TLD + c0);
		uint4* dst = reinterpret_cast<uint4*>(g.Cb + (int64_t)(m0 + row) * g.ldcb + n0 + c0);
		dst[0] = src[0];
		dst[1] = src[1];
	}
	if (g.Ct) {
		const uint4* src = reinterpret_cast<const uint4*>(TT + row * TLD + c0);
		uint4* dst = reinterpret_cast<uint4*>(g.Ct + (int64_t)(n0 + row) * g.ldct + m0 + c0);
		dst[0] = src[0];
		dst[1] = src[1];
	}
}

/* ---- features, loss, reductions, optimiser ---------------------------------- */
/* Rotation of the normalised quaternion (Eigen toRotationMatrix on q/|q|) and
 * translation - 0.5; rows >= n (padding up to the tile size) are zero. */
__global__ void k_features(const double* __restrict__ q, int64_t n, int64_t rows, uint16_t* __restrict__ X,
                           uint16_t* __restrict__ XT)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= rows)
		return;
	float f[IN];
#pragma unroll
	for (int k = 0; k < IN; k++)
		f[k] = 0.0f;
	if (i < n) {
		const double* s = q + 7 * i;
		double w = s[3], x = s[4], y = s[5], z = s[6];
		const double nn = sqrt((x * x + z * z) + (y * y + w * w));
		if (nn > 0.0) {
			w /= nn; x /= nn; y /= nn; z /= nn;
		} else {
			w = 1.0;
		}
		const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
		const double R[9] = {1 - (ty * y + tz * z), tx * y - tz * w, tx * z + ty * w,
		                     tx * y + tz * w, 1 - (tx * x + tz * z), ty * z - tx * w,
		                     tx * z - ty * w, ty * z + tx * w, 1 - (tx * x + ty * y)};
#pragma unroll
		for (int k = 0; k < 9; k++)
			f[k] = (float)R[k];
#pragma unroll
		for (int k = 0; k < 3; k++)
			f[9 + k] = (float)(s[k] - 0.5);
	}
#pragma unroll
	for (int k = 0; k < IN; k++) {
		const uint16_t h = f2bf(f[k]);
		X[i * IN + k] = h;
		if (XT)
			XT[(int64_t)k * rows + i] = h;
	}
}

/* Adam on the fp32 master copies of every parameter tensor of the network
 * in one launch (grid row t takes tensor t); refreshes the bf16 copies of
 * the weight tensors, W [rows][cols] and its transpose, both fragment-major.
 * Entries outside `used` rows stay zero.  The gradient of element i is the
 * sum of `parts` partials g[t][z * size + i] (size = rows x cols), added in
 * z order. */
enum { ADAM_TENSORS = 2 * LAYERS };
struct AdamSet {
	float* w[ADAM_TENSORS];
	const float* g[ADAM_TENSORS];
	float* m[ADAM_TENSORS];
	float* v[ADAM_TENSORS];
	uint16_t* Wf[ADAM_TENSORS];
	uint16_t* WTf[ADAM_TENSORS];
	int rows[ADAM_TENSORS], cols[ADAM_TENSORS], used[ADAM_TENSORS];
	int parts;
};

__global__ void k_adam(AdamSet a, float lr, float b1, float b2, float eps, float c1, float c2)
{
	const int t = blockIdx.y;
	const int cols = a.cols[t];
	const int size = a.rows[t] * cols;
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= size)
		return;
	const int r = i / cols, c = i - r * cols;
	float* w = a.w[t];
	if (r < a.used[t]) {
		const float* gp = a.g[t] + i;
		float gi = gp[0];
		for (int z = 1; z < a.parts; z++)
			gi += gp[z * size];
		const float mi = b1 * a.m[t][i] + (1.0f - b1) * gi;
		const float vi = b2 * a.v[t][i] + (1.0f - b2) * gi * gi;
		a.m[t][i] = mi;
		a.v[t][i] = vi;
		w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
	}
	if (a.Wf[t]) {
		const uint16_t h = f2bf(w[i]);
		a.Wf[t][frag_at(r, c, cols)] = h;
		a.WTf[t][frag_at(c, r, a.rows[t])] = h;
	}
}

/* the bf16 copies of one fp32 weight tensor [rows][cols]: W and its
 * transpose, fragment-major */
__global__ void k_cast(const float* __restrict__ w, int rows, int cols, uint16_t* __restrict__ Wf,
                       uint16_t* __restrict__ WTf)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (int64_t)rows * cols)
		return;
	const uint16_t h = f2bf(w[i]);
	const int r = (int)(i / cols), c = (int)(i % cols);
	Wf[frag_at(r, c, cols)] = h;
	WTf[frag_at(c, r, rows)] = h;
}

/* ---- fused inference ------------------------------------------------------ */
/*
 * The whole network for 128 states per block of four wavefronts: features,
 * four hidden layers and the head, with the activations of the tile kept in
 * LDS (rows padded by 8 bf16 so the 32 rows a fragment load touches spread
 * over the banks; 67.6 KB, two blocks per CU).  Wave w computes output
 * features [64w, 64w + 64) of every hidden layer for all 128 states (2 x 4
 * MFMA tiles of 32 x 32), so each block reads each weight matrix once from
 * L2 and every weight fragment feeds four MFMAs.  The tiles are computed
 * transposed (weights as the A operand, activations as B), so each lane's
 * accumulator holds four consecutive features of one state: the epilogue
 * converts them with v_cvt_pk_bf16_f32 (round to nearest even, two per
 * instruction) and writes them with one 8-byte LDS store.  The next
 * k-step's weight fragments are loaded while the current step's MFMAs run.
 * The head (32 outputs, 2 used) gives each wave one 32-state tile.
 *
 * Measured on 64M samples: 71.7 ms with 64 states per block, 47.8 ms with
 * 128, and 33.7 ms once the epilogue left software rounding (which had made
 * the kernel VALU-bound: 17 VALU instructions per MFMA).  Loading
 * activation fragments a step ahead, or weights two steps ahead, measured
 * slower.
 */
#ifndef SE3MLP_INFER_AHEAD
#define SE3MLP_INFER_AHEAD 3 /* 1 / 2 / 3 steps ahead: 5.73 / 5.73 / 5.72 ms per 16M samples (5.89 with a rolled k-loop) */
#endif
enum { FM = 128, ALD = HID + 8 };
static const size_t MLP_LDS = (size_t)FM * ALD * sizeof(uint16_t);

struct MlpW {
	const uint16_t* W[LAYERS]; /* [kOut][kIn], fragment-major (frag_at) */
	const float* b[LAYERS];
};

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p)
{
	const uint4 v = *reinterpret_cast<const uint4*>(p);
	bf16x8 r;
	std::memcpy(&r, &v, 16);
	return r;
}

/*
 * Prescreen modes of the same kernel.  INFER_PLAIN writes P(free), margin or
 * logits per row.  INFER_SCREEN_STATES also appends the index of every row
 * with P(free) > threshold to idx (one atomic per block reserves the block's
 * run of slots in *count).  INFER_SCREEN_SEGMENTS evaluates `points` rows per
 * segment, the states at t = j / points (j = 1 .. points) of the segment
 * from -> to (translation interpolated linearly, the quaternion blended
 * linearly on the short arc and normalised: a cheap stand-in for the slerp
 * that is only used to rank), and appends the segment when every one of its
 * rows passes.  The order of the appended indices is not fixed.
 */
enum { INFER_PLAIN = 0, INFER_SCREEN_STATES = 1, INFER_SCREEN_SEGMENTS = 2 };

struct ScreenArgs {
	const double* to;  /* segments: end states (q holds the start states) */
	int points;        /* segments: rows per segment, a divisor of 32 */
	float threshold;
	int32_t* idx;
	uint32_t* count;
	int64_t base;      /* index of this launch's first row (state) or segment */
};

template <int MODE>
__global__ void __launch_bounds__(256, 2) k_mlp_infer(const double* __restrict__ q, int64_t n, MlpW w,
                                                      float* __restrict__ pfree, float* __restrict__ margin,
                                                      float* __restrict__ logits, ScreenArgs sa)
{
	extern __shared__ __align__(16) uint16_t act[]; /* FM x ALD */
	__shared__ uint32_t wcount[4], wbase[4];
	__shared__ float head[4][2 * 32]; /* per wave: the 32 rows' two logits */
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;
	const int r = lane & 31, kq = (lane >> 5) * 8;
	const int64_t row0 = (int64_t)blockIdx.x * FM;
	if (threadIdx.x < FM) {
		const int64_t i = row0 + threadIdx.x;
		float f[IN];
#pragma unroll
		for (int k = 0; k < IN; k++)
			f[k] = 0.0f;
		if (i < n) {
			double st[7];
			if (MODE == INFER_SCREEN_SEGMENTS) {
				const int64_t seg = i / sa.points;
				const double t = (double)(i - seg * sa.points + 1) / (double)sa.points;
				const double* a = q + 7 * seg;
				const double* b = sa.to + 7 * seg;
				const double sg = ((a[3] * b[3] + a[4] * b[4]) + (a[5] * b[5] + a[6] * b[6])) < 0.0 ? -1.0 : 1.0;
#pragma unroll
				for (int k = 0; k < 3; k++)
					st[k] = a[k] + t * (b[k] - a[k]);
#pragma unroll
				for (int k = 3; k < 7; k++)
					st[k] = a[k] + t * (sg * b[k] - a[k]);
			} else {
#pragma unroll
				for (int k = 0; k < 7; k++)
					st[k] = q[7 * i + k];
			}
			const double* s = st;
			double qw = s[3], qx = s[4], qy = s[5], qz = s[6];
			const double nn = sqrt((qx * qx + qz * qz) + (qy * qy + qw * qw));
			if (nn > 0.0) {
				qw /= nn; qx /= nn; qy /= nn; qz /= nn;
			} else {
				qw = 1.0;
			}
			const double tx = 2 * qx, ty = 2 * qy, tz = 2 * qz;
			f[0] = (float)(1 - (ty * qy + tz * qz));
			f[1] = (float)(tx * qy - tz * qw);
			f[2] = (float)(tx * qz + ty * qw);
			f[3] = (float)(tx * qy + tz * qw);
			f[4] = (float)(1 - (tx * qx + tz * qz));
			f[5] = (float)(ty * qz - tx * qw);
			f[6] = (float)(tx * qz - ty * qw);
			f[7] = (float)(ty * qz + tx * qw);
			f[8] = (float)(1 - (tx * qx + ty * qy));
#pragma unroll
			for (int k = 0; k < 3; k++)
				f[9 + k] = (float)(s[k] - 0.5);
		}
#pragma unroll
		for (int k = 0; k < IN; k += 2) {
			const bf16x2 p = {(__bf16)f[k], (__bf16)f[k + 1]};
			std::memcpy(act + threadIdx.x * ALD + k, &p, 4);
		}
	}
	__syncthreads();
	const int nb = 64 * wave;
	/* layer l's bias tile (lane (r, half) of tile h holds features nb + 32h
	 * + 8(i/4) + 4 half + (i%4) in element i) and the weight fragments of
	 * its first PD + 1 k-steps, loaded while the previous layer finishes;
	 * within a layer, the fragments of step j + PD + 1 are loaded into the
	 * slot that step j has just used */
	constexpr int PD = SE3MLP_INFER_AHEAD;
	f32x16 pbt[2];
	bf16x8 wq[PD + 1][2];
	/* this lane's part of the first fragment of weight rows nb + [0, 32) of
	 * layer l (fragment-major: k-step j is 512 values further, rows nb + 32
	 * + [0, 32) are 32 K values further) */
	auto wrow = [&](int l, int K) { return w.W[l] + (int64_t)nb * K + lane * 8; };
	auto prefetch = [&](int l) {
		const int K = l == 0 ? IN : HID;
		const uint16_t* W0 = wrow(l, K);
#pragma unroll
		for (int j = 0; j <= PD; j++) {
			if (16 * j < K) {
				wq[j][0] = ld8(W0 + 512 * j);
				wq[j][1] = ld8(W0 + 32 * K + 512 * j);
			}
		}
		const float* bl = w.b[l] + nb + 4 * (lane >> 5);
#pragma unroll
		for (int h = 0; h < 2; h++) {
#pragma unroll
			for (int g = 0; g < 4; g++) {
				const float4 bias = *reinterpret_cast<const float4*>(bl + 32 * h + 8 * g);
				pbt[h][4 * g] = bias.x;
				pbt[h][4 * g + 1] = bias.y;
				pbt[h][4 * g + 2] = bias.z;
				pbt[h][4 * g + 3] = bias.w;
			}
		}
	};
	/* c[h][t]: features nb + 32h + [0, 32) x states 32t + [0, 32).  The
	 * first k-step adds its products to the bias tile, so no accumulator is
	 * initialised separately. */
	f32x16 c[2][4];
	auto layer = [&](auto nk_c, int l) {
		constexpr int NK = decltype(nk_c)::value;
		const uint16_t* W0 = wrow(l, 16 * NK);
		const uint16_t* W1 = W0 + 32 * 16 * NK;
#pragma unroll
		for (int j = 0; j < NK; j++) {
			const int sl = j % (PD + 1);
#pragma unroll
			for (int t = 0; t < 4; t++) {
				const bf16x8 a = ld8(act + (32 * t + r) * ALD + 16 * j + kq);
				c[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wq[sl][0], a, j == 0 ? pbt[0] : c[0][t], 0, 0, 0);
				c[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wq[sl][1], a, j == 0 ? pbt[1] : c[1][t], 0, 0, 0);
			}
			if (j + PD + 1 < NK) {
				wq[sl][0] = ld8(W0 + 512 * (j + PD + 1));
				wq[sl][1] = ld8(W1 + 512 * (j + PD + 1));
			}
		}
	};
	/* lane (r, half) of tile (h, t): state 32t + r.  Each pair is rounded to
	 * bf16 (v_cvt_pk_bf16_f32, nearest even) and then put through the ReLU
	 * as two int16: a bf16 with its sign bit set is a negative int16, so
	 * max(., 0) maps it to +0 and keeps the rest (the same values as
	 * rounding max(x, 0)). */
	auto epilogue = [&]() {
#pragma unroll
		for (int h = 0; h < 2; h++) {
#pragma unroll
			for (int g = 0; g < 4; g++) {
#pragma unroll
				for (int t = 0; t < 4; t++) {
					const f32x16& v = c[h][t];
					const bf16x2 p0 = {(__bf16)v[4 * g], (__bf16)v[4 * g + 1]};
					const bf16x2 p1 = {(__bf16)v[4 * g + 2], (__bf16)v[4 * g + 3]};
					uint2 u;
					std::memcpy(&u.x, &p0, 4);
					std::memcpy(&u.y, &p1, 4);
					/* one instruction per pair; written out so that the
					 * compiler does not split it into masks */
					asm("v_pk_max_i16 %0, %0, 0" : "+v"(u.x));
					asm("v_pk_max_i16 %0, %0, 0" : "+v"(u.y));
					*reinterpret_cast<uint2*>(act + (32 * t + r) * ALD + nb + 32 * h + 8 * g + 4 * (lane >> 5)) = u;
				}
			}
		}
	};
	prefetch(0);
	layer(std::integral_constant<int, IN / 16>{}, 0);
	prefetch(1);
	__syncthreads(); /* every wave has read this layer's input */
	epilogue();
	__syncthreads();
	for (int l = 1; l < LAYERS - 1; l++) {
		layer(std::integral_constant<int, HID / 16>{}, l);
		if (l + 1 < LAYERS - 1)
			prefetch(l + 1);
		__syncthreads();
		epilogue();
		__syncthreads();
	}
	{
		const uint16_t* W = w.W[LAYERS - 1] + lane * 8;
		/* all sixteen weight fragments at once: one load latency, not 16 */
		bf16x8 hw[HID / 16];
#pragma unroll
		for (int j = 0; j < HID / 16; j++)
			hw[j] = ld8(W + 512 * j);
		f32x16 c = {};
#pragma unroll
		for (int j = 0; j < HID / 16; j++)
			c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(act + (32 * wave + r) * ALD + 16 * j + kq), hw[j], c, 0,
			                                            0, 0);
		const float bias = w.b[LAYERS - 1][r];
		/* the two used logit columns of the tile's 32 rows go to LDS (lanes
		 * of columns 0 and 1: element i is row 8(i/4) + 4 half + i%4), then
		 * lane p < 32 takes row 32 wave + p */
		float* zs = head[wave];
		if (r < OUT_USED) {
#pragma unroll
			for (int i = 0; i < 16; i++)
				zs[2 * (8 * (i >> 2) + 4 * (lane >> 5) + (i & 3)) + r] = c[i] + bias;
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		const int64_t m = row0 + 32 * wave + lane;
		bool ok = false;
		if (lane < 32 && m < n) {
			const float z = zs[2 * lane], z1 = zs[2 * lane + 1];
			const float d = z - z1;
			const float pf = 1.0f / (1.0f + expf(-d));
			if (pfree)
				pfree[m] = pf;
			if (margin)
				margin[m] = d;
			if (logits) {
				logits[2 * m] = z;
				logits[2 * m + 1] = z1;
			}
			ok = pf > sa.threshold;
		}
		if (MODE != INFER_PLAIN) {
			/* bit p: row 32 wave + p of the tile passes */
			const uint32_t rows = (uint32_t)__ballot(ok);
			/* items of this wave's tile that pass: rows, or segments whose
			 * rows all pass (a segment's rows never straddle tiles) */
			uint32_t sel = rows;
			int per_tile = 32;
			if (MODE == INFER_SCREEN_SEGMENTS) {
				const int k = sa.points;
				const uint32_t all = k == 32 ? 0xffffffffu : ((1u << k) - 1u);
				per_tile = 32 / k;
				sel = 0;
				for (int g = 0; g < per_tile; g++)
					if (((rows >> (g * k)) & all) == all)
						sel |= 1u << g;
			}
			if (lane == 0)
				wcount[wave] = (uint32_t)__popc(sel);
			__syncthreads();
			if (threadIdx.x == 0) {
				const uint32_t tot = wcount[0] + wcount[1] + wcount[2] + wcount[3];
				const uint32_t b0 = tot ? atomicAdd(sa.count, tot) : 0u;
				wbase[0] = b0;
				wbase[1] = b0 + wcount[0];
				wbase[2] = b0 + wcount[0] + wcount[1];
				wbase[3] = b0 + wcount[0] + wcount[1] + wcount[2];
			}
			__syncthreads();
			if (lane < per_tile && ((sel >> lane) & 1u)) {
				const uint32_t slot = wbase[wave] + (uint32_t)__popc(sel & ((1u << lane) - 1u));
				const int64_t first = (MODE == INFER_SCREEN_STATES) ? row0 + 32 * wave : (row0 + 32 * wave) / sa.points;
				sa.idx[slot] = (int32_t)(sa.base + first + lane);
			}
		}
	}
}

/* ---- fused training pass ---------------------------------------------------- */
/*
 * Forward, loss and backward of one training batch down to every layer's
 * pre-activation gradient, TM = 64 rows per block of four wavefronts, in
 * one launch (a batch of 16,384 rows is 256 blocks, one per CU).
 *
 * All four hidden activation tiles of the block stay in LDS (hb(l), 64 rows
 * x 256 features, rows padded as in k_mlp_infer; 145 KB with the output
 * gradient and the biases): layer l reads hb(l - 1) and writes hb(l), so a
 * layer's epilogue needs no barrier before it, and the gradient pass that
 * produces D_l reads the ReLU mask [H_l > 0] from hb(l) and overwrites it
 * with D_l, which the next gradient pass reads.  The features sit in hb(1)
 * until layer 1 writes it.  Wave w computes features [64w, 64w + 64) of a
 * layer for all rows (2 x TM/32 MFMA tiles, weights as the A operand, so
 * lane (r, half) of tile (h, t) holds features 64w + 32h + 8g + 4 half +
 * [0, 4) of row 32t + r in elements 4g + [0, 4)).
 *
 * The nine matrix passes (four hidden layers, the head, then the gradient
 * through W4, W3, W2 and W1) read their A fragments from wf[16][2], one per
 * k-step and feature half; the moment a pass has used fragment j it loads
 * fragment j of the next pass into the same registers, so every weight load
 * has a whole pass to arrive.  The weights are fragment-major (frag_at).
 *
 * Written for the weight gradients, each copied out of LDS once its tile is
 * complete: the bf16 features XT [IN][rows], hidden activations HT[l]
 * [HID][rows] (after the ReLU) and pre-activation gradients DT[l]
 * [kOut[l]][rows], laid out by act_at; and the per-row loss.  The loss is
 * the two-class softmax cross-entropy on logit columns 0 (free) and 1
 * (colliding), label 1 = free, averaged over the n real rows; rows in [n,
 * rows) get zero gradient.
 */
#ifndef SE3MLP_TRAIN_TM
#define SE3MLP_TRAIN_TM 64 /* 32 (two blocks per CU) took 32.4 against 27.5 us: each block reads all weights */
#endif
enum { TM = SE3MLP_TRAIN_TM, TT = TM / 32, DZLD = OUT + 8 };
static const size_t TRAIN_LDS = (size_t)(LAYERS - 1) * TM * ALD * 2 + (size_t)TM * DZLD * 2 + 2 * 64 * 4 +
                                (size_t)(LAYERS - 1) * HID * 4 + 16;

struct TrainArgs {
	const double* q;
	const uint8_t* label;
	int64_t n, rows;
	const uint16_t* Wf[LAYERS];  /* [kOut][kIn], fragment-major */
	const uint16_t* WTf[LAYERS]; /* [kIn][kOut], fragment-major */
	const float* b[LAYERS];
	uint16_t* XT;
	uint16_t* HT[LAYERS - 1];
	uint16_t* DT[LAYERS];
	const int64_t* perm; /* row i of the batch is state perm[i] (null: state i) */
	float* loss_part;    /* null, or per block the sum of its rows' losses */
};

/* lane (r, half) of bias tile h: features nb + 32h + 8g + 4 half + [0, 4)
 * in elements 4g + [0, 4) (b in LDS, so that a pass never waits for a
 * global load issued just before it) */
__device__ __forceinline__ void bias_tile(const float* b, int nb, int lane, f32x16 (&pb)[2])
{
	const float* bl = b + nb + 4 * (lane >> 5);
#pragma unroll
	for (int h = 0; h < 2; h++) {
#pragma unroll
		for (int g = 0; g < 4; g++) {
			const float4 v = *reinterpret_cast<const float4*>(bl + 32 * h + 8 * g);
			pb[h][4 * g] = v.x;
			pb[h][4 * g + 1] = v.y;
			pb[h][4 * g + 2] = v.z;
			pb[h][4 * g + 3] = v.w;
		}
	}
}

/*
 * c[h][t] += A_h . in^T over NK k-steps, A_h's fragments in wf[j][h], the
 * B operand rows 32t + r of the LDS tile `in` (row stride ld).  After using
 * wf[j] the lane loads fragment j of the next pass from nx0 / nx1 (this
 * lane's part of the next pass's first fragment of each half, in a
 * fragment-major matrix, so fragment j is 512 values further); fragments
 * the next pass has beyond NK are loaded first.  H_NEXT = 1 loads only nx0.
 */
template <int NK, int NK_NEXT, int H_NEXT>
__device__ __forceinline__ void train_pass(bf16x8 (&wf)[16][2], f32x16 (&c)[2][TT], const uint16_t* in, int ld,
                                           const uint16_t* nx0, const uint16_t* nx1, int r, int kq)
{
#pragma unroll
	for (int j = NK; j < NK_NEXT; j++) {
		wf[j][0] = ld8(nx0 + 512 * j);
		if (H_NEXT == 2)
			wf[j][1] = ld8(nx1 + 512 * j);
	}
	/* the B fragments of step j + 1 are read from LDS before step j's MFMAs
	 * (reading them 2 or 3 steps ahead measured the same) */
	bf16x8 b[TT];
#pragma unroll
	for (int t = 0; t < TT; t++)
		b[t] = ld8(in + (32 * t + r) * ld + kq);
#pragma unroll
	for (int j = 0; j < NK; j++) {
		bf16x8 nb[TT];
#pragma unroll
		for (int t = 0; t < TT; t++)
			nb[t] = j + 1 < NK ? ld8(in + (32 * t + r) * ld + 16 * (j + 1) + kq) : b[t];
#pragma unroll
		for (int h = 0; h < 2; h++)
#pragma unroll
			for (int t = 0; t < TT; t++)
				c[h][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wf[j][h], b[t], c[h][t], 0, 0, 0);
		if (j < NK_NEXT) {
			wf[j][0] = ld8(nx0 + 512 * j);
			if (H_NEXT == 2)
				wf[j][1] = ld8(nx1 + 512 * j);
		}
#pragma unroll
		for (int t = 0; t < TT; t++)
			b[t] = nb[t];
	}
}

/* The block's 64 rows x NF features of an LDS tile (row stride ld) to the
 * [features][rows] matrix FT laid out by act_at.  An item is 8 rows x 4
 * features: eight 8-byte LDS reads, transposed with byte permutes into four
 * 16-byte stores of 8 rows of one feature; store j of the 8 items of a
 * 32-feature block goes to fragment rows 8j .. 8j + 7, one 128-byte run.
 * The trip count is a constant, so the compiler counts the stores exactly
 * and does not wait for the next pass's weight fragments along with them. */
template <int NF>
__device__ __forceinline__ void copy_out(const uint16_t* tile, int ld, uint16_t* FT, int64_t rows, int64_t row0)
{
	constexpr int GROUPS = NF / 4, ITEMS = GROUPS * (TM / 8);
#pragma unroll
	for (int k = 0; k < (ITEMS + 255) / 256; k++) {
		const int q = threadIdx.x + 256 * k;
		if (ITEMS % 256 == 0 || q < ITEMS) {
			const int fg = q % GROUPS, s0 = 8 * (q / GROUPS);
			uint2 x[8];
#pragma unroll
			for (int i = 0; i < 8; i++)
				x[i] = *reinterpret_cast<const uint2*>(tile + (s0 + i) * ld + 4 * fg);
#pragma unroll
			for (int j = 0; j < 4; j++) {
				/* feature 4 fg + j: the low (j even) or high half of dword j / 2
				 * of each row; rows 2i and 2i + 1 form output dword i */
				const uint32_t sel = (j & 1) ? 0x07060302u : 0x05040100u;
				uint32_t w[4];
#pragma unroll
				for (int i = 0; i < 4; i++) {
					const uint32_t lo = (j >> 1) ? x[2 * i].y : x[2 * i].x;
					const uint32_t hi = (j >> 1) ? x[2 * i + 1].y : x[2 * i + 1].x;
					w[i] = __builtin_amdgcn_perm(hi, lo, sel);
				}
				*reinterpret_cast<uint4*>(FT + act_at(4 * fg + j, row0 + s0, rows)) = make_uint4(w[0], w[1], w[2], w[3]);
			}
		}
	}
}

/* hidden layer epilogue: bf16 rounding and ReLU as in k_mlp_infer, the
 * activations to the layer's LDS tile */
__device__ __forceinline__ void train_fwd_epi(const f32x16 (&c)[2][TT], uint16_t* hl, int wave, int lane)
{
	const int r = lane & 31, half = lane >> 5;
#pragma unroll
	for (int h = 0; h < 2; h++) {
#pragma unroll
		for (int t = 0; t < TT; t++) {
			const f32x16& v = c[h][t];
#pragma unroll
			for (int g = 0; g < 4; g++) {
				const bf16x2 p0 = {(__bf16)v[4 * g], (__bf16)v[4 * g + 1]};
				const bf16x2 p1 = {(__bf16)v[4 * g + 2], (__bf16)v[4 * g + 3]};
				uint2 u;
				std::memcpy(&u.x, &p0, 4);
				std::memcpy(&u.y, &p1, 4);
				asm("v_pk_max_i16 %0, %0, 0" : "+v"(u.x));
				asm("v_pk_max_i16 %0, %0, 0" : "+v"(u.y));
				*reinterpret_cast<uint2*>(hl + (32 * t + r) * ALD + 64 * wave + 32 * h + 8 * g + 4 * half) = u;
			}
		}
	}
}

/* 0xffff in each 16-bit half of x that is not zero (x holds two bf16 after
 * the ReLU, so each half is in [0, 0x7fff]): min(half, 1), then 0 - that */
__device__ __forceinline__ uint32_t nonzero_halves(uint32_t x)
{
	uint32_t m;
	asm("v_pk_min_u16 %0, %1, %2" : "=v"(m) : "v"(x), "v"(0x00010001u));
	asm("v_pk_sub_u16 %0, 0, %0" : "+v"(m));
	return m;
}

/* gradient epilogue: round to bf16, zero where the layer's activation in hl
 * is not positive, and overwrite the activation with the gradient */
__device__ __forceinline__ void train_bwd_epi(const f32x16 (&c)[2][TT], uint16_t* hl, int wave, int lane)
{
	const int r = lane & 31, half = lane >> 5;
#pragma unroll
	for (int h = 0; h < 2; h++) {
#pragma unroll
		for (int t = 0; t < TT; t++) {
			const f32x16& v = c[h][t];
#pragma unroll
			for (int g = 0; g < 4; g++) {
				uint2* p = reinterpret_cast<uint2*>(hl + (32 * t + r) * ALD + 64 * wave + 32 * h + 8 * g + 4 * half);
				const uint2 act = *p;
				const bf16x2 p0 = {(__bf16)v[4 * g], (__bf16)v[4 * g + 1]};
				const bf16x2 p1 = {(__bf16)v[4 * g + 2], (__bf16)v[4 * g + 3]};
				uint2 u;
				std::memcpy(&u.x, &p0, 4);
				std::memcpy(&u.y, &p1, 4);
				u.x &= nonzero_halves(act.x);
				u.y &= nonzero_halves(act.y);
				*p = u;
			}
		}
	}
}

__global__ void __launch_bounds__(256, 64 / TM) k_mlp_train(TrainArgs a)
{
	extern __shared__ __align__(16) uint16_t lds[];
	auto hb = [&](int l) { return lds + l * TM * ALD; }; /* hidden tile l */
	uint16_t* dz = lds + (LAYERS - 1) * TM * ALD;        /* output gradient, columns >= 2 zero */
	float* head = reinterpret_cast<float*>(dz + TM * DZLD); /* [2][64]: the two logits of 32 rows per wave */
	float* bias = head + 2 * 64;                           /* [4][HID]: the hidden layers' biases */
	float* head_loss = bias + (LAYERS - 1) * HID;          /* [2]: the loss sums of the head's waves */
	uint16_t* xs = hb(1);
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;
	const int r = lane & 31, kq = (lane >> 5) * 8;
	const int64_t rows = a.rows;
	const int64_t row0 = (int64_t)blockIdx.x * TM;
	/* this lane's part of the first fragment of rows 64 wave + 32h + [0, 32)
	 * of a fragment-major [*][K] matrix */
	auto arow = [&](const uint16_t* M, int K, int h) {
		return M + ((int64_t)(2 * wave + h) * (K / 16) * 64 + lane) * 8;
	};

	bf16x8 wf[16][2];
	f32x16 c[2][TT], pb[2];
	auto clear = [&]() {
#pragma unroll
		for (int h = 0; h < 2; h++)
#pragma unroll
			for (int t = 0; t < TT; t++)
				c[h][t] = f32x16{};
	};
#pragma unroll
	for (int j = 0; j < IN / 16; j++) {
		wf[j][0] = ld8(arow(a.Wf[0], IN, 0) + 512 * j);
		wf[j][1] = ld8(arow(a.Wf[0], IN, 1) + 512 * j);
	}
#pragma unroll
	for (int l = 0; l < LAYERS - 1; l++)
		bias[l * HID + threadIdx.x] = a.b[l][threadIdx.x];
	/* what the head needs from global memory, loaded now */
	const float head_bias = a.b[LAYERS - 1][r];
	const int64_t head_row = row0 + 32 * wave + (lane & 31);
	const bool head_real = wave < TT && lane < 32 && head_row < a.n;
	const float head_y = head_real && a.label[a.perm ? a.perm[head_row] : head_row] ? 1.0f : 0.0f;
	for (int i = threadIdx.x; i < TM * DZLD / 2; i += 256)
		reinterpret_cast<uint32_t*>(dz)[i] = 0u;
	if (threadIdx.x < TM) {
		/* features of row row0 + threadIdx.x, as k_features computes them */
		const int64_t i = row0 + threadIdx.x;
		float f[IN];
#pragma unroll
		for (int k = 0; k < IN; k++)
			f[k] = 0.0f;
		if (i < a.n) {
			const double* s = a.q + 7 * (a.perm ? a.perm[i] : i);
			double qw = s[3], qx = s[4], qy = s[5], qz = s[6];
			const double nn = sqrt((qx * qx + qz * qz) + (qy * qy + qw * qw));
			if (nn > 0.0) {
				qw /= nn; qx /= nn; qy /= nn; qz /= nn;
			} else {
				qw = 1.0;
			}
			const double tx = 2 * qx, ty = 2 * qy, tz = 2 * qz;
			f[0] = (float)(1 - (ty * qy + tz * qz));
			f[1] = (float)(tx * qy - tz * qw);
			f[2] = (float)(tx * qz + ty * qw);
			f[3] = (float)(tx * qy + tz * qw);
			f[4] = (float)(1 - (tx * qx + tz * qz));
			f[5] = (float)(ty * qz - tx * qw);
			f[6] = (float)(tx * qz - ty * qw);
			f[7] = (float)(ty * qz + tx * qw);
			f[8] = (float)(1 - (tx * qx + ty * qy));
#pragma unroll
			for (int k = 0; k < 3; k++)
				f[9 + k] = (float)(s[k] - 0.5);
		}
#pragma unroll
		for (int k = 0; k < IN; k += 2) {
			const bf16x2 p = {(__bf16)f[k], (__bf16)f[k + 1]};
			uint32_t u;
			std::memcpy(&u, &p, 4);
			*reinterpret_cast<uint32_t*>(xs + threadIdx.x * ALD + k) = u;
		}
	}
	__syncthreads();
	copy_out<IN>(xs, ALD, a.XT, rows, row0);

	/* hidden layers; layer 3 is followed by the head, whose B fragments
	 * (output columns [0, 32) of W4) take wf[j][0].  Layer 1 overwrites the
	 * features in hb(1), which layer 0 and the copy above have read before
	 * the barrier that ends layer 0.  Layers 1 and 2 run the same code (a
	 * loop that is not unrolled: 26 KB of code instead of 33 KB, at the
	 * same speed). */
	auto fwd_begin = [&](int l) {
		bias_tile(bias + l * HID, 64 * wave, lane, pb);
#pragma unroll
		for (int h = 0; h < 2; h++)
#pragma unroll
			for (int t = 0; t < TT; t++)
				c[h][t] = pb[h];
	};
	auto fwd_end = [&](int l) {
		train_fwd_epi(c, hb(l), wave, lane);
		__syncthreads();
		copy_out<HID>(hb(l), ALD, a.HT[l], rows, row0);
	};
	fwd_begin(0);
	train_pass<IN / 16, HID / 16, 2>(wf, c, xs, ALD, arow(a.Wf[1], HID, 0), arow(a.Wf[1], HID, 1), r, kq);
	fwd_end(0);
#pragma unroll 1
	for (int l = 1; l < LAYERS - 2; l++) {
		fwd_begin(l);
		train_pass<HID / 16, HID / 16, 2>(wf, c, hb(l - 1), ALD, arow(a.Wf[l + 1], HID, 0), arow(a.Wf[l + 1], HID, 1),
		                                  r, kq);
		fwd_end(l);
	}
	fwd_begin(LAYERS - 2);
	train_pass<HID / 16, HID / 16, 1>(wf, c, hb(LAYERS - 3), ALD, a.Wf[LAYERS - 1] + lane * 8, nullptr, r, kq);
	fwd_end(LAYERS - 2);

	/* head and loss: wave w < TT takes rows 32 w + [0, 32); every wave then
	 * loads the A fragments of the first gradient pass (WT4, K = 32) */
	if (wave < TT) {
		f32x16 ch = {};
		const uint16_t* h3 = hb(LAYERS - 2);
#pragma unroll
		for (int j = 0; j < HID / 16; j++) {
			ch = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(h3 + (32 * wave + r) * ALD + 16 * j + kq), wf[j][0], ch, 0,
			                                             0, 0);
			if (j < OUT / 16) {
				wf[j][0] = ld8(arow(a.WTf[LAYERS - 1], OUT, 0) + 512 * j);
				wf[j][1] = ld8(arow(a.WTf[LAYERS - 1], OUT, 1) + 512 * j);
			}
		}
		float* zs = head + 64 * wave;
		if (r < OUT_USED) {
#pragma unroll
			for (int e = 0; e < 16; e++)
				zs[2 * (8 * (e >> 2) + 4 * (lane >> 5) + (e & 3)) + r] = ch[e] + head_bias;
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		float l = 0.0f;
		if (lane < 32) {
			const int s = 32 * wave + lane;
			float g0 = 0.0f;
			if (head_real) {
				const float z0 = zs[2 * lane], z1 = zs[2 * lane + 1];
				const float d = z1 - z0;                   /* log-odds of "colliding" */
				const float p0 = 1.0f / (1.0f + expf(d)); /* P(free) */
				const float y0 = head_y;
				g0 = (p0 - y0) * (1.0f / (float)a.n);
				/* -log p(label), computed stably */
				const float tt = y0 > 0.0f ? d : -d;
				l = tt > 0.0f ? tt + log1pf(expf(-tt)) : log1pf(expf(tt));
			}
			dz[s * DZLD] = f2bf(g0);
			dz[s * DZLD + 1] = f2bf(-g0);
		}
		/* the wave's 32 losses added in a fixed butterfly order */
#pragma unroll
		for (int off = 16; off > 0; off >>= 1)
			l += __shfl_xor(l, off);
		if (lane == 0)
			head_loss[wave] = l;
	} else {
#pragma unroll
		for (int j = 0; j < OUT / 16; j++) {
			wf[j][0] = ld8(arow(a.WTf[LAYERS - 1], OUT, 0) + 512 * j);
			wf[j][1] = ld8(arow(a.WTf[LAYERS - 1], OUT, 1) + 512 * j);
		}
	}
	__syncthreads();
	if (a.loss_part && threadIdx.x == 0)
		a.loss_part[blockIdx.x] = TT == 2 ? head_loss[0] + head_loss[1] : head_loss[0];
	copy_out<OUT>(dz, DZLD, a.DT[LAYERS - 1], rows, row0);

	/* D3 = (D4 . W4) * [H3 > 0] from dz, into hb(3) */
	clear();
	train_pass<OUT / 16, HID / 16, 2>(wf, c, dz, DZLD, arow(a.WTf[3], HID, 0), arow(a.WTf[3], HID, 1), r, kq);
	train_bwd_epi(c, hb(3), wave, lane);
	__syncthreads();
	copy_out<HID>(hb(3), ALD, a.DT[3], rows, row0);
	/* D_{l-1} = (D_l . W_l) * [H_{l-1} > 0] from hb(l) into hb(l - 1); l = 3
	 * and 2 run the same code */
	auto bwd_end = [&](int l) {
		train_bwd_epi(c, hb(l - 1), wave, lane);
		__syncthreads();
		copy_out<HID>(hb(l - 1), ALD, a.DT[l - 1], rows, row0);
	};
#pragma unroll 1
	for (int l = 3; l >= 2; l--) {
		clear();
		train_pass<HID / 16, HID / 16, 2>(wf, c, hb(l), ALD, arow(a.WTf[l - 1], HID, 0), arow(a.WTf[l - 1], HID, 1),
		                                  r, kq);
		bwd_end(l);
	}
	clear();
	train_pass<HID / 16, 0, 2>(wf, c, hb(1), ALD, nullptr, nullptr, r, kq);
	bwd_end(1);
}

/*
 * Weight and bias gradients of every layer in one launch.  Layer l's
 * gradient is gW[o][i] = sum over rows s of DT_l[o][s] XT_l[i][s] (XT_0 the
 * features, XT_l = HT[l - 1]) and gb[o] = sum_s DT_l[o][s].  Block b takes a
 * 64 x 64 tile of one layer (32-wide tiles at the narrow layers compute the
 * missing half on duplicated rows and drop it) over slice z of WG_ROWS
 * rows, each wavefront a quarter of the slice with WG_KCH k-steps'
 * fragments in flight.  DT_l and XT_l are [features][rows] matrices laid
 * out by act_at, so each fragment load reads 1 KB of consecutive bytes, and
 * fragment row p of a block is feature pos_feat(p).  The four wavefronts
 * add their tiles in LDS in wave order and write P_l[z][o][i]; blocks of
 * tile column 0 also write the bias partials
 * Pb_l[z][o].  k_adam adds the slices in z order, so the sums do not depend
 * on scheduling.
 */
#ifndef SE3MLP_WG_XCD
#define SE3MLP_WG_XCD 1
#endif
#ifndef SE3MLP_WG_KCH
#define SE3MLP_WG_KCH 4 /* 8 took 17.0 against 16.4 us per launch */
#endif
#ifndef SE3MLP_WG_ROWS
#define SE3MLP_WG_ROWS 2048 /* 1024 / 2048 / 4096: 55.8 / 50.2 / 54.5 us per training step */
#endif
enum { WG_ROWS = SE3MLP_WG_ROWS, WG_KCH = SE3MLP_WG_KCH };
static const size_t WG_LDS = (size_t)4 * 64 * 64 * 4 + (size_t)4 * 64 * 4;

struct WgradArgs {
	const uint16_t* A[LAYERS]; /* DT_l: [M][rows], act_at layout */
	const uint16_t* B[LAYERS]; /* XT_l: [N][rows], act_at layout */
	float* P[LAYERS];          /* [parts][M][N] */
	float* Pb[LAYERS];         /* [parts][M] */
	int M[LAYERS], N[LAYERS];
	int first[LAYERS + 1];     /* first block of each layer; first[LAYERS] = grid */
	int64_t rows;
	const float* loss_part; /* the training kernel's per-block loss sums */
	int loss_blocks;
	float loss_scale;
	float* loss_out;        /* null, or where block 0 writes loss_scale x their sum */
};

__global__ void __launch_bounds__(256, 2) k_wgrad(WgradArgs g)
{
	extern __shared__ __align__(16) float red[]; /* [4][64][64] tiles, then [4][64] bias sums */
	if (g.loss_out && blockIdx.x == 0 && threadIdx.x == 0) {
		/* the batch's mean loss: the training blocks' sums in block order */
		float sum = 0.0f;
		for (int k = 0; k < g.loss_blocks; k++)
			sum += g.loss_part[k];
		*g.loss_out = sum * g.loss_scale;
	}
	int l = 0;
	while ((int)blockIdx.x >= g.first[l + 1])
		l++;
	const int M = g.M[l], N = g.N[l];
	const int tn = (N + 63) / 64, tiles = ((M + 63) / 64) * tn;
	const int b = (int)blockIdx.x - g.first[l];
#if SE3MLP_WG_XCD
	/* consecutive blocks go to different XCDs, so with 8 slices every tile
	 * of slice z runs on one XCD and reads the slice's rows from its L2 */
	const int parts = (g.first[l + 1] - g.first[l]) / tiles;
	const int z = b % parts, tile = b / parts;
#else
	const int z = b / tiles, tile = b % tiles;
#endif
	const int o0 = (tile / tn) * 64, i0 = (tile % tn) * 64;
	const bool m2 = M - o0 >= 64, n2 = N - i0 >= 64, bias = i0 == 0;
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;
	const int r = lane & 31, kq = (lane >> 5) * 8;
	const int64_t rows = g.rows;
	const int64_t s0 = (int64_t)z * WG_ROWS;
	const int64_t s1 = rows - s0 < WG_ROWS ? rows : s0 + WG_ROWS;
	const int64_t q = (s1 - s0) / 4; /* a multiple of 16: rows is a multiple of 64 */
	const int64_t k0 = s0 + wave * q, k1 = k0 + q;
	/* this lane's part of the fragments of row blocks o0 / 32 (+ 1) and
	 * i0 / 32 (+ 1); the fragment of k-step k (a multiple of 16) is 32 k
	 * values further */
	const int64_t blk = rows / 16 * 64 * 8; /* one row block */
	const uint16_t* pa[2];
	const uint16_t* pb[2];
	pa[0] = g.A[l] + (o0 / 32) * blk + lane * 8;
	pa[1] = m2 ? pa[0] + blk : pa[0];
	pb[0] = g.B[l] + (i0 / 32) * blk + lane * 8;
	pb[1] = n2 ? pb[0] + blk : pb[0];
	f32x16 c[2][2] = {};
	float sa[2] = {0.0f, 0.0f};
	/* one k-step on loaded fragments (A rows o0 + 32x + r, B rows i0 + 32y + r) */
	auto step = [&](const uint4 (&va)[2], const uint4 (&vb)[2]) {
		bf16x8 fa[2], fb[2];
#pragma unroll
		for (int x = 0; x < 2; x++) {
			std::memcpy(&fa[x], &va[x], 16);
			std::memcpy(&fb[x], &vb[x], 16);
		}
#pragma unroll
		for (int x = 0; x < 2; x++)
#pragma unroll
			for (int y = 0; y < 2; y++)
				c[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[x], fb[y], c[x][y], 0, 0, 0);
		if (bias) {
#pragma unroll
			for (int x = 0; x < 2; x++) {
				const uint32_t w4[4] = {va[x].x, va[x].y, va[x].z, va[x].w};
#pragma unroll
				for (int e = 0; e < 4; e++)
					sa[x] += bf2f((uint16_t)(w4[e] & 0xffffu)) + bf2f((uint16_t)(w4[e] >> 16));
			}
		}
	};
	int64_t k = k0;
	for (; k + 16 * WG_KCH <= k1; k += 16 * WG_KCH) {
		uint4 va[WG_KCH][2], vb[WG_KCH][2];
#pragma unroll
		for (int j = 0; j < WG_KCH; j++) {
#pragma unroll
			for (int x = 0; x < 2; x++) {
				va[j][x] = *reinterpret_cast<const uint4*>(pa[x] + 32 * (k + 16 * j));
				vb[j][x] = *reinterpret_cast<const uint4*>(pb[x] + 32 * (k + 16 * j));
			}
		}
#pragma unroll
		for (int j = 0; j < WG_KCH; j++)
			step(va[j], vb[j]);
	}
	for (; k < k1; k += 16) {
		uint4 va[2], vb[2];
#pragma unroll
		for (int x = 0; x < 2; x++) {
			va[x] = *reinterpret_cast<const uint4*>(pa[x] + 32 * k);
			vb[x] = *reinterpret_cast<const uint4*>(pb[x] + 32 * k);
		}
		step(va, vb);
	}
	/* element e of lane (r, half) of tile (x, y) is fragment row 8(e/4) +
	 * 4 half + e%4 of A block x against fragment row r of B block y */
	float* mine = red + wave * 4096;
	const int ir = pos_feat(r);
#pragma unroll
	for (int x = 0; x < 2; x++)
#pragma unroll
		for (int y = 0; y < 2; y++)
#pragma unroll
			for (int e = 0; e < 16; e++)
				mine[(32 * x + pos_feat(8 * (e >> 2) + 4 * (lane >> 5) + (e & 3))) * 64 + 32 * y + ir] = c[x][y][e];
	float* rb = red + 4 * 4096;
	if (bias) {
#pragma unroll
		for (int x = 0; x < 2; x++)
			sa[x] += __shfl_xor(sa[x], 32);
		if (lane < 32) {
			rb[wave * 64 + ir] = sa[0];
			rb[wave * 64 + 32 + ir] = sa[1];
		}
	}
	__syncthreads();
	float* P = g.P[l] + (int64_t)z * M * N;
#pragma unroll
	for (int v = 0; v < 4; v++) {
		const int e = 4 * (threadIdx.x + 256 * v); /* element of the 64 x 64 tile */
		const int ol = e >> 6, il = e & 63;
		float4 s = *reinterpret_cast<const float4*>(red + e);
#pragma unroll
		for (int w = 1; w < 4; w++) {
			const float4 t = *reinterpret_cast<const float4*>(red + w * 4096 + e);
			s.x += t.x;
			s.y += t.y;
			s.z += t.z;
			s.w += t.w;
		}
		if ((ol < 32 || m2) && (il < 32 || n2))
			*reinterpret_cast<float4*>(P + (int64_t)(o0 + ol) * N + i0 + il) = s;
	}
	if (bias && threadIdx.x < 64 && ((int)threadIdx.x < 32 || m2)) {
		const float s = ((rb[threadIdx.x] + rb[64 + threadIdx.x]) + rb[128 + threadIdx.x]) + rb[192 + threadIdx.x];
		g.Pb[l][(int64_t)z * M + o0 + threadIdx.x] = s;
	}
}

__global__ void k_zero_u32(uint32_t* p)
{
	if (threadIdx.x == 0)
		*p = 0u;
}

/* ---- host side ------------------------------------------------------------ */
struct Buf {
	void* p = nullptr;
	size_t n = 0;
	~Buf() { release(); }
	void release()
	{
		if (p)
			(void)hipFree(p);
		p = nullptr;
		n = 0;
	}
	bool alloc(size_t bytes)
	{
		release();
		if (hipMalloc(&p, bytes) != hipSuccess) {
			p = nullptr;
			return false;
		}
		n = bytes;
		return hipMemset(p, 0, bytes) == hipSuccess;
	}
	template <class T> T* as() const { return (T*)p; }
};

static const int kIn[LAYERS] = {IN, HID, HID, HID, HID};
static const int kOut[LAYERS] = {HID, HID, HID, HID, OUT};
static const int kUsed[LAYERS] = {HID, HID, HID, HID, OUT_USED};
static const int kFanIn[LAYERS] = {12, HID, HID, HID, HID};

} // namespace se3mlp

using namespace se3mlp;

struct se3mlp_net {
	int64_t rows = 0; /* batch capacity, a multiple of 64 */
	int step = 0;
	Buf w[LAYERS], b[LAYERS], mw[LAYERS], vw[LAYERS], mb[LAYERS], vb[LAYERS];
	Buf Wf[LAYERS], WTf[LAYERS]; /* bf16 weights and their transposes, fragment-major */
	Buf XT;                     /* bf16 features, feature-major */
	Buf HT[LAYERS - 1];         /* bf16 hidden activations, feature-major */
	Buf DT[LAYERS];             /* bf16 pre-activation gradients, feature-major */
	Buf lpart, lsum; /* per training block loss sums; a one-float loss slot */
	Buf part[LAYERS], bpart[LAYERS]; /* per row slice: weight and bias gradient partials */
	hipStream_t stream = nullptr;
	std::mutex mu;
};

namespace {
thread_local std::string g_err;
int fail(int rc, const std::string& m)
{
	g_err = m;
	return rc;
}
#define HCHK(x)                                                                 \
	do {                                                                        \
		hipError_t e_ = (x);                                                    \
		if (e_ != hipSuccess)                                                   \
			return fail(SE3MLP_EDEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
	} while (0)

uint64_t splitmix(uint64_t& s)
{
	uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

hipError_t gemm(hipStream_t st, const uint16_t* A, int64_t lda, const uint16_t* Bt, int64_t ldb, int M, int N, int K,
                const float* bias, int relu, const uint16_t* mask, int64_t ldm, float* C, int64_t ldc, uint16_t* Cb,
                int64_t ldcb, uint16_t* Ct, int64_t ldct)
{
	Gemm g{A, Bt, lda, ldb, M, N, K, bias, relu, mask, ldm, C, ldc, Cb, ldcb, Ct, ldct};
	dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
	hipLaunchKernelGGL(k_gemm_nt, grid, dim3(256), 0, st, g);
	return hipGetLastError();
}

inline unsigned blocks(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

} // namespace

extern "C" {

const char* se3mlp_last_error(void) { return g_err.c_str(); }

int se3mlp_create(int64_t max_batch, uint64_t seed, se3mlp_net** out)
{
	if (!out || max_batch <= 0)
		return fail(SE3MLP_EINVAL, "se3mlp_create: bad arguments");
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess)
		return fail(SE3MLP_EDEVICE, "no HIP device available (the surrogate has no CPU fallback)");
	auto* net = new se3mlp_net();
	net->rows = (max_batch + 63) / 64 * 64;
	const int64_t R = net->rows;
	const int64_t slices = (R + WG_ROWS - 1) / WG_ROWS;
	bool ok = true;
	uint64_t s = seed;
	for (int l = 0; l < LAYERS && ok; l++) {
		const size_t nw = (size_t)kOut[l] * kIn[l];
		ok = net->w[l].alloc(nw * 4) && net->b[l].alloc(kOut[l] * 4) && net->mw[l].alloc(nw * 4) &&
		     net->vw[l].alloc(nw * 4) && net->mb[l].alloc(kOut[l] * 4) && net->vb[l].alloc(kOut[l] * 4) &&
		     net->Wf[l].alloc(nw * 2) && net->WTf[l].alloc(nw * 2) &&
		     (l == LAYERS - 1 || net->HT[l].alloc((size_t)R * HID * 2)) &&
		     net->DT[l].alloc((size_t)R * kOut[l] * 2) && net->part[l].alloc((size_t)slices * nw * 4) &&
		     net->bpart[l].alloc((size_t)slices * kOut[l] * 4);
		if (!ok)
			break;
		/* Kaiming-uniform weights over the used fan-in; padding rows and
		 * columns stay zero */
		std::vector<float> hw(nw, 0.0f);
		const float bound = std::sqrt(6.0f / (float)kFanIn[l]);
		for (int r = 0; r < kUsed[l]; r++)
			for (int c = 0; c < kFanIn[l]; c++) {
				const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
				hw[(size_t)r * kIn[l] + c] = (float)((2.0 * u - 1.0) * bound);
			}
		ok = hipMemcpy(net->w[l].p, hw.data(), nw * 4, hipMemcpyHostToDevice) == hipSuccess;
	}
	ok = ok && net->XT.alloc((size_t)R * IN * 2) && net->lpart.alloc((size_t)(R / 32) * 4) && net->lsum.alloc(4);
	if (!ok) {
		delete net;
		return fail(SE3MLP_ENOMEM, "se3mlp_create: device allocation failed");
	}
	for (int l = 0; l < LAYERS; l++) {
		hipLaunchKernelGGL(k_cast, dim3(blocks((int64_t)kOut[l] * kIn[l])), dim3(256), 0, nullptr, net->w[l].as<float>(),
		                   kOut[l], kIn[l], net->Wf[l].as<uint16_t>(),
		                   net->WTf[l].as<uint16_t>());
	}
	if (hipDeviceSynchronize() != hipSuccess) {
		delete net;
		return fail(SE3MLP_EDEVICE, "se3mlp_create: weight upload failed");
	}
	*out = net;
	return SE3MLP_OK;
}

int se3mlp_destroy(se3mlp_net* net)
{
	delete net;
	return SE3MLP_OK;
}

int se3mlp_param_layout(int64_t* sizes)
{
	/* per layer: weight rows, cols, used rows, used cols */
	for (int l = 0; l < LAYERS; l++) {
		sizes[4 * l] = kOut[l];
		sizes[4 * l + 1] = kIn[l];
		sizes[4 * l + 2] = kUsed[l];
		sizes[4 * l + 3] = kFanIn[l];
	}
	return LAYERS;
}

int se3mlp_get_params(se3mlp_net* net, int layer, float* w, float* b)
{
	if (!net || layer < 0 || layer >= LAYERS)
		return fail(SE3MLP_EINVAL, "se3mlp_get_params: bad layer");
	std::lock_guard<std::mutex> lk(net->mu);
	HCHK(hipDeviceSynchronize());
	if (w)
		HCHK(hipMemcpy(w, net->w[layer].p, (size_t)kOut[layer] * kIn[layer] * 4, hipMemcpyDeviceToHost));
	if (b)
		HCHK(hipMemcpy(b, net->b[layer].p, (size_t)kOut[layer] * 4, hipMemcpyDeviceToHost));
	return SE3MLP_OK;
}

int se3mlp_set_params(se3mlp_net* net, int layer, const float* w, const float* b)
{
	if (!net || layer < 0 || layer >= LAYERS)
		return fail(SE3MLP_EINVAL, "se3mlp_set_params: bad layer");
	std::lock_guard<std::mutex> lk(net->mu);
	HCHK(hipDeviceSynchronize());
	if (w)
		HCHK(hipMemcpy(net->w[layer].p, w, (size_t)kOut[layer] * kIn[layer] * 4, hipMemcpyHostToDevice));
	if (b)
		HCHK(hipMemcpy(net->b[layer].p, b, (size_t)kOut[layer] * 4, hipMemcpyHostToDevice));
	hipLaunchKernelGGL(k_cast, dim3(blocks((int64_t)kOut[layer] * kIn[layer])), dim3(256), 0, nullptr,
	                   net->w[layer].as<float>(), kOut[layer], kIn[layer], net->Wf[layer].as<uint16_t>(),
	                   net->WTf[layer].as<uint16_t>());
	HCHK(hipDeviceSynchronize());
	return SE3MLP_OK;
}

int se3mlp_forward_device(se3mlp_net* net, const double* d_states, int64_t n, float* d_pfree, float* d_margin,
                          float* d_logits, void* stream)
{
	if (!net || n < 0 || (n > 0 && !d_states))
		return fail(SE3MLP_EINVAL, "se3mlp_forward_device: bad arguments");
	std::lock_guard<std::mutex> lk(net->mu);
	if (n == 0)
		return SE3MLP_OK;
	MlpW mw;
	for (int l = 0; l < LAYERS; l++) {
		mw.W[l] = net->Wf[l].as<uint16_t>();
		mw.b[l] = net->b[l].as<float>();
	}
	/* one block per 128 states; grids above 2^31 blocks are split */
	const int64_t per_launch = (int64_t)FM * 0x40000000LL;
	for (int64_t off = 0; off < n; off += per_launch) {
		const int64_t cnt = (n - off < per_launch) ? (n - off) : per_launch;
		hipLaunchKernelGGL(k_mlp_infer<INFER_PLAIN>, dim3((unsigned)((cnt + FM - 1) / FM)), dim3(256), MLP_LDS,
		                   (hipStream_t)stream, d_states + 7 * off, cnt, mw, d_pfree ? d_pfree + off : nullptr,
		                   d_margin ? d_margin + off : nullptr, d_logits ? d_logits + 2 * off : nullptr, ScreenArgs{});
		HCHK(hipGetLastError());
	}
	return SE3MLP_OK;
}

int se3mlp_screen_states_device(se3mlp_net* net, const double* d_states, int64_t n, float threshold, int32_t* d_idx,
                                uint32_t* d_count, float* d_pfree, void* stream)
{
	if (!net || n < 0 || n >= ((int64_t)1 << 31) || !d_count || (n > 0 && (!d_states || !d_idx)) ||
	    !std::isfinite(threshold))
		return fail(SE3MLP_EINVAL, "se3mlp_screen_states_device: bad arguments");
	std::lock_guard<std::mutex> lk(net->mu);
	hipStream_t st = (hipStream_t)stream;
	hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, d_count);
	HCHK(hipGetLastError());
	if (n == 0)
		return SE3MLP_OK;
	MlpW mw;
	for (int l = 0; l < LAYERS; l++) {
		mw.W[l] = net->Wf[l].as<uint16_t>();
		mw.b[l] = net->b[l].as<float>();
	}
	const ScreenArgs sa{nullptr, 1, threshold, d_idx, d_count, 0};
	hipLaunchKernelGGL(k_mlp_infer<INFER_SCREEN_STATES>, dim3((unsigned)((n + FM - 1) / FM)), dim3(256), MLP_LDS, st,
	                   d_states, n, mw, d_pfree, nullptr, nullptr, sa);
	HCHK(hipGetLastError());
	return SE3MLP_OK;
}

int se3mlp_screen_segments_device(se3mlp_net* net, const double* d_from, const double* d_to, int64_t n, int points,
                                  float threshold, int32_t* d_idx, uint32_t* d_count, void* stream)
{
	if (!net || n < 0 || !d_count || (n > 0 && (!d_from || !d_to || !d_idx)) || !std::isfinite(threshold) ||
	    points < 1 || points > 32 || (32 % points) != 0 || n * (int64_t)points >= ((int64_t)1 << 31))
		return fail(SE3MLP_EINVAL, "se3mlp_screen_segments_device: bad arguments (points must divide 32, "
		                           "n * points must stay below 2^31)");
	std::lock_guard<std::mutex> lk(net->mu);
	hipStream_t st = (hipStream_t)stream;
	hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, d_count);
	HCHK(hipGetLastError());
	if (n == 0)
		return SE3MLP_OK;
	MlpW mw;
	for (int l = 0; l < LAYERS; l++) {
		mw.W[l] = net->Wf[l].as<uint16_t>();
		mw.b[l] = net->b[l].as<float>();
	}
	const int64_t rows = n * points;
	const ScreenArgs sa{d_to, points, threshold, d_idx, d_count, 0};
	hipLaunchKernelGGL(k_mlp_infer<INFER_SCREEN_SEGMENTS>, dim3((unsigned)((rows + FM - 1) / FM)), dim3(256), MLP_LDS,
	                   st, d_from, rows, mw, nullptr, nullptr, nullptr, sa);
	HCHK(hipGetLastError());
	return SE3MLP_OK;
}

namespace {
/* Queues one training step on st; if d_loss is set, the batch's mean loss
 * (before the step) goes there.  The caller holds net->mu. */
int train_step(se3mlp_net* net, const double* d_states, const uint8_t* d_labels, const int64_t* d_perm, int64_t n,
               float lr, float* d_loss, hipStream_t st)
{
	const int64_t rows = (n + 63) / 64 * 64;
	/* forward, loss and pre-activation gradients */
	TrainArgs ta{};
	ta.q = d_states;
	ta.label = d_labels;
	ta.n = n;
	ta.rows = rows;
	for (int l = 0; l < LAYERS; l++) {
		ta.Wf[l] = net->Wf[l].as<uint16_t>();
		ta.WTf[l] = net->WTf[l].as<uint16_t>();
		ta.b[l] = net->b[l].as<float>();
		ta.DT[l] = net->DT[l].as<uint16_t>();
		if (l < LAYERS - 1)
			ta.HT[l] = net->HT[l].as<uint16_t>();
	}
	ta.XT = net->XT.as<uint16_t>();
	ta.perm = d_perm;
	ta.loss_part = d_loss ? net->lpart.as<float>() : nullptr;
	hipLaunchKernelGGL(k_mlp_train, dim3((unsigned)(rows / TM)), dim3(256), TRAIN_LDS, st, ta);
	HCHK(hipGetLastError());
	/* weight and bias gradient partials per slice of WG_ROWS rows */
	const int parts = (int)((rows + WG_ROWS - 1) / WG_ROWS);
	WgradArgs wg{};
	wg.rows = rows;
	int nblk = 0;
	for (int l = 0; l < LAYERS; l++) {
		wg.A[l] = net->DT[l].as<uint16_t>();
		wg.B[l] = l == 0 ? net->XT.as<uint16_t>() : net->HT[l - 1].as<uint16_t>();
		wg.P[l] = net->part[l].as<float>();
		wg.Pb[l] = net->bpart[l].as<float>();
		wg.M[l] = kOut[l];
		wg.N[l] = kIn[l];
		wg.first[l] = nblk;
		nblk += ((kOut[l] + 63) / 64) * ((kIn[l] + 63) / 64) * parts;
	}
	wg.first[LAYERS] = nblk;
	wg.loss_part = net->lpart.as<float>();
	wg.loss_blocks = (int)(rows / TM);
	wg.loss_scale = 1.0f / (float)n;
	wg.loss_out = d_loss;
	hipLaunchKernelGGL(k_wgrad, dim3((unsigned)nblk), dim3(256), WG_LDS, st, wg);
	HCHK(hipGetLastError());
	net->step++;
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	const float c1 = 1.0f / (1.0f - std::pow(b1, (float)net->step));
	const float c2 = 1.0f / (1.0f - std::pow(b2, (float)net->step));
	{
		AdamSet a{};
		int64_t biggest = 0;
		for (int l = 0; l < LAYERS; l++) {
			/* tensor 2l: the weights of layer l, tensor 2l + 1 its biases */
			const int tw = 2 * l, tb = 2 * l + 1;
			a.w[tw] = net->w[l].as<float>();
			a.g[tw] = net->part[l].as<float>();
			a.m[tw] = net->mw[l].as<float>();
			a.v[tw] = net->vw[l].as<float>();
			a.Wf[tw] = net->Wf[l].as<uint16_t>();
			a.WTf[tw] = net->WTf[l].as<uint16_t>();
			a.rows[tw] = kOut[l];
			a.cols[tw] = kIn[l];
			a.used[tw] = kUsed[l];
			biggest = std::max(biggest, (int64_t)kOut[l] * kIn[l]);
			a.w[tb] = net->b[l].as<float>();
			a.g[tb] = net->bpart[l].as<float>();
			a.m[tb] = net->mb[l].as<float>();
			a.v[tb] = net->vb[l].as<float>();
			a.Wf[tb] = nullptr;
			a.WTf[tb] = nullptr;
			a.rows[tb] = kOut[l];
			a.cols[tb] = 1;
			a.used[tb] = kUsed[l];
		}
		a.parts = parts;
		hipLaunchKernelGGL(k_adam, dim3(blocks(biggest), ADAM_TENSORS), dim3(256), 0, st, a, lr, b1, b2, eps, c1, c2);
		HCHK(hipGetLastError());
	}
	return SE3MLP_OK;
}
} // namespace

int se3mlp_train_step_device(se3mlp_net* net, const double* d_states, const uint8_t* d_labels, int64_t n, float lr,
                             float* loss_out, void* stream)
{
	if (!net || n <= 0 || n > net->rows || !d_states || !d_labels)
		return fail(SE3MLP_EINVAL, "se3mlp_train_step_device: batch must hold 1..max_batch states");
	std::lock_guard<std::mutex> lk(net->mu);
	hipStream_t st = (hipStream_t)stream;
	const int rc = train_step(net, d_states, d_labels, nullptr, n, lr, loss_out ? net->lsum.as<float>() : nullptr, st);
	if (rc)
		return rc;
	if (loss_out) {
		HCHK(hipMemcpyAsync(loss_out, net->lsum.p, 4, hipMemcpyDeviceToHost, st));
		HCHK(hipStreamSynchronize(st));
	}
	return SE3MLP_OK;
}

int se3mlp_train_epoch_device(se3mlp_net* net, const double* d_states, const uint8_t* d_labels, const int64_t* d_perm,
                              int64_t n, int64_t batch, float lr, int loss_every, float* d_losses, void* stream)
{
	if (!net || batch <= 0 || batch > net->rows || n < batch || !d_states || !d_labels || (d_losses && loss_every <= 0))
		return fail(SE3MLP_EINVAL, "se3mlp_train_epoch_device: need max_batch >= batch > 0, n >= batch and, "
		                           "with a loss buffer, loss_every > 0");
	std::lock_guard<std::mutex> lk(net->mu);
	hipStream_t st = (hipStream_t)stream;
	const int64_t steps = n / batch;
	for (int64_t k = 0; k < steps; k++) {
		float* slot = (d_losses && k % loss_every == 0) ? d_losses + k / loss_every : nullptr;
		const int rc = d_perm ? train_step(net, d_states, d_labels, d_perm + k * batch, batch, lr, slot, st)
		                      : train_step(net, d_states + 7 * k * batch, d_labels + k * batch, nullptr, batch, lr, slot, st);
		if (rc)
			return rc;
	}
	return SE3MLP_OK;
}

int se3mlp_gemm_device(const uint16_t* d_A, int64_t lda, const uint16_t* d_Bt, int64_t ldb, int M, int N, int K,
                       float* d_C, int64_t ldc, void* stream)
{
	if (M <= 0 || N <= 0 || K <= 0 || M % 32 || N % 32 || K % 16 || lda % 8 || ldb % 8)
		return fail(SE3MLP_EINVAL, "se3mlp_gemm_device: M, N must be multiples of 32, K of 16, lda/ldb of 8");
	HCHK(gemm((hipStream_t)stream, d_A, lda, d_Bt, ldb, M, N, K, nullptr, 0, nullptr, 0, d_C, ldc, nullptr, 0, nullptr,
	          0));
	return SE3MLP_OK;
}

int se3mlp_features_device(const double* d_states, int64_t n, uint16_t* d_X, void* stream)
{
	if (n <= 0)
		return n == 0 ? SE3MLP_OK : fail(SE3MLP_EINVAL, "negative count");
	hipLaunchKernelGGL(k_features, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, d_states, n, n, d_X,
	                   (uint16_t*)nullptr);
	HCHK(hipGetLastError());
	return SE3MLP_OK;
}

} // extern "C"

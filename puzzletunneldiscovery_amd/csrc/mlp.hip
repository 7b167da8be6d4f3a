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

/*
 * Split-K form of the same product for long reductions (the weight
 * gradients reduce over the batch, K = rows, while M x N is only 256 x 256:
 * one wavefront per tile would leave the device nearly idle).  Block
 * (x, y, z) computes its four 32 x 32 tiles over k in [z SPLIT_K,
 * (z + 1) SPLIT_K) and stores the f32 partial P[z][m][n]; k_sum_parts adds
 * the partials in z order, so the sum does not depend on scheduling.
 */
enum { SPLIT_K = 256 };
__global__ void __launch_bounds__(256) k_gemm_nt_part(Gemm g, float* __restrict__ P)
{
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;
	const int m0 = (blockIdx.y * 2 + (wave >> 1)) * 32;
	const int n0 = (blockIdx.x * 2 + (wave & 1)) * 32;
	if (m0 >= g.M || n0 >= g.N)
		return;
	const int k0 = blockIdx.z * SPLIT_K;
	const int k1 = (g.K - k0 < SPLIT_K) ? g.K : k0 + SPLIT_K;
	const int r = lane & 31, kq = (lane >> 5) * 8;
	f32x16 acc = {};
	tile_k(g.A + (int64_t)(m0 + r) * g.lda + kq, g.Bt + (int64_t)(n0 + r) * g.ldb + kq, k0, k1, acc);
	float* out = P + (int64_t)blockIdx.z * g.M * g.N;
	const int n = n0 + r;
#pragma unroll
	for (int i = 0; i < 16; i++) {
		const int m = m0 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
		out[(int64_t)m * g.N + n] = acc[i];
	}
}

/* C[m][n] = sum over z < parts of P[z][m][n], in z order; four consecutive
 * elements per thread (M x N and N multiples of 4), eight partials' loads in
 * flight. */
__global__ void __launch_bounds__(64) k_sum_parts(const float* __restrict__ P, int parts, int M, int N,
                                                  float* __restrict__ C, int64_t ldc)
{
	const int64_t mn = (int64_t)M * N;
	const int64_t e = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
	if (e >= mn)
		return;
	float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll 8
	for (int z = 0; z < parts; z++) {
		const float4 v = *reinterpret_cast<const float4*>(P + (int64_t)z * mn + e);
		s.x += v.x;
		s.y += v.y;
		s.z += v.z;
		s.w += v.w;
	}
	*reinterpret_cast<float4*>(C + (e / N) * ldc + e % N) = s;
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

/* Two-class softmax cross-entropy on columns 0 (free) and 1 (colliding) of
 * the output logits; label 1 = free.  Writes the gradient with respect to the
 * logits (scaled by 1/n) as bf16, row-major and transposed, and the per-row
 * loss. */
__global__ void k_xent(const float* __restrict__ Z, int64_t n, int64_t rows, const uint8_t* __restrict__ label,
                       uint16_t* __restrict__ dZ, uint16_t* __restrict__ dZT, float* __restrict__ loss)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= rows)
		return;
	float g0 = 0.0f, g1 = 0.0f, l = 0.0f;
	if (i < n) {
		const float z0 = Z[i * OUT], z1 = Z[i * OUT + 1];
		const float d = z1 - z0; /* log-odds of "colliding" */
		const float p0 = 1.0f / (1.0f + expf(d)); /* P(free) */
		const float y0 = label[i] ? 1.0f : 0.0f;
		const float inv = 1.0f / (float)n;
		g0 = (p0 - y0) * inv;
		g1 = -g0;
		/* -log p(label), computed stably */
		const float t = y0 > 0.0f ? d : -d;
		l = t > 0.0f ? t + log1pf(expf(-t)) : log1pf(expf(t));
	}
	for (int k = 0; k < OUT; k++) {
		const float v = k == 0 ? g0 : (k == 1 ? g1 : 0.0f);
		const uint16_t h = f2bf(v);
		dZ[i * OUT + k] = h;
		dZT[(int64_t)k * rows + i] = h;
	}
	loss[i] = l;
}

/* out[j] = sum_i XT[j][i] over i < rows (bias gradient; rows a multiple of
 * 8): one block per j, each thread summing runs of eight, then a tree of
 * fixed order in LDS. */
__global__ void __launch_bounds__(256) k_rowsum_bf16(const uint16_t* __restrict__ XT, int64_t rows, int cols,
                                                     float* __restrict__ out)
{
	__shared__ float part[256];
	const int j = blockIdx.x;
	if (j >= cols)
		return;
	const uint16_t* x = XT + (int64_t)j * rows;
	float s = 0.0f;
	for (int64_t i = 8 * (int64_t)threadIdx.x; i < rows; i += 8 * 256) {
		const uint4 v = *reinterpret_cast<const uint4*>(x + i);
		const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
		for (int k = 0; k < 4; k++)
			s += bf2f((uint16_t)(w[k] & 0xffffu)) + bf2f((uint16_t)(w[k] >> 16));
	}
	part[threadIdx.x] = s;
	__syncthreads();
	for (int off = 128; off > 0; off >>= 1) {
		if ((int)threadIdx.x < off)
			part[threadIdx.x] += part[threadIdx.x + off];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		out[j] = part[0];
}

__global__ void k_sum_f32(const float* __restrict__ x, int64_t n, float* __restrict__ out)
{
	__shared__ float part[256];
	float s = 0.0f;
	for (int64_t i = threadIdx.x; i < n; i += 256)
		s += x[i];
	part[threadIdx.x] = s;
	__syncthreads();
	for (int off = 128; off > 0; off >>= 1) {
		if ((int)threadIdx.x < off)
			part[threadIdx.x] += part[threadIdx.x + off];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		out[0] = part[0];
}

/* Adam on the fp32 master copies of every parameter tensor of the network
 * in one launch (tensor t holds elements [start[t], start[t + 1]) of the
 * grid); refreshes the bf16 W[rows x cols] and its transpose WT[cols x rows]
 * of the weight tensors.  Entries outside `used` rows stay zero. */
enum { ADAM_TENSORS = 2 * LAYERS };
struct AdamSet {
	float* w[ADAM_TENSORS];
	const float* g[ADAM_TENSORS];
	float* m[ADAM_TENSORS];
	float* v[ADAM_TENSORS];
	uint16_t* W[ADAM_TENSORS];
	uint16_t* WT[ADAM_TENSORS];
	int rows[ADAM_TENSORS], cols[ADAM_TENSORS], used[ADAM_TENSORS];
	int64_t start[ADAM_TENSORS + 1];
};

__global__ void k_adam(AdamSet a, float lr, float b1, float b2, float eps, float c1, float c2)
{
	const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= a.start[ADAM_TENSORS])
		return;
	int t = 0;
	while (e >= a.start[t + 1])
		t++;
	const int64_t i = e - a.start[t];
	const int cols = a.cols[t];
	const int r = (int)(i / cols), c = (int)(i % cols);
	float* w = a.w[t];
	if (r < a.used[t]) {
		const float gi = a.g[t][i];
		const float mi = b1 * a.m[t][i] + (1.0f - b1) * gi;
		const float vi = b2 * a.v[t][i] + (1.0f - b2) * gi * gi;
		a.m[t][i] = mi;
		a.v[t][i] = vi;
		w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
	}
	if (a.W[t]) {
		const uint16_t h = f2bf(w[i]);
		a.W[t][i] = h;
		if (a.WT[t])
			a.WT[t][(int64_t)c * a.rows[t] + r] = h;
	}
}

__global__ void k_cast(const float* __restrict__ w, int rows, int cols, uint16_t* __restrict__ W,
                       uint16_t* __restrict__ WT)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (int64_t)rows * cols)
		return;
	const uint16_t h = f2bf(w[i]);
	W[i] = h;
	if (WT)
		WT[(int64_t)(i % cols) * rows + i / cols] = h;
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
enum { FM = 128, ALD = HID + 8 };
static const size_t MLP_LDS = (size_t)FM * ALD * sizeof(uint16_t);

struct MlpW {
	const uint16_t* W[LAYERS];
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
	 * its first two k-steps, loaded while the previous layer finishes */
	f32x16 pbt[2];
	bf16x8 pw[4];
	auto prefetch = [&](int l) {
		const int K = l == 0 ? IN : HID;
		const uint16_t* W0 = w.W[l] + (int64_t)(nb + r) * K + kq;
		pw[0] = ld8(W0);
		pw[1] = ld8(W0 + 32 * K);
		pw[2] = ld8(W0 + 16); /* K >= 32 */
		pw[3] = ld8(W0 + 32 * K + 16);
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
	prefetch(0);
	for (int l = 0; l < LAYERS - 1; l++) {
		const int K = l == 0 ? IN : HID;
		const uint16_t* W0 = w.W[l] + (int64_t)(nb + r) * K + kq;
		const uint16_t* W1 = W0 + 32 * K;
		/* c[h][t]: features nb + 32h + [0, 32) x states 32t + [0, 32).
		 * The first k-step adds its products to the bias tile, so no
		 * accumulator is initialised separately. */
		f32x16 c[2][4];
		bf16x8 w0 = pw[2], w1 = pw[3];
#pragma unroll
		for (int t = 0; t < 4; t++) {
			const bf16x8 a = ld8(act + (32 * t + r) * ALD + kq);
			c[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pw[0], a, pbt[0], 0, 0, 0);
			c[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pw[1], a, pbt[1], 0, 0, 0);
		}
		for (int k = 16; k < K; k += 16) {
			bf16x8 nw0 = w0, nw1 = w1;
			if (k + 16 < K) {
				nw0 = ld8(W0 + k + 16);
				nw1 = ld8(W1 + k + 16);
			}
#pragma unroll
			for (int t = 0; t < 4; t++) {
				const bf16x8 a = ld8(act + (32 * t + r) * ALD + k + kq);
				c[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w0, a, c[0][t], 0, 0, 0);
				c[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, a, c[1][t], 0, 0, 0);
			}
			w0 = nw0;
			w1 = nw1;
		}
		if (l + 1 < LAYERS - 1)
			prefetch(l + 1);
		__syncthreads(); /* every wave has read this layer's input */
		/* lane (r, half) of tile (h, t): state 32t + r.  Each pair is
		 * rounded to bf16 (v_cvt_pk_bf16_f32, nearest even) and then put
		 * through the ReLU as two int16: a bf16 with its sign bit set is a
		 * negative int16, so max(., 0) maps it to +0 and keeps the rest
		 * (the same values as rounding max(x, 0)). */
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
		__syncthreads();
	}
	{
		const uint16_t* W = w.W[LAYERS - 1] + (int64_t)r * HID + kq;
		/* all sixteen weight fragments at once: one load latency, not 16 */
		bf16x8 hw[HID / 16];
#pragma unroll
		for (int j = 0; j < HID / 16; j++)
			hw[j] = ld8(W + 16 * j);
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
	Buf w[LAYERS], b[LAYERS], mw[LAYERS], vw[LAYERS], mb[LAYERS], vb[LAYERS], gw[LAYERS], gb[LAYERS];
	Buf W[LAYERS], WT[LAYERS];
	Buf X, XT;                  /* features */
	Buf H[LAYERS], HT[LAYERS];  /* bf16 activations after layer l (H[4] unused) */
	Buf Z;                      /* f32 output logits */
	Buf D[LAYERS], DT[LAYERS];  /* bf16 gradient w.r.t. the pre-activation of layer l */
	Buf loss, lsum;
	Buf part;                   /* split-K partials of the weight gradients */
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

/* forward over rows (already featurised into X); fills H/HT and Z */
int forward_rows(se3mlp_net* net, int64_t rows, bool keep_t, hipStream_t st)
{
	const uint16_t* in = net->X.as<uint16_t>();
	int64_t ldin = IN;
	for (int l = 0; l < LAYERS; l++) {
		const bool last = l == LAYERS - 1;
		HCHK(gemm(st, in, ldin, net->W[l].as<uint16_t>(), kIn[l], (int)rows, kOut[l], kIn[l], net->b[l].as<float>(),
		          last ? 0 : 1, nullptr, 0, last ? net->Z.as<float>() : nullptr, OUT,
		          last ? nullptr : net->H[l].as<uint16_t>(), HID,
		          (last || !keep_t) ? nullptr : net->HT[l].as<uint16_t>(), rows));
		in = net->H[l].as<uint16_t>();
		ldin = HID;
	}
	return SE3MLP_OK;
}
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
	bool ok = true;
	uint64_t s = seed;
	for (int l = 0; l < LAYERS && ok; l++) {
		const size_t nw = (size_t)kOut[l] * kIn[l];
		ok = net->w[l].alloc(nw * 4) && net->b[l].alloc(kOut[l] * 4) && net->mw[l].alloc(nw * 4) &&
		     net->vw[l].alloc(nw * 4) && net->mb[l].alloc(kOut[l] * 4) && net->vb[l].alloc(kOut[l] * 4) &&
		     net->gw[l].alloc(nw * 4) && net->gb[l].alloc(kOut[l] * 4) && net->W[l].alloc(nw * 2) &&
		     net->WT[l].alloc(nw * 2) && net->H[l].alloc((size_t)R * HID * 2) &&
		     net->HT[l].alloc((size_t)R * HID * 2) && net->D[l].alloc((size_t)R * HID * 2) &&
		     net->DT[l].alloc((size_t)R * HID * 2);
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
	ok = ok && net->X.alloc((size_t)R * IN * 2) && net->XT.alloc((size_t)R * IN * 2) &&
	     net->Z.alloc((size_t)R * OUT * 4) && net->loss.alloc((size_t)R * 4) && net->lsum.alloc(4) &&
	     net->part.alloc((size_t)((R + SPLIT_K - 1) / SPLIT_K) * HID * HID * 4);
	if (!ok) {
		delete net;
		return fail(SE3MLP_ENOMEM, "se3mlp_create: device allocation failed");
	}
	for (int l = 0; l < LAYERS; l++) {
		hipLaunchKernelGGL(k_cast, dim3(blocks((int64_t)kOut[l] * kIn[l])), dim3(256), 0, nullptr, net->w[l].as<float>(),
		                   kOut[l], kIn[l], net->W[l].as<uint16_t>(), net->WT[l].as<uint16_t>());
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
	                   net->w[layer].as<float>(), kOut[layer], kIn[layer], net->W[layer].as<uint16_t>(),
	                   net->WT[layer].as<uint16_t>());
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
		mw.W[l] = net->W[l].as<uint16_t>();
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
		mw.W[l] = net->W[l].as<uint16_t>();
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
		mw.W[l] = net->W[l].as<uint16_t>();
		mw.b[l] = net->b[l].as<float>();
	}
	const int64_t rows = n * points;
	const ScreenArgs sa{d_to, points, threshold, d_idx, d_count, 0};
	hipLaunchKernelGGL(k_mlp_infer<INFER_SCREEN_SEGMENTS>, dim3((unsigned)((rows + FM - 1) / FM)), dim3(256), MLP_LDS,
	                   st, d_from, rows, mw, nullptr, nullptr, nullptr, sa);
	HCHK(hipGetLastError());
	return SE3MLP_OK;
}

int se3mlp_train_step_device(se3mlp_net* net, const double* d_states, const uint8_t* d_labels, int64_t n, float lr,
                             float* loss_out, void* stream)
{
	if (!net || n <= 0 || n > net->rows || !d_states || !d_labels)
		return fail(SE3MLP_EINVAL, "se3mlp_train_step_device: batch must hold 1..max_batch states");
	std::lock_guard<std::mutex> lk(net->mu);
	hipStream_t st = (hipStream_t)stream;
	const int64_t rows = (n + 63) / 64 * 64;
	hipLaunchKernelGGL(k_features, dim3(blocks(rows)), dim3(256), 0, st, d_states, n, rows, net->X.as<uint16_t>(),
	                   net->XT.as<uint16_t>());
	HCHK(hipGetLastError());
	int rc = forward_rows(net, rows, true, st);
	if (rc)
		return rc;
	/* output gradient into D[4] / DT[4] (32 columns) */
	hipLaunchKernelGGL(k_xent, dim3(blocks(rows)), dim3(256), 0, st, net->Z.as<float>(), n, rows, d_labels,
	                   net->D[4].as<uint16_t>(), net->DT[4].as<uint16_t>(), net->loss.as<float>());
	HCHK(hipGetLastError());
	for (int l = LAYERS - 1; l >= 0; l--) {
		const int no = kOut[l], ni = kIn[l];
		const uint16_t* inT = l == 0 ? net->XT.as<uint16_t>() : net->HT[l - 1].as<uint16_t>();
		/* weight gradient: gW[no x ni] = D_l^T (no x rows) . X_l (rows x ni),
		 * split over the rows */
		{
			Gemm g{net->DT[l].as<uint16_t>(), inT, rows, rows, no, ni, (int)rows, nullptr, 0, nullptr, 0,
			       nullptr, 0, nullptr, 0, nullptr, 0};
			const int parts = (int)((rows + SPLIT_K - 1) / SPLIT_K);
			hipLaunchKernelGGL(k_gemm_nt_part, dim3((unsigned)((ni + 63) / 64), (unsigned)((no + 63) / 64), (unsigned)parts),
			                   dim3(256), 0, st, g, net->part.as<float>());
			hipLaunchKernelGGL(k_sum_parts, dim3(blocks((int64_t)no * ni / 4, 64)), dim3(64), 0, st,
			                   net->part.as<float>(), parts, no, ni, net->gw[l].as<float>(), (int64_t)ni);
		}
		hipLaunchKernelGGL(k_rowsum_bf16, dim3((unsigned)no), dim3(256), 0, st, net->DT[l].as<uint16_t>(), rows, no,
		                   net->gb[l].as<float>());
		HCHK(hipGetLastError());
		if (l > 0) {
			/* input gradient through ReLU: D_{l-1} = (D_l . W_l) * [H_{l-1} > 0] */
			HCHK(gemm(st, net->D[l].as<uint16_t>(), no, net->WT[l].as<uint16_t>(), no, (int)rows, ni, no, nullptr, 0,
			          net->H[l - 1].as<uint16_t>(), HID, nullptr, 0, net->D[l - 1].as<uint16_t>(), HID,
			          net->DT[l - 1].as<uint16_t>(), rows));
		}
	}
	net->step++;
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	const float c1 = 1.0f / (1.0f - std::pow(b1, (float)net->step));
	const float c2 = 1.0f / (1.0f - std::pow(b2, (float)net->step));
	{
		AdamSet a{};
		int64_t at = 0;
		for (int l = 0; l < LAYERS; l++) {
			/* tensor 2l: the weights of layer l, tensor 2l + 1 its biases */
			const int tw = 2 * l, tb = 2 * l + 1;
			a.w[tw] = net->w[l].as<float>();
			a.g[tw] = net->gw[l].as<float>();
			a.m[tw] = net->mw[l].as<float>();
			a.v[tw] = net->vw[l].as<float>();
			a.W[tw] = net->W[l].as<uint16_t>();
			a.WT[tw] = net->WT[l].as<uint16_t>();
			a.rows[tw] = kOut[l];
			a.cols[tw] = kIn[l];
			a.used[tw] = kUsed[l];
			a.start[tw] = at;
			at += (int64_t)kOut[l] * kIn[l];
			a.w[tb] = net->b[l].as<float>();
			a.g[tb] = net->gb[l].as<float>();
			a.m[tb] = net->mb[l].as<float>();
			a.v[tb] = net->vb[l].as<float>();
			a.W[tb] = nullptr;
			a.WT[tb] = nullptr;
			a.rows[tb] = kOut[l];
			a.cols[tb] = 1;
			a.used[tb] = kUsed[l];
			a.start[tb] = at;
			at += kOut[l];
		}
		a.start[ADAM_TENSORS] = at;
		hipLaunchKernelGGL(k_adam, dim3(blocks(at)), dim3(256), 0, st, a, lr, b1, b2, eps, c1, c2);
		HCHK(hipGetLastError());
	}
	if (loss_out) {
		hipLaunchKernelGGL(k_sum_f32, dim3(1), dim3(256), 0, st, net->loss.as<float>(), n, net->lsum.as<float>());
		HCHK(hipGetLastError());
		HCHK(hipMemcpyAsync(loss_out, net->lsum.p, 4, hipMemcpyDeviceToHost, st));
		HCHK(hipStreamSynchronize(st));
		*loss_out /= (float)n;
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

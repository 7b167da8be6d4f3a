/*
 * se3mlp.h -- C ABI of the learned prescreen in libse3cq.so: a
 * 12 -> 256 -> 256 -> 256 -> 256 -> 2 multilayer perceptron over SE(3)
 * states, trained and run with gfx950 BF16 matrix cores (FP32 accumulation,
 * FP32 master weights, Adam).
 *
 * No reference counterpart: the reference's learned components are
 * TensorFlow models outside the collision path (src/GP/pipeline/hourglass_tiny.py).
 * The prescreen only ranks candidate states.  Every state it accepts must be
 * re-verified with se3cq_valid_states_* (include/se3cq.h), which stays the
 * judge of the surrogate's recall and precision.
 *
 * Features per state: the rotation matrix of the normalised quaternion
 * (row-major), then the translation minus 0.5.  Outputs: logits z_free and
 * z_coll; P(free) = 1 / (1 + exp(z_coll - z_free)); margin = z_free - z_coll.
 * Label convention for training: 1 = collision free.
 *
 * Conventions as in se3cq.h: 0 on success, a negative status otherwise, with
 * se3mlp_last_error() holding the message; "_device" calls take device
 * pointers and a hipStream_t (NULL = the default stream).
 */
#ifndef SE3MLP_H
#define SE3MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct se3mlp_net se3mlp_net;

#define SE3MLP_OK 0
#define SE3MLP_EINVAL (-1)
#define SE3MLP_EDEVICE (-4)
#define SE3MLP_ENOMEM (-5)

const char* se3mlp_last_error(void);

/* A network with seeded Kaiming-uniform weights, able to take batches of up
 * to max_batch states (training) or to stream any number (inference). */
int se3mlp_create(int64_t max_batch, uint64_t seed, se3mlp_net** out);
int se3mlp_destroy(se3mlp_net* net);

/* Weight shapes: for each of the 5 layers, 4 int64 (rows, cols, used rows,
 * used cols) into sizes[20]; returns the number of layers.  Storage is padded
 * (layer 0 has 32 input columns of which 12 are used, the last layer 32 rows
 * of which 2 are used); padding stays zero. */
int se3mlp_param_layout(int64_t* sizes);
/* Host copies of one layer's fp32 weights [rows x cols] and bias [rows]. */
int se3mlp_get_params(se3mlp_net* net, int layer, float* w, float* b);
int se3mlp_set_params(se3mlp_net* net, int layer, const float* w, const float* b);

/* Inference over n states [n, 7] (binary64, [t, qw, qx, qy, qz]).  Any output
 * may be NULL; d_logits receives [n, 2] (z_free, z_coll). */
int se3mlp_forward_device(se3mlp_net* net, const double* d_states, int64_t n, float* d_pfree, float* d_margin,
                          float* d_logits, void* stream);

/* The prescreen of n < 2^31 states in one kernel: inference, then the index
 * of every state with P(free) > threshold appended to d_idx (room for n
 * int32), *d_count (device uint32, zeroed by the call) their number.  The
 * order of the indices is not fixed.  d_pfree (may be NULL) receives P(free)
 * of every state.  Re-verify the survivors with
 * se3cq_valid_states_subset_device, which reads the count on the device. */
int se3mlp_screen_states_device(se3mlp_net* net, const double* d_states, int64_t n, float threshold, int32_t* d_idx,
                                uint32_t* d_count, float* d_pfree, void* stream);
/* The prescreen of n segments from d_from[i] to d_to[i]: `points` (a divisor
 * of 32) states per segment at t = j / points, j = 1 .. points, with the
 * translation interpolated linearly and the quaternion blended linearly on
 * the short arc (an approximation of the slerp, for ranking only); a
 * segment is appended to d_idx when all its points have P(free) >
 * threshold.  n * points < 2^31.  Re-verify with
 * se3cq_valid_transitions_subset_device. */
int se3mlp_screen_segments_device(se3mlp_net* net, const double* d_from, const double* d_to, int64_t n, int points,
                                  float threshold, int32_t* d_idx, uint32_t* d_count, void* stream);

/* One Adam step (lr; betas 0.9, 0.999; eps 1e-8) on the mean two-class
 * cross-entropy of a batch of n <= max_batch labelled states.  loss_out, if
 * not NULL, receives the batch's mean loss before the step (synchronises). */
int se3mlp_train_step_device(se3mlp_net* net, const double* d_states, const uint8_t* d_labels, int64_t n, float lr,
                             float* loss_out, void* stream);
/* n / batch such steps (batch <= max_batch) over the consecutive batches of
 * the given n states, queued without waiting for the device.  With d_perm
 * (a permutation of 0 .. n-1, int64, on the device) batch k is the states
 * d_perm[k * batch .. (k + 1) * batch), read in place (a shuffled epoch
 * without a shuffled copy).  If d_losses is not NULL, step k with k %
 * loss_every == 0 writes its batch's mean loss (before the step) to
 * d_losses[k / loss_every], device memory the caller sizes to
 * ceil((n / batch) / loss_every) floats. */
int se3mlp_train_epoch_device(se3mlp_net* net, const double* d_states, const uint8_t* d_labels, const int64_t* d_perm,
                              int64_t n, int64_t batch, float lr, int loss_every, float* d_losses, void* stream);

/* The matrix-core kernel alone, for testing: C[M x N] (fp32, row stride ldc)
 * = A[M x K] . Bt[N x K]^T with bf16 A and Bt (row strides lda, ldb).  M and N
 * multiples of 32, K of 16, lda and ldb of 8. */
int se3mlp_gemm_device(const uint16_t* d_A, int64_t lda, const uint16_t* d_Bt, int64_t ldb, int M, int N, int K,
                       float* d_C, int64_t ldc, void* stream);
/* The 32 bf16 input features of n states, [n x 32]. */
int se3mlp_features_device(const double* d_states, int64_t n, uint16_t* d_X, void* stream);

#ifdef __cplusplus
}
#endif

#endif

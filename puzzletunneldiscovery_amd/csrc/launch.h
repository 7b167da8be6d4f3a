/*
 * launch.h -- host-callable launch wrappers of the gfx950 kernels.
 */
#ifndef SE3CQ_LAUNCH_H
#define SE3CQ_LAUNCH_H

#include <hip/hip_runtime.h>

#include "engine.h"
#include "../../include/se3cq.h"

namespace se3cq {

/* Per-segment outputs; any pointer may be null except done. */
struct SegOut {
	uint8_t* done;     /* 1 = collision free */
	double* last_tau;  /* binary64 */
	double* tau;       /* binary64 */
	double* last_free; /* [n,7] */
	double* state;     /* [n,7] */
	double* lo;        /* touch search: bisection bracket */
	double* hi;
	int32_t* steps;    /* isValid calls under the reference's semantics (+ probes) */
};

hipError_t launch_valid_states(const DevScene& s, const double* q, int64_t n, uint8_t* out, hipStream_t st);
/* mode: 0 transit (fixed verify_delta), 1 transition (min(d/2, ivd)), 2 touch (transit + bisection).
 * With order_scratch (segment_order_bytes(n) bytes of device memory, or
 * null), a launch of more segments than resident wavefronts first sorts
 * them longest first into that scratch and walks them in that order; with
 * rec_scratch as well (segment_record_bytes(n) bytes, 256-byte aligned, or
 * null), it also computes every segment's record beforehand, one per lane. */
hipError_t launch_segments(const DevScene& s, int mode, const double* from, const double* to, int64_t n,
                           double vd, int levels, const SegOut& o, hipStream_t st, void* order_scratch = nullptr,
                           void* rec_scratch = nullptr);
/* Bytes of order scratch a launch of n segments uses (0: it walks in input order). */
size_t segment_order_bytes(int64_t n);
/* Bytes of record scratch a launch of n segments uses. */
size_t segment_record_bytes(int64_t n);
/* Order scratch followed by record scratch, and where the records start. */
inline size_t segment_scratch_bytes(int64_t n)
{
	return segment_order_bytes(n) ? ((segment_order_bytes(n) + 255) & ~(size_t)255) + segment_record_bytes(n) : 0;
}
inline void* segment_records_in(void* scratch, int64_t n)
{
	return scratch && segment_record_bytes(n) ? (char*)scratch + ((segment_order_bytes(n) + 255) & ~(size_t)255)
	                                          : nullptr;
}

/* The allocation-free touch search's scratch on the current device: the
 * work counter in the first NOALLOC_COUNTER_BYTES, then order scratch for
 * up to NOALLOC_MAX_SEGMENTS segments.  Looked up once per device. */
constexpr int64_t NOALLOC_MAX_SEGMENTS = 1 << 20;
constexpr size_t NOALLOC_COUNTER_BYTES = 256;
hipError_t noalloc_scratch(void** out);
/* vis: 0 pair (row i vs row i), 1 symmetric matrix (fi < ti, mirrored), 2 full m x n matrix */
hipError_t launch_visibility(const DevScene& s, int vis, const double* qs0, const double* qs1, int64_t m,
                             int64_t n, double ivd, int8_t* out8, uint32_t* bits, int32_t* steps, hipStream_t st);

/* All intersecting triangle pairs per state (up to cap; count[i] = how many,
 * -1 after a traversal overflow) and each listed pair's intersection segment
 * (segs [n*cap*6]) and flags (bit 0: the segment test meets, bit 1: coplanar). */
hipError_t launch_contacts(const DevScene& s, const double* q, int64_t n, int cap, int32_t* pairs, int32_t* count,
                           double* segs, int8_t* flags, hipStream_t st);

/* A mixed batch for k_batch: segments under transitStateToWithContact
 * semantics (outputs as SegOut, lo/hi unused) and single states (valid: 1 =
 * free).  The work counter is the scene's queue word, which the launch
 * zeroes first. */
struct BatchArgs {
	const double* seg_from;
	const double* seg_to;
	int64_t n_seg;
	double vd;
	SegOut seg;
	const double* states;
	int64_t n_states;
	uint8_t* valid;
	int states_per_lane;
};
hipError_t launch_batch(const DevScene& s, const BatchArgs& a, int block, int grid, hipStream_t st);
/* Largest block (whose traversal stacks fit in LDS when the batch has
 * states), and how many such blocks one CU keeps resident. */
int batch_default_block(const DevScene& s, bool states);
int batch_blocks_per_cu(const DevScene& s, int block, bool states);

/* out[idx[k]] = 1 if state idx[k] is free, for k < min(*count, n); indices
 * outside [0, n) are skipped; other entries of out are not written. */
hipError_t launch_valid_states_subset(const DevScene& s, const double* q, int64_t n, const int32_t* idx,
                                      const uint32_t* count, uint8_t* out, hipStream_t st);

/* OMPL-convention validity (OmplDriver::validateStates / validateMotionPairs):
 * out 1 valid, 0 invalid; motions: -1 when the sample count is not finite
 * below 2^31, samples = max(nd, 1) (may be null). */
hipError_t launch_ompl_states(const DevScene& s, const OmplXf& x, const double* q, int64_t n, int32_t* out,
                              hipStream_t st);
hipError_t launch_ompl_motions(const DevScene& s, const OmplXf& x, const double* q0, const double* q1, int64_t n,
                               double lvs, int32_t* out, int32_t* samples, hipStream_t st);

/* Local AABBs of the two models (centre, half extents), for isDisentangled. */
struct BoxPair {
	double ce[3], he[3]; /* environment */
	double cr[3], hr[3]; /* robot */
};
hipError_t launch_disentangled(const BoxPair& b, const PertData* pert, const double* q, int64_t n, uint8_t* out,
                               hipStream_t st);

/* Fill kernels (used instead of hipMemsetAsync so callers may capture the
 * device entry points in a HIP graph). */
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st);
hipError_t launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t st);
/* *out = *flag; *flag = 0 (one atomic exchange on the device). */
hipError_t launch_take_u32(uint32_t* flag, uint32_t* out, hipStream_t st);
/* *flag |= 1 if the n bytes at a and b differ. */
hipError_t launch_differ(const void* a, const void* b, int64_t n, uint32_t* flag, hipStream_t st);

} // namespace se3cq

#endif

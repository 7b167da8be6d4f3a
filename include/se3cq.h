/*
 * se3cq.h -- C ABI of the MI355X (gfx950) SE(3) collision-query engine.
 *
 * The engine replaces the configuration-space validity path of
 * PuzzleTunnelDiscovery's pyosr.UnitWorld (lib/pyosr/module.cc:127-250 ->
 * lib/osr/unit_world.cc -> lib/osr/cdmodel.cc -> FCL).  Each entry point
 * names the reference member it stands in for.  Plain pointers and sizes only;
 * states are binary64 rows [tx, ty, tz, qw, qx, qy, qz] (lib/osr/osr_state.h:20-33).
 *
 * Conventions
 *   - Every function returns 0 on success and a negative status on failure;
 *     se3cq_last_error() then returns a message for the calling thread.
 *     Nothing throws across the boundary.
 *   - "_host" entry points take host buffers and block until results are
 *     back in them.  "_device" entry points take device pointers (HBM) and
 *     enqueue on `stream` (a hipStream_t; NULL = the world's own stream)
 *     without synchronising; call se3cq_check(w) after synchronising to
 *     surface (and clear) device-side errors of _device calls.  A _host call
 *     reports its own launches' errors.
 *   - A world is immutable after se3cq_angle_model.  Queries on one world may
 *     come from several threads at once (the reference releases the GIL on
 *     them, lib/pyosr/module.cc:138-184): each _host call runs on a stream
 *     and staging buffers of its own.  Set-up calls must not overlap queries.
 *   - geo: 0 = environment (UnitWorld::GEO_ENV), 1 = robot (GEO_ROB).
 */
#ifndef SE3CQ_H
#define SE3CQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct se3cq_world se3cq_world;

#define SE3CQ_OK 0
#define SE3CQ_EINVAL (-1)
#define SE3CQ_ESTATE (-2)
#define SE3CQ_EIO (-3)
#define SE3CQ_EDEVICE (-4)
#define SE3CQ_ENOMEM (-5)

const char* se3cq_last_error(void);
const char* se3cq_version(void);

/* ---- world lifecycle ------------------------------------------------- */
/* UnitWorld::UnitWorld (lib/osr/unit_world.cc:65-70): identity perturbation. */
int se3cq_world_create(se3cq_world** out);
int se3cq_world_destroy(se3cq_world* w);
/* UnitWorld::copyFrom (unit_world.cc:76-92), pyosr "setupFrom". */
int se3cq_world_copy_from(se3cq_world* dst, const se3cq_world* src);

/* ---- geometry ingest and unit-world normalisation ---------------------- */
/* UnitWorld::loadModelFromFile / loadRobotFromFile (unit_world.cc:94-110) ->
 * Scene::load (lib/osr/scene.cc:78-156).  OBJ files only. */
int se3cq_load_file(se3cq_world* w, int geo, const char* path);
/* Same, from in-memory binary32 vertices [nv,3] and triangle indices [nf,3]. */
int se3cq_load_mesh(se3cq_world* w, int geo, const float* V, int64_t nv, const uint32_t* F, int64_t nf);
/* Scene::getCenter / overrideCenter (scene.h:63-65), UnitWorld::enforceRobotCenter (unit_world.cc:113-117). */
int se3cq_get_center(const se3cq_world* w, int geo, float* c3);
int se3cq_get_ompl_center(const se3cq_world* w, int geo, float* c3);
int se3cq_enforce_robot_center(se3cq_world* w, const double* c3);
/* UnitWorld::scaleToUnit (unit_world.cc:123-131). */
int se3cq_scale_to_unit(se3cq_world* w);
/* UnitWorld::angleModel (unit_world.cc:137-161): builds both BVHs and uploads
 * them to the current HIP device.  Both angles must be 0. */
int se3cq_angle_model(se3cq_world* w, float latitude, float longitude);
/* scene_scale property (module.cc:248). */
int se3cq_get_scene_scale(const se3cq_world* w, float* scale);
/* scene_matrix / robot_matrix properties (module.cc:249-250), row-major 4x4. */
int se3cq_get_matrix(const se3cq_world* w, int geo, double* m16);
/* Mesh sizes after ingest and the unit-world (collision) vertices,
 * binary64 [nv,3] (CDModel::addVF, lib/osr/cdmodel.cc:120-160). */
int se3cq_get_mesh_size(const se3cq_world* w, int geo, int64_t* nv, int64_t* nf);
int se3cq_get_mesh(const se3cq_world* w, int geo, float* V, uint32_t* F);
int se3cq_get_cd_vertices(const se3cq_world* w, int geo, double* V);
/* Depth of the engine's BVH over one model (diagnostics). */
int se3cq_get_bvh_info(const se3cq_world* w, int geo, int32_t* depth, int64_t* nnodes);
/* set_perturbation / perturbation (module.cc:135-136, unit_world.cc:164-175). */
int se3cq_set_perturbation(se3cq_world* w, const double* q7);
int se3cq_get_perturbation(const se3cq_world* w, double* q7);
/* recommended_cres property (module.cc:247). */
int se3cq_set_recommended_cres(se3cq_world* w, double cres);
int se3cq_get_recommended_cres(const se3cq_world* w, double* cres);

/* ---- the hot path ------------------------------------------------------ */
/* UnitWorld::isValid (unit_world.cc:215-247), batched: out[i] = 1 if state i
 * is collision free.  Without both models loaded every state is valid
 * (unit_world.cc:218-219). */
int se3cq_valid_states_host(se3cq_world* w, const double* q, int64_t n, uint8_t* out);
int se3cq_valid_states_device(se3cq_world* w, const double* d_q, int64_t n, uint8_t* d_out, void* stream);

/* isValid on the states a prescreen kept (no reference counterpart: it
 * replaces the samplers' isValid calls on accepted candidates, e.g.
 * src/GP/uw_random.py:46-52): for k < min(*d_count, n), with i = d_idx[k],
 * d_out[i] = 1 if state i is free.  Entries of d_out not listed are not
 * written; indices outside [0, n) are skipped.  The count is read on the
 * device, so a prescreen's output list needs no synchronisation. */
int se3cq_valid_states_subset_device(se3cq_world* w, const double* d_states, int64_t n, const int32_t* d_idx,
                                     const uint32_t* d_count, uint8_t* d_out, void* stream);

/* UnitWorld::transitStateToWithContact (unit_world.cc:405-474), batched over
 * n (from, to) rows with one verify_delta.  Outputs per row: last_free [7],
 * first colliding state [7], done (1 = free), last_tau and tau in binary64
 * (the reference narrows both to binary32 in its tuple), steps = isValid
 * calls made.  last_free/state/steps may be NULL. */
int se3cq_transit_host(se3cq_world* w, const double* from, const double* to, int64_t n, double verify_delta,
                       double* last_free, double* state, uint8_t* done, double* last_tau, double* tau,
                       int32_t* steps);
int se3cq_transit_device(se3cq_world* w, const double* d_from, const double* d_to, int64_t n,
                         double verify_delta, double* d_last_free, double* d_state, uint8_t* d_done,
                         double* d_last_tau, double* d_tau, int32_t* d_steps, void* stream);

/* UnitWorld::isValidTransition (unit_world.cc:477-486), batched. */
int se3cq_valid_transitions_host(se3cq_world* w, const double* from, const double* to, int64_t n,
                                 double initial_verify_delta, uint8_t* out);

int se3cq_valid_transitions_device(se3cq_world* w, const double* d_from, const double* d_to, int64_t n,
                                   double initial_verify_delta, uint8_t* d_out, int32_t* d_steps, void* stream);

/* isValidTransition on the segments a prescreen kept (unit_world.cc:477-486,
 * listed as se3cq_valid_states_subset_device lists states): d_out[i] (and
 * d_steps[i], may be NULL) for i = d_idx[k], k < min(*d_count, n). */
int se3cq_valid_transitions_subset_device(se3cq_world* w, const double* d_from, const double* d_to, int64_t n,
                                          const int32_t* d_idx, const uint32_t* d_count, double initial_verify_delta,
                                          uint8_t* d_out, int32_t* d_steps, void* stream);

/* Touch-configuration search: transitStateToWithContact followed, for rows
 * that hit an obstacle, by `levels` bisection halvings of [last_tau, tau]
 * (mid = 0.5*(lo+hi); lo = mid if interpolate(from,to,mid) is valid, else
 * hi = mid).  No reference counterpart: it extends the touch search of
 * src/GP/pipeline/touchq_util.py:8-27.  lo/hi: final bracket (free rows get
 * lo = last_tau, hi = 1).  steps counts isValid calls plus probes. */
int se3cq_touch_search_host(se3cq_world* w, const double* from, const double* to, int64_t n,
                            double verify_delta, int levels, uint8_t* done, double* last_tau, double* tau,
                            double* lo, double* hi, int32_t* steps);
int se3cq_touch_search_device(se3cq_world* w, const double* d_from, const double* d_to, int64_t n,
                              double verify_delta, int levels, uint8_t* d_done, double* d_last_tau,
                              double* d_tau, double* d_lo, double* d_hi, int32_t* d_steps, void* stream);

/* se3cq_touch_search_device without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * launch's scratch (its work counter and the longest-first order of its
 * segments) lives in one static array per device in the library's code
 * object (5 MiB), not in memory allocated for the call; n may be at most
 * SE3CQ_NOALLOC_MAX_SEGMENTS (larger n fails with SE3CQ_EINVAL).  The world
 * must already be on the current device (any earlier query on it uploads
 * it); otherwise the call fails with SE3CQ_ESTATE.  Device errors go to the
 * world's error word, as for the other _device calls (se3cq_check).  Every
 * call on one device shares that array: calls that overlap on one device
 * (other streams, other threads) race and give wrong results. */
#define SE3CQ_NOALLOC_MAX_SEGMENTS 1048576
int se3cq_touch_search_device_noalloc(se3cq_world* w, const double* d_from, const double* d_to, int64_t n,
                                      double verify_delta, int levels, uint8_t* d_done, double* d_last_tau,
                                      double* d_tau, double* d_lo, double* d_hi, int32_t* d_steps, void* stream);

/* Bisection alone on given brackets (one per row). */
int se3cq_bisect_contact_host(se3cq_world* w, const double* from, const double* to, int64_t n,
                              const double* lo_in, const double* hi_in, int levels, double* lo, double* hi);

/* UnitWorld::calculateVisibilityMatrix (unit_world.cc:740-762): int8 n x n,
 * symmetric, diagonal 0 (the reference leaves it unwritten). */
int se3cq_visibility_matrix_host(se3cq_world* w, const double* qs, int64_t n, int qs_are_unit,
                                 double verify_magnitude, int8_t* out);
/* calculateVisibilityMatrix2 (unit_world.cc:764-797): int8 m x n. */
int se3cq_visibility_matrix2_host(se3cq_world* w, const double* qs0, int64_t m, int qs0_are_unit,
                                  const double* qs1, int64_t n, int qs1_are_unit, double verify_magnitude,
                                  int8_t* out);
/* calculateVisibilityPair (unit_world.cc:799-840): row i of qs0 against row i
 * of qs1; both must have n rows. */
int se3cq_visibility_pair_host(se3cq_world* w, const double* qs0, int qs0_are_unit, const double* qs1,
                               int qs1_are_unit, int64_t n, double verify_magnitude, int8_t* out);
/* Bit-packed symmetric visibility matrix on the device: bit (i*n+j) of the
 * uint32 words d_bits (ceil(n*n/32) words, zeroed by the call) is 1 when the
 * segment i -> j (i < j, mirrored) is free.  States must be unit states. */
int se3cq_visibility_bits_device(se3cq_world* w, const double* d_qs, int64_t n, double verify_magnitude,
                                 uint32_t* d_bits, int32_t* d_steps, void* stream);

/* UnitWorld::isDisentangled (unit_world.cc:249-264 -> CDModel::collideBB,
 * cdmodel.cc:180-198): out[i] = 1 when the robot's and the environment's
 * bounding boxes (each model's AABB, placed by its transform) are disjoint at
 * state i.  Both models must be loaded (the reference throws otherwise). */
int se3cq_disentangled_host(se3cq_world* w, const double* q, int64_t n, uint8_t* out);
int se3cq_disentangled_device(se3cq_world* w, const double* d_q, int64_t n, uint8_t* d_out, void* stream);

/* UnitWorld::intersectingSegments (unit_world.cc:959-1012) for n states:
 * every (environment triangle, robot triangle) pair that intersects
 * (CDModel::collideForDetails, cdmodel.cc:201-229), and each pair's
 * intersection segment in world coordinates (Moller's triangle-triangle test
 * with end points, lib/tritri).  For state i, counts[i] is the number of
 * intersecting pairs; the first min(counts[i], cap) of them, in (env, robot)
 * order, are written to pairs[(i*cap + k)*2 + {0,1}], the segment end points
 * to segs[(i*cap + k)*6 + {0..5}] and flags[i*cap + k] (bit 0: the segment
 * test meets; bit 1: the triangles are coplanar, segment left at 0).  Which
 * pairs are listed when counts[i] > cap depends on the traversal order. */
int se3cq_contacts_host(se3cq_world* w, const double* q, int64_t n, int32_t cap, int32_t* pairs, double* segs,
                        int8_t* flags, int32_t* counts);

/* Surface device-side failures of earlier _device calls (after sync). */
int se3cq_check(se3cq_world* w);

/* ---- state arithmetic (host) ------------------------------------------ */
/* translateToUnitState (unit_world.cc:526-539) / translateFromUnitState (541-555). */
int se3cq_translate_to_unit(const se3cq_world* w, const double* q, int64_t n, double* out);
int se3cq_translate_from_unit(const se3cq_world* w, const double* q, int64_t n, double* out);
/* OMPL and "vanilla" state conventions (module.cc:159-167).  OMPL states
 * are [t, qx, qy, qz, qw] (w last) and place the robot by its OMPL centre
 * (the mean of its mesh vertices, before enforceRobotCenter).
 *   translateOMPLStateToUnitState (unit_world.cc:689-711)
 *   translateUnitStateToOMPLState (558-605); to_angle_axis writes
 *     [t, angle, axis] instead of a quaternion
 *   translateVanillaStateToOMPLState / translateOMPLStateToVanillaState (606-660)
 *   translateVanillaStateToUnitState (662-666)
 *   translateVanillaPointsToUnitPoints (668-687): pts [n, 3] */
int se3cq_translate_ompl_to_unit(const se3cq_world* w, const double* q, int64_t n, double* out);
int se3cq_translate_unit_to_ompl(const se3cq_world* w, const double* q, int64_t n, int to_angle_axis, double* out);
int se3cq_translate_vanilla_to_ompl(const se3cq_world* w, const double* q, int64_t n, double* out);
int se3cq_translate_ompl_to_vanilla(const se3cq_world* w, const double* q, int64_t n, double* out);
int se3cq_translate_vanilla_to_unit(const se3cq_world* w, const double* q, int64_t n, double* out);
int se3cq_translate_vanilla_pts_to_unit(const se3cq_world* w, int geo, const double* pts, int64_t n, double* out);

/* Mixed batches in one launch (no single reference counterpart: they
 * replace the samplers' and touch search's loops of isValid and
 * transitStateToWithContact calls, e.g. src/GP/uw_random.py:46-52 and
 * src/GP/touchq_util.py:8-27, with one device call).  Segments follow
 * se3cq_transit_* (unit_world.cc:405-474; outputs as there, last_free/state
 * optional, last_tau/tau/steps optional), states follow
 * se3cq_valid_states_* (valid[i] = 1 if free).  Persistent wavefronts take
 * segments and chunks of states from a work queue; the device form is one
 * fill kernel plus one kernel on the stream, so it can be captured in a HIP
 * graph.  A split batch (split = 1) runs the persistent segment kernel of
 * se3cq_transit_* and the state kernel of se3cq_valid_states_* instead,
 * which wins when the state traversal's LDS stacks would limit how many
 * segment wavefronts a CU holds; a batch without states always does.
 *
 * se3cq_batch_tune (batches with states) searches, one knob at a time and
 * keeping the best value of each before the next: the world's leaf size
 * (1, 2, 4), the launch shapes (the split form; the one-kernel form with
 * waves per block 1, 2, 4, resident blocks per CU all or half, states per
 * lane 1, 2, 4), the pool chunk (64, 128, 256) and the queue depth (1, 2,
 * 4) of se3cq_tuning.  Each candidate is timed over `repeats` launches on
 * the world's own stream, and its outputs are compared byte for byte on the
 * device with the first run's; a difference fails the call with
 * SE3CQ_EDEVICE (every candidate computes the same answers).  The fastest
 * configuration stays set (se3cq_batch_get_config, se3cq_get_tuning), the
 * outputs hold the batch's results, *ms = the time per launch of the
 * choice. */
typedef struct se3cq_batch_config {
	int32_t block_waves;     /* waves per block (1-4; 0 = as many as the LDS allows) */
	int32_t blocks_per_cu;   /* resident blocks per CU to launch (0 = as many as fit) */
	int32_t states_per_lane; /* states one lane checks per state work item (>= 1) */
	int32_t split;           /* 1: the segment and state kernels, one after the other on the stream;
	                          * 0: one kernel takes both kinds of work */
} se3cq_batch_config;
int se3cq_batch_device(se3cq_world* w, const double* d_seg_from, const double* d_seg_to, int64_t n_seg,
                       double verify_delta, double* d_last_free, double* d_state, uint8_t* d_done, double* d_last_tau,
                       double* d_tau, int32_t* d_steps, const double* d_states, int64_t n_states, uint8_t* d_valid,
                       void* stream);
int se3cq_batch_host(se3cq_world* w, const double* seg_from, const double* seg_to, int64_t n_seg, double verify_delta,
                     double* last_free, double* state, uint8_t* done, double* last_tau, double* tau, int32_t* steps,
                     const double* states, int64_t n_states, uint8_t* valid);
int se3cq_batch_tune(se3cq_world* w, const double* d_seg_from, const double* d_seg_to, int64_t n_seg,
                     double verify_delta, double* d_last_free, double* d_state, uint8_t* d_done, double* d_last_tau,
                     double* d_tau, int32_t* d_steps, const double* d_states, int64_t n_states, uint8_t* d_valid,
                     int32_t repeats, se3cq_batch_config* chosen, float* ms);
int se3cq_batch_get_config(const se3cq_world* w, se3cq_batch_config* cfg);
int se3cq_batch_set_config(se3cq_world* w, const se3cq_batch_config* cfg);

/* World-wide traversal knobs (no reference counterpart: FCL's BVH, built by
 * CDModel, lib/osr/cdmodel.cc:85-114, holds one triangle per leaf and the
 * reference has no work queue).  Results never depend on them.
 *   leaf_size   triangles per BVH leaf, at most (1-8, default 1); a change
 *               rebuilds both trees (about 2 s for 215k triangles) after
 *               waiting for the device.  A set-up call: no query of the
 *               world may run meanwhile, and HIP graphs captured before it
 *               must be captured again (a deeper tree than the one they were
 *               captured with can overflow their traversal stacks, which is
 *               reported as an error).
 *   pool_chunk  states a wavefront of the state kernels takes per counter
 *               add (16-1024, default 64).
 *   queue_depth segments (or batch items) a wavefront of the segment and
 *               batch kernels takes per counter add (1-16, default 1). */
typedef struct se3cq_tuning {
	int32_t leaf_size;
	int32_t pool_chunk;
	int32_t queue_depth;
} se3cq_tuning;
int se3cq_get_tuning(const se3cq_world* w, se3cq_tuning* t);
int se3cq_set_tuning(se3cq_world* w, const se3cq_tuning* t);

/* OMPL-side validity, the two queries of pyse3ompl's OmplDriver that do not
 * plan (lib/pyse3ompl/pyse3ompl.cc:77-81), over this world's geometry.
 * States are OMPL states (x, y, z, qx, qy, qz, qw) with unit quaternions.
 *   se3cq_ompl_longest_valid_segment: the bounding box diagonal plus pi/2,
 *     times the collision resolution (OmplDriver::set_bb / set_cdres,
 *     ompldriver.cc:676-685); the resolution must lie in (0, 1).
 *   validate_states (ompldriver.cc:566-589): out[i] = 1 if state i is free.
 *   validate_motions (ompldriver.cc:592-620): OMPL's discrete motion
 *     validator on (q0[i], q1[i]); out[i] = 1 if the end state and every
 *     interpolate(k / nd), 0 < k < nd, nd = ceil(distance / lvs), are free.
 *     The start state is not tested.  samples[i] (may be NULL) = max(nd, 1).
 *     The device form writes -1 for a motion whose nd is not finite below
 *     2^31; the host form fails with SE3CQ_EINVAL. */
int se3cq_ompl_longest_valid_segment(const double* lo, const double* hi, double resolution, double* out);
int se3cq_ompl_validate_states_host(se3cq_world* w, const double* q, int64_t n, int32_t* out);
int se3cq_ompl_validate_states_device(se3cq_world* w, const double* d_q, int64_t n, int32_t* d_out, void* stream);
int se3cq_ompl_validate_motions_host(se3cq_world* w, const double* q0, const double* q1, int64_t n, double lvs,
                                     int32_t* out, int32_t* samples);
int se3cq_ompl_validate_motions_device(se3cq_world* w, const double* d_q0, const double* d_q1, int64_t n, double lvs,
                                       int32_t* d_out, int32_t* d_samples, void* stream);

/* distance (lib/osr/osr_state.cc:130-154) and multi_distance (156-166). */
int se3cq_distance(const double* a7, const double* b7, double* out);
int se3cq_multi_distance(const double* origin7, const double* targets, int64_t n, double* out);
/* interpolate (osr_state.cc:27-44). */
int se3cq_interpolate(const double* a7, const double* b7, double tau, double* out7);
/* apply (osr_state.cc:243-259). */
int se3cq_apply(const double* from7, const double* tr3, const double* aa3, double* out7);
/* apply over n rows: out[i] = apply(from[i], tr[i], aa[i]), as the pipeline's
 * touch sampling calls it once per sample (src/GP/pipeline/touchq_util.py:8-13). */
int se3cq_apply_batch(const double* from, const double* tr, const double* aa, int64_t n, double* out);
/* decompose_2 (osr_state.cc:94-99): translation and rotation matrix (row-major). */
int se3cq_decompose_2(const double* q7, double* t3, double* R9);
/* decompose_3 (osr_state.cc:101-107, bound at lib/pyosr/module.cc:81-82):
 * translation, rotation angle in [0, pi] and unit axis of the normalised
 * quaternion (Eigen's AngleAxis of it). */
int se3cq_decompose_3(const double* q7, double* t3, double* angle, double* axis3);
/* compose_from_angleaxis (osr_state.cc:121-128). */
int se3cq_compose_from_angleaxis(const double* t3, double angle, const double* axis3, double* out7);

#ifdef __cplusplus
}
#endif

#endif

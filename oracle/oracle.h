/*
 * oracle.h -- CPU oracle for the SE(3) collision-query path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; it exists so tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py can check and time the GPU path against a
 * plain C restatement of the reference's algorithm.
 *
 * It restates, function by function (citations are to the reference tree):
 *   - mesh ingest and unit-world normalisation
 *       lib/osr/scene.cc:40-50,78-156,211-230, lib/osr/bbox.h:22-51,
 *       lib/osr/unit_world.cc:123-161, lib/osr/cdmodel.cc:120-160
 *   - state arithmetic (Eigen 3.3 formulas)
 *       lib/osr/osr_state.cc:16-44,85-154,243-259
 *   - FCL's BVH-vs-BVH collide on OBB nodes with the PQP 17-axis triangle
 *     separating-axis test (FCL >= 0.6, external; called from
 *     lib/osr/cdmodel.cc:162-178)
 *   - UnitWorld::isValid / transitStateToWithContact / isValidTransition /
 *     calculateVisibility{Matrix,Matrix2,Pair} / translateToUnitState
 *       lib/osr/unit_world.cc:215-247,405-486,526-539,713-724,740-840
 *   - a bisection refinement of the contact parameter (no reference
 *     counterpart; defined here as the contract for the library's
 *     bisect_contact)
 *   - a brute-force test over all triangle pairs that pins the BVH traversal.
 *
 * All arithmetic is binary64 with contraction off (-ffp-contract=off).
 */
#ifndef SE3_ORACLE_H
#define SE3_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrWorld OrWorld;

/* sin, cos, acos, atan2 of slerp, angularDistance and AngleAxis:
 * 0 = correctly rounded (what the reference platform's glibc 2.27 returns),
 * 1 = this host's glibc, 2 = as 0 except the slerp weights of the states a
 * walk only tests, which use the GPU kernels' faster sine. */
void or_set_libm(int mode);
int or_get_libm(void);

OrWorld* or_world_new(void);
void or_world_free(OrWorld* w);

/* Mesh ingest.  geo: 0 = environment, 1 = robot.  Vertices binary32 xyz. */
int or_load_arrays(OrWorld* w, int geo, const float* V, int64_t nv, const uint32_t* F, int64_t nf);
int or_load_obj(OrWorld* w, int geo, const char* path);
/* Parsed mesh sizes and copies (after ingest). */
int64_t or_mesh_nv(const OrWorld* w, int geo);
int64_t or_mesh_nf(const OrWorld* w, int geo);
void or_mesh_copy(const OrWorld* w, int geo, float* V, uint32_t* F);
void or_mesh_center(const OrWorld* w, int geo, float* c3);
void or_enforce_robot_center(OrWorld* w, const double* c3);
void or_scale_to_unit(OrWorld* w);
int or_angle_model(OrWorld* w, float latitude, float longitude);
float or_scene_scale(const OrWorld* w);
void or_set_perturbation(OrWorld* w, const double* q7);
/* Unit-world (collision) vertices, binary64, nv*3. */
void or_cd_vertices(const OrWorld* w, int geo, double* out);
/* Calibration matrix (row-major 4x4, binary64 of the binary32 matrix). */
void or_calibration(const OrWorld* w, double* m16);
int or_bvh_depth(const OrWorld* w, int geo);

/* State arithmetic. */
void or_state_to_transform(const double* q7, double* R9, double* t3);
double or_distance(const double* a7, const double* b7);
void or_interpolate(const double* a7, const double* b7, double tau, double* out7);
void or_apply(const double* from7, const double* tr3, const double* aa3, double* out7);
void or_translate_to_unit_state(const OrWorld* w, const double* q7, double* out7);
/* translateOMPLStateToUnitState (lib/osr/unit_world.cc:689-711). */
void or_translate_ompl_to_unit(const OrWorld* w, const double* q7, double* out7);
void or_ompl_center(const OrWorld* w, float* c3);

/* OMPL-side validity (OMPL 1.4.1's SE3StateSpace and discrete motion
 * validator, restated; states (x, y, z, qx, qy, qz, qw)). */
double or_ompl_distance(const double* a7, const double* b7);
double or_ompl_longest_segment(const double* lo3, const double* hi3, double resolution);
void or_ompl_interpolate(const double* from7, const double* to7, double t, double* out7);
int or_ompl_is_valid(const OrWorld* w, const double* q7);
int or_ompl_check_motion(const OrWorld* w, const double* s1, const double* s2, double lvs, int64_t* nd_out);
void or_ompl_validate_states(const OrWorld* w, const double* q, int64_t n, int32_t* out, int nthreads);
void or_ompl_validate_motions(const OrWorld* w, const double* q0, const double* q1, int64_t n, double lvs,
                              int32_t* out, int32_t* samples, int nthreads);

/* Triangle-triangle separating-axis test (FCL/PQP, 17 axes), binary64.
 * P: 3 env vertices; Q: 3 robot vertices already in the env frame. */
int or_tri_intersect(const double* P9, const double* Q9);

/* Validity.  Returns 1 if free (valid), 0 if colliding. */
int or_is_valid(const OrWorld* w, const double* q7);
int or_is_valid_counted(const OrWorld* w, const double* q7, int64_t* n_bv, int64_t* n_tri);
int or_is_valid_bruteforce(const OrWorld* w, const double* q7);
void or_valid_states(const OrWorld* w, const double* q, int64_t n, uint8_t* out,
                     int64_t* n_bv, int64_t* n_tri, int nthreads);

/* transitStateToWithContact.  Taus are returned in binary64 (the reference
 * narrows them to binary32 in its tuple); nsteps = isValid calls made. */
void or_transit(const OrWorld* w, const double* from7, const double* to7, double vd,
                double* last_free7, double* state7, int* done,
                double* last_tau, double* tau, int64_t* nsteps);
int or_is_valid_transition(const OrWorld* w, const double* from7, const double* to7,
                           double initial_vd, int64_t* nsteps);
/* Bisection between a free tau lo and a colliding tau hi, `levels` halvings. */
void or_bisect(const OrWorld* w, const double* from7, const double* to7,
               double lo, double hi, int levels, double* lo_out, double* hi_out);

/* Batched segment search: transit, then (if a collision was bracketed)
 * `levels` bisection probes.  Per segment outputs: done (1 = free),
 * last_tau/tau (binary64), contact lo/hi, steps evaluated (+ probes). */
void or_touch_batch(const OrWorld* w, const double* from, const double* to, int64_t n,
                    double vd, int levels, uint8_t* done, double* last_tau, double* tau,
                    double* lo, double* hi, int64_t* queries, int nthreads);

void or_visibility_matrix(const OrWorld* w, const double* qs, int64_t n, int is_unit,
                          double vd, int8_t* out, int nthreads);
void or_visibility_matrix2(const OrWorld* w, const double* qs0, int64_t m, int unit0,
                           const double* qs1, int64_t n, int unit1, double vd,
                           int8_t* out, int nthreads);
void or_visibility_pair(const OrWorld* w, const double* qs0, const double* qs1, int64_t n,
                        int unit0, int unit1, double vd, int8_t* out, int nthreads);

/* intersectingSegments: all intersecting triangle pairs at q by brute force,
 * (env, robot) order, first `cap` written with their world-frame segments
 * (flags bit 0: the segment test meets, bit 1: coplanar); returns the count. */
int or_contacts(const OrWorld* w, const double* q7, int cap, int32_t* pairs, double* segs, int8_t* flags);
int or_tri_isect_segment(const double* V9, const double* U9, int* coplanar, double* p1, double* p2);

/* isDisentangled: 1 if the two models' bounding boxes are disjoint at q,
 * 0 if they intersect, -1 if a model is missing. */
int or_is_disentangled(const OrWorld* w, const double* q7);
void or_disentangled_states(const OrWorld* w, const double* q, int64_t n, int8_t* out, int nthreads);
/* The box-box test alone (box i: row-major rotation, centre, half extents). */
int or_boxes_disjoint(const double* R1, const double* p1, const double* A, const double* R2, const double* p2,
                      const double* B);
/* A model's local AABB centre and half extents. */
void or_model_box(const OrWorld* w, int geo, double* c3, double* h3);

/* The library's portable transcendental functions (se3_math.h). */
double or_se3_sin(double x);
double or_se3_cos(double x);
double or_se3_acos(double x);
double or_se3_atan2(double y, double x);

#ifdef __cplusplus
}
#endif

#endif

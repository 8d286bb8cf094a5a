/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.  CPU restatement of the reference's
 * AVBD step (LuckyIYI/avbd-metal) used as the parity checker for the HIP
 * backend and as the CPU baseline.  Nothing in the product links this.
 *
 * Serial, templated on float / double.  It consumes the same scene wire
 * format as the backend (include/avbd.h) so both start from identical inputs.
 */
#ifndef AVBD_ORACLE_H
#define AVBD_ORACLE_H
#include <stdint.h>
#include "../include/avbd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_solver oracle_solver;

/* precision: 32 or 64 */
int oracle_create(const avbd_scene_desc* scene, int precision, int max_pairs_per_body, oracle_solver** out);
void oracle_destroy(oracle_solver* o);
/* split each color's primal over `threads` threads (identical results) */
int oracle_set_threads(oracle_solver* o, int threads);
int oracle_step(oracle_solver* o);
/* one frame iterated until the residual <= tol (chunks of check_every sweeps, at most max_iterations) */
int oracle_step_to_residual(oracle_solver* o, double tol, int max_iterations, int check_every, int* iterations,
                            double* residual);
int oracle_set_settings(oracle_solver* o, const avbd_settings* s);
int oracle_get_body_states(oracle_solver* o, uint32_t n, float* pos3, float* rot4, float* vel3, float* angvel3);
int oracle_get_body_states_f64(oracle_solver* o, uint32_t n, double* pos3, double* rot4, double* vel3,
                               double* angvel3);
int oracle_set_body_states(oracle_solver* o, const uint32_t* idx, uint32_t n, const float* pos3, const float* rot4,
                           const float* vel3, const float* angvel3);
int oracle_get_colors(oracle_solver* o, uint32_t n, uint32_t* colors);
int oracle_get_cell_lists(oracle_solver* o, uint32_t* hash_size, uint32_t* num_hashed, float* cell_size,
                          uint32_t* cell_start, uint32_t* cell_count, uint32_t* cell_bodies);
int oracle_get_pairs(oracle_solver* o, uint32_t capacity, uint32_t* n, uint32_t* pairs);
/* manifold summary of the last frame: per pair (bodyA, bodyB, numContacts) and contact normals */
int oracle_get_manifold_records(oracle_solver* o, uint32_t capacity, uint32_t* n, void* records704);
int oracle_get_manifold_keys(oracle_solver* o, uint32_t capacity, uint32_t* n, uint32_t* keys8);
/* isotropic limiter state after the last frame: tracked triangles [4 nb],
 * tracked edges [4 num_contact_edges], squared clearance bounds [nb], anchor
 * poses [3 nb], the frame's truncations and refreshes [2]; -1 when the mode
 * is off */
int oracle_get_iso_state(oracle_solver* o, uint32_t* vt_track, uint32_t* ee_track, float* bound, float* anchor,
                         uint32_t* counts);
int oracle_get_manifolds(oracle_solver* o, uint32_t capacity, uint32_t* n, uint32_t* header3, float* normal3,
                         float* lambda_pen);
int oracle_residual(oracle_solver* o, double* out);
int oracle_stats(oracle_solver* o, int32_t* num_pairs, int32_t* active_manifolds, int32_t* num_colors,
                 int32_t* pair_candidates, int32_t* color_conflicts);
int oracle_map_find(oracle_solver* o, const uint32_t* pairs, uint32_t n, int32_t* out);
/* cloth: derived topology / parameters, the V-T / E-E pair stream (kind, owner, other),
 * element contacts of the last frame (112-byte records as the backend's) and counts */
int oracle_get_cloth_info(oracle_solver* o, avbd_cloth_info* info, uint32_t* edges, uint32_t* particles);
int oracle_get_collision_tris(oracle_solver* o, uint32_t capacity, uint32_t* n, uint32_t* tris);
int oracle_get_motor_targets(oracle_solver* o, uint32_t n, float* targets);
int oracle_get_dat_pairs(oracle_solver* o, uint32_t capacity, uint32_t* n, uint32_t* pairs3);
int oracle_get_soft_contacts(oracle_solver* o, uint32_t capacity, uint32_t* n, void* records);
int oracle_dat_pairs_at(oracle_solver* o, const float* pos3, uint32_t capacity, uint32_t* n, uint32_t* pairs3);
int oracle_cloth_stats(oracle_solver* o, int32_t* dat_pairs, int32_t* dat_vt, int32_t* soft_candidates,
                       int32_t* soft_rt, int32_t* overflow, int32_t* nonfinite);
/* Planar-DAT truncation counters of the last frame: truncations, invalid
 * anchors, V-T and E-E degeneracies, non-finite values, grid overflows. */
int oracle_dat_stats(oracle_solver* o, int32_t* out6);
/* Test hooks of the Planar-DAT fixtures: query radius override (> 0), and
 * pair / soft-contact capacity caps (> 0), as the backend's. */
/* setJointWorldAnchors: world points of world-anchored joints. */
int oracle_set_joint_world_anchors(oracle_solver* o, const uint32_t* joints, uint32_t n, const float* pts);
int oracle_set_dat_test_hooks(oracle_solver* o, float query_radius, int32_t pair_capacity, int32_t soft_capacity);

/* unit restatements for pinning */
void oracle_splitmix64(uint64_t seed, uint32_t n, uint64_t* out_u64, float* out_float);
/* SNH tet: positions x[12], DmInv rows r[9], mu, lambda, vol, self slot (0..3):
 * energy, gradient on self (3) and the per-vertex Hessian approximation (9) */
void oracle_snh_tet(const double* x, const double* dminv, double mu, double lam, double vol, int self, double* energy,
                    double* grad3, double* hess9);
void oracle_solve6x6(const double* A36, const double* b6, double* x6, int precision);
void oracle_membrane(const double* x9, const double* dm4, double muA, double lamA, int self, double* energy,
                     double* grad3, double* hess9);
int oracle_hinge(const float* x12, float k1, float k2, float* K4, float* s);
uint32_t oracle_cell_hash(float px, float py, float pz, float cell_size, uint32_t group, uint32_t hash_size);

/* The reference's CPU AVBD path (CPUSolver semantics, cpu_path.cpp): rigid
 * boxes / spheres, joints and springs; returns 4 for scenes outside it. */
typedef struct oracle_cpu_solver oracle_cpu_solver;
int oracle_cpu_create(const avbd_scene_desc* scene, int precision, oracle_cpu_solver** out);
void oracle_cpu_destroy(oracle_cpu_solver* o);
int oracle_cpu_step(oracle_cpu_solver* o, int n);
int oracle_cpu_get_body_states_f64(oracle_cpu_solver* o, uint32_t n, double* pos3, double* rot4, double* vel3,
                                   double* angvel3);
int oracle_cpu_num_manifolds(oracle_cpu_solver* o);
double oracle_cpu_max_constraint_error(oracle_cpu_solver* o);

#ifdef __cplusplus
}
#endif
#endif

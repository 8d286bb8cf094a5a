/*
 * avbd.h -- C ABI of the MI355X (gfx950) AVBD step backend.
 *
 * This is the drop-in boundary for the per-step AVBD solve of
 * LuckyIYI/avbd-metal.  Every entry point below replaces one member of the
 * reference `GPUSolver` class (Sources/PhysicsAVBD/GPU/GPUSolver.swift); the
 * citation above each declaration names the member it stands in for.
 *
 * Conventions
 *  - Plain C: pointers + sizes, no C++ or device types cross the boundary.
 *  - The library owns all device memory.  Callers hand over and receive host
 *    arrays (copy-in / copy-out), exactly like the reference's shared-buffer
 *    accessors after `sync()`.
 *  - Every function returns an `avbd_status` (0 = OK) and never aborts.
 *  - Runtime failures are latched sticky, earliest first: once a solver
 *    reports a failure it stays failed and must be destroyed
 *    (GPUSolver.swift:2221-2223 "Rebuild the solver; never clear and
 *    continue").
 *  - A solver handle is not thread-safe; use one handle per thread.
 *  - Vectors are packed float arrays: 3 floats per F3, 4 per quaternion
 *    stored (x, y, z, w) with w the real part (GPUSolver.swift:952).
 */
#ifndef AVBD_H
#define AVBD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AVBD_ABI_VERSION 6
#define AVBD_MAX_COLORS 64

/* ---- status codes ------------------------------------------------------ */
typedef enum avbd_status {
    AVBD_OK = 0,
    AVBD_ERR_INVALID_ARGUMENT = 1,  /* precondition violated (Scene.swift preconditions) */
    AVBD_ERR_NO_DEVICE = 2,         /* AVBDError.noDevice          GPUSolver.swift:641 */
    AVBD_ERR_ALLOC = 3,             /* AVBDError.allocFailed       GPUSolver.swift:642 */
    AVBD_ERR_UNSUPPORTED = 4,       /* request outside the scene's configuration (snapshots of soft scenes; isotropic state with the mode off) */
    AVBD_ERR_RUNTIME_FAILURE = 5,   /* a RuntimeFailure is latched; see avbd_last_failure */
    AVBD_ERR_HIP = 6,               /* HIP API error outside a latched frame */
    AVBD_ERR_TOO_LARGE = 7          /* request exceeds a fixed capacity (avbd_set_body_states_noalloc) */
} avbd_status;

/* RuntimeFailure kinds, GPUSolver.swift:663-758 */
typedef enum avbd_failure_kind {
    AVBD_FAIL_NONE = 0,
    AVBD_FAIL_COMMAND_EXECUTION = 1,     /* .commandExecution (HIP launch / completion error) */
    AVBD_FAIL_RIGID_PAIR_CAPACITY = 2,   /* .rigidPairCapacity */
    AVBD_FAIL_SOFT_CONTACT_CAPACITY = 3, /* .softContactCapacity */
    AVBD_FAIL_STATIC_COLOR_CAPACITY = 4, /* .staticColorCapacity (thrown by create) */
    AVBD_FAIL_UNRESOLVED_COLORING = 5,   /* .unresolvedColoring */
    AVBD_FAIL_STAGE_DENIED = 6,          /* .commandEncoderCreation (test hook) */
    AVBD_FAIL_BARRIER_TIMEOUT = 7,       /* persistent kernel grid barrier exceeded its spin bound */
    AVBD_FAIL_DAT_PAIR_CAPACITY = 8,     /* .planarDATPairCapacity (commandExecution, PlanarDAT code 1) */
    AVBD_FAIL_DAT_GRID_SPAN = 9,         /* .planarDATElementGridSpan (PlanarDAT code 2) */
    AVBD_FAIL_DAT_INVALID_ANCHOR = 10    /* .planarDATInvalidAnchor, non-finite anchors (PlanarDAT code 3) */
} avbd_failure_kind;

typedef struct avbd_failure {
    int32_t kind;          /* avbd_failure_kind */
    int64_t frame;         /* frame the failure belongs to (-1 at init) */
    int64_t required;      /* demand (pairs, colors, conflicting bodies ...) */
    int64_t capacity;      /* capacity that was exceeded */
    int32_t hip_error;     /* hipError_t for COMMAND_EXECUTION, else 0 */
    char message[160];
} avbd_failure;

/* ---- scene wire format ------------------------------------------------- */

/* SimSettings, Scene.swift:485-545 (defaults: avbd_settings_default) */
typedef struct avbd_settings {
    float dt;                     /* 1/60 */
    float gravity;                /* -10, along z */
    int32_t iterations;           /* 10 */
    float collision_margin;       /* 0.01 */
    uint32_t friction_combine_mode; /* 0 geometric .. 4 average */
    float alpha;                  /* 0.99 */
    float beta_lin;               /* 1e4 */
    float beta_ang;               /* 100 */
    float gamma;                  /* 0.999 */
    float lambda_max;             /* 1e6 */
    float rod_decay_pow;          /* 0 */
    float particle_damping;       /* 0 */
    float rigid_linear_damping;   /* 0 */
    float rigid_angular_damping;  /* 0 */
    float cloth_viscosity;        /* 0 */
} avbd_settings;

/* BodyShape, Scene.swift:6-11 */
enum { AVBD_SHAPE_BOX = 0, AVBD_SHAPE_SPHERE = 1, AVBD_SHAPE_TORUS = 2, AVBD_SHAPE_CAPSULE = 3 };

/*
 * Packed SoA mirror of PhysicsScene (Scene.swift:547-752).  Optional arrays
 * may be NULL and then take the Scene.swift default.  All upload-time
 * precompute (mass/inertia, Dm^-1, topology, coloring) happens inside the
 * library, as GPUSolver.upload does (GPUSolver.swift:901-2081).
 */
typedef struct avbd_scene_desc {
    uint32_t abi_version;                 /* AVBD_ABI_VERSION */
    avbd_settings settings;

    /* SceneBody, Scene.swift:34-113 */
    uint32_t num_bodies;
    const float* body_size;               /* [n*3] */
    const float* body_density;            /* [n]   0 = static */
    const float* body_friction;           /* [n]   static friction */
    const float* body_dynamic_friction;   /* [n]   NULL -> friction */
    const float* body_position;           /* [n*3] */
    const float* body_rotation;           /* [n*4] NULL -> identity */
    const float* body_velocity;           /* [n*3] NULL -> 0 */
    const uint32_t* body_shape;           /* [n]   AVBD_SHAPE_* */
    const uint8_t* body_is_particle;      /* [n]   NULL -> 0 */
    const float* body_mass;               /* [n]   NULL or NaN -> derive from density */
    const float* body_inertia;            /* [n*3] used where body_mass is given */
    const float* body_gravity_scale;      /* [n]   NULL -> 1 */

    /* SceneCollider, Scene.swift:121-184 */
    uint32_t num_colliders;
    const uint32_t* collider_body;        /* [m] */
    const float* collider_size;           /* [m*3] */
    const float* collider_friction;       /* [m] */
    const float* collider_dynamic_friction; /* [m] NULL -> friction */
    const float* collider_local_position; /* [m*3] NULL -> 0 */
    const float* collider_local_rotation; /* [m*4] NULL -> identity */
    const uint32_t* collider_shape;       /* [m] AVBD_SHAPE_* */
    const uint32_t* collider_group;       /* [m] NULL -> 0 */
    const uint8_t* collider_shared;       /* [m] collidesWithSharedGeometry, NULL -> 1 */
    const uint8_t* collider_enabled;      /* [m] collisionEnabled, NULL -> 1 */
    const uint8_t* collider_world_round_anchor; /* [m] usesWorldSpaceRoundAnchor, NULL -> 0 */

    /* SceneJoint, Scene.swift:206-299 */
    uint32_t num_joints;
    const int32_t* joint_body_a;          /* [j] -1 = world */
    const int32_t* joint_body_b;          /* [j] */
    const float* joint_ra;                /* [j*3] */
    const float* joint_rb;                /* [j*3] */
    const float* joint_stiffness_lin;     /* [j] +inf = hard */
    const float* joint_stiffness_ang;     /* [j] +inf = hard, 0 = free */
    const float* joint_fracture;          /* [j] +inf = unbreakable */
    const uint8_t* joint_fracture_linear; /* [j] NULL -> 0 */
    const float* joint_hinge_axis;        /* [j*3] NaN x -> no hinge */
    const float* joint_motor;             /* [j*6] target, torque limit, kp, kd, armature, mode(0 implicit PD,1 explicit PD,2 velocity) ; NULL -> off */
    const float* joint_limits;            /* [j*2] lo, hi ; NULL -> (1,-1) disabled */

    /* SceneSpring, Scene.swift:433-452 */
    uint32_t num_springs;
    const uint32_t* spring_body_a;
    const uint32_t* spring_body_b;
    const float* spring_ra;               /* [s*3] */
    const float* spring_rb;               /* [s*3] */
    const float* spring_stiffness;        /* [s] */
    const float* spring_rest;             /* [s] <0 -> spawn distance */
    const uint8_t* spring_hard;           /* [s] */

    /* SceneTet, Scene.swift:337-361 */
    uint32_t num_tets;
    const uint32_t* tet_ids;              /* [t*4] */
    const float* tet_mu;                  /* [t] */
    const float* tet_lambda;              /* [t] */

    /* SceneTri, Scene.swift:321-334 */
    uint32_t num_tris;
    const uint32_t* tri_ids;              /* [r*3] */
    const float* tri_mu;
    const float* tri_lambda;
    const float* tri_bend;

    /* SceneCollisionExclusion, Scene.swift:466-475 */
    uint32_t num_exclusions;
    const uint32_t* exclusion_pairs;      /* [e*2] */

    /* SceneCollider.convexHullVertices, Scene.swift:127: per collider a
     * vertex count (0 = its primitive shape, at most 64), then the vertices
     * of every hull collider in collider order, collider-local.  A hull
     * collides with boxes only (30_narrowphase.metal:404-410).  NULL -> none. */
    const uint32_t* collider_hull_count;  /* [m] */
    const float* hull_vertices;           /* [sum(count)*3] */
    /* legacy servo rate (rad/s) of implicit-PD hinge motors: before every
     * frame the target advances by rate*dt, wrapped to [-pi, pi]
     * (SceneJoint.motorRate, Scene.swift:300-314; GPUSolver.swift:1154-1158,
     * 4359-4416); NULL -> 0 */
    const float* joint_motor_rate;        /* [j] */
    /* SceneSpinner, Scene.swift:459-469: kinematic spinners.  Before every
     * frame the body's orientation turns by omega*dt about the unit axis
     * (GPUSolver.swift:4406-4430), and contact targets against it move
     * with its surface (30_narrowphase.metal:270-284). */
    uint32_t num_spinners;
    const uint32_t* spinner_body;         /* [k] */
    const float* spinner_axis_omega;      /* [k*4] unit axis xyz, omega (rad/s) */
} avbd_scene_desc;

/* Scene files: a versioned on-disk form of avbd_scene_desc (no reference
 * counterpart; for fixtures and for large scenes that should load without
 * being rebuilt).  avbd_scene_write stores every present array with the
 * settings and counts; avbd_scene_read checks the magic, the file version
 * and every array's id, element size and length against the counts, and
 * returns a descriptor owned by the library (release it with
 * avbd_scene_free; avbd_create copies what it needs).  Unreadable,
 * truncated or inconsistent files give AVBD_ERR_INVALID_ARGUMENT. */
#define AVBD_SCENE_FILE_VERSION 2
avbd_status avbd_scene_write(const avbd_scene_desc* d, const char* path);
avbd_status avbd_scene_read(const char* path, avbd_scene_desc** out);
void avbd_scene_free(avbd_scene_desc* d);

/* Solver creation options (GPUSolver.init maxPairsPerBody, GPUSolver.swift:331) */
typedef struct avbd_options {
    int32_t max_pairs_per_body;        /* 16; maxPairs = max(4096, colliders * this), negative -> 16 */
    int32_t lanes_per_body;            /* 0 = autotune; else 4, 8, 16 or 32 lanes per body in the primal sweep */
    int32_t sweep_mode;                /* 0 = autotune, 1 = one launch per color, 2 = persistent cooperative sweep */
    int32_t use_graph;                 /* 1 = capture each step in a hipGraph (default), 0 = eager launches */
    int32_t pair_capacity_override;    /* test hook (planarDATPairCapacityForTesting analogue): >0 caps rigid pairs */
    int32_t dat_pair_capacity_override;  /* test hook, planarDATPairCapacityForTesting (GPUSolver.swift:2119): >0 caps the V-T/E-E stream */
    int32_t soft_capacity_override;      /* test hook, planarDATSoftCapacityForTesting (GPUSolver.swift:2128): >0 caps soft contacts */
} avbd_options;

typedef struct avbd_solver avbd_solver;

/* ---- lifecycle ----------------------------------------------------------- */

/* Fills Scene.swift:485-545 defaults. */
void avbd_settings_default(avbd_settings* out);
/* Fills avbd_options defaults. */
void avbd_options_default(avbd_options* out);

/* GPUSolver.init(scene:device:maxPairsPerBody:) throws -- GPUSolver.swift:331-338.
 * Builds device buffers, runs the upload precompute and the static coloring.
 * On AVBD_ERR_RUNTIME_FAILURE (static palette > 64 colors, GPUSolver.swift:2008)
 * *out is still set so avbd_last_failure can be read; destroy it afterwards. */
avbd_status avbd_create(const avbd_scene_desc* scene, const avbd_options* opts,
                        avbd_solver** out);
/* deinit of GPUSolver */
void avbd_destroy(avbd_solver* s);

/* ---- stepping ------------------------------------------------------------ */

/* GPUSolver.submitStep() throws -- GPUSolver.swift:3107.  Enqueues one frame
 * (at most 2 frames in flight, GPUSolver.swift:3059-3074). */
avbd_status avbd_submit_step(avbd_solver* s);
/* GPUSolver.step() -- GPUSolver.swift:3095 (submit + no fatal wrapper: returns status). */
avbd_status avbd_step(avbd_solver* s);
/* GPUSolver.synchronize() throws -- GPUSolver.swift:2924.  Retires all
 * frames, publishes stats, surfaces the sticky failure. */
avbd_status avbd_synchronize(avbd_solver* s);

/* mutable `settings: SimSettings` -- GPUSolver.swift:45 / syncParams :2097 */
avbd_status avbd_get_settings(const avbd_solver* s, avbd_settings* out);
avbd_status avbd_set_settings(avbd_solver* s, const avbd_settings* in);

/* ---- state readback / reset ---------------------------------------------- */

/* bodyPosition/Rotation/Velocity/AngularVelocity, GPUSolver.swift:4431-4469,
 * bodyStates(_:) :4471.  idx NULL -> bodies 0..n-1.  Any output may be NULL. */
avbd_status avbd_get_body_states(avbd_solver* s, const uint32_t* idx, uint32_t n,
                                 float* pos3, float* rot4, float* vel3, float* angvel3);
/* bodyMass / bodyDiagonalInertia, GPUSolver.swift:4443-4455 */
avbd_status avbd_get_body_mass(avbd_solver* s, uint32_t n, float* mass, float* inertia3);
/* setBodyStates(_:) -- GPUSolver.swift:2409-2473, including the warm-start
 * clearing of incident joints, springs and manifolds (:2433-2472). */
avbd_status avbd_set_body_states(avbd_solver* s, const uint32_t* idx, uint32_t n,
                                 const float* pos3, const float* rot4,
                                 const float* vel3, const float* angvel3);

/* No counterpart in the reference.  avbd_set_body_states for callers that
 * cannot allocate device memory (code under HIP graph capture, tight RL
 * reset loops, harnesses that forbid hipMalloc): identical effect, but the
 * staging lives in a fixed 4 MiB array per GPU (resolved at the first
 * avbd_create on that device) and nothing is allocated or freed.  Capacity:
 * 14 * n + (bodies in the solver) <= 1048576 words, else AVBD_ERR_TOO_LARGE.
 * The array is shared by every solver on the GPU and keeps the last call's
 * contents: calls that overlap on one GPU (other threads) race. */
avbd_status avbd_set_body_states_noalloc(avbd_solver* s, const uint32_t* idx, uint32_t n,
                                         const float* pos3, const float* rot4, const float* vel3,
                                         const float* angvel3);
/* setBodyPoses(_:) -- GPUSolver.swift:2399-2404: setBodyStates with zero velocities. */
avbd_status avbd_set_body_poses(avbd_solver* s, const uint32_t* idx, uint32_t n, const float* pos3,
                                const float* rot4);
/* setMotorTargets / setMotorVelocities / setMotorTorques -- GPUSolver.swift:2496-2590.
 * The reference's preconditions return AVBD_ERR_INVALID_ARGUMENT (nothing is
 * written): a joint index out of range or a non-finite value; targets need a
 * position-PD motor (joints without a motor ignore them); velocities need a
 * velocity motor; effort limits need an authored motor and a value >= 0. */
avbd_status avbd_set_motor_targets(avbd_solver* s, const uint32_t* joints, uint32_t n,
                                   const float* target);
avbd_status avbd_set_motor_velocities(avbd_solver* s, const uint32_t* joints, uint32_t n,
                                      const float* rad_per_s);
avbd_status avbd_set_motor_torques(avbd_solver* s, const uint32_t* joints, uint32_t n,
                                   const float* torque);
/* setJointWorldAnchors(_:) -- GPUSolver.swift:2481-2490: moves the world point
 * of world-anchored joints (point3[3*n]). */
avbd_status avbd_set_joint_world_anchors(avbd_solver* s, const uint32_t* joints, uint32_t n,
                                         const float* point3);
/* motorStates(_:) -- GPUSolver.swift:2605-2637: hinge twist angle (2 atan2 of
 * the rest-relative rotation about the axis) and the relative angular
 * velocity projected on the child's world hinge axis. */
avbd_status avbd_get_motor_states(avbd_solver* s, const uint32_t* joints, uint32_t n, float* angle,
                                  float* velocity);
/* Motor targets (JointGPU.motor.x) of joints [0, n): motorTargetForTesting. */
avbd_status avbd_get_motor_targets(avbd_solver* s, uint32_t n, float* targets);
/* debugBrokenJoints / debugMaxLambda -- GPUSolver.swift:2652-2672 */
avbd_status avbd_get_joint_diagnostics(avbd_solver* s, int32_t* broken, float* max_lambda_lin,
                                       float* max_lambda_ang);

/* ---- stats / debug ------------------------------------------------------- */

/* lastNumPairs/lastPairCandidates/lastNumSoft/lastColorCounts/lastMaxColorUsed/
 * frameIndex -- GPUSolver.swift:314-328, published by retire :2829-2880. */
typedef struct avbd_stats {
    int64_t frame_index;
    int32_t num_pairs;
    int32_t pair_candidates;
    int32_t num_soft;
    int32_t soft_candidates;
    int32_t num_active_manifolds;
    int32_t max_color_used;
    int32_t num_colors;
    int32_t color_counts[AVBD_MAX_COLORS];
    int32_t color_conflicts;
    int32_t uses_dynamic_coloring;
    int32_t lanes_per_body;        /* primal lane split in use */
    int32_t sweep_mode;            /* 1 per-color launches, 2 persistent */
    int32_t dat_pairs;             /* lastPlanarDATPairs: V-T + E-E pair demand */
    int32_t dat_vt_pairs;          /* lastPlanarDATVertexTrianglePairs */
    int32_t dat_ee_pairs;          /* lastPlanarDATEdgeEdgePairs */
    int32_t num_soft_rt;           /* rigid-feature vs triangle records (prefix of the soft array) */
    int32_t ws_accepted;           /* learned warm start kept (1) or fell back to the plain guess (0) */
    float ws_residual_plain;       /* residual at the plain adaptive guess (learned stage active) */
    float ws_residual_learned;     /* residual at the learned start */
    int32_t dat_truncations;       /* lastPlanarDATTruncations: particles truncated by the Planar-DAT passes */
    int32_t dat_invalid_anchors;   /* invalid Planar-DAT anchors (any > 0 latches AVBD_FAIL_DAT_INVALID_ANCHOR) */
    int32_t dat_vt_degeneracies;   /* lastPlanarDATVertexTriangleDegeneracies */
    int32_t dat_ee_degeneracies;   /* lastPlanarDATEdgeEdgeDegeneracies */
    int32_t dat_nonfinite;         /* lastPlanarDATNonfiniteValues */
} avbd_stats;
avbd_status avbd_get_stats(avbd_solver* s, avbd_stats* out);

/* debugColors() -- GPUSolver.swift:2674.  colors[n] */
avbd_status avbd_get_colors(avbd_solver* s, uint32_t n, uint32_t* colors);

/* Spatial hash of the last frame (bp_count/scan/scatter/sort, 20_broadphase.metal:69-131):
 * cell_start/cell_count [hash_size], cell_bodies [num_hashed] (collider ids sorted per cell).
 * Pass NULL arrays to query sizes. */
avbd_status avbd_get_cell_lists(avbd_solver* s, uint32_t* hash_size, uint32_t* num_hashed,
                                float* cell_size, uint32_t* cell_start, uint32_t* cell_count,
                                uint32_t* cell_bodies);
/* Rigid contact-pair stream of the last frame (bp_emit_pairs_deterministic,
 * 20_broadphase.metal:457): pairs[2*n] collider ids in emission order. */
avbd_status avbd_get_pairs(avbd_solver* s, uint32_t capacity, uint32_t* n, uint32_t* pairs);

/* Manifolds of the last frame (ManifoldGPU, GPUTypes.swift:164): raw 704-byte
 * records, `capacity` records at most. */
avbd_status avbd_get_manifolds(avbd_solver* s, uint32_t capacity, uint32_t* n, void* records);

/* Residual-driven step: runs one frame whose iteration loop stops once the
 * residual r = |g| / |m g_grav| over dynamic bodies (the block right-hand
 * side of primal_one, 40_solver.metal:1384-1425, at the current iterate)
 * reaches `tol`, checking every `check_every` sweeps, at most
 * `max_iterations`.  No reference counterpart (the reference always runs
 * settings.iterations sweeps, GPUSolver.swift:3138-4405); it measures the
 * "step to residual 1e-4" rate.  Completes and retires the frame. */
avbd_status avbd_step_to_residual(avbd_solver* s, float tol, int32_t max_iterations, int32_t check_every,
                                  int32_t* iterations_used, float* residual);

/* Collision-surface topology derived at create (GPUSolver.swift:384-390,
 * 1289-1640): counts of collision triangles (the authored ones, then the
 * boundary faces of the tet meshes), unique edges, surface particles,
 * membrane and bend elements, and the grid / query parameters.
 * num_contact_edges: the edges of authored triangles, a prefix of edges[]
 * (E-E contacts; the tet-boundary edges after it only bound motion);
 * pair_stream: 1 when the V-T / E-E pair stream and its contacts run (with
 * the Planar-DAT limiter); dat_on: 1 when the Planar-DAT limiter runs (a
 * self-colliding, i.e. authored, surface and no tets); iso_on: 1 when the isotropic OGC limiter and its tracked V-T /
 * E-E contacts run (tets and either two or more surfaces or an authored
 * one) -- the reference's automatic selection, GPUSolver.swift:93-99.
 * Any output pointer may be NULL;
 * edges[2*num_edges] (ascending vertex ids, edge order of the pair stream),
 * particles[num_particles]. */
typedef struct avbd_cloth_info {
    uint32_t num_tris, num_edges, num_particles, num_membranes, num_bends;
    uint32_t elem_hash_size, max_soft, soft_map_capacity, max_dat_pairs;
    float elem_cell_size, elem_margin, dat_query_radius, surface_contact_cell_size;
    uint32_t num_contact_edges, pair_stream, dat_on, iso_on;
} avbd_cloth_info;
avbd_status avbd_get_cloth_info(avbd_solver* s, avbd_cloth_info* info, uint32_t* edges, uint32_t* particles);

/* Collision triangles (trisBuf, GPUSolver.swift:1289-1310): tris[3*n], the
 * authored ones in scene order, then the tet boundary faces wound outward in
 * ascending (a, b, c) order (tetBoundaryFaces, GPUSolver.swift:872-897). */
avbd_status avbd_get_collision_tris(avbd_solver* s, uint32_t capacity, uint32_t* n, uint32_t* tris);

/* Planar-DAT pair stream of the last frame (dat_build_vt_pairs /
 * dat_build_ee_pairs, 27_planar_dat.metal:35-229), the query at the accepted
 * predictor pose that feeds the contacts and the per-color truncation
 * passes: pairs[3*n] = (kind 0 V-T | 1 E-E, owner vertex / edge, triangle /
 * edge), producer-major with the partner ascending.  avbd_stats.dat_pairs
 * is the larger demand of the frame's two queries. */
avbd_status avbd_get_dat_pairs(avbd_solver* s, uint32_t capacity, uint32_t* n, uint32_t* pairs);

/* Element contacts of the last frame (SoftContactGPU, GPUTypes.swift:127-136):
 * raw 112-byte records -- ids u32[4], normal+friction, anchorA+flags, weights,
 * C0+lambdaCap, lambda+keyA, penalty+keyB (float4 each). */
avbd_status avbd_get_soft_contacts(avbd_solver* s, uint32_t capacity, uint32_t* n, void* records);

/* No counterpart in the reference.  Wall time of avbd_create's phases
 * (host clock; each phase's device work is drained before it closes), in
 * order: phase `index` -> a static name and its milliseconds;
 * AVBD_ERR_INVALID_ARGUMENT past the last phase. */
avbd_status avbd_get_create_profile(avbd_solver* s, uint32_t index, const char** name, float* ms);

/* Isotropic limiter state after the last frame (vtTrackBuf, eeTrackBuf,
 * boundsBuf, ogcPrevBuf of GPUSolver.swift:1336-1460, 25_cloth.metal:
 * 350-940): vt_track[4*num_bodies] tracked triangles per body (0xFFFFFFFF:
 * none), ee_track[4*num_contact_edges] tracked edges per authored edge,
 * bound[num_bodies] squared clearance bounds (3.3961514e38: none),
 * anchor[3*num_bodies] the poses the bounds hold around, counts[2] the
 * truncations and refreshes of the frame.  Any output may be NULL;
 * AVBD_ERR_UNSUPPORTED when the mode is off (avbd_cloth_info.iso_on). */
avbd_status avbd_get_iso_state(avbd_solver* s, uint32_t* vt_track, uint32_t* ee_track, float* bound, float* anchor,
                               uint32_t* counts);

/* Persistence map lookup (pairMapFind, 20_broadphase.metal:596): for each
 * (a,b) collider pair returns the manifold index of the map built by the last
 * frame, or -1. */
avbd_status avbd_map_find(avbd_solver* s, const uint32_t* pairs, uint32_t n, int32_t* out);

/* Element-contact persistence map lookup (softMapFind, 25_cloth.metal:87-99):
 * for each (keyA, keyB) returns the index of the last frame's element
 * contact (avbd_get_soft_contacts order) carrying that key, or -1. */
avbd_status avbd_soft_map_find(avbd_solver* s, const uint32_t* keys, uint32_t n, int32_t* out);

/* maxConstraintErrorChecked() -- GPUSolver.swift:4829 (diag_error kernel 40_solver.metal:2189) */
avbd_status avbd_max_constraint_error(avbd_solver* s, float* out);

/* Block residual r = ||g||_2 / ||m g_grav||_2 at the current iterate, g the
 * per-body block right-hand side primal_one assembles (40_solver.metal:1384-1425).
 * Non-mutating.  The reference defines no residual; this is the backend's. */
avbd_status avbd_residual(avbd_solver* s, float* out);

/* runtimeFailure -- GPUSolver.swift:2224 */
avbd_status avbd_last_failure(const avbd_solver* s, avbd_failure* out);

/* profiling/profileNS -- GPUSolver.swift:2200-2202.  Per stage GPU ms of the last
 * retired frame, names comma separated. */
avbd_status avbd_get_profile(avbd_solver* s, int32_t enable, char* names, uint32_t names_cap,
                             float* ms, uint32_t* n_stages);

/* Times `reps` full primal color sweeps (every color once) on the solver's stream:
 * the sweeps are captured in a hipGraph and one replay is bracketed by HIP
 * events; reports the average duration of one per-color primal launch and the
 * algorithmic bytes one sweep touches (per-stamp record model).  Mutates
 * positions (runs extra sweeps); call it after the timed region. */
avbd_status avbd_time_primal(avbd_solver* s, int32_t reps, float* avg_launch_us,
                             double* sweep_bytes, int32_t* launches_per_sweep);

/* Test hooks, GPUSolver.swift:2233-2241 (deniedEncoderStageForTesting,
 * completionFailureForTesting): deny the named stage ("broadphase",
 * "narrowphase", "warmstart", "adjacency", "coloring", "solve", "finalize")
 * on frame `frame` (-1 = every frame).  NULL clears. */
avbd_status avbd_test_deny_stage(avbd_solver* s, const char* stage, int64_t frame);
/* completionFailureForTesting (GPUSolver.swift:2239): the frame `frame`
 * (-1 = the next one retired) completes with failure `kind` (an
 * avbd_failure_kind), latched like a real one; one shot, 0 clears. */
avbd_status avbd_test_fail_completion(avbd_solver* s, int32_t kind, int64_t frame);
/* params.planarDATQueryRadius assignment of the Planar-DAT tests
 * (PlanarDATTests.swift:930, 960, 990): replaces the query radius rq of
 * later frames (pair query, element padding, truncation cap); the element
 * grid keeps the cell size derived at create.  r <= 0 restores the derived rq. */
avbd_status avbd_test_set_dat_query_radius(avbd_solver* s, float r);

/* ---- Speculation snapshots ------------------------------------------------
 * RigidSpeculationSnapshot (GPUSolver.swift:2948-2982): the temporal solver
 * state of a rigid scene -- poses, velocities and the adaptive velocity
 * history, inertial targets, joint / spring multipliers and penalties, both
 * contact-manifold buffers and the persistence map, the incremental colors,
 * the frame counters, frame index and stats (plus the learned warm start's
 * correction history when one is attached).  Authored scene data and
 * settings are not included.  The snapshot lives in device memory and is
 * bound to the solver that captured it. */
typedef struct avbd_snapshot avbd_snapshot;
/* captureRigidSpeculationSnapshot() -- GPUSolver.swift:2984-3012.  Waits for
 * the frames in flight.  *out == NULL allocates a snapshot; a snapshot of
 * this solver passed in *out is overwritten in place (no allocation).  A
 * scene with cloth triangles or tets returns AVBD_ERR_UNSUPPORTED (the
 * reference's precondition). */
avbd_status avbd_capture_speculation_snapshot(avbd_solver* s, avbd_snapshot** out);
/* restoreRigidSpeculationSnapshot(_:) -- GPUSolver.swift:3014-3057.  Later
 * frames replay the captured rollout bit for bit given the same inputs. */
avbd_status avbd_restore_speculation_snapshot(avbd_solver* s, const avbd_snapshot* snap);
void avbd_destroy_speculation_snapshot(avbd_snapshot* snap);

/* ---- Learned warm start (no counterpart in the reference) ----------------
 * A per-body MLP 24 -> 128 -> 128 -> 3 (particles) or 6 (rigid bodies:
 * linear, angular), ReLU, BF16 weights and activations, FP32 accumulation on
 * the matrix cores, predicts the step's converged correction relative to the
 * adaptive warm-start guess of warmstart_bodies (40_solver.metal:657-722), in
 * units of L = |g| dt^2 (rotations as the displacement at the body's bounding
 * radius r: rotation vector x r / L).
 * Particle features (24, units of L): v dt, (v - v_prev) dt, the
 * correction of two steps back c', x_guess - x, the previous step's
 * correction c, the 1-ring means of v dt
 * and c relative to the particle's own, the contact-entry count / 8, the
 * 1-ring size / 16 and |c|.  Rigid-body features: v dt, w dt r,
 * (v - v_prev) dt, x_guess - x, the previous linear and angular corrections,
 * the joint partners' mean linear correction relative to the body's own, the
 * contact-entry count / 8, the joint count / 4 and |c_lin|.
 * The parameters are a flat f32 array: W1[24][128], b1[128], W2[128][128],
 * b2[128], W3[128][outputs], b3[outputs] (W[in][out]). */
#define AVBD_MLP_INPUTS 24
#define AVBD_MLP_OUTPUTS 3
#define AVBD_MLP_OUTPUTS_RIGID 6
#define AVBD_MLP_PARAMS 20099        /* 3 outputs */
#define AVBD_MLP_PARAMS_RIGID 20486  /* 6 outputs */
typedef struct avbd_mlp avbd_mlp;
/* outputs: AVBD_MLP_OUTPUTS or AVBD_MLP_OUTPUTS_RIGID.  He-uniform
 * initialisation from `seed` (SplitMix64), output layer scaled by 0.1, zero
 * biases. */
avbd_status avbd_mlp_create(uint64_t seed, int32_t outputs, avbd_mlp** out);
void avbd_mlp_destroy(avbd_mlp* m);
avbd_status avbd_mlp_info(const avbd_mlp* m, int32_t* outputs, int32_t* num_params);
avbd_status avbd_mlp_get_params(avbd_mlp* m, float* params);
/* also resets the Adam moments */
avbd_status avbd_mlp_set_params(avbd_mlp* m, const float* params);
/* y[n][outputs] = MLP(x[n][24]); params arrays hold num_params floats */
avbd_status avbd_mlp_forward(avbd_mlp* m, const float* x, uint32_t n, float* y);
/* loss = sum |y - t|^2 / (outputs n) and its gradient grad[num_params] over
 * the n rows (one batch, the training kernels) */
avbd_status avbd_mlp_loss_grad(avbd_mlp* m, const float* x, const float* t, uint32_t n, float* loss,
                               float* grad);
/* Adam (lr, beta 0.9 / 0.999) over `epochs` passes of shuffled batches of
 * `batch` rows (the last partial batch of an epoch is dropped unless it is the
 * only one); epoch_loss[epochs] = mean batch loss per epoch (may be NULL). */
avbd_status avbd_mlp_train(avbd_mlp* m, const float* x, const float* t, uint32_t n, int32_t epochs,
                           uint32_t batch, float lr, uint64_t seed, float* epoch_loss);
/* average duration of one forward launch over n rows (HIP events on the
 * network's stream, `reps` launches) */
avbd_status avbd_mlp_time_forward(avbd_mlp* m, uint32_t n, int32_t reps, float* avg_us);

/* Attach a network to a solver: a 3-output network serves the dynamic
 * particles, a 6-output network the dynamic rigid bodies (call once per
 * network; the mode is shared).  mode bits: AVBD_WS_LEARNED use the learned
 * start (the plain guess is kept whenever the learned start's residual is
 * higher), AVBD_WS_LOG record every step's features and converged
 * corrections, AVBD_WS_FORCE keep the learned start unconditionally
 * (testing).  m may be NULL for AVBD_WS_LOG alone; the network must outlive
 * the attachment.  mode 0 detaches every network.  The learned start of
 * either kind is kept or rejected as a whole (one residual comparison). */
#define AVBD_WS_LEARNED 1
#define AVBD_WS_LOG 2
#define AVBD_WS_FORCE 4
/* bits 8-11: passes of c <- (c + <c_j>) / 2 smoothing the predicted
 * correction over the 1-ring before it is applied (0-15) */
#define AVBD_WS_SMOOTH_SHIFT 8
avbd_status avbd_set_warm_start(avbd_solver* s, avbd_mlp* m, int32_t mode);
/* Samples of the last retired frame (AVBD_WS_LOG): x[n][24] features and
 * t[n][3] converged corrections of the dynamic particles in index order;
 * idx[n] their body ids (any output may be NULL). */
avbd_status avbd_get_warm_start_samples(avbd_solver* s, uint32_t capacity, uint32_t* n, float* x, float* t,
                                        uint32_t* idx);
/* The same for the dynamic rigid bodies: x[n][24], t[n][6], idx[n]. */
avbd_status avbd_get_warm_start_samples_rigid(avbd_solver* s, uint32_t capacity, uint32_t* n, float* x, float* t,
                                              uint32_t* idx);

/* Library / device information string. */
const char* avbd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* AVBD_H */

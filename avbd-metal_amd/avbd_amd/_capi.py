"""ctypes mirror of include/avbd.h and the loader for lib/libavbd_hip.so.

The HIP library is the only compute path: if it is missing this module raises,
it never substitutes a CPU implementation.
"""
import atexit
import ctypes as C
import os
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libavbd_hip.so")

ABI_VERSION = 6
MAX_COLORS = 64

# avbd_status
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_NO_DEVICE = 2
ERR_ALLOC = 3
ERR_UNSUPPORTED = 4
ERR_RUNTIME_FAILURE = 5
ERR_HIP = 6
ERR_TOO_LARGE = 7
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "NO_DEVICE", 3: "ALLOC", 4: "UNSUPPORTED",
                5: "RUNTIME_FAILURE", 6: "HIP", 7: "TOO_LARGE"}

# avbd_failure_kind
FAILURE_NAMES = {0: "none", 1: "commandExecution", 2: "rigidPairCapacity", 3: "softContactCapacity",
                 4: "staticColorCapacity", 5: "unresolvedColoring", 6: "commandEncoderCreation",
                 7: "barrierTimeout", 8: "planarDATPairCapacity", 9: "planarDATElementGridSpan",
                 10: "planarDATInvalidAnchor"}

FP = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)
U8P = C.POINTER(C.c_uint8)


class Settings(C.Structure):
    _fields_ = [("dt", C.c_float), ("gravity", C.c_float), ("iterations", C.c_int32),
                ("collision_margin", C.c_float), ("friction_combine_mode", C.c_uint32),
                ("alpha", C.c_float), ("beta_lin", C.c_float), ("beta_ang", C.c_float),
                ("gamma", C.c_float), ("lambda_max", C.c_float), ("rod_decay_pow", C.c_float),
                ("particle_damping", C.c_float), ("rigid_linear_damping", C.c_float),
                ("rigid_angular_damping", C.c_float), ("cloth_viscosity", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("settings", Settings),
                ("num_bodies", C.c_uint32), ("body_size", FP), ("body_density", FP),
                ("body_friction", FP), ("body_dynamic_friction", FP), ("body_position", FP),
                ("body_rotation", FP), ("body_velocity", FP), ("body_shape", U32P),
                ("body_is_particle", U8P), ("body_mass", FP), ("body_inertia", FP),
                ("body_gravity_scale", FP),
                ("num_colliders", C.c_uint32), ("collider_body", U32P), ("collider_size", FP),
                ("collider_friction", FP), ("collider_dynamic_friction", FP),
                ("collider_local_position", FP), ("collider_local_rotation", FP),
                ("collider_shape", U32P), ("collider_group", U32P), ("collider_shared", U8P),
                ("collider_enabled", U8P), ("collider_world_round_anchor", U8P),
                ("num_joints", C.c_uint32), ("joint_body_a", I32P), ("joint_body_b", I32P),
                ("joint_ra", FP), ("joint_rb", FP), ("joint_stiffness_lin", FP),
                ("joint_stiffness_ang", FP), ("joint_fracture", FP), ("joint_fracture_linear", U8P),
                ("joint_hinge_axis", FP), ("joint_motor", FP), ("joint_limits", FP),
                ("num_springs", C.c_uint32), ("spring_body_a", U32P), ("spring_body_b", U32P),
                ("spring_ra", FP), ("spring_rb", FP), ("spring_stiffness", FP), ("spring_rest", FP),
                ("spring_hard", U8P),
                ("num_tets", C.c_uint32), ("tet_ids", U32P), ("tet_mu", FP), ("tet_lambda", FP),
                ("num_tris", C.c_uint32), ("tri_ids", U32P), ("tri_mu", FP), ("tri_lambda", FP),
                ("tri_bend", FP),
                ("num_exclusions", C.c_uint32), ("exclusion_pairs", U32P),
                ("collider_hull_count", U32P), ("hull_vertices", FP), ("joint_motor_rate", FP),
                ("num_spinners", C.c_uint32), ("spinner_body", U32P), ("spinner_axis_omega", FP)]


class Options(C.Structure):
    _fields_ = [("max_pairs_per_body", C.c_int32), ("lanes_per_body", C.c_int32),
                ("sweep_mode", C.c_int32), ("use_graph", C.c_int32),
                ("pair_capacity_override", C.c_int32), ("dat_pair_capacity_override", C.c_int32),
                ("soft_capacity_override", C.c_int32)]


class Failure(C.Structure):
    _fields_ = [("kind", C.c_int32), ("frame", C.c_int64), ("required", C.c_int64),
                ("capacity", C.c_int64), ("hip_error", C.c_int32), ("message", C.c_char * 160)]


class Stats(C.Structure):
    _fields_ = [("frame_index", C.c_int64), ("num_pairs", C.c_int32), ("pair_candidates", C.c_int32),
                ("num_soft", C.c_int32), ("soft_candidates", C.c_int32),
                ("num_active_manifolds", C.c_int32), ("max_color_used", C.c_int32),
                ("num_colors", C.c_int32), ("color_counts", C.c_int32 * MAX_COLORS),
                ("color_conflicts", C.c_int32), ("uses_dynamic_coloring", C.c_int32),
                ("lanes_per_body", C.c_int32), ("sweep_mode", C.c_int32),
                ("dat_pairs", C.c_int32), ("dat_vt_pairs", C.c_int32), ("dat_ee_pairs", C.c_int32),
                ("num_soft_rt", C.c_int32), ("ws_accepted", C.c_int32),
                ("ws_residual_plain", C.c_float), ("ws_residual_learned", C.c_float),
                ("dat_truncations", C.c_int32), ("dat_invalid_anchors", C.c_int32),
                ("dat_vt_degeneracies", C.c_int32), ("dat_ee_degeneracies", C.c_int32),
                ("dat_nonfinite", C.c_int32)]


class ClothInfo(C.Structure):
    _fields_ = [("num_tris", C.c_uint32), ("num_edges", C.c_uint32), ("num_particles", C.c_uint32),
                ("num_membranes", C.c_uint32), ("num_bends", C.c_uint32), ("elem_hash_size", C.c_uint32),
                ("max_soft", C.c_uint32), ("soft_map_capacity", C.c_uint32), ("max_dat_pairs", C.c_uint32),
                ("elem_cell_size", C.c_float), ("elem_margin", C.c_float), ("dat_query_radius", C.c_float),
                ("surface_contact_cell_size", C.c_float), ("num_contact_edges", C.c_uint32),
                ("pair_stream", C.c_uint32), ("dat_on", C.c_uint32),
                ("iso_on", C.c_uint32)]


# name -> (restype, argtypes); the complete exported surface of avbd.h
SIGNATURES = {
    "avbd_settings_default": (None, [C.POINTER(Settings)]),
    "avbd_options_default": (None, [C.POINTER(Options)]),
    "avbd_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "avbd_destroy": (None, [C.c_void_p]),
    "avbd_submit_step": (C.c_int, [C.c_void_p]),
    "avbd_step": (C.c_int, [C.c_void_p]),
    "avbd_synchronize": (C.c_int, [C.c_void_p]),
    "avbd_get_settings": (C.c_int, [C.c_void_p, C.POINTER(Settings)]),
    "avbd_set_settings": (C.c_int, [C.c_void_p, C.POINTER(Settings)]),
    "avbd_get_body_states": (C.c_int, [C.c_void_p, U32P, C.c_uint32, FP, FP, FP, FP]),
    "avbd_get_body_mass": (C.c_int, [C.c_void_p, C.c_uint32, FP, FP]),
    "avbd_set_body_states": (C.c_int, [C.c_void_p, U32P, C.c_uint32, FP, FP, FP, FP]),
    "avbd_set_body_states_noalloc": (C.c_int, [C.c_void_p, U32P, C.c_uint32, FP, FP, FP, FP]),
    "avbd_set_motor_targets": (C.c_int, [C.c_void_p, U32P, C.c_uint32, FP]),
    "avbd_set_motor_torques": (C.c_int, [C.c_void_p, U32P, C.c_uint32, FP]),
    "avbd_set_motor_velocities": (C.c_int, [C.c_void_p, U32P, C.c_uint32, FP]),
    "avbd_set_joint_world_anchors": (C.c_int, [C.c_void_p, U32P, C.c_uint32, FP]),
    "avbd_set_body_poses": (C.c_int, [C.c_void_p, U32P, C.c_uint32, FP, FP]),
    "avbd_get_motor_states": (C.c_int, [C.c_void_p, U32P, C.c_uint32, FP, FP]),
    "avbd_get_joint_diagnostics": (C.c_int, [C.c_void_p, I32P, FP, FP]),
    "avbd_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "avbd_get_colors": (C.c_int, [C.c_void_p, C.c_uint32, U32P]),
    "avbd_get_cell_lists": (C.c_int, [C.c_void_p, U32P, U32P, FP, U32P, U32P, U32P]),
    "avbd_get_pairs": (C.c_int, [C.c_void_p, C.c_uint32, U32P, U32P]),
    "avbd_get_manifolds": (C.c_int, [C.c_void_p, C.c_uint32, U32P, C.c_void_p]),
    "avbd_step_to_residual": (C.c_int, [C.c_void_p, C.c_float, C.c_int32, C.c_int32, I32P, FP]),
    "avbd_get_cloth_info": (C.c_int, [C.c_void_p, C.POINTER(ClothInfo), U32P, U32P]),
    "avbd_get_dat_pairs": (C.c_int, [C.c_void_p, C.c_uint32, U32P, U32P]),
    "avbd_get_collision_tris": (C.c_int, [C.c_void_p, C.c_uint32, U32P, U32P]),
    "avbd_get_motor_targets": (C.c_int, [C.c_void_p, C.c_uint32, FP]),
    "avbd_get_soft_contacts": (C.c_int, [C.c_void_p, C.c_uint32, U32P, C.c_void_p]),
    "avbd_get_iso_state": (C.c_int, [C.c_void_p, U32P, U32P, FP, FP, U32P]),
    "avbd_get_create_profile": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_char_p), FP]),
    "avbd_scene_write": (C.c_int, [C.c_void_p, C.c_char_p]),
    "avbd_scene_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "avbd_scene_free": (None, [C.c_void_p]),
    "avbd_map_find": (C.c_int, [C.c_void_p, U32P, C.c_uint32, I32P]),
    "avbd_soft_map_find": (C.c_int, [C.c_void_p, U32P, C.c_uint32, I32P]),
    "avbd_max_constraint_error": (C.c_int, [C.c_void_p, FP]),
    "avbd_residual": (C.c_int, [C.c_void_p, FP]),
    "avbd_last_failure": (C.c_int, [C.c_void_p, C.POINTER(Failure)]),
    "avbd_get_profile": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.c_uint32, FP, U32P]),
    "avbd_time_primal": (C.c_int, [C.c_void_p, C.c_int32, FP, C.POINTER(C.c_double), I32P]),
    "avbd_test_deny_stage": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "avbd_test_fail_completion": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64]),
    "avbd_test_set_dat_query_radius": (C.c_int, [C.c_void_p, C.c_float]),
    "avbd_capture_speculation_snapshot": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "avbd_restore_speculation_snapshot": (C.c_int, [C.c_void_p, C.c_void_p]),
    "avbd_destroy_speculation_snapshot": (None, [C.c_void_p]),
    "avbd_mlp_create": (C.c_int, [C.c_uint64, C.c_int32, C.POINTER(C.c_void_p)]),
    "avbd_mlp_info": (C.c_int, [C.c_void_p, I32P, I32P]),
    "avbd_mlp_destroy": (None, [C.c_void_p]),
    "avbd_mlp_get_params": (C.c_int, [C.c_void_p, FP]),
    "avbd_mlp_set_params": (C.c_int, [C.c_void_p, FP]),
    "avbd_mlp_forward": (C.c_int, [C.c_void_p, FP, C.c_uint32, FP]),
    "avbd_mlp_loss_grad": (C.c_int, [C.c_void_p, FP, FP, C.c_uint32, FP, FP]),
    "avbd_mlp_train": (C.c_int, [C.c_void_p, FP, FP, C.c_uint32, C.c_int32, C.c_uint32, C.c_float, C.c_uint64,
                                 FP]),
    "avbd_mlp_time_forward": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, FP]),
    "avbd_set_warm_start": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "avbd_get_warm_start_samples": (C.c_int, [C.c_void_p, C.c_uint32, U32P, FP, FP, U32P]),
    "avbd_get_warm_start_samples_rigid": (C.c_int, [C.c_void_p, C.c_uint32, U32P, FP, FP, U32P]),
    "avbd_version": (C.c_char_p, []),
}

# learned warm start (avbd.h)
MLP_INPUTS = 24
MLP_OUTPUTS = 3
MLP_OUTPUTS_RIGID = 6
MLP_PARAMS = 20099
MLP_PARAMS_RIGID = 20486
WS_LEARNED = 1
WS_LOG = 2
WS_FORCE = 4
WS_SMOOTH_SHIFT = 8

_lib = None


# Handles still alive when the interpreter exits are closed by an atexit
# hook, snapshots first, then solvers, then networks (a solver may hold a
# network attached as its warm start), while the HIP runtime is still up.
_live = {"snapshot": weakref.WeakSet(), "solver": weakref.WeakSet(), "network": weakref.WeakSet()}


def track(kind, obj):
    _live[kind].add(obj)


@atexit.register
def _close_all():
    for kind in ("snapshot", "solver", "network"):
        for obj in list(_live[kind]):
            try:
                obj.close()
            except Exception:
                pass


def load(path=None):
    """Load the HIP backend library; raises OSError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"HIP backend library missing: {p} (run __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class AVBDError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")


class RuntimeFailure(AVBDError):
    """Sticky solver failure (GPUSolver.RuntimeFailure, GPUSolver.swift:663)."""

    def __init__(self, kind, frame, required, capacity, message):
        self.kind = FAILURE_NAMES.get(kind, str(kind))
        self.frame = frame
        self.required = required
        self.capacity = capacity
        self.message = message
        AVBDError.__init__(self, ERR_RUNTIME_FAILURE, f"{self.kind}: {message}")

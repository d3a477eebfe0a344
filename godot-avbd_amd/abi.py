"""ctypes view of the C ABI in include/avbd_hip.h.

The structs mirror the header field for field; `load_library()` opens the in-tree
libavbd_hip.so and raises if it is missing (there is no CPU fallback).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libavbd_hip.so")

AVBD_OK = 0
AVBD_ERR_INVALID = -1
AVBD_ERR_UNSUPPORTED = -2
AVBD_ERR_CAPACITY = -3
AVBD_ERR_DEVICE = -4
AVBD_ERR_NO_DEVICE = -5
AVBD_ERR_WORLD = -6
AVBD_WORLD_SLOTS, AVBD_WORLD_TREE, AVBD_WORLD_NONFINITE = 1, 12, 16
MAX_CONTACTS = 8

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
u64p = C.POINTER(C.c_uint64)


class Tuning(C.Structure):
    _fields_ = [("block_dual", C.c_int32), ("block_primal", C.c_int32), ("block_stamp", C.c_int32),
                ("block_projection", C.c_int32), ("block_traverse", C.c_int32), ("live_blocks", C.c_int32),
                ("colour_lanes", C.c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("gravity", C.c_float), ("iterations", C.c_int32), ("alpha", C.c_float),
        ("beta_lin", C.c_float), ("beta_ang", C.c_float), ("gamma", C.c_float),
        ("newton_ratio", C.c_float), ("stiffness_decay", C.c_float), ("adaptive_beta", C.c_int32),
        ("adaptive_beta_error_floor", C.c_float), ("convergence_threshold", C.c_float),
        ("jacobi_relaxation", C.c_float), ("contact_only_overflow", C.c_int32),
        ("stick_drift_heal", C.c_float), ("max_penetration", C.c_float),
        ("jacobian_rebuild_distance", C.c_float), ("speculative_margin", C.c_float),
        ("position_projection", C.c_int32), ("projection_iterations", C.c_int32),
        ("projection_beta", C.c_float), ("projection_slop", C.c_float),
        ("projection_max_correction", C.c_float), ("contact_grace_ticks", C.c_int32),
        ("threads", C.c_int32), ("sleep_linear_threshold", C.c_float),
        ("sleep_angular_threshold", C.c_float), ("sleep_frames", C.c_int32), ("penalty_max", C.c_float),
    ]


def default_params(**overrides):
    """The reference Solver defaults (src/avbd/solver.h:54-144), with overrides."""
    p = Params(
        dt=1.0 / 60.0, gravity=-10.0, iterations=8, alpha=0.99, beta_lin=10000.0, beta_ang=100.0,
        gamma=0.999, newton_ratio=1.0, stiffness_decay=1.0, adaptive_beta=0,
        adaptive_beta_error_floor=1e-3, convergence_threshold=0.0, jacobi_relaxation=0.7,
        contact_only_overflow=1, stick_drift_heal=0.002, max_penetration=0.02,
        jacobian_rebuild_distance=0.0, speculative_margin=0.01, position_projection=1,
        projection_iterations=4, projection_beta=1.0, projection_slop=0.008,
        projection_max_correction=0.0, contact_grace_ticks=0, threads=0,
        sleep_linear_threshold=0.1, sleep_angular_threshold=0.139626, sleep_frames=0, penalty_max=1e10)
    # dt = 1/60 must be the float32 the reference computes (1.0f / 60.0f)
    p.dt = float(np.float32(1.0) / np.float32(60.0))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


class Bodies(C.Structure):
    _fields_ = [
        ("count", C.c_int32), ("size", f32p), ("mass", f32p), ("moment", f32p), ("friction", f32p),
        ("radius", f32p), ("gravity", f32p), ("layer", u32p), ("mask", u32p), ("lock_linear", u8p),
        ("lock_angular", u8p), ("position", f32p), ("rotation", f32p), ("velocity", f32p),
        ("angular_velocity", f32p), ("prev_velocity", f32p), ("shape", u8p), ("sleep_mode", u8p),
        ("sleeping", u8p), ("still_frames", i32p),
    ]


class Joints(C.Structure):
    _fields_ = [
        ("count", C.c_int32), ("body_a", i32p), ("body_b", i32p), ("r_a", f32p), ("r_b", f32p),
        ("stiffness_lin", f32p), ("stiffness_ang", f32p), ("fracture", f32p), ("torque_arm", f32p),
        ("c0_lin", f32p), ("c0_ang", f32p), ("penalty_lin", f32p), ("penalty_ang", f32p),
        ("lambda_lin", f32p), ("lambda_ang", f32p), ("broken", u8p), ("kind", u8p), ("rest", f32p),
        ("rest_rotation", f32p), ("axis_mode", u8p), ("axis_lower", f32p), ("axis_upper", f32p),
        ("axis_spring", f32p), ("axis_equilibrium", f32p), ("axis_damping", f32p), ("axis_lambda", f32p),
        ("axis_penalty", f32p), ("axis_initial", f32p),
    ]


class Manifolds(C.Structure):
    _fields_ = [
        ("count", C.c_int32), ("body_a", i32p), ("body_b", i32p), ("basis", f32p), ("friction", f32p),
        ("num_contacts", i32p), ("feature", i32p), ("r_a", f32p), ("r_b", f32p), ("c0", f32p),
        ("penalty", f32p), ("lambda_", f32p), ("stick", u8p), ("beta_scale", f32p),
    ]


class RayHits(C.Structure):
    _fields_ = [("body", i32p), ("t", f32p), ("point", f32p), ("normal", f32p), ("inside", u8p)]


class QueryProfile(C.Structure):
    _fields_ = [("bounds_ms", C.c_float), ("kernel_ms", C.c_float)]


class StepStats(C.Structure):
    _fields_ = [
        ("bodies", C.c_int32), ("joints", C.c_int32), ("manifolds", C.c_int32), ("contacts", C.c_int32),
        ("candidate_pairs", C.c_int32), ("new_manifolds", C.c_int32), ("colours", C.c_int32),
        ("overflow_bodies", C.c_int32), ("brake_levels", C.c_int32), ("ms_broad", C.c_float),
        ("ms_narrow", C.c_float), ("ms_solve", C.c_float), ("ms_finish", C.c_float), ("ms_total", C.c_float),
        ("ms_broad_tree", C.c_float), ("ms_broad_emit", C.c_float),
    ]


K_CLASSES = ("primal", "dual", "narrow", "brake_levels", "brake_replay", "traverse", "colour_pass",
             "proj_manifold", "proj_body", "adjacency", "sort", "body", "stamp", "refit")


class KernelProfile(C.Structure):
    _fields_ = [("steps", C.c_int32), ("avg_ms", C.c_float * len(K_CLASSES)), ("launches", C.c_int32 * len(K_CLASSES)),
                ("step_ms", C.c_float)]

    def as_dict(self):
        return {"steps": self.steps, "step_ms": self.step_ms,
                "avg_ms": dict(zip(K_CLASSES, list(self.avg_ms))),
                "launches": dict(zip(K_CLASSES, list(self.launches)))}


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


BODY_FIELDS = {
    "size": (np.float32, 3), "mass": (np.float32, 1), "moment": (np.float32, 3),
    "friction": (np.float32, 1), "radius": (np.float32, 1), "gravity": (np.float32, 1),
    "layer": (np.uint32, 1), "mask": (np.uint32, 1), "lock_linear": (np.uint8, 1),
    "lock_angular": (np.uint8, 1), "position": (np.float32, 3), "rotation": (np.float32, 4),
    "velocity": (np.float32, 3), "angular_velocity": (np.float32, 3), "prev_velocity": (np.float32, 3),
    "shape": (np.uint8, 1), "sleep_mode": (np.uint8, 1), "sleeping": (np.uint8, 1), "still_frames": (np.int32, 1),
}
SLEEP_NEVER, SLEEP_IDLE, SLEEP_START_ASLEEP = 0, 1, 2
SHAPE_BOX, SHAPE_SPHERE, SHAPE_CYLINDER = 0, 1, 2
JOINT_FIELDS = {
    "body_a": (np.int32, 1), "body_b": (np.int32, 1), "r_a": (np.float32, 3), "r_b": (np.float32, 3),
    "stiffness_lin": (np.float32, 1), "stiffness_ang": (np.float32, 1), "fracture": (np.float32, 1),
    "torque_arm": (np.float32, 1), "c0_lin": (np.float32, 3), "c0_ang": (np.float32, 3),
    "penalty_lin": (np.float32, 3), "penalty_ang": (np.float32, 3), "lambda_lin": (np.float32, 3),
    "lambda_ang": (np.float32, 3), "broken": (np.uint8, 1), "kind": (np.uint8, 1), "rest": (np.float32, 1),
    "rest_rotation": (np.float32, 4), "axis_mode": (np.uint8, 6), "axis_lower": (np.float32, 6),
    "axis_upper": (np.float32, 6), "axis_spring": (np.float32, 6), "axis_equilibrium": (np.float32, 6),
    "axis_damping": (np.float32, 6), "axis_lambda": (np.float32, 6), "axis_penalty": (np.float32, 6),
    "axis_initial": (np.float32, 6),
}
FORCE_JOINT, FORCE_SPRING, FORCE_IGNORE_COLLISION, FORCE_GENERIC = 0, 1, 2, 3
AXIS_FREE, AXIS_LOCKED, AXIS_LIMITED = 0, 1, 2
MANIFOLD_FIELDS = {
    "body_a": (np.int32, 1), "body_b": (np.int32, 1), "basis": (np.float32, 9), "friction": (np.float32, 1),
    "num_contacts": (np.int32, 1), "feature": (np.int32, 8), "r_a": (np.float32, 24), "r_b": (np.float32, 24),
    "c0": (np.float32, 24), "penalty": (np.float32, 24), "lambda_": (np.float32, 24), "stick": (np.uint8, 8),
    "beta_scale": (np.float32, 8),
}
_CT = {np.float32: f32p, np.int32: i32p, np.uint32: u32p, np.uint8: u8p}


def _struct(cls, fields, arrays, count):
    s = cls()
    s.count = count
    keep = {}
    for name, (dt, w) in fields.items():
        a = arrays.get(name)
        if a is None:
            setattr(s, name, None)
            continue
        a = np.ascontiguousarray(a, dtype=dt)
        keep[name] = a
        setattr(s, name, _ptr(a, _CT[dt]))
    s._keep = keep
    return s


def ray_hits(n):
    """Output arrays for n rays and the RayHits view over them."""
    out = {"body": np.zeros(n, np.int32), "t": np.zeros(n, np.float32), "point": np.zeros((n, 3), np.float32),
           "normal": np.zeros((n, 3), np.float32), "inside": np.zeros(n, np.uint8)}
    h = RayHits(_ptr(out["body"], i32p), _ptr(out["t"], f32p), _ptr(out["point"], f32p), _ptr(out["normal"], f32p),
                _ptr(out["inside"], u8p))
    return out, h


def report_arrays(n, stride):
    """Output arrays of a contact report for n bodies."""
    return (np.zeros(n, np.int32), np.zeros((n, stride, 3), np.float32), np.zeros((n, stride, 3), np.float32),
            np.full((n, stride), -1, np.int32))


def bodies_struct(arrays):
    n = len(arrays["mass"])
    return _struct(Bodies, BODY_FIELDS, arrays, n)


DYNAMIC_FIELDS = ("position", "rotation", "velocity", "angular_velocity", "prev_velocity", "sleeping", "still_frames")


def dynamic_bodies_struct(arrays, count):
    """A Bodies view over only the dynamic fields in `arrays` (others null), for write_bodies."""
    return _struct(Bodies, BODY_FIELDS, {k: v for k, v in arrays.items() if k in DYNAMIC_FIELDS}, count)


def joints_struct(arrays):
    if arrays is None or len(arrays.get("body_b", [])) == 0:
        s = Joints()
        s.count = 0
        return s
    return _struct(Joints, JOINT_FIELDS, arrays, len(arrays["body_b"]))


def empty_arrays(fields, n):
    return {k: np.zeros((n, w) if w > 1 else (n,), dtype=dt) for k, (dt, w) in fields.items()}


def manifolds_struct(arrays):
    return _struct(Manifolds, MANIFOLD_FIELDS, arrays, len(arrays["body_a"]))


_lib = None


def load_library(path=LIB_PATH):
    """Load libavbd_hip.so; raise if it is not built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() or godot-avbd_amd/build.sh")
    lib = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "avbd_params_default": (None, [C.POINTER(Params)]),
        "avbd_box_properties": (None, [f32p, C.c_float, f32p, f32p, f32p]),
        "avbd_shape_properties": (C.c_int, [C.c_int, f32p, C.c_float, f32p, f32p, f32p]),
        "avbd_state_digest": (C.c_uint64, [C.POINTER(Bodies)]),
        "avbd_last_error": (C.c_char_p, []),
        "avbd_device_count": (C.c_int32, []),
        "avbd_world_create": (C.c_int32, [C.POINTER(Params), C.c_int32, C.c_int32, C.c_int32, C.POINTER(vp)]),
        "avbd_world_destroy": (C.c_int32, [vp]),
        "avbd_world_set_params": (C.c_int32, [vp, C.POINTER(Params)]),
        "avbd_world_upload": (C.c_int32, [vp, C.POINTER(Bodies), C.POINTER(Joints)]),
        "avbd_world_step": (C.c_int32, [vp]),
        "avbd_world_write_bodies": (C.c_int32, [vp, C.POINTER(Bodies)]),
        "avbd_world_step_interleaved": (C.c_int32, [vp, C.c_int32, C.c_int32]),
        "avbd_world_download": (C.c_int32, [vp, C.POINTER(Bodies), C.POINTER(Joints)]),
        "avbd_world_get_manifolds": (C.c_int32, [vp, C.POINTER(Manifolds)]),
        "avbd_world_record_residuals": (C.c_int32, [vp, C.c_int32]),
        "avbd_world_get_residuals": (C.c_int32, [vp, i32p, f32p, i32p, f32p]),
        "avbd_world_raycast": (C.c_int32, [vp, C.c_int32, f32p, f32p, u32p, C.c_int32, C.POINTER(RayHits)]),
        "avbd_world_raycast_noalloc": (C.c_int32, [vp, C.c_int32, f32p, f32p, u32p, C.c_int32, C.POINTER(RayHits)]),
        "avbd_world_collide_pair": (C.c_int32, [vp, C.c_int32, C.c_int32, C.c_float, i32p, i32p, f32p, f32p, f32p]),
        "avbd_world_collide_shape": (C.c_int32, [vp, C.c_int32, i32p, f32p, f32p, f32p, f32p, u32p, C.c_int32, i32p,
                                                 u8p, f32p]),
        "avbd_world_rest_info": (C.c_int32, [vp, C.c_int32, i32p, f32p, f32p, f32p, f32p, u32p, i32p, f32p, f32p,
                                             f32p]),
        "avbd_world_closest_point": (C.c_int32, [vp, C.c_int32, i32p, f32p, f32p]),
        "avbd_batch_raycast": (C.c_int32, [vp, C.c_int32, C.c_int32, f32p, f32p, u32p, C.c_int32,
                                           C.POINTER(RayHits)]),
        "avbd_world_pick": (C.c_int32, [vp, C.c_int32, f32p, f32p, u32p, i32p, f32p]),
        "avbd_world_contact_report": (C.c_int32, [vp, i32p, C.c_int32, i32p, f32p, f32p, i32p]),
        "avbd_world_intersect_point": (C.c_int32, [vp, C.c_int32, f32p, u32p, C.c_int32, i32p, i32p]),
        "avbd_world_cast_motion": (C.c_int32, [vp, C.c_int32, i32p, f32p, f32p, f32p, f32p, f32p, u32p, i32p, f32p,
                                               f32p, f32p]),
        "avbd_world_intersect_shape": (C.c_int32, [vp, C.c_int32, i32p, f32p, f32p, f32p, f32p, u32p, C.c_int32, i32p,
                                                   i32p]),
        "avbd_batch_contact_report": (C.c_int32, [vp, C.c_int32, i32p, C.c_int32, i32p, f32p, f32p, i32p]),
        "avbd_world_profile_raycast": (C.c_int32, [vp, C.c_int32, f32p, f32p, C.c_int32, C.POINTER(QueryProfile)]),
        "avbd_world_collect_aabb": (C.c_int32, [vp, f32p, f32p, C.c_uint32, i32p, C.c_int32, i32p]),
        "avbd_world_set_manifolds": (C.c_int32, [vp, C.POINTER(Manifolds)]),
        "avbd_world_stats": (C.c_int32, [vp, C.POINTER(StepStats)]),
        "avbd_world_colours": (C.c_int32, [vp, u8p]),
        "avbd_world_probe_broadphase": (C.c_int32, [vp, u64p, i32p, u64p, i32p]),
        "avbd_world_profile": (C.c_int32, [vp, C.c_int32, C.POINTER(KernelProfile)]),
        "avbd_batch_profile": (C.c_int32, [vp, C.c_int32, C.POINTER(KernelProfile)]),
        "avbd_batch_create": (C.c_int32, [C.POINTER(Params), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                          C.POINTER(vp)]),
        "avbd_batch_destroy": (C.c_int32, [vp]),
        "avbd_batch_autotune": (C.c_int32, [vp, C.c_int32, C.POINTER(Tuning)]),
        "avbd_world_autotune": (C.c_int32, [vp, C.c_int32, C.POINTER(Tuning)]),
        "avbd_batch_set_tuning": (C.c_int32, [vp, C.POINTER(Tuning)]),
        "avbd_batch_footprint": (C.c_int32, [C.POINTER(Params), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                             C.POINTER(C.c_int64)]),
        "avbd_batch_device_bytes": (C.c_int32, [vp, C.POINTER(C.c_int64)]),
        "avbd_batch_add_world": (C.c_int32, [vp, C.POINTER(Bodies), C.POINTER(Joints), i32p]),
        "avbd_batch_add_world_async": (C.c_int32, [vp, C.POINTER(Bodies), C.POINTER(Joints), i32p]),
        "avbd_batch_step_launch": (C.c_int32, [vp]),
        "avbd_batch_step_wait": (C.c_int32, [vp]),
        "avbd_batch_remove_world": (C.c_int32, [vp, C.c_int32]),
        "avbd_batch_step": (C.c_int32, [vp]),
        "avbd_batch_failed_worlds": (C.c_int32, [vp, C.c_int32, i32p, i32p, i32p]),
        "avbd_batch_fits": (C.c_int32, [vp, C.c_int32, C.c_int32, i32p, i32p]),
        "avbd_batch_use_graph": (C.c_int32, [vp, C.c_int32]),
        "avbd_world_use_graph": (C.c_int32, [vp, C.c_int32]),
        "avbd_world_use_query_tree": (C.c_int32, [vp, C.c_int32]),
        "avbd_batch_use_query_tree": (C.c_int32, [vp, C.c_int32]),
        "avbd_batch_download_world": (C.c_int32, [vp, C.c_int32, C.POINTER(Bodies), C.POINTER(Joints)]),
        "avbd_batch_world_count": (C.c_int32, [vp, i32p]),
        "avbd_batch_get_manifolds": (C.c_int32, [vp, C.c_int32, C.POINTER(Manifolds)]),
        "avbd_batch_set_manifolds": (C.c_int32, [vp, C.c_int32, C.POINTER(Manifolds)]),
        "avbd_batch_list_stats": (C.c_int32, [vp, i32p]),
        "avbd_world_list_stats": (C.c_int32, [vp, i32p]),
        "avbd_batch_totals": (C.c_int32, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class AvbdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"avbd error {code}: {msg}")
        self.code = code


def check(code):
    if code != AVBD_OK:
        raise AvbdError(code, load_library().avbd_last_error().decode())
    return code

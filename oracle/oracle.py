"""ORACLE - test infrastructure only.  ctypes binding of liboracle.so (the serial CPU
restatement of the reference AVBD step) and of oracle/_ref/libavbd_ref.so (the reference's
own collide.cpp and maths.h, built from the reference tree when it is present).

Only tests/, the smoke check and the bench's CPU baseline may import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libavbd_ref.so")


def build(ref=False):
    """Build liboracle.so (and the reference pieces when `ref` and the tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def _abi():
    import importlib
    import sys
    root = os.path.dirname(HERE)
    if root not in sys.path:
        sys.path.insert(0, root)
    return importlib.import_module("godot-avbd_amd.abi")


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    a = _abi()
    L = C.CDLL(LIB)
    vp = C.c_void_p
    sig = {
        "oracle_create": (vp, [C.POINTER(a.Params)]),
        "oracle_destroy": (None, [vp]),
        "oracle_set_params": (C.c_int, [vp, C.POINTER(a.Params)]),
        "oracle_upload": (C.c_int, [vp, C.POINTER(a.Bodies), C.POINTER(a.Joints)]),
        "oracle_write_bodies": (C.c_int, [vp, C.POINTER(a.Bodies)]),
        "oracle_force_stamp": (C.c_int, [vp, C.c_int, C.c_int, a.f32p]),
        "oracle_force_update": (C.c_int, [vp, C.c_int, C.c_int]),
        "oracle_step": (C.c_int, [vp]),
        "oracle_step_interleaved": (C.c_int, [vp, C.c_int, C.c_int]),
        "oracle_download": (C.c_int, [vp, C.POINTER(a.Bodies), C.POINTER(a.Joints)]),
        "oracle_manifold_count": (C.c_int, [vp]),
        "oracle_get_manifolds": (C.c_int, [vp, C.POINTER(a.Manifolds)]),
        "oracle_set_manifolds": (C.c_int, [vp, C.POINTER(a.Manifolds)]),
        "oracle_digest": (C.c_uint64, [vp]),
        "oracle_plant": (None, [C.c_int]),
        "oracle_cache_counters": (None, [vp, C.POINTER(C.c_longlong)]),
        "oracle_get_residuals": (C.c_int, [vp, a.f32p, a.f32p]),
        "oracle_collide_pair": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, a.i32p, a.f32p, a.f32p, a.f32p]),
        "oracle_stats": (None, [vp, a.i32p]),
        "oracle_colours": (None, [vp, a.u8p]),
        "oracle_swept_aabbs": (None, [vp, a.f32p, a.f32p]),
        "oracle_candidate_pairs": (C.c_int, [vp, a.u64p, C.c_int]),
        "oracle_morton": (None, [a.f32p, C.c_int, a.u64p, a.i32p]),
        "oracle_solve6": (C.c_int, [a.f32p, a.f32p, a.f32p]),
        "oracle_collide_boxes": (C.c_int, [a.f32p, a.f32p, a.f32p, a.f32p, a.f32p, a.f32p, C.c_float,
                                           a.f32p, a.i32p, a.f32p, a.f32p]),
        "oracle_collide_shapes": (C.c_int, [C.c_int, a.f32p, a.f32p, a.f32p, C.c_int, a.f32p, a.f32p, a.f32p,
                                            C.c_float, a.f32p, a.i32p, a.f32p, a.f32p]),
        "oracle_atan2": (C.c_float, [C.c_float, C.c_float]),
        "oracle_intersect_ray": (C.c_int, [C.c_int, a.f32p, a.f32p, a.f32p, a.f32p, a.f32p, C.c_float, a.f32p,
                                           a.f32p, a.f32p, a.u8p]),
        "oracle_collect_aabb": (C.c_int, [vp, a.f32p, a.f32p, C.c_uint32, a.i32p, C.c_int]),
        "oracle_raycast": (C.c_int, [vp, C.c_int, a.f32p, a.f32p, a.u32p, C.c_int, a.i32p, a.f32p, a.f32p, a.f32p,
                                     a.u8p]),
        "oracle_pick": (C.c_int, [vp, C.c_int, a.f32p, a.f32p, a.u32p, a.i32p, a.f32p]),
        "oracle_contact_report": (C.c_int, [vp, a.i32p, C.c_int, a.i32p, a.f32p, a.f32p, a.i32p]),
        "oracle_contains_point": (C.c_int, [C.c_int, a.f32p, a.f32p, a.f32p, a.f32p, C.c_float]),
        "oracle_closest_point_shape": (None, [C.c_int, a.f32p, a.f32p, a.f32p, a.f32p, a.f32p]),
        "oracle_closest_point": (C.c_int, [vp, C.c_int, a.i32p, a.f32p, a.f32p]),
        "oracle_collide_shape": (C.c_int, [vp, C.c_int, a.i32p, a.f32p, a.f32p, a.f32p, a.f32p, a.u32p, C.c_int,
                                           a.i32p, a.u8p, a.f32p]),
        "oracle_rest_info": (C.c_int, [vp, C.c_int, a.i32p, a.f32p, a.f32p, a.f32p, a.f32p, a.u32p, a.i32p, a.f32p,
                                       a.f32p, a.f32p]),
        "oracle_sweep_shape": (C.c_int, [C.c_int, a.f32p, a.f32p, a.f32p, a.f32p, C.c_int, a.f32p, a.f32p, a.f32p,
                                         C.c_float, a.f32p]),
        "oracle_cast_motion": (C.c_int, [vp, C.c_int, a.i32p, a.f32p, a.f32p, a.f32p, a.f32p, a.f32p, a.u32p, a.i32p,
                                         a.f32p, a.f32p, a.f32p]),
        "oracle_intersect_shape": (C.c_int, [vp, C.c_int, a.i32p, a.f32p, a.f32p, a.f32p, a.f32p, a.u32p, C.c_int,
                                             a.i32p, a.i32p]),
        "oracle_intersect_point": (C.c_int, [vp, C.c_int, a.f32p, a.u32p, C.c_int, a.i32p, a.i32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def ref_lib():
    """The reference's own collide.cpp / maths.h solves, or None when not built."""
    global _ref
    if _ref is not None:
        return _ref
    if not os.path.exists(REF_LIB):
        return None
    a = _abi()
    R = C.CDLL(REF_LIB)
    R.ref_collide_boxes.restype = C.c_int
    R.ref_collide_boxes.argtypes = [a.f32p, a.f32p, a.f32p, a.f32p, a.f32p, a.f32p, C.c_float,
                                    a.f32p, a.i32p, a.f32p, a.f32p]
    R.ref_collide_shapes.restype = C.c_int
    R.ref_collide_shapes.argtypes = [C.c_int, a.f32p, a.f32p, a.f32p, C.c_int, a.f32p, a.f32p, a.f32p, C.c_float,
                                     a.f32p, a.i32p, a.f32p, a.f32p]
    R.ref_intersect_ray.restype = C.c_int
    R.ref_intersect_ray.argtypes = [C.c_int, a.f32p, a.f32p, a.f32p, a.f32p, a.f32p, C.c_float, a.f32p, a.f32p,
                                    a.f32p, a.u8p]
    R.ref_sweep_shape.restype = C.c_int
    R.ref_sweep_shape.argtypes = [C.c_int, a.f32p, a.f32p, a.f32p, a.f32p, C.c_int, a.f32p, a.f32p, a.f32p, C.c_float,
                                  a.f32p]
    R.ref_closest_point.restype = None
    R.ref_closest_point.argtypes = [C.c_int, a.f32p, a.f32p, a.f32p, a.f32p, a.f32p]
    R.ref_contains_point.restype = C.c_int
    R.ref_contains_point.argtypes = [C.c_int, a.f32p, a.f32p, a.f32p, a.f32p, C.c_float]
    R.ref_solve6.restype = C.c_int
    R.ref_solve6.argtypes = [a.f32p, a.f32p, a.f32p]
    R.ref_morton3D.restype = C.c_uint64
    R.ref_morton3D.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]
    _ref = R
    return R


class OracleWorld:
    def __init__(self, params):
        self.a = _abi()
        self.L = lib()
        self.h = self.L.oracle_create(C.byref(params))
        if not self.h:
            raise ValueError("oracle rejected the parameters (unsupported feature)")
        self.n = 0
        self.nj = 0

    def __del__(self):
        try:
            if self.h:
                self.L.oracle_destroy(self.h)
        except Exception:
            pass

    def set_params(self, params):
        assert self.L.oracle_set_params(self.h, C.byref(params)) == 0

    def upload(self, bodies, joints=None):
        bs = self.a.bodies_struct(bodies)
        js = self.a.joints_struct(joints)
        assert self.L.oracle_upload(self.h, C.byref(bs), C.byref(js)) == 0
        self.n = bs.count
        self.nj = js.count

    def write_bodies(self, arrays):
        """Overwrite the dynamic state of the uploaded bodies (what a host does between steps)."""
        bs = self.a.dynamic_bodies_struct(arrays, self.n)
        assert self.L.oracle_write_bodies(self.h, C.byref(bs)) == 0

    def force_stamp(self, joint, body):
        """(lhs 6x6 over [lin, ang], rhs 6) that force `joint` adds to `body`'s primal block now."""
        out = np.zeros(33, np.float32)
        assert self.L.oracle_force_stamp(self.h, joint, body, out.ctypes.data_as(self.a.f32p)) == 0
        ll, aa, al = out[0:9].reshape(3, 3), out[9:18].reshape(3, 3), out[18:27].reshape(3, 3)
        lhs = np.block([[ll, al.T], [al, aa]]).astype(np.float64)
        return lhs, out[27:33].astype(np.float64)

    def force_update(self, joint, what):
        """Apply one force's warm start ("init") or dual update ("dual") alone."""
        assert self.L.oracle_force_update(self.h, joint, {"init": 0, "dual": 1}[what]) == 0

    def collect_aabb(self, mn, mx, mask=0xFFFFFFFF):
        """Solver::collectAABB: list indices of the bodies whose bounds overlap [mn, mx]."""
        a = self.a
        out = np.zeros(max(self.n, 1), np.int32)
        lo, hi = np.ascontiguousarray(mn, np.float32), np.ascontiguousarray(mx, np.float32)
        k = self.L.oracle_collect_aabb(self.h, lo.ctypes.data_as(a.f32p), hi.ctypes.data_as(a.f32p), int(mask),
                                       out.ctypes.data_as(a.i32p), len(out))
        return out[:k]

    def raycast(self, frm, to, masks=None, hit_from_inside=False):
        """Segments frm -> to (n x 3): dict of body (-1 none), t, point, normal, inside."""
        a = self.a
        frm, to = np.ascontiguousarray(frm, np.float32), np.ascontiguousarray(to, np.float32)
        n = len(frm)
        out = {"body": np.zeros(n, np.int32), "t": np.zeros(n, np.float32), "point": np.zeros((n, 3), np.float32),
               "normal": np.zeros((n, 3), np.float32), "inside": np.zeros(n, np.uint8)}
        m = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        self.L.oracle_raycast(self.h, n, frm.ctypes.data_as(a.f32p), to.ctypes.data_as(a.f32p),
                              None if m is None else m.ctypes.data_as(a.u32p), int(bool(hit_from_inside)),
                              out["body"].ctypes.data_as(a.i32p), out["t"].ctypes.data_as(a.f32p),
                              out["point"].ctypes.data_as(a.f32p), out["normal"].ctypes.data_as(a.f32p),
                              out["inside"].ctypes.data_as(a.u8p))
        return out

    def pick(self, origin, direction, masks=None):
        """Solver::pick for n rays: (body list index or -1, local hit point)."""
        a = self.a
        o, d = np.ascontiguousarray(origin, np.float32), np.ascontiguousarray(direction, np.float32)
        n = len(o)
        body, local = np.zeros(n, np.int32), np.zeros((n, 3), np.float32)
        m = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        self.L.oracle_pick(self.h, n, o.ctypes.data_as(a.f32p), d.ctypes.data_as(a.f32p),
                           None if m is None else m.ctypes.data_as(a.u32p), body.ctypes.data_as(a.i32p),
                           local.ctypes.data_as(a.f32p))
        return body, local

    def intersect_point(self, points, masks=None, max_results=8):
        """The server's intersect_point for n points: (count, body indices n x max_results)."""
        a = self.a
        pts = np.ascontiguousarray(points, np.float32)
        n = len(pts)
        count, out = np.zeros(n, np.int32), np.full((n, max_results), -1, np.int32)
        m = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        self.L.oracle_intersect_point(self.h, n, pts.ctypes.data_as(a.f32p), None if m is None else m.ctypes.data_as(a.u32p),
                                      max_results, count.ctypes.data_as(a.i32p), out.ctypes.data_as(a.i32p))
        return count, out

    def intersect_shape(self, types, centers, rotations, sizes, margins, masks=None, max_results=8):
        """The server's intersect_shape for n query shapes: (count, body indices n x max_results)."""
        a = self.a
        t = np.ascontiguousarray(types, np.int32)
        c, q, s3, m = (np.ascontiguousarray(x, np.float32) for x in (centers, rotations, sizes, margins))
        n = len(t)
        count, out = np.zeros(n, np.int32), np.full((n, max_results), -1, np.int32)
        mk = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        self.L.oracle_intersect_shape(self.h, n, t.ctypes.data_as(a.i32p), c.ctypes.data_as(a.f32p),
                                      q.ctypes.data_as(a.f32p), s3.ctypes.data_as(a.f32p), m.ctypes.data_as(a.f32p),
                                      None if mk is None else mk.ctypes.data_as(a.u32p), max_results,
                                      count.ctypes.data_as(a.i32p), out.ctypes.data_as(a.i32p))
        return count, out

    def _shape_args(self, types, centers, rotations, sizes, margins, masks):
        a = self.a
        t = np.ascontiguousarray(types, np.int32)
        arrs = [np.ascontiguousarray(x, np.float32) for x in (centers, rotations, sizes, margins)]
        mk = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        keep = (t, arrs, mk)
        args = [len(t), t.ctypes.data_as(a.i32p)] + [x.ctypes.data_as(a.f32p) for x in arrs] + \
               [None if mk is None else mk.ctypes.data_as(a.u32p)]
        return keep, args

    def collide_shape(self, types, centers, rotations, sizes, margins, masks=None, max_pairs=8):
        """The server's collide_shape for n query shapes: (pair count, any, points n x max_pairs x 2 x 3)."""
        a = self.a
        keep, args = self._shape_args(types, centers, rotations, sizes, margins, masks)
        n = args[0]
        count, anyhit = np.zeros(n, np.int32), np.zeros(n, np.uint8)
        pts = np.zeros((n, max_pairs, 2, 3), np.float32)
        self.L.oracle_collide_shape(self.h, *args, max_pairs, count.ctypes.data_as(a.i32p),
                                    anyhit.ctypes.data_as(a.u8p), pts.ctypes.data_as(a.f32p))
        return count, anyhit, pts

    def rest_info(self, types, centers, rotations, sizes, margins, masks=None):
        """The server's rest_info for n query shapes: (body or -1, point, normal, depth)."""
        a = self.a
        keep, args = self._shape_args(types, centers, rotations, sizes, margins, masks)
        n = args[0]
        body, point, normal, depth = (np.zeros(n, np.int32), np.zeros((n, 3), np.float32),
                                      np.zeros((n, 3), np.float32), np.zeros(n, np.float32))
        self.L.oracle_rest_info(self.h, *args, body.ctypes.data_as(a.i32p), point.ctypes.data_as(a.f32p),
                                normal.ctypes.data_as(a.f32p), depth.ctypes.data_as(a.f32p))
        return body, point, normal, depth

    def closest_point(self, bodies, points):
        """The server's closest point to each body's volume (bodies, points n x 3)."""
        a = self.a
        b = np.ascontiguousarray(bodies, np.int32)
        p = np.ascontiguousarray(points, np.float32)
        out = np.zeros((len(b), 3), np.float32)
        assert self.L.oracle_closest_point(self.h, len(b), b.ctypes.data_as(a.i32p), p.ctypes.data_as(a.f32p),
                                           out.ctypes.data_as(a.f32p)) == 0
        return out

    def cast_motion(self, types, centers, rotations, sizes, motions, margins, masks=None):
        """The server's cast_motion for n query shapes: dict of body (-1: none), safe, unsafe,
        point, normal."""
        a = self.a
        t = np.ascontiguousarray(types, np.int32)
        c, q, s3, mv, m = (np.ascontiguousarray(x, np.float32) for x in (centers, rotations, sizes, motions, margins))
        n = len(t)
        body, frac = np.zeros(n, np.int32), np.zeros((n, 2), np.float32)
        point, normal = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        mk = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        self.L.oracle_cast_motion(self.h, n, t.ctypes.data_as(a.i32p), c.ctypes.data_as(a.f32p), q.ctypes.data_as(a.f32p),
                                  s3.ctypes.data_as(a.f32p), mv.ctypes.data_as(a.f32p), m.ctypes.data_as(a.f32p),
                                  None if mk is None else mk.ctypes.data_as(a.u32p), body.ctypes.data_as(a.i32p),
                                  frac.ctypes.data_as(a.f32p), point.ctypes.data_as(a.f32p), normal.ctypes.data_as(a.f32p))
        return {"body": body, "safe": frac[:, 0], "unsafe": frac[:, 1], "point": point, "normal": normal}

    def contact_report(self, max_reported):
        """Per-body contact reports: (count, position, normal, other) with a stride of
        max(max_reported)."""
        a = self.a
        mr = np.ascontiguousarray(max_reported, np.int32)
        stride = max(1, int(mr.max()) if len(mr) else 1)
        count = np.zeros(self.n, np.int32)
        pos = np.zeros((self.n, stride, 3), np.float32)
        nrm = np.zeros((self.n, stride, 3), np.float32)
        other = np.full((self.n, stride), -1, np.int32)
        self.L.oracle_contact_report(self.h, mr.ctypes.data_as(a.i32p), stride, count.ctypes.data_as(a.i32p),
                                     pos.ctypes.data_as(a.f32p), nrm.ctypes.data_as(a.f32p), other.ctypes.data_as(a.i32p))
        return count, pos, nrm, other

    def step(self, n=1):
        for _ in range(n):
            self.L.oracle_step(self.h)

    def step_interleaved(self, substeps, rounds):
        self.L.oracle_step_interleaved(self.h, substeps, rounds)

    def download(self, joints=False):
        a = self.a
        out = a.empty_arrays({k: a.BODY_FIELDS[k] for k in
                              ("position", "rotation", "velocity", "angular_velocity", "prev_velocity", "sleeping",
                                 "still_frames")}, self.n)
        bs = a.bodies_struct({**out, "mass": np.zeros(self.n, np.float32)})
        bs.mass = None
        js = None
        jout = None
        if joints and self.nj:
            jout = a.empty_arrays({k: a.JOINT_FIELDS[k] for k in
                                   ("body_a", "body_b", "c0_lin", "c0_ang", "penalty_lin", "penalty_ang",
                                    "lambda_lin", "lambda_ang", "broken", "axis_lambda", "axis_penalty",
                                    "axis_initial")}, self.nj)
            js = a.joints_struct(jout)
        self.L.oracle_download(self.h, C.byref(bs), C.byref(js) if js is not None else None)
        if jout is not None:
            jout = {k: v[:js.count] for k, v in jout.items()}
        return (out, jout) if joints else out

    def manifolds(self):
        a = self.a
        m = self.L.oracle_manifold_count(self.h)
        arr = a.empty_arrays(a.MANIFOLD_FIELDS, m)
        ms = a.manifolds_struct(arr)
        assert self.L.oracle_get_manifolds(self.h, C.byref(ms)) == 0
        return arr

    def set_manifolds(self, arrays):
        ms = self.a.manifolds_struct(arrays)
        assert self.L.oracle_set_manifolds(self.h, C.byref(ms)) == 0

    def digest(self):
        return int(self.L.oracle_digest(self.h))

    def residuals(self):
        """The last dual's constraint values: (live manifolds x 8 normal values C.x, joints x 3
        linear C), in the order of manifolds() and download(joints=True)."""
        nm = int(self.L.oracle_manifold_count(self.h))
        _, jt = self.download(joints=True)
        nj = 0 if jt is None else len(jt["body_b"])
        cn = np.zeros((nm, 8), np.float32)
        cl = np.zeros((max(nj, 1), 3), np.float32)
        assert self.L.oracle_get_residuals(self.h, cn.ctypes.data_as(self.a.f32p), cl.ctypes.data_as(self.a.f32p)) == 0
        return cn, cl[:nj]

    def cache_counters(self):
        """(dual evaluations that reused cached arms, primal touches served by the force cache)
        since the world was created."""
        out = (C.c_longlong * 2)()
        self.L.oracle_cache_counters(self.h, out)
        return int(out[0]), int(out[1])

    def collide_pair(self, a, b, hz):
        """avbd::Manifold::collide for bodies a and b now: (count, keys, rA, rB, basis 3x3 rows)."""
        keys, ra, rb = np.zeros(8, np.int32), np.zeros((8, 3), np.float32), np.zeros((8, 3), np.float32)
        bas = np.zeros(9, np.float32)
        n = self.L.oracle_collide_pair(self.h, int(a), int(b), float(hz), keys.ctypes.data_as(self.a.i32p),
                                       ra.ctypes.data_as(self.a.f32p), rb.ctypes.data_as(self.a.f32p),
                                       bas.ctypes.data_as(self.a.f32p))
        assert n >= 0
        return n, keys[:n], ra[:n], rb[:n], bas.reshape(3, 3)

    def stats(self):
        out = (C.c_int32 * 6)()
        self.L.oracle_stats(self.h, out)
        keys = ("manifolds", "contacts", "candidate_pairs", "new_manifolds", "colours", "overflow_bodies")
        return dict(zip(keys, list(out)))

    def colours(self):
        out = np.zeros(self.n, np.uint8)
        self.L.oracle_colours(self.h, out.ctypes.data_as(_abi().u8p))
        return out

    def swept_aabbs(self):
        c = np.zeros((self.n, 3), np.float32)
        h = np.zeros((self.n, 3), np.float32)
        self.L.oracle_swept_aabbs(self.h, c.ctypes.data_as(self.a.f32p), h.ctypes.data_as(self.a.f32p))
        return c, h

    def candidate_pairs(self):
        cap = 16 * self.n + 64
        p = np.zeros(cap, np.uint64)
        k = self.L.oracle_candidate_pairs(self.h, p.ctypes.data_as(self.a.u64p), cap)
        if k > cap:
            p = np.zeros(k, np.uint64)
            self.L.oracle_candidate_pairs(self.h, p.ctypes.data_as(self.a.u64p), k)
        return p[:k]


def morton(centers):
    a = _abi()
    c = np.ascontiguousarray(centers, np.float32)
    n = len(c)
    codes = np.zeros(n, np.uint64)
    order = np.zeros(n, np.int32)
    lib().oracle_morton(c.ctypes.data_as(a.f32p), n, codes.ctypes.data_as(a.u64p), order.ctypes.data_as(a.i32p))
    return codes, order


def ref_morton_codes(centers):
    """Codes from the reference's own morton3D with the scene bounds of `centers`
    (bvh_broadphase.cpp:332-348), or None when oracle/_ref is not built."""
    R = ref_lib()
    if R is None:
        return None
    c = np.asarray(centers, np.float32)
    lo = c.min(axis=0)
    hi = c.max(axis=0)
    ext = hi - lo
    scale = np.float32(2097151.0) / np.float32(max(float(ext[0]), float(ext[1]), float(ext[2]), float(np.float32(1e-6))))
    return np.array([R.ref_morton3D(float(x), float(y), float(z), float(lo[0]), float(scale)) for x, y, z in c],
                    np.uint64)


def _collide(fn, pa, qa, sa, pb, qb, sb, margin):
    a = _abi()
    args = [np.ascontiguousarray(x, np.float32) for x in (pa, qa, sa, pb, qb, sb)]
    basis = np.zeros(9, np.float32)
    keys = np.zeros(8, np.int32)
    ra = np.zeros(24, np.float32)
    rb = np.zeros(24, np.float32)
    n = fn(*[x.ctypes.data_as(a.f32p) for x in args], float(margin), basis.ctypes.data_as(a.f32p),
           keys.ctypes.data_as(a.i32p), ra.ctypes.data_as(a.f32p), rb.ctypes.data_as(a.f32p))
    return n, basis, keys[:n], ra[:3 * n].reshape(-1, 3), rb[:3 * n].reshape(-1, 3)


def collide_boxes(pa, qa, sa, pb, qb, sb, margin):
    return _collide(lib().oracle_collide_boxes, pa, qa, sa, pb, qb, sb, margin)


def ref_collide_boxes(pa, qa, sa, pb, qb, sb, margin):
    R = ref_lib()
    return None if R is None else _collide(R.ref_collide_boxes, pa, qa, sa, pb, qb, sb, margin)


def _collide_typed(fn, ta, pa, qa, sa, tb, pb, qb, sb, margin):
    def call(*args):
        return fn(int(ta), args[0], args[1], args[2], int(tb), *args[3:])
    return _collide(call, pa, qa, sa, pb, qb, sb, margin)


def collide_shapes(ta, pa, qa, sa, tb, pb, qb, sb, margin):
    """(count, basis row-major, keys, rA, rB) for any pair of shapes (abi.SHAPE_*)."""
    return _collide_typed(lib().oracle_collide_shapes, ta, pa, qa, sa, tb, pb, qb, sb, margin)


def ref_collide_shapes(ta, pa, qa, sa, tb, pb, qb, sb, margin):
    R = ref_lib()
    return None if R is None else _collide_typed(R.ref_collide_shapes, ta, pa, qa, sa, tb, pb, qb, sb, margin)


def _ray(fn, t, pos, rot, size, origin, direction, max_t):
    a = _abi()
    args = [np.ascontiguousarray(x, np.float32) for x in (pos, rot, size, origin, direction)]
    tt, point, normal, inside = np.zeros(1, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.uint8)
    hit = fn(int(t), *[x.ctypes.data_as(a.f32p) for x in args], float(max_t), tt.ctypes.data_as(a.f32p),
             point.ctypes.data_as(a.f32p), normal.ctypes.data_as(a.f32p), inside.ctypes.data_as(a.u8p))
    return bool(hit), tt[0], point, normal, bool(inside[0])


def intersect_ray(t, pos, rot, size, origin, direction, max_t):
    """query::intersectRay on one shape: (hit, t, point, normal, inside)."""
    return _ray(lib().oracle_intersect_ray, t, pos, rot, size, origin, direction, max_t)


def _contains(fn, t, pos, rot, size, point, margin):
    a = _abi()
    args = [np.ascontiguousarray(x, np.float32) for x in (pos, rot, size, point)]
    return bool(fn(int(t), *[x.ctypes.data_as(a.f32p) for x in args], float(margin)))


def contains_point(t, pos, rot, size, point, margin=0.0):
    """query::containsPoint on one shape."""
    return _contains(lib().oracle_contains_point, t, pos, rot, size, point, margin)


def ref_contains_point(t, pos, rot, size, point, margin=0.0):
    R = ref_lib()
    return None if R is None else _contains(R.ref_contains_point, t, pos, rot, size, point, margin)


def _sweep(fn, ta, pa, qa, sa, motion, tb, pb, qb, sb, margin):
    a = _abi()
    f = [np.ascontiguousarray(x, np.float32) for x in (pa, qa, sa, motion, pb, qb, sb)]
    p = [x.ctypes.data_as(a.f32p) for x in f]
    out = np.zeros(9, np.float32)
    hit = fn(int(ta), p[0], p[1], p[2], p[3], int(tb), p[4], p[5], p[6], float(margin), out.ctypes.data_as(a.f32p))
    return bool(hit), out


def sweep_shape(ta, pa, qa, sa, motion, tb, pb, qb, sb, margin=0.0):
    """query::sweepShape: (hit, [safe, unsafe, depth, point xyz, normal xyz])."""
    return _sweep(lib().oracle_sweep_shape, ta, pa, qa, sa, motion, tb, pb, qb, sb, margin)


def ref_sweep_shape(ta, pa, qa, sa, motion, tb, pb, qb, sb, margin=0.0):
    R = ref_lib()
    return None if R is None else _sweep(R.ref_sweep_shape, ta, pa, qa, sa, motion, tb, pb, qb, sb, margin)


def ref_intersect_ray(t, pos, rot, size, origin, direction, max_t):
    R = ref_lib()
    return None if R is None else _ray(R.ref_intersect_ray, t, pos, rot, size, origin, direction, max_t)


def _solve6(fn, A, b):
    a = _abi()
    A = np.ascontiguousarray(A, np.float32).reshape(36)
    b = np.ascontiguousarray(b, np.float32)
    x = np.zeros(6, np.float32)
    path = fn(A.ctypes.data_as(a.f32p), b.ctypes.data_as(a.f32p), x.ctypes.data_as(a.f32p))
    return path, x


def solve6(A, b):
    return _solve6(lib().oracle_solve6, A, b)


def ref_solve6(A, b):
    R = ref_lib()
    return None if R is None else _solve6(R.ref_solve6, A, b)


def atan2(y, x):
    """The generic joint's atan2, evaluated by the oracle library (one float32 pair)."""
    return float(lib().oracle_atan2(float(y), float(x)))


class planted:
    """Context manager for the oracle's own tests: plants mistake `pid` (see avbd_oracle.cpp)
    in the restatement for the duration of the block."""

    def __init__(self, pid):
        self.pid = pid

    def __enter__(self):
        lib().oracle_plant(int(self.pid))
        return self

    def __exit__(self, *exc):
        lib().oracle_plant(0)
        return False


def _closest(fn, t, pos, rot, size, point):
    a = _abi()
    args = [np.ascontiguousarray(x, np.float32) for x in (pos, rot, size, point)]
    out = np.zeros(3, np.float32)
    fn(int(t), *[x.ctypes.data_as(a.f32p) for x in args], out.ctypes.data_as(a.f32p))
    return out


def closest_point(t, pos, rot, size, point):
    """query::closestPoint on one shape (the oracle's restatement)."""
    return _closest(lib().oracle_closest_point_shape, t, pos, rot, size, point)


def ref_closest_point(t, pos, rot, size, point):
    """The reference's own query::closestPoint, or None when oracle/_ref is not built."""
    R = ref_lib()
    return None if R is None else _closest(R.ref_closest_point, t, pos, rot, size, point)

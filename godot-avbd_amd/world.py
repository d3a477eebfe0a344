"""Python handles over the C ABI: a plain single world and a batch of worlds.

These are thin: every call goes straight to libavbd_hip.so, which owns the device state.
"""
import ctypes as C

import numpy as np

from . import abi


def _rows(a, width, name, dtype=np.float32, n=None):
    """`a` as a contiguous (n, width) array ((n,) for width 1); ValueError when its shape is
    not that (the C side reads n rows of width values through a raw pointer)."""
    a = np.ascontiguousarray(a, dtype)
    if width == 1:
        ok = a.ndim == 1
    else:
        ok = a.ndim == 2 and a.shape[1] == width
    if not ok or (n is not None and len(a) != n):
        want = f"({n if n is not None else 'n'}{'' if width == 1 else f', {width}'})"
        raise ValueError(f"{name} has shape {a.shape}, expected {want}")
    return a


class World:
    """One world, stepped by the plain single-world path (avbd_world_*)."""

    def __init__(self, params, max_bodies, max_joints=0, max_manifolds=0):
        self.lib = abi.load_library()
        self.params = params
        h = C.c_void_p()
        abi.check(self.lib.avbd_world_create(C.byref(params), int(max_bodies), int(max_joints),
                                             int(max_manifolds), C.byref(h)))
        self.h = h
        self.n = 0
        self.nj = 0

    def close(self):
        if self.h:
            self.lib.avbd_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        self.params = params
        abi.check(self.lib.avbd_world_set_params(self.h, C.byref(params)))

    def upload(self, bodies, joints=None):
        bs = abi.bodies_struct(bodies)
        js = abi.joints_struct(joints)
        abi.check(self.lib.avbd_world_upload(self.h, C.byref(bs), C.byref(js)))
        self.n = bs.count
        self.nj = js.count

    def step(self, n=1):
        for _ in range(n):
            abi.check(self.lib.avbd_world_step(self.h))

    def use_graph(self, enable):
        """Step as one captured HIP graph (default) or kernel by kernel (for tracing tools)."""
        abi.check(self.lib.avbd_world_use_graph(self.h, 1 if enable else 0))

    def list_stats(self):
        """The broad phase's candidate lists in the last step."""
        out = np.zeros(4, np.int32)
        abi.check(self.lib.avbd_world_list_stats(self.h, out.ctypes.data_as(abi.i32p)))
        return dict(zip(("moved_bodies", "busy_worlds", "exact_walk_worlds", "relisted_worlds"), out.tolist()))

    def use_query_tree(self, enable):
        """Space queries walk the refit step tree (default) or scan the body list; same results."""
        abi.check(self.lib.avbd_world_use_query_tree(self.h, 1 if enable else 0))

    def step_interleaved(self, substeps, rounds):
        abi.check(self.lib.avbd_world_step_interleaved(self.h, int(substeps), int(rounds)))

    def write_bodies(self, arrays):
        """Overwrite the dynamic state of the uploaded bodies, keeping manifolds and joints."""
        bs = abi.dynamic_bodies_struct(arrays, self.n)
        abi.check(self.lib.avbd_world_write_bodies(self.h, C.byref(bs)))

    def download(self, joints=False):
        out = abi.empty_arrays({k: abi.BODY_FIELDS[k] for k in
                                ("position", "rotation", "velocity", "angular_velocity", "prev_velocity", "sleeping",
                                 "still_frames")}, self.n)
        bs = abi.bodies_struct({**out, "mass": np.zeros(self.n, np.float32)})
        bs.mass = None
        bs.count = self.n
        js = None
        jout = None
        if joints and self.nj:
            jout = abi.empty_arrays({k: abi.JOINT_FIELDS[k] for k in
                                     ("body_a", "body_b", "c0_lin", "c0_ang", "penalty_lin", "penalty_ang",
                                      "lambda_lin", "lambda_ang", "broken", "axis_lambda", "axis_penalty",
                                      "axis_initial")}, self.nj)
            js = abi.joints_struct({**jout, "body_b": jout["body_b"]})
            js.count = self.nj
        abi.check(self.lib.avbd_world_download(self.h, C.byref(bs), C.byref(js) if js is not None else None))
        if jout is not None:
            k = js.count
            jout = {key: v[:k] for key, v in jout.items()}
        return (out, jout) if joints else out

    def raycast(self, frm, to, masks=None, hit_from_inside=False, no_alloc=False):
        """Segments frm -> to (n x 3) against the bodies: dict of body (-1 none), t, point,
        normal, inside (the server's intersect_ray rules).  no_alloc: through
        avbd_world_raycast_noalloc (fixed device scratch, no device allocation)."""
        frm, to = _rows(frm, 3, "frm"), _rows(to, 3, "to")
        n = len(frm)
        if len(to) != n:
            raise ValueError(f"to has {len(to)} rows, frm {n}")
        out, h = abi.ray_hits(n)
        m = None if masks is None else _rows(masks, 1, "masks", np.uint32, n)
        fn = self.lib.avbd_world_raycast_noalloc if no_alloc else self.lib.avbd_world_raycast
        abi.check(fn(self.h, n, frm.ctypes.data_as(abi.f32p), to.ctypes.data_as(abi.f32p),
                     None if m is None else m.ctypes.data_as(abi.u32p), int(bool(hit_from_inside)), C.byref(h)))
        return out

    def pick(self, origin, direction, masks=None):
        """Solver::pick for n rays: (body list index or -1, local hit point)."""
        o, d = np.ascontiguousarray(origin, np.float32), np.ascontiguousarray(direction, np.float32)
        body, local = np.zeros(len(o), np.int32), np.zeros((len(o), 3), np.float32)
        m = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        abi.check(self.lib.avbd_world_pick(self.h, len(o), o.ctypes.data_as(abi.f32p), d.ctypes.data_as(abi.f32p),
                                           None if m is None else m.ctypes.data_as(abi.u32p),
                                           body.ctypes.data_as(abi.i32p), local.ctypes.data_as(abi.f32p)))
        return body, local

    def intersect_point(self, points, masks=None, max_results=8):
        """The server's intersect_point for n points: (count, body indices n x max_results)."""
        pts = np.ascontiguousarray(points, np.float32)
        n = len(pts)
        count, out = np.zeros(n, np.int32), np.full((n, max_results), -1, np.int32)
        m = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        abi.check(self.lib.avbd_world_intersect_point(self.h, n, pts.ctypes.data_as(abi.f32p),
                                                      None if m is None else m.ctypes.data_as(abi.u32p), max_results,
                                                      count.ctypes.data_as(abi.i32p), out.ctypes.data_as(abi.i32p)))
        return count, out

    def intersect_shape(self, types, centers, rotations, sizes, margins, masks=None, max_results=8):
        """The server's intersect_shape for n query shapes: (count, body indices n x max_results)."""
        t = np.ascontiguousarray(types, np.int32)
        c, q, s3, m = (np.ascontiguousarray(x, np.float32) for x in (centers, rotations, sizes, margins))
        n = len(t)
        count, out = np.zeros(n, np.int32), np.full((n, max_results), -1, np.int32)
        mk = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        abi.check(self.lib.avbd_world_intersect_shape(self.h, n, t.ctypes.data_as(abi.i32p), c.ctypes.data_as(abi.f32p),
                                                      q.ctypes.data_as(abi.f32p), s3.ctypes.data_as(abi.f32p),
                                                      m.ctypes.data_as(abi.f32p),
                                                      None if mk is None else mk.ctypes.data_as(abi.u32p), max_results,
                                                      count.ctypes.data_as(abi.i32p), out.ctypes.data_as(abi.i32p)))
        return count, out

    def cast_motion(self, types, centers, rotations, sizes, motions, margins, masks=None):
        """The server's cast_motion for n query shapes: dict of body (-1: none), safe, unsafe,
        point, normal."""
        t = np.ascontiguousarray(types, np.int32)
        c, q, s3, mv, m = (np.ascontiguousarray(x, np.float32) for x in (centers, rotations, sizes, motions, margins))
        n = len(t)
        body, frac = np.zeros(n, np.int32), np.zeros((n, 2), np.float32)
        point, normal = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        mk = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        abi.check(self.lib.avbd_world_cast_motion(self.h, n, t.ctypes.data_as(abi.i32p), c.ctypes.data_as(abi.f32p),
                                                  q.ctypes.data_as(abi.f32p), s3.ctypes.data_as(abi.f32p),
                                                  mv.ctypes.data_as(abi.f32p), m.ctypes.data_as(abi.f32p),
                                                  None if mk is None else mk.ctypes.data_as(abi.u32p),
                                                  body.ctypes.data_as(abi.i32p), frac.ctypes.data_as(abi.f32p),
                                                  point.ctypes.data_as(abi.f32p), normal.ctypes.data_as(abi.f32p)))
        return {"body": body, "safe": frac[:, 0], "unsafe": frac[:, 1], "point": point, "normal": normal}

    def _shapes(self, types, centers, rotations, sizes, margins, masks):
        t = _rows(types, 1, "types", np.int32)
        n = len(t)
        c, q, s3 = _rows(centers, 3, "centers", n=n), _rows(rotations, 4, "rotations", n=n), _rows(sizes, 3, "sizes", n=n)
        m = _rows(margins, 1, "margins", n=n)
        mk = None if masks is None else _rows(masks, 1, "masks", np.uint32, n)
        keep = (t, c, q, s3, m, mk)
        return keep, [n, t.ctypes.data_as(abi.i32p), c.ctypes.data_as(abi.f32p), q.ctypes.data_as(abi.f32p),
                      s3.ctypes.data_as(abi.f32p), m.ctypes.data_as(abi.f32p),
                      None if mk is None else mk.ctypes.data_as(abi.u32p)]

    def collide_shape(self, types, centers, rotations, sizes, margins, masks=None, max_pairs=8):
        """The server's collide_shape for n query shapes: (pair count, any, points n x max_pairs x 2 x 3)."""
        keep, args = self._shapes(types, centers, rotations, sizes, margins, masks)
        n = args[0]
        count, anyhit = np.zeros(n, np.int32), np.zeros(n, np.uint8)
        pts = np.zeros((n, max_pairs, 2, 3), np.float32)
        abi.check(self.lib.avbd_world_collide_shape(self.h, *args, int(max_pairs), count.ctypes.data_as(abi.i32p),
                                                    anyhit.ctypes.data_as(abi.u8p), pts.ctypes.data_as(abi.f32p)))
        return count, anyhit, pts

    def rest_info(self, types, centers, rotations, sizes, margins, masks=None):
        """The server's rest_info for n query shapes: (body or -1, point, normal, depth)."""
        keep, args = self._shapes(types, centers, rotations, sizes, margins, masks)
        n = args[0]
        body, point, normal, depth = (np.zeros(n, np.int32), np.zeros((n, 3), np.float32),
                                      np.zeros((n, 3), np.float32), np.zeros(n, np.float32))
        abi.check(self.lib.avbd_world_rest_info(self.h, *args, body.ctypes.data_as(abi.i32p),
                                                point.ctypes.data_as(abi.f32p), normal.ctypes.data_as(abi.f32p),
                                                depth.ctypes.data_as(abi.f32p)))
        return body, point, normal, depth

    def closest_point(self, bodies, points):
        """The server's closest point to each body's volume (body list indices, points n x 3)."""
        b = _rows(bodies, 1, "bodies", np.int32)
        p = _rows(points, 3, "points", n=len(b))
        out = np.zeros((len(b), 3), np.float32)
        abi.check(self.lib.avbd_world_closest_point(self.h, len(b), b.ctypes.data_as(abi.i32p),
                                                    p.ctypes.data_as(abi.f32p), out.ctypes.data_as(abi.f32p)))
        return out

    def collide_pair(self, a, b, collision_dt):
        """avbd::Manifold::collide for bodies a and b at their current pose on the device:
        (count, keys, rA, rB, basis 3x3 rows; zeros when the pair is beyond the margin)."""
        n = C.c_int32()
        keys, ra, rb = np.zeros(8, np.int32), np.zeros((8, 3), np.float32), np.zeros((8, 3), np.float32)
        bas = np.zeros(9, np.float32)
        abi.check(self.lib.avbd_world_collide_pair(self.h, int(a), int(b), float(collision_dt), C.byref(n),
                                                   keys.ctypes.data_as(abi.i32p), ra.ctypes.data_as(abi.f32p),
                                                   rb.ctypes.data_as(abi.f32p), bas.ctypes.data_as(abi.f32p)))
        k = n.value
        return k, keys[:k], ra[:k], rb[:k], bas.reshape(3, 3)

    def contact_report(self, max_reported):
        """Per-body contact reports: (count, position, normal, other), stride max(max_reported)."""
        mr = np.ascontiguousarray(max_reported, np.int32)
        stride = max(1, int(mr.max()) if len(mr) else 1)
        count, pos, nrm, other = abi.report_arrays(self.n, stride)
        abi.check(self.lib.avbd_world_contact_report(self.h, mr.ctypes.data_as(abi.i32p), stride,
                                                     count.ctypes.data_as(abi.i32p), pos.ctypes.data_as(abi.f32p),
                                                     nrm.ctypes.data_as(abi.f32p), other.ctypes.data_as(abi.i32p)))
        return count, pos, nrm, other

    def collect_aabb(self, mn, mx, mask=0xFFFFFFFF):
        """Solver::collectAABB: list indices of the bodies whose bounds overlap [mn, mx]."""
        out = np.zeros(max(self.n, 1), np.int32)
        count = C.c_int32()
        lo, hi = np.ascontiguousarray(mn, np.float32), np.ascontiguousarray(mx, np.float32)
        abi.check(self.lib.avbd_world_collect_aabb(self.h, lo.ctypes.data_as(abi.f32p), hi.ctypes.data_as(abi.f32p),
                                                   int(mask), out.ctypes.data_as(abi.i32p), len(out), C.byref(count)))
        return out[:count.value]

    def manifolds(self, capacity=None):
        cap = capacity or max(64, 8 * self.n)
        arr = abi.empty_arrays(abi.MANIFOLD_FIELDS, cap)
        ms = abi.manifolds_struct(arr)
        rc = self.lib.avbd_world_get_manifolds(self.h, C.byref(ms))
        if rc == abi.AVBD_ERR_CAPACITY:
            return self.manifolds(ms.count + 16)
        abi.check(rc)
        m = ms.count
        return {k: v[:m] for k, v in arr.items()}

    def record_residuals(self, on=True):
        """Keep the constraint values of each step's last dual (see residuals())."""
        abi.check(self.lib.avbd_world_record_residuals(self.h, 1 if on else 0))

    def residuals(self):
        """The last step's dual constraint values: (manifolds x 8 contact normal rows C.x in the
        order of manifolds(), joints x 3 linear rows in the order of download(joints=True))."""
        nm, nj = C.c_int32(0), C.c_int32(0)
        rc = self.lib.avbd_world_get_residuals(self.h, C.byref(nm), None, C.byref(nj), None)
        if rc != abi.AVBD_ERR_CAPACITY:
            abi.check(rc)
        cn = np.zeros((max(nm.value, 1), 8), np.float32)
        cl = np.zeros((max(nj.value, 1), 3), np.float32)
        abi.check(self.lib.avbd_world_get_residuals(self.h, C.byref(nm), cn.ctypes.data_as(abi.f32p), C.byref(nj),
                                                    cl.ctypes.data_as(abi.f32p)))
        return cn[:nm.value], cl[:nj.value]

    def set_manifolds(self, arrays):
        ms = abi.manifolds_struct(arrays)
        abi.check(self.lib.avbd_world_set_manifolds(self.h, C.byref(ms)))

    def stats(self):
        s = abi.StepStats()
        abi.check(self.lib.avbd_world_stats(self.h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in abi.StepStats._fields_}

    def colours(self):
        """Per body (list order): 0-7 colour, 8 Jacobi bucket, 9 unsolved overflow, 255 not updated."""
        out = np.zeros(self.n, np.uint8)
        abi.check(self.lib.avbd_world_colours(self.h, out.ctypes.data_as(abi.u8p)))
        return out

    def probe_broadphase(self, pair_capacity=None):
        n = self.n
        codes = np.zeros(n, np.uint64)
        order = np.zeros(n, np.int32)
        cap = pair_capacity or 16 * n + 64
        pairs = np.zeros(cap, np.uint64)
        cnt = C.c_int32(cap)
        abi.check(self.lib.avbd_world_probe_broadphase(self.h, codes.ctypes.data_as(abi.u64p),
                                                       order.ctypes.data_as(abi.i32p),
                                                       pairs.ctypes.data_as(abi.u64p), C.byref(cnt)))
        return codes, order, pairs[:cnt.value]

    def digest(self):
        st = self.download()
        return state_digest(st)

    def profile(self, steps):
        out = abi.KernelProfile()
        abi.check(self.lib.avbd_world_profile(self.h, int(steps), C.byref(out)))
        return out.as_dict()

    def autotune(self, steps_per_trial=2):
        out = abi.Tuning()
        abi.check(self.lib.avbd_world_autotune(self.h, int(steps_per_trial), C.byref(out)))
        return out.as_dict()


class Batch:
    """Many independent worlds stepped together (avbd_batch_*)."""

    def __init__(self, params, max_worlds, max_bodies, max_joints=0, max_manifolds=0):
        self.lib = abi.load_library()
        h = C.c_void_p()
        abi.check(self.lib.avbd_batch_create(C.byref(params), int(max_worlds), int(max_bodies),
                                             int(max_joints), int(max_manifolds), C.byref(h)))
        self.h = h
        self.sizes = {}

    def close(self):
        if self.h:
            self.lib.avbd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, bodies, joints=None, overlap=False):
        """Add a world (active on return).  overlap: upload it on the second stream, returning
        before the copies land (it may be called while a step is in flight); the world is
        stepped from the next step launch."""
        bs = abi.bodies_struct(bodies)
        js = abi.joints_struct(joints)
        wid = C.c_int32(-1)
        fn = self.lib.avbd_batch_add_world_async if overlap else self.lib.avbd_batch_add_world
        abi.check(fn(self.h, C.byref(bs), C.byref(js), C.byref(wid)))
        self.sizes[wid.value] = bs.count
        return wid.value

    def remove(self, wid):
        abi.check(self.lib.avbd_batch_remove_world(self.h, int(wid)))
        self.sizes.pop(wid, None)

    def step(self):
        """One step of every world.  Raises AvbdError with code AVBD_ERR_WORLD when some worlds
        failed (the others stepped): `failed_worlds()` lists them."""
        abi.check(self.lib.avbd_batch_step(self.h))

    def step_launch(self):
        """Queue one step and return; only add(..., overlap=True), fits and failed_worlds may be
        called before step_wait."""
        abi.check(self.lib.avbd_batch_step_launch(self.h))

    def step_wait(self):
        """Wait for the launched step; raises as step() does."""
        abi.check(self.lib.avbd_batch_step_wait(self.h))

    def failed_worlds(self):
        """[(world id, AVBD_WORLD_* reason bits)] of the worlds that failed the last step."""
        n = C.c_int32()
        abi.check(self.lib.avbd_batch_failed_worlds(self.h, 0, None, None, C.byref(n)))
        ids, why = np.zeros(max(n.value, 1), np.int32), np.zeros(max(n.value, 1), np.int32)
        abi.check(self.lib.avbd_batch_failed_worlds(self.h, n.value, ids.ctypes.data_as(abi.i32p),
                                                    why.ctypes.data_as(abi.i32p), C.byref(n)))
        return [(int(ids[k]), int(why[k])) for k in range(n.value)]

    def fits(self, n_bodies, n_joints=0):
        """(fits now, fits an empty batch) for a world of this size."""
        now, ever = C.c_int32(), C.c_int32()
        abi.check(self.lib.avbd_batch_fits(self.h, int(n_bodies), int(n_joints), C.byref(now), C.byref(ever)))
        return bool(now.value), bool(ever.value)

    def use_graph(self, enable):
        abi.check(self.lib.avbd_batch_use_graph(self.h, 1 if enable else 0))

    def use_query_tree(self, enable):
        abi.check(self.lib.avbd_batch_use_query_tree(self.h, 1 if enable else 0))

    def download(self, wid):
        n = self.sizes[wid]
        out = abi.empty_arrays({k: abi.BODY_FIELDS[k] for k in
                                ("position", "rotation", "velocity", "angular_velocity", "prev_velocity", "sleeping",
                                 "still_frames")}, n)
        bs = abi.bodies_struct({**out, "mass": np.zeros(n, np.float32)})
        bs.mass = None
        bs.count = n
        abi.check(self.lib.avbd_batch_download_world(self.h, int(wid), C.byref(bs), None))
        return out


    def raycast(self, wid, frm, to, masks=None, hit_from_inside=False):
        frm, to = np.ascontiguousarray(frm, np.float32), np.ascontiguousarray(to, np.float32)
        out, h = abi.ray_hits(len(frm))
        m = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        abi.check(self.lib.avbd_batch_raycast(self.h, int(wid), len(frm), frm.ctypes.data_as(abi.f32p),
                                              to.ctypes.data_as(abi.f32p), None if m is None else m.ctypes.data_as(abi.u32p),
                                              int(bool(hit_from_inside)), C.byref(h)))
        return out

    def manifolds(self, wid, capacity=None):
        cap = capacity or max(64, 8 * self.sizes[wid])
        arr = abi.empty_arrays(abi.MANIFOLD_FIELDS, cap)
        ms = abi.manifolds_struct(arr)
        rc = self.lib.avbd_batch_get_manifolds(self.h, int(wid), C.byref(ms))
        if rc == abi.AVBD_ERR_CAPACITY:
            return self.manifolds(wid, ms.count + 16)
        abi.check(rc)
        return {k: v[:ms.count] for k, v in arr.items()}

    def set_manifolds(self, wid, arrays):
        ms = abi.manifolds_struct(arrays)
        abi.check(self.lib.avbd_batch_set_manifolds(self.h, int(wid), C.byref(ms)))

    def autotune(self, steps_per_trial=2):
        out = abi.Tuning()
        abi.check(self.lib.avbd_batch_autotune(self.h, int(steps_per_trial), C.byref(out)))
        return out.as_dict()

    def set_tuning(self, tuning):
        """Pin launch shapes (a dict as returned by autotune())."""
        t = abi.Tuning(**{k: int(v) for k, v in tuning.items()})
        abi.check(self.lib.avbd_batch_set_tuning(self.h, C.byref(t)))

    def profile(self, steps):
        out = abi.KernelProfile()
        abi.check(self.lib.avbd_batch_profile(self.h, int(steps), C.byref(out)))
        return out.as_dict()

    def list_stats(self):
        """The broad phase's candidate lists in the last step."""
        out = np.zeros(4, np.int32)
        abi.check(self.lib.avbd_batch_list_stats(self.h, out.ctypes.data_as(abi.i32p)))
        return dict(zip(("moved_bodies", "busy_worlds", "exact_walk_worlds", "relisted_worlds"), out.tolist()))

    def totals(self):
        v = [C.c_int64() for _ in range(4)]
        abi.check(self.lib.avbd_batch_totals(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("movable_bodies", "manifolds", "contacts", "joints"), [x.value for x in v]))


def state_digest(state):
    """FNV-1a over position, rotation, velocity, angular velocity per body (list order),
    as the reference's test digest (test/core_test.cpp:133-149)."""
    n = len(state["position"])
    buf = np.concatenate([state["position"].astype(np.float32).reshape(n, 3).view(np.uint8),
                          state["rotation"].astype(np.float32).reshape(n, 4).view(np.uint8),
                          state["velocity"].astype(np.float32).reshape(n, 3).view(np.uint8),
                          state["angular_velocity"].astype(np.float32).reshape(n, 3).view(np.uint8)], axis=1)
    h = 1469598103934665603
    for byte in buf.reshape(-1).tolist():
        h ^= byte
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h

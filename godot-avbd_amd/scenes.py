"""Scene builders: body and joint arrays in the reference's list order.

A scene is built in creation order (the order a reference user calls `new Rigid` /
`new Joint`) and then reversed, because the reference prepends each new body and force to
its intrusive list (src/avbd/rigid.cpp:28-30, force.cpp:19-21): list index 0 is the body
created last.  Mass, inertia and bounding radius follow the reference's box constructor
(src/avbd/rigid.cpp:59-70) in float32.

Builders:
  digest_stack        ground + 10 cubes (test/core_test.cpp:391-399)
  digest_mixed        30-link Joint(1000, 250) chain + 40 cubes (test/core_test.cpp:612-631)
  pyramid             C1: 28-row triangular stack of 406 unit cubes on a 100x1x100 ground
  pile                C2: N boxes, U[0.25,1] half extents, log-normal density, random
                      orientation, dropped from a jittered 3-D grid
  chains              C3: chains of unit boxes linked by ball joints hanging from anchors
  batch_piles         C4: many independent piles with log-uniform body counts
"""
import math

import numpy as np

F = np.float32


class SceneBuilder:
    def __init__(self, gravity=-10.0):
        self.gravity = F(gravity)
        self.b = []   # dicts in creation order
        self.j = []

    def box(self, size, density, friction, pos, rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0)):
        size = np.asarray(size, dtype=F)
        d = F(density)
        m = size[0] * size[1] * size[2] * d
        moment = np.array([
            (size[1] * size[1] + size[2] * size[2]) / F(12.0) * m,
            (size[0] * size[0] + size[2] * size[2]) / F(12.0) * m,
            (size[0] * size[0] + size[1] * size[1]) / F(12.0) * m], dtype=F)
        h = size * F(0.5)
        radius = np.sqrt(((F(0.0) + h[0] * h[0]) + h[1] * h[1]) + h[2] * h[2])
        self.b.append(dict(size=size, mass=F(m), moment=moment, friction=F(friction), radius=F(radius),
                           gravity=self.gravity, position=np.asarray(pos, dtype=F),
                           rotation=np.asarray(rot, dtype=F), velocity=np.asarray(vel, dtype=F)))
        return len(self.b) - 1

    def sphere(self, radius, density, friction, pos, rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0)):
        """Reference `Rigid(..., ShapeType::Sphere, ...)` (rigid.cpp:36-44): size = (r, 0, 0)."""
        r, d = F(radius), F(density)
        m = F(4.0 / 3.0) * F(3.14159265358979) * r * r * r * d
        i = F(0.4) * m * r * r
        return self._add(np.array([r, 0, 0], F), m, np.array([i, i, i], F), friction, r, pos, rot, vel, 1)

    def cylinder(self, radius, height, density, friction, pos, rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0)):
        """Reference `Rigid(..., ShapeType::Cylinder, ...)` (rigid.cpp:46-57): size = (r, r, h),
        axis local +Z."""
        r, h, d = F(radius), F(height), F(density)
        m = F(3.14159265358979) * r * r * h * d
        axial = F(0.5) * m * r * r
        radial = m * (F(3.0) * r * r + h * h) / F(12.0)
        hh = h * F(0.5)
        rad = np.sqrt(((F(0.0) + r * r) + hh * hh) + r * r)
        return self._add(np.array([r, r, h], F), m, np.array([radial, radial, axial], F), friction, rad, pos,
                         rot, vel, 2)

    def _add(self, size, m, moment, friction, radius, pos, rot, vel, shape):
        self.b.append(dict(size=size, mass=F(m), moment=moment, friction=F(friction), radius=F(radius),
                           gravity=self.gravity, position=np.asarray(pos, dtype=F),
                           rotation=np.asarray(rot, dtype=F), velocity=np.asarray(vel, dtype=F), shape=shape))
        return len(self.b) - 1

    def joint(self, a, b, r_a, r_b, k_lin=math.inf, k_ang=0.0, fracture=math.inf):
        sa = self.b[a]["size"] if a is not None and a >= 0 else np.zeros(3, F)
        s = sa + self.b[b]["size"]
        arm = ((F(0.0) + s[0] * s[0]) + s[1] * s[1]) + s[2] * s[2]
        self.j.append(dict(a=-1 if a is None else a, b=b, r_a=np.asarray(r_a, F), r_b=np.asarray(r_b, F),
                           k_lin=F(k_lin), k_ang=F(k_ang), fracture=F(fracture), arm=F(arm)))
        return len(self.j) - 1

    def spring(self, a, b, r_a, r_b, stiffness, rest=-1.0):
        """Reference `Spring` (spring.cpp:16-28) between bodies a and b; a negative rest length
        takes the anchors' current distance (computed here in float64, then rounded)."""
        r_a, r_b = np.asarray(r_a, F), np.asarray(r_b, F)
        if rest < 0:
            pa = self.b[a]["position"].astype(np.float64) + _qrot64(self.b[a]["rotation"], r_a)
            pb = self.b[b]["position"].astype(np.float64) + _qrot64(self.b[b]["rotation"], r_b)
            rest = float(np.linalg.norm(pa - pb))
        self.j.append(dict(kind=1, a=a, b=b, r_a=r_a, r_b=r_b, k_lin=F(stiffness), k_ang=F(0.0),
                           fracture=F(math.inf), arm=F(0.0), rest=F(rest)))
        return len(self.j) - 1

    def generic(self, a, b, r_a, r_b, linear=None, angular=None):
        """Reference `GenericJoint` (generic_joint.cpp:24-30): rest = conj(frame) * rot_b with
        frame = rot_a (identity when a is None), in float32 with the reference's product order.
        linear / angular: three (mode, lower, upper, spring, equilibrium, damping) tuples
        (default: every axis free)."""
        qa = self.b[a]["rotation"] if a is not None and a >= 0 else np.array([0, 0, 0, 1], F)
        rest = _qmul32(np.array([-qa[0], -qa[1], -qa[2], qa[3]], F), self.b[b]["rotation"])
        axes = list(linear or [(0, 0, 0, 0, 0, 0)] * 3) + list(angular or [(0, 0, 0, 0, 0, 0)] * 3)
        z = np.zeros(3, F)
        self.j.append(dict(kind=3, a=-1 if a is None else a, b=b, r_a=np.asarray(r_a, F), r_b=np.asarray(r_b, F),
                           k_lin=F(0.0), k_ang=F(0.0), fracture=F(math.inf), arm=F(0.0), rest=F(0.0),
                           rest_rotation=rest, axes=[tuple(x) for x in axes]))
        return len(self.j) - 1

    def ignore_collision(self, a, b):
        """Reference `IgnoreCollision` (forces.hpp:104-113): the pair skips the broad phase."""
        z = np.zeros(3, F)
        self.j.append(dict(kind=2, a=a, b=b, r_a=z, r_b=z, k_lin=F(0.0), k_ang=F(0.0),
                           fracture=F(math.inf), arm=F(0.0), rest=F(0.0)))
        return len(self.j) - 1

    def build(self):
        """(bodies, joints) arrays in list order."""
        n = len(self.b)
        order = list(range(n - 1, -1, -1))
        pos_of = {c: n - 1 - c for c in range(n)}
        bodies = {
            "size": np.stack([self.b[c]["size"] for c in order]) if n else np.zeros((0, 3), F),
            "mass": np.array([self.b[c]["mass"] for c in order], F),
            "moment": np.stack([self.b[c]["moment"] for c in order]) if n else np.zeros((0, 3), F),
            "friction": np.array([self.b[c]["friction"] for c in order], F),
            "radius": np.array([self.b[c]["radius"] for c in order], F),
            "gravity": np.array([self.b[c]["gravity"] for c in order], F),
            "position": np.stack([self.b[c]["position"] for c in order]) if n else np.zeros((0, 3), F),
            "rotation": np.stack([self.b[c]["rotation"] for c in order]) if n else np.zeros((0, 4), F),
            "velocity": np.stack([self.b[c]["velocity"] for c in order]) if n else np.zeros((0, 3), F),
            "angular_velocity": (np.stack([np.asarray(self.b[c].get("ang", (0.0, 0.0, 0.0)), F) for c in order])
                                 if n else np.zeros((0, 3), F)),
        }
        bodies["prev_velocity"] = bodies["velocity"].copy()
        if any(self.b[c].get("shape", 0) for c in order):
            bodies["shape"] = np.array([self.b[c].get("shape", 0) for c in order], np.uint8)
        if any(self.b[c].get("sleep_mode", 0) for c in order):
            bodies["sleep_mode"] = np.array([self.b[c].get("sleep_mode", 0) for c in order], np.uint8)
        nj = len(self.j)
        jorder = list(range(nj - 1, -1, -1))
        joints = None
        if nj:
            joints = {
                "body_a": np.array([pos_of[self.j[c]["a"]] if self.j[c]["a"] >= 0 else -1 for c in jorder], np.int32),
                "body_b": np.array([pos_of[self.j[c]["b"]] for c in jorder], np.int32),
                "r_a": np.stack([self.j[c]["r_a"] for c in jorder]),
                "r_b": np.stack([self.j[c]["r_b"] for c in jorder]),
                "stiffness_lin": np.array([self.j[c]["k_lin"] for c in jorder], F),
                "stiffness_ang": np.array([self.j[c]["k_ang"] for c in jorder], F),
                "fracture": np.array([self.j[c]["fracture"] for c in jorder], F),
                "torque_arm": np.array([self.j[c]["arm"] for c in jorder], F),
                "kind": np.array([self.j[c].get("kind", 0) for c in jorder], np.uint8),
                "rest": np.array([self.j[c].get("rest", 0.0) for c in jorder], F),
            }
            if any(self.j[c].get("kind", 0) == 3 for c in jorder):
                none = [(0, 0, 0, 0, 0, 0)] * 6
                joints["rest_rotation"] = np.stack([self.j[c].get("rest_rotation", np.array([0, 0, 0, 1], F))
                                                    for c in jorder]).astype(F)
                ax = np.array([self.j[c].get("axes", none) for c in jorder], np.float64)   # (nj, 6, 6)
                joints["axis_mode"] = ax[:, :, 0].astype(np.uint8)
                for q, name in enumerate(("axis_lower", "axis_upper", "axis_spring", "axis_equilibrium",
                                          "axis_damping"), start=1):
                    joints[name] = ax[:, :, q].astype(F)
                for name in ("axis_lambda", "axis_penalty", "axis_initial"):
                    joints[name] = np.zeros((nj, 6), F)
            for k in ("c0_lin", "c0_ang", "penalty_lin", "penalty_ang", "lambda_lin", "lambda_ang"):
                joints[k] = np.zeros((nj, 3), F)
            joints["broken"] = np.zeros(nj, np.uint8)
        return bodies, joints


def _qmul32(a, b):
    """Hamilton product in float32 with the reference's term order (maths.h:87-95)."""
    a = np.asarray(a, F)
    b = np.asarray(b, F)
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([((aw * bx + ax * bw) + ay * bz) - az * by, ((aw * by + ay * bw) + az * bx) - ax * bz,
                     ((aw * bz + az * bw) + ax * by) - ay * bx, ((aw * bw - ax * bx) - ay * by) - az * bz], F)


def _qrot64(q, v):
    x, y, z, w = (float(t) for t in q)
    u = np.array([x, y, z])
    v = np.asarray(v, np.float64)
    return 2.0 * np.dot(u, v) * u + (w * w - np.dot(u, u)) * v + 2.0 * w * np.cross(u, v)


def springs_and_filters(seed=5):
    """Forces of every kind in one world: a pile whose boxes are linked in pairs by springs
    (some with off-centre anchors), boxes hanging from static anchors by springs, jointed
    pairs, and collision filters between boxes that start interpenetrating (they pass through
    each other) and between a box and the ground (it falls through)."""
    rng = np.random.default_rng(seed)
    s = SceneBuilder()
    ground = s.box((40, 40, 1), 0.0, 0.5, (0, 0, 0))
    ids = []
    for i in range(24):
        p = (F(rng.uniform(-4, 4)), F(rng.uniform(-4, 4)), F(1.5 + 1.1 * i / 3))
        ids.append(s.box((1, 1, 1), 1.0, 0.5, p))
    for k in range(0, 20, 2):
        off = (0.3, 0.0, 0.2) if k % 4 == 0 else (0.0, 0.0, 0.0)
        s.spring(ids[k], ids[k + 1], off, (0, 0, 0), 200.0 + 50.0 * k)
    s.joint(ids[20], ids[21], (0.6, 0, 0), (-0.6, 0, 0))
    for h in range(3):
        anchor = s.box((0.5, 0.5, 0.5), 0.0, 0.5, (F(8 + 2 * h), 0, F(8)))
        bob = s.box((1, 1, 1), 1.0, 0.5, (F(8 + 2 * h), 0, F(6)))
        s.spring(anchor, bob, (0, 0, 0), (0.2 * h, 0, 0.5), 300.0, rest=1.0)
    a = s.box((1, 1, 1), 1.0, 0.5, (F(-8), 0, F(2)))
    b = s.box((1, 1, 1), 1.0, 0.5, (F(-7.6), 0, F(2.3)))
    s.ignore_collision(a, b)
    faller = s.box((1, 1, 1), 1.0, 0.5, (F(-12), 0, F(1.2)))
    s.ignore_collision(ground, faller)
    return s.build()


def generic_joints():
    """The reference's GenericJoint rigs (core_test.cpp:703-1000) side by side in one world: a
    hinge spinning about its free axle, a prismatic strut with a spring, an axle with a ±0.4 rad
    limit, a ball socket anchored off the body's centre and swinging, one joint to a world point
    (no body A), and one built from a rotated body; then a hinged chain between dynamic bodies
    dropped onto a ground box."""
    L, K, M = 1, 2, 0   # locked, limited, free
    lock = (L, 0, 0, 0, 0, 0)
    free = (M, 0, 0, 0, 0, 0)
    s = SceneBuilder()
    rigs = []
    for r in range(6):
        x = F(6.0 * r)
        anchor = s.box((1, 1, 1), 0.0, 0.5, (x, 0, 10))
        rigs.append((anchor, x))
    # hinge: axle Y free
    a, x = rigs[0]
    body = s.box((1, 1, 1), 1.0, 0.5, (x, 0, 10), vel=(0, 0, 0))
    s.b[body]["ang"] = (0.0, 3.0, 0.0)
    s.generic(a, body, (0, 0, 0), (0, 0, 0), [lock] * 3, [lock, free, lock])
    # prismatic strut: linear Z limited with a spring
    a, x = rigs[1]
    body = s.box((1, 1, 1), 1.0, 0.5, (x, 0, 10))
    s.generic(a, body, (0, 0, 0), (0, 0, 0), [lock, lock, (K, -2.0, 2.0, 500.0, 0.0, 5.0)], [lock] * 3)
    # limited axle
    a, x = rigs[2]
    body = s.box((1, 1, 1), 1.0, 0.5, (x, 0, 10))
    s.b[body]["ang"] = (0.0, 6.0, 0.0)
    s.generic(a, body, (0, 0, 0), (0, 0, 0), [lock] * 3, [lock, (K, -0.4, 0.4, 0, 0, 0), lock])
    # ball socket off the centre, swinging
    a, x = rigs[3]
    body = s.box((1, 1, 1), 1.0, 0.5, (x, 0, 9), vel=(1.5, 0, 0))
    s.generic(a, body, (0, 0, 0), (0, 0, 1.0), [lock] * 3, [free] * 3)
    # to a world point, with an angular spring about Z
    _, x = rigs[4]
    body = s.box((1, 1, 1), 1.0, 0.5, (x, 0, 8))
    s.b[body]["ang"] = (0.0, 0.0, 2.0)
    s.generic(None, body, (x, 0, 9), (0, 0, 1.0), [lock] * 3, [free, free, (M, 0, 0, 40.0, 0.0, 1.0)])
    # a body created rotated (rest orientation not the identity), axle X free
    a, x = rigs[5]
    body = s.box((1.0, 0.6, 0.6), 1.0, 0.5, (x, 0, 9), rot=(0.0, -0.38268343, 0.0, 0.9238795))
    s.b[body]["ang"] = (4.0, 0.0, 0.0)
    s.generic(a, body, (0, 0, -1.0), (0, 0, 0), [lock] * 3, [free, lock, lock])
    # a chain of hinged links between dynamic bodies, dropped onto a ground box: joints and
    # contacts on the same bodies, the middle hinge limited
    s.box((40, 10, 1), 0.0, 0.5, (8.0, -8.0, -0.5))
    prev = None
    for i in range(6):
        link = s.box((1.0, 0.4, 0.4), 1.0, 0.5, (F(i) * F(1.1), -8.0, F(2.0) + F(0.3) * F(i)),
                     rot=(0.0, 0.0, 0.0998334, 0.9950042))
        if prev is not None:
            hinge = (K, -0.6, 0.6, 0, 0, 0) if i == 3 else free
            s.generic(prev, link, (0.55, 0, 0), (-0.55, 0, 0), [lock] * 3, [lock, hinge, lock])
        prev = link
    return s.build()


def digest_stack():
    s = SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    for i in range(10):
        s.box((1, 1, 1), 1.0, 0.5, (0, 0, F(i) * F(1.5) + F(1.0)))
    return s.build()


def digest_mixed():
    s = SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    prev = None
    for i in range(30):
        link = s.box((1, 1, 1), 1.0, 0.5, (0, 0, F(20.0) - F(i) * F(1.5)))
        if prev is not None:
            s.joint(prev, link, (0, 0, -0.5), (0, 0, 0.5), 1000.0, 250.0)
        prev = link
    for i in range(40):
        s.box((1, 1, 1), 1.0, 0.5, (F(5.0) + F(i) * F(1.2), 0, F(3.0) + F(i % 5)))
    return s.build()


def hard_chain():
    """test/core_test.cpp:271-300: 4 links on infinitely stiff ball joints under an anchor."""
    s = SceneBuilder()
    anchor = s.box((1, 1, 1), 0.0, 0.5, (0, 0, 10))
    prev = anchor
    for i in range(4):
        link = s.box((1, 1, 1), 1.0, 0.5, (0, 0, F(9.0) - F(i)))
        s.joint(prev, link, (0, 0, -0.5), (0, 0, 0.5), math.inf, 0.0)
        prev = link
    return s.build()


def _shoemake(rng, n):
    u = rng.random((n, 3))
    q = np.stack([np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]),
                  np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
                  np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]),
                  np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])], axis=1)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(F)


def pyramid(rows=28, seed=1):
    """C1: `rows`-row triangular stack (406 unit cubes for 28 rows) on a 100x1x100 ground."""
    rng = np.random.default_rng(seed)
    s = SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, -0.5))
    for r in range(rows):
        for i in range(rows - r):
            jx, jy = (rng.random(2) * 2.0 - 1.0) * 1e-3
            x = (i - (rows - 1 - r) / 2.0) * 1.002 + jx
            s.box((1, 1, 1), 1.0, 0.5, (x, jy, 0.5 + r))
    return s.build()


def box_arrays(size, density, friction, pos, rot, gravity=-10.0):
    """Vectorised form of SceneBuilder.box for many boxes (same float32 expressions)."""
    size = np.asarray(size, F)
    d = np.asarray(density, F)
    m = size[:, 0] * size[:, 1] * size[:, 2] * d
    moment = np.stack([(size[:, 1] * size[:, 1] + size[:, 2] * size[:, 2]) / F(12.0) * m,
                       (size[:, 0] * size[:, 0] + size[:, 2] * size[:, 2]) / F(12.0) * m,
                       (size[:, 0] * size[:, 0] + size[:, 1] * size[:, 1]) / F(12.0) * m], axis=1).astype(F)
    h = size * F(0.5)
    radius = np.sqrt(((F(0.0) + h[:, 0] * h[:, 0]) + h[:, 1] * h[:, 1]) + h[:, 2] * h[:, 2])
    n = len(m)
    return dict(size=size, mass=m.astype(F), moment=moment, friction=np.full(n, friction, F) if np.isscalar(friction)
                else np.asarray(friction, F), radius=radius.astype(F), gravity=np.full(n, gravity, F),
                position=np.asarray(pos, F), rotation=np.asarray(rot, F), velocity=np.zeros((n, 3), F),
                angular_velocity=np.zeros((n, 3), F))


def pile(n=8192, seed=2, grid=None, pitch=4.0, ground=400.0):
    """C2: n boxes with half extents U[0.25, 1], log-normal density (sigma 0.7, clipped to
    2 sigma), uniform random orientation, on a jittered 3-D grid (lowest centre z = 2), created
    after a static ground box; body k of the grid is at x = k % nx, y = (k // nx) % ny,
    z = k // (nx ny)."""
    rng = np.random.default_rng(seed)
    if grid is None:
        nx = max(1, int(round((n / 2) ** (1.0 / 3.0))))
        ny = nx
        nz = int(math.ceil(n / (nx * ny)))
        grid = (nx, ny, nz)
    nx, ny, nz = grid
    half = rng.uniform(0.25, 1.0, size=(n, 3))
    dens = np.exp(np.clip(rng.normal(0.0, 0.7, size=n), -1.4, 1.4))
    quats = _shoemake(rng, n)
    jit = rng.uniform(-0.25, 0.25, size=(n, 3))
    k = np.arange(n)
    x, y, z = k % nx, (k // nx) % ny, k // (nx * ny)
    pos = np.stack([(x - (nx - 1) / 2.0) * pitch + jit[:, 0], (y - (ny - 1) / 2.0) * pitch + jit[:, 1],
                    2.0 + z * pitch + jit[:, 2]], axis=1)
    boxes = box_arrays((2.0 * half).astype(F), dens.astype(F), 0.5, pos.astype(F), quats)
    gnd = box_arrays(np.array([[ground, ground, 1.0]], F), np.zeros(1, F), 0.5, np.array([[0, 0, -0.5]], F),
                     np.array([[0, 0, 0, 1]], F))
    # creation order: ground, then boxes; list order is the reverse
    bodies = {key: np.concatenate([gnd[key], boxes[key]])[::-1].copy() for key in gnd}
    bodies["prev_velocity"] = bodies["velocity"].copy()
    return bodies, None


def sleepy_pile(n=240, seed=6, modes=(1, 1, 1, 0, 2)):
    """A settling pile with sleep enabled: sleep modes cycle through `modes` (IDLE, NEVER,
    START_ASLEEP), plus a late box held high above the pile (it falls when the pile sleeps and
    wakes the sleepers it lands on) and two boxes created asleep."""
    b, j = pile(n, seed=seed, pitch=2.2)
    m = len(b["mass"])
    b["sleep_mode"] = np.array([modes[k % len(modes)] if b["mass"][k] > 0 else 0 for k in range(m)], np.uint8)
    b["sleeping"] = np.zeros(m, np.uint8)
    b["sleeping"][[1, 4]] = 1
    b["position"][0] = (0.3, 0.2, 40.0)     # the last-created box: high above the pile
    b["velocity"][0] = (0.0, 0.0, 0.0)
    b["prev_velocity"][0] = (0.0, 0.0, 0.0)
    return b, j


def round_pile(n=300, seed=4, pitch=2.2):
    """Boxes, spheres and capped cylinders (a third each, random sizes, densities and
    orientations) dropped on a jittered grid onto a static ground box, plus a sphere resting on
    a sphere, a cylinder on its side given spin and a static cylinder and sphere as obstacles."""
    rng = np.random.default_rng(seed)
    s = SceneBuilder()
    s.box((200, 200, 1), 0.0, 0.5, (0, 0, -0.5))
    s.cylinder(1.0, 3.0, 0.0, 0.6, (14.0, 0.0, 1.5))
    s.sphere(1.5, 0.0, 0.6, (-14.0, 0.0, 1.0))
    s.sphere(0.5, 1000.0, 0.5, (20.0, 20.0, 1.0))
    s.sphere(0.5, 1000.0, 0.5, (20.0, 20.0, 2.5))
    roll = s.cylinder(0.5, 1.0, 500.0, 0.9, (-20.0, 20.0, 0.52), rot=(0.70710677, 0.0, 0.0, 0.70710677))
    s.b[roll]["ang"] = (8.0, 0.0, 0.0)
    nx = max(1, int(round(n ** (1.0 / 3.0))))
    q = _shoemake(rng, n)
    for k in range(n):
        x, y, z = k % nx, (k // nx) % nx, k // (nx * nx)
        pos = ((x - (nx - 1) / 2.0) * pitch + rng.uniform(-0.2, 0.2), (y - (nx - 1) / 2.0) * pitch + rng.uniform(-0.2, 0.2),
               2.0 + z * pitch + rng.uniform(-0.2, 0.2))
        dens = float(np.exp(np.clip(rng.normal(0.0, 0.5), -1.0, 1.0)))
        kind = k % 3
        if kind == 0:
            s.box(rng.uniform(0.4, 1.4, 3), dens, 0.5, pos, rot=q[k])
        elif kind == 1:
            s.sphere(rng.uniform(0.25, 0.7), dens, 0.5, pos, rot=q[k])
        else:
            s.cylinder(rng.uniform(0.25, 0.6), rng.uniform(0.4, 1.4), dens, 0.5, pos, rot=q[k])
    return s.build()


def chains(n_chains=4096, links=32, seed=3, grid=None, pitch_y=1.5, pitch_x=34.0, anchor_z=30.0):
    """C3: chains of unit boxes along +x linked by ball joints (k_lin = inf, k_ang = 0),
    each hanging from a static anchor, with a ground box below."""
    rng = np.random.default_rng(seed)
    if grid is None:
        gx = max(1, int(round(math.sqrt(n_chains))))
        gy = int(math.ceil(n_chains / gx))
        grid = (gx, gy)
    gx, gy = grid
    s = SceneBuilder()
    span_x = gx * pitch_x + 2 * links
    span_y = gy * pitch_y + 10
    s.box((span_x, span_y, 1), 0.0, 0.5, ((gx - 1) * pitch_x / 2.0, (gy - 1) * pitch_y / 2.0, -0.5))
    c = 0
    for iy in range(gy):
        for ix in range(gx):
            if c >= n_chains:
                break
            x0, y0 = ix * pitch_x, iy * pitch_y
            prev = s.box((1, 1, 1), 0.0, 0.5, (x0, y0, anchor_z))
            for k in range(links):
                jx, jy, jz = rng.uniform(-1e-3, 1e-3, size=3)
                link = s.box((1, 1, 1), 1.0, 0.5, (x0 + 1.0 + k + jx, y0 + jy, anchor_z + jz))
                s.joint(prev, link, (0.5, 0, 0), (-0.5, 0, 0))
                prev = link
            c += 1
    return s.build()


def batch_sizes(worlds=4096, lo=8, hi=2048, seed=4):
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), size=worlds))).astype(int)


def batch_piles(worlds=4096, lo=8, hi=2048, seed=4):
    """C4: independent C2-style piles with log-uniform body counts; world w uses seed 4000+w."""
    sizes = batch_sizes(worlds, lo, hi, seed)
    return [pile(int(n), seed=4000 + w, pitch=4.0, ground=max(40.0, 8.0 * math.sqrt(n))) for w, n in enumerate(sizes)]


def greedy_colours(bodies, joints, manifolds, contact_only_overflow=True, max_colours=8):
    """Colouring of solver_solve.cpp:365-424 restated from its definition: movable bodies in list
    order each take the smallest colour in [0, 8) not used by a lower-ordered neighbour (a body
    sharing a joint or a manifold), else overflow; bodies touching no joint are moved to the
    Jacobi bucket (8) when contact_only_overflow; an overflow body that is not contact-only is
    left unsolved (9).  Bodies outside the update set get 255."""
    n = len(bodies["mass"])
    movable = np.asarray(bodies["mass"]) > 0
    order = np.nonzero(movable)[0]
    rank = np.full(n, -1)
    rank[order] = np.arange(len(order))
    nbr = [set() for _ in range(n)]
    jointed = np.zeros(n, bool)
    pairs = []
    if joints is not None:
        for a, b in zip(joints["body_a"], joints["body_b"]):
            pairs.append((int(a), int(b)))
            if a >= 0:
                jointed[a] = True
            jointed[b] = True
    for a, b in zip(manifolds["body_a"], manifolds["body_b"]):
        pairs.append((int(a), int(b)))
    for a, b in pairs:
        if a >= 0 and b >= 0:
            nbr[a].add(b)
            nbr[b].add(a)
    out = np.full(n, 255, np.uint8)
    colour = {}
    for i in order:
        used = {colour[o] for o in nbr[i] if o in colour and rank[o] >= 0 and rank[o] < rank[i]}
        c = next((k for k in range(max_colours) if k not in used), max_colours)
        colour[i] = c
    for i in order:
        c = colour[i]
        if contact_only_overflow and not jointed[i]:
            out[i] = 8
        else:
            out[i] = c if c < max_colours else 9
    return out


def world_energy(state):
    """Mechanical energy of a world's movable bodies, E = sum 1/2 m |v|^2 + 1/2 w^T (R diag(I) R^T) w
    - m g z (in float64; the reference has no energy metric, so both implementations use this)."""
    m = np.asarray(state["mass"], np.float64)
    mov = m > 0
    v = np.asarray(state["velocity"], np.float64).reshape(-1, 3)
    w = np.asarray(state["angular_velocity"], np.float64).reshape(-1, 3)
    inertia = np.asarray(state["moment"], np.float64).reshape(-1, 3)
    rot = np.asarray(state["rotation"], np.float64).reshape(-1, 4)
    x, y, z, s_ = rot[:, 0], rot[:, 1], rot[:, 2], rot[:, 3]
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - s_ * z), 2 * (x * z + s_ * y)], -1),
                  np.stack([2 * (x * y + s_ * z), 1 - 2 * (x * x + z * z), 2 * (y * z - s_ * x)], -1),
                  np.stack([2 * (x * z - s_ * y), 2 * (y * z + s_ * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    wl = np.einsum("nji,nj->ni", R, w)   # body-frame angular velocity
    ke = 0.5 * m * (v * v).sum(1) + 0.5 * (inertia * wl * wl).sum(1)
    pe = -m * np.asarray(state["gravity"], np.float64) * np.asarray(state["position"], np.float64).reshape(-1, 3)[:, 2]
    return float((ke + pe)[mov].sum())


def world_statistics(state, manifolds, residuals=None, initial_state=None):
    """The long-run statistics compared between implementations:
    * penetration depth per contact, max(0, 1 mm - C0.x) after the rebase
      (avbd_physics_server3d.cpp:1547-1560), as p50/p95/p99/max;
    * constraint residual |C| from the step's last dual (`residuals` = World.residuals() with
      residual recording on): mean and max |C_n| over contacts (manifold.cpp:626-630) and of
      |C_lin| over joints (joint.cpp:176-182); without `residuals`, |C0.x| after the rebase;
    * energy E (world_energy) and, given the run's `initial_state`, its drift E - E(initial)."""
    num = np.asarray(manifolds["num_contacts"])
    c0 = np.asarray(manifolds["c0"]).reshape(len(num), 8, 3)
    used = np.arange(8)[None, :] < num[:, None]
    c0n = c0[..., 0][used]
    depth = np.maximum(0.0, 1.0e-3 - c0n.astype(np.float64))
    q = lambda a, k: float(np.percentile(a, k)) if len(a) else 0.0
    out = {"depth_p50": q(depth, 50), "depth_p95": q(depth, 95), "depth_p99": q(depth, 99),
           "depth_max": float(depth.max()) if len(depth) else 0.0, "contacts": int(num.sum())}
    if residuals is not None:
        cn, cl = residuals
        rn = np.abs(np.asarray(cn, np.float64).reshape(len(num), 8)[used])
        rj = np.linalg.norm(np.asarray(cl, np.float64).reshape(-1, 3), axis=1)
        out["residual_source"] = "dual"
    else:
        rn = np.abs(c0n.astype(np.float64))
        rj = np.zeros(0)
        out["residual_source"] = "c0"
    out.update({"residual_mean": float(rn.mean()) if len(rn) else 0.0,
                "residual_max": float(rn.max()) if len(rn) else 0.0,
                "joint_residual_mean": float(rj.mean()) if len(rj) else 0.0,
                "joint_residual_max": float(rj.max()) if len(rj) else 0.0,
                "energy": world_energy(state)})
    if initial_state is not None:
        out["energy_drift"] = out["energy"] - world_energy(initial_state)
    return out

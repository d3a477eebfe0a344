"""CPU tests that pin the oracle (the serial restatement of the reference step) before it
is trusted as the checker of the HIP path.

Pins, each chosen so that a plausible restatement mistake fails one of them:
  - the reference's own whole-step digests (tools/baseline/digests.txt), which exercise
    contacts, the brake, projection, the Jacobi primal, colour Gauss-Seidel and
    finite-stiffness joints (test/core_test.cpp:391-414, 612-631);
  - narrow-phase outputs of the reference's own collide.cpp (golden fixture, and live when
    oracle/_ref is built);
  - the reference's own 6x6 solves (live when oracle/_ref is built) and a float64 solve;
  - Morton codes against the reference's own encoder, candidate pairs against brute force;
  - the reference's closed-form scenario checks (test/core_test.cpp:155-300, 1624-1638).
"""
import math
import os

import numpy as np
import pytest

from helpers import (DIGEST_INTERLEAVED_300, DIGEST_MIXED_240, DIGEST_STACK_300, ROOT, dt120, oracle,
                     pkg, scenes)

P = pkg.default_params


def run(bodies, joints, params, steps, inter=None):
    o = oracle.OracleWorld(params)
    o.upload(bodies, joints)
    for _ in range(steps):
        if inter:
            o.step_interleaved(*inter)
        else:
            o.step()
    return o


def test_digest_stack_300():
    assert run(*scenes.digest_stack(), P(), 300).digest() == DIGEST_STACK_300


def test_digest_interleaved_300():
    p = P()
    p.dt = dt120()
    assert run(*scenes.digest_stack(), p, 300, inter=(2, 3)).digest() == DIGEST_INTERLEAVED_300


def test_digest_mixed_240():
    assert run(*scenes.digest_mixed(), P(), 240).digest() == DIGEST_MIXED_240


def test_python_digest_matches_oracle_digest():
    o = run(*scenes.digest_stack(), P(), 30)
    assert pkg.state_digest(o.download()) == o.digest()


def _fixture():
    return np.load(os.path.join(ROOT, "tests", "golden", "collide_boxes.npz"))


def test_collide_matches_reference_fixture():
    f = _fixture()
    n = len(f["count"])
    assert n > 300 and int((f["count"] > 0).sum()) > 100
    for i in range(n):
        c, b, k, ra, rb = oracle.collide_boxes(f["pa"][i], f["qa"][i], f["sa"][i], f["pb"][i], f["qb"][i],
                                               f["sb"][i], f["margin"])
        assert c == f["count"][i], i
        if c:
            assert np.array_equal(k, f["keys"][i, :c]), i
            assert np.array_equal(ra.view(np.uint32), f["ra"][i, :c].view(np.uint32)), i
            assert np.array_equal(rb.view(np.uint32), f["rb"][i, :c].view(np.uint32)), i
            assert np.array_equal(b.view(np.uint32), f["basis"][i].view(np.uint32)), i


def test_collide_matches_reference_live():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent); the fixture test covers this")
    rng = np.random.default_rng(99)
    hits = 0
    for _ in range(2000):
        sa, sb = rng.uniform(0.3, 2.5, 3), rng.uniform(0.3, 2.5, 3)
        qa, qb = rng.normal(size=4), rng.normal(size=4)
        qa /= np.linalg.norm(qa)
        qb /= np.linalg.norm(qb)
        pa = rng.uniform(-3, 3, 3)
        d = rng.normal(size=3)
        pb = pa + d / np.linalg.norm(d) * rng.uniform(0.2, 1.1) * (np.linalg.norm(sa) + np.linalg.norm(sb)) / 2
        margin = rng.choice([0.0, 0.01, 0.05])
        a = oracle.collide_boxes(pa, qa, sa, pb, qb, sb, margin)
        r = oracle.ref_collide_boxes(pa, qa, sa, pb, qb, sb, margin)
        assert a[0] == r[0]
        if a[0]:
            hits += 1
            for x, y in zip(a[1:], r[1:]):
                assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
    assert hits > 200


def _spd_block(rng, scale=1.0):
    m = rng.normal(size=(6, 6))
    a = m @ m.T + np.eye(6) * rng.uniform(0.5, 5.0)
    return (a * scale).astype(np.float32)


def test_solve6_against_float64():
    rng = np.random.default_rng(5)
    for _ in range(200):
        a = _spd_block(rng, 100.0)
        x_true = rng.uniform(-0.05, 0.05, 6)
        b = (a.astype(np.float64) @ x_true).astype(np.float32)
        path, x = oracle.solve6(a, b)
        assert path == 1
        ref = np.linalg.solve(a.astype(np.float64), b.astype(np.float64))
        assert np.max(np.abs(x - ref)) <= 1e-4 * (1 + np.max(np.abs(ref)))


def test_solve6_fallbacks():
    # indefinite: unpivoted fails, pivoted LDL^T solves it
    a = np.diag([4.0, -3.0, 2.0, 5.0, 1.0, 7.0]).astype(np.float32)
    b = np.array([0.4, -0.3, 0.2, 0.5, 0.1, 0.7], np.float32)
    path, x = oracle.solve6(a, b)
    assert path == 2 and np.allclose(x, np.ones(6) * 0.1, atol=1e-6)
    # an |x| >= 10 answer is rejected by both and frozen to zero
    a = np.eye(6, dtype=np.float32)
    path, x = oracle.solve6(a, np.full(6, 20.0, np.float32))
    assert path == 0 and not x.any()


def test_solve6_matches_reference_live():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(6)
    mats = [_spd_block(rng, 1e3) for _ in range(200)]
    mats += [np.diag(rng.uniform(-2, 2, 6)).astype(np.float32) for _ in range(50)]
    for a in mats:
        b = rng.uniform(-1, 1, 6).astype(np.float32)
        p1, x1 = oracle.solve6(a, b)
        p2, x2 = oracle.ref_solve6(a, b)
        assert p1 == p2 and np.array_equal(x1.view(np.uint32), x2.view(np.uint32))


def test_morton_matches_reference_fixture():
    """Codes of the reference's own morton3D (golden fixture) and the (code, index) order."""
    f = np.load(os.path.join(ROOT, "tests", "golden", "morton.npz"))
    for k in range(int(f["sets"])):
        c = f[f"centers{k}"]
        codes, order = oracle.morton(c)
        assert np.array_equal(codes, f[f"codes{k}"])
        want = np.array(sorted(range(len(c)), key=lambda i: (int(codes[i]), i)), np.int32)
        assert np.array_equal(order, want)


def test_morton_matches_reference_live():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent); the fixture test covers this")
    rng = np.random.default_rng(7)
    for n in (2, 17, 300):
        c = rng.uniform(-50, 80, (n, 3)).astype(np.float32)
        c[: n // 3] = c[0]  # ties
        codes, _ = oracle.morton(c)
        assert np.array_equal(codes, oracle.ref_morton_codes(c))


def test_candidate_pairs_against_brute_force():
    b, j = scenes.pile(300, seed=11, pitch=2.2)
    o = oracle.OracleWorld(P())
    o.upload(b, j)
    for _ in range(3):
        c, h = o.swept_aabbs()
        lo, hi = c - h, c + h
        n = len(c)
        want = []
        for i in range(n):
            ov = ~np.any((lo[i + 1:] > hi[i]) | (hi[i + 1:] < lo[i]), axis=1)
            for k in np.nonzero(ov)[0]:
                want.append((i << 32) | (i + 1 + int(k)))
        assert np.array_equal(o.candidate_pairs(), np.array(sorted(want), np.uint64))
        o.step(20)


def test_rest_contact_closed_form():
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 4))
    o = run(*s.build(), P(), 240)
    st = o.download()
    z = st["position"][0, 2]           # the box is list index 0 (created last)
    assert abs(1.0 - z) <= 0.01 and abs(st["velocity"][0, 2]) <= 0.05


def test_stack_stability_closed_form():
    o = run(*scenes.digest_stack(), P(), 300)
    z = o.download()["position"][:, 2][:10][::-1]   # creation order, bottom first
    assert 1.0 - z[0] <= 0.01
    assert np.max(1.0 - np.diff(z)) <= 0.02


def test_hard_joint_closed_form():
    o = run(*scenes.hard_chain(), P(), 300)
    st = o.download()
    pos, rot = st["position"], st["rotation"]

    def rot_v(q, v):
        x, y, z, w = q
        u = np.array([x, y, z])
        return v + 2 * w * np.cross(u, v) + 2 * np.cross(u, np.cross(u, v))
    # list order: links (last created first) then the anchor at index 4
    bodies = [4, 3, 2, 1, 0]
    worst = 0.0
    for a, b in zip(bodies[:-1], bodies[1:]):
        ca = pos[a] + rot_v(rot[a], np.array([0, 0, -0.5]))
        cb = pos[b] + rot_v(rot[b], np.array([0, 0, 0.5]))
        worst = max(worst, float(np.linalg.norm(ca - cb)))
    assert worst <= 0.01 and pos[0, 2] < pos[4, 2]


def test_static_pairs_make_no_manifold():
    s = scenes.SceneBuilder()
    s.box((10, 10, 1), 0.0, 0.5, (0, 0, 0))
    s.box((1, 1, 1), 0.0, 0.5, (0, 0, 0.9))
    o = run(*s.build(), P(), 5)
    assert o.stats()["manifolds"] == 0


@pytest.mark.parametrize("scene,overflow", [("chains", True), ("pile", True), ("pile", False)])
def test_colouring_matches_sequential_greedy(scene, overflow):
    """The oracle's colouring equals greedy colouring restated from its definition (smallest
    colour unused by lower-ordered neighbours, contact-only bodies to the Jacobi bucket)."""
    if scene == "chains":
        bodies, joints = scenes.chains(n_chains=8, links=12)
        steps = 90
    else:
        bodies, joints = scenes.pile(n=300, seed=5)
        steps = 100
    params = P(iterations=10)
    params.contact_only_overflow = 1 if overflow else 0
    o = run(bodies, joints, params, steps)
    state = {**bodies, **o.download()}
    ref = scenes.greedy_colours(state, joints, o.manifolds(), contact_only_overflow=overflow)
    assert np.array_equal(o.colours(), ref)


# ---------------------------------------------------------------- springs, filters ----
# spring.cpp:31-68 and forces.hpp:104-113 have no recorded digest or fixture in the reference;
# the oracle's restatement is pinned by closed forms: a spring's static equilibrium, where its
# pull k (|d| - rest) equals the body's weight and the anchor point hangs straight above the
# centre of mass; a collision filter lets a box fall through the ground it would rest on.

def _hanging(r_b, steps):
    s = scenes.SceneBuilder()
    anchor = s.box((0.5, 0.5, 0.5), 0.0, 0.5, (0, 0, 10))
    bob = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 8.5))
    s.spring(anchor, bob, (0, 0, 0), r_b, 100.0, rest=1.0)
    o = run(*s.build(), P(iterations=10), steps)
    st = o.download()
    return st["position"][0].astype(np.float64), st["rotation"][0].astype(np.float64)


def test_spring_equilibrium_centred():
    """m g = k (L - rest): a 1 kg box on a k = 100 spring of rest length 1 hangs 1.1 below."""
    p, q = _hanging((0, 0, 0), 600)
    assert abs(p[2] - (10.0 - 1.1)) < 1e-3 and abs(p[0]) < 1e-6 and abs(p[1]) < 1e-6
    assert np.allclose(q, [0, 0, 0, 1], atol=1e-6)


def test_spring_equilibrium_off_centre():
    """Anchored off the centre, the box turns until the anchor point is straight above the
    centre of mass, 1.1 below the static anchor (wrong arm or torque signs never settle there)."""
    r_b = np.array([0.3, 0.0, 0.5])
    p, q = _hanging(r_b, 2400)
    attach = p + scenes._qrot64(q, r_b)
    assert abs(np.linalg.norm(attach - np.array([0, 0, 10.0])) - 1.1) < 2e-3
    assert abs(attach[0] - p[0]) < 2e-2 and abs(attach[1] - p[1]) < 1e-6
    assert abs((attach[2] - p[2]) - np.linalg.norm(r_b)) < 2e-3


@pytest.mark.parametrize("filtered", [False, True])
def test_ignore_collision_drops_the_pair(filtered):
    s = scenes.SceneBuilder()
    ground = s.box((40, 40, 1), 0.0, 0.5, (0, 0, 0))
    box = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 1.0))
    if filtered:
        s.ignore_collision(ground, box)
    o = run(*s.build(), P(iterations=10), 30)
    z = o.download()["position"][0][2]
    if filtered:
        assert len(o.manifolds()["body_a"]) == 0 and z < 0.0   # fell through the ground
    else:
        assert len(o.manifolds()["body_a"]) == 1 and abs(z - 1.0) < 5e-3


# ---------------------------------------------------------------- generic joints ----
# generic_joint.cpp has no recorded digest; the oracle's restatement is pinned by the
# reference's own GenericJoint scenarios (test/core_test.cpp:703-1000), measured here from the
# bodies' poses (not from the joint's formulas), and its atan2 against numpy's.
LOCK, FREE = (1, 0, 0, 0, 0, 0), (0, 0, 0, 0, 0, 0)


def _rig(body_pos, r_b, linear, angular, ang_vel=(0, 0, 0), lin_vel=(0, 0, 0), params=None):
    s = scenes.SceneBuilder()
    anchor = s.box((1, 1, 1), 0.0, 0.5, (0, 0, 0))
    body = s.box((1, 1, 1), 1.0, 0.5, body_pos, vel=lin_vel)
    s.b[body]["ang"] = ang_vel
    s.generic(anchor, body, (0, 0, 0), r_b, linear, angular)
    o = oracle.OracleWorld(params or P(iterations=10))
    o.upload(*s.build())
    return o


def test_atan2_poly_against_numpy():
    rng = np.random.default_rng(3)
    y = rng.standard_normal(4000).astype(np.float32) * np.float32(3)
    x = rng.standard_normal(4000).astype(np.float32) * np.float32(3)
    worst = 0.0
    for a, b in zip(y, x):
        got = oracle.atan2(a, b)
        ref = np.arctan2(np.float64(a), np.float64(b))
        worst = max(worst, abs(got - ref) / float(np.spacing(np.float32(abs(ref)))))
    assert worst <= 4.0, worst
    for a, b, want in [(0.0, 1.0, 0.0), (-0.0, 1.0, -0.0), (0.0, -1.0, math.pi), (-0.0, -1.0, -math.pi),
                       (1.0, 0.0, math.pi / 2), (-1.0, -0.0, -math.pi / 2), (1.0, 1.0, math.pi / 4)]:
        got = oracle.atan2(a, b)
        assert abs(got - want) <= 4e-7 * max(1.0, abs(want)) and math.copysign(1, got) == math.copysign(1, want)


def test_generic_hinge_holds_its_locked_axes():
    """core_test.cpp:742-790: a wheel spinning about its free axle; the two locked tilt axes
    and the hub hold (0.01 rad / 0.01 m), and a slow spin is kept."""
    o = _rig((0, 0, 0), (0, 0, 0), [LOCK] * 3, [LOCK, FREE, LOCK], ang_vel=(0, 30, 0))
    worst_tilt = worst_hub = 0.0
    for _ in range(300):
        o.step()
        st = o.download()
        q, p = st["rotation"][0], st["position"][0]
        worst_tilt = max(worst_tilt, abs(float(q[0])), abs(float(q[2])))
        worst_hub = max(worst_hub, float(np.linalg.norm(p)))
    assert worst_tilt <= 0.005 and worst_hub <= 0.01, (worst_tilt, worst_hub)
    o = _rig((0, 0, 0), (0, 0, 0), [LOCK] * 3, [LOCK, FREE, LOCK], ang_vel=(0, 1, 0))
    o.step(120)
    assert o.download()["angular_velocity"][0][1] >= 0.98


def test_generic_prismatic_spring_settles_at_mg_over_k():
    """core_test.cpp:880-908: a strut free along Z (limited to ±2) with a k = 500 spring hangs
    m g / k below the anchor."""
    strut = (2, -2.0, 2.0, 500.0, 0.0, 0.0)
    o = _rig((0, 0, 0), (0, 0, 0), [LOCK, LOCK, strut], [LOCK] * 3)
    o.step(600)
    travel = -float(o.download()["position"][0][2])   # anchor A minus anchor B along Z
    expected = 1.0 * 10.0 / 500.0
    assert abs(travel - expected) <= 0.2 * expected + 0.002, travel


@pytest.mark.parametrize("beta_ang", [100.0, 1000.0])
def test_generic_axis_limit_holds(beta_ang):
    """core_test.cpp:800-878: kicked repeatedly, an axle limited to ±0.4 rad reaches the bound
    and stays within 0.1 rad past it; the locked axes stay put."""
    o = _rig((0, 0, 0), (0, 0, 0), [LOCK] * 3, [LOCK, (2, -0.4, 0.4, 0, 0, 0), LOCK],
             params=P(iterations=10, beta_ang=beta_ang))
    worst = reached = 0.0
    for i in range(240):
        if i % 10 == 0 and i < 60:
            v = o.download()["angular_velocity"]
            v[0, 1] += np.float32(6.0)
            o.write_bodies({"angular_velocity": v})
        o.step()
        q = o.download()["rotation"][0].astype(np.float64)
        angle = abs(2.0 * math.atan2(q[1], q[3]))
        angle = min(angle, 2 * math.pi - angle)
        reached = max(reached, angle)
        worst = max(worst, angle - 0.4)
        assert abs(q[0]) <= 0.01 and abs(q[2]) <= 0.01
    assert reached >= 0.35 and worst <= 0.1, (reached, worst)


def test_generic_ball_socket_holds_off_centre():
    """core_test.cpp:915-947: a socket anchored at the body's top holds while it swings."""
    o = _rig((0, 0, -1.0), (0, 0, 1.0), [LOCK] * 3, [FREE] * 3, lin_vel=(1.5, 0, 0))
    worst = 0.0
    for _ in range(300):
        o.step()
        st = o.download()
        p, q = st["position"][0].astype(np.float64), st["rotation"][0].astype(np.float64)
        worst = max(worst, float(np.linalg.norm(p + scenes._qrot64(q, (0, 0, 1.0)))))
    assert worst <= 0.01 and o.download()["position"][0][2] < -0.5, worst


def _q64mul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def _qexp(v):
    th = float(np.linalg.norm(v))
    if th == 0.0:
        return np.array([0.0, 0.0, 0.0, 1.0])
    return np.concatenate([np.asarray(v) / th * math.sin(th / 2), [math.cos(th / 2)]])


def _rotvec(q):
    q = q if q[3] >= 0 else -q
    s = float(np.linalg.norm(q[:3]))
    return np.zeros(3) if s == 0.0 else q[:3] / s * 2.0 * math.atan2(s, q[3])


def _conj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]])


def _generic_values(pa, qa, ra, pb, qb, rb, rest):
    """A generic joint's six measurements, from their definition (generic_joint.cpp:69-96):
    A's anchor minus B's in A's frame, then the orientation of A relative to B (its rest
    orientation taken out) in A's frame, as a rotation vector."""
    anchor_a = pa + scenes._qrot64(qa, ra)
    anchor_b = pb + scenes._qrot64(qb, rb)
    lin = scenes._qrot64(_conj(qa), anchor_a - anchor_b)
    ang = _rotvec(_q64mul(rest, _conj(_q64mul(_conj(qa), qb))))
    return np.concatenate([lin, ang])


def _expected_stamp(st, jt, k, side, h=1e-6):
    """The penalty-method stamp of every active axis, J^T K J and J^T F, with J and the axis
    rates taken by finite differences of _generic_values (float64)."""
    a, b = jt["body_a"][k], jt["body_b"][k]
    pose = {i: (st["position"][i].astype(np.float64), st["rotation"][i].astype(np.float64)) for i in (a, b)}

    def values(pa, qa, pb, qb):
        return _generic_values(pa, qa, jt["r_a"][k].astype(np.float64), pb, qb,
                               jt["r_b"][k].astype(np.float64), jt["rest_rotation"][k].astype(np.float64))

    base = values(*pose[a], *pose[b])
    J = np.zeros((6, 6))
    for d in range(6):
        p, q = pose[side]
        e = np.zeros(3)
        e[d % 3] = h
        moved = (p + e, q) if d < 3 else (p, _q64mul(_qexp(e), q))
        args = (*moved, *pose[b]) if side == a else (*pose[a], *moved)
        J[:, d] = (values(*args) - base) / h
    step = {i: (pose[i][0] + h * st["velocity"][i], _q64mul(_qexp(h * st["angular_velocity"][i]), pose[i][1]))
            for i in (a, b)}
    rate = (values(*step[a], *step[b]) - base) / h
    lhs, rhs = np.zeros((6, 6)), np.zeros(6)
    alpha = 0.99
    for i in range(6):
        mode, lo, hi = jt["axis_mode"][k][i], jt["axis_lower"][k][i], jt["axis_upper"][k][i]
        pen, lam = float(jt["axis_penalty"][k][i]), float(jt["axis_lambda"][k][i])
        v = base[i]
        if mode == 1 or (mode == 2 and not lo <= v <= hi):
            C = v - jt["axis_initial"][k][i] * alpha if mode == 1 else (v - hi if v > hi else v - lo)
            F = pen * C + lam
            if mode == 2:
                F = max(F, 0.0) if C > 0 else min(F, 0.0)
            lhs += pen * np.outer(J[i], J[i])
            rhs += J[i] * F
        ks = float(jt["axis_spring"][k][i])
        if ks > 0:
            lhs += ks * np.outer(J[i], J[i])
            rhs += J[i] * (ks * (v - jt["axis_equilibrium"][k][i]) - jt["axis_damping"][k][i] * rate[i])
    return lhs, rhs


def _stamp_rig(a_dynamic):
    s = scenes.SceneBuilder()
    qa = _qexp(np.array([0.3, -0.5, 0.8]) * 0.9).astype(np.float32)
    qb = _qexp(np.array([-0.7, 0.2, 0.4]) * 1.3).astype(np.float32)
    pa = np.array([0.4, -0.2, 1.0], np.float32)
    ra, rb = np.array([0.3, 0.5, -0.5], np.float32), np.array([-0.5, 0.25, 0.4], np.float32)
    if a_dynamic:    # anchors coincide: the joint's Jacobian is the exact gradient there on both sides
        pb = (pa + scenes._qrot64(qa, ra) - scenes._qrot64(qb, rb)).astype(np.float32)
    else:
        pb = np.array([0.9, 0.35, -0.6], np.float32)
    A = s.box((1.0, 0.8, 1.2), 1.0 if a_dynamic else 0.0, 0.5, pa, rot=qa, vel=(0.2, -0.1, 0.3))
    B = s.box((0.6, 1.0, 0.7), 2.0, 0.5, pb, rot=qb, vel=(-0.4, 0.3, 0.5))
    s.b[A]["ang"] = (0.3, 0.1, -0.2) if a_dynamic else (0.0, 0.0, 0.0)
    s.b[B]["ang"] = (0.5, -0.8, 0.6)
    s.generic(A, B, ra, rb)
    bodies, joints = s.build()
    st = {k: bodies[k] for k in ("position", "rotation", "velocity", "angular_velocity")}
    vals = _generic_values(st["position"][1].astype(np.float64), st["rotation"][1].astype(np.float64),
                           ra.astype(np.float64), st["position"][0].astype(np.float64),
                           st["rotation"][0].astype(np.float64), rb.astype(np.float64),
                           joints["rest_rotation"][0].astype(np.float64))
    # linear X locked, Y limited past its upper bound, Z a damped spring; angular X locked,
    # Y limited (inside its range: no force), Z a damped spring
    joints["axis_mode"][0] = [1, 2, 0, 1, 2, 0]
    joints["axis_lower"][0] = [0, -1.0, 0, 0, -0.3, 0]
    joints["axis_upper"][0] = [0, vals[1] - 0.05, 0, 0, 0.3, 0]
    joints["axis_spring"][0] = [0, 0, 50.0, 0, 0, 7.0]
    joints["axis_equilibrium"][0] = [0, 0, 0.2, 0, 0, -0.4]
    joints["axis_damping"][0] = [0, 0, 3.0, 0, 0, 0.6]
    joints["axis_penalty"] = np.array([[300.0, 200.0, 0, 40.0, 60.0, 0]], np.float32)
    joints["axis_lambda"] = np.array([[-2.0, 1.5, 0, 0.5, 0, 0]], np.float32)
    joints["axis_initial"] = np.array([[0.05, 0, 0, 0.02, 0, 0]], np.float32)
    o = oracle.OracleWorld(P())
    o.upload(bodies, joints)
    return o, st, joints


@pytest.mark.parametrize("a_dynamic", [False, True])
def test_generic_stamp_is_the_constraint_gradient(a_dynamic):
    """Each active axis of a generic joint must add penalty * J J^T and J F to the primal
    block (generic_joint.cpp:157-277), where J is the gradient of the axis' measurement: here
    J and the spring rates come from finite differences of the measurements' definition,
    on a rotated frame with both anchors off-centre."""
    o, st, jt = _stamp_rig(a_dynamic)
    sides = (0, 1) if a_dynamic else (0,)
    for side in sides:
        lhs, rhs = o.force_stamp(0, side)
        el, er = _expected_stamp(st, jt, 0, side)
        assert np.abs(lhs - el).max() <= 1e-4 * np.abs(el).max(), (side, lhs - el)
        assert np.abs(rhs - er).max() <= 1e-4 * np.abs(er).max() + 1e-6, (side, rhs, er)


def _quat_matrix(q):
    x, y, z, w = (float(c) for c in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_generic_dual_and_warm_start_follow_the_penalty_method():
    """The augmented-Lagrangian updates of a generic joint's axes (generic_joint.cpp:191-232,
    279-319): the dual sets lambda to the clamped force and ramps the penalty by beta |C| up to
    what one body can resolve in a step, (0.25 / dt)^2 times its mass or its inertia about the
    axis (world inertia R I R^T); the warm start decays lambda by alpha gamma and the penalty
    by gamma (clamped to [1, penalty_max]) and records the current measurement."""
    o, st, jt = _stamp_rig(False)
    a, b = jt["body_a"][0], jt["body_b"][0]
    vals = _generic_values(st["position"][a].astype(np.float64), st["rotation"][a].astype(np.float64),
                           jt["r_a"][0].astype(np.float64), st["position"][b].astype(np.float64),
                           st["rotation"][b].astype(np.float64), jt["r_b"][0].astype(np.float64),
                           jt["rest_rotation"][0].astype(np.float64))
    p = P()
    dt, alpha, gamma = float(np.float32(p.dt)), 0.99, 0.999
    mass = 0.6 * 1.0 * 0.7 * 2.0
    moment = np.diag([(1.0 + 0.49) / 12 * mass, (0.36 + 0.49) / 12 * mass, (0.36 + 1.0) / 12 * mass])
    R_a, R_b = _quat_matrix(st["rotation"][a]), _quat_matrix(st["rotation"][b])
    capped = []
    o.force_update(0, "dual")
    _, j1 = o.download(joints=True)
    for i in range(6):
        mode, lo, hi = jt["axis_mode"][0][i], jt["axis_lower"][0][i], jt["axis_upper"][0][i]
        pen, lam, v = float(jt["axis_penalty"][0][i]), float(jt["axis_lambda"][0][i]), vals[i]
        axis = R_a[:, i % 3]
        if i < 3:
            cap, beta = mass * (0.25 / dt) ** 2, 1.0e4
        else:
            cap, beta = float(axis @ (R_b @ moment @ R_b.T) @ axis) * (0.25 / dt) ** 2, 100.0
        if mode == 1 or (mode == 2 and not lo <= v <= hi):
            C = v - jt["axis_initial"][0][i] * alpha if mode == 1 else (v - hi if v > hi else v - lo)
            F = pen * C + lam
            if mode == 2:
                F = max(F, 0.0) if C > 0 else min(F, 0.0)
            want_lam, want_pen = F, min(pen + abs(C) * beta, cap)
            capped.append(i) if cap < pen + abs(C) * beta else None
        else:
            want_lam, want_pen = (0.0 if mode == 2 else lam), pen
        assert abs(j1["axis_lambda"][0][i] - want_lam) <= 1e-4 * max(1.0, abs(want_lam)), (i, j1["axis_lambda"][0])
        assert abs(j1["axis_penalty"][0][i] - want_pen) <= 1e-4 * max(1.0, want_pen), (i, j1["axis_penalty"][0], want_pen)
    assert 0 in capped and 3 in capped   # a linear and an angular cap bind
    o.force_update(0, "init")
    _, j2 = o.download(joints=True)
    want_lam = j1["axis_lambda"][0].astype(np.float64) * alpha * gamma
    want_pen = np.clip(j1["axis_penalty"][0].astype(np.float64) * gamma, 1.0, float(p.penalty_max))
    assert np.allclose(j2["axis_lambda"][0], want_lam, rtol=1e-5, atol=1e-7)
    assert np.allclose(j2["axis_penalty"][0], want_pen, rtol=1e-5)
    assert np.allclose(j2["axis_initial"][0], vals, rtol=1e-4, atol=1e-5)


# ---------------------------------------------------------------- round shapes ----
def test_collide_shapes_matches_reference_fixture():
    """Spheres and cylinders against every shape: the oracle's narrow phase gives the bits the
    reference's own collide.cpp gave (tests/golden/collide_shapes.npz, 616 poses)."""
    f = np.load(os.path.join(ROOT, "tests", "golden", "collide_shapes.npz"))
    n = len(f["count"])
    for i in range(n):
        c, b, k, ra, rb = oracle.collide_shapes(f["ta"][i], f["pa"][i], f["qa"][i], f["sa"][i], f["tb"][i],
                                                f["pb"][i], f["qb"][i], f["sb"][i], f["margin"][i])
        assert c == f["count"][i], i
        if c:
            assert np.array_equal(b.view(np.uint32), f["basis"][i].view(np.uint32)), i
            assert np.array_equal(k, f["keys"][i, :c]), i
            assert np.array_equal(ra.view(np.uint32), f["ra"][i, :c].view(np.uint32)), i
            assert np.array_equal(rb.view(np.uint32), f["rb"][i, :c].view(np.uint32)), i
    assert (f["count"] > 0).sum() > 400 and set(zip(f["ta"], f["tb"])) >= {(0, 1), (1, 0), (0, 2), (2, 0), (1, 1),
                                                                            (1, 2), (2, 1), (2, 2)}


def test_collide_shapes_matches_reference_live():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(77)
    hits = 0
    for _ in range(3000):
        ta, tb = (int(x) for x in rng.integers(0, 3, 2))
        sa, sb = (rng.random(3) * 1.5 + 0.2).astype(np.float32), (rng.random(3) * 1.5 + 0.2).astype(np.float32)
        qa, qb = rng.standard_normal(4), rng.standard_normal(4)
        qa, qb = (qa / np.linalg.norm(qa)).astype(np.float32), (qb / np.linalg.norm(qb)).astype(np.float32)
        pb = rng.standard_normal(3).astype(np.float32)
        m = float(rng.choice([0.0, 0.01, 0.05]))
        a = oracle.collide_shapes(ta, np.zeros(3, np.float32), qa, sa, tb, pb, qb, sb, m)
        r = oracle.ref_collide_shapes(ta, np.zeros(3, np.float32), qa, sa, tb, pb, qb, sb, m)
        assert a[0] == r[0]
        hits += a[0] > 0
        for x, y in zip(a[1:], r[1:]):
            assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
    assert hits > 1000


def _rest_height(shape, size, rot=(0.0, 0.0, 0.0, 1.0), steps=400):
    s = scenes.SceneBuilder()
    s.box((200, 200, 1), 0.0, 0.5, (0, 0, -0.5))
    if shape == "sphere":
        s.sphere(size[0], 1000.0, 0.5, (0, 0, 4), rot=rot)
    else:
        s.cylinder(size[0], size[2], 1000.0, 0.5, (0, 0, 4), rot=rot)
    o = run(*s.build(), P(), steps)
    return float(o.download()["position"][0][2])


def _side():
    h = np.float32(np.float32(1.5707963) / np.float32(2.0))
    return (float(np.float32(math.sin(h))), 0.0, 0.0, float(np.float32(math.cos(h))))


# tools/baseline/core_test.txt was recorded at another revision of the reference than
# tools/baseline/digests.txt: its interleaved_300 digest is 0xbf1ce46c667164f3 where digests.txt
# (which the oracle reproduces) has 0xb03adeb37742850a.  Runs it recorded that the oracle
# matches to every printed digit are asserted to the print; the others to the reference's own
# acceptance bound and within 1e-3 of the recorded value.
@pytest.mark.parametrize("shape,size,rot,recorded,exact", [
    ("sphere", (0.5, 0, 0), (0.0, 0.0, 0.0, 1.0), 0.4920, True),
    ("cylinder", (0.5, 0.5, 1.0), (0.0, 0.0, 0.0, 1.0), 0.4921, False),
    ("cylinder", (0.5, 0.5, 1.0), "side", 0.4591, False),
])
def test_round_shape_rest_height_matches_recorded_run(shape, size, rot, recorded, exact):
    """core_test.cpp:1017-1043 (tools/baseline/core_test.txt:81-83): a sphere, a cylinder on
    its cap and one on its side come to rest on a slab one radius / half height up."""
    z = _rest_height(shape, size, _side() if rot == "side" else rot)
    assert abs(z - 0.5) <= (0.05 if shape == "sphere" else 0.06)
    assert abs(z - recorded) <= (5.01e-5 if exact else 1e-3), z


def test_sphere_stack_and_rolling_match_recorded_runs():
    """core_test.cpp:1045-1090 (tools/baseline/core_test.txt:84-87): a sphere resting on a
    sphere (centres 0.9898 apart, recorded) and a spun sphere rolling on the ground (spin
    20 -> 12.546 rad/s, travel -3.9950 m in 2 s, both matched to the print)."""
    s = scenes.SceneBuilder()
    s.box((200, 200, 1), 0.0, 0.5, (0, 0, -0.5))
    s.sphere(0.5, 1000.0, 0.5, (0, 0, 1.0))
    s.sphere(0.5, 1000.0, 0.5, (0, 0, 2.5))
    o = run(*s.build(), P(), 500)
    z = o.download()["position"][:, 2]
    gap = float(z[0] - z[1])
    assert abs(gap - 1.0) <= 0.12 and abs(gap - 0.9898) <= 1e-3, z

    s = scenes.SceneBuilder()
    s.box((200, 200, 1), 0.0, 0.9, (0, 0, -0.5))
    s.sphere(0.25, 1000.0, 1.0, (0, 0, 0.25))
    o = run(*s.build(), P(), 120)
    st = o.download()
    st["angular_velocity"][0] = (20.0, 0.0, 0.0)
    o.write_bodies({"angular_velocity": st["angular_velocity"]})
    y0 = float(st["position"][0][1])
    o.step(120)
    st = o.download()
    assert abs(float(st["angular_velocity"][0][0]) - 12.546) <= 5.01e-4, st["angular_velocity"][0]
    assert abs((float(st["position"][0][1]) - y0) - (-3.9950)) <= 5.01e-5, st["position"][0]


def test_cylinder_speculative_contacts_match_recorded_runs():
    """core_test.cpp:1352-1430 (tools/baseline/core_test.txt:120-124): a cylinder parked 5 mm
    above a box holds 8 speculative contacts (max normal C0 0.0060) and does not move; one
    thrown down at 15 m/s with no gravity never sinks past the slab and ends at z = 1.4665."""
    s = scenes.SceneBuilder()
    s.box((1, 1, 1), 0.0, 0.5, (0, 0, 0))
    s.cylinder(0.5, 1.0, 1.0, 0.5, (0, 0, 1.005))
    b, j = s.build()
    b["gravity"][0] = 0.0
    o = oracle.OracleWorld(P(speculative_margin=0.01))
    o.upload(b, j)
    o.step()
    m = o.manifolds()
    assert list(m["num_contacts"]) == [8]
    assert abs(float(m["c0"].reshape(-1, 8, 3)[0, :, 0].max()) - 0.0060) <= 5.01e-5
    assert abs(float(o.download()["position"][0][2]) - 1.005) < 1e-4

    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    s.cylinder(0.5, 1.0, 1.0, 0.5, (0, 0, 1.008), vel=(0, 0, -15.0))
    b, j = s.build()
    b["gravity"][0] = 0.0
    o = oracle.OracleWorld(P(speculative_margin=0.01))
    o.upload(b, j)
    lowest = 9.0
    for _ in range(60):
        o.step()
        lowest = min(lowest, float(o.download()["position"][0][2]))
    assert lowest - 0.5 >= 0.5 - 0.01
    assert abs(float(o.download()["position"][0][2]) - 1.4665) <= 5.01e-5


# ---------------------------------------------------------------- space queries ----
def test_intersect_ray_matches_reference_fixture():
    """query::intersectRay (query_shape.cpp:258-274) against tests/golden/ray_shapes.npz, made
    by the reference's own query_shape.cpp: hit flag, t, point, normal and the inside flag."""
    f = np.load(os.path.join(ROOT, "tests", "golden", "ray_shapes.npz"))
    for i in range(len(f["hit"])):
        hit, t, point, normal, inside = oracle.intersect_ray(f["type"][i], f["pos"][i], f["rot"][i], f["size"][i],
                                                             f["origin"][i], f["dir"][i], f["max_t"][i])
        assert hit == bool(f["hit"][i]), i
        if hit:
            assert np.float32(t).view(np.uint32) == f["t"][i].view(np.uint32), i
            assert np.array_equal(point.view(np.uint32), f["point"][i].view(np.uint32)), i
            assert np.array_equal(normal.view(np.uint32), f["normal"][i].view(np.uint32)), i
            assert inside == bool(f["inside"][i]), i
    assert f["hit"].sum() > 500 and f["inside"].sum() > 20


def test_intersect_ray_matches_reference_live():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(91)
    for _ in range(4000):
        t = int(rng.integers(0, 3))
        size = (rng.random(3) * 1.5 + 0.2).astype(np.float32)
        q = rng.standard_normal(4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        pos = rng.standard_normal(3).astype(np.float32)
        org = (rng.standard_normal(3) * 3).astype(np.float32)
        d = (pos - org + rng.standard_normal(3).astype(np.float32) * 0.5).astype(np.float32)
        max_t = float(rng.choice([0.0, 3.0]))
        a = oracle.intersect_ray(t, pos, q, size, org, d, max_t)
        r = oracle.ref_intersect_ray(t, pos, q, size, org, d, max_t)
        assert a[0] == r[0]
        if a[0]:
            assert np.float32(a[1]) == np.float32(r[1]) and np.array_equal(a[2], r[2]) and np.array_equal(a[3], r[3])
            assert a[4] == r[4]


@pytest.mark.parametrize("shape,size,origin,direction,max_t,t,normal,inside", [
    (1, (1, 0, 0), (-5, 0, 0), (1, 0, 0), 10.0, 4.0, (-1, 0, 0), False),    # query_test.cpp:79-85
    (1, (1, 0, 0), (-5, 0, 0), (1, 0, 0), 3.0, None, None, None),           # :87-90 maxT rejects
    (1, (1, 0, 0), (0, 0, 0), (1, 0, 0), 10.0, 1.0, (1, 0, 0), True),       # :92-98 from inside
    (0, (2, 2, 2), (-5, 0, 0), (1, 0, 0), 10.0, 4.0, (-1, 0, 0), False),    # :100-103 box entry
    (2, (1, 1, 4), (-5, 0, 0), (1, 0, 0), 10.0, 4.0, (-1, 0, 0), False),    # :105-109 cylinder side
    (2, (1, 1, 4), (0, 0, 5), (0, 0, -1), 10.0, 3.0, (0, 0, 1), False),     # :111-118 cylinder cap
    (1, (1, 0, 0), (0, 2, 0), (0, 0, 1), 10.0, None, None, None),           # :120-124 miss
])
def test_intersect_ray_known_answers(shape, size, origin, direction, max_t, t, normal, inside):
    """The reference's own ray checks (test/query_test.cpp:69-125) on unit shapes at the origin."""
    hit, tt, point, n, ins = oracle.intersect_ray(shape, (0, 0, 0), (0, 0, 0, 1), size, origin, direction, max_t)
    assert hit == (t is not None)
    if hit:
        assert abs(tt - t) <= 1e-5 and np.allclose(n, normal, atol=1e-5) and ins == inside
        assert np.allclose(point, np.asarray(origin) + np.asarray(direction) * t, atol=1e-5)


def _query_scene():
    s = scenes.SceneBuilder()
    s.box((20, 20, 1), 0.0, 0.5, (0, 0, -0.5))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 0.5))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 1.5))
    s.sphere(0.5, 1.0, 0.5, (3, 0, 0.5))
    s.cylinder(0.5, 2.0, 1.0, 0.5, (-3, 0, 1.0))
    b, j = s.build()
    b["layer"] = np.array([1, 1, 2, 1, 1], np.uint32)   # list order: cylinder, sphere, upper box, lower box, ground
    o = oracle.OracleWorld(P())
    o.upload(b, j)
    return o, b


def test_raycast_and_pick_closed_forms():
    """Rays straight down onto the stack, the sphere, the cylinder and the ground hit the top
    surfaces (z = 2, 1, 2, 0) with upward normals; a mask that excludes the upper box's layer
    hits the lower box at z = 1; a ray starting inside a body hits at t = 0 only when asked."""
    o, b = _query_scene()
    frm = np.array([[0, 0, 5], [3, 0, 5], [-3, 0, 5], [6, 6, 5], [0, 0, 5], [0, 0, 1.5]], np.float32)
    to = frm - np.array([0, 0, 10], np.float32)
    masks = np.array([0xFFFFFFFF] * 4 + [1, 0xFFFFFFFF], np.uint32)
    r = o.raycast(frm, to, masks)
    assert list(r["body"]) == [2, 1, 0, 4, 3, 3]
    assert np.allclose(r["point"][:, 2], [2, 1, 2, 0, 1, 1], atol=1e-5)
    assert np.allclose(r["t"][:5], [3, 4, 3, 5, 4], atol=1e-5) and np.allclose(r["normal"][:5], [0, 0, 1], atol=1e-5)
    assert r["inside"][5] == 0 and abs(r["t"][5] - 0.5) <= 1e-5   # from inside, not asked: the next body below
    ri = o.raycast(frm[5:], to[5:], None, hit_from_inside=True)
    assert ri["body"][0] == 2 and ri["inside"][0] == 1 and ri["t"][0] == 0.0
    assert np.allclose(ri["normal"][0], [0, 0, 1]) and np.allclose(ri["point"][0], frm[5])
    assert o.raycast(np.array([[0, 0, 5]], np.float32), np.array([[0, 0, 4]], np.float32))["body"][0] == -1
    down = np.tile(np.array([0, 0, -1], np.float32), (4, 1))   # pick's sphere test assumes a unit direction
    body, local = o.pick(frm[:4], down)
    assert list(body) == [2, 1, 0, 4]
    assert np.allclose(local[:, 2], [0.5, 0.5, 1.0, 0.5], atol=1e-5)
    body, _ = o.pick(frm[:1], down[:1], np.array([1], np.uint32))
    assert body[0] == 3


def test_collect_aabb_against_brute_force():
    """Solver::collectAABB: list order, mask, and each body's tight world bound computed here
    from its rotation matrix."""
    b, j = scenes.round_pile(60, seed=12)
    o = oracle.OracleWorld(P())
    o.upload(b, j)
    o.step(30)
    st = o.download()
    rng = np.random.default_rng(4)
    shape = b["shape"]
    for _ in range(50):
        c = rng.uniform(-6, 6, 3)
        h = rng.uniform(0.2, 3.0, 3)
        mask = int(rng.choice([1, 2, 3, 0xFFFFFFFF]))
        got = o.collect_aabb(c - h, c + h, mask)
        strict, loose = [], []
        for i in range(len(b["mass"])):
            if (int(b.get("layer", np.ones(len(b["mass"]), np.uint32))[i]) & mask) == 0:
                continue
            x, y, z, w = st["rotation"][i].astype(np.float64)
            R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                          [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                          [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
            s3 = b["size"][i].astype(np.float64)
            half = {0: s3 / 2, 1: np.array([s3[0]] * 3), 2: np.array([s3[0], s3[0], s3[2] / 2])}[int(shape[i])]
            wx = np.abs(R) @ half
            p = st["position"][i].astype(np.float64)
            if np.all(p - wx <= c + h - 1e-4) and np.all(c - h <= p + wx - 1e-4):
                strict.append(i)
            if np.all(p - wx <= c + h + 1e-4) and np.all(c - h <= p + wx + 1e-4):
                loose.append(i)
        assert list(got) == sorted(got) and set(strict) <= set(got) <= set(loose)


def test_contact_report_closed_form():
    """The server's contact report (avbd_physics_server3d.cpp:1593-1652) on a box resting on a
    slab: the box gets its four bottom corners with the normal pointing up out of the slab, the
    slab (asking for two) its first two with the normal pointing down; a body asking for none
    gets none; the other body is reported."""
    s = scenes.SceneBuilder()
    s.box((10, 10, 1), 0.0, 0.5, (0, 0, -0.5))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 0.5))
    s.box((1, 1, 1), 1.0, 0.5, (3, 0, 0.5))
    o = run(*s.build(), P(), 60)               # list order: right box, centre box, slab
    count, pos, nrm, other = o.contact_report([0, 8, 2])
    assert list(count) == [0, 4, 2]
    assert np.allclose(pos[1, :4, 2], 0.0, atol=2e-3) and np.allclose(np.abs(pos[1, :4, :2]), 0.5, atol=1e-3)
    assert np.allclose(nrm[1, :4], [0, 0, 1], atol=1e-5) and list(other[1, :4]) == [2] * 4
    assert np.allclose(nrm[2, :2], [0, 0, -1], atol=1e-5) and set(other[2, :2]) <= {0, 1}
    assert np.allclose(pos[2, :2, 2], 0.0, atol=2e-3)
    m = o.manifolds()
    first = [a if a != 2 else b for a, b in zip(m["body_a"], m["body_b"])][0]
    assert other[2, 0] == first               # the slab's first contacts come from the first live manifold


# ---------------------------------------------------------------- sleeping ----
# solver_solve.cpp:143-339, solver_broadphase.cpp:67-80 and manifold.cpp:168-172 have no
# recorded digest; the restatement is pinned by the reference's idle-sleep check
# (core_test.cpp:1645-1659, core_test.txt:139-140) and by the rules' observable effects.
def _sleep_world(p, *bodies_spec):
    s = scenes.SceneBuilder()
    s.box((20, 20, 1), 0.0, 0.5, (0, 0, -0.5))
    for pos, mode, vel in bodies_spec:
        k = s.box((1, 1, 1), 1.0, 0.5, pos, vel=vel)
        s.b[k]["sleep_mode"] = mode
    o = oracle.OracleWorld(p)
    o.upload(*s.build())
    return o


def test_idle_bodies_fall_asleep_and_hold_still():
    """core_test.cpp:1645-1659: a two-box stack with sleep_mode 1 and sleep_frames 5 is asleep
    after 120 steps.  Sleeping bodies keep zero velocity and an unchanged pose, and the count of
    still steps reaches exactly sleep_frames when a body falls asleep."""
    o = _sleep_world(P(sleep_frames=5), ((0, 0, 0.5), 1, (0, 0, 0)), ((0, 0, 1.5), 1, (0, 0, 0)))
    o.step(120)
    st = o.download()
    assert list(st["sleeping"][:2]) == [1, 1] and list(st["still_frames"][:2]) == [5, 5]
    assert np.all(st["velocity"][:2] == 0) and np.all(st["angular_velocity"][:2] == 0)
    o.step(30)
    st2 = o.download()
    assert np.array_equal(st2["position"], st["position"]) and np.array_equal(st2["rotation"], st["rotation"])
    assert o.stats()["manifolds"] == 2     # parked, not released


def test_sleep_needs_mode_and_frames():
    o = _sleep_world(P(sleep_frames=5), ((0, 0, 0.5), 0, (0, 0, 0)))      # mode never
    o.step(60)
    assert o.download()["sleeping"][0] == 0
    o = _sleep_world(P(sleep_frames=0), ((0, 0, 0.5), 1, (0, 0, 0)))      # frames 0: no idle detection
    o.step(60)
    assert o.download()["sleeping"][0] == 0


def test_falling_box_wakes_sleeper_on_bounds_contact():
    """A box dropped onto a sleeping one wakes it in the step their swept bounds first meet (the
    step that creates their manifold), not before; the two then rest on each other."""
    o = _sleep_world(P(sleep_frames=5), ((0, 0, 0.5), 1, (0, 0, 0)))
    o.step(60)
    assert o.download()["sleeping"][0] == 1
    st = o.download()
    s = scenes.SceneBuilder()
    s.box((20, 20, 1), 0.0, 0.5, (0, 0, -0.5))
    k = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 0.5))
    s.b[k]["sleep_mode"] = 1
    s.box((1, 1, 1), 1.0, 0.5, (0.2, 0, 3.0))
    b, j = s.build()
    b["position"][1] = st["position"][0]
    b["rotation"][1] = st["rotation"][0]
    b["sleeping"] = np.array([0, 1, 0], np.uint8)
    b["still_frames"] = np.array([0, 5, 0], np.int32)
    o = oracle.OracleWorld(P(sleep_frames=5))
    o.upload(b, j)
    woke = None
    for t in range(120):
        o.step()
        d = o.download()
        pair = any({a, c} == {0, 1} for a, c in zip(o.manifolds()["body_a"], o.manifolds()["body_b"]))
        if woke is None:
            assert (d["sleeping"][1] == 0) == pair, t
            if pair:
                woke = t
    assert woke is not None and woke > 5
    d = o.download()
    assert abs(float(d["position"][0][2]) - 1.5) < 0.05 and abs(float(d["position"][1][2]) - 0.5) < 0.02


def test_moving_neighbour_wakes_sleeper_and_start_asleep():
    """A sleeper touching an awake body that moves faster than the thresholds wakes; a body
    created asleep (mode 2) with no velocity and no forces or contacts is left where it is
    (gravity does not move it) until something touches it."""
    s = scenes.SceneBuilder()
    s.box((20, 20, 1), 0.0, 0.5, (0, 0, -0.5))
    a = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 0.5))
    s.b[a]["sleep_mode"] = 1
    s.box((1, 1, 1), 1.0, 0.5, (1.0, 0, 0.5))
    floater = s.box((1, 1, 1), 1.0, 0.5, (5, 5, 6.0))
    s.b[floater]["sleep_mode"] = 2
    b, j = s.build()                      # list order: floater, neighbour, sleeper, ground
    o = oracle.OracleWorld(P(sleep_frames=10))
    o.upload(b, j)
    o.step(40)                            # the sleeper falls asleep touching its awake neighbour
    d = o.download()
    assert d["sleeping"][2] == 1 and d["sleeping"][1] == 0
    assert any({x, y} == {1, 2} for x, y in zip(o.manifolds()["body_a"], o.manifolds()["body_b"]))
    assert np.array_equal(d["position"][0], b["position"][0])          # mode 2, untouched: not solved
    o.step(5)
    assert o.download()["sleeping"][2] == 1
    v = d["velocity"].copy()
    v[1] = (-2.0, 0, 0)                   # push the neighbour into the sleeper
    o.write_bodies({"velocity": v})
    o.step(2)
    assert o.download()["sleeping"][2] == 0
    v = o.download()["velocity"].copy()
    v[0] = (0, 0, -1.0)                   # the floater gets a velocity: it is solved from now on
    o.write_bodies({"velocity": v})
    o.step(10)
    assert o.download()["position"][0][2] < 5.9


def test_contains_point_matches_reference_live():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(33)
    inside = 0
    for _ in range(4000):
        t = int(rng.integers(0, 3))
        size = (rng.random(3) * 1.5 + 0.2).astype(np.float32)
        q = rng.standard_normal(4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        pos = rng.standard_normal(3).astype(np.float32)
        pt = (pos + rng.standard_normal(3) * 0.8).astype(np.float32)
        m = float(rng.choice([0.0, 0.05]))
        a = oracle.contains_point(t, pos, q, size, pt, m)
        assert a == oracle.ref_contains_point(t, pos, q, size, pt, m)
        inside += a
    assert inside > 800


def test_intersect_point_closed_forms():
    """The server's intersect_point: a point inside the lower of two stacked boxes finds it, a
    point on their shared face finds both (list order: upper first), a mask skips a layer, a
    point in the air finds nothing; shapes are tested exactly (a point in a sphere's bounding
    box corner is outside the sphere)."""
    o, b = _query_scene()     # list order: cylinder, sphere, upper box, lower box, ground
    pts = np.array([[0, 0, 0.5], [0, 0, 1.0], [0, 0, 1.0], [0, 0, 9.0], [3.45, 0.45, 0.95], [3, 0, 0.5]], np.float32)
    masks = np.array([0xFFFFFFFF, 0xFFFFFFFF, 1, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF], np.uint32)
    count, out = o.intersect_point(pts, masks, max_results=4)
    assert list(count) == [1, 2, 1, 0, 0, 1]
    assert out[0, 0] == 3 and list(out[1, :2]) == [2, 3] and out[2, 0] == 3 and out[5, 0] == 1
    count, out = o.intersect_point(pts[1:2], None, max_results=1)
    assert count[0] == 1 and out[0, 0] == 2


def test_intersect_shape_closed_forms():
    """The server's intersect_shape (avbd_direct_space_state3d.cpp:261-306): a sphere sunk 0.05
    into the upper box of the stack reaches the lower one too (list order: upper first); a mask
    skips the upper box's layer; a sphere 0.05 above the stack touches nothing until the margin
    is 0.1; a thin slab at the ground's top face finds every body resting on it and the ground;
    max_results truncates in list order; a query in the air finds nothing."""
    o, b = _query_scene()     # list order: cylinder, sphere, upper box, lower box, ground
    I = [0, 0, 0, 1]
    q = [(1, [0, 0, 1.05], [0.25] * 3, 0.0, 0xFFFFFFFF), (1, [0, 0, 1.05], [0.25] * 3, 0.0, 1),
         (1, [0, 0, 2.3], [0.25] * 3, 0.0, 0xFFFFFFFF), (1, [0, 0, 2.3], [0.25] * 3, 0.1, 0xFFFFFFFF),
         (0, [3, 0, 0.5], [0.5] * 3, 0.0, 0xFFFFFFFF), (2, [-3, 0, 2.1], [0.3, 0.3, 0.4], 0.0, 0xFFFFFFFF),
         (0, [0, 0, 0.05], [10, 10, 0.2], 0.0, 0xFFFFFFFF), (0, [6, 6, 5], [1, 1, 1], 0.0, 0xFFFFFFFF)]
    t, c, s, m, mk = (np.array([x[k] for x in q]) for k in range(5))
    count, out = o.intersect_shape(t, c, [I] * len(q), s, m, mk.astype(np.uint32), max_results=6)
    expect = [[2, 3], [3], [], [2], [1], [0], [0, 1, 3, 4], []]
    assert [list(out[i, :count[i]]) for i in range(len(q))] == expect
    count, out = o.intersect_shape(t[6:7], c[6:7], [I], s[6:7], m[6:7], None, max_results=2)
    assert count[0] == 2 and list(out[0]) == [0, 1]


def test_intersect_shape_is_the_narrow_phase_over_collected_bodies():
    """Random query shapes over a settled round pile: each result is exactly the bodies that
    collect_aabb returns for the query's bounds grown by the margin, filtered by the narrow
    phase (collide_shapes, query first) finding a contact, in list order."""
    b, j = scenes.round_pile(60, seed=3)
    o = oracle.OracleWorld(P())
    o.upload(b, j)
    o.step(20)
    st = o.download()
    rng = np.random.default_rng(8)
    n = 60
    t = rng.integers(0, 3, n).astype(np.int32)
    c = (st["position"][rng.integers(0, len(st["position"]), n)] + rng.normal(0, 0.6, (n, 3))).astype(np.float32)
    qr = rng.normal(size=(n, 4)).astype(np.float32)
    qr /= np.linalg.norm(qr, axis=1, keepdims=True)
    s = rng.uniform(0.2, 1.2, (n, 3)).astype(np.float32)
    m = np.where(rng.random(n) < 0.5, 0.0, 0.05).astype(np.float32)
    count, out = o.intersect_shape(t, c, qr, s, m, max_results=64)
    hits = 0
    for r in range(n):
        x, y, z, w = qr[r].astype(np.float64)
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        half = {0: s[r] * 0.5, 1: np.full(3, s[r][0]), 2: np.array([s[r][0], s[r][0], s[r][2] * 0.5])}[int(t[r])]
        wx = np.abs(R) @ half + m[r] + 1e-4
        want = [k for k in range(len(b["mass"]))
                if oracle.collide_shapes(t[r], c[r], qr[r], s[r], b["shape"][k], st["position"][k], st["rotation"][k],
                                         b["size"][k], m[r])[0] > 0]
        # a contact within the margin implies the grown bounds overlap, so the bounds pre-filter
        # drops no hit; every hit is also among the bodies whose bounds the query's reach
        near = set(o.collect_aabb(c[r] - wx, c[r] + wx))
        assert set(want) <= near
        assert list(out[r, :count[r]]) == want, r
        hits += count[r]
    assert hits > 20


def test_sweep_shape_matches_reference_live():
    """query::sweepShape restated: safe/unsafe fractions, depth, point and normal equal the
    reference's own (built from its sources) bit for bit over random shape pairs, motions
    (a tenth of them zero) and margins."""
    if oracle.ref_lib() is None:
        pytest.skip("reference library not built")
    rng = np.random.default_rng(1)
    hits = 0
    for i in range(300):
        ta, tb = rng.integers(0, 3, 2)
        qa, qb = rng.normal(size=4), rng.normal(size=4)
        qa, qb = qa / np.linalg.norm(qa), qb / np.linalg.norm(qb)
        pa, pb = rng.normal(0, 1.5, 3), rng.normal(0, 0.5, 3)
        mv = np.zeros(3) if i % 10 == 0 else (pb - pa) * rng.uniform(0.5, 2.5) + rng.normal(0, 0.5, 3)
        sa, sb = rng.uniform(0.3, 1.2, 3), rng.uniform(0.3, 1.2, 3)
        m = [0.0, 0.05][i % 2]
        o = oracle.sweep_shape(ta, pa, qa, sa, mv, tb, pb, qb, sb, m)
        r = oracle.ref_sweep_shape(ta, pa, qa, sa, mv, tb, pb, qb, sb, m)
        assert o[0] == r[0] and (not o[0] or np.array_equal(o[1].view(np.uint32), r[1].view(np.uint32))), i
        hits += r[0]
    assert hits > 200


def test_cast_motion_closed_forms():
    """The server's cast_motion: a unit box dropped 3 from 1.5 above the stack (gap 1) touches the
    upper box at safe = unsafe = 1/3 (the swept bounds' entry) with a vertical normal; a mask that skips the
    upper box's layer passes through it to the lower box at 2/3; a sideways motion that misses everything returns (1, 1) and no body."""
    o, b = _query_scene()     # list order: cylinder, sphere, upper box, lower box, ground
    I = [0, 0, 0, 1]
    n = 3
    r = o.cast_motion([0] * n, [[0, 0, 3.5]] * n, [I] * n, [[1, 1, 1]] * n, [[0, 0, -3], [0, 0, -3], [0, 0, 0.5]],
                      [0.0] * n, np.array([0xFFFFFFFF, 1, 0xFFFFFFFF], np.uint32))
    assert r["body"][0] == 2 and abs(r["safe"][0] - 1 / 3) < 2e-3 and 0 <= r["unsafe"][0] - r["safe"][0] < 1e-3
    assert np.allclose(np.abs(r["normal"][0]), [0, 0, 1], atol=1e-5) and abs(r["point"][0][2] - 2.0) < 5e-3
    assert r["body"][1] == 3 and abs(r["safe"][1] - 2 / 3) < 2e-3
    assert r["body"][2] == -1 and r["safe"][2] == 1 and r["unsafe"][2] == 1


# ---------------------------------------------------------------- contact grace ----
def test_contact_grace_parks_and_revives():
    """manifold.cpp:249-378 with solver_solve.cpp:66-86: with contact_grace_ticks = 3, a box that
    hops off the ground keeps its manifold parked (no contacts, no force) for up to 3 ticks; when
    it lands within them the contacts re-match by feature key and keep their penalty; a box that
    stays away longer has its manifold released; fresh contacts start at 5 % of beta_lin."""
    def world(vz, grace):
        s = scenes.SceneBuilder()
        s.box((20, 20, 1), 0.0, 0.5, (0, 0, -0.5))
        s.box((1, 1, 1), 1.0, 0.5, (0, 0, 0.5))
        o = oracle.OracleWorld(P(contact_grace_ticks=grace))
        o.upload(*s.build())
        o.step(30)
        st = o.download()
        v = st["velocity"].copy()
        v[0] = (0, 0, vz)
        o.write_bodies({"velocity": v})
        return o
    o = world(0.75, 3)                  # a short hop: about 9 steps in the air at 60 Hz
    counts = []
    for _ in range(30):
        o.step()
        m = o.manifolds()
        counts.append((len(m["body_a"]), int(m["num_contacts"][0]) if len(m["body_a"]) else -1))
    parked = [c for c in counts if c == (1, 0)]
    assert 1 <= len(parked) <= 3 and counts[-1][0] == 1 and counts[-1][1] == 4, counts
    o = world(3.0, 3)                   # a long flight: released after 3 parked ticks
    seen = []
    for _ in range(12):
        o.step()
        seen.append(len(o.manifolds()["body_a"]))
    assert seen[:3] == [1, 1, 1] and 0 in seen, seen
    # a fresh contact's penalty starts at 5 % of beta_lin (then the warm-start decay and one
    # dual ramp); without grace it starts at the floor
    s = scenes.SceneBuilder()
    s.box((20, 20, 1), 0.0, 0.5, (0, 0, -0.5))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 0.505))
    first = {}
    for grace in (3, 0):
        o = oracle.OracleWorld(P(contact_grace_ticks=grace, iterations=1))
        o.upload(*s.build())
        o.step()
        m = o.manifolds()
        first[grace] = float(m["penalty"][0][0])
    seed = float(np.float32(np.float32(0.05) * np.float32(1e4)) * np.float32(0.999))
    assert first[3] >= seed and first[0] < seed, first


# ---------------------------------------------------------------- collide_shape, rest_info, closest point ----

def test_closest_point_matches_reference_fixture():
    f = np.load(os.path.join(ROOT, "tests", "golden", "closest_point.npz"))
    assert len(f["type"]) > 1000
    for i in range(len(f["type"])):
        got = oracle.closest_point(f["type"][i], f["pos"][i], f["rot"][i], f["size"][i], f["point"][i])
        assert np.array_equal(got.view(np.uint32), f["closest"][i].view(np.uint32)), i


def test_closest_point_matches_reference_live():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent); the fixture test covers this")
    rng = np.random.default_rng(17)
    for _ in range(3000):
        t = int(rng.integers(0, 3))
        size = rng.uniform(0.2, 2.5, 3).astype(np.float32)
        q = rng.normal(size=4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        pos = rng.uniform(-3, 3, 3).astype(np.float32)
        p = (pos + rng.normal(size=3) * rng.choice([0.01, 0.3, 1.5, 4.0])).astype(np.float32)
        a = oracle.closest_point(t, pos, q, size, p)
        b = oracle.ref_closest_point(t, pos, q, size, p)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_closest_point_known_answers():
    I = np.array([0, 0, 0, 1], np.float32)
    box = (0, (1, 2, 3), I, (2, 4, 6))
    assert np.allclose(oracle.closest_point(*box, (5, 2, 3)), (2, 2, 3))          # clamped to the +x face
    assert np.allclose(oracle.closest_point(*box, (1.5, 2.5, 3.5)), (1.5, 2.5, 3.5))   # inside: itself
    sph = (1, (0, 0, 0), I, (2, 0, 0))
    assert np.allclose(oracle.closest_point(*sph, (0, 0, 5)), (0, 0, 2))           # onto the surface
    assert np.allclose(oracle.closest_point(*sph, (0, 0, 0)), (2, 0, 0))           # the centre: +x
    cyl = (2, (0, 0, 0), I, (1, 1, 4))
    assert np.allclose(oracle.closest_point(*cyl, (3, 0, 1)), (1, 0, 1))           # side
    assert np.allclose(oracle.closest_point(*cyl, (0.2, 0, 1.9)), (0.2, 0, 2))     # nearer cap
    assert np.allclose(oracle.closest_point(*cyl, (0, 0, 0.5)), (1, 0, 0.5))       # on the axis, side nearer


def _stacked(second_z=0.5):
    s = scenes.SceneBuilder()
    s.box((20, 20, 1), 0.0, 0.5, (0, 0, -0.5))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 0.5))
    s.box((1, 1, 1), 1.0, 0.5, (3, 0, second_z))
    o = oracle.OracleWorld(P())
    o.upload(*s.build())
    return o


def test_collide_shape_closed_form():
    """A unit box query sunk 1 cm into the top of a resting unit box (list index 1): four pairs,
    the query's bottom corners and the box's top corners, 1 cm apart; the cap on pairs holds and
    `any` still says it touched; masks and a far query give nothing."""
    o = _stacked()
    I = np.array([[0, 0, 0, 1]], np.float32)
    cnt, anyhit, pts = o.collide_shape([0], [[0, 0, 1.49]], I, [[1, 1, 1]], [0.0], max_pairs=8)
    assert cnt[0] == 4 and anyhit[0] == 1
    q, b = pts[0, :4, 0], pts[0, :4, 1]
    assert np.allclose(q[:, 2], 0.99, atol=1e-6) and np.allclose(b[:, 2], 1.0, atol=1e-6)
    assert np.allclose(np.abs(q[:, :2]), 0.5, atol=1e-6) and np.allclose(q[:, :2], b[:, :2], atol=1e-6)
    cnt, anyhit, _ = o.collide_shape([0], [[0, 0, 1.49]], I, [[1, 1, 1]], [0.0], max_pairs=2)
    assert cnt[0] == 2 and anyhit[0] == 1
    cnt, anyhit, _ = o.collide_shape([0], [[0, 0, 1.49]], I, [[1, 1, 1]], [0.0], masks=[2])
    assert cnt[0] == 0 and anyhit[0] == 0
    cnt, anyhit, _ = o.collide_shape([0], [[0, 0, 1.6]], I, [[1, 1, 1]], [0.05])
    assert cnt[0] == 0 and anyhit[0] == 0
    cnt, anyhit, _ = o.collide_shape([0], [[0, 0, 1.6]], I, [[1, 1, 1]], [0.2])
    assert cnt[0] == 4 and anyhit[0] == 1   # inside the margin: contacts 10 cm apart
    assert np.allclose(_[0, :4, 0, 2] - _[0, :4, 1, 2], 0.1, atol=1e-5)


def test_rest_info_closed_form():
    """A slab over two boxes, 1 cm into the one at x = 0 (list index 1) and 3 cm into the raised
    one at x = 3 (list index 0), rests on the deeper: depth 0.03 (query::penetrationDepth along
    the contact normal), the normal points up out of it, the point is on the slab's bottom face;
    a unit box sunk 1 cm into the first alone gets depth 0.01; nothing touched gives body -1."""
    o = _stacked(second_z=0.52)
    I = np.array([[0, 0, 0, 1]], np.float32)
    body, point, normal, depth = o.rest_info([0], [[1.5, 0, 1.49]], I, [[4.2, 1, 1]], [0.0])
    assert body[0] == 0 and np.allclose(normal[0], (0, 0, 1), atol=1e-6), (body, normal)
    assert abs(depth[0] - 0.03) < 1e-5 and abs(point[0][2] - 0.99) < 1e-6 and point[0][0] > 2.4, (depth, point)
    o = _stacked()
    body, point, normal, depth = o.rest_info([0], [[0, 0, 1.49]], I, [[1, 1, 1]], [0.0])
    assert body[0] == 1 and abs(depth[0] - 0.01) < 1e-5 and np.allclose(normal[0], (0, 0, 1), atol=1e-6)
    assert abs(point[0][2] - 0.99) < 1e-6
    body, _, _, _ = o.rest_info([0], [[0, 0, 3.0]], I, [[1, 1, 1]], [0.0])
    assert body[0] == -1


def test_collide_shape_and_rest_info_are_the_narrow_phase_over_collected_bodies():
    """Brute force over every body (narrow phase query first, the list order, the caps): the
    bounds pre-filter drops no contact, and the deepest is the maximum over all contacts."""
    b, j = scenes.round_pile(80, seed=9, pitch=2.2)
    o = oracle.OracleWorld(P(iterations=10))
    o.upload(b, j)
    o.step(60)
    st = o.download()
    rng = np.random.default_rng(3)
    pos = st["position"]
    n = 120
    types = rng.integers(0, 3, n).astype(np.int32)
    centers = (pos[rng.integers(0, len(pos), n)] + rng.normal(0, 0.6, (n, 3))).astype(np.float32)
    rots = rng.normal(size=(n, 4)).astype(np.float32)
    rots /= np.linalg.norm(rots, axis=1, keepdims=True)
    sizes = rng.uniform(0.3, 1.5, (n, 3)).astype(np.float32)
    margins = rng.choice([0.0, 0.02], n).astype(np.float32)
    cnt, anyhit, pts = o.collide_shape(types, centers, rots, sizes, margins, max_pairs=64)
    body, point, normal, depth = o.rest_info(types, centers, rots, sizes, margins)
    touched = 0
    for r in range(n):
        want_pairs, best = [], None
        for i in range(len(pos)):
            c = oracle.collide_shapes(types[r], centers[r], rots[r], sizes[r], int(b.get("shape", np.zeros(len(pos)))[i]),
                                      st["position"][i], st["rotation"][i], b["size"][i], margins[r])
            for k in range(c[0]):
                want_pairs.append(i)
        assert cnt[r] == len(want_pairs) and bool(anyhit[r]) == bool(want_pairs), r
        touched += bool(want_pairs)
        assert (body[r] >= 0) == bool(want_pairs)
    assert touched > 30

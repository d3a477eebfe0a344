"""CPU pins for the oracle's functions that no recorded digest or golden fixture covers:
collision layers and masks, axis locks, the speculative contact paths, friction, the update set,
the colour-overflow bucket, the position projection and its cap, adaptive beta, Newton rounds and
stiffness decay, hard (infinite-stiffness) joints with the penalty cap, joint fracture, and the
manifold caches (the primal's force cache, drift-gated arm reuse).

Each pin is either one of the reference's own scenario checks, restated with its thresholds
(test/core_test.cpp, cited per function), or a closed form the mechanics fix (a hanging body's
joint force equals its weight; the stabilised joint error shrinks by alpha per step; a penalty
that the ramp overshoots equals its cap exactly; a step with Newton ratio r equals a step with
ceil(r N) plain iterations; ...).  The pins are checked against planted mistakes: the oracle
library carries a switch (oracle_plant, test-only) that swaps one expression for a plausible wrong
one, and `test_planted_mistake_is_caught` shows that every planted mistake fails some pin.
"""
import math

import numpy as np
import pytest

from helpers import bits_equal, oracle, pkg, scenes

P = pkg.default_params
F = np.float32


def _pos(created, n):
    """List index of the body created `created`-th of n (the list is in reverse creation order)."""
    return n - 1 - created


def _run(bodies, joints, params, steps):
    o = oracle.OracleWorld(params)
    o.upload(bodies, joints)
    o.step(steps)
    return o


# ------------------------------------------------------------------- layers, locks ----

def pin_layer_mask():
    """core_test.cpp:1506-1527: a box on layer 4 with an open mask rests on the ground (layer 1,
    mask 1: the pair collides through one direction of the test); a box on layer 2 with mask 2
    falls through."""
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    compatible = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 4))
    filtered = s.box((1, 1, 1), 1.0, 0.5, (5, 0, 4))
    b, j = s.build()
    n = len(s.b)
    b["layer"] = np.ones(n, np.uint32)
    b["mask"] = np.ones(n, np.uint32)
    b["layer"][_pos(compatible, n)], b["mask"][_pos(compatible, n)] = 4, 0xFFFFFFFF
    b["layer"][_pos(filtered, n)], b["mask"][_pos(filtered, n)] = 2, 2
    z = _run(b, j, P(), 240).download()["position"][:, 2]
    assert abs(z[_pos(compatible, n)] - 1.0) <= 0.02, z
    assert z[_pos(filtered, n)] < 0.5, z


def pin_axis_lock():
    """core_test.cpp:1533-1558: a body locked on z hovers with no velocity, an unlocked control
    falls, and a body locked about x loses its initial spin about x."""
    s = scenes.SceneBuilder()
    locked = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 5.0))
    control = s.box((1, 1, 1), 1.0, 0.5, (3, 0, 5.0))
    spin = s.box((1, 1, 1), 1.0, 0.5, (-3, 0, 5.0))
    s.b[spin]["ang"] = (5.0, 0.0, 0.0)
    b, j = s.build()
    n = len(s.b)
    b["lock_linear"] = np.zeros(n, np.uint8)
    b["lock_angular"] = np.zeros(n, np.uint8)
    b["lock_linear"][_pos(locked, n)] = 0x4
    b["lock_angular"][_pos(spin, n)] = 0x1
    st = _run(b, j, P(), 240).download()
    assert abs(st["position"][_pos(locked, n), 2] - 5.0) <= 1e-4
    assert abs(st["velocity"][_pos(locked, n), 2]) <= 1e-4
    assert st["position"][_pos(control, n), 2] < 5.0 - 3.0
    assert abs(st["rotation"][_pos(spin, n), 0]) <= 1e-4 and abs(st["angular_velocity"][_pos(spin, n), 0]) <= 1e-4


def pin_locked_axis_reports_no_velocity():
    """A constraint may move a body along a locked axis inside the iterations (the locks act on
    velocities: solver_solve.cpp:109-176), but the recovered velocity on that axis is exactly zero
    every step (the Godot BodyAxis promise): a spring pulls a box locked on x towards an anchor."""
    s = scenes.SceneBuilder(gravity=0.0)
    anchor = s.box((0.5, 0.5, 0.5), 0.0, 0.5, (3, 0, 5))
    box = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 5))
    s.spring(anchor, box, (0, 0, 0), (0, 0, 0), 50.0, rest=1.0)
    b, j = s.build()
    n = len(s.b)
    b["lock_linear"] = np.zeros(n, np.uint8)
    b["lock_linear"][_pos(box, n)] = 0x1
    o = oracle.OracleWorld(P())
    o.upload(b, j)
    for _ in range(30):
        o.step()
        st = o.download()
        assert st["velocity"][_pos(box, n), 0] == 0.0
    assert st["position"][_pos(box, n), 0] > 0.3      # the pull did move it


# ------------------------------------------------------------------- speculative ----

def _gap_box(vz, margin, z0):
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    box = s.box((1, 1, 1), 1.0, 0.5, (0, 0, z0), vel=(0, 0, vz))
    b, j = s.build()
    b["gravity"][_pos(box, 2)] = 0.0
    o = oracle.OracleWorld(P(speculative_margin=margin))
    o.upload(b, j)
    return o, _pos(box, 2)


def pin_speculative_contacts():
    """core_test.cpp:1166-1209: a 3 cm gap gives no contact at margin 0 and contacts at margin
    0.05; at rest inside the margin the manifold exists, its normal C0 is the gap, and it does
    not push; closing the gap at 10 m/s, the body is stopped at the surface."""
    q = np.array([0, 0, 0, 1], F)
    h = np.ones(3, F)
    c = np.array([0, 0, 1.03], F)
    assert oracle.collide_boxes(np.zeros(3, F), q, h, c, q, h, 0.0)[0] == 0
    assert oracle.collide_boxes(np.zeros(3, F), q, h, c, q, h, 0.05)[0] > 0
    o, i = _gap_box(0.0, 0.05, 1.03)
    o.step()
    m = o.manifolds()
    assert m["num_contacts"].sum() > 0 and m["c0"].reshape(-1, 8, 3)[:, :, 0].max() > 0.02
    assert abs(o.download()["position"][i, 2] - 1.03) < 1e-4
    o, i = _gap_box(-10.0, 0.05, 1.03)
    o.step()
    assert o.download()["position"][i, 2] > 0.95


def pin_speculative_gap_clamp():
    """core_test.cpp:1216-1235: from a 5 mm gap at 10 m/s the body stops at the surface, with the
    face contact symmetric (no spin), and holds there the next step."""
    o, i = _gap_box(-10.0, 0.01, 1.005)
    o.step()
    st = o.download()
    assert abs(st["position"][i, 2] - 1.0) < 1e-4 and np.linalg.norm(st["angular_velocity"][i]) < 1e-6
    o.step()
    assert o.download()["position"][i, 2] > 0.95


def _horizon(substeps, projection, margin=0.01):
    s = scenes.SceneBuilder(gravity=0.0)
    s.box((100, 100, 0.02), 0.0, 0.5, (0, 0, -0.01))
    s.sphere(0.05, 1.0, 0.5, (0, 0, 0.3), vel=(0, 0, -30))
    b, j = s.build()
    p = P(dt=float(F(1.0) / F(60.0 * substeps)), position_projection=int(projection), speculative_margin=margin)
    o = oracle.OracleWorld(p)
    o.upload(b, j)
    return o


def pin_speculative_horizon():
    """core_test.cpp:1237-1268: a 30 m/s sphere is intercepted within the tick at 1 and 4
    substeps, with and without projection, because the margin grows with the relative motion;
    with margin 0 nothing is intercepted."""
    for substeps in (1, 4):
        for projection in (False, True):
            o = _horizon(substeps, projection)
            lowest = 0.3
            for _ in range(5):
                o.step_interleaved(substeps, 2)
                lowest = min(lowest, float(o.download()["position"][0, 2]))
            assert lowest >= 0.0, (substeps, projection, lowest)
    s = scenes.SceneBuilder(gravity=0.0)
    s.box((100, 100, 0.02), 0.0, 0.5, (0, 0, -0.01))
    s.sphere(0.05, 1.0, 0.5, (0, 0, 0.3), vel=(0, 0, -30))
    b, j = s.build()
    o = oracle.OracleWorld(P(dt=float(F(1.0) / F(240.0)), position_projection=0, speculative_margin=0.0))
    o.upload(b, j)
    o.step_interleaved(4, 2)
    assert o.download()["position"][0, 2] < 0.0 and o.stats()["manifolds"] == 0


# ------------------------------------------------------------------- friction ----

def _ramp(mu, contact_only_overflow=1, seconds=4.0):
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, mu, (0, 0, 0))
    angle = float(F(20.0) * F(math.pi) / F(180.0))
    q = np.array([0, math.sin(angle / 2), 0, math.cos(angle / 2)], F)
    s.box((40, 24, 1), 0.0, mu, (0, 0, 6), rot=q)
    tangent, normal = scenes._qrot64(q, (1, 0, 0)), scenes._qrot64(q, (0, 0, 1))
    start = (np.array([0, 0, 6.0]) + tangent * -5.0 + normal * 1.05).astype(F)
    box = s.box((1, 1, 1), 1.0, mu, start)
    b, j = s.build()
    i = _pos(box, len(s.b))
    o = oracle.OracleWorld(P(contact_only_overflow=contact_only_overflow))
    o.upload(b, j)
    total, late = int(seconds * 60), 30
    o.step(total - late)
    before = o.download()["position"][i].copy()
    o.step(late)
    st = o.download()
    p = st["position"][i]
    return (math.hypot(*(p - start)[:2]), math.hypot(*(p - before)[:2]), float(np.linalg.norm(st["velocity"][i])))


@pytest.mark.parametrize("contact_only_overflow", [1, 0])
def test_friction_on_a_ramp(contact_only_overflow):
    """core_test.cpp:213-265: on a 20 degree ramp mu = 0.5 holds the box (creep <= 2 cm in the
    last 0.5 s, speed <= 0.05) and mu = 0.05 lets it slide (>= 1 m in the last 0.5 s, >= 2 m in
    all); also through the colour sweep (contact-only bodies not sent to the Jacobi bucket)."""
    stuck = _ramp(0.5, contact_only_overflow)
    slid = _ramp(0.05, contact_only_overflow)
    assert stuck[1] <= 0.02 and stuck[2] <= 0.05, stuck
    assert slid[1] >= 1.0 and slid[0] >= 2.0 and slid[1] > stuck[1] * 50.0, slid


def test_rest_and_stack_through_the_colour_sweep():
    """core_test.cpp:155-211 with contactOnlyOverflow = false: the resting box and the 10-box
    stack hold through colour Gauss-Seidel too."""
    p = P(contact_only_overflow=0)
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 4))
    st = _run(*s.build(), p, 240).download()
    assert abs(1.0 - st["position"][0, 2]) <= 0.01 and abs(st["velocity"][0, 2]) <= 0.05
    z = _run(*scenes.digest_stack(), p, 300).download()["position"][:, 2][:10][::-1]
    assert 1.0 - z[0] <= 0.01 and np.max(1.0 - np.diff(z)) <= 0.02


# ------------------------------------------------------------------- update set, colours ----

def test_update_set_covers_every_awake_constraint():
    """core_test.cpp:1848-1875: two boxes meet mid-run; every awake movable side of every
    manifold and force is in the update set (coloured, or in the Jacobi bucket)."""
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 2))
    s.box((1, 1, 1), 1.0, 0.5, (0.25, 0, 6))
    b, j = s.build()
    o = _run(b, j, P(), 300)
    m, col = o.manifolds(), o.colours()
    assert len(m["body_a"]) >= 2
    for side in np.concatenate([m["body_a"], m["body_b"]]):
        if b["mass"][side] > 0:
            assert col[side] != 255, (side, col)
    st = o.download()
    assert all(np.isfinite(st[k]).all() for k in ("position", "rotation", "velocity"))


def test_colour_overflow_clique_stays_bounded():
    """core_test.cpp:1884-1907: ten bodies jointed pairwise (Joint k 1000 / 250) need ten colours;
    with gravity off the clique stays finite, within 2 m of its start in x and |z| < 100 over
    600 steps.  (As in the reference, bodies past the eighth colour that touch a non-contact force
    join neither a colour run nor the Jacobi list.)"""
    s = scenes.SceneBuilder(gravity=0.0)
    ids = [s.box((1, 1, 1), 1.0, 0.5, (0.1 * i, 0, 5.0 + 0.05 * i)) for i in range(10)]
    for i in range(10):
        for k in range(i + 1, 10):
            s.joint(ids[i], ids[k], (0.5, 0, 0), (-0.5, 0, 0), 1000.0, 250.0)
    o = _run(*s.build(), P(), 600)
    p = o.download()["position"]
    assert np.isfinite(p).all()
    assert max(abs(p[_pos(ids[i], 10), 0] - 0.1 * i) for i in range(10)) < 2.0 and np.abs(p[:, 2]).max() < 100.0
    assert o.colours().max() == 9          # two bodies left without a colour


def pin_contact_only_overflow_off():
    """With contactOnlyOverflow = false the colouring is plain greedy colouring (smallest colour
    unused by lower-ordered neighbours), restated from its definition (solver_solve.cpp:365-424)."""
    bodies, joints = scenes.pile(n=200, seed=5)
    o = _run(bodies, joints, P(iterations=10, contact_only_overflow=0), 60)
    ref = scenes.greedy_colours({**bodies, **o.download()}, joints, o.manifolds(), contact_only_overflow=False)
    assert np.array_equal(o.colours(), ref)


# ------------------------------------------------------------------- projection ----

def test_overlap_projection_preserves_velocity():
    """core_test.cpp:1270-1300: one primal iteration with alpha 1, then one projection pass: the
    box ends on the surface (z = 1) with the primal's velocity and no spin, either creation order."""
    for reverse in (False, True):
        out = {}
        for projection in (False, True):
            s = scenes.SceneBuilder(gravity=0.0)
            box = None
            for i in range(2):
                if (i == 0) == reverse:
                    box = s.box((1, 1, 1), 1.0, 0.0, (0, 0, 0.9), vel=(0, 0, -1))
                else:
                    s.box((100, 100, 1), 0.0, 0.0, (0, 0, 0))
            p = P(alpha=1.0, iterations=1, position_projection=int(projection), projection_iterations=1,
                  projection_slop=0.0)
            st = _run(*s.build(), p, 1).download()
            out[projection] = st, _pos(box, 2)
        (a, i), (b, k) = out[False], out[True]
        assert np.linalg.norm(b["angular_velocity"][k]) < 1e-6
        assert np.linalg.norm(b["velocity"][k] - a["velocity"][i]) < 1e-6
        assert abs(b["position"][k, 2] - 1.0) < 1e-5


def pin_projection_shared_plane():
    """core_test.cpp:1301-1345: a sphere 10 cm into 1, 2 or 4 coincident slabs is lifted to the
    surface in one pass (the per-contact corrections are averaged, not summed), or by exactly the
    cap when projectionMaxCorrection is set; the slabs stay put; opposed slabs keep it centred."""
    for slabs in (1, 2, 4):
        for density in (1.0, 10.0):
            for cap in (0.0, 0.01):
                s = scenes.SceneBuilder(gravity=0.0)
                for _ in range(slabs):
                    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
                sphere = s.sphere(0.5, density, 0.5, (0, 0, 0.9))
                p = P(alpha=1.0, iterations=1, projection_iterations=1, projection_beta=1.0, projection_slop=0.0,
                      projection_max_correction=cap)
                st = _run(*s.build(), p, 1).download()
                k = _pos(sphere, len(s.b))
                want = 0.9 + cap if cap > 0 else 1.0
                assert abs(st["position"][k, 2] - want) < 1e-5, (slabs, density, cap, st["position"][k])
                assert float(np.sum(st["velocity"][k].astype(np.float64) ** 2)) < 1e-10
                others = [i for i in range(len(s.b)) if i != k]
                assert not st["position"][others].any()
    s = scenes.SceneBuilder(gravity=0.0)
    s.box((1, 100, 100), 0.0, 0.5, (-0.9, 0, 0))
    s.box((1, 100, 100), 0.0, 0.5, (0.9, 0, 0))
    s.sphere(0.5, 1.0, 0.5, (0, 0, 0))
    p = P(alpha=1.0, iterations=1, projection_iterations=1, projection_beta=1.0, projection_slop=0.0)
    assert np.linalg.norm(_run(*s.build(), p, 1).download()["position"][0]) < 1e-5


# ------------------------------------------------------------------- adaptive beta ----

def _flat_box(adaptive, beta, iterations):
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 1.0005))
    b, j = s.build()
    b["gravity"][0] = 0.0
    o = oracle.OracleWorld(P(iterations=iterations, adaptive_beta=adaptive, beta_lin=beta, position_projection=0))
    o.upload(b, j)
    return o


def _teleport(o, z):
    zero = np.zeros((2, 3), F)
    o.write_bodies({"position": np.array([[0, 0, z], [0, 0, 0]], F), "rotation": np.array([[0, 0, 0, 1]] * 2, F),
                    "velocity": zero, "angular_velocity": zero, "prev_velocity": zero})


def _beta_rule(scale, err, prev, floor=1e-3, cap=64.0):
    """The per-contact multiplier rule as solver.h:74-79 documents it: a solved contact (error at
    or below the floor) decays towards 1; one that was solved and is not doubles; a penetration
    that did not shrink to half doubles; one that shrank below a tenth halves; else it is kept."""
    if err <= floor:
        return max(scale * 0.5, 1.0), "solved"
    if prev <= floor:
        return min(scale * 2.0, cap), "reopened"
    if err > 0.5 * prev:
        return min(scale * 2.0, cap), "not shrinking"
    if err < 0.1 * prev:
        return max(scale * 0.5, 1.0), "converging"
    return scale, "kept"


def pin_adaptive_beta_rule():
    """A flat box is put back on the ground at a chosen depth before every step (gravity off,
    projection off, weak ramp so that penetration persists); each contact's beta multiplier after
    the step follows the rule from its previous multiplier, its error at the step's start (the
    chosen depth) and its error at the previous step's end (its rebased C0, read back).  Every
    branch of the rule is taken."""
    rng = np.random.default_rng(1)
    o = _flat_box(1, 1.0, 1)
    prev, seen = None, {}
    for _ in range(80):
        z = F(1.0 - float(rng.choice([0.0, 0.0005, 0.0015, 0.003, 0.008, 0.012, 0.016, 0.019])))
        _teleport(o, z)
        o.step()
        m = o.manifolds()
        nc = int(m["num_contacts"][0])
        keys, scale, c0 = m["feature"][0][:nc], m["beta_scale"][0][:nc], m["c0"].reshape(-1, 8, 3)[0, :nc, 0]
        err = max(0.0, 1.0 - float(z))
        assert nc == 4
        for i in range(nc):
            if prev is None or keys[i] not in prev:
                want, branch = 1.0, "fresh"
            else:
                want, branch = _beta_rule(*prev[keys[i]][:1], err, prev[keys[i]][1])
            seen[branch] = seen.get(branch, 0) + 1
            assert scale[i] == want, (branch, scale[i], want)
        prev = {keys[i]: (float(scale[i]), max(0.0, 0.001 - float(c0[i]))) for i in range(nc)}
    assert set(seen) == {"fresh", "solved", "reopened", "not shrinking", "converging", "kept"}, seen


def pin_adaptive_scale_multiplies_beta():
    """When every contact of a step carries multiplier s, the step equals (bit for bit) the same
    step with adaptive beta off and beta_lin multiplied by s: the multiplier scales the whole
    penalty ramp and nothing else (manifold.cpp:583-671)."""
    rng = np.random.default_rng(1)
    o = _flat_box(1, 1.0e4, 10)
    for _ in range(40):
        z = F(1.0 - float(rng.choice([0.0005, 0.008, 0.016])))
        before, man = o.download(), o.manifolds()
        _teleport(o, z)
        o.step()
        m = o.manifolds()
        nc = int(m["num_contacts"][0])
        scale = m["beta_scale"][0][:nc]
        if nc and np.all(scale == scale[0]) and scale[0] > 1:
            other = _flat_box(0, 1.0e4 * float(scale[0]), 10)
            other.write_bodies(before)
            other.set_manifolds(man)
            _teleport(other, z)
            other.step()
            a, b = o.download(), other.download()
            assert all(bits_equal(a[k], b[k]) for k in ("position", "rotation", "velocity", "angular_velocity"))
            mb = other.manifolds()
            assert bits_equal(m["penalty"], mb["penalty"]) and bits_equal(m["lambda_"], mb["lambda_"])
            return
    raise AssertionError("no step with a common multiplier above 1")


# ------------------------------------------------------------------- Newton rounds, decay ----

def _pile_state():
    b, j = scenes.pile(120, seed=3, pitch=2.0)
    o = _run(b, j, P(iterations=10), 90)
    return b, j, o.download(), o.manifolds()


def _one_step(state, **kw):
    b, j, st, m = state
    o = oracle.OracleWorld(P(**kw))
    o.upload({**b, **st}, j)
    o.set_manifolds(m)
    o.step()
    return o.download(), o.manifolds()


def pin_newton_rounds():
    """solver_step.cpp:199-202: the first ceil(r N) rounds are Newton rounds and the rest move no
    body, so a step of 7 iterations at Newton ratio 0.5 leaves every body exactly where a step of
    4 plain iterations does (the dual-only rounds change lambda and k, not poses)."""
    state = _pile_state()
    a, _ = _one_step(state, iterations=7, newton_ratio=0.5)
    c, _ = _one_step(state, iterations=4)
    assert all(bits_equal(a[k], c[k]) for k in ("position", "rotation", "velocity", "angular_velocity"))


def pin_stiffness_decay():
    """solver_step.cpp:277-296: in a dual-only round the contact penalties are scaled by the decay
    before the dual.  From one state: R1 = 1 iteration, R2 = 2 iterations (Newton ratio 0.5, decay
    1), R3 = the same with decay 0.5.  Round 2 sees the poses of round 1 in both, so for every
    pushing contact whose normal penalty stays above the floor: k(R2) - k(R3) = 0.5 k(R1) (the ramp
    adds the same beta |C|), and lambda(R3) - lambda(R1) = 0.5 (lambda(R2) - lambda(R1)) (the force
    k C + lambda with half the penalty)."""
    state = _pile_state()
    _, m1 = _one_step(state, iterations=1)
    s2, m2 = _one_step(state, iterations=2, newton_ratio=0.5, stiffness_decay=1.0)
    s3, m3 = _one_step(state, iterations=2, newton_ratio=0.5, stiffness_decay=0.5)
    assert bits_equal(s2["position"], s3["position"])
    k = [m["penalty"].reshape(-1, 8, 3)[:, :, 0].astype(np.float64) for m in (m1, m2, m3)]
    lam = [m["lambda_"].reshape(-1, 8, 3)[:, :, 0].astype(np.float64) for m in (m1, m2, m3)]
    checked = 0
    for s_, nc in enumerate(m1["num_contacts"]):
        for i in range(nc):
            if lam[1][s_, i] < 0 and lam[2][s_, i] < 0 and k[0][s_, i] * 0.5 >= 1.0:
                checked += 1
                assert abs((k[1][s_, i] - k[2][s_, i]) - 0.5 * k[0][s_, i]) <= 1e-5 * k[0][s_, i]
                d2, d3 = lam[1][s_, i] - lam[0][s_, i], lam[2][s_, i] - lam[0][s_, i]
                assert abs(d3 - 0.5 * d2) <= 1e-3 * abs(d2) + 1e-5 * abs(lam[0][s_, i]) + 1e-9
    assert checked > 100


def pin_convergence_probe():
    """solver_step.cpp:88-89, 299-314: the probe starts each substep from a previous residual of
    0 and reads the step-start C0 (manifold.cpp:672-692), so the mean penetration it sees is the
    same every round and never negative: with any positive threshold the first round ends the
    iterations.  A penetrating pile stepped with 10 rounds and threshold 1e-6 (and 1e3) equals, bit
    for bit, the same step with 1 round; with threshold 0 it does not (the other rounds matter)."""
    state = _pile_state()
    one, m1 = _one_step(state, iterations=1)
    assert float(np.max(-m1["c0"].reshape(-1, 8, 3)[:, :, 0])) > 1e-4      # it does penetrate
    for thr in (1e-6, 1e3):
        got, mg = _one_step(state, iterations=10, convergence_threshold=thr)
        assert all(bits_equal(got[k], one[k]) for k in ("position", "rotation", "velocity", "angular_velocity"))
        assert bits_equal(mg["lambda_"], m1["lambda_"]) and bits_equal(mg["penalty"], m1["penalty"])
    ten, _ = _one_step(state, iterations=10)
    assert not bits_equal(ten["position"], one["position"])


# ------------------------------------------------------------------- hard joints ----

def _anchor_error(st, r_anchor, body, r_b):
    p, q = st["position"][body].astype(np.float64), st["rotation"][body]
    return np.asarray(r_anchor, np.float64) - (p + scenes._qrot64(q, r_b))


def pin_hard_joint_force_is_the_weight():
    """A 2 kg box hanging from a static anchor by a ball joint (k_lin infinite): at rest the
    joint's lambda is the force that holds the weight, (0, 0, m |g|), and the anchors meet."""
    s = scenes.SceneBuilder()
    anchor = s.box((1, 1, 1), 0.0, 0.5, (0, 0, 10))
    box = s.box((1, 1, 1), 2.0, 0.5, (0, 0, 9))
    s.joint(anchor, box, (0, 0, -0.5), (0, 0, 0.5))
    o = _run(*s.build(), P(iterations=10), 300)
    st, jt = o.download(joints=True)
    lam = jt["lambda_lin"][0].astype(np.float64)
    assert np.abs(lam - np.array([0, 0, 20.0])).max() <= 2e-3 * 20.0, lam
    assert np.linalg.norm(_anchor_error(st, (0, 0, 9.5), 0, (0, 0, 0.5))) < 5e-3


def pin_hard_joint_error_decays_by_alpha():
    """joint.cpp:80-190: a hard joint solves C(x) - alpha C(x-) = 0 each step, so with gravity off
    an initial 1 cm anchor offset shrinks by the factor alpha = 0.99 per step, keeping its sign."""
    s = scenes.SceneBuilder(gravity=0.0)
    anchor = s.box((1, 1, 1), 0.0, 0.5, (0, 0, 10))
    box = s.box((1, 1, 1), 1.0, 0.5, (0.01, 0, 9))
    s.joint(anchor, box, (0, 0, -0.5), (0, 0, 0.5))
    o = oracle.OracleWorld(P(iterations=10, beta_lin=1e5))
    o.upload(*s.build())
    err = []
    for _ in range(30):
        o.step()
        err.append(_anchor_error(o.download(), (0, 0, 9.5), 0, (0, 0, 0.5))[0])
    assert all(e < 0 for e in err), err
    for k in range(10, 29):
        assert abs(err[k + 1] / err[k] - 0.99) < 1e-3, (k, err[k + 1] / err[k])


def pin_joint_penalty_cap():
    """joint.cpp:185-188: the linear penalty ramps to min(stiffness, penalty_max).  Two hard joints
    pull one box towards anchors 2 m apart, so C never vanishes; with a ramp (beta_lin 1e12) far
    past the cap, every linear penalty is exactly penalty_max (1e8) after a step.  A finite-
    stiffness joint (k_lin 500) under the same ramp sits exactly at its stiffness."""
    s = scenes.SceneBuilder()
    a1 = s.box((1, 1, 1), 0.0, 0.5, (-1, 0, 10))
    a2 = s.box((1, 1, 1), 0.0, 0.5, (1, 0, 10))
    box = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 10))
    s.joint(a1, box, (0, 0, 0), (0, 0, 0))
    s.joint(a2, box, (0, 0, 0), (0, 0, 0))
    o = _run(*s.build(), P(iterations=10, beta_lin=1e12, penalty_max=1e8), 3)
    pen = o.download(joints=True)[1]["penalty_lin"]
    assert pen[:, 0].tolist() == [1e8, 1e8], pen
    s = scenes.SceneBuilder()
    anchor = s.box((1, 1, 1), 0.0, 0.5, (0, 0, 10))
    box = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 9))
    s.joint(anchor, box, (0, 0, -0.5), (0, 0, 0.5), k_lin=500.0)
    pen = _run(*s.build(), P(iterations=10, beta_lin=1e12, penalty_max=1e8), 3).download(joints=True)[1]["penalty_lin"]
    assert pen[0, 2] == 500.0, pen


def _weld(fracture):
    s = scenes.SceneBuilder()
    anchor = s.box((1, 1, 1), 0.0, 0.5, (0, 0, 10))
    box = s.box((1, 1, 1), 1.0, 0.5, (1, 0, 10))
    s.joint(anchor, box, (0.5, 0, 0), (-0.5, 0, 0), k_lin=math.inf, k_ang=math.inf, fracture=fracture)
    return _run(*s.build(), P(iterations=10), 300)


def pin_fracture():
    """joint.cpp:221-229: a weld (k_lin, k_ang infinite) holds a 1 kg box cantilevered 0.5 m past
    its joint.  At rest its angular lambda carries the torque m g 0.5 = 5 N m through the torque
    arm |size_a + size_b|^2 = 12, so |lambda_ang| = 5/12, and lambda_lin = (0, 0, m |g|).  The
    joint breaks when |lambda_ang| passes the fracture threshold: at 0.7 (above the largest
    |lambda_ang| of the run, 0.63 while it settles) it holds; at 0.5 it breaks and the box falls."""
    st, jt = _weld(math.inf).download(joints=True)
    assert abs(float(np.linalg.norm(jt["lambda_ang"][0])) - 5.0 / 12.0) < 1e-3, jt["lambda_ang"]
    assert np.abs(jt["lambda_lin"][0].astype(np.float64) - [0, 0, 10.0]).max() < 1e-2, jt["lambda_lin"]
    st, jt = _weld(0.7).download(joints=True)
    assert len(jt["broken"]) == 1 and not jt["broken"][0] and abs(st["position"][0, 2] - 10.0) < 1e-2
    st, jt = _weld(0.5).download(joints=True)
    assert len(jt["broken"]) == 0 and st["position"][0, 2] < 0.0


# ------------------------------------------------------------------- arm caching ----
# manifold.cpp:583-671 (jacobianRebuildDistance): the dual reuses the world arms computed last
# while neither body's first rotation column has moved farther than the distance from where it
# was when the arms were captured; a new contact set (every tick) drops the capture.

def _both(bodies, joints, params_a, params_b, steps, interleaved=None):
    out = []
    for prm in (params_a, params_b):
        o = oracle.OracleWorld(prm)
        o.upload(bodies, joints)
        for _ in range(steps):
            if interleaved:
                o.step_interleaved(*interleaved)
            else:
                o.step()
        out.append(o)
    return out


def _same_run(x, y):
    sx, sy = x.download(), y.download()
    mx, my = x.manifolds(), y.manifolds()
    return all(bits_equal(sx[k], sy[k]) for k in ("position", "rotation", "velocity", "angular_velocity")) and \
        all(bits_equal(mx[k], my[k]) for k in ("penalty", "lambda_", "c0"))


def pin_arm_cache_tiny_distance_is_no_cache():
    """A distance so small that only an unchanged first rotation column passes the drift test:
    the cached arms are then the arms a recomputation gives, so a tumbling pile (Jacobi path) and
    a pile through the colour sweep (contact-only bodies coloured) end bit-identical to a run
    without caching."""
    b, j = scenes.pile(40, seed=9, pitch=1.6, ground=30.0)
    for overflow in (1, 0):
        x, y = _both(b, j, P(contact_only_overflow=overflow), P(contact_only_overflow=overflow,
                     jacobian_rebuild_distance=1e-19), 60)
        assert _same_run(x, y), overflow


def _sliders(friction=0.0):
    """Boxes alone on the ground (resting, sliding frictionless): their rotations stay within
    1e-8 of the identity, so their world arms do not move in float32."""
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, friction, (0, 0, 0))
    s.box((1, 1, 1), 1.0, friction, (0.0, 0.0, 1.0), vel=(2.0, 0.0, 0))
    s.box((1, 2, 1), 2.0, friction, (6.0, 0.0, 1.0), vel=(0.0, -1.0, 0))
    s.box((1, 1, 1), 1.0, friction, (-6.0, 3.0, 1.2), vel=(1.0, 0.5, 0))
    return s.build()


def pin_arm_cache_reuses_the_arms_it_stored():
    """With a distance past any drift (2: the first columns are unit vectors) every dual after a
    tick's first reuses stored arms.  For boxes whose rotations stay within 1e-8 of the identity
    the stored arms are the recomputed ones up to the last bits, so the run stays within 1e-6 of
    one without caching, through the Jacobi path and the colour sweep, while the cache is used (a
    swapped side or contact would change the constraint error by centimetres)."""
    b, j = _sliders()
    for overflow in (1, 0):
        x, y = _both(b, j, P(contact_only_overflow=overflow), P(contact_only_overflow=overflow,
                     jacobian_rebuild_distance=2.0), 40)
        sx, sy = x.download(), y.download()
        assert np.abs(sy["rotation"][:, :3].astype(np.float64)).max() < 1e-8, sy["rotation"]
        hits = y.cache_counters()[0]
        assert hits >= 40 * 2 * 9, hits            # 9 cached duals of 10 per step and manifold
        for k in ("position", "rotation", "velocity"):
            d = np.abs(sx[k].astype(np.float64) - sy[k].astype(np.float64)).max()
            assert d < 1e-6, (overflow, k, d)


def pin_arm_cache_changes_a_tumbling_pile():
    """The same large distance on a pile whose boxes tumble: the reused arms lag the rotations,
    so the run departs from the uncached one (the cache is in effect)."""
    b, j = scenes.pile(40, seed=9, pitch=1.6, ground=30.0)
    x, y = _both(b, j, P(), P(jacobian_rebuild_distance=2.0), 60)
    assert y.cache_counters()[0] > 0 and not _same_run(x, y)


def _spinner(omega):
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.0, (0, 0, 0))
    s.box((1, 1, 1), 1.0, 0.0, (0, 0, 1.0))
    b, j = s.build()
    b["angular_velocity"] = np.array([[0, 0, omega], [0, 0, 0]], np.float32)
    return b, j


def _predicted_hits(theta, dist, substeps, rounds):
    """Drift rule on a body turning theta per substep about the contact normal (its first column
    turns in the ground plane, |col0(a) - col0(b)| = 2 sin(|a - b| / 2)): per tick the capture is
    dropped, then each substep's duals hit while the chord to the capture is within the distance
    and recapture on a miss."""
    hits, cap = 0, None
    for sub in range(substeps):
        ang = sub * theta
        for _ in range(rounds):
            if cap is not None and 2.0 * math.sin(abs(ang - cap) / 2.0) <= dist:
                hits += 1
            else:
                cap = ang
    return hits


def pin_arm_cache_hits_follow_the_drift_rule():
    """A frictionless box spinning about the vertical on the ground, four interleaved substeps
    of two rounds: the rounds never turn it about another axis, so whether a dual reuses the arms
    follows from the chord its first column has swept since the capture.  Distances between the
    one-, two- and three-substep chords give 4, 6 and 7 hits per tick, and a tick's first dual never
    hits (its contact set is new)."""
    omega, dt = 12.0, 1.0 / 240.0
    theta = omega * dt
    b, j = _spinner(omega)
    ticks = 20
    for dist, want in ((0.03, 4), (0.075, 6), (0.12, 6), (0.2, 7)):
        o = oracle.OracleWorld(P(dt=dt, jacobian_rebuild_distance=dist))
        o.upload(b, j)
        for _ in range(ticks):
            o.step_interleaved(4, 2)
        st = o.download()
        spin = float(st["angular_velocity"][0, 2])
        # the spin slows a little (the normal rows' warm start); the prediction holds over the range
        assert 0.95 * omega < spin <= omega and np.abs(st["rotation"][0, :2]).max() < 1e-6, st["rotation"][0]
        assert _predicted_hits(theta, dist, 4, 2) == want == _predicted_hits(spin * dt, dist, 4, 2)
        assert o.cache_counters()[0] == ticks * want, (dist, o.cache_counters())


def pin_primal_cache_serves_the_first_touch():
    """solver_step.cpp:215-226 with manifold.cpp:570-581: in the colour sweep a manifold's first
    touch of every round after the first is served from the force cache the dual left; a box on
    the ground through the colour sweep (three rounds) has two such touches per step, a chain of
    boxes standing on each other one per manifold and round."""
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 1.0))
    o = _run(*s.build(), P(iterations=3, contact_only_overflow=0), 10)
    assert o.cache_counters()[1] == 10 * 2, o.cache_counters()
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    for k in range(4):
        s.box((1, 1, 1), 1.0, 0.5, (0, 0, 1.0 + k))
    o = _run(*s.build(), P(iterations=3, contact_only_overflow=0), 10)
    assert o.stats()["manifolds"] == 4 and o.cache_counters()[1] == 10 * 2 * 4, o.cache_counters()


def _tight_stack_params(**kw):
    return P(dt=1.0 / 240.0, gravity=-9.8, iterations=12, beta_lin=1e5, speculative_margin=0.02,
             jacobian_rebuild_distance=0.02, projection_max_correction=0.1, sleep_frames=30,
             sleep_linear_threshold=0.05, sleep_angular_threshold=0.05, **kw)


def pin_arm_cache_reference_tight_stack():
    """core_test.cpp:426-456: ten unit cubes stacked on the ground (sleep when idle), arm caching
    at 2 cm, 600 ticks of stepInterleaved(4, 2): height error <= 5 cm, interface penetration
    <= 1 cm, lateral drift <= 1 cm."""
    s = scenes.SceneBuilder(gravity=-9.8)
    s.box((100, 100, 1), 0.0, 0.8, (0, 0, 0))
    ids = [s.box((1, 1, 1), 1.0, 0.8, (0, 0, 1.0 + i)) for i in range(10)]
    b, j = s.build()
    n = len(s.b)
    b["sleep_mode"] = np.where(b["mass"] > 0, 1, 0).astype(np.uint8)
    o = oracle.OracleWorld(_tight_stack_params())
    o.upload(b, j)
    for _ in range(600):
        o.step_interleaved(4, 2)
    pos = o.download()["position"].astype(np.float64)
    z = [pos[_pos(k, n)] for k in ids]
    height = max(abs(z[i][2] - (1.0 + i)) for i in range(10))
    pen = max((1.0 - z[i][2]) if i == 0 else 1.0 - (z[i][2] - z[i - 1][2]) for i in range(10))
    drift = max(math.hypot(z[i][0], z[i][1]) for i in range(10))
    assert height <= 0.05 and pen <= 0.01 and drift <= 0.01, (height, pen, drift)
    assert o.cache_counters()[0] > 0


def pin_arm_cache_reference_impact():
    """core_test.cpp:536-575: a cube arriving at 18 m/s on a resting cube, arm caching at 2 cm,
    240 ticks of stepInterleaved(4, 2): the overlap stays <= 2 cm, no body passes 20 m/s, the base
    stays within 5 cm of z = 1."""
    s = scenes.SceneBuilder(gravity=-9.8)
    s.box((100, 100, 1), 0.0, 0.8, (0, 0, 0))
    target = s.box((1, 1, 1), 1.0, 0.8, (0, 0, 1.0))
    striker = s.box((1, 1, 1), 1.0, 0.8, (0, 0, 3.0), vel=(0, 0, -18.0))
    b, j = s.build()
    n = len(s.b)
    b["sleep_mode"] = np.where(b["mass"] > 0, 1, 0).astype(np.uint8)
    o = oracle.OracleWorld(_tight_stack_params())
    o.upload(b, j)
    worst = speed = 0.0
    for _ in range(240):
        o.step_interleaved(4, 2)
        st = o.download()
        zs, zt = float(st["position"][_pos(striker, n), 2]), float(st["position"][_pos(target, n), 2])
        worst = max(worst, 1.0 - (zs - zt))
        speed = max(speed, float(np.linalg.norm(st["velocity"].astype(np.float64), axis=1).max()))
    assert worst <= 0.02 and speed <= 20.0 and abs(zt - 1.0) <= 0.05, (worst, speed, zt)
    assert o.cache_counters()[0] > 0


# ------------------------------------------------------------------- residuals ----
# The constraint values the step's last dual computes (manifold.cpp:626-630, joint.cpp:176-182),
# kept for the long-run residual statistics.

def pin_dual_residual_of_a_resting_contact():
    """A box resting on the ground keeps its stuck contacts, so a step starts from the contact
    error the previous step's rebase left (C0_prev), and the rebase after the dual measures the
    pose the dual saw (C0_now; nothing moves the resting box in between): the dual's normal row
    C = (1 - alpha) C0_prev + n . bracket equals C0_now - alpha C0_prev to rounding."""
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 1.0))
    o = oracle.OracleWorld(P(iterations=10))
    o.upload(*s.build())
    o.step(40)
    prev = o.manifolds()["c0"].reshape(-1, 8, 3)[0, :4, 0].astype(np.float64)
    for _ in range(10):
        o.step()
        m = o.manifolds()
        assert m["num_contacts"].tolist() == [4]
        now = m["c0"].reshape(-1, 8, 3)[0, :4, 0].astype(np.float64)
        cn, _ = o.residuals()
        assert np.abs(cn[0, :4] - (now - 0.99 * prev)).max() < 1e-9, (cn[0, :4], now - 0.99 * prev)
        assert np.abs(cn[0, :4]).max() > 1e-8
        prev = now


def pin_dual_residual_of_a_hard_joint():
    """joint.cpp:176-182: a hard joint's dual value is C(x) - alpha C0.  A box hanging from a static
    anchor: its final pose is the pose of the last dual (no contact moves it afterwards), so the
    kept value equals the anchor gap at the final pose minus alpha times the step's C0, to
    rounding; a spring records nothing."""
    s = scenes.SceneBuilder()
    anchor = s.box((1, 1, 1), 0.0, 0.5, (0, 0, 10))
    box = s.box((1, 1, 1), 2.0, 0.5, (0.3, 0, 9))
    s.joint(anchor, box, (0, 0, -0.5), (0, 0, 0.5))
    a2 = s.box((1, 1, 1), 0.0, 0.5, (5, 0, 10))
    b2 = s.box((1, 1, 1), 1.0, 0.5, (5, 0, 8.5))
    s.spring(a2, b2, (0, 0, -0.5), (0, 0, 0.5), 50.0, rest=1.0)
    bodies, joints = s.build()
    n = len(s.b)
    o = oracle.OracleWorld(P(iterations=10))
    o.upload(bodies, joints)
    for _ in range(25):
        o.step()
        st, jt = o.download(joints=True)
        _, cl = o.residuals()
        jb = int(np.nonzero(joints["kind"] == pkg.abi.FORCE_JOINT)[0][0])     # the force list is newest first
        sp = int(np.nonzero(joints["kind"] == pkg.abi.FORCE_SPRING)[0][0])
        pb, qb = st["position"][_pos(box, n)].astype(np.float64), st["rotation"][_pos(box, n)]
        gap = np.array([0.0, 0.0, 9.5]) - (pb + scenes._qrot64(qb, (0, 0, 0.5)))
        want = gap - 0.99 * jt["c0_lin"][jb].astype(np.float64)
        assert np.abs(cl[jb] - want).max() < 1e-6 + 1e-5 * np.abs(want).max(), (cl[jb], want)
        assert np.abs(cl[jb]).max() > 0 and not cl[sp].any()


# ------------------------------------------------------------------- test entry points ----

PINS = {name: fn for name, fn in globals().items() if name.startswith("pin_")}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    PINS[name]()


# planted mistake id (avbd_oracle.cpp) -> the pins that must catch it
PLANTS = {
    1: ["pin_layer_mask"], 2: ["pin_layer_mask"],
    3: ["pin_axis_lock"], 4: ["pin_locked_axis_reports_no_velocity"], 5: ["pin_axis_lock"],
    6: ["pin_adaptive_beta_rule"], 7: ["pin_adaptive_beta_rule"], 8: ["pin_adaptive_scale_multiplies_beta"],
    9: ["pin_newton_rounds"], 10: ["pin_stiffness_decay"], 11: ["pin_stiffness_decay"],
    12: ["pin_hard_joint_force_is_the_weight"], 13: ["pin_hard_joint_error_decays_by_alpha"],
    14: ["pin_joint_penalty_cap"], 15: ["pin_fracture"], 16: ["pin_fracture"],
    17: ["pin_projection_shared_plane"], 18: ["pin_projection_shared_plane"],
    19: ["pin_contact_only_overflow_off"], 20: ["pin_speculative_horizon"], 21: ["pin_convergence_probe"],
    22: ["pin_arm_cache_reuses_the_arms_it_stored"], 23: ["pin_arm_cache_tiny_distance_is_no_cache"],
    24: ["pin_arm_cache_hits_follow_the_drift_rule"], 25: ["pin_arm_cache_hits_follow_the_drift_rule"],
    26: ["pin_primal_cache_serves_the_first_touch"],
    27: ["pin_dual_residual_of_a_hard_joint"], 28: ["pin_dual_residual_of_a_resting_contact"],
}


@pytest.mark.parametrize("pid", sorted(PLANTS))
def test_planted_mistake_is_caught(pid):
    failed = []
    with oracle.planted(pid):
        for name in PLANTS[pid]:
            try:
                PINS[name]()
            except AssertionError:
                failed.append(name)
    assert failed, f"planted mistake {pid} passed {PLANTS[pid]}"


def test_pins_pass_without_plants():
    """The plant switch is off by default and back off after a planted run."""
    with oracle.planted(9):
        pass
    pin_newton_rounds()

"""The HIP path against the oracle on the scenes that pin the oracle's less common features
(tests/test_oracle_pins.py): collision layers, axis locks under a constraint, the projection and
its cap, the colour sweep for contact bodies, adaptive beta with its per-contact multipliers,
Newton rounds with stiffness decay, hard joints at the penalty cap, joint fracture, and drift-gated
arm reuse (plain, interleaved, sleeping, round shapes with grace, switched mid-run, batched).  Bar:
bodies, every manifold field and every joint field bit-equal, every step.
"""
import math

import numpy as np
import pytest

from helpers import bits_equal, oracle, pkg, scenes, state_mismatch

pytestmark = pytest.mark.gpu
P = pkg.default_params
F = np.float32
MANIFOLD_KEYS = ("body_a", "body_b", "basis", "friction", "num_contacts", "feature", "r_a", "r_b", "c0", "penalty",
                 "lambda_", "stick", "beta_scale")
JOINT_KEYS = ("body_a", "body_b", "c0_lin", "c0_ang", "penalty_lin", "penalty_ang", "lambda_lin", "lambda_ang", "broken")


def _lockstep(bodies, joints, params, steps, before_step=None, inter=None):
    nj = 0 if joints is None else len(joints["body_b"])
    w = pkg.World(params, len(bodies["mass"]), nj)
    w.upload(bodies, joints)
    o = oracle.OracleWorld(params)
    o.upload(bodies, joints)
    for s in range(steps):
        if before_step:
            before_step(s, w, o)
        if inter:
            w.step_interleaved(*inter)
            o.step_interleaved(*inter)
        else:
            w.step()
            o.step()
        bad = state_mismatch(w.download(), o.download())
        assert not bad, f"step {s}: {bad}"
        gm, om = w.manifolds(), o.manifolds()
        assert len(gm["body_a"]) == len(om["body_a"]), f"step {s}"
        for k in MANIFOLD_KEYS:
            assert bits_equal(gm[k], om[k]), f"step {s}: manifold {k}"
        if nj:
            gj, oj = w.download(joints=True)[1], o.download(joints=True)[1]
            for k in JOINT_KEYS:
                assert bits_equal(gj[k], oj[k]), f"step {s}: joint {k}"
    return w, o


def _pos(created, n):
    return n - 1 - created


def test_layers_and_locks_under_a_spring():
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    a = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 4))
    f = s.box((1, 1, 1), 1.0, 0.5, (5, 0, 4))
    anchor = s.box((0.5, 0.5, 0.5), 0.0, 0.5, (3, 3, 5))
    pulled = s.box((1, 1, 1), 1.0, 0.5, (0, 3, 5))
    spin = s.box((1, 1, 1), 1.0, 0.5, (-3, 0, 5.0))
    s.b[spin]["ang"] = (5.0, 1.0, 0.0)
    s.spring(anchor, pulled, (0, 0, 0), (0.2, 0, 0), 50.0, rest=1.0)
    b, j = s.build()
    n = len(s.b)
    b["layer"], b["mask"] = np.ones(n, np.uint32), np.ones(n, np.uint32)
    b["layer"][_pos(a, n)], b["mask"][_pos(a, n)] = 4, 0xFFFFFFFF
    b["layer"][_pos(f, n)], b["mask"][_pos(f, n)] = 2, 2
    b["lock_linear"], b["lock_angular"] = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    b["lock_linear"][_pos(pulled, n)] = 0x5
    b["lock_angular"][_pos(spin, n)] = 0x1
    _lockstep(b, j, P(iterations=10), 120)


@pytest.mark.parametrize("cap", [0.0, 0.01])
def test_projection_cap_on_shared_planes(cap):
    s = scenes.SceneBuilder(gravity=0.0)
    for _ in range(3):
        s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    s.sphere(0.5, 10.0, 0.5, (0, 0, 0.9))
    s.box((1, 1, 1), 2.0, 0.5, (3, 0, 0.85))
    p = P(alpha=1.0, iterations=1, projection_iterations=2, projection_beta=1.0, projection_slop=0.0,
          projection_max_correction=cap)
    _lockstep(*s.build(), p, 5)


def test_pile_through_the_colour_sweep_with_projection_cap():
    b, j = scenes.pile(200, seed=12, pitch=2.0)
    _lockstep(b, j, P(iterations=10, contact_only_overflow=0, projection_max_correction=0.004), 90)


def test_adaptive_beta_multipliers_under_teleports():
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    s.box((1, 1, 1), 1.0, 0.5, (0, 0, 1.0005))
    b, j = s.build()
    b["gravity"][0] = 0.0
    rng = np.random.default_rng(1)
    depths = rng.choice([0.0, 0.0005, 0.0015, 0.003, 0.008, 0.012, 0.016, 0.019], 60)
    zero = np.zeros((2, 3), F)

    def teleport(k, w, o):
        z = F(1.0 - float(depths[k]))
        arrays = {"position": np.array([[0, 0, z], [0, 0, 0]], F), "rotation": np.array([[0, 0, 0, 1]] * 2, F),
                  "velocity": zero, "angular_velocity": zero, "prev_velocity": zero}
        w.write_bodies(arrays)
        o.write_bodies(arrays)
    _lockstep(b, j, P(iterations=1, adaptive_beta=1, beta_lin=1.0, position_projection=0), 60, teleport)


def test_newton_ratio_and_stiffness_decay():
    b, j = scenes.pile(150, seed=3, pitch=2.0)
    _lockstep(b, j, P(iterations=7, newton_ratio=0.5, stiffness_decay=0.5), 60)


def test_hard_joints_at_the_penalty_cap():
    s = scenes.SceneBuilder()
    a1 = s.box((1, 1, 1), 0.0, 0.5, (-1, 0, 10))
    a2 = s.box((1, 1, 1), 0.0, 0.5, (1, 0, 10))
    box = s.box((1, 1, 1), 1.0, 0.5, (0, 0, 10))
    s.joint(a1, box, (0, 0, 0), (0, 0, 0))
    s.joint(a2, box, (0, 0, 0), (0, 0, 0))
    hang = s.box((1, 1, 1), 2.0, 0.5, (0, 0, 8.9))
    s.joint(box, hang, (0, 0, -0.5), (0, 0, 0.5), k_lin=500.0)
    w, o = _lockstep(*s.build(), P(iterations=10, beta_lin=1e12, penalty_max=1e8), 20)
    assert o.download(joints=True)[1]["penalty_lin"][1:, 0].tolist() == [1e8, 1e8]


@pytest.mark.parametrize("fracture", [math.inf, 0.7, 0.5])
def test_weld_fracture(fracture):
    s = scenes.SceneBuilder()
    anchor = s.box((1, 1, 1), 0.0, 0.5, (0, 0, 10))
    box = s.box((1, 1, 1), 1.0, 0.5, (1, 0, 10))
    s.joint(anchor, box, (0.5, 0, 0), (-0.5, 0, 0), k_lin=math.inf, k_ang=math.inf, fracture=fracture)
    s.box((40, 40, 1), 0.0, 0.5, (0, 0, 0))
    w, o = _lockstep(*s.build(), P(iterations=10), 240)
    broken = len(o.download(joints=True)[1]["broken"]) == 0
    assert broken == (fracture == 0.5)


@pytest.mark.parametrize("inter", [None, (2, 3)])
def test_convergence_threshold(inter):
    b, j = scenes.pile(150, seed=9, pitch=2.0)
    p = P(iterations=10, convergence_threshold=1e-5)
    if inter:
        p.dt = float(F(1.0) / F(120.0))
    _lockstep(b, j, p, 40, inter=inter)


def test_convergence_threshold_in_a_batch():
    p = P(iterations=10, convergence_threshold=1e-5)
    worlds = [scenes.pile(60 + 40 * k, seed=30 + k, ground=60.0) for k in range(4)]
    batch = pkg.Batch(p, 4, sum(len(w[0]["mass"]) for w in worlds), 0)
    ids = [batch.add(bb, jj) for bb, jj in worlds]
    orc = []
    for bb, jj in worlds:
        o = oracle.OracleWorld(p)
        o.upload(bb, jj)
        orc.append(o)
    for _ in range(30):
        batch.step()
        for o in orc:
            o.step()
    for wid, o in zip(ids, orc):
        assert not state_mismatch(batch.download(wid), o.download())
    batch.close()


# ---- arm reuse (jacobianRebuildDistance, manifold.cpp:583-671) ----

@pytest.mark.parametrize("overflow,dist", [(1, 0.02), (1, 2.0), (0, 0.02), (0, 2.0)])
def test_arm_cache_pile_lockstep(overflow, dist):
    """Arm reuse through the Jacobi path and through the colour sweep (contact bodies coloured),
    at the reference's 2 cm and at a distance past any drift."""
    b, j = scenes.pile(150, seed=9, pitch=2.0)
    _lockstep(b, j, P(iterations=10, contact_only_overflow=overflow, jacobian_rebuild_distance=dist), 60)


@pytest.mark.parametrize("dist", [0.02, 2.0])
def test_arm_cache_chains_and_mixed_lockstep(dist):
    """Jointed chains (colour sweep) colliding with each other and the ground, and the reference's
    mixed digest scene, with arm reuse: first touches served from the stored arms."""
    b, j = scenes.chains(n_chains=6, links=16, seed=5)
    _lockstep(b, j, P(iterations=10, jacobian_rebuild_distance=dist), 80)
    b, j = scenes.digest_mixed()
    _lockstep(b, j, P(iterations=10, jacobian_rebuild_distance=dist), 80)


def test_arm_cache_interleaved_spinner_and_pile():
    """Interleaved substeps keep the captured arms across substeps of a tick: the spinning
    frictionless box of the drift-rule pin, and a pile."""
    s = scenes.SceneBuilder()
    s.box((100, 100, 1), 0.0, 0.0, (0, 0, 0))
    s.box((1, 1, 1), 1.0, 0.0, (0, 0, 1.0))
    b, j = s.build()
    b["angular_velocity"] = np.array([[0, 0, 12.0], [0, 0, 0]], np.float32)
    dt = float(F(1.0) / F(240.0))
    for dist in (0.03, 0.075, 0.12, 0.2):
        _lockstep(b, j, P(dt=dt, jacobian_rebuild_distance=dist), 20, inter=(4, 2))
    b, j = scenes.pile(120, seed=4, pitch=2.0)
    _lockstep(b, j, P(dt=dt, jacobian_rebuild_distance=0.05), 40, inter=(4, 2))


def test_arm_cache_with_sleep_round_shapes_and_grace():
    """Frozen manifolds keep their captured arms until a neighbour wakes their bodies (the
    sleepy pile); round shapes and contact grace take the narrow passes that drop the capture."""
    b, j = scenes.sleepy_pile(160, seed=6)
    _lockstep(b, j, P(iterations=10, sleep_frames=20, jacobian_rebuild_distance=2.0), 160)
    b, j = scenes.round_pile(120, seed=4)
    for dist in (0.0, 0.02):    # 0: also the basis of box pairs parked without contacts
        _lockstep(b, j, P(iterations=10, contact_grace_ticks=3, jacobian_rebuild_distance=dist), 80)


def test_arm_cache_switched_on_mid_run():
    """set_params turns arm reuse on after 20 steps (the arrays are allocated then) and off again."""
    b, j = scenes.pile(100, seed=12, pitch=2.0)
    on = P(iterations=10, jacobian_rebuild_distance=0.5)

    def switch(s, w, o):
        if s in (20, 45):
            prm = on if s == 20 else P(iterations=10)
            w.set_params(prm)
            o.set_params(prm)
    _lockstep(b, j, P(iterations=10), 60, before_step=switch)


def test_arm_cache_in_a_batch():
    p = P(iterations=10, jacobian_rebuild_distance=0.02)
    worlds = [scenes.pile(60 + 40 * k, seed=40 + k, ground=60.0) for k in range(3)]
    worlds.append(scenes.chains(n_chains=4, links=16, seed=44))
    batch = pkg.Batch(p, 4, sum(len(w[0]["mass"]) for w in worlds), sum(0 if w[1] is None else len(w[1]["body_b"]) for w in worlds))
    ids = [batch.add(bb, jj) for bb, jj in worlds]
    orc = []
    for bb, jj in worlds:
        o = oracle.OracleWorld(p)
        o.upload(bb, jj)
        orc.append(o)
    for _ in range(40):
        batch.step()
        for o in orc:
            o.step()
    for wid, o in zip(ids, orc):
        assert not state_mismatch(batch.download(wid), o.download())
    batch.close()

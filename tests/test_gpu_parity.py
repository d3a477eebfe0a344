"""GPU parity of the HIP path against the oracle, through the C ABI.

Bar: integer outputs (Morton codes, sort order, candidate pairs, feature keys, live
manifold order, colours) bit-exact; FP32 state bit-exact as well, because the HIP path
keeps the reference's evaluation order (the stated fallback tolerance, used only where a
test says so, is |dp| <= 1e-5 (1 + |p|), |dq| <= 1e-6, |d lambda|, |d k| <= 1e-5 relative).
"""
import numpy as np
import pytest

from helpers import (DIGEST_INTERLEAVED_300, DIGEST_MIXED_240, DIGEST_STACK_300, bits_equal, dt120, oracle, pkg,
                     scenes, state_mismatch)

pytestmark = pytest.mark.gpu
P = pkg.default_params
MANIFOLD_KEYS = ("body_a", "body_b", "basis", "friction", "num_contacts", "feature", "r_a", "r_b", "c0", "penalty",
                 "lambda_", "stick")


def gpu_world(bodies, joints, params):
    w = pkg.World(params, len(bodies["mass"]), 0 if joints is None else len(joints["body_b"]))
    w.upload(bodies, joints)
    return w


def lockstep(bodies, joints, params, steps, inter=None, every=1):
    w = gpu_world(bodies, joints, params)
    o = oracle.OracleWorld(params)
    o.upload(bodies, joints)
    for s in range(steps):
        if inter:
            w.step_interleaved(*inter)
            o.step_interleaved(*inter)
        else:
            w.step()
            o.step()
        if (s + 1) % every == 0 or s + 1 == steps:
            bad = state_mismatch(w.download(), o.download())
            assert not bad, f"step {s}: {bad}"
    return w, o


def manifolds_equal(w, o):
    gm, om = w.manifolds(), o.manifolds()
    assert len(gm["body_a"]) == len(om["body_a"])
    for k in MANIFOLD_KEYS:
        assert bits_equal(gm[k], om[k]), k


def test_gpu_digest_stack_300():
    w = gpu_world(*scenes.digest_stack(), P())
    w.step(300)
    assert w.digest() == DIGEST_STACK_300


def test_gpu_digest_interleaved_300():
    p = P()
    p.dt = dt120()
    w = gpu_world(*scenes.digest_stack(), p)
    for _ in range(300):
        w.step_interleaved(2, 3)
    assert w.digest() == DIGEST_INTERLEAVED_300


def test_gpu_digest_mixed_240():
    w = gpu_world(*scenes.digest_mixed(), P())
    w.step(240)
    assert w.digest() == DIGEST_MIXED_240


def test_pyramid_lockstep_and_manifolds():
    w, o = lockstep(*scenes.pyramid(), P(iterations=10), 90, every=10)
    manifolds_equal(w, o)
    gs, os_ = w.stats(), o.stats()
    for k in ("manifolds", "contacts", "candidate_pairs", "new_manifolds"):
        assert gs[k] == os_[k], k


def test_pile_lockstep_and_manifolds():
    w, o = lockstep(*scenes.pile(600, seed=21), P(iterations=10), 150, every=15)
    manifolds_equal(w, o)


def test_springs_and_collision_filters_lockstep():
    """Springs (centred and off-centre anchors, between bodies and from static anchors), a
    joint and collision filters in one force list, bit for bit against the oracle; the
    filtered pairs make no manifold."""
    b, j = scenes.springs_and_filters()
    w, o = lockstep(b, j, P(iterations=10), 120, every=10)
    manifolds_equal(w, o)
    gs, os_ = w.stats(), o.stats()
    for k in ("manifolds", "contacts", "candidate_pairs", "colours", "overflow_bodies"):
        assert gs[k] == os_[k], k
    kinds = j["kind"]
    pairs = {(min(a, c), max(a, c)) for a, c, kk in zip(j["body_a"], j["body_b"], kinds)
             if kk == pkg.abi.FORCE_IGNORE_COLLISION}
    m = w.manifolds()
    assert not pairs & {(min(a, c), max(a, c)) for a, c in zip(m["body_a"], m["body_b"])}


def test_generic_joints_lockstep():
    """Generic 6-DOF joints (locked, limited, free and sprung axes; to a world point, from a
    rotated body, between dynamic bodies resting on a ground) bit for bit against the oracle,
    the axes' dual state included."""
    b, j = scenes.generic_joints()
    w, o = lockstep(b, j, P(iterations=10), 150, every=10)
    manifolds_equal(w, o)
    gs, os_ = w.stats(), o.stats()
    assert gs["manifolds"] > 0
    for k in ("manifolds", "contacts", "candidate_pairs", "colours", "overflow_bodies"):
        assert gs[k] == os_[k], k
    _, gj = w.download(joints=True)
    _, oj = o.download(joints=True)
    for k in ("axis_lambda", "axis_penalty", "axis_initial"):
        assert bits_equal(gj[k], oj[k]), k


def test_round_shapes_lockstep():
    """Spheres and capped cylinders among boxes (every shape pair: sphere-sphere, -box,
    -cylinder, cylinder-box, -cylinder), with rolling, stacking and static round obstacles,
    bit for bit against the oracle, manifolds included."""
    b, j = scenes.round_pile(300)
    w, o = lockstep(b, j, P(iterations=10), 150, every=10)
    manifolds_equal(w, o)
    gs, os_ = w.stats(), o.stats()
    assert gs["manifolds"] > 300
    for k in ("manifolds", "contacts", "candidate_pairs", "new_manifolds"):
        assert gs[k] == os_[k], k
    m = w.manifolds()
    shape = b["shape"]
    kinds = {(min(shape[a], shape[c]), max(shape[a], shape[c])) for a, c in zip(m["body_a"], m["body_b"])}
    assert kinds >= {(0, 1), (0, 2), (1, 1), (1, 2), (2, 2)}, kinds


def test_shape_validation():
    b, j = scenes.round_pile(30)
    b["shape"][3] = 7
    with pytest.raises(pkg.abi.AvbdError):
        gpu_world(b, j, P())


def test_chains_colour_sweep_lockstep():
    p = P(iterations=10, beta_lin=1e5, penalty_max=1e9)
    w, o = lockstep(*scenes.chains(16, links=32, grid=(4, 4), pitch_x=4.0), p, 150, every=15)
    manifolds_equal(w, o)
    gs, os_ = w.stats(), o.stats()
    assert gs["colours"] == os_["colours"] and gs["overflow_bodies"] == os_["overflow_bodies"]
    _, gj = w.download(joints=True)
    _, oj = o.download(joints=True)
    for k in ("penalty_lin", "lambda_lin", "c0_lin"):
        assert bits_equal(gj[k], oj[k]), k


def test_interleaved_pile_lockstep():
    p = P(iterations=10)
    p.dt = dt120()
    lockstep(*scenes.pile(200, seed=5, pitch=2.5), p, 80, inter=(2, 0), every=10)


def test_locks_layers_and_adaptive_beta():
    b, j = scenes.pile(150, seed=8, pitch=2.0)
    n = len(b["mass"])
    rng = np.random.default_rng(1)
    b["lock_linear"] = (rng.random(n) < 0.2).astype(np.uint8) * 1
    b["lock_angular"] = (rng.random(n) < 0.2).astype(np.uint8) * 6
    b["layer"] = np.where(rng.random(n) < 0.3, 2, 1).astype(np.uint32)
    b["mask"] = np.ones(n, np.uint32)
    b["mask"][-1] = 3          # the ground collides with both layers
    lockstep(b, j, P(iterations=6, adaptive_beta=1, newton_ratio=0.5, stiffness_decay=0.9), 100, every=10)


def test_single_step_from_identical_state():
    b, j = scenes.pile(500, seed=31)
    o = oracle.OracleWorld(P(iterations=10))
    o.upload(b, j)
    o.step(150)
    snap = o.download()
    m = o.manifolds()
    assert len(m["body_a"]) > 150
    start = {**b, **{k: snap[k] for k in ("position", "rotation", "velocity", "angular_velocity", "prev_velocity")}}
    w = gpu_world(start, j, P(iterations=10))
    w.set_manifolds(m)
    o2 = oracle.OracleWorld(P(iterations=10))
    o2.upload(start, j)
    o2.set_manifolds(m)
    w.step()
    o2.step()
    g, r = w.download(), o2.download()
    # stated tolerance (the result is in fact bit-exact, asserted below)
    assert np.all(np.abs(g["position"] - r["position"]) <= 1e-5 * (1 + np.abs(r["position"])))
    assert np.all(np.abs(g["rotation"] - r["rotation"]) <= 1e-6)
    assert not state_mismatch(g, r)
    manifolds_equal(w, o2)


def test_broadphase_probe_codes_order_pairs():
    b, j = scenes.pile(700, seed=41, pitch=2.5)
    w = gpu_world(b, j, P())
    o = oracle.OracleWorld(P())
    o.upload(b, j)
    for _ in range(3):
        codes, order, pairs = w.probe_broadphase()
        c, _ = o.swept_aabbs()
        oc, oo = oracle.morton(c)
        assert np.array_equal(codes, oc)
        assert np.array_equal(order, oo)
        assert np.array_equal(pairs, o.candidate_pairs())
        w.step(25)
        o.step(25)


def aabbs_overlap(f, i):
    def box(p, q, size):
        x, y, z, w = q.astype(np.float64)
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        h = np.abs(R) @ (size.astype(np.float64) / 2) + 0.01 + 1e-3
        return p - h, p + h
    a0, a1 = box(f["pa"][i], f["qa"][i], f["sa"][i])
    b0, b1 = box(f["pb"][i], f["qb"][i], f["sb"][i])
    return bool(np.all(a0 <= b1) and np.all(b0 <= a1))


def test_narrow_phase_against_reference_fixture():
    """Each fixture pair becomes a two-body world; the first step's manifold holds the
    narrow-phase output (feature keys, local arms, basis), which the rest of the step does
    not modify.  Expected values come from the reference's own collide.cpp."""
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "collide_boxes.npz"))
    n = len(f["count"])
    batch = pkg.Batch(P(), n, 2 * n)
    ids = []
    for i in range(n):
        s = scenes.SceneBuilder()
        s.box(f["sb"][i], 1.0, 0.5, f["pb"][i], rot=f["qb"][i])
        s.box(f["sa"][i], 1.0, 0.5, f["pa"][i], rot=f["qa"][i])   # list index 0 = A
        ids.append(batch.add(*s.build()))
    batch.step()
    checked = 0
    for i, wid in enumerate(ids):
        m = batch.manifolds(wid)
        c = int(f["count"][i])
        if c == 0:
            assert len(m["body_a"]) == 0, i
            continue
        if len(m["body_a"]) == 0:
            # SAT within the margin on all 15 axes does not imply swept-AABB overlap; such a
            # pair is legitimately dropped by the broad phase before the narrow phase runs
            assert not aabbs_overlap(f, i), i
            continue
        assert m["num_contacts"][0] == c, i
        assert np.array_equal(m["feature"][0, :c], f["keys"][i, :c]), i
        assert bits_equal(m["r_a"][0].reshape(8, 3)[:c], f["ra"][i, :c]), i
        assert bits_equal(m["r_b"][0].reshape(8, 3)[:c], f["rb"][i, :c]), i
        assert bits_equal(m["basis"][0], f["basis"][i]), i
        checked += 1
    assert checked > 100


def test_round_narrow_phase_against_reference_fixture():
    """The same for the sphere and cylinder pairs of tests/golden/collide_shapes.npz (every
    pair of shapes, margins 0 / 0.01 / 0.05: one batch per margin, the solver's speculative
    margin set to it; the bodies rest, so the margin is not grown)."""
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "collide_shapes.npz"))
    checked = 0

    def add(s, t, size, pos, rot):
        if t == pkg.abi.SHAPE_SPHERE:
            s.sphere(size[0], 1.0, 0.5, pos, rot=rot)
        elif t == pkg.abi.SHAPE_CYLINDER:
            s.cylinder(size[0], size[2], 1.0, 0.5, pos, rot=rot)
        else:
            s.box(size, 1.0, 0.5, pos, rot=rot)

    for margin in sorted(set(float(x) for x in f["margin"])):
        cases = [i for i in range(len(f["count"])) if float(f["margin"][i]) == margin]
        batch = pkg.Batch(P(speculative_margin=margin), len(cases), 2 * len(cases))
        ids = []
        for i in cases:
            s = scenes.SceneBuilder()
            add(s, f["tb"][i], f["sb"][i], f["pb"][i], f["qb"][i])
            add(s, f["ta"][i], f["sa"][i], f["pa"][i], f["qa"][i])   # list index 0 = A
            ids.append(batch.add(*s.build()))
        batch.step()
        for i, wid in zip(cases, ids):
            m = batch.manifolds(wid)
            c = int(f["count"][i])
            if c == 0:
                assert len(m["body_a"]) == 0, i
                continue
            if len(m["body_a"]) == 0:
                continue   # dropped by the swept-AABB broad phase before the narrow phase
            assert m["num_contacts"][0] == c, i
            assert np.array_equal(m["feature"][0, :c], f["keys"][i, :c]), i
            assert bits_equal(m["r_a"][0].reshape(8, 3)[:c], f["ra"][i, :c]), i
            assert bits_equal(m["r_b"][0].reshape(8, 3)[:c], f["rb"][i, :c]), i
            assert bits_equal(m["basis"][0], f["basis"][i]), i
            checked += 1
    assert checked > 300


def test_batch_alone_equals_in_mix():
    worlds = [scenes.pyramid(rows=12, seed=3), scenes.pile(300, seed=51), scenes.digest_mixed(),
              scenes.chains(4, links=16, grid=(2, 2), pitch_x=3.0), scenes.pile(40, seed=52), scenes.digest_stack(),
              scenes.generic_joints(), scenes.springs_and_filters(), scenes.round_pile(120, seed=9)]
    p = P(iterations=10)
    batch = pkg.Batch(p, 16, 4000, 2000)
    ids = [batch.add(b, j) for b, j in worlds]
    alone = [oracle.OracleWorld(p) for _ in worlds]
    for o, (b, j) in zip(alone, worlds):
        o.upload(b, j)
    for s in range(60):
        batch.step()
        for o in alone:
            o.step()
        if s == 29:
            # a world leaves and a new one arrives mid-run; the others must not notice
            batch.remove(ids[1])
            nb, nj = scenes.pile(120, seed=53)
            new_id = batch.add(nb, nj)
            o_new = oracle.OracleWorld(p)
            o_new.upload(nb, nj)
            ids[1], alone[1] = new_id, o_new
        if (s + 1) % 10 == 0:
            for wid, o in zip(ids, alone):
                assert not state_mismatch(batch.download(wid), o.download()), (s, wid)


def test_pinned_launch_shapes_keep_results():
    """Launch shapes set by the caller (here small blocks and a tiny grid, so every live-list
    kernel strides many times) give the same bits as the defaults."""
    worlds = [scenes.pile(500, seed=61), scenes.pyramid(rows=10, seed=62), scenes.pile(90, seed=63)]
    p = P(iterations=10)
    a, b = pkg.Batch(p, 4, 2000), pkg.Batch(p, 4, 2000)
    ia = [a.add(bo, j) for bo, j in worlds]
    ib = [b.add(bo, j) for bo, j in worlds]
    b.set_tuning({"block_dual": 128, "block_primal": 256, "block_stamp": 128, "block_projection": 64,
                  "block_traverse": 128, "live_blocks": 8})
    for _ in range(40):
        a.step()
        b.step()
    for x, y in zip(ia, ib):
        assert not state_mismatch(a.download(x), b.download(y))
    with pytest.raises(pkg.abi.AvbdError):
        b.set_tuning({"block_dual": 96, "block_primal": 64, "block_stamp": 64, "block_projection": 64,
                      "block_traverse": 64, "live_blocks": 8})


def test_colour_lanes_keep_results():
    """The colour passes with four lanes per body (a body's adjacency entries in parallel, the
    block summed in adjacency order by its first lane) give the one-lane bits: jointed chains that
    collide, the mixed digest scene with springs and generic joints, and a pile through the colour
    sweep; both equal the oracle."""
    p = P(iterations=10)
    worlds = [scenes.chains(n_chains=6, links=16, seed=71), scenes.springs_and_filters(), scenes.generic_joints()]
    pc = P(iterations=10, contact_only_overflow=0)
    runs = []
    for lanes in (1, 4):
        for prm, ws in ((p, worlds), (pc, [scenes.pile(300, seed=72, pitch=2.0)])):
            nb = sum(len(w[0]["mass"]) for w in ws)
            nj = sum(0 if w[1] is None else len(w[1]["body_b"]) for w in ws)
            batch = pkg.Batch(prm, 4, nb, nj)
            ids = [batch.add(bb, jj) for bb, jj in ws]
            batch.set_tuning({"block_dual": 64, "block_primal": 64, "block_stamp": 64, "block_projection": 64,
                              "block_traverse": 64, "live_blocks": 64, "colour_lanes": lanes})
            for _ in range(60):
                batch.step()
            runs.append([batch.download(i) for i in ids])
            batch.close()
    for one, four in zip(runs[:2], runs[2:]):
        for x, y in zip(one, four):
            assert not state_mismatch(x, y)
    for (bb, jj), st in zip(worlds, runs[2]):
        o = oracle.OracleWorld(p)
        o.upload(bb, jj)
        o.step(60)
        assert not state_mismatch(st, o.download())


def test_full_size_pile_digest_matches_oracle():
    """C2 at its full size (8192 boxes), 40 steps."""
    b, j = scenes.pile(8192, seed=2)
    w = gpu_world(b, j, P(iterations=10))
    o = oracle.OracleWorld(P(iterations=10))
    o.upload(b, j)
    w.step(40)
    o.step(40)
    assert pkg.state_digest(w.download()) == o.digest()


def test_statistics_match():
    """Penetration distribution (depth = max(0, 1 mm - C0.n) after rebase) and energy
    agree between the HIP path and the oracle after a long run."""
    b, j = scenes.pyramid(rows=14, seed=9)
    w, o = lockstep(b, j, P(iterations=10), 180, every=60)
    gm, om = w.manifolds(), o.manifolds()

    def depths(m):
        n = m["num_contacts"]
        c0 = m["c0"].reshape(len(n), 8, 3)[:, :, 0]
        mask = np.arange(8)[None, :] < n[:, None]
        return np.maximum(0.0, 0.001 - c0[mask])
    dg, do = depths(gm), depths(om)
    assert np.array_equal(np.percentile(dg, [50, 99]), np.percentile(do, [50, 99]))


def test_colours_valid_deterministic_and_equal_to_oracle():
    """Colourings: the Jones-Plassmann result equals the oracle's sequential greedy colouring
    and the greedy restatement, no joint or manifold joins two bodies of one colour, and a
    second run gives the same colours."""
    bodies, joints = scenes.chains(n_chains=32, links=16)
    params = P(iterations=10)
    params.beta_lin = 1.0e5
    params.penalty_max = 1.0e9
    runs = []
    for rep in range(2):
        w = gpu_world(bodies, joints, params)
        o = oracle.OracleWorld(params) if rep == 0 else None
        if o is not None:
            o.upload(bodies, joints)
        seq = []
        for s in range(120):
            w.step()
            if o is not None:
                o.step()
            if (s + 1) % 20 == 0:
                col = w.colours()
                seq.append(col)
                if o is not None:
                    assert np.array_equal(col, o.colours()), f"step {s}"
                    m = w.manifolds()
                    assert np.array_equal(col, scenes.greedy_colours({**bodies, **w.download()}, joints, m))
                    pairs = list(zip(joints["body_a"], joints["body_b"])) + list(zip(m["body_a"], m["body_b"]))
                    for a, b in pairs:
                        if a >= 0 and col[a] < 8 and col[b] < 8:
                            assert col[a] != col[b]
        runs.append(seq)
    assert all(np.array_equal(x, y) for x, y in zip(*runs))


@pytest.mark.parametrize("scene", ["pyramid", "chains"])
def test_long_run_statistics_match_oracle(scene):
    """Long runs (C1 pyramid, 300 steps; reduced C3 chains, 120 steps): penetration-depth
    distribution, the last dual's residuals |C_n| over contacts and |C_lin| over joints, and the
    energy drift over the run equal the oracle's (tolerance stated: 1e-6 relative; the states
    are bit-identical, so they are equal)."""
    if scene == "pyramid":
        bodies, joints = scenes.pyramid()
        params, steps = P(iterations=10), 300
    else:
        bodies, joints = scenes.chains(16, links=32, grid=(4, 4), pitch_x=4.0)
        params, steps = P(iterations=10, beta_lin=1e5, penalty_max=1e9), 120
    w = gpu_world(bodies, joints, params)
    w.record_residuals(True)
    o = oracle.OracleWorld(params)
    o.upload(bodies, joints)
    for _ in range(steps):
        w.step()
        o.step()
    gs = scenes.world_statistics({**bodies, **w.download()}, w.manifolds(), w.residuals(), bodies)
    os_ = scenes.world_statistics({**bodies, **o.download()}, o.manifolds(), o.residuals(), bodies)
    assert gs["contacts"] == os_["contacts"] and gs["contacts"] > 0 and gs["residual_source"] == "dual"
    if joints is not None:
        assert os_["joint_residual_max"] > 0
    for k, v in os_.items():
        if isinstance(v, float):
            assert abs(gs[k] - v) <= 1e-6 * max(1.0, abs(v)), (k, gs[k], v)
    gn, gl = w.residuals()
    on, ol = o.residuals()
    assert bits_equal(gn, on) and bits_equal(gl, ol)


def test_autotune_keeps_results():
    """The autotuner's trial steps are ordinary steps under other launch shapes: a world tuned
    mid-run ends bit-identical to one that is not, and to the oracle."""
    bodies, joints = scenes.pile(n=400, seed=21, ground=80.0)
    params = P(iterations=10)
    a = gpu_world(bodies, joints, params)
    b = gpu_world(bodies, joints, params)
    o = oracle.OracleWorld(params)
    o.upload(bodies, joints)
    for _ in range(10):
        a.step()
    tuning = a.autotune(1)           # 8 trial steps: 3 block sizes, 5 live grids (no colour passes here)
    assert all(v > 0 for v in tuning.values())
    for _ in range(10):
        a.step()
    for _ in range(28):
        b.step()
        o.step()
    assert not state_mismatch(a.download(), b.download())
    assert not state_mismatch(a.download(), o.download())


def _sleep_equal(g, o):
    return np.array_equal(g["sleeping"], o["sleeping"]) and np.array_equal(g["still_frames"], o["still_frames"])


def test_sleeping_lockstep():
    """Island sleep and wake (idle detection, wake on first bounds contact, wake by a moving
    neighbour, START_ASLEEP bodies, bodies uploaded asleep), bit for bit with the oracle,
    the sleep state included."""
    b, j = scenes.sleepy_pile()
    p = P(iterations=10, sleep_frames=10)
    w = gpu_world(b, j, p)
    o = oracle.OracleWorld(p)
    o.upload(b, j)
    slept = 0
    for s in range(300):
        w.step()
        o.step()
        if (s + 1) % 20 == 0:
            g, r = w.download(), o.download()
            assert not state_mismatch(g, r), s
            assert _sleep_equal(g, r), s
            slept = max(slept, int(r["sleeping"].sum()))
    manifolds_equal(w, o)
    assert slept >= 10


def test_sleeping_interleaved_and_batch():
    p = P(iterations=10, sleep_frames=8)
    p.dt = dt120()
    worlds = [scenes.sleepy_pile(120, seed=7), scenes.pile(60, seed=8), scenes.sleepy_pile(90, seed=9, modes=(1,))]
    batch = pkg.Batch(p, 4, 2000)
    ids = [batch.add(b, j) for b, j in worlds]
    alone = []
    for b, j in worlds:
        o = oracle.OracleWorld(p)
        o.upload(b, j)
        alone.append(o)
    for s in range(200):
        batch.step()
        for o in alone:
            o.step()
        if (s + 1) % 25 == 0:
            for wid, o in zip(ids, alone):
                assert not state_mismatch(batch.download(wid), o.download()), (s, wid)
    w = gpu_world(*worlds[0], p)
    o = oracle.OracleWorld(p)
    o.upload(*worlds[0])
    for s in range(100):
        w.step_interleaved(2, 0)
        o.step_interleaved(2, 0)
    g, r = w.download(), o.download()
    assert not state_mismatch(g, r) and _sleep_equal(g, r)


def test_contact_grace_lockstep():
    """Contact grace (manifolds parked for up to contact_grace_ticks after their pair separates,
    revived by feature key, fresh contacts seeded at 5 % of beta_lin) on a bouncing mixed pile,
    bit for bit with the oracle, plain and in a batch with sleeping on."""
    b, j = scenes.round_pile(150, seed=19)
    b["velocity"][:, 2] += np.where(np.arange(len(b["mass"])) % 4 == 0, 3.0, 0.0).astype(np.float32)
    p = P(iterations=10, contact_grace_ticks=3)
    w, o = lockstep(b, j, p, 150, every=10)
    manifolds_equal(w, o)
    assert w.stats()["manifolds"] == o.stats()["manifolds"]
    p2 = P(iterations=8, contact_grace_ticks=2, sleep_frames=10)
    worlds = [scenes.sleepy_pile(100, seed=21), scenes.pile(80, seed=22)]
    batch = pkg.Batch(p2, 4, 1000)
    ids = [batch.add(bb, jj) for bb, jj in worlds]
    alone = []
    for bb, jj in worlds:
        ox = oracle.OracleWorld(p2)
        ox.upload(bb, jj)
        alone.append(ox)
    for s in range(120):
        batch.step()
        for ox in alone:
            ox.step()
    for wid, ox in zip(ids, alone):
        assert not state_mismatch(batch.download(wid), ox.download()), wid

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


def test_unsupported_parameters_rejected():
    with pytest.raises(ValueError):
        oracle.OracleWorld(P(sleep_frames=30))


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

"""GPU: the broad phase's candidate lists against the oracle's pair search.  Each body lists the
bodies whose boxes, grown by a margin, meet its own, and walks the tree again only when its box
leaves its grown box (a moved body adds itself to the lists of the lower bodies it now meets);
a world where most bodies move walks the exact boxes until it calms down.  None of this may
change the pairs: bodies, every manifold field and the pair order stay bit-equal to the oracle's
every step, through teleports next to lower and higher bodies, lists that overflow, worlds
that go busy and calm down again, and a wide body (a ground spanning its world) whose pairs
are written by the bodies that meet it instead of by its own walk."""
import numpy as np
import pytest

from helpers import bits_equal, oracle, pkg, scenes, state_mismatch

pytestmark = pytest.mark.gpu
P = pkg.default_params
F = np.float32
MANIFOLD_KEYS = ("body_a", "body_b", "basis", "friction", "num_contacts", "feature", "r_a", "r_b", "c0", "penalty",
                 "lambda_", "stick")
DYN = ("position", "rotation", "velocity", "angular_velocity", "prev_velocity")


def _check(w, o, s):
    bad = state_mismatch(w.download(), o.download())
    assert not bad, f"step {s}: {bad}"
    gm, om = w.manifolds(), o.manifolds()
    assert len(gm["body_a"]) == len(om["body_a"]), f"step {s}"
    for k in MANIFOLD_KEYS:
        assert bits_equal(gm[k], om[k]), f"step {s}: manifold {k}"


def test_lists_follow_teleports_to_lower_and_higher_bodies():
    """A settled pile; then each step a few bodies jump onto other bodies (of lower and of higher
    list index), so their new neighbours' lists must gain them, and later jump away again."""
    b, j = scenes.pile(400, seed=21, pitch=2.0, ground=60.0)
    p = P(iterations=10)
    w = pkg.World(p, len(b["mass"]), 0)
    w.upload(b, j)
    o = oracle.OracleWorld(p)
    o.upload(b, j)
    for s in range(90):
        w.step()
        o.step()
    _check(w, o, 90)
    rng = np.random.default_rng(8)
    n = len(b["mass"]) - 1   # the ground is last
    listed = 0
    for s in range(60):
        st = w.download()
        arrays = {k: st[k].copy() for k in DYN}
        for _ in range(3):
            a, t = rng.integers(0, n, 2)
            arrays["position"][a] = arrays["position"][t] + np.array([0, 0, 2.2], F)
            arrays["velocity"][a] = 0
            arrays["angular_velocity"][a] = 0
            arrays["prev_velocity"][a] = 0
        w.write_bodies(arrays)
        o.write_bodies(arrays)
        w.step()
        o.step()
        _check(w, o, 91 + s)
        st = w.list_stats()
        listed += st["moved_bodies"] > 0 and st["busy_worlds"] == 0 and st["exact_walk_worlds"] == 0
    assert listed >= 10          # steps where moved bodies re-listed and appended
    w.close()


def test_overflowing_list_of_a_low_body():
    """A platform with list index 0 under 60 small boxes: its list holds more than 16
    candidates, so it walks the tree every step; a free-falling swarm meanwhile makes the world
    busy, then calm again."""
    s = scenes.SceneBuilder()
    rng = np.random.default_rng(3)
    for k in range(60):
        s.box((0.4, 0.4, 0.4), 1.0, 0.5, (-3.0 + (k % 10) * 0.6, -1.5 + (k // 10) * 0.6, 1.21 + rng.uniform(0, 0.01)))
    for k in range(70):
        s.box((0.5, 0.5, 0.5), 1.0, 0.5, (-4.5 + (k % 10) * 1.0, -3.0 + (k // 10) * 1.0, 8.0 + rng.uniform(0, 0.1)))
    s.box((40, 40, 1), 0.0, 0.5, (0, 0, -2.0))
    s.box((8, 5, 1), 0.0, 0.5, (0, 0, 0.5))      # created last: list index 0
    b, j = s.build()
    p = P(iterations=10)
    w = pkg.World(p, len(b["mass"]), 0)
    w.upload(b, j)
    o = oracle.OracleWorld(p)
    o.upload(b, j)
    seen = set()
    for step in range(220):
        w.step()
        o.step()
        _check(w, o, step)
        st = w.list_stats()
        seen.add((st["busy_worlds"], st["exact_walk_worlds"]))
    assert any(busy for busy, _ in seen)          # busy while the swarm falls
    assert (0, 0) in seen                         # lists again once it has landed
    w.close()


def test_batch_lists_with_busy_and_calm_worlds():
    """Worlds that settle next to worlds that keep falling (a tall column) in one batch: each
    equals its own oracle world every step."""
    p = P(iterations=8)
    worlds = [scenes.pile(120, seed=50, ground=40.0), scenes.pile(60, seed=51, pitch=6.0, ground=40.0),
              scenes.pile(200, seed=52, ground=40.0)]
    batch = pkg.Batch(p, 4, sum(len(x[0]["mass"]) for x in worlds) + 8, 0)
    ids = [batch.add(bb, jj) for bb, jj in worlds]
    orc = []
    for bb, jj in worlds:
        o = oracle.OracleWorld(p)
        o.upload(bb, jj)
        orc.append(o)
    for s in range(120):
        batch.step()
        for o in orc:
            o.step()
        if s % 10 == 9:
            for wid, o in zip(ids, orc):
                assert not state_mismatch(batch.download(wid), o.download()), (s, wid)
    batch.close()


def _ground_first(bodies):
    """The pile's arrays with the ground (last) moved to list index 0."""
    n = len(bodies["mass"])
    order = np.r_[n - 1, np.arange(n - 1)]
    return {k: v[order].copy() for k, v in bodies.items()}


def test_wide_ground_with_list_index_zero():
    """A ground box at list index 0 under a pile: the ground spans its world, so it does not walk
    the tree; the bodies that meet it write themselves into its pair range, ranked by index.
    Its pairs (far more than 16) and everything downstream stay bit-equal to the oracle while the
    pile falls (exact walks) and after it settles (lists), in a plain world and in a batch."""
    b, _ = scenes.pile(300, seed=23, pitch=2.0, ground=60.0)
    b = _ground_first(b)
    p = P(iterations=10)
    w = pkg.World(p, len(b["mass"]), 0)
    w.upload(b, None)
    o = oracle.OracleWorld(p)
    o.upload(b, None)
    for s in range(150):
        w.step()
        o.step()
        _check(w, o, s)
    on_ground = int(np.sum(w.manifolds()["body_a"] == 0))
    assert on_ground > 16, on_ground
    w.close()

    worlds = [_ground_first(scenes.pile(n, seed=60 + n, pitch=2.0, ground=40.0)[0]) for n in (40, 150, 90)]
    batch = pkg.Batch(p, 4, sum(len(x["mass"]) for x in worlds) + 8, 0)
    ids = [batch.add(bb, None) for bb in worlds]
    orc = []
    for bb in worlds:
        o = oracle.OracleWorld(p)
        o.upload(bb, None)
        orc.append(o)
    for s in range(100):
        batch.step()
        for o in orc:
            o.step()
        if s % 10 == 9:
            for wid, o in zip(ids, orc):
                assert not state_mismatch(batch.download(wid), o.download()), (s, wid)
    batch.close()

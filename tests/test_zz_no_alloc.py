"""avbd_world_raycast_noalloc (fixed device scratch, no device allocation) against
avbd_world_raycast, bit for bit, on two calls in a row (the second smaller), with masks and
hits from inside, on a settled pile of boxes, spheres and cylinders."""
import numpy as np
import pytest

from helpers import bits_equal, pkg, scenes

pytestmark = pytest.mark.gpu


def _rays(rng, pos, n):
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    frm = rng.uniform(lo - 1.0, hi + 1.0, (n, 3)).astype(np.float32)
    to = (frm + rng.normal(0, 5, (n, 3))).astype(np.float32)
    to[: n // 3, 2] = lo[2] - 2.0
    masks = np.where(rng.random(n) < 0.2, 2, 0xFFFFFFFF).astype(np.uint32)
    return frm, to, masks


def test_raycast_noalloc_equals_raycast():
    b, j = scenes.round_pile(600, seed=31, pitch=2.4)
    n = len(b["mass"])
    b["layer"] = np.where(np.random.default_rng(3).random(n) < 0.5, 1, 2).astype(np.uint32)
    w = pkg.World(pkg.default_params(iterations=10), n, 0)
    w.upload(b, j)
    w.step(40)
    pos = w.download()["position"]
    rng = np.random.default_rng(8)
    for count, inside in ((12000, True), (3000, False)):
        frm, to, masks = _rays(rng, pos, count)
        want = w.raycast(frm, to, masks, hit_from_inside=inside)
        got = w.raycast(frm, to, masks, hit_from_inside=inside, no_alloc=True)
        assert (want["body"] >= 0).sum() > count // 10
        for k in ("body", "t", "point", "normal", "inside"):
            assert bits_equal(got[k], want[k]), (count, k)
    w.close()


def test_raycast_noalloc_rejects_sizes_past_its_limits():
    b, j = scenes.pile(50, seed=2, ground=40.0)
    w = pkg.World(pkg.default_params(), len(b["mass"]), 0)
    w.upload(b, j)
    n = 16385
    frm = np.zeros((n, 3), np.float32)
    with pytest.raises(pkg.abi.AvbdError) as e:
        w.raycast(frm, frm + 1, no_alloc=True)
    assert e.value.code == pkg.abi.AVBD_ERR_CAPACITY
    w.close()

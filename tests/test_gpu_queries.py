"""GPU: space queries (ray casts, picks, AABB collection) on the device state against the
oracle, which is pinned to the reference's own query_shape.cpp.  Bar: bit-exact."""
import numpy as np
import pytest

from helpers import bits_equal, oracle, pkg, scenes

pytestmark = pytest.mark.gpu
P = pkg.default_params


def _worlds(steps=60):
    b, j = scenes.round_pile(200, seed=17)
    n = len(b["mass"])
    b["layer"] = np.where(np.arange(n) % 5 == 0, 2, 1).astype(np.uint32)
    b["mask"] = np.full(n, 3, np.uint32)
    w = pkg.World(P(iterations=10), n, 0)
    w.upload(b, j)
    o = oracle.OracleWorld(P(iterations=10))
    o.upload(b, j)
    w.step(steps)
    o.step(steps)
    return w, o, b


def _rays(rng, n, inside_from=None):
    frm = rng.uniform(-12, 12, (n, 3)).astype(np.float32)
    frm[:, 2] = rng.uniform(0.2, 14, n)
    to = (frm + rng.normal(0, 8, (n, 3))).astype(np.float32)
    to[: n // 3, 2] = -2.0                       # a third straight-ish down onto the pile
    if inside_from is not None:
        frm[-len(inside_from):] = inside_from    # some start at body centres
    return frm, to


def test_raycast_matches_oracle():
    w, o, b = _worlds()
    rng = np.random.default_rng(5)
    centres = w.download()["position"][:40]
    frm, to = _rays(rng, 3000, centres)
    masks = rng.choice(np.array([1, 2, 3, 0xFFFFFFFF], np.uint32), 3000)
    for inside in (False, True):
        g = w.raycast(frm, to, masks, hit_from_inside=inside)
        r = o.raycast(frm, to, masks, hit_from_inside=inside)
        assert np.array_equal(g["body"], r["body"])
        for k in ("t", "point", "normal"):
            assert bits_equal(g[k], r[k]), k
        assert np.array_equal(g["inside"], r["inside"])
        assert (g["body"] >= 0).sum() > 1000
    assert r["inside"].sum() >= 30


def test_pick_and_collect_match_oracle():
    w, o, b = _worlds()
    rng = np.random.default_rng(6)
    org = rng.uniform(-10, 10, (2000, 3)).astype(np.float32)
    org[:, 2] = rng.uniform(1, 12, 2000)
    d = rng.normal(0, 1, (2000, 3)).astype(np.float32)
    d[:1000] /= np.linalg.norm(d[:1000], axis=1, keepdims=True)
    d[:500] = (0, 0, -1)
    masks = rng.choice(np.array([1, 2, 0xFFFFFFFF], np.uint32), 2000)
    gb, gl = w.pick(org, d, masks)
    ob, ol = o.pick(org, d, masks)
    assert np.array_equal(gb, ob) and bits_equal(gl, ol)
    assert (gb >= 0).sum() > 500
    for _ in range(40):
        c = rng.uniform(-8, 8, 3).astype(np.float32)
        h = rng.uniform(0.3, 4, 3).astype(np.float32)
        mask = int(rng.choice([1, 2, 0xFFFFFFFF]))
        assert np.array_equal(w.collect_aabb(c - h, c + h, mask), o.collect_aabb(c - h, c + h, mask))


def test_batch_raycast_equals_plain():
    b, j = scenes.round_pile(120, seed=23)
    p = P(iterations=10)
    batch = pkg.Batch(p, 4, 2000)
    other = batch.add(*scenes.pile(200, seed=3))
    wid = batch.add(b, j)
    w = pkg.World(p, len(b["mass"]), 0)
    w.upload(b, j)
    for _ in range(40):
        batch.step()
        w.step()
    rng = np.random.default_rng(7)
    frm, to = _rays(rng, 1000)
    g, r = batch.raycast(wid, frm, to), w.raycast(frm, to)
    assert np.array_equal(g["body"], r["body"]) and bits_equal(g["t"], r["t"]) and bits_equal(g["normal"], r["normal"])
    assert other != wid


def test_contact_report_matches_oracle():
    w, o, b = _worlds(90)
    n = len(b["mass"])
    rng = np.random.default_rng(9)
    maxrep = rng.choice(np.array([0, 0, 1, 3, 8]), n).astype(np.int32)
    maxrep[-1] = 16                      # the ground, touched by many manifolds
    g = w.contact_report(maxrep)
    r = o.contact_report(maxrep)
    assert np.array_equal(g[0], r[0]) and g[0].sum() > 100
    assert bits_equal(g[1], r[1]) and bits_equal(g[2], r[2]) and np.array_equal(g[3], r[3])
    w.step(3)                            # reports leave the simulation alone
    o.step(3)
    from helpers import state_mismatch
    assert not state_mismatch(w.download(), o.download())


def test_intersect_point_matches_oracle():
    w, o, b = _worlds()
    rng = np.random.default_rng(10)
    centres = w.download()["position"]
    pts = np.concatenate([centres + rng.normal(0, 0.3, centres.shape), rng.uniform(-10, 10, (700, 3))]).astype(np.float32)
    masks = rng.choice(np.array([1, 2, 0xFFFFFFFF], np.uint32), len(pts))
    for k in (1, 4):
        gc, go = w.intersect_point(pts, masks, max_results=k)
        oc, oo = o.intersect_point(pts, masks, max_results=k)
        assert np.array_equal(gc, oc) and np.array_equal(go, oo)
    assert gc.sum() > 80


def test_intersect_shape_matches_oracle():
    """The server's intersect_shape on the device: random boxes, spheres and cylinders (random
    rotations, margins 0 or 0.05, layer masks) over the settled round pile, two result caps;
    a second call with fewer queries reuses the grown buffers."""
    w, o, b = _worlds()
    rng = np.random.default_rng(12)
    centres = w.download()["position"]
    n = 900
    t = rng.integers(0, 3, n).astype(np.int32)
    c = np.concatenate([centres[rng.integers(0, len(centres), n - 200)] + rng.normal(0, 0.5, (n - 200, 3)),
                        rng.uniform(-10, 10, (200, 3))]).astype(np.float32)
    qr = rng.normal(size=(n, 4)).astype(np.float32)
    qr /= np.linalg.norm(qr, axis=1, keepdims=True)
    s = rng.uniform(0.2, 1.5, (n, 3)).astype(np.float32)
    m = np.where(rng.random(n) < 0.5, 0.0, 0.05).astype(np.float32)
    masks = rng.choice(np.array([1, 2, 3, 0xFFFFFFFF], np.uint32), n)
    for k in (2, 16):
        gc, go = w.intersect_shape(t, c, qr, s, m, masks, max_results=k)
        oc, oo = o.intersect_shape(t, c, qr, s, m, masks, max_results=k)
        assert np.array_equal(gc, oc) and np.array_equal(go, oo)
    assert gc.sum() > 500 and (gc >= 3).sum() > 20
    gc, go = w.intersect_shape(t[:50], c[:50], qr[:50], s[:50], m[:50], None, max_results=4)
    oc, oo = o.intersect_shape(t[:50], c[:50], qr[:50], s[:50], m[:50], None, max_results=4)
    assert np.array_equal(gc, oc) and np.array_equal(go, oo)
    with pytest.raises(Exception):
        w.intersect_shape([7], c[:1], qr[:1], s[:1], m[:1])


def test_cast_motion_matches_oracle():
    """The server's cast_motion on the device: random boxes, spheres and cylinders swept by
    random motions (some zero, some straight down onto the pile) with margins 0 or 0.05 and
    layer masks over the settled round pile: body, safe, unsafe, point and normal bit-exact."""
    w, o, b = _worlds()
    rng = np.random.default_rng(21)
    centres = w.download()["position"]
    n = 400
    t = rng.integers(0, 3, n).astype(np.int32)
    c = (centres[rng.integers(0, len(centres), n)] + rng.normal(0, 1.5, (n, 3))).astype(np.float32)
    c[: n // 2, 2] += 4.0
    qr = rng.normal(size=(n, 4)).astype(np.float32)
    qr /= np.linalg.norm(qr, axis=1, keepdims=True)
    s = rng.uniform(0.2, 1.0, (n, 3)).astype(np.float32)
    mv = rng.normal(0, 2.0, (n, 3)).astype(np.float32)
    mv[: n // 2] = (0, 0, -8)
    mv[::10] = 0
    m = np.where(rng.random(n) < 0.5, 0.0, 0.05).astype(np.float32)
    masks = rng.choice(np.array([1, 2, 3, 0xFFFFFFFF], np.uint32), n)
    g = w.cast_motion(t, c, qr, s, mv, m, masks)
    r = o.cast_motion(t, c, qr, s, mv, m, masks)
    assert np.array_equal(g["body"], r["body"])
    for k in ("safe", "unsafe", "point", "normal"):
        assert bits_equal(g[k], r[k]), k
    assert (r["body"] >= 0).sum() > 150 and ((r["safe"] > 0) & (r["safe"] < 1)).sum() > 50


def test_collide_pair_matches_oracle():
    """avbd_world_collide_pair (avbd::Manifold::collide) against the oracle over touching,
    near and far pairs of a moving mixed pile (boxes, spheres, cylinders), both orders, at two
    collision horizons: counts, feature keys, arms and basis bit-exact."""
    b, j = scenes.round_pile(300, seed=13, pitch=2.2)
    n = len(b["mass"])
    p = pkg.default_params(iterations=10)
    w = pkg.World(p, n, 0)
    w.upload(b, j)
    w.step(45)
    o = oracle.OracleWorld(p)
    o.upload({**b, **w.download()}, j)
    rng = np.random.default_rng(4)
    m = w.manifolds()
    pairs = list(zip(m["body_a"], m["body_b"]))[:150]
    pairs += [tuple(x) for x in rng.choice(n, (150, 2))]
    hits = 0
    for a, c in pairs:
        if a == c:
            continue
        for x, y in ((a, c), (c, a)):
            for hz in (p.dt, 2 * p.dt):
                g = w.collide_pair(x, y, hz)
                r = o.collide_pair(x, y, hz)
                assert g[0] == r[0], (x, y)
                if g[0]:
                    hits += 1
                    for u, v in zip(g[1:], r[1:]):
                        assert np.array_equal(np.asarray(u).view(np.uint32), np.asarray(v).view(np.uint32)), (x, y)
    assert hits > 300
    w.close()


def _settled_round_pile(n=400, steps=60):
    b, j = scenes.round_pile(n, seed=23, pitch=2.2)
    p = pkg.default_params(iterations=10)
    w = pkg.World(p, len(b["mass"]), 0)
    w.upload(b, j)
    w.step(steps)
    o = oracle.OracleWorld(p)
    o.upload({**b, **w.download()}, j)
    return b, w, o


def _random_shapes(rng, pos, n):
    types = rng.integers(0, 3, n).astype(np.int32)
    centers = (pos[rng.integers(0, len(pos), n)] + rng.normal(0, 0.7, (n, 3))).astype(np.float32)
    rots = rng.normal(size=(n, 4)).astype(np.float32)
    rots /= np.linalg.norm(rots, axis=1, keepdims=True)
    sizes = rng.uniform(0.3, 1.6, (n, 3)).astype(np.float32)
    margins = rng.choice([0.0, 0.0, 0.03], n).astype(np.float32)
    masks = np.where(rng.random(n) < 0.15, 2, 0xFFFFFFFF).astype(np.uint32)
    return types, centers, rots, sizes, margins, masks


def test_collide_shape_and_rest_info_match_oracle():
    """The server's collide_shape and rest_info on the device against the oracle over 600 random
    query shapes on a settled mixed pile: pair counts, the any flags, every point, the deepest
    body, its point, normal and depth bit for bit."""
    b, w, o = _settled_round_pile()
    rng = np.random.default_rng(12)
    args = _random_shapes(rng, w.download()["position"], 600)
    for max_pairs in (3, 32):
        g = w.collide_shape(*args, max_pairs=max_pairs)
        r = o.collide_shape(*args, max_pairs=max_pairs)
        assert bits_equal(g[0], r[0]) and bits_equal(g[1], r[1])
        for q in range(len(args[0])):
            assert bits_equal(g[2][q, :g[0][q]], r[2][q, :r[0][q]]), q
        assert int((g[0] > 0).sum()) > 150
    g = w.rest_info(*args)
    r = o.rest_info(*args)
    assert bits_equal(g[0], r[0])
    hit = g[0] >= 0
    assert int(hit.sum()) > 150
    for x, y in zip(g[1:], r[1:]):
        assert bits_equal(x[hit], y[hit])
    w.close()


def test_closest_point_matches_oracle():
    b, w, o = _settled_round_pile()
    rng = np.random.default_rng(5)
    n = len(b["mass"])
    idx = rng.integers(0, n, 4000).astype(np.int32)
    pos = w.download()["position"]
    pts = (pos[idx] + rng.normal(0, 1.0, (4000, 3)) * rng.choice([0.05, 1.0, 3.0], (4000, 1))).astype(np.float32)
    assert bits_equal(w.closest_point(idx, pts), o.closest_point(idx, pts))
    w.close()


def _all_queries(q, rng_seed, pos, wide):
    """Every space query of q (a World or the oracle) on one seeded set of inputs; `wide` adds
    query boxes that span most of the pile (hundreds of candidates each)."""
    rng = np.random.default_rng(rng_seed)
    frm, to = _rays(rng, 1500, pos[:30])
    out = {"ray": q.raycast(frm, to, hit_from_inside=True), "ray_out": q.raycast(frm, to)}
    pts = (pos[rng.integers(0, len(pos), 600)] + rng.normal(0, 0.4, (600, 3))).astype(np.float32)
    out["point"] = q.intersect_point(pts, max_results=3)
    t, c, r, s, m, masks = _random_shapes(rng, pos, 300)
    if wide:
        s[:40] = rng.uniform(8.0, 20.0, (40, 3))
        c[:40] = pos.mean(axis=0)
    out["shape"] = q.intersect_shape(t, c, r, s, m, masks, max_results=80 if wide else 6)
    out["collide"] = q.collide_shape(t, c, r, s, m, masks, max_pairs=120 if wide else 8)
    out["rest"] = q.rest_info(t, c, r, s, m, masks)
    mv = rng.normal(0, 3.0, (300, 3)).astype(np.float32)
    out["cast"] = q.cast_motion(t, c, r, s, mv, m, masks)
    return out


def _same(key, a, b):
    """Bit equality of one query's results; collide_shape points past each count and rest_info
    fields of queries without a hit are unspecified."""
    if key == "collide":
        return (bits_equal(a[0], b[0]) and bits_equal(a[1], b[1])
                and all(bits_equal(a[2][q, :a[0][q]], b[2][q, :b[0][q]]) for q in range(len(a[0]))))
    if key == "rest":
        hit = a[0] >= 0
        return bits_equal(a[0], b[0]) and all(bits_equal(x[hit], y[hit]) for x, y in zip(a[1:], b[1:]))
    if isinstance(a, dict):
        return all(bits_equal(a[k], b[k]) for k in a)
    return all(bits_equal(x, y) for x, y in zip(a, b))


def test_query_tree_equals_list_scan_and_oracle():
    """Queries walk the step's LBVH refit to the current bounds: the same bits as the list scan
    and as the oracle, also after teleports that leave the tree's topology far from the bodies,
    for query boxes with hundreds of candidates (several walks of 32), and on a fresh upload of
    the same state (no tree yet: the scan)."""
    b, w, o = _settled_round_pile(300, 45)
    n = len(b["mass"])
    st = w.download()
    moved = {k: st[k].copy() for k in ("position", "rotation", "velocity", "angular_velocity", "prev_velocity")}
    moved["position"][::7] += np.float32(6.0)          # every 7th body jumps: bounds move, topology stays
    moved["position"][3::11, 2] += np.float32(9.0)
    w.write_bodies(moved)
    o.write_bodies(moved)
    pos = moved["position"]
    for wide in (False, True):
        tree = _all_queries(w, 40 + wide, pos, wide)
        w.use_query_tree(False)
        scan = _all_queries(w, 40 + wide, pos, wide)
        w.use_query_tree(True)
        ref = _all_queries(o, 40 + wide, pos, wide)
        for k in tree:
            assert _same(k, tree[k], scan[k]), k
            assert _same(k, tree[k], ref[k]), k
    assert int((tree["shape"][0] >= 32).sum()) >= 20      # boxes that took several walks
    fresh = pkg.World(pkg.default_params(iterations=10), n, 0)
    fresh.upload({**b, **moved}, None)
    again = _all_queries(fresh, 41, pos, True)
    for k in again:
        assert _same(k, again[k], tree[k]), k
    fresh.close()
    w.close()


def test_batch_queries_through_the_tree_after_world_changes():
    """A batch world's queries walk its own subtree (the other worlds' leaves are empty); after
    another world is removed the next query scans, the step after walks again: all equal to
    the plain world's."""
    b, j = scenes.round_pile(150, seed=29)
    p = P(iterations=10)
    batch = pkg.Batch(p, 4, 3000)
    first = batch.add(*scenes.pile(300, seed=4))
    wid = batch.add(b, j)
    last = batch.add(*scenes.round_pile(90, seed=8))
    w = pkg.World(p, len(b["mass"]), 0)
    w.upload(b, j)
    for _ in range(30):
        batch.step()
        w.step()
    rng = np.random.default_rng(31)
    frm, to = _rays(rng, 1200)
    ref = w.raycast(frm, to)

    def check():
        g = batch.raycast(wid, frm, to)
        assert np.array_equal(g["body"], ref["body"]) and bits_equal(g["t"], ref["t"])
        assert bits_equal(g["normal"], ref["normal"])

    check()
    batch.remove(first)
    check()
    batch.use_query_tree(False)
    check()
    batch.use_query_tree(True)
    batch.step()
    w.step()
    ref = w.raycast(frm, to)
    check()
    assert (ref["body"] >= 0).sum() > 300 and last != wid
    batch.close()
    w.close()

"""Writes tests/golden/morton.npz (codes of the reference's own Morton encoder,
src/avbd/bvh_broadphase.cpp:19-36) and tests/golden/collide_boxes.npz: box-box
narrow-phase outputs (contact count,
basis, feature keys, local arms) produced by the reference's own src/avbd/collide.cpp,
compiled by oracle/Makefile into oracle/_ref/libavbd_ref.so.  The reference tree exists
only in the build container, so its outputs travel as this fixture.

Cases: resting face contacts (stacked cubes, cube on a slab), tilted and edge-on pairs,
random orientations and sizes near touching, separated pairs inside and outside the
margin, all with the solver's speculative margin.

tests/golden/collide_shapes.npz holds the same outputs for pairs with a sphere or a
cylinder (collide.cpp:505-1032): the reference's own scenario poses (resting on a slab, on
the cap and on the side, the 5 mm speculative gaps, a sphere centred inside a box) and
random near-touching pairs of every shape combination, at margins 0, 0.01 and 0.05.
`python make_fixtures.py shapes` writes only that file.

tests/golden/ray_shapes.npz holds query::intersectRay outputs (hit, t, point, normal, inside)
of the reference's own src/avbd/query/query_shape.cpp for rays against boxes, spheres and
cylinders: axis-aligned and oblique, from outside and inside, unnormalised directions, a zero
direction component, bounded and unbounded.  `python make_fixtures.py rays` writes only that file.

tests/golden/closest_point.npz holds query::closestPoint outputs of the same query_shape.cpp
(`python make_fixtures.py closest` writes only that file).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402


def quat_axis(axis, ang):
    axis = np.asarray(axis, np.float64)
    axis /= np.linalg.norm(axis)
    s = np.sin(ang / 2)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, np.cos(ang / 2)], np.float32)


def cases(rng, n_random=400):
    I = np.array([0, 0, 0, 1], np.float32)
    out = []
    out.append(((0, 0, 1.0), I, (1, 1, 1), (0, 0, 0), I, (100, 100, 1)))
    out.append(((0, 0, 0.999), I, (1, 1, 1), (0, 0, 0), I, (100, 100, 1)))
    out.append(((0, 0, 2.0), I, (1, 1, 1), (0, 0, 1.0), I, (1, 1, 1)))
    out.append(((0.3, 0.2, 1.995), I, (1, 1, 1), (0, 0, 1.0), I, (1, 1, 1)))
    out.append(((0, 0, 1.1), quat_axis((0, 0, 1), 0.7), (1, 1, 1), (0, 0, 0), I, (10, 10, 1)))
    out.append(((0, 0, 1.2), quat_axis((1, 0, 0), np.pi / 4), (1, 1, 1), (0, 0, 0), I, (10, 10, 1)))
    out.append(((0, 0, 1.6), quat_axis((1, 1, 0), 0.95), (1, 1, 1), (0, 0, 0), I, (10, 10, 1)))
    out.append(((0.9, 0.0, 0.0), quat_axis((0, 0, 1), np.pi / 4), (1, 1, 1), (0, 0, 0), quat_axis((1, 0, 0), np.pi / 4),
                (1, 1, 1)))
    for _ in range(n_random):
        sa = rng.uniform(0.5, 2.0, 3).astype(np.float32)
        sb = rng.uniform(0.5, 2.0, 3).astype(np.float32)
        qa = rng.normal(size=4); qa /= np.linalg.norm(qa)
        qb = rng.normal(size=4); qb /= np.linalg.norm(qb)
        d = rng.normal(size=3); d /= np.linalg.norm(d)
        dist = rng.uniform(0.3, 1.2) * (np.linalg.norm(sa) + np.linalg.norm(sb)) / 2
        pa = rng.uniform(-5, 5, 3)
        out.append((pa, qa.astype(np.float32), sa, pa + d * dist, qb.astype(np.float32), sb))
    return out


def shape_cases(rng, n_random=600):
    I = np.array([0, 0, 0, 1], np.float32)
    side = quat_axis((1, 0, 0), np.pi / 2)
    slab = (200, 200, 1)
    S, C, B = 1, 2, 0
    out = [   # (typeA, pa, qa, sa, typeB, pb, qb, sb, margin)
        (B, (0, 0, -0.5), I, slab, S, (0, 0, 0.492), I, (0.5, 0, 0), 0.01),
        (S, (0, 0, 0.49), I, (0.5, 0, 0), B, (0, 0, -0.5), I, slab, 0.01),
        (B, (0, 0, -0.5), I, slab, C, (0, 0, 0.4921), I, (0.5, 0.5, 1.0), 0.01),
        (B, (0, 0, -0.5), I, slab, C, (0, 0, 0.4591), side, (0.5, 0.5, 1.0), 0.01),
        (S, (0, 0, 1.0), I, (0.5, 0, 0), S, (0, 0, 1.9898), I, (0.5, 0, 0), 0.01),
        (B, (0, 0, 0), I, (1, 1, 1), S, (0.1, 0, 0.55), I, (0.5, 0, 0), 0.0),
        (S, (0.1, 0, 0.55), I, (0.5, 0, 0), B, (0, 0, 0), I, (1, 1, 1), 0.0),
        (B, (0, 0, 0), I, (1, 1, 1), S, (0.1, 0.05, 0.0), I, (0.2, 0, 0), 0.0),   # centre inside
        (B, (0, 0, 0), I, (2, 2, 1), C, (0, 0, 1.005), I, (0.5, 0.5, 1.0), 0.0),
        (B, (0, 0, 0), I, (2, 2, 1), C, (0, 0, 1.005), I, (0.5, 0.5, 1.0), 0.01),
        (C, (0, 0, 0), I, (0.5, 0.5, 1.0), C, (0, 0, 1.005), I, (0.5, 0.5, 1.0), 0.0),
        (C, (0, 0, 0), I, (0.5, 0.5, 1.0), C, (0, 0, 1.005), I, (0.5, 0.5, 1.0), 0.01),
        (C, (0, 0, 0), I, (0.5, 0.5, 2.0), S, (0.9, 0, 0.3), I, (0.45, 0, 0), 0.01),    # sphere on the side
        (C, (0, 0, 0), I, (0.5, 0.5, 2.0), S, (0.0, 0, 1.4), I, (0.45, 0, 0), 0.01),    # on the cap
        (S, (0.6, 0, 1.3), I, (0.3, 0, 0), C, (0, 0, 0), I, (0.5, 0.5, 2.0), 0.01),     # at the rim
        (C, (0, 0, 0), side, (0.3, 0.3, 0.16), C, (0.0, 0.1, 0.55), I, (0.3, 0.3, 0.16), 0.05),
    ]
    for _ in range(n_random):
        ta, tb = 0, 0
        while ta == 0 and tb == 0:
            ta, tb = (int(x) for x in rng.integers(0, 3, 2))
        sa = rng.uniform(0.2, 1.6, 3).astype(np.float32)
        sb = rng.uniform(0.2, 1.6, 3).astype(np.float32)
        qa = rng.normal(size=4); qa /= np.linalg.norm(qa)
        qb = rng.normal(size=4); qb /= np.linalg.norm(qb)
        d = rng.normal(size=3); d /= np.linalg.norm(d)
        dist = rng.uniform(0.3, 1.1) * (np.linalg.norm(sa) + np.linalg.norm(sb)) / 2
        pa = rng.uniform(-5, 5, 3)
        out.append((ta, pa, qa.astype(np.float32), sa, tb, pa + d * dist, qb.astype(np.float32), sb,
                    float(rng.choice([0.0, 0.01, 0.05]))))
    return out


def write_shapes():
    rng = np.random.default_rng(2468)
    cs = shape_cases(rng)
    n = len(cs)
    rec = {k: [] for k in ("ta", "pa", "qa", "sa", "tb", "pb", "qb", "sb", "margin")}
    count = np.zeros(n, np.int32)
    basis = np.zeros((n, 9), np.float32)
    keys = np.zeros((n, 8), np.int32)
    ra = np.zeros((n, 8, 3), np.float32)
    rb = np.zeros((n, 8, 3), np.float32)
    for i, case in enumerate(cs):
        for k, v in zip(rec, case):
            rec[k].append(np.asarray(v, np.int32 if k in ("ta", "tb") else np.float32))
        ta, pa, qa, sa, tb, pb, qb, sb, m = case
        c, b, kk, a, bb = oracle.ref_collide_shapes(ta, pa, qa, sa, tb, pb, qb, sb, m)
        count[i] = c
        basis[i] = b
        keys[i, :c] = kk
        ra[i, :c] = a
        rb[i, :c] = bb
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "collide_shapes.npz"),
                        count=count, basis=basis, keys=keys, ra=ra, rb=rb, **{k: np.stack(v) for k, v in rec.items()})
    print(f"shapes: {n} cases, {int((count > 0).sum())} with contacts, {int(count.sum())} contacts")


def write_rays(n=1500):
    rng = np.random.default_rng(1357)
    rec = {k: [] for k in ("type", "pos", "rot", "size", "origin", "dir", "max_t")}
    out = {k: [] for k in ("hit", "t", "point", "normal", "inside")}
    I = np.array([0, 0, 0, 1], np.float32)
    fixed = [(1, (0, 0, 0), I, (1, 0, 0), (-5, 0, 0), (1, 0, 0), 10.0), (1, (0, 0, 0), I, (1, 0, 0), (-5, 0, 0), (1, 0, 0), 3.0),
             (1, (0, 0, 0), I, (1, 0, 0), (0, 0, 0), (1, 0, 0), 10.0), (0, (0, 0, 0), I, (2, 2, 2), (-5, 0, 0), (1, 0, 0), 10.0),
             (2, (0, 0, 0), I, (1, 1, 4), (-5, 0, 0), (1, 0, 0), 10.0), (2, (0, 0, 0), I, (1, 1, 4), (0, 0, 5), (0, 0, -1), 10.0),
             (1, (0, 0, 0), I, (1, 0, 0), (0, 2, 0), (0, 0, 1), 10.0)]
    cases = list(fixed)
    for _ in range(n):
        t = int(rng.integers(0, 3))
        size = (rng.random(3) * 1.5 + 0.2).astype(np.float32)
        q = rng.standard_normal(4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        pos = rng.standard_normal(3).astype(np.float32)
        org = (rng.standard_normal(3) * (3 if rng.random() < 0.8 else 0.3)).astype(np.float32)
        d = (pos + rng.standard_normal(3).astype(np.float32) * 0.5 - org).astype(np.float32)
        if rng.random() < 0.5:
            d = (d / np.linalg.norm(d)).astype(np.float32)
        if rng.random() < 0.1:
            d[int(rng.integers(0, 3))] = 0
        cases.append((t, pos, q, size, org, d, float(rng.choice([0.0, 2.0, 10.0]))))
    for c in cases:
        for k, v in zip(rec, c):
            rec[k].append(np.asarray(v, np.int32 if k == "type" else np.float32))
        hit, tt, point, normal, inside = oracle.ref_intersect_ray(*c)
        out["hit"].append(np.uint8(hit)); out["t"].append(np.float32(tt)); out["point"].append(point)
        out["normal"].append(normal); out["inside"].append(np.uint8(inside))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ray_shapes.npz"),
                        **{k: np.stack(v) for k, v in rec.items()}, **{k: np.stack(v) for k, v in out.items()})
    print(f"rays: {len(cases)} cases, {int(np.sum(out['hit']))} hits, {int(np.sum(out['inside']))} from inside")


def write_closest(n=1200):
    """query::closestPoint of the reference's own query_shape.cpp for points against boxes,
    spheres and cylinders: outside, inside, on the axis and at the centre."""
    rng = np.random.default_rng(2468)
    rec = {k: [] for k in ("type", "pos", "rot", "size", "point")}
    out = []
    I = np.array([0, 0, 0, 1], np.float32)
    fixed = [(1, (0, 0, 0), I, (1, 0, 0), (0, 0, 0)), (2, (0, 0, 0), I, (1, 1, 4), (0, 0, 0.5)),
             (2, (0, 0, 0), I, (1, 1, 1), (0, 0, 0.2)), (2, (0, 0, 0), I, (1, 1, 4), (0.2, 0, 1.9)),
             (0, (0, 0, 0), I, (2, 2, 2), (0.5, 0.2, -0.1)), (0, (0, 0, 0), I, (2, 2, 2), (3, -4, 0.5))]
    cases = list(fixed)
    for _ in range(n):
        t = int(rng.integers(0, 3))
        size = (rng.random(3) * 2.0 + 0.2).astype(np.float32)
        q = rng.standard_normal(4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        pos = rng.standard_normal(3).astype(np.float32)
        pt = (pos + rng.standard_normal(3) * rng.choice([0.05, 0.5, 2.0])).astype(np.float32)
        cases.append((t, pos, q, size, pt))
    for c in cases:
        for k, v in zip(rec, c):
            rec[k].append(np.asarray(v, np.int32 if k == "type" else np.float32))
        out.append(oracle.ref_closest_point(*c))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "closest_point.npz"),
                        **{k: np.stack(v) for k, v in rec.items()}, closest=np.stack(out))
    print(f"closest points: {len(cases)} cases")


def main():
    oracle.build(ref=True)
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libavbd_ref.so is not built (needs the reference tree)")
    if sys.argv[1:] == ["shapes"]:
        write_shapes()
        return
    if sys.argv[1:] == ["rays"]:
        write_rays()
        return
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libavbd_ref.so is not built (needs the reference tree)")
    rng = np.random.default_rng(12345)
    cs = cases(rng)
    n = len(cs)
    rec = {k: [] for k in ("pa", "qa", "sa", "pb", "qb", "sb")}
    count = np.zeros(n, np.int32)
    basis = np.zeros((n, 9), np.float32)
    keys = np.zeros((n, 8), np.int32)
    ra = np.zeros((n, 8, 3), np.float32)
    rb = np.zeros((n, 8, 3), np.float32)
    margin = np.float32(0.01)
    for i, (pa, qa, sa, pb, qb, sb) in enumerate(cs):
        for k, v in zip(rec, (pa, qa, sa, pb, qb, sb)):
            rec[k].append(np.asarray(v, np.float32))
        c, b, kk, a, bb = oracle.ref_collide_boxes(pa, qa, sa, pb, qb, sb, margin)
        count[i] = c
        basis[i] = b
        keys[i, :c] = kk
        ra[i, :c] = a
        rb[i, :c] = bb
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "collide_boxes.npz"),
                        margin=margin, count=count, basis=basis, keys=keys, ra=ra, rb=rb,
                        **{k: np.stack(v) for k, v in rec.items()})
    print(f"{n} cases, {int((count > 0).sum())} with contacts, {int(count.sum())} contacts")
    # Morton codes of the reference's own encoder for a few centre sets (ties included)
    sets = {}
    for k, (n_c, lo, hi) in enumerate([(2, -1, 1), (64, -50, 80), (500, 0, 400), (300, -3, 3)]):
        c = rng.uniform(lo, hi, (n_c, 3)).astype(np.float32)
        c[: n_c // 4] = c[0]
        sets[f"centers{k}"] = c
        sets[f"codes{k}"] = oracle.ref_morton_codes(c)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "morton.npz"),
                        sets=np.int32(4), **sets)
    write_shapes()
    write_rays()
    write_closest()


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "closest":
        write_closest()
    else:
        main()

// queries.hpp - space queries on the device state: ray casts (the server's intersect_ray and
// Solver::pick) and AABB collection, evaluated against each body's current pose.
//
// One thread per ray walks the world's bodies in list order, so the nearest hit and its
// tie-break (the earlier body) come out as in the reference's serial loops.  Each body's bounds
// (with its layer bits) are computed once per launch; the body index is uniform across a
// wavefront, so they are read once per wavefront, and only bodies whose bounds meet a segment's
// box are read for the exact test.
#pragma once

#include "kernels.hpp"

namespace avbd_dev {

constexpr float QEPS = 1.0e-6f;
AD float qinf() { return __int_as_float(0x7f800000); }

struct QHit {
    float t; v3 point, normal; bool inside;
};

// query_shape.cpp:27-44
AD bool insideLocal(int type, const RShp &s, v3 p, float m)
{
    if (type == AVBD_SHAPE_SPHERE) return length2(p) <= (s.r + m) * (s.r + m);
    if (type == AVBD_SHAPE_CYLINDER)
        return (p.x * p.x + p.y * p.y) <= (s.r + m) * (s.r + m) && fabsf(p.z) <= s.hh + m;
    return fabsf(p.x) <= s.half.x + m && fabsf(p.y) <= s.half.y + m && fabsf(p.z) <= s.half.z + m;
}
// query_shape.cpp:46-56
AD v3 boxNormal(v3 p, v3 h)
{
    const float qx = fabsf(p.x) / (h.x > 0.0f ? h.x : 1.0f), qy = fabsf(p.y) / (h.y > 0.0f ? h.y : 1.0f),
                qz = fabsf(p.z) / (h.z > 0.0f ? h.z : 1.0f);
    if (qx >= qy && qx >= qz) return v3{p.x < 0.0f ? -1.0f : 1.0f, 0, 0};
    if (qy >= qz) return v3{0, p.y < 0.0f ? -1.0f : 1.0f, 0};
    return v3{0, 0, p.z < 0.0f ? -1.0f : 1.0f};
}
// query_shape.cpp:58-67
AD v3 cylinderNormal(v3 p, float hh)
{
    if (fabsf(fabsf(p.z) - hh) <= 1.0e-4f) return v3{0, 0, p.z < 0.0f ? -1.0f : 1.0f};
    const float radial = sqrtf(p.x * p.x + p.y * p.y);
    if (radial <= QEPS) return v3{1, 0, 0};
    return v3{p.x / radial, p.y / radial, 0};
}
// query_shape.cpp:69-96
AD bool raySphere(const RShp &s, v3 o, v3 d, float maxT, QHit &h)
{
    const float a = dot(d, d);
    if (a <= QEPS) return false;
    const float b = dot(o, d);
    const float c = dot(o, o) - s.r * s.r;
    const float disc = b * b - a * c;
    if (disc < 0.0f) return false;
    const float root = sqrtf(disc);
    float t0 = (-b - root) / a, t1 = (-b + root) / a;
    if (t0 > t1) { const float x = t0; t0 = t1; t1 = x; }
    const bool inside = c < 0.0f;
    float t = inside ? t1 : t0;
    if (t < 0.0f) t = t1;
    if (t < 0.0f || t > maxT) return false;
    const v3 n = o + d * t;
    h = QHit{t, s.c + qrot(s.q, n), qrot(s.q, normalize(n)), inside};
    return true;
}
// query_shape.cpp:98-130
AD bool rayBox(const RShp &s, v3 o, v3 d, float maxT, QHit &h)
{
    float tEnter = 0.0f, tExit = qinf();
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float di = d[i], oi = o[i], hi = s.half[i];
        if (fabsf(di) <= QEPS) {
            if (oi < -hi || oi > hi) return false;
            continue;
        }
        const float inv = 1.0f / di;
        float t0 = (-hi - oi) * inv, t1 = (hi - oi) * inv;
        if (t0 > t1) { const float x = t0; t0 = t1; t1 = x; }
        tEnter = smax(tEnter, t0);
        tExit = smin(tExit, t1);
        if (tEnter > tExit) return false;
    }
    const bool inside = insideLocal(AVBD_SHAPE_BOX, s, o, 0.0f);
    const float t = tEnter > 0.0f ? tEnter : tExit;
    if (t < 0.0f || t > maxT) return false;
    const v3 p = o + d * t;
    h = QHit{t, s.c + qrot(s.q, p), qrot(s.q, boxNormal(p, s.half)), inside};
    return true;
}
// query_shape.cpp:132-210: side hits clipped to the caps, cap hits clipped to the radius, the
// (at most four) hits sorted; the first, or the last when the origin is inside
AD bool rayCylinder(const RShp &s, v3 o, v3 d, float maxT, QHit &h)
{
    const float r = s.r, hh = s.hh;
    float hits[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int count = 0;
    auto add = [&](float t) {
        if (t >= -QEPS && t <= maxT + QEPS && count < 4) {
            const float v = smax(0.0f, t);
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (k == count) hits[k] = v;
            count++;
        }
    };
    const float a = d.x * d.x + d.y * d.y;
    if (a > QEPS) {
        const float b = o.x * d.x + o.y * d.y;
        const float c = o.x * o.x + o.y * o.y - r * r;
        const float disc = b * b - a * c;
        if (disc >= 0.0f) {
            const float root = sqrtf(disc);
            const float t0 = (-b - root) / a, t1 = (-b + root) / a;
            if (fabsf(o.z + d.z * t0) <= hh + QEPS) add(t0);
            if (fabsf(o.z + d.z * t1) <= hh + QEPS) add(t1);
        }
    } else if (o.x * o.x + o.y * o.y > r * r + QEPS) {
        return false;
    }
    if (fabsf(d.z) > QEPS) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const float z = e == 0 ? hh : -hh;
            const float t = (z - o.z) / d.z;
            const float px = o.x + d.x * t, py = o.y + d.y * t;
            if (px * px + py * py <= r * r + QEPS) add(t);
        }
    }
    if (count == 0) return false;
    // ascending insertion sort (stable; a sorting network would give the same values)
#pragma unroll
    for (int i = 1; i < 4; i++) {
        if (i >= count) break;
#pragma unroll
        for (int j = i; j > 0; j--)
            if (hits[j - 1] > hits[j]) { const float x = hits[j]; hits[j] = hits[j - 1]; hits[j - 1] = x; }
    }
    const bool inside = insideLocal(AVBD_SHAPE_CYLINDER, s, o, 0.0f);
    float t = hits[0];
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (inside && k == count - 1) t = hits[k];
    if (t < 0.0f || t > maxT) return false;
    const v3 p = o + d * t;
    h = QHit{t, s.c + qrot(s.q, p), qrot(s.q, cylinderNormal(p, hh)), inside};
    return true;
}
// query_shape.cpp:258-274
AD bool intersectRay(int type, const RShp &s, v3 origin, v3 dir, float rayMaxT, QHit &h)
{
    const float maxT = rayMaxT > 0.0f ? rayMaxT : qinf();
    const v3 o = qrot(qconj(s.q), origin - s.c);
    const v3 d = qrot(qconj(s.q), dir);
    if (type == AVBD_SHAPE_SPHERE) return raySphere(s, o, d, maxT, h);
    if (type == AVBD_SHAPE_CYLINDER) return rayCylinder(s, o, d, maxT, h);
    return rayBox(s, o, d, maxT, h);
}

// rigid.cpp:83-99: world half extents of body i at its current pose
AD v3 tightHalf(const State &S, int i, int type, float4 sr)
{
    const v3 h = type == AVBD_SHAPE_SPHERE     ? v3{sr.x, sr.x, sr.x}
                 : type == AVBD_SHAPE_CYLINDER ? v3{sr.x, sr.x, sr.z * 0.5f}
                                               : v3{sr.x, sr.y, sr.z} * 0.5f;
    const m3 R = qmat(ldq(S.rot[i]));
    return v3{(fabsf(R.c[0][0]) * h.x + fabsf(R.c[1][0]) * h.y) + fabsf(R.c[2][0]) * h.z,
              (fabsf(R.c[0][1]) * h.x + fabsf(R.c[1][1]) * h.y) + fabsf(R.c[2][1]) * h.z,
              (fabsf(R.c[0][2]) * h.x + fabsf(R.c[1][2]) * h.y) + fabsf(R.c[2][2]) * h.z};
}
AD bool boxesOverlap(v3 lo, v3 hi, v3 mn, v3 mx)
{
    return !(lo.x > mx.x || lo.y > mx.y || lo.z > mx.z || mn.x > hi.x || mn.y > hi.y || mn.z > hi.z);
}

// Each body's current world bounds (position -/+ tight half extents), computed once per query
// launch: bounds[2j] = lower corner, bounds[2j+1] = upper corner of body j of the world.
// ---- the query tree: the last step's LBVH topology refit to the current tight bounds ----
// A query walks its world's subtree instead of scanning the body list; the results stay those
// of the scan (candidates are re-tested on their exact bounds, and taken in list order).
struct QTree {
    const float4 *box;   // node boxes of the refit (nullptr: no tree, scan the list)
    int n;               // leaves (bodies) of the tree
    int w;               // the world whose subtree is walked
};

// Leaf boxes: each body's exact tight bounds (bounds[2j], [2j + 1] of world-local j) as a
// centre and a half extent grown by 4e-6 (|c| + h) + 1e-6, so that the float corners c -/+ h
// enclose them; bodies outside [base, base + count) get an empty box.
__global__ void k_qtree_leaves(State S, int base, int count, const float4 *bounds, float4 *c, float4 *h)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    const int j = i - base;
    if (S.bodyWorld[i] < 0 || j < 0 || j >= count) {
        c[i] = make_float4(0, 0, 0, 0);
        h[i] = make_float4(-1, -1, -1, 0);
        return;
    }
    const v3 lo = ld3(bounds[2 * j]), hi = ld3(bounds[2 * j + 1]);
    const v3 cc = (lo + hi) * 0.5f, hh = (hi - lo) * 0.5f;
    const v3 slack{(fabsf(cc.x) + hh.x) * 4.0e-6f + 1.0e-6f, (fabsf(cc.y) + hh.y) * 4.0e-6f + 1.0e-6f,
                   (fabsf(cc.z) + hh.z) * 4.0e-6f + 1.0e-6f};
    c[i] = st3(cc);
    h[i] = st3(hh + slack);
}

// Stackless walk of world w's subtree (ropes): visit(body) for every leaf whose box overlaps
// [mn, mx].  A walk that does not end within 2n steps, or a world without a root (a corrupt
// tree), returns false.
template <typename T, typename F>
AD bool walkTree(const State &S, const QTree &t, T &&test, F &&visit)
{
    int node = S.wRoot[t.w];
    if (node < 0) return false;
    for (int guard = 2 * t.n; node >= 0; guard--) {
        if (guard == 0) return false;
        const float4 lo = t.box[2 * node], hi = t.box[2 * node + 1];
        const bool ov = test(ld3(lo), ld3(hi));
        const int down = __float_as_int(lo.w), rope = __float_as_int(hi.w);
        if (node >= t.n - 1) {
            if (ov) visit(down);
            node = rope;
        } else {
            node = ov ? down : rope;
        }
    }
    return true;
}

template <typename F>
AD bool walkBox(const State &S, const QTree &t, v3 mn, v3 mx, F &&visit)
{
    return walkTree(S, t, [&](v3 lo, v3 hi) { return aabbOverlap(mn, mx, lo, hi); }, visit);
}

// Does the segment o + s d, s in [0, 1], meet the box [lo, hi] grown by `slack` plus 1e-5 of
// the box's largest coordinate per axis?  Monotone in the box (a box inside another passes only
// if the outer one does, in float too), so it prunes a walk.  A ray hit on a shape lies on the
// segment and inside the shape's bounds up to rounding, far below the growth.
AD bool segmentMeetsBox(v3 o, v3 d, float slack, v3 lo, v3 hi)
{
    float s0 = 0.0f, s1 = 1.0f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float g = slack + 1.0e-5f * smax(fabsf(lo[k]), fabsf(hi[k]));
        const float a = lo[k] - g, b = hi[k] + g, ok = o[k], dk = d[k];
        if (dk == 0.0f) {
            if (ok < a || ok > b) return false;
            continue;
        }
        const float inv = 1.0f / dk;
        float ta = (a - ok) * inv, tb = (b - ok) * inv;
        if (ta > tb) { const float x = ta; ta = tb; tb = x; }
        s0 = smax(s0, ta);
        s1 = smin(s1, tb);
    }
    return s0 <= s1;
}

// World-local candidates of a query box in list order: the bodies whose layer passes the mask
// and whose exact bounds overlap [mn, mx] (closed intervals); visit(j) returns false to stop.
// With a tree, each walk keeps the QCAND smallest candidates above the last one visited,
// sorted, so a box with many candidates walks once per QCAND of them; without one (or on a
// corrupt tree) the rest of the list is scanned.
constexpr int QCAND = 32;
template <typename F>
AD void forCandidates(const State &S, const QTree &t, int base, int count, const float4 *bounds, v3 mn, v3 mx,
                      uint32_t mask, F &&visit)
{
    auto passes = [&](int j) {
        const float4 lo = bounds[2 * j], hi = bounds[2 * j + 1];
        return (__float_as_uint(lo.w) & mask) != 0 && boxesOverlap(ld3(lo), ld3(hi), mn, mx);
    };
    int after = -1;
    if (t.box) {
        while (true) {
            int cand[QCAND];
            int nc = 0;
            bool over = false;
            const bool ok = walkBox(S, t, mn, mx, [&](int i) {
                const int j = i - base;
                if (j <= after || j >= count || !passes(j)) return;
                if (nc == QCAND) {
                    over = true;
                    if (j > cand[QCAND - 1]) return;
                    nc--;
                }
                int b = nc++;
                while (b > 0 && cand[b - 1] > j) { cand[b] = cand[b - 1]; b--; }
                cand[b] = j;
            });
            if (!ok) break;
            for (int k = 0; k < nc; k++)
                if (!visit(cand[k])) return;
            if (!over) return;
            after = cand[QCAND - 1];
        }
    }
    for (int j = after + 1; j < count; j++)
        if (passes(j) && !visit(j)) return;
}

__global__ void k_query_bounds(State S, int base, int count, float4 *bounds)
{
    const int j = GID;
    if (j >= count) return;
    const int i = base + j;
    const v3 p = ld3(S.pos[i]);
    const v3 wx = tightHalf(S, i, S.shape[i], S.sizeR[i]);
    bounds[2 * j] = st3(p - wx, __uint_as_float(S.layer[i]));   // w: the body's layer bits
    bounds[2 * j + 1] = st3(p + wx);
}

// The bodies part of the server's intersect_ray (avbd_direct_space_state3d.cpp:103-195): ray r
// runs from rays[2r] to rays[2r+1] (w of the first: the mask's bits); candidates are the bodies
// whose tight bounds overlap the segment's box padded by 1e-4 (Solver::collectAABB), in list
// order.  Output: body list index or -1; hitA = (point, t); hitB = (normal, inside).
__global__ void k_raycast(State S, int base, int count, int n, const float4 *rays, const float4 *bounds,
                          int hitFromInside, int *body, float4 *hitA, float4 *hitB, QTree tree)
{
    const int r = GID;
    if (r >= n) return;
    const float4 f4 = rays[2 * r], t4 = rays[2 * r + 1];
    const v3 origin = ld3(f4), end = ld3(t4);
    const uint32_t mask = __float_as_uint(f4.w);
    int best = -1;
    QHit bh{0.0f, v3{0, 0, 0}, v3{0, 0, 0}, false};
    v3 dir = end - origin;
    const float maxD = length(dir);
    if (maxD > 1.0e-6f) {
        dir = dir * (1.0f / maxD);
        const float pad = 1.0e-4f;
        const v3 mn{smin(origin.x, end.x) - pad, smin(origin.y, end.y) - pad, smin(origin.z, end.z) - pad};
        const v3 mx{smax(origin.x, end.x) + pad, smax(origin.y, end.y) + pad, smax(origin.z, end.z) + pad};
        // The list-order scan keeps the first body the segment starts inside (hit from inside),
        // else the nearest hit, the earlier body on a tie: a minimum, so the tree may visit the
        // candidates in any order.
        int firstInside = -1;
        auto consider = [&](int j) {
            const float4 lo = bounds[2 * j], hi = bounds[2 * j + 1];
            if ((__float_as_uint(lo.w) & mask) == 0) return;
            if (!boxesOverlap(ld3(lo), ld3(hi), mn, mx)) return;
            const int i = base + j;
            QHit h;
            if (!intersectRay(S.shape[i], makeRShp(ld3(S.pos[i]), ldq(S.rot[i]), S.sizeR[i]), origin, dir, maxD, h)) return;
            if (h.inside && !hitFromInside) return;
            if (h.inside) {
                if (firstInside < 0 || j < firstInside) firstInside = j;
                return;
            }
            if (h.t > maxD) return;
            if (best >= 0 && (h.t > bh.t || (h.t == bh.t && j > best))) return;
            best = j;
            bh = QHit{h.t, h.point, h.normal, false};
        };
        // The walk also skips nodes the segment itself misses (the long segments' boxes hold many
        // bodies the ray passes far from): such bodies have no hit the scan would keep.
        const v3 seg = end - origin;
        const float slack = 1.0e-4f + 1.0e-5f * (smax(smax(fabsf(origin.x), fabsf(origin.y)), fabsf(origin.z)) +
                                                  smax(smax(fabsf(seg.x), fabsf(seg.y)), fabsf(seg.z)));
        bool walked = false;
        if (tree.box)
            walked = walkTree(S, tree, [&](v3 lo, v3 hi) {
                return aabbOverlap(mn, mx, lo, hi) && segmentMeetsBox(origin, seg, slack, lo, hi);
            }, [&](int i) {
                const int j = i - base;
                if (j >= 0 && j < count) consider(j);
            });
        if (!walked) {
            best = -1;
            firstInside = -1;
            for (int j = 0; j < count; j++) {   // body bounds uniform across the wavefront
                consider(j);
                if (firstInside >= 0) break;
            }
        }
        if (firstInside >= 0) {
            best = firstInside;
            bh = QHit{0.0f, origin, dir * -1.0f, true};
        }
    }
    body[r] = best;
    hitA[r] = st3(bh.point, bh.t);
    hitB[r] = st3(bh.normal, bh.inside ? 1.0f : 0.0f);
}

// The bodies part of the server's intersect_point (avbd_direct_space_state3d.cpp:225-259):
// pts[r] = (point, mask bits); candidates are the bodies whose bounds overlap the point's box
// padded by 1e-4, in list order; out[r * maxRes + c] = the c-th body containing the point
// (query::containsPoint, query_shape.cpp:334-337), up to maxRes.
__global__ void k_intersect_point(State S, int base, int count, int n, const float4 *pts, const float4 *bounds,
                                  int maxRes, int *cnt, int *out, QTree tree)
{
    const int r = GID;
    if (r >= n) return;
    const float4 p4 = pts[r];
    const v3 p = ld3(p4);
    const uint32_t mask = __float_as_uint(p4.w);
    const float pad = 1.0e-4f;
    const v3 mn{p.x - pad, p.y - pad, p.z - pad}, mx{p.x + pad, p.y + pad, p.z + pad};
    int c = 0;
    if (maxRes > 0)
        forCandidates(S, tree, base, count, bounds, mn, mx, mask, [&](int j) {
            const int i = base + j;
            const RShp s = makeRShp(ld3(S.pos[i]), ldq(S.rot[i]), S.sizeR[i]);
            if (!insideLocal(S.shape[i], s, qrot(qconj(s.q), p - s.c), 0.0f)) return true;
            out[static_cast<size_t>(r) * maxRes + c] = j;
            c++;
            return c < maxRes;
        });
    cnt[r] = c;
}

// The bodies part of the server's intersect_shape (avbd_direct_space_state3d.cpp:261-306):
// query r is qin[4r] = (centre, type bits), qin[4r+1] = rotation, qin[4r+2] = (size, margin),
// qin[4r+3].x = mask bits.  Candidates: bodies whose bounds overlap the query's bounds
// (query_shape.cpp:371-394) grown by the margin, in list order; a hit is a body the narrow
// phase (query first) finds within the margin.  Contact rows and polygon spill rows per query
// are staging-shaped scratch (row stageRow(r, i)).
__global__ void __launch_bounds__(NARROW_TPB) k_intersect_shape(State S, int base, int count, int n, const float4 *qin,
                                                               const float4 *bounds, int maxRes, int *cnt, int *out,
                                                               float4 *cra, float4 *crb, float4 *cmid, float4 *spA,
                                                               float4 *spB, QTree tree)
{
    __shared__ NarrowPolys poly;
    const int r = GID;
    if (r >= n) return;
    const float4 c4 = qin[4 * r], r4 = qin[4 * r + 1], s4 = qin[4 * r + 2], m4 = qin[4 * r + 3];
    const int type = __float_as_int(c4.w);
    const v3 qc = ld3(c4);
    const qt qq = ldq(r4);
    const float m = smax(0.0f, s4.w);
    const uint32_t mask = __float_as_uint(m4.x);
    const v3 h = type == AVBD_SHAPE_SPHERE     ? v3{s4.x, s4.x, s4.x}
                 : type == AVBD_SHAPE_CYLINDER ? v3{s4.x, s4.x, s4.z * 0.5f}
                                               : v3{s4.x, s4.y, s4.z} * 0.5f;
    const m3 R = qmat(qq);
    const v3 wx{(fabsf(R.c[0][0]) * h.x + fabsf(R.c[1][0]) * h.y) + fabsf(R.c[2][0]) * h.z,
                (fabsf(R.c[0][1]) * h.x + fabsf(R.c[1][1]) * h.y) + fabsf(R.c[2][1]) * h.z,
                (fabsf(R.c[0][2]) * h.x + fabsf(R.c[1][2]) * h.y) + fabsf(R.c[2][2]) * h.z};
    v3 mn = qc - wx, mx = qc + wx;
    if (m > 0.0f) { mn = mn - v3{m, m, m}; mx = mx + v3{m, m, m}; }
    const size_t d = stageRow(r, 0);
    const int l = threadIdx.x;
    int c = 0;
    if (maxRes > 0) forCandidates(S, tree, base, count, bounds, mn, mx, mask, [&](int j) {
        const int i = base + j;
        const int bt = S.shape[i];
        const v3 bp = ld3(S.pos[i]);
        const qt bq = ldq(S.rot[i]);
        const float4 bs = S.sizeR[i];
        Contacts cs{0, cra + d, crb + d, cmid + d};
        m3 B;
        bool wrote;
        int hits;
        if (type != AVBD_SHAPE_BOX || bt != AVBD_SHAPE_BOX) {
            hits = collideRound(type, bt, makeRShp(qc, qq, make_float4(s4.x, s4.y, s4.z, 0.0f)), makeRShp(bp, bq, bs),
                                S.rim, m, B, wrote, cs);
        } else {
            const Poly c0{&poly[0][0][l], &poly[0][1][l], &poly[0][2][l], spA + d};
            const Poly c1{&poly[1][0][l], &poly[1][1][l], &poly[1][2][l], spB + d};
            hits = collideBoxes(qc, qq, v3{s4.x, s4.y, s4.z} * 0.5f, bp, bq, v3{bs.x, bs.y, bs.z} * 0.5f, m, B, wrote, cs,
                                c0, c1);
        }
        if (hits <= 0) return true;
        out[static_cast<size_t>(r) * maxRes + c] = j;
        c++;
        return c < maxRes;
    });
    cnt[r] = c;
}

// avbd::Manifold::collide (collide.cpp:1100-1120) for bodies a (as A) and b at their current
// pose: the speculative margin grown by the pair's predicted relative motion over hz (the same
// expression as the step's narrow phase), then the step's narrow phase.  One work item, run by
// lane 0 of one workgroup in the staging rows of query 0.  out[0] = (count, basis written),
// out[1..3] = basis rows, out[4 + 2k] = (rA, feature key bits), out[5 + 2k] = (rB, 0).
__global__ void __launch_bounds__(NARROW_TPB) k_collide_pair(State S, int a, int b, float hz, float4 *out, float4 *cra,
                                                            float4 *crb, float4 *cmid, float4 *spA, float4 *spB)
{
    __shared__ NarrowPolys poly;
    if (threadIdx.x != 0) return;
    const float4 ma4 = S.massI[a], mb4 = S.massI[b];
    const float4 sa = S.sizeR[a], sb = S.sizeR[b];
    const v3 pa = ld3(S.pos[a]), pb = ld3(S.pos[b]);
    const qt qa = ldq(S.rot[a]), qb = ldq(S.rot[b]);
    float margin = S.P.specMargin;
    if (margin > 0.0f) {
        const float gA = ma4.x > 0.0f ? S.grav[a] : 0.0f;
        const float gB = mb4.x > 0.0f ? S.grav[b] : 0.0f;
        const v3 disp = (ld3(S.vel[a]) - ld3(S.vel[b])) * hz + v3{0, 0, (gA - gB) * hz * hz};
        margin += length(disp) + (sa.w * length(ld3(S.angVel[a])) + sb.w * length(ld3(S.angVel[b]))) * hz;
    }
    const int shA = S.shape[a], shB = S.shape[b];
    Contacts cs{0, cra, crb, cmid};
    m3 B;
    bool wrote = false;
    int n;
    if (shA != AVBD_SHAPE_BOX || shB != AVBD_SHAPE_BOX) {
        n = collideRound(shA, shB, makeRShp(pa, qa, sa), makeRShp(pb, qb, sb), S.rim, margin, B, wrote, cs);
    } else {
        const Poly c0{&poly[0][0][0], &poly[0][1][0], &poly[0][2][0], spA};
        const Poly c1{&poly[1][0][0], &poly[1][1][0], &poly[1][2][0], spB};
        n = collideBoxes(pa, qa, v3{sa.x, sa.y, sa.z} * 0.5f, pb, qb, v3{sb.x, sb.y, sb.z} * 0.5f, margin, B, wrote, cs,
                         c0, c1);
    }
    n = n < 0 ? 0 : n;
    out[0] = make_float4(static_cast<float>(n), wrote ? 1.0f : 0.0f, 0.0f, 0.0f);
    out[1] = st3(row(B, 0));
    out[2] = st3(row(B, 1));
    out[3] = st3(row(B, 2));
    for (int k = 0; k < n; k++) {
        const float4 x = cra[crow(k)], y = crb[crow(k)];
        out[4 + 2 * k] = make_float4(x.x, x.y, x.z, y.w);
        out[5 + 2 * k] = make_float4(y.x, y.y, y.z, 0.0f);
    }
}

// query_shape.cpp:340-369: a shape's extent along an axis, and the overlap of two shapes along
// a contact normal (type, rotation, size as a body's: box extents, sphere (r, -, -), cylinder
// (r, -, height) about local +Z)
AD float supportRadius(int type, qt q, v3 size, v3 axis)
{
    const v3 n = normalize(axis);
    if (type == AVBD_SHAPE_SPHERE) return size.x;
    if (type == AVBD_SHAPE_CYLINDER) {
        const float axial = fabsf(dot(qrot(q, v3{0, 0, 1}), n));
        const float radial = sqrtf(smax(0.0f, 1.0f - axial * axial));
        return size.x * radial + (size.z * 0.5f) * axial;
    }
    const m3 R = qmat(q);
    const v3 h = size * 0.5f;
    return (h.x * fabsf(dot(row(R, 0), n)) + h.y * fabsf(dot(row(R, 1), n))) + h.z * fabsf(dot(row(R, 2), n));
}
AD float penetrationDepth(int ta, v3 ca, qt qa, v3 sa, int tb, v3 cb, qt qb, v3 sb, v3 normal)
{
    const float len = length(normal);
    if (len <= 1.0e-6f) return 0.0f;
    const v3 n = normal * (1.0f / len);
    return (supportRadius(ta, qa, sa, n) + supportRadius(tb, qb, sb, n)) - fabsf(dot(cb - ca, n));
}

// One query shape against one candidate body: the narrow phase, query first (its contacts in
// the query's staging rows of lane l), returning the contact count and the basis.
AD int queryContacts(const State &S, int type, v3 qc, qt qq, float4 s4, float m, int i, NarrowPolys &poly, int l,
                     size_t d, float4 *cra, float4 *crb, float4 *cmid, float4 *spA, float4 *spB, m3 &B)
{
    const int bt = S.shape[i];
    const v3 bp = ld3(S.pos[i]);
    const qt bq = ldq(S.rot[i]);
    const float4 bs = S.sizeR[i];
    Contacts cs{0, cra + d, crb + d, cmid + d};
    bool wrote;
    if (type != AVBD_SHAPE_BOX || bt != AVBD_SHAPE_BOX)
        return collideRound(type, bt, makeRShp(qc, qq, make_float4(s4.x, s4.y, s4.z, 0.0f)), makeRShp(bp, bq, bs),
                            S.rim, m, B, wrote, cs);
    const Poly c0{&poly[0][0][l], &poly[0][1][l], &poly[0][2][l], spA + d};
    const Poly c1{&poly[1][0][l], &poly[1][1][l], &poly[1][2][l], spB + d};
    return collideBoxes(qc, qq, v3{s4.x, s4.y, s4.z} * 0.5f, bp, bq, v3{bs.x, bs.y, bs.z} * 0.5f, m, B, wrote, cs, c0, c1);
}

// The query's candidate bounds: its tight bounds grown by the margin (query::shapeAabb).
AD void queryBox(int type, qt qq, v3 qc, float4 s4, float m, v3 &mn, v3 &mx)
{
    const v3 h = type == AVBD_SHAPE_SPHERE     ? v3{s4.x, s4.x, s4.x}
                 : type == AVBD_SHAPE_CYLINDER ? v3{s4.x, s4.x, s4.z * 0.5f}
                                               : v3{s4.x, s4.y, s4.z} * 0.5f;
    const m3 R = qmat(qq);
    const v3 wx{(fabsf(R.c[0][0]) * h.x + fabsf(R.c[1][0]) * h.y) + fabsf(R.c[2][0]) * h.z,
                (fabsf(R.c[0][1]) * h.x + fabsf(R.c[1][1]) * h.y) + fabsf(R.c[2][1]) * h.z,
                (fabsf(R.c[0][2]) * h.x + fabsf(R.c[1][2]) * h.y) + fabsf(R.c[2][2]) * h.z};
    mn = qc - wx;
    mx = qc + wx;
    if (m > 0.0f) { mn = mn - v3{m, m, m}; mx = mx + v3{m, m, m}; }
}

// The bodies part of the server's collide_shape (avbd_direct_space_state3d.cpp:386-447): query
// r (laid out as for k_intersect_shape) against its candidates in list order; every contact as
// a pair of world points (on the query, on the body) until maxPairs pairs are out:
// pts[2 (r maxPairs + p)] and [.. + 1]; cnt[r] pairs, anyHit[r] = some candidate touched.
__global__ void __launch_bounds__(NARROW_TPB) k_collide_shape(State S, int base, int count, int n, const float4 *qin,
                                                             const float4 *bounds, int maxPairs, int *cnt,
                                                             uint8_t *anyHit, float4 *pts, float4 *cra, float4 *crb,
                                                             float4 *cmid, float4 *spA, float4 *spB, QTree tree)
{
    __shared__ NarrowPolys poly;
    const int r = GID;
    if (r >= n) return;
    const float4 c4 = qin[4 * r], r4 = qin[4 * r + 1], s4 = qin[4 * r + 2], m4 = qin[4 * r + 3];
    const int type = __float_as_int(c4.w);
    const v3 qc = ld3(c4);
    const qt qq = ldq(r4);
    const float m = smax(0.0f, s4.w);
    const uint32_t mask = __float_as_uint(m4.x);
    v3 mn, mx;
    queryBox(type, qq, qc, s4, m, mn, mx);
    const size_t d = stageRow(r, 0);
    int pairs = 0;
    bool any = false;
    forCandidates(S, tree, base, count, bounds, mn, mx, mask, [&](int j) {
        const int i = base + j;
        m3 B;
        const int hits = queryContacts(S, type, qc, qq, s4, m, i, poly, threadIdx.x, d, cra, crb, cmid, spA, spB, B);
        if (hits <= 0) return true;
        any = true;
        const v3 bp = ld3(S.pos[i]);
        const qt bq = ldq(S.rot[i]);
        for (int k = 0; k < hits && pairs < maxPairs; k++) {
            const v3 wa = qc + qrot(qq, ld3(cra[d + crow(k)]));
            const v3 wb = bp + qrot(bq, ld3(crb[d + crow(k)]));
            const size_t o = 2 * (static_cast<size_t>(r) * maxPairs + pairs);
            pts[o] = st3(wa);
            pts[o + 1] = st3(wb);
            pairs++;
        }
        return pairs < maxPairs;
    });
    cnt[r] = pairs;
    anyHit[r] = any ? 1 : 0;
}

// The bodies part of the server's rest_info (avbd_direct_space_state3d.cpp:449-513): over the
// candidates in list order and their contacts, the deepest by penetrationDepth along the
// contact normal (a later one must be strictly deeper).  body[r] (-1: none); hitA[r] = (point
// on the query, depth); hitB[r] = (normal = the basis' first row, 0).
__global__ void __launch_bounds__(NARROW_TPB) k_rest_info(State S, int base, int count, int n, const float4 *qin,
                                                         const float4 *bounds, int *body, float4 *hitA, float4 *hitB,
                                                         float4 *cra, float4 *crb, float4 *cmid, float4 *spA,
                                                         float4 *spB, QTree tree)
{
    __shared__ NarrowPolys poly;
    const int r = GID;
    if (r >= n) return;
    const float4 c4 = qin[4 * r], r4 = qin[4 * r + 1], s4 = qin[4 * r + 2], m4 = qin[4 * r + 3];
    const int type = __float_as_int(c4.w);
    const v3 qc = ld3(c4);
    const qt qq = ldq(r4);
    const float m = smax(0.0f, s4.w);
    const uint32_t mask = __float_as_uint(m4.x);
    v3 mn, mx;
    queryBox(type, qq, qc, s4, m, mn, mx);
    const size_t d = stageRow(r, 0);
    float best = -1.0f;
    v3 bestPoint{0, 0, 0}, bestNormal{0, 0, 0};
    int bestBody = -1;
    forCandidates(S, tree, base, count, bounds, mn, mx, mask, [&](int j) {
        const int i = base + j;
        m3 B;
        const int hits = queryContacts(S, type, qc, qq, s4, m, i, poly, threadIdx.x, d, cra, crb, cmid, spA, spB, B);
        if (hits <= 0) return true;
        const v3 nrm = row(B, 0);
        const float dep = penetrationDepth(type, qc, qq, v3{s4.x, s4.y, s4.z}, S.shape[i], ld3(S.pos[i]), ldq(S.rot[i]),
                                          ld3(S.sizeR[i]), nrm);
        for (int k = 0; k < hits; k++) {
            if (dep <= best) continue;
            best = dep;
            bestNormal = nrm;
            bestPoint = qc + qrot(qq, ld3(cra[d + crow(k)]));
            bestBody = j;
        }
        return true;
    });
    body[r] = bestBody;
    hitA[r] = st3(bestPoint, best);
    hitB[r] = st3(bestNormal, 0.0f);
}

// query_shape.cpp:280-323: the closest point of body i's shape at its current pose to p.
AD v3 closestPoint(int type, v3 c, qt q, float4 size, v3 point)
{
    const v3 p = qrot(qconj(q), point - c);
    v3 local = p;
    if (type == AVBD_SHAPE_SPHERE) {
        const float len = length(p);
        local = len > 1.0e-6f ? p * (size.x / len) : v3{size.x, 0, 0};
    } else if (type == AVBD_SHAPE_CYLINDER) {
        const float r = size.x, hh = size.z * 0.5f;
        const float radial = sqrtf(p.x * p.x + p.y * p.y);
        const float clampedZ = sclamp(p.z, -hh, hh);
        if (radial > r) {
            const float scale = r / radial;
            local = v3{p.x * scale, p.y * scale, clampedZ};
        } else if (radial <= 1.0e-6f) {
            const float capDistance = hh - fabsf(p.z);
            local = capDistance < r ? v3{0, 0, p.z < 0.0f ? -hh : hh} : v3{r, 0, p.z};
        } else {
            const float sideDistance = r - radial;
            const float capDistance = hh - fabsf(p.z);
            if (capDistance < sideDistance) {
                local = v3{p.x, p.y, p.z < 0.0f ? -hh : hh};
            } else {
                const float scale = r / radial;
                local = v3{p.x * scale, p.y * scale, p.z};
            }
        }
    } else {
        const v3 h = v3{size.x, size.y, size.z} * 0.5f;
        local = v3{sclamp(p.x, -h.x, h.x), sclamp(p.y, -h.y, h.y), sclamp(p.z, -h.z, h.z)};
    }
    return c + qrot(q, local);
}

// The server's get_closest_point_to_object_volume for bodies: out[r] = the closest point of
// body (base + idx[r])'s volume to pts[r] (w of pts unused).
__global__ void k_closest_point(State S, int base, int n, const int *idx, const float4 *pts, float4 *out)
{
    const int r = GID;
    if (r >= n) return;
    const int i = base + idx[r];
    out[r] = st3(closestPoint(S.shape[i], ld3(S.pos[i]), ldq(S.rot[i]), S.sizeR[i], ld3(pts[r])));
}

// ---- shape sweeps (query_shape.cpp:213-258, 402-498; the server's cast_motion) ----
struct SweepShape {
    int type;
    v3 c;
    qt q;
    float4 size;   // w = 0
};

// query::shapeAabb's half extent (query_shape.cpp:371-394)
AD v3 shapeHalf(int type, qt q, float4 s)
{
    const v3 h = type == AVBD_SHAPE_SPHERE     ? v3{s.x, s.x, s.x}
                 : type == AVBD_SHAPE_CYLINDER ? v3{s.x, s.x, s.z * 0.5f}
                                               : v3{s.x, s.y, s.z} * 0.5f;
    const m3 R = qmat(q);
    return v3{(fabsf(R.c[0][0]) * h.x + fabsf(R.c[1][0]) * h.y) + fabsf(R.c[2][0]) * h.z,
              (fabsf(R.c[0][1]) * h.x + fabsf(R.c[1][1]) * h.y) + fabsf(R.c[2][1]) * h.z,
              (fabsf(R.c[0][2]) * h.x + fabsf(R.c[1][2]) * h.y) + fabsf(R.c[2][2]) * h.z};
}

// sweptAabbInterval (query_shape.cpp:213-241)
AD bool sweptInterval(v3 aMin, v3 aMax, v3 bMin, v3 bMax, v3 mv, float &rEnter, float &rExit)
{
    float enter = 0.0f, exit = 1.0f;
    for (int i = 0; i < 3; i++) {
        const float m = mv[i];
        if (fabsf(m) <= 1.0e-6f) {
            if (aMax[i] < bMin[i] || aMin[i] > bMax[i]) return false;
            continue;
        }
        float t0 = (bMin[i] - aMax[i]) / m, t1 = (bMax[i] - aMin[i]) / m;
        if (t0 > t1) { const float x = t0; t0 = t1; t1 = x; }
        enter = smax(enter, t0);
        exit = smin(exit, t1);
        if (enter > exit) return false;
    }
    if (exit < 0.0f || enter > 1.0f) return false;
    rEnter = smax(0.0f, enter);
    rExit = smin(1.0f, exit);
    return rEnter <= rExit;
}

// contactAt (query_shape.cpp:243-258): the narrow phase with the query moved to c + mv * t; the
// contact with the largest |rA - rB| (the first of equals) and the basis normal.
AD bool sweepContact(const State &S, const SweepShape &a, v3 mv, float t, int bt, v3 bp, qt bq, float4 bs, float margin,
                     Contacts &cs, const Poly &c0, const Poly &c1, v3 &normal, v3 &rA)
{
    const v3 c = a.c + mv * t;
    m3 B;
    bool wrote;
    const int n = a.type != AVBD_SHAPE_BOX || bt != AVBD_SHAPE_BOX
                      ? collideRound(a.type, bt, makeRShp(c, a.q, a.size), makeRShp(bp, bq, bs), S.rim, margin, B, wrote, cs)
                      : collideBoxes(c, a.q, v3{a.size.x, a.size.y, a.size.z} * 0.5f, bp, bq, v3{bs.x, bs.y, bs.z} * 0.5f,
                                     margin, B, wrote, cs, c0, c1);
    if (n <= 0) return false;
    int best = 0;
    float bestL = -1.0f;
    for (int k = 0; k < n; k++) {
        const float l = length2(ld3(cs.ra[crow(k)]) - ld3(cs.rb[crow(k)]));
        if (k == 0 || bestL < l) { best = k; bestL = l; }
    }
    normal = v3{B.c[0][0], B.c[1][0], B.c[2][0]};
    rA = ld3(cs.ra[crow(best)]);
    return true;
}

// sweepShape (query_shape.cpp:402-498) without the depth: the AABB interval, 32 samples over
// it, 12 bisections of the first bracket, the contact re-read at the first colliding pose.
AD bool sweepShape(const State &S, const SweepShape &a, v3 mv, int bt, v3 bp, qt bq, float4 bs, float margin,
                   Contacts &cs, const Poly &c0, const Poly &c1, float &safe, float &unsafe, v3 &point, v3 &normal)
{
    v3 rA;
    if (length2(mv) <= 1.0e-6f * 1.0e-6f) {
        if (!sweepContact(S, a, v3{0.0f, 0.0f, 0.0f}, 0.0f, bt, bp, bq, bs, margin, cs, c0, c1, normal, rA)) return false;
        safe = unsafe = 0.0f;
        point = a.c + qrot(a.q, rA);
        return true;
    }
    const v3 ha = shapeHalf(a.type, a.q, a.size), hb = shapeHalf(bt, bq, bs);
    v3 aMin = a.c - ha, aMax = a.c + ha;
    if (margin > 0.0f) { aMin = aMin - v3{margin, margin, margin}; aMax = aMax + v3{margin, margin, margin}; }
    float enter = 0.0f, exit = 1.0f;
    if (!sweptInterval(aMin, aMax, bp - hb, bp + hb, mv, enter, exit)) return false;
    float low = enter, high = exit;
    if (sweepContact(S, a, mv, enter, bt, bp, bq, bs, margin, cs, c0, c1, normal, rA)) {
        safe = unsafe = enter;
        point = (a.c + mv * enter) + qrot(a.q, rA);
        return true;
    }
    bool found = false;
    float previous = enter;
    for (int i = 1; i <= 32; i++) {
        const float t = enter + (exit - enter) * (static_cast<float>(i) / 32.0f);
        if (sweepContact(S, a, mv, t, bt, bp, bq, bs, margin, cs, c0, c1, normal, rA)) {
            low = previous;
            high = t;
            found = true;
            break;
        }
        previous = t;
    }
    if (!found && sweepContact(S, a, mv, exit, bt, bp, bq, bs, margin, cs, c0, c1, normal, rA)) {
        low = previous;
        high = exit;
        found = true;
    }
    if (!found) return false;
    for (int i = 0; i < 12; i++) {
        const float mid = (low + high) * 0.5f;
        if (sweepContact(S, a, mv, mid, bt, bp, bq, bs, margin, cs, c0, c1, normal, rA)) high = mid;
        else low = mid;
    }
    if (!sweepContact(S, a, mv, high, bt, bp, bq, bs, margin, cs, c0, c1, normal, rA)) return false;
    safe = low;
    unsafe = high;
    point = (a.c + mv * high) + qrot(a.q, rA);
    return true;
}

// The bodies part of the server's cast_motion (avbd_direct_space_state3d.cpp:310-385): query r
// is qin[4r] = (centre, type bits), qin[4r+1] = rotation, qin[4r+2] = (size, margin),
// qin[4r+3] = (motion, mask bits).  Candidates: bodies whose bounds overlap the query's bounds
// swept by the motion and grown by the margin, in list order; the sweep with the smallest safe
// fraction (the first of equals) wins: body[r] (-1: none), hitA[r] = (point, safe),
// hitB[r] = (normal, unsafe), safe = unsafe = 1 when none.
__global__ void __launch_bounds__(NARROW_TPB) k_cast_motion(State S, int base, int count, int n, const float4 *qin,
                                                           const float4 *bounds, int *body, float4 *hitA, float4 *hitB,
                                                           float4 *cra, float4 *crb, float4 *cmid, float4 *spA, float4 *spB,
                                                           QTree tree)
{
    __shared__ NarrowPolys poly;
    const int r = GID;
    if (r >= n) return;
    const float4 c4 = qin[4 * r], r4 = qin[4 * r + 1], s4 = qin[4 * r + 2], m4 = qin[4 * r + 3];
    const SweepShape a{__float_as_int(c4.w), ld3(c4), ldq(r4), make_float4(s4.x, s4.y, s4.z, 0.0f)};
    const v3 mv = ld3(m4);
    const float m = smax(0.0f, s4.w);
    const uint32_t mask = __float_as_uint(m4.w);
    const v3 wx = shapeHalf(a.type, a.q, a.size);
    v3 mn = a.c - wx, mx = a.c + wx;
    mn = v3{smin(mn.x, mn.x + mv.x), smin(mn.y, mn.y + mv.y), smin(mn.z, mn.z + mv.z)};
    mx = v3{smax(mx.x, mx.x + mv.x), smax(mx.y, mx.y + mv.y), smax(mx.z, mx.z + mv.z)};
    if (m > 0.0f) { mn = mn - v3{m, m, m}; mx = mx + v3{m, m, m}; }
    const size_t d = stageRow(r, 0);
    const int l = threadIdx.x;
    const Poly c0{&poly[0][0][l], &poly[0][1][l], &poly[0][2][l], spA + d};
    const Poly c1{&poly[1][0][l], &poly[1][1][l], &poly[1][2][l], spB + d};
    int bestBody = -1;
    float bestSafe = 1.0f, bestUnsafe = 1.0f;
    v3 bestPoint{0.0f, 0.0f, 0.0f}, bestNormal{0.0f, 0.0f, 0.0f};
    forCandidates(S, tree, base, count, bounds, mn, mx, mask, [&](int j) {
        const int i = base + j;
        Contacts cs{0, cra + d, crb + d, cmid + d};
        float safe, unsafe;
        v3 point, normal;
        if (!sweepShape(S, a, mv, S.shape[i], ld3(S.pos[i]), ldq(S.rot[i]), S.sizeR[i], m, cs, c0, c1, safe, unsafe, point,
                        normal))
            return true;
        if (bestBody < 0 || safe < bestSafe) {
            bestBody = j;
            bestSafe = safe;
            bestUnsafe = unsafe;
            bestPoint = point;
            bestNormal = normal;
        }
        return true;
    });
    body[r] = bestBody;
    hitA[r] = st3(bestPoint, bestSafe);
    hitB[r] = st3(bestNormal, bestUnsafe);
}

// Solver::pick (solver_query.cpp:29-155): rays[2r] = (origin, mask bits), rays[2r+1] = dir.
// Every body that passes the mask: an exact sphere test, else the box bounding the shape; the
// nearest t (the earlier body on a tie).  local[r] = hit point in the body's frame.
__global__ void k_pick(State S, int base, int count, int n, const float4 *rays, int *body, float4 *local)
{
    const int r = GID;
    if (r >= n) return;
    const float4 o4 = rays[2 * r], d4 = rays[2 * r + 1];
    const v3 org = ld3(o4), dr = ld3(d4);
    const uint32_t mask = __float_as_uint(o4.w);
    float bestT = qinf();
    int best = -1;
    v3 bestLocal{0, 0, 0};
    for (int j = 0; j < count; j++) {
        const int i = base + j;
        if ((S.layer[i] & mask) == 0) continue;
        const int type = S.shape[i];
        const float4 sr = S.sizeR[i];
        const qt inv = qconj(ldq(S.rot[i]));
        const v3 ol = qrot(inv, org - ld3(S.pos[i])), d = qrot(inv, dr);
        if (type == AVBD_SHAPE_SPHERE) {
            const float rr = sr.x;
            const float bb = dot(ol, d);
            const float c = length2(ol) - rr * rr;
            const float disc = bb * bb - c;
            if (disc < 0.0f) continue;
            const float sq = sqrtf(disc);
            float t = -bb - sq;
            if (t < 0.0f) t = -bb + sq;
            if (t < 0.0f || t >= bestT) continue;
            bestT = t; best = j; bestLocal = ol + d * t;
            continue;
        }
        const v3 half = type == AVBD_SHAPE_CYLINDER ? v3{sr.x, sr.x, sr.z * 0.5f} : v3{sr.x, sr.y, sr.z} * 0.5f;
        float tEnter = 0.0f, tExit = qinf();
        bool hit = true;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float dk = d[k], ok = ol[k], hk = half[k];
            if (hit) {
                if (fabsf(dk) < 1.0e-6f) {
                    if (ok < -hk || ok > hk) hit = false;
                } else {
                    const float invD = 1.0f / dk;
                    float t0 = (-hk - ok) * invD, t1 = (hk - ok) * invD;
                    if (t0 > t1) { const float x = t0; t0 = t1; t1 = x; }
                    tEnter = smax(tEnter, t0);
                    tExit = smin(tExit, t1);
                    if (tEnter > tExit) hit = false;
                }
            }
        }
        if (!hit) continue;
        const float tHit = tEnter >= 0.0f ? tEnter : tExit;
        if (tHit < 0.0f) continue;
        if (tHit < bestT) { bestT = tHit; best = j; bestLocal = ol + d * tHit; }
    }
    body[r] = best;
    local[r] = st3(best >= 0 ? bestLocal : v3{0, 0, 0});
}

// Solver::collectAABB (solver_query.cpp:158-171): flag[j] = body j of the world passes the mask
// and its tight bounds overlap [mn, mx].
__global__ void k_collect_aabb(State S, int base, int count, const float4 *bounds, float4 mn4, float4 mx4, uint32_t mask,
                               uint8_t *flag)
{
    const int j = GID;
    if (j >= count) return;
    const int i = base + j;
    flag[j] = (S.layer[i] & mask) != 0 && boxesOverlap(ld3(bounds[2 * j]), ld3(bounds[2 * j + 1]), ld3(mn4), ld3(mx4))
                  ? 1 : 0;   // bounds from k_query_bounds
}

// ------------------------------------------------------------ contact reports ----
// The server's per-body contact report (avbd_physics_server3d.cpp:1593-1652), from the live
// manifolds of world w (contacts unpacked to the slot layout first).  maxRep[j] > 0 asks for
// body j's first maxRep[j] contacts in live order.
// 1) entries per asking body, 2) their positions (any order: each entry carries its live rank),
// 3) per body, the entries taken in live-rank order until it holds its contacts.
__global__ void k_report_count(State S, int w, const int *maxRep, uint32_t *cnt)
{
    const int k = GID;
    if (k >= S.wLive[w]) return;
    const int sb = S.wSlotBase[w], base = S.wBodyBase[w];
    const int slot = S.live[sb + k];
    if (S.mNum[slot] <= 0) return;
    const int a = S.mA[slot] - base, b = S.mB[slot] - base;
    if (maxRep[a] > 0) atomicAdd(cnt + a, 1u);
    if (maxRep[b] > 0) atomicAdd(cnt + b, 1u);
}
__global__ void k_report_fill(State S, int w, const int *maxRep, const uint32_t *off, uint32_t *cursor, int *entries)
{
    const int k = GID;
    if (k >= S.wLive[w]) return;
    const int sb = S.wSlotBase[w], base = S.wBodyBase[w];
    const int slot = S.live[sb + k];
    if (S.mNum[slot] <= 0) return;
    const int a = S.mA[slot] - base, b = S.mB[slot] - base;
    if (maxRep[a] > 0) entries[off[a] + atomicAdd(cursor + a, 1u)] = 2 * k;
    if (maxRep[b] > 0) entries[off[b] + atomicAdd(cursor + b, 1u)] = 2 * k + 1;
}
__global__ void k_report_emit(State S, int w, int n, const int *maxRep, int stride, const uint32_t *off,
                              const int *entries, int *count, float4 *pos, float4 *normal)
{
    const int j = GID;
    if (j >= n) return;
    const int lim = maxRep[j] < stride ? maxRep[j] : stride;
    int emitted = 0;
    if (lim > 0) {
        const int sb = S.wSlotBase[w], base = S.wBodyBase[w], i = base + j;
        const v3 p = ld3(S.pos[i]);
        const qt q = ldq(S.rot[i]);
        const int e0 = static_cast<int>(off[j]), e1 = static_cast<int>(off[j + 1]);
        int last = -1;
        while (emitted < lim) {
            int key = 0x7fffffff;
            for (int e = e0; e < e1; e++) {
                const int x = entries[e];
                if (x > last && x < key) key = x;
            }
            if (key == 0x7fffffff) break;
            last = key;
            const int slot = S.live[sb + (key >> 1)], side = key & 1;
            const int nc = S.mNum[slot];
            const float4 b0 = S.mRec[4 * static_cast<size_t>(slot)];
            const float sgn = side == 0 ? 1.0f : -1.0f;
            const int otherBody = (side == 0 ? S.mB[slot] : S.mA[slot]) - base;
            for (int c = 0; c < nc && emitted < lim; c++) {
                const size_t row = static_cast<size_t>(slot) * MAXC + c;
                const v3 arm = ld3(side == 0 ? S.sRA[row] : S.sRB[row]);
                const size_t o = static_cast<size_t>(j) * stride + emitted;
                pos[o] = st3(p + qrot(q, arm), __int_as_float(otherBody));
                normal[o] = st3(ld3(b0) * sgn);
                emitted++;
            }
        }
    }
    count[j] = emitted;
}

} // namespace avbd_dev

// ORACLE - test infrastructure only.  C entry points over the reference's own sources,
// compiled from where they lie in the reference tree (see oracle/Makefile): the narrow phase
// of src/avbd/collide.cpp (every shape pair), the ray test of src/avbd/query/query_shape.cpp
// and the header-only 6x6 solves of
// src/avbd/maths.h.  They pin the restatement in avbd_oracle.cpp directly.
#include "avbd/solver.h"
#include "avbd/query/query_shape.hpp"

// Any pair of shapes (0 box, 1 sphere, 2 cylinder), each built from a body's size as the
// solver's makeShape does.
extern "C" int ref_collide_shapes(int typeA, const float posA[3], const float rotA[4], const float sizeA[3],
                                  int typeB, const float posB[3], const float rotB[4], const float sizeB[3],
                                  float margin, float basis[9], int32_t keys[8], float rA[24], float rB[24])
{
    using namespace avbd;
    auto shape = [](int type, const float *p, const float *q, const float *s) {
        Shape sh;
        sh.type = static_cast<ShapeType>(type);
        sh.center = float3{p[0], p[1], p[2]};
        sh.rotation = quat{q[0], q[1], q[2], q[3]};
        sh.half = float3{s[0], s[1], s[2]} * 0.5f;
        sh.radius = s[0];
        sh.halfHeight = s[2] * 0.5f;
        sh.axis = qrot(sh.rotation, float3{0, 0, 1});
        return sh;
    };
    std::array<Manifold::Contact, 8> cs{};
    float3x3 b{};
    const int n = collideShapes(shape(typeA, posA, rotA, sizeA), shape(typeB, posB, rotB, sizeB),
                                std::span<Manifold::Contact>(cs), b, margin);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) basis[3 * r + c] = b[c][r];
    for (int k = 0; k < n; k++) {
        keys[k] = cs[k].feature.key;
        for (int c = 0; c < 3; c++) {
            rA[3 * k + c] = cs[k].rA[c];
            rB[3 * k + c] = cs[k].rB[c];
        }
    }
    return n;
}

extern "C" int ref_collide_boxes(const float posA[3], const float rotA[4], const float sizeA[3],
                                 const float posB[3], const float rotB[4], const float sizeB[3],
                                 float margin, float basis[9], int32_t keys[8], float rA[24],
                                 float rB[24])
{
    return ref_collide_shapes(0, posA, rotA, sizeA, 0, posB, rotB, sizeB, margin, basis, keys, rA, rB);
}

extern "C" int ref_solve6(const float *a36, const float *b6, float *x6)
{
    std::array<std::array<float, 6>, 6> a{};
    std::array<float, 6> b{}, x{};
    for (int r = 0; r < 6; r++) {
        b[r] = b6[r];
        for (int c = 0; c < 6; c++) a[r][c] = a36[6 * r + c];
    }
    int path = 0;
    if (avbd::ldlt6_solve(a, b, x)) {
        path = 1;
    } else {
        x = {};
        if (avbd::ldlt6_pivoted_solve(a, b, x)) path = 2;
        else x = {};
    }
    for (int i = 0; i < 6; i++) x6[i] = x[i];
    return path;
}

// query::containsPoint on a shape built from a body's size as the solver's makeShape does.
extern "C" int ref_contains_point(int type, const float pos[3], const float rot[4], const float size[3],
                                  const float point[3], float margin)
{
    using namespace avbd;
    Shape sh;
    sh.type = static_cast<ShapeType>(type);
    sh.center = float3{pos[0], pos[1], pos[2]};
    sh.rotation = quat{rot[0], rot[1], rot[2], rot[3]};
    sh.half = float3{size[0], size[1], size[2]} * 0.5f;
    sh.radius = size[0];
    sh.halfHeight = size[2] * 0.5f;
    sh.axis = qrot(sh.rotation, float3{0, 0, 1});
    return query::containsPoint(sh, float3{point[0], point[1], point[2]}, margin) ? 1 : 0;
}

// query::closestPoint on a shape built from a body's size as the solver's makeShape does.
extern "C" void ref_closest_point(int type, const float pos[3], const float rot[4], const float size[3],
                                  const float point[3], float out[3])
{
    using namespace avbd;
    Shape sh;
    sh.type = static_cast<ShapeType>(type);
    sh.center = float3{pos[0], pos[1], pos[2]};
    sh.rotation = quat{rot[0], rot[1], rot[2], rot[3]};
    sh.half = float3{size[0], size[1], size[2]} * 0.5f;
    sh.radius = size[0];
    sh.halfHeight = size[2] * 0.5f;
    sh.axis = qrot(sh.rotation, float3{0, 0, 1});
    const float3 c = query::closestPoint(sh, float3{point[0], point[1], point[2]});
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

// query::sweepShape of shape a moved by motion against shape b; out = safe, unsafe, depth,
// point, normal.
extern "C" int ref_sweep_shape(int typeA, const float posA[3], const float rotA[4], const float sizeA[3],
                               const float motion[3], int typeB, const float posB[3], const float rotB[4],
                               const float sizeB[3], float margin, float out[9])
{
    using namespace avbd;
    auto shape = [](int type, const float *p, const float *q, const float *s) {
        Shape sh;
        sh.type = static_cast<ShapeType>(type);
        sh.center = float3{p[0], p[1], p[2]};
        sh.rotation = quat{q[0], q[1], q[2], q[3]};
        sh.half = float3{s[0], s[1], s[2]} * 0.5f;
        sh.radius = s[0];
        sh.halfHeight = s[2] * 0.5f;
        sh.axis = qrot(sh.rotation, float3{0, 0, 1});
        return sh;
    };
    const query::SweepHit h = query::sweepShape(shape(typeA, posA, rotA, sizeA), float3{motion[0], motion[1], motion[2]},
                                                shape(typeB, posB, rotB, sizeB), margin);
    if (!h.hit) return 0;
    const float v[9] = {h.safe, h.unsafe, h.depth, h.point.x, h.point.y, h.point.z, h.normal.x, h.normal.y, h.normal.z};
    for (int k = 0; k < 9; k++) out[k] = v[k];
    return 1;
}

// query::intersectRay on a shape built from a body's size as the solver's makeShape does.
extern "C" int ref_intersect_ray(int type, const float pos[3], const float rot[4], const float size[3],
                                 const float origin[3], const float dir[3], float maxT, float *t, float point[3],
                                 float normal[3], uint8_t *inside)
{
    using namespace avbd;
    Shape sh;
    sh.type = static_cast<ShapeType>(type);
    sh.center = float3{pos[0], pos[1], pos[2]};
    sh.rotation = quat{rot[0], rot[1], rot[2], rot[3]};
    sh.half = float3{size[0], size[1], size[2]} * 0.5f;
    sh.radius = size[0];
    sh.halfHeight = size[2] * 0.5f;
    sh.axis = qrot(sh.rotation, float3{0, 0, 1});
    query::Ray ray{float3{origin[0], origin[1], origin[2]}, float3{dir[0], dir[1], dir[2]}, maxT};
    query::RayHit hit{};
    if (!query::intersectRay(sh, ray, hit)) return 0;
    *t = hit.t;
    for (int c = 0; c < 3; c++) { point[c] = hit.point[c]; normal[c] = hit.normal[c]; }
    *inside = hit.inside ? 1 : 0;
    return 1;
}

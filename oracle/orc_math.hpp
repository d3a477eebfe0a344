// ORACLE - test infrastructure only.  CPU restatement of the reference's math layer
// (src/avbd/maths.h, src/avbd/linalg.h) used to check the HIP path.  Nothing in the
// product may include, link or call this file.
//
// Every expression keeps the reference's evaluation order, because the build contract is
// bit-exact FP32 (compiled with -ffp-contract=off, no fast-math):
//  - linalg folds a sum as ((0 + x) + y) + z            (linalg.h:365-372, dot at :501)
//  - mul(matrix, vector) is (c0*v.x + c1*v.y) + c2*v.z  (linalg.h:533)
//  - linalg min/max/clamp are `a<b ? ...` ternaries     (linalg.h:155-157)
//  - quaternion product / rotation / matrix use Eigen's term order (maths.h:63-96)
#pragma once

#include <cmath>
#include <cstdint>

namespace orc {

struct f3 {
    float x, y, z;
    float &operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
};
struct quat {
    float x, y, z, w;
};
// Column-major 3x3 (c[col][row]), the layout of linalg::mat<float,3,3>.
struct m3 {
    f3 c[3];
    f3 &operator[](int i) { return c[i]; }
    const f3 &operator[](int i) const { return c[i]; }
};

inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline f3 &operator+=(f3 &a, f3 b) { a = a + b; return a; }
inline f3 &operator-=(f3 &a, f3 b) { a = a - b; return a; }
inline f3 cmul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline f3 fabs3(f3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }

inline float dot(f3 a, f3 b) { return ((0.0f + a.x * b.x) + a.y * b.y) + a.z * b.z; }
inline float length2(f3 a) { return dot(a, a); }
inline float length(f3 a) { return std::sqrt(length2(a)); }
inline f3 normalize(f3 a) { return a / length(a); }
inline float length2d(float a, float b) { return std::sqrt((0.0f + a * a) + b * b); }
inline f3 cross(f3 a, f3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// linalg ternaries (NaN-sensitive, kept as written).
inline float lmin(float a, float b) { return a < b ? a : b; }
inline float lmax(float a, float b) { return a < b ? b : a; }
inline float lclamp(float a, float lo, float hi) { return a < lo ? lo : a < hi ? a : hi; }
inline f3 lmin3(f3 a, f3 b) { return {lmin(a.x, b.x), lmin(a.y, b.y), lmin(a.z, b.z)}; }
inline f3 lclamp3(f3 a, f3 lo, f3 hi)
{
    return {lclamp(a.x, lo.x, hi.x), lclamp(a.y, lo.y, hi.y), lclamp(a.z, lo.z, hi.z)};
}
// std::min / std::max / std::clamp as libstdc++ defines them.
inline float smin(float a, float b) { return b < a ? b : a; }
inline float smax(float a, float b) { return a < b ? b : a; }
inline float sclamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

inline m3 operator+(const m3 &a, const m3 &b) { return {{a[0] + b[0], a[1] + b[1], a[2] + b[2]}}; }
inline m3 &operator+=(m3 &a, const m3 &b) { a = a + b; return a; }
inline m3 operator*(const m3 &a, float s) { return {{a[0] * s, a[1] * s, a[2] * s}}; }
inline m3 operator/(const m3 &a, float s) { return {{a[0] / s, a[1] / s, a[2] / s}}; }
inline m3 operator-(const m3 &a) { return {{-a[0], -a[1], -a[2]}}; }
inline f3 mul(const m3 &a, f3 v) { return (a[0] * v.x + a[1] * v.y) + a[2] * v.z; }
inline m3 mul(const m3 &a, const m3 &b) { return {{mul(a, b[0]), mul(a, b[1]), mul(a, b[2])}}; }
inline m3 mul(const m3 &a, const m3 &b, const m3 &c) { return mul(mul(a, b), c); }
inline f3 row(const m3 &m, int i) { return {m[0][i], m[1][i], m[2][i]}; }
inline m3 transpose(const m3 &m) { return {{row(m, 0), row(m, 1), row(m, 2)}}; }
inline m3 diagonal(float a, float b, float c)
{
    return {{f3{a, 0.0f, 0.0f}, f3{0.0f, b, 0.0f}, f3{0.0f, 0.0f, c}}};
}
inline m3 identity3() { return diagonal(1.0f, 1.0f, 1.0f); }
inline m3 zero3() { return {{f3{0, 0, 0}, f3{0, 0, 0}, f3{0, 0, 0}}}; }

// Quaternions (maths.h:48-152).
inline float qlength2(quat q) { return (q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w); }
inline quat qdiv(quat q, float s) { return {q.x / s, q.y / s, q.z / s, q.w / s}; }
inline quat qnormalize(quat q)
{
    const float z = qlength2(q);
    return z > 0.0f ? qdiv(q, std::sqrt(z)) : q;
}
inline quat qconj(quat q) { return {-q.x, -q.y, -q.z, q.w}; }
inline quat qmul(quat a, quat b)
{
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
            a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
// The generic joint's atan2 (a polynomial the device evaluates with the same operations, so
// both agree bit for bit): octant reduction, atan(a) = a P(a^2) on [0, 1] with a degree-19 odd
// polynomial fitted to arctan by relative least squares (<= 1.3 ulp there), IEEE zero signs.
// It stands in for std::atan2 in generic_joint.cpp:64; test_atan2_poly bounds the difference.
inline float atan2Poly(float y, float x)
{
    const float ax = std::fabs(x), ay = std::fabs(y);
    const bool swap = ay > ax;
    const float mx = swap ? ay : ax, mn = swap ? ax : ay;
    float r = 0.0f;
    if (mx > 0.0f) {
        const float a = mn / mx, s = a * a;
        float p = -0.0017904319101944566f;
        p = p * s + 0.010887323878705502f;
        p = p * s + -0.0310975369066f;
        p = p * s + 0.05783987045288086f;
        p = p * s + -0.0839373916387558f;
        p = p * s + 0.10947515815496445f;
        p = p * s + -0.14262957870960236f;
        p = p * s + 0.199983611702919f;
        p = p * s + -0.3333328664302826f;
        p = p * s + 1.0f;
        r = p * a;
    }
    if (swap) r = 1.5707963705062866f - r;
    if (std::signbit(x)) r = 3.1415927410125732f - r;
    return std::copysign(r, y);
}

inline f3 qrot(quat q, f3 v)
{
    const f3 z{q.x, q.y, q.z};
    f3 uv = cross(z, v);
    uv = uv + uv;
    return (v + uv * q.w) + cross(z, uv);
}
inline m3 qmat(quat q)
{
    const float tx = 2.0f * q.x, ty = 2.0f * q.y, tz = 2.0f * q.z;
    const float twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const float txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const float tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    return {{f3{1.0f - (tyy + tzz), txy + twz, txz - twy},
             f3{txy - twz, 1.0f - (txx + tzz), tyz + twx},
             f3{txz + twy, tyz - twx, 1.0f - (txx + tyy)}}};
}
// Rotation vector from b to a: 2 * vec(a * conj(b)).
inline f3 qsub(quat a, quat b)
{
    const quat d = qmul(a, qconj(b));
    return f3{d.x, d.y, d.z} * 2.0f;
}
// First-order integration: normalize(a + 0.5 * (b,0) * a).
inline quat qadd(quat a, f3 b)
{
    const quat p = qmul(quat{b.x, b.y, b.z, 0.0f}, a);
    return qnormalize(quat{a.x + p.x * 0.5f, a.y + p.y * 0.5f, a.z + p.z * 0.5f, a.w + p.w * 0.5f});
}

// Right-handed frame, rows = (normal, t1, t2) (maths.h:108-116).
inline m3 orthonormal(f3 n)
{
    f3 t1 = std::fabs(n.x) > std::fabs(n.z) ? f3{-n.y, n.x, 0.0f} : f3{0.0f, -n.z, n.y};
    t1 = normalize(t1);
    const f3 t2 = cross(n, t1);
    return transpose(m3{{n, t1, t2}});
}

} // namespace orc

// Device math for the AVBD step on gfx950.
//
// The results of this library are bit-identical to the reference CPU solver because every
// expression below keeps the reference's evaluation order (src/avbd/maths.h:48-152,
// src/avbd/linalg.h:155-157, 365-372, 501-533) and the library is compiled with
// -ffp-contract=off, correctly rounded fp32 division and square root, and denormals kept:
//   dot(a,b)      = ((0 + ax*bx) + ay*by) + az*bz
//   mul(M, v)     = (col0*v.x + col1*v.y) + col2*v.z       (M column-major, c[col][row])
//   quaternion    = Eigen's term order for product, rotation and matrix
//   min/max/clamp = the reference's ternaries (linalg `a<b?a:b`, libstdc++ `b<a?b:a`)
// Matrices multiply literally, zero entries included, so signed zeros propagate as on
// the CPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace avbd_dev {

struct v3 {
    float x, y, z;
    __device__ __host__ float &operator[](int i) { return (&x)[i]; }
    // select chain over values: a run-time index then stays in registers (a conditional over
    // the members themselves selects their addresses, which keeps the vector in memory)
    __device__ __host__ float operator[](int i) const
    {
        const float a = x, b = y, c = z;
        return i == 0 ? a : (i == 1 ? b : c);
    }
};
struct qt {
    float x, y, z, w;
};
struct m3 {
    v3 c[3];
    __device__ v3 &operator[](int i) { return c[i]; }
    __device__ const v3 &operator[](int i) const { return c[i]; }
};

#define AD __device__ __forceinline__

AD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
AD v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
AD v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
AD v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
AD v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
AD v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
AD v3 &operator+=(v3 &a, v3 b) { a = a + b; return a; }
AD v3 &operator-=(v3 &a, v3 b) { a = a - b; return a; }
AD v3 cmul(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
AD v3 fabs3(v3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
AD float dot(v3 a, v3 b) { return ((0.0f + a.x * b.x) + a.y * b.y) + a.z * b.z; }
AD float length2(v3 a) { return dot(a, a); }
AD float length(v3 a) { return sqrtf(length2(a)); }
AD v3 normalize(v3 a) { return a / length(a); }
AD float length2d(float a, float b) { return sqrtf((0.0f + a * a) + b * b); }
AD v3 cross(v3 a, v3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// linalg ternaries and libstdc++ std::min/max/clamp, kept apart because they differ on
// equal operands and NaN.
AD float lmin(float a, float b) { return a < b ? a : b; }
AD float lclamp(float a, float lo, float hi) { return a < lo ? lo : a < hi ? a : hi; }
AD v3 lmin3(v3 a, v3 b) { return {lmin(a.x, b.x), lmin(a.y, b.y), lmin(a.z, b.z)}; }
AD v3 lclamp3(v3 a, float lo, float hi) { return {lclamp(a.x, lo, hi), lclamp(a.y, lo, hi), lclamp(a.z, lo, hi)}; }
AD float smin(float a, float b) { return b < a ? b : a; }
AD float smax(float a, float b) { return a < b ? b : a; }
AD float sclamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

AD m3 madd(const m3 &a, const m3 &b) { return {{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }
AD m3 mscale(const m3 &a, float s) { return {{a.c[0] * s, a.c[1] * s, a.c[2] * s}}; }
AD m3 mdiv(const m3 &a, float s) { return {{a.c[0] / s, a.c[1] / s, a.c[2] / s}}; }
AD m3 mneg(const m3 &a) { return {{-a.c[0], -a.c[1], -a.c[2]}}; }
AD v3 mul(const m3 &a, v3 v) { return (a.c[0] * v.x + a.c[1] * v.y) + a.c[2] * v.z; }
AD m3 mul(const m3 &a, const m3 &b) { return {{mul(a, b.c[0]), mul(a, b.c[1]), mul(a, b.c[2])}}; }
AD v3 row(const m3 &m, int i) { return {m.c[0][i], m.c[1][i], m.c[2][i]}; }
AD m3 transpose(const m3 &m) { return {{row(m, 0), row(m, 1), row(m, 2)}}; }
AD m3 diag(float a, float b, float c) { return {{v3{a, 0.0f, 0.0f}, v3{0.0f, b, 0.0f}, v3{0.0f, 0.0f, c}}}; }
AD m3 ident() { return diag(1.0f, 1.0f, 1.0f); }
AD m3 zero3() { return {{v3{0, 0, 0}, v3{0, 0, 0}, v3{0, 0, 0}}}; }

AD float qlength2(qt q) { return (q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w); }
AD qt qnormalize(qt q)
{
    const float z = qlength2(q);
    if (z > 0.0f) {
        const float s = sqrtf(z);
        return {q.x / s, q.y / s, q.z / s, q.w / s};
    }
    return q;
}
AD qt qconj(qt q) { return {-q.x, -q.y, -q.z, q.w}; }
AD qt qmul(qt a, qt b)
{
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
            a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
// atan2 for the generic joint's twist angles: the odd polynomial atan(a) = a P(a^2) on
// [0, 1] (degree 19, least-squares fit to arctan with relative weights, <= 1.3 ulp there) after
// the usual octant reduction; IEEE signs of zero as atan2.  The oracle evaluates the same
// expression, so both sides agree bit for bit.
AD float atan2Poly(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
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
    if (__builtin_signbit(x)) r = 3.1415927410125732f - r;
    return __builtin_copysignf(r, y);
}

AD v3 qrot(qt q, v3 v)
{
    const v3 z{q.x, q.y, q.z};
    v3 uv = cross(z, v);
    uv = uv + uv;
    return (v + uv * q.w) + cross(z, uv);
}
AD m3 qmat(qt q)
{
    const float tx = 2.0f * q.x, ty = 2.0f * q.y, tz = 2.0f * q.z;
    const float twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const float txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const float tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    return {{v3{1.0f - (tyy + tzz), txy + twz, txz - twy},
             v3{txy - twz, 1.0f - (txx + tzz), tyz + twx},
             v3{txz + twy, tyz - twx, 1.0f - (txx + tyy)}}};
}
// 2 * vec(a * conj(b)): the rotation vector from b to a.
AD v3 qsub(qt a, qt b)
{
    const qt d = qmul(a, qconj(b));
    return v3{d.x, d.y, d.z} * 2.0f;
}
// normalize(a + 0.5 * (b, 0) * a)
AD qt qadd(qt a, v3 b)
{
    const qt p = qmul(qt{b.x, b.y, b.z, 0.0f}, a);
    return qnormalize(qt{a.x + p.x * 0.5f, a.y + p.y * 0.5f, a.z + p.z * 0.5f, a.w + p.w * 0.5f});
}
// Frame with rows (n, t1, t2).
AD m3 orthonormal(v3 n)
{
    v3 t1 = fabsf(n.x) > fabsf(n.z) ? v3{-n.y, n.x, 0.0f} : v3{0.0f, -n.z, n.y};
    t1 = normalize(t1);
    const v3 t2 = cross(n, t1);
    return transpose(m3{{n, t1, t2}});
}

AD v3 ld3(const float4 &f) { return {f.x, f.y, f.z}; }
AD qt ldq(const float4 &f) { return {f.x, f.y, f.z, f.w}; }
AD float4 st3(v3 v, float w = 0.0f) { return make_float4(v.x, v.y, v.z, w); }
AD float4 stq(qt q) { return make_float4(q.x, q.y, q.z, q.w); }

} // namespace avbd_dev

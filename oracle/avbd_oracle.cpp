// ORACLE - test infrastructure only.  A serial CPU restatement of the reference's
// `avbd::Solver::step()` / `stepInterleaved()` for box bodies and `Joint`s, used by the
// tests, the smoke check and the bench's CPU baseline as the checker of the HIP path.
// Nothing in the product may include, link or call this file.
//
// It restates (file:line in the reference tree):
//   bodies            src/avbd/types.hpp:75-139, rigid.cpp:22-108
//   broad phase       src/avbd/solver_broadphase.cpp:30-186 (pair set of bvh_broadphase.cpp:225-290)
//   Morton codes      src/avbd/bvh_broadphase.cpp:19-36, 61-96
//   narrow phase      src/avbd/collide.cpp:71-490, 1001-1120 (box-box only)
//   manifold store    src/avbd/manifold.cpp:20-88, 123-166, 168-671
//   joints            src/avbd/joint.cpp:28-230
//   step sequencing   src/avbd/solver_step.cpp:44-316, solver_solve.cpp:31-424
//   brake/projection  src/avbd/solver_projection.cpp:33-281
//   6x6 solves        src/avbd/maths.h:154-343, types.hpp:147-201
// in the reference's floating-point evaluation order, serially (the reference's results do
// not depend on its thread count: test/core_test.cpp:612-681).
//
// The manifold primal follows the reference's cache state machine literally (force cache,
// world arms, the PRIMAL_READY and ARMS_CACHED flags: manifold.cpp:440-671, solver_step.cpp:
// 209-296), including drift-gated arm reuse (jacobianRebuildDistance > 0).
#include "avbd_oracle.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cstring>
#include <limits>
#include <vector>

#include "orc_math.hpp"

using namespace orc;

namespace {

// Planted mistakes, for the oracle's own tests only: a nonzero id swaps one expression of the
// restatement for a plausible wrong one, so that tests/test_oracle_pins.py can show that some
// pin fails for each.  Zero (the default, and the only value anything but those tests uses)
// is the restatement itself.  The ids and what they change:
//    1 layer filter needs both directions      2 layer filter tests one direction only
//    3 a linear lock on z keeps gravity        4 velocity recovery ignores the locks
//    5 angular locks ignored                   6 adaptive beta doubles when converging
//    7 adaptive beta tests the old error       8 adaptive beta scale left out of the dual
//    9 Newton rounds rounded down             10 stiffness decay after the dual
//   11 stiffness decay ignored                12 hard joint keeps its lambda
//   13 hard joint stabilisation sign flipped  14 joint penalty cap ignores penalty_max
//   15 fracture tests the linear lambda       16 fracture threshold not squared
//   17 projection cap ignored                 18 projection sums not averaged
//   19 contact-only bodies always to Jacobi   20 speculative margin not extended by motion
//   21 convergence probe compares the wrong way
int g_plant = 0;

constexpr float PENALTY_MIN = 1.0f;
constexpr float COLLISION_MARGIN = 0.001f;
constexpr float STICK_THRESH = 0.00001f;
constexpr float CONTACT_RELEASE_GAP = 2.0e-3f;
constexpr float ADAPTIVE_BETA_MAX = 64.0f;
constexpr int MAXC = 8;
constexpr int kMaxColours = 8;
constexpr uint8_t kOverflow = kMaxColours;

struct Body {
    f3 pos; quat q;
    f3 initLin; quat initAng;
    f3 inertLin; quat inertAng;
    f3 vel, angVel, prevVel;
    f3 size; float mass; f3 moment; float friction, radius, gravity;
    uint32_t layer, mask; uint8_t lockLin, lockAng; uint8_t shape = AVBD_SHAPE_BOX;
    uint8_t sleepMode = AVBD_SLEEP_NEVER; bool sleeping = false; int stillFrames = 0;
    bool inUpdate = false; int updateIndex = -1; bool contactOnly = true;
};

// One entry of the force list (forces.hpp:19-113): a joint, a spring (spring.cpp) or a
// collision filter (IgnoreCollision: in the list, acts on nothing).
// One degree of freedom of a generic joint (forces.hpp:123-139).
struct JointAxis {
    uint8_t mode = AVBD_AXIS_FREE; float lower = 0, upper = 0, spring = 0, equilibrium = 0, damping = 0;
    float lambda = 0, penalty = 0, initial = 0;
};
struct Joint {
    int a, b; f3 rA, rB, C0Lin, C0Ang, penLin, penAng, lamLin, lamAng;
    float stiffLin, stiffAng, fracture, torqueArm; bool broken;
    uint8_t kind = AVBD_FORCE_JOINT; float rest = 0.0f;
    quat restQ{0, 0, 0, 1}; JointAxis axis[6];   // generic joints: 0-2 linear, 3-5 angular
    f3 resLin{0, 0, 0};   // the last dual's linear constraint value C (joint.cpp:178), 0 if none
};

struct Contact { int32_t key; f3 rA, rB; };

struct Block {
    m3 lhsLin, lhsAng, lhsCross; f3 rhsLin, rhsAng;
};

// ---------------------------------------------------------------- 6x6 solves ----
// maths.h:156-198 (unpivoted LDL^T, sanity |x| < 10).
bool ldlt6(const float a[6][6], const float b[6], float x[6])
{
    float l[6][6] = {}; float d[6] = {};
    for (int j = 0; j < 6; j++) {
        float dj = a[j][j];
        for (int s = 0; s < j; s++) dj -= l[j][s] * l[j][s] * d[s];
        if (!(dj > 1.0e-12f)) return false;
        d[j] = dj; l[j][j] = 1.0f;
        for (int i = j + 1; i < 6; i++) {
            float lij = a[i][j];
            for (int s = 0; s < j; s++) lij -= l[i][s] * l[j][s] * d[s];
            l[i][j] = lij / d[j];
        }
    }
    float z[6] = {};
    for (int i = 0; i < 6; i++) {
        float v = b[i];
        for (int s = 0; s < i; s++) v -= l[i][s] * z[s];
        z[i] = v;
    }
    for (int i = 0; i < 6; i++) x[i] = z[i] / d[i];
    for (int i = 5; i >= 0; i--) {
        float v = x[i];
        for (int s = i + 1; s < 6; s++) v -= l[s][i] * x[s];
        x[i] = v;
    }
    float worst = 0.0f;
    for (int i = 0; i < 6; i++) worst = smax(worst, std::fabs(x[i]));
    return std::isfinite(worst) && worst < 1.0e1f;
}

// maths.h:202-307 (Eigen-order pivoted LDL^T).
bool ldlt6Pivoted(const float ain[6][6], const float b[6], float x[6])
{
    float a[6][6]; std::memcpy(a, ain, sizeof(a));
    float d[6] = {}; int t[6] = {0, 1, 2, 3, 4, 5};
    bool zeroPivot = false;
    for (int k = 0; k < 6; k++) {
        int j = k; float best = std::fabs(a[k][k]);
        for (int i = k + 1; i < 6; i++) {
            const float c = std::fabs(a[i][i]);
            if (c > best) { best = c; j = i; }
        }
        t[k] = j;
        if (j != k) {
            for (int c = 0; c < 6; c++) std::swap(a[k][c], a[j][c]);
            for (int r = 0; r < 6; r++) std::swap(a[r][k], a[r][j]);
        }
        if (k > 0) {
            float temp[6] = {}, prod[6] = {};
            for (int i = 0; i < k; i++) temp[i] = d[i] * a[k][i];
            for (int i = 0; i < k; i++) prod[i] = a[k][i] * temp[i];
            float dt;
            switch (k) {
            case 1: dt = prod[0]; break;
            case 2: dt = prod[0] + prod[1]; break;
            case 3: dt = (prod[0] + prod[1]) + prod[2]; break;
            case 4: dt = (prod[0] + prod[2]) + (prod[1] + prod[3]); break;
            default: dt = ((prod[0] + prod[2]) + (prod[1] + prod[3])) + prod[4]; break;
            }
            a[k][k] -= dt;
            for (int i = 0; i < k; i++) {
                const float ti = temp[i];
                for (int r = k + 1; r < 6; r++) a[r][k] -= a[r][i] * ti;
            }
        }
        const float pivot = a[k][k];
        const bool valid = std::fabs(pivot) > 0.0f;
        if (!valid) {
            for (int r = k + 1; r < 6; r++)
                if (a[r][k] != 0.0f) return false;
            zeroPivot = true;
        } else if (zeroPivot) {
            return false;
        }
        if (valid)
            for (int r = k + 1; r < 6; r++) a[r][k] /= pivot;
        d[k] = pivot;
    }
    for (int i = 0; i < 6; i++) x[i] = b[i];
    for (int k = 0; k < 6; k++) std::swap(x[k], x[t[k]]);
    for (int jj = 0; jj < 5; jj++) {
        const float v = x[jj];
        if (v != 0.0f)
            for (int i = jj + 1; i < 6; i++) x[i] -= a[i][jj] * v;
    }
    const float tol = std::numeric_limits<float>::min();
    for (int i = 0; i < 6; i++) x[i] = std::fabs(d[i]) > tol ? x[i] / d[i] : 0.0f;
    for (int jj = 5; jj >= 1; jj--) {
        const float v = x[jj];
        if (v != 0.0f)
            for (int i = jj - 1; i >= 0; i--) x[i] -= a[jj][i] * v;
    }
    for (int k = 5; k >= 0; k--) std::swap(x[k], x[t[k]]);
    float worst = 0.0f;
    for (int i = 0; i < 6; i++) worst = smax(worst, std::fabs(x[i]));
    return std::isfinite(worst) && worst < 1.0e1f;
}

// types.hpp:176-197 + maths.h:315-343: LDL^T, else pivoted, else zero.
void solveDelta(const Block &blk, f3 &dLin, f3 &dAng)
{
    float a[6][6] = {};
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            a[r][c] = blk.lhsLin[c][r];
            a[r + 3][c + 3] = blk.lhsAng[c][r];
            a[r + 3][c] = blk.lhsCross[c][r];
            a[r][c + 3] = blk.lhsCross[r][c];
        }
    const float b[6] = {-blk.rhsLin.x, -blk.rhsLin.y, -blk.rhsLin.z,
                        -blk.rhsAng.x, -blk.rhsAng.y, -blk.rhsAng.z};
    float x[6] = {};
    if (ldlt6(a, b, x) || ldlt6Pivoted(a, b, x)) {
        dLin = f3{x[0], x[1], x[2]};
        dAng = f3{x[3], x[4], x[5]};
        return;
    }
    dLin = f3{0, 0, 0};
    dAng = f3{0, 0, 0};
}

// types.hpp:157-172
Block makeBlock(const Body &b, float dtSq)
{
    Block k;
    const f3 mo = f3{b.mass, b.mass, b.mass} / dtSq;
    k.lhsLin = diagonal(mo.x, mo.y, mo.z);
    k.lhsCross = zero3();
    k.rhsLin = cmul(mo, b.pos - b.inertLin);
    const m3 rot = qmat(b.q);
    const m3 iw = mul(rot, mul(diagonal(b.moment.x, b.moment.y, b.moment.z), transpose(rot)));
    k.lhsAng = iw / dtSq;
    k.rhsAng = mul(k.lhsAng, qsub(b.q, b.inertAng));
    return k;
}

// ---------------------------------------------------------------- narrow phase ----
// collide.cpp box-box path.
struct OBB { f3 center; quat rot; f3 half; f3 axis[3]; };
struct SatAxis { int type; int ia, ib; float sep; f3 n; bool valid; };
struct FaceFrame { int axisIndex; f3 normal, center, u, v; float eu, ev; };
enum { AXIS_FACE_A = 0, AXIS_FACE_B = 1, AXIS_EDGE = 2 };
constexpr float SAT_EPS = 1.0e-6f;
constexpr float PLANE_EPS = 1.0e-5f;
constexpr float MERGE_DIST_SQ = 1.0e-6f;
constexpr int MAX_POLY = 16;

OBB makeOBB(f3 center, quat rot, f3 half)
{
    OBB o{center, rot, half, {}};
    o.axis[0] = qrot(rot, f3{1, 0, 0});
    o.axis[1] = qrot(rot, f3{0, 1, 0});
    o.axis[2] = qrot(rot, f3{0, 0, 1});
    return o;
}

f3 supportPoint(const OBB &b, f3 dir)
{
    const float sx = dot(dir, b.axis[0]) >= 0.0f ? 1.0f : -1.0f;
    const float sy = dot(dir, b.axis[1]) >= 0.0f ? 1.0f : -1.0f;
    const float sz = dot(dir, b.axis[2]) >= 0.0f ? 1.0f : -1.0f;
    return ((b.center + b.axis[0] * (b.half.x * sx)) + b.axis[1] * (b.half.y * sy)) + b.axis[2] * (b.half.z * sz);
}

void faceAxes(const OBB &b, int i, f3 &u, f3 &v, float &eu, float &ev)
{
    if (i == 0) { u = b.axis[1]; v = b.axis[2]; eu = b.half.y; ev = b.half.z; }
    else if (i == 1) { u = b.axis[0]; v = b.axis[2]; eu = b.half.x; ev = b.half.z; }
    else { u = b.axis[0]; v = b.axis[1]; eu = b.half.x; ev = b.half.y; }
}

int clipPoly(const f3 *in, int n, f3 pn, float off, f3 *out)
{
    if (n <= 0) return 0;
    int m = 0;
    f3 a = in[n - 1];
    float da = dot(pn, a) - off;
    for (int i = 0; i < n; i++) {
        const f3 b = in[i];
        const float db = dot(pn, b) - off;
        const bool ai = da <= PLANE_EPS, bi = db <= PLANE_EPS;
        if (ai != bi) {
            float t = 0.0f;
            const float den = da - db;
            if (std::fabs(den) > SAT_EPS) t = sclamp(da / den, 0.0f, 1.0f);
            if (m < MAX_POLY) out[m++] = a + (b - a) * t;
        }
        if (bi && m < MAX_POLY) out[m++] = b;
        a = b; da = db;
    }
    return m;
}

struct ShapeIn { f3 center; quat rot; };

bool addContact(const ShapeIn &A, const ShapeIn &B, Contact *cs, int &cnt, f3 *mids, f3 xA, f3 xB, int key)
{
    const f3 mid = (xA + xB) * 0.5f;
    for (int i = 0; i < cnt; i++)
        if (length2(mid - mids[i]) < MERGE_DIST_SQ) return false;
    if (cnt >= MAXC) return false;
    Contact &c = cs[cnt];
    c.key = key;
    c.rA = qrot(qconj(A.rot), xA - A.center);
    c.rB = qrot(qconj(B.rot), xB - B.center);
    mids[cnt] = mid;
    ++cnt;
    return true;
}

int faceManifold(const ShapeIn &sA, const ShapeIn &sB, const OBB &bA, const OBB &bB, bool refIsA, int refAxis,
                 f3 nAB, float margin, Contact *cs)
{
    const OBB &ref = refIsA ? bA : bB;
    const OBB &inc = refIsA ? bB : bA;
    const f3 refOut = refIsA ? nAB : -nAB;
    FaceFrame fr;
    {
        const float sign = dot(refOut, ref.axis[refAxis]) >= 0.0f ? 1.0f : -1.0f;
        fr.axisIndex = refAxis;
        fr.normal = ref.axis[refAxis] * sign;
        fr.center = ref.center + fr.normal * ref.half[refAxis];
        faceAxes(ref, refAxis, fr.u, fr.v, fr.eu, fr.ev);
    }
    int incAxis = 0;
    {
        float best = -FLT_MAX;
        for (int i = 0; i < 3; i++) {
            const float d = std::fabs(dot(inc.axis[i], fr.normal));
            if (d > best) { best = d; incAxis = i; }
        }
    }
    f3 c0[MAX_POLY], c1[MAX_POLY];
    {
        const float sign = dot(inc.axis[incAxis], fr.normal) > 0.0f ? -1.0f : 1.0f;
        const f3 fn = inc.axis[incAxis] * sign;
        const f3 fc = inc.center + fn * inc.half[incAxis];
        f3 u, v; float eu, ev;
        faceAxes(inc, incAxis, u, v, eu, ev);
        c0[0] = (fc + u * eu) + v * ev;
        c0[1] = (fc - u * eu) + v * ev;
        c0[2] = (fc - u * eu) - v * ev;
        c0[3] = (fc + u * eu) - v * ev;
    }
    int n = 4;
    f3 p;
    p = fr.u;  n = clipPoly(c0, n, p, dot(p, fr.center) + fr.eu, c1); if (!n) return 0;
    p = -fr.u; n = clipPoly(c1, n, p, dot(p, fr.center) + fr.eu, c0); if (!n) return 0;
    p = fr.v;  n = clipPoly(c0, n, p, dot(p, fr.center) + fr.ev, c1); if (!n) return 0;
    p = -fr.v; n = clipPoly(c1, n, p, dot(p, fr.center) + fr.ev, c0); if (!n) return 0;

    int cnt = 0;
    f3 mids[MAXC];
    int prefix = (refIsA ? AXIS_FACE_A : AXIS_FACE_B) << 24;
    prefix |= (refAxis & 0xFF) << 16;
    prefix |= (incAxis & 0xFF) << 8;
    const float accept = smax(margin, PLANE_EPS);
    for (int i = 0; i < n && cnt < MAXC; i++) {
        const f3 pi = c0[i];
        const float dist = dot(pi - fr.center, fr.normal);
        if (dist > accept) continue;
        const f3 pr = pi - fr.normal * dist;
        addContact(sA, sB, cs, cnt, mids, refIsA ? pr : pi, refIsA ? pi : pr, prefix | (i & 0xFF));
    }
    if (!cnt) {
        const f3 xA = supportPoint(bA, nAB);
        f3 xB = supportPoint(bB, -nAB);
        xB = xA - nAB * dot(xA - xB, nAB);
        addContact(sA, sB, cs, cnt, mids, xA, xB, prefix);
    }
    return cnt;
}

void supportEdge(const OBB &b, int ai, f3 dir, f3 &ea, f3 &eb)
{
    const int a1 = (ai + 1) % 3, a2 = (ai + 2) % 3;
    const float s1 = dot(dir, b.axis[a1]) >= 0.0f ? 1.0f : -1.0f;
    const float s2 = dot(dir, b.axis[a2]) >= 0.0f ? 1.0f : -1.0f;
    const f3 ec = (b.center + b.axis[a1] * (b.half[a1] * s1)) + b.axis[a2] * (b.half[a2] * s2);
    ea = ec - b.axis[ai] * b.half[ai];
    eb = ec + b.axis[ai] * b.half[ai];
}

void closestSeg(f3 p0, f3 p1, f3 q0, f3 q1, f3 &c0, f3 &c1)
{
    const f3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
    const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    float s = 0.0f, t = 0.0f;
    if (a <= SAT_EPS && e <= SAT_EPS) { c0 = p0; c1 = q0; return; }
    if (a <= SAT_EPS) {
        t = sclamp(f / e, 0.0f, 1.0f);
    } else {
        const float c = dot(d1, r);
        if (e <= SAT_EPS) {
            s = sclamp(-c / a, 0.0f, 1.0f);
        } else {
            const float b = dot(d1, d2);
            const float den = a * e - b * b;
            if (std::fabs(den) > SAT_EPS) s = sclamp((b * f - c * e) / den, 0.0f, 1.0f);
            t = (b * s + f) / e;
            if (t < 0.0f) { t = 0.0f; s = sclamp(-c / a, 0.0f, 1.0f); }
            else if (t > 1.0f) { t = 1.0f; s = sclamp((b - c) / a, 0.0f, 1.0f); }
        }
    }
    c0 = p0 + d1 * s;
    c1 = q0 + d2 * t;
}

int edgeContact(const ShapeIn &sA, const ShapeIn &sB, const OBB &bA, const OBB &bB, int axA, int axB, f3 nAB,
                Contact *cs)
{
    f3 a0, a1, b0, b1;
    supportEdge(bA, axA, nAB, a0, a1);
    supportEdge(bB, axB, -nAB, b0, b1);
    f3 xA, xB;
    closestSeg(a0, a1, b0, b1, xA, xB);
    int cnt = 0;
    f3 mids[MAXC];
    const int key = (AXIS_EDGE << 24) | ((axA & 0xFF) << 8) | (axB & 0xFF);
    addContact(sA, sB, cs, cnt, mids, xA, xB, key);
    if (!cnt) {
        xA = supportPoint(bA, nAB);
        xB = supportPoint(bB, -nAB);
        addContact(sA, sB, cs, cnt, mids, xA, xB, key);
    }
    return cnt;
}

// ------------------------------------------------------------ round shapes ----
// collide.cpp:505-998: spheres and capped cylinders.  A shape carries what makeShape derives
// from a body (collide.cpp:505-516): half extents, radius = size.x, half height = size.z / 2,
// the cylinder axis = local +Z in world space.
struct RShape { int type; f3 c; quat q; f3 half; float r, hh; f3 axis; };

RShape makeRShape(int type, f3 p, quat q, f3 size)
{
    return RShape{type, p, q, size * 0.5f, size.x, size.z * 0.5f, qrot(q, f3{0, 0, 1})};
}

enum { AXIS_CYL_SAMPLE = 3, AXIS_CYL_CORNER = 4 };
constexpr int RIM = 8, RINGS = 3;
constexpr int CYL_SAMPLES = 2 * (RIM + 1) + RIM * (RINGS - 1);   // 34
constexpr int MAX_CAND = 2 * CYL_SAMPLES + 8;

// collide.cpp:733-753: rim directions at angle steps i*5 mod 8 (coprime stride), from the
// C library's cosf / sinf evaluated at run time.
struct Rim { float c[RIM], s[RIM]; };
const Rim &rim()
{
    static const Rim t = [] {
        Rim r{};
        for (int i = 0; i < RIM; i++) {
            volatile float a = 6.28318530718f * static_cast<float>(i * 5 % RIM) / static_cast<float>(RIM);
            r.c[i] = std::cos(static_cast<float>(a));
            r.s[i] = std::sin(static_cast<float>(a));
        }
        return r;
    }();
    return t;
}

// collide.cpp:520-546
bool pointInBox(const RShape &b, f3 p, f3 &push, float &depth)
{
    const f3 l = qrot(qconj(b.q), p - b.c);
    float best = FLT_MAX; int axis = 0; float sign = 1.0f;
    for (int i = 0; i < 3; i++) {
        const float d = b.half[i] - std::fabs(l[i]);
        if (d <= 0.0f) return false;
        if (d < best) { best = d; axis = i; sign = l[i] >= 0.0f ? 1.0f : -1.0f; }
    }
    f3 lp{0, 0, 0};
    lp[axis] = sign;
    push = qrot(b.q, lp);
    depth = best;
    return true;
}

// collide.cpp:549-577
bool pointNearBox(const RShape &b, f3 p, float margin, f3 &push, float &depth)
{
    const f3 l = qrot(qconj(b.q), p - b.c);
    bool outside = false;
    for (int i = 0; i < 3; i++)
        if (std::fabs(l[i]) >= b.half[i]) { outside = true; break; }
    if (!outside) return false;
    const f3 cl{sclamp(l.x, -b.half.x, b.half.x), sclamp(l.y, -b.half.y, b.half.y), sclamp(l.z, -b.half.z, b.half.z)};
    const f3 d = l - cl;
    const float gap = length(d);
    if (gap > margin || gap <= 1.0e-9f) return false;
    push = qrot(b.q, d / gap);
    depth = -gap;
    return true;
}

// collide.cpp:580-614
f3 closestOnCylinder(const RShape &cy, f3 p, f3 &outward, float &distance)
{
    const f3 l = qrot(qconj(cy.q), p - cy.c);
    const float t = sclamp(l.z, -cy.hh, cy.hh);
    const float lenSq = (0.0f + l.x * l.x) + l.y * l.y;
    f3 cl{0, 0, t};
    if (lenSq > cy.r * cy.r) {
        const float scale = cy.r / std::sqrt(lenSq);
        cl.x = l.x * scale; cl.y = l.y * scale;
    } else if (lenSq > 1.0e-12f) {
        cl.x = l.x; cl.y = l.y;
    } else {
        cl.x = cy.r;
    }
    const f3 w = qrot(cy.q, cl) + cy.c;
    const f3 d = p - w;
    distance = length(d);
    outward = distance > 1.0e-9f ? d / distance : qrot(cy.q, f3{1, 0, 0});
    return w;
}

// collide.cpp:617-651
bool pointInCylinder(const RShape &cy, f3 p, f3 &push, float &surface)
{
    const f3 l = qrot(qconj(cy.q), p - cy.c);
    const float rs = l.x * l.x + l.y * l.y;
    if (rs >= cy.r * cy.r || std::fabs(l.z) >= cy.hh) return false;
    const float radialDepth = cy.r - std::sqrt(rs);
    const float capDepth = cy.hh - std::fabs(l.z);
    f3 pl{0, 0, 0};
    if (capDepth < radialDepth) {
        pl.z = l.z >= 0.0f ? 1.0f : -1.0f;
        surface = capDepth;
    } else {
        const float len = std::sqrt(rs);
        if (len > 1.0e-6f) { pl.x = l.x / len; pl.y = l.y / len; }
        else pl.x = 1.0f;
        surface = radialDepth;
    }
    push = qrot(cy.q, pl);
    return true;
}

// collide.cpp:654-671
bool pointNearCylinder(const RShape &cy, f3 p, float margin, f3 &push, float &depth)
{
    const f3 l = qrot(qconj(cy.q), p - cy.c);
    const float rs = l.x * l.x + l.y * l.y;
    if (rs < cy.r * cy.r && std::fabs(l.z) < cy.hh) return false;
    f3 out; float dist;
    closestOnCylinder(cy, p, out, dist);
    if (dist > margin) return false;
    push = out;
    depth = -dist;
    return true;
}

// collide.cpp:674-684: the exact inside test first, the speculative one only when it fails
bool insideBox(const RShape &b, f3 p, float margin, f3 &push, float &depth)
{
    return pointInBox(b, p, push, depth) || (margin > 0.0f && pointNearBox(b, p, margin, push, depth));
}
bool insideCylinder(const RShape &cy, f3 p, float margin, f3 &push, float &depth)
{
    return pointInCylinder(cy, p, push, depth) || (margin > 0.0f && pointNearCylinder(cy, p, margin, push, depth));
}

// collide.cpp:755-784: both caps (centre, then the rim), then the inner rings
int cylinderSamples(const RShape &cy, f3 *out)
{
    int n = 0;
    const f3 x = qrot(cy.q, f3{1, 0, 0}), y = qrot(cy.q, f3{0, 1, 0});
    const Rim &t = rim();
    for (int end = -1; end <= 1; end += 2) {
        const f3 cc = cy.c + cy.axis * (cy.hh * static_cast<float>(end));
        out[n++] = cc;
        for (int i = 0; i < RIM; i++) out[n++] = cc + (x * t.c[i] + y * t.s[i]) * cy.r;
    }
    for (int ring = 1; ring < RINGS; ring++) {
        const float tt = -cy.hh + 2.0f * cy.hh * static_cast<float>(ring) / static_cast<float>(RINGS);
        const f3 rc = cy.c + cy.axis * tt;
        for (int i = 0; i < RIM; i++) out[n++] = rc + (x * t.c[i] + y * t.s[i]) * cy.r;
    }
    return n;
}

struct Cand { f3 xA, xB, n; float depth; int32_t key; };

// collide.cpp:702-730: stable by depth (deepest first), the first MAXC offered in turn,
// stopping at the first one the manifold refuses
int emitRanked(const ShapeIn &A, const ShapeIn &B, const Cand *cand, int nc, Contact *cs)
{
    Cand ranked[MAX_CAND];
    for (int i = 0; i < nc; i++) {
        int j = i;
        while (j > 0 && ranked[j - 1].depth < cand[i].depth) { ranked[j] = ranked[j - 1]; --j; }
        ranked[j] = cand[i];
    }
    int cnt = 0;
    f3 mids[MAXC];
    const int emit = nc < MAXC ? nc : MAXC;
    for (int i = 0; i < emit; i++)
        if (!addContact(A, B, cs, cnt, mids, ranked[i].xA, ranked[i].xB, ranked[i].key)) break;
    return cnt;
}

// collide.cpp:786-806
int sphereSphere(const RShape &a, const RShape &b, Contact *cs, m3 &basis, float margin)
{
    const f3 d = b.c - a.c;
    const float distSq = length2(d);
    const float sumPlus = a.r + b.r + margin;
    if (distSq >= sumPlus * sumPlus) return 0;
    const float dist = std::sqrt(distSq);
    const f3 n = dist > 1.0e-6f ? d / dist : f3{0, 0, 1};
    basis = orthonormal(-n);
    int cnt = 0; f3 mids[MAXC];
    addContact(ShapeIn{a.c, a.q}, ShapeIn{b.c, b.q}, cs, cnt, mids, a.c + n * a.r, b.c - n * b.r, 1);
    return cnt;
}

// collide.cpp:809-862
int sphereBox(const RShape &a, const RShape &b, bool sphereIsA, Contact *cs, m3 &basis, float margin)
{
    const RShape &sp = sphereIsA ? a : b;
    const RShape &bx = sphereIsA ? b : a;
    const f3 l = qrot(qconj(bx.q), sp.c - bx.c);
    const f3 cl{sclamp(l.x, -bx.half.x, bx.half.x), sclamp(l.y, -bx.half.y, bx.half.y), sclamp(l.z, -bx.half.z, bx.half.z)};
    const f3 onBox = qrot(bx.q, cl) + bx.c;
    const f3 d = sp.c - onBox;
    const float distSq = length2(d);
    const float radPlus = sp.r + margin;
    if (distSq >= radPlus * radPlus) return 0;
    const float dist = std::sqrt(distSq);
    f3 n;
    if (dist > 1.0e-6f) {
        n = d / dist;
    } else {
        float depth;
        if (!pointInBox(bx, sp.c, n, depth)) return 0;
    }
    basis = orthonormal(-(sphereIsA ? -n : n));
    const f3 onSphere = sp.c - n * sp.r;
    int cnt = 0; f3 mids[MAXC];
    addContact(ShapeIn{a.c, a.q}, ShapeIn{b.c, b.q}, cs, cnt, mids, sphereIsA ? onSphere : onBox,
               sphereIsA ? onBox : onSphere, 1);
    return cnt;
}

// collide.cpp:865-891
int sphereCylinder(const RShape &a, const RShape &b, bool sphereIsA, Contact *cs, m3 &basis, float margin)
{
    const RShape &sp = sphereIsA ? a : b;
    const RShape &cy = sphereIsA ? b : a;
    f3 out; float dist;
    const f3 onCyl = closestOnCylinder(cy, sp.c, out, dist);
    if (dist >= sp.r + margin) return 0;
    basis = orthonormal(-(sphereIsA ? -out : out));
    const f3 onSphere = sp.c - out * sp.r;
    int cnt = 0; f3 mids[MAXC];
    addContact(ShapeIn{a.c, a.q}, ShapeIn{b.c, b.q}, cs, cnt, mids, sphereIsA ? onSphere : onCyl,
               sphereIsA ? onCyl : onSphere, 1);
    return cnt;
}

// collide.cpp:894-954
int cylinderBox(const RShape &a, const RShape &b, bool cylIsA, Contact *cs, m3 &basis, float margin)
{
    const RShape &cy = cylIsA ? a : b;
    const RShape &bx = cylIsA ? b : a;
    Cand cand[MAX_CAND];
    int nc = 0;
    f3 nAB{0, 0, 0};
    f3 smp[CYL_SAMPLES];
    const int ns = cylinderSamples(cy, smp);
    for (int i = 0; i < ns && nc < MAX_CAND; i++) {
        f3 push; float depth;
        if (!insideBox(bx, smp[i], margin, push, depth)) continue;
        Cand &c = cand[nc++];
        c.xA = cylIsA ? smp[i] : smp[i] + push * depth;
        c.xB = cylIsA ? smp[i] + push * depth : smp[i];
        c.n = cylIsA ? -push : push;
        c.depth = depth;
        c.key = static_cast<int32_t>((uint32_t(AXIS_CYL_SAMPLE) << 24) | (uint32_t(cylIsA ? 1u : 0u) << 20) |
                                     static_cast<uint32_t>(i));
        nAB += c.n;
    }
    for (int k = 0; k < 8 && nc < MAX_CAND; k++) {
        const f3 cl{(k & 1) ? bx.half.x : -bx.half.x, (k & 2) ? bx.half.y : -bx.half.y,
                    (k & 4) ? bx.half.z : -bx.half.z};
        const f3 cw = qrot(bx.q, cl) + bx.c;
        f3 push; float sd;
        if (!insideCylinder(cy, cw, margin, push, sd)) continue;
        Cand &c = cand[nc++];
        c.xA = cylIsA ? cw + push * sd : cw;
        c.xB = cylIsA ? cw : cw + push * sd;
        c.n = cylIsA ? push : -push;
        c.depth = sd;
        c.key = static_cast<int32_t>((uint32_t(AXIS_CYL_CORNER) << 24) | (uint32_t(cylIsA ? 0u : 1u) << 20) |
                                     static_cast<uint32_t>(k));
        nAB += c.n;
    }
    if (!nc) return 0;
    basis = orthonormal(-(length2(nAB) > 1.0e-12f ? normalize(nAB) : cy.axis));
    return emitRanked(ShapeIn{a.c, a.q}, ShapeIn{b.c, b.q}, cand, nc, cs);
}

// collide.cpp:957-998
int cylinderCylinder(const RShape &a, const RShape &b, Contact *cs, m3 &basis, float margin)
{
    Cand cand[MAX_CAND];
    int nc = 0;
    f3 nAB{0, 0, 0};
    f3 smp[CYL_SAMPLES];
    for (int pass = 0; pass < 2; pass++) {
        const RShape &from = pass == 0 ? a : b;
        const RShape &into = pass == 0 ? b : a;
        const bool fromIsA = pass == 0;
        const int ns = cylinderSamples(from, smp);
        for (int i = 0; i < ns && nc < MAX_CAND; i++) {
            f3 push; float sd;
            if (!insideCylinder(into, smp[i], margin, push, sd)) continue;
            Cand &c = cand[nc++];
            c.xA = fromIsA ? smp[i] : smp[i] + push * sd;
            c.xB = fromIsA ? smp[i] + push * sd : smp[i];
            c.n = fromIsA ? -push : push;
            c.depth = sd;
            c.key = static_cast<int32_t>((uint32_t(AXIS_CYL_SAMPLE) << 24) | (uint32_t(fromIsA ? 1u : 0u) << 20) |
                                         static_cast<uint32_t>(i));
            nAB += c.n;
        }
    }
    if (!nc) return 0;
    basis = orthonormal(-(length2(nAB) > 1.0e-12f ? normalize(nAB) : a.axis));
    return emitRanked(ShapeIn{a.c, a.q}, ShapeIn{b.c, b.q}, cand, nc, cs);
}

// collide.cpp:1001-1032: the dispatch for pairs with a round shape
int collideRound(const RShape &a, const RShape &b, Contact *cs, m3 &basis, float margin)
{
    const int S = AVBD_SHAPE_SPHERE, C = AVBD_SHAPE_CYLINDER;
    if (a.type == S && b.type == S) return sphereSphere(a, b, cs, basis, margin);
    if (a.type == S || b.type == S) {
        const bool sphereIsA = a.type == S;
        if (a.type == C || b.type == C) return sphereCylinder(a, b, sphereIsA, cs, basis, margin);
        return sphereBox(a, b, sphereIsA, cs, basis, margin);
    }
    if (a.type == C && b.type == C) return cylinderCylinder(a, b, cs, basis, margin);
    return cylinderBox(a, b, a.type == C, cs, basis, margin);
}

// ------------------------------------------------------------ space queries ----
// query/query_shape.cpp:8-270: exact ray against one shape.
constexpr float QEPS = 1.0e-6f;
struct RayHit { float t; f3 point, normal; bool inside; };

bool insideLocal(const RShape &s, f3 p, float m)
{
    if (s.type == AVBD_SHAPE_SPHERE) return length2(p) <= (s.r + m) * (s.r + m);
    if (s.type == AVBD_SHAPE_CYLINDER) return (p.x * p.x + p.y * p.y) <= (s.r + m) * (s.r + m) && std::fabs(p.z) <= s.hh + m;
    return std::fabs(p.x) <= s.half.x + m && std::fabs(p.y) <= s.half.y + m && std::fabs(p.z) <= s.half.z + m;
}
f3 boxNormal(f3 p, f3 h)
{
    const f3 q{std::fabs(p.x) / (h.x > 0.0f ? h.x : 1.0f), std::fabs(p.y) / (h.y > 0.0f ? h.y : 1.0f),
               std::fabs(p.z) / (h.z > 0.0f ? h.z : 1.0f)};
    if (q.x >= q.y && q.x >= q.z) return f3{p.x < 0.0f ? -1.0f : 1.0f, 0, 0};
    if (q.y >= q.z) return f3{0, p.y < 0.0f ? -1.0f : 1.0f, 0};
    return f3{0, 0, p.z < 0.0f ? -1.0f : 1.0f};
}
f3 cylinderNormal(f3 p, float hh)
{
    if (std::fabs(std::fabs(p.z) - hh) <= 1.0e-4f) return f3{0, 0, p.z < 0.0f ? -1.0f : 1.0f};
    const float radial = std::sqrt(p.x * p.x + p.y * p.y);
    if (radial <= QEPS) return f3{1, 0, 0};
    return f3{p.x / radial, p.y / radial, 0};
}
bool raySphere(const RShape &s, f3 o, f3 d, float maxT, RayHit &h)
{
    const float a = dot(d, d);
    if (a <= QEPS) return false;
    const float b = dot(o, d);
    const float c = dot(o, o) - s.r * s.r;
    const float disc = b * b - a * c;
    if (disc < 0.0f) return false;
    const float root = std::sqrt(disc);
    float t0 = (-b - root) / a, t1 = (-b + root) / a;
    if (t0 > t1) std::swap(t0, t1);
    const bool inside = c < 0.0f;
    float t = inside ? t1 : t0;
    if (t < 0.0f) t = t1;
    if (t < 0.0f || t > maxT) return false;
    const f3 n = o + d * t;
    h = RayHit{t, s.c + qrot(s.q, n), qrot(s.q, normalize(n)), inside};
    return true;
}
bool rayBox(const RShape &s, f3 o, f3 d, float maxT, RayHit &h)
{
    float tEnter = 0.0f, tExit = std::numeric_limits<float>::infinity();
    for (int i = 0; i < 3; i++) {
        if (std::fabs(d[i]) <= QEPS) {
            if (o[i] < -s.half[i] || o[i] > s.half[i]) return false;
            continue;
        }
        const float inv = 1.0f / d[i];
        float t0 = (-s.half[i] - o[i]) * inv, t1 = (s.half[i] - o[i]) * inv;
        if (t0 > t1) std::swap(t0, t1);
        tEnter = smax(tEnter, t0);
        tExit = smin(tExit, t1);
        if (tEnter > tExit) return false;
    }
    const bool inside = insideLocal(s, o, 0.0f);
    const float t = tEnter > 0.0f ? tEnter : tExit;
    if (t < 0.0f || t > maxT) return false;
    const f3 p = o + d * t;
    h = RayHit{t, s.c + qrot(s.q, p), qrot(s.q, boxNormal(p, s.half)), inside};
    return true;
}
bool rayCylinder(const RShape &s, f3 o, f3 d, float maxT, RayHit &h)
{
    const float r = s.r, hh = s.hh;
    float hits[4] = {};
    int count = 0;
    auto add = [&](float t) {
        if (t >= -QEPS && t <= maxT + QEPS && count < 4) hits[count++] = smax(0.0f, t);
    };
    const float a = d.x * d.x + d.y * d.y;
    if (a > QEPS) {
        const float b = o.x * d.x + o.y * d.y;
        const float c = o.x * o.x + o.y * o.y - r * r;
        const float disc = b * b - a * c;
        if (disc >= 0.0f) {
            const float root = std::sqrt(disc);
            const float ts[2] = {(-b - root) / a, (-b + root) / a};
            for (float t : ts)
                if (std::fabs(o.z + d.z * t) <= hh + QEPS) add(t);
        }
    } else if (o.x * o.x + o.y * o.y > r * r + QEPS) {
        return false;
    }
    if (std::fabs(d.z) > QEPS) {
        const float zs[2] = {hh, -hh};
        for (float z : zs) {
            const float t = (z - o.z) / d.z;
            const float px = o.x + d.x * t, py = o.y + d.y * t;
            if (px * px + py * py <= r * r + QEPS) add(t);
        }
    }
    if (count == 0) return false;
    for (int i = 1; i < count; i++) {
        const float v = hits[i];
        int j = i - 1;
        while (j >= 0 && hits[j] > v) { hits[j + 1] = hits[j]; j--; }
        hits[j + 1] = v;
    }
    const bool inside = insideLocal(s, o, 0.0f);
    const float t = inside ? hits[count - 1] : hits[0];
    if (t < 0.0f || t > maxT) return false;
    const f3 p = o + d * t;
    h = RayHit{t, s.c + qrot(s.q, p), qrot(s.q, cylinderNormal(p, hh)), inside};
    return true;
}
// query_shape.cpp:334-337
bool containsPoint(const RShape &s, f3 p, float margin)
{
    return insideLocal(s, qrot(qconj(s.q), p - s.c), margin);
}
// query_shape.cpp:258-274
bool intersectRay(const RShape &s, f3 origin, f3 dir, float rayMaxT, RayHit &h)
{
    const float maxT = rayMaxT > 0.0f ? rayMaxT : std::numeric_limits<float>::infinity();
    const f3 o = qrot(qconj(s.q), origin - s.c);
    const f3 d = qrot(qconj(s.q), dir);
    if (s.type == AVBD_SHAPE_SPHERE) return raySphere(s, o, d, maxT, h);
    if (s.type == AVBD_SHAPE_CYLINDER) return rayCylinder(s, o, d, maxT, h);
    return rayBox(s, o, d, maxT, h);
}

} // namespace

// collide.cpp:1034-1097 (box-box SAT + clipping).  Returns the contact count; `basis` is
// written only when the boxes are within the margin on every axis.
extern "C" int oracle_collide_boxes(const float posA[3], const float rotA[4], const float sizeA[3],
                                    const float posB[3], const float rotB[4], const float sizeB[3],
                                    float margin, float basis[9], int32_t keys[8], float rA[24],
                                    float rB[24])
{
    const ShapeIn sA{{posA[0], posA[1], posA[2]}, {rotA[0], rotA[1], rotA[2], rotA[3]}};
    const ShapeIn sB{{posB[0], posB[1], posB[2]}, {rotB[0], rotB[1], rotB[2], rotB[3]}};
    const f3 hA = f3{sizeA[0], sizeA[1], sizeA[2]} * 0.5f;
    const f3 hB = f3{sizeB[0], sizeB[1], sizeB[2]} * 0.5f;
    const OBB bA = makeOBB(sA.center, sA.rot, hA);
    const OBB bB = makeOBB(sB.center, sB.rot, hB);
    const f3 delta = bB.center - bA.center;
    float M[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[i][j] = dot(bA.axis[i], bB.axis[j]);
    const f3 dA{dot(delta, bA.axis[0]), dot(delta, bA.axis[1]), dot(delta, bA.axis[2])};
    const f3 dB{dot(delta, bB.axis[0]), dot(delta, bB.axis[1]), dot(delta, bB.axis[2])};
    SatAxis bf{}; bf.sep = -FLT_MAX; bf.valid = false;
    SatAxis be{}; be.sep = -FLT_MAX; be.valid = false;
    for (int i = 0; i < 3; i++) {
        const float sep = std::fabs(dA[i]) -
                (((bA.half[i] + bB.half.x * std::fabs(M[i][0])) + bB.half.y * std::fabs(M[i][1])) +
                        bB.half.z * std::fabs(M[i][2]));
        if (sep > margin) return 0;
        if (!bf.valid || sep > bf.sep) {
            bf.valid = true; bf.type = AXIS_FACE_A; bf.ia = i; bf.ib = -1; bf.sep = sep;
            bf.n = bA.axis[i] * (dA[i] < 0.0f ? -1.0f : 1.0f);
        }
    }
    for (int j = 0; j < 3; j++) {
        const float sep = std::fabs(dB[j]) -
                (((bB.half[j] + bA.half.x * std::fabs(M[0][j])) + bA.half.y * std::fabs(M[1][j])) +
                        bA.half.z * std::fabs(M[2][j]));
        if (sep > margin) return 0;
        if (!bf.valid || sep > bf.sep) {
            bf.valid = true; bf.type = AXIS_FACE_B; bf.ia = -1; bf.ib = j; bf.sep = sep;
            bf.n = bB.axis[j] * (dB[j] < 0.0f ? -1.0f : 1.0f);
        }
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const f3 axis = cross(bA.axis[i], bB.axis[j]);
            const float lenSq = length2(axis);
            if (lenSq < SAT_EPS) continue;
            const float invLen = 1.0f / std::sqrt(lenSq);
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            const float ra = (bA.half[i1] * std::fabs(M[i2][j]) + bA.half[i2] * std::fabs(M[i1][j])) * invLen;
            const float rb = (bB.half[j1] * std::fabs(M[i][j2]) + bB.half[j2] * std::fabs(M[i][j1])) * invLen;
            const float sd = dot(delta, axis) * invLen;
            const float sep = std::fabs(sd) - (ra + rb);
            if (sep > margin) return 0;
            if (!be.valid || sep > be.sep) {
                be.valid = true; be.type = AXIS_EDGE; be.ia = i; be.ib = j; be.sep = sep;
                be.n = axis * (sd < 0.0f ? -invLen : invLen);
            }
        }
    if (!bf.valid) return 0;
    SatAxis best = bf;
    if (be.valid && 0.95f * be.sep > bf.sep + 0.01f) best = be;
    const m3 B = orthonormal(-best.n);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) basis[r * 3 + c] = B[c][r];
    Contact cs[MAXC] = {};
    int n;
    if (best.type == AXIS_EDGE) n = edgeContact(sA, sB, bA, bB, best.ia, best.ib, best.n, cs);
    else if (best.type == AXIS_FACE_A) n = faceManifold(sA, sB, bA, bB, true, best.ia, best.n, margin, cs);
    else n = faceManifold(sA, sB, bA, bB, false, best.ib, best.n, margin, cs);
    for (int k = 0; k < n; k++) {
        keys[k] = cs[k].key;
        rA[3 * k] = cs[k].rA.x; rA[3 * k + 1] = cs[k].rA.y; rA[3 * k + 2] = cs[k].rA.z;
        rB[3 * k] = cs[k].rB.x; rB[3 * k + 1] = cs[k].rB.y; rB[3 * k + 2] = cs[k].rB.z;
    }
    return n;
}

// collide.cpp:1001-1097: any pair of shapes (AVBD_SHAPE_*); box-box takes the SAT path above.
extern "C" int oracle_collide_shapes(int typeA, const float posA[3], const float rotA[4], const float sizeA[3],
                                     int typeB, const float posB[3], const float rotB[4], const float sizeB[3],
                                     float margin, float basis[9], int32_t keys[8], float rA[24], float rB[24])
{
    if (typeA == AVBD_SHAPE_BOX && typeB == AVBD_SHAPE_BOX)
        return oracle_collide_boxes(posA, rotA, sizeA, posB, rotB, sizeB, margin, basis, keys, rA, rB);
    const RShape a = makeRShape(typeA, f3{posA[0], posA[1], posA[2]}, quat{rotA[0], rotA[1], rotA[2], rotA[3]},
                                f3{sizeA[0], sizeA[1], sizeA[2]});
    const RShape b = makeRShape(typeB, f3{posB[0], posB[1], posB[2]}, quat{rotB[0], rotB[1], rotB[2], rotB[3]},
                                f3{sizeB[0], sizeB[1], sizeB[2]});
    Contact cs[MAXC] = {};
    m3 B{};
    const int n = collideRound(a, b, cs, B, margin);
    if (n > 0)
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) basis[r * 3 + c] = B[c][r];
    for (int k = 0; k < n; k++) {
        keys[k] = cs[k].key;
        rA[3 * k] = cs[k].rA.x; rA[3 * k + 1] = cs[k].rA.y; rA[3 * k + 2] = cs[k].rA.z;
        rB[3 * k] = cs[k].rB.x; rB[3 * k + 1] = cs[k].rB.y; rB[3 * k + 2] = cs[k].rB.z;
    }
    return n;
}

namespace {

// ---------------------------------------------------------------- the world ----
struct World {
    avbd_params P;
    float collisionDt;
    int roundsThisStep = 8;
    std::vector<Body> bodies;
    std::vector<Joint> joints;
    // Manifold store: stable slots, live walk order, LIFO free list (manifold.cpp:20-88).
    std::vector<int> mA, mB;
    std::vector<m3> basis;
    std::vector<int> numC;
    std::vector<float> mFric;
    std::vector<int> dormant, saved;   // contact grace (manifold.hpp): ticks parked, contacts kept
    std::vector<Contact> cont;
    std::vector<f3> C0, pen, lam;
    std::vector<uint8_t> stick;
    std::vector<float> betaScale;
    // Drift-gated arm cache (manifold.hpp:76-97): per slot the flags (ARMS_CACHED = 1,
    // PRIMAL_READY = 2) and the first columns of both bodies' rotation matrices when the arms
    // were captured; per contact the world arms last computed and the primal's force cache.
    std::vector<uint8_t> mFlags;
    std::vector<f3> capA, capB;
    std::vector<f3> rAW, rBW, fCache;
    // the last dual's normal constraint value per contact (manifold.cpp:626-630; 0 while frozen)
    std::vector<float> resC;
    // test hooks: dual evaluations that reused cached arms, primal touches served by the cache
    long long armHits = 0, cachedTouches = 0;
    std::vector<uint32_t> live, freeList;
    // Adjacency CSR by body index: joints (list order) then manifolds (live order).
    // An entry >= 0 is a manifold slot, < 0 is joint -(j+1).
    std::vector<uint32_t> adjStart;
    std::vector<int> adj;
    std::vector<int> updateOrder;
    std::vector<uint8_t> colour;
    std::vector<int> colourBegins, colourOrder, overflowOrder;
    int colourCount = 0;
    std::vector<f3> dLin, dAng;
    // statistics of the last step
    int lastPairs = 0, lastNew = 0, lastLevels = 0;

    float pmax() const { return P.penalty_max; }

    uint32_t acquire(int a, int b)
    {
        uint32_t s;
        if (!freeList.empty()) { s = freeList.back(); freeList.pop_back(); }
        else {
            s = static_cast<uint32_t>(mA.size());
            mA.push_back(0); mB.push_back(0); basis.push_back(zero3()); numC.push_back(0); mFric.push_back(0);
            dormant.push_back(0); saved.push_back(0);
            mFlags.push_back(0); capA.push_back(f3{0, 0, 0}); capB.push_back(f3{0, 0, 0});
            cont.resize(cont.size() + MAXC); C0.resize(C0.size() + MAXC); pen.resize(pen.size() + MAXC);
            lam.resize(lam.size() + MAXC); stick.resize(stick.size() + MAXC); betaScale.resize(betaScale.size() + MAXC);
            rAW.resize(rAW.size() + MAXC); rBW.resize(rBW.size() + MAXC); fCache.resize(fCache.size() + MAXC);
            resC.resize(resC.size() + MAXC);
        }
        mA[s] = a; mB[s] = b; basis[s] = zero3(); numC[s] = 0; mFric[s] = 0.0f;
        dormant[s] = 0; saved[s] = 0;
        mFlags[s] = 0; capA[s] = capB[s] = f3{0, 0, 0};   // manifold.cpp:55-80
        for (int k = 0; k < MAXC; k++) {
            const int i = s * MAXC + k;
            cont[i] = Contact{0, {0, 0, 0}, {0, 0, 0}};
            C0[i] = pen[i] = lam[i] = f3{0, 0, 0};
            rAW[i] = rBW[i] = fCache[i] = f3{0, 0, 0};
            resC[i] = 0.0f;
            stick[i] = 0; betaScale[i] = 1.0f;
        }
        live.push_back(s);
        return s;
    }

    void buildAdjacency()
    {
        const int n = static_cast<int>(bodies.size());
        adjStart.assign(n + 1, 0);
        for (const Joint &j : joints) {
            if (j.a >= 0) adjStart[j.a + 1]++;
            adjStart[j.b + 1]++;
        }
        for (uint32_t s : live) { adjStart[mA[s] + 1]++; adjStart[mB[s] + 1]++; }
        for (int i = 0; i < n; i++) adjStart[i + 1] += adjStart[i];
        adj.assign(adjStart[n], 0);
        std::vector<uint32_t> cur(adjStart.begin(), adjStart.begin() + n);
        for (size_t k = 0; k < joints.size(); k++) {
            const int ref = -static_cast<int>(k) - 1;
            if (joints[k].a >= 0) adj[cur[joints[k].a]++] = ref;
            adj[cur[joints[k].b]++] = ref;
        }
        for (uint32_t s : live) {
            adj[cur[mA[s]]++] = static_cast<int>(s);
            adj[cur[mB[s]]++] = static_cast<int>(s);
        }
    }

    int otherOf(int ref, int body) const
    {
        int a, b;
        if (ref >= 0) { a = mA[ref]; b = mB[ref]; }
        else { const Joint &j = joints[-ref - 1]; a = j.a; b = j.b; }
        if (a == body) return b;
        if (b == body) return a;
        return -1;
    }

    // ------------------------------------------------------------ broad phase ----
    // rigid.cpp:83-99: world-space half extents of the body's shape at its current pose
    static f3 tightHalf(const Body &b)
    {
        const f3 h = b.shape == AVBD_SHAPE_SPHERE     ? f3{b.size.x, b.size.x, b.size.x}
                     : b.shape == AVBD_SHAPE_CYLINDER ? f3{b.size.x, b.size.x, b.size.z * 0.5f}
                                                      : b.size * 0.5f;
        const m3 R = qmat(b.q);
        return f3{(std::fabs(R[0][0]) * h.x + std::fabs(R[1][0]) * h.y) + std::fabs(R[2][0]) * h.z,
                  (std::fabs(R[0][1]) * h.x + std::fabs(R[1][1]) * h.y) + std::fabs(R[2][1]) * h.z,
                  (std::fabs(R[0][2]) * h.x + std::fabs(R[1][2]) * h.y) + std::fabs(R[2][2]) * h.z};
    }
    void sweptAabbs(std::vector<f3> &cen, std::vector<f3> &half) const
    {
        const int n = static_cast<int>(bodies.size());
        cen.resize(n); half.resize(n);
        for (int i = 0; i < n; i++) {
            const Body &b = bodies[i];
            const f3 tight = tightHalf(b);
            f3 disp = b.vel * collisionDt;
            if (b.mass > 0) disp += f3{0, 0, b.gravity * collisionDt * collisionDt};
            const float rotSweep = b.radius * length(b.angVel) * collisionDt;
            const float grow = rotSweep + P.speculative_margin;
            cen[i] = b.pos + disp * 0.5f;
            half[i] = (tight + fabs3(disp) * 0.5f) + f3{grow, grow, grow};
        }
    }

    static bool overlap(f3 ac, f3 ah, f3 bc, f3 bh)
    {
        const f3 aMin = ac - ah, aMax = ac + ah, bMin = bc - bh, bMax = bc + bh;
        return !(bMin.x > aMax.x || bMin.y > aMax.y || bMin.z > aMax.z || aMin.x > bMax.x || aMin.y > bMax.y ||
                 aMin.z > bMax.z);
    }

    // The candidate set the reference's LBVH self-query emits: every i < j whose swept
    // AABBs overlap on closed intervals (bvh_broadphase.cpp:231-290), sorted and unique.
    // Found here by a sweep over x, which yields the same set.
    void candidatePairs(const std::vector<f3> &cen, const std::vector<f3> &half, std::vector<uint64_t> &out) const
    {
        const int n = static_cast<int>(cen.size());
        out.clear();
        std::vector<int> ord(n);
        std::vector<float> lo(n);
        for (int i = 0; i < n; i++) { ord[i] = i; lo[i] = cen[i].x - half[i].x; }
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return lo[a] < lo[b]; });
        for (int p = 0; p < n; p++) {
            const int i = ord[p];
            const float hi = cen[i].x + half[i].x;
            for (int q = p + 1; q < n; q++) {
                const int j = ord[q];
                if (lo[j] > hi) break;
                if (overlap(cen[i], half[i], cen[j], half[j])) {
                    const uint64_t a = static_cast<uint32_t>(std::min(i, j));
                    const uint64_t b = static_cast<uint32_t>(std::max(i, j));
                    out.push_back((a << 32) | b);
                }
            }
        }
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    }

    // solver_internal.hpp:67-84: a force or a live manifold joins the pair; a manifold parked by
    // contact grace does not (its pair can still wake sleepers on bounds overlap)
    bool constrained(int a, int b) const
    {
        for (uint32_t k = adjStart[a]; k < adjStart[a + 1]; k++) {
            if (otherOf(adj[k], a) != b) continue;
            if (adj[k] >= 0 && numC[adj[k]] == 0 && dormant[adj[k]] > 0) continue;
            return true;
        }
        return false;
    }

    // solver_broadphase.cpp:67-80: a sleeper whose bounds touch an awake movable body's
    // (a pair about to get its first manifold) wakes, in pair order
    static void wakePair(Body &a, Body &b)
    {
        if (a.sleeping && !b.sleeping && b.mass > 0.0f) { a.sleeping = false; a.stillFrames = 0; }
        else if (b.sleeping && !a.sleeping && a.mass > 0.0f) { b.sleeping = false; b.stillFrames = 0; }
    }

    void broadPhase()
    {
        buildAdjacency();
        const int n = static_cast<int>(bodies.size());
        lastPairs = 0; lastNew = 0;
        if (n < 2) return;
        std::vector<f3> cen, half;
        sweptAabbs(cen, half);
        std::vector<uint64_t> pairs;
        candidatePairs(cen, half, pairs);
        lastPairs = static_cast<int>(pairs.size());
        std::vector<uint8_t> valid(pairs.size());
        for (size_t k = 0; k < pairs.size(); k++) {
            const int i = static_cast<int>(pairs[k] >> 32), j = static_cast<int>(pairs[k] & 0xFFFFFFFFu);
            const Body &A = bodies[i], &B = bodies[j];
            const bool ov = overlap(cen[i], half[i], cen[j], half[j]);
            const bool layers = g_plant == 1   ? ((A.layer & B.mask) != 0) && ((B.layer & A.mask) != 0)
                                : g_plant == 2 ? ((A.layer & B.mask) != 0)
                                               : ((A.layer & B.mask) != 0) || ((B.layer & A.mask) != 0);
            const bool bothStatic = A.mass <= 0.0f && B.mass <= 0.0f;
            valid[k] = ov && layers && !constrained(i, j) && !bothStatic;
        }
        for (size_t k = 0; k < pairs.size(); k++) {
            if (!valid[k]) continue;
            const int i = static_cast<int>(pairs[k] >> 32), j = static_cast<int>(pairs[k] & 0xFFFFFFFFu);
            wakePair(bodies[i], bodies[j]);
            bool exists = false;
            for (uint32_t e = adjStart[i]; e < adjStart[i + 1]; e++)
                if (adj[e] >= 0 && otherOf(adj[e], i) == j) { exists = true; break; }
            if (exists) continue;
            acquire(i, j);
            lastNew++;
        }
    }

    // ------------------------------------------------------------ manifolds ----
    f3 constraintError(f3 c0, f3 bracket, const m3 &B, float oma) const
    {
        const bool spec = P.speculative_margin > 0.0f;
        const bool separated = spec && c0.x > COLLISION_MARGIN + CONTACT_RELEASE_GAP;
        return c0 * (separated ? 1.0f : oma) + mul(B, bracket);
    }

    struct Clamp { float bounds, scale; };
    static Clamp clampForce(f3 &F, float mu)
    {
        F[0] = smin(F[0], 0.0f);
        const float bounds = std::fabs(F[0]) * mu;
        const float fs = length2d(F[1], F[2]);
        if (fs > bounds && fs > 0) {
            F[1] *= bounds / fs;
            F[2] *= bounds / fs;
        }
        return {bounds, fs};
    }

    struct Hess { float s00, s01, s02, s11, s12, s22; };
    static Hess basisHessian(const m3 &B, f3 k)
    {
        const f3 b0 = row(B, 0), b1 = row(B, 1), b2 = row(B, 2);
        const float b0x2 = b0.x * b0.x, b0y2 = b0.y * b0.y, b0z2 = b0.z * b0.z;
        const float b0xy = b0.x * b0.y, b0xz = b0.x * b0.z, b0yz = b0.y * b0.z;
        const float b1x2 = b1.x * b1.x, b1y2 = b1.y * b1.y, b1z2 = b1.z * b1.z;
        const float b1xy = b1.x * b1.y, b1xz = b1.x * b1.z, b1yz = b1.y * b1.z;
        const float b2x2 = b2.x * b2.x, b2y2 = b2.y * b2.y, b2z2 = b2.z * b2.z;
        const float b2xy = b2.x * b2.y, b2xz = b2.x * b2.z, b2yz = b2.y * b2.z;
        return {k.x * b0x2 + k.y * b1x2 + k.z * b2x2, k.x * b0xy + k.y * b1xy + k.z * b2xy,
                k.x * b0xz + k.y * b1xz + k.z * b2xz, k.x * b0y2 + k.y * b1y2 + k.z * b2y2,
                k.x * b0yz + k.y * b1yz + k.z * b2yz, k.x * b0z2 + k.y * b1z2 + k.z * b2z2};
    }

    float marginFor(const Body &a, const Body &b) const
    {
        float m = P.speculative_margin;
        if (m > 0.0f && g_plant != 20) {
            const float h = collisionDt;
            const float gA = a.mass > 0.0f ? a.gravity : 0.0f;
            const float gB = b.mass > 0.0f ? b.gravity : 0.0f;
            const f3 disp = (a.vel - b.vel) * h + f3{0, 0, (gA - gB) * h * h};
            m += length(disp) + (a.radius * length(a.angVel) + b.radius * length(b.angVel)) * h;
        }
        return m;
    }

    // manifold.cpp:237-435 (grace ticks off).
    bool initialize(uint32_t s)
    {
        const Body &a = bodies[mA[s]], &b = bodies[mB[s]];
        const int first = static_cast<int>(s) * MAXC;
        mFric[s] = std::sqrt(a.friction * b.friction);
        // manifold.cpp:246-247: a new contact set drops the cached arms and forces
        mFlags[s] = static_cast<uint8_t>(mFlags[s] & (g_plant == 24 ? ~PRIMAL_READY : ~(ARMS_CACHED | PRIMAL_READY)));
        // manifold.cpp:249-253: a parked manifold re-enters with the contacts it kept
        if (numC[s] == 0 && saved[s] > 0) numC[s] = saved[s];
        const int liveCount = numC[s];
        Contact nc[MAXC] = {};
        float bas[9];
        bas[0] = std::numeric_limits<float>::quiet_NaN();   // written only by a collision that writes the basis
        int32_t keys[8]; float ra[24], rb[24];
        const float pa[3] = {a.pos.x, a.pos.y, a.pos.z}, qa[4] = {a.q.x, a.q.y, a.q.z, a.q.w},
                    sa[3] = {a.size.x, a.size.y, a.size.z};
        const float pb[3] = {b.pos.x, b.pos.y, b.pos.z}, qb[4] = {b.q.x, b.q.y, b.q.z, b.q.w},
                    sb[3] = {b.size.x, b.size.y, b.size.z};
        const int nn = oracle_collide_shapes(a.shape, pa, qa, sa, b.shape, pb, qb, sb, marginFor(a, b), bas, keys, ra, rb);
        for (int k = 0; k < nn; k++)
            nc[k] = Contact{keys[k], {ra[3 * k], ra[3 * k + 1], ra[3 * k + 2]}, {rb[3 * k], rb[3 * k + 1], rb[3 * k + 2]}};
        // The reference's box pairs write the basis once the pair is within the margin on every
        // axis, even when clipping leaves no contact (collide.cpp:1089); round pairs write it
        // with their contact.  A pair left without contacts is released, or parked under
        // contact grace with this basis.
        if (!std::isnan(bas[0]))
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) basis[s][c][r] = bas[r * 3 + c];

        f3 nPen[MAXC] = {}, nLam[MAXC] = {};
        uint8_t nStick[MAXC] = {};
        float nBeta[MAXC], prevErr[MAXC];
        for (int i = 0; i < MAXC; i++) { nBeta[i] = 1.0f; prevErr[i] = -1.0f; }
        for (int i = 0; i < nn; i++) {
            for (int j = 0; j < numC[s]; j++) {
                if (nc[i].key != cont[first + j].key) continue;
                const f3 newRA = nc[i].rA, newRB = nc[i].rB;
                nc[i] = cont[first + j];
                prevErr[i] = smax(0.0f, COLLISION_MARGIN - C0[first + j].x);
                nPen[i] = pen[first + j];
                nLam[i] = lam[first + j];
                nStick[i] = stick[first + j];
                nBeta[i] = betaScale[first + j];
                if (!stick[first + j]) {
                    nc[i].rA = newRA; nc[i].rB = newRB;
                } else {
                    const f3 oldA = a.pos + qrot(a.q, cont[first + j].rA);
                    const f3 oldB = b.pos + qrot(b.q, cont[first + j].rB);
                    const f3 newA = a.pos + qrot(a.q, newRA);
                    const f3 newB = b.pos + qrot(b.q, newRB);
                    const float drift = smax(length(oldA - newA), length(oldB - newB));
                    if (drift > P.stick_drift_heal) {
                        nc[i].rA = newRA; nc[i].rB = newRB;
                        nStick[i] = 0;
                        nLam[i] = f3{nLam[i].x, 0.0f, 0.0f};
                    }
                }
                break;
            }
        }
        // manifold.cpp:338-347: with grace, unmatched contacts start at 5 % of beta_lin
        if (P.contact_grace_ticks > 0) {
            const float seed = smax(PENALTY_MIN, 0.05f * P.beta_lin);
            for (int i = 0; i < nn; i++)
                if (prevErr[i] < 0.0f) nPen[i] = f3{seed, seed, seed};
        }
        numC[s] = nn;
        for (int i = 0; i < nn; i++) {
            cont[first + i] = nc[i]; pen[first + i] = nPen[i]; lam[first + i] = nLam[i];
            stick[first + i] = nStick[i]; betaScale[first + i] = nBeta[i];
        }
        // manifold.cpp:359-378: a separated pair is parked (not released) for up to grace ticks,
        // its contacts kept with no force
        if (nn == 0) {
            if (P.contact_grace_ticks > 0 && dormant[s] < P.contact_grace_ticks) {
                for (int i = 0; i < liveCount; i++) { lam[first + i] = f3{0, 0, 0}; stick[first + i] = 0; }
                dormant[s]++;
                saved[s] = liveCount;
                return false;
            }
            saved[s] = 0;
        }
        dormant[s] = 0;
        for (int i = 0; i < numC[s]; i++) {
            const f3 xA = a.pos + qrot(a.q, cont[first + i].rA);
            const f3 xB = b.pos + qrot(b.q, cont[first + i].rB);
            C0[first + i] = mul(basis[s], xA - xB) + f3{COLLISION_MARGIN, 0, 0};
            if (C0[first + i].x < -P.max_penetration) C0[first + i].x = -P.max_penetration;
            if (P.adaptive_beta && prevErr[i] >= 0.0f) {
                const float err = smax(0.0f, COLLISION_MARGIN - C0[first + i].x);
                const float floor = P.adaptive_beta_error_floor;
                float sc = nBeta[i];
                if ((g_plant == 7 ? prevErr[i] : err) <= floor) sc = smax(sc * 0.5f, 1.0f);
                else if ((g_plant == 7 ? err : prevErr[i]) <= floor) sc = smin(sc * 2.0f, ADAPTIVE_BETA_MAX);
                else if (err > 0.5f * prevErr[i]) sc = smin(sc * 2.0f, ADAPTIVE_BETA_MAX);
                else if (err < 0.1f * prevErr[i]) sc = g_plant == 6 ? smin(sc * 2.0f, ADAPTIVE_BETA_MAX) : smax(sc * 0.5f, 1.0f);
                betaScale[first + i] = sc;
            } else {
                betaScale[first + i] = 1.0f;
            }
            if (C0[first + i][0] > COLLISION_MARGIN + CONTACT_RELEASE_GAP) {
                lam[first + i] = f3{0, 0, 0};
                stick[first + i] = 0;
            }
            lam[first + i] = lam[first + i] * P.alpha * P.gamma;
            pen[first + i] = lclamp3(pen[first + i] * P.gamma, f3{PENALTY_MIN, PENALTY_MIN, PENALTY_MIN},
                                     f3{pmax(), pmax(), pmax()});
        }
        return numC[s] > 0;
    }

    bool frozen(uint32_t s) const
    {
        auto awake = [](const Body &b) { return b.mass > 0.0f && !b.sleeping; };
        return !awake(bodies[mA[s]]) && !awake(bodies[mB[s]]);
    }

    // Contact forces of one manifold at the current poses (manifold.cpp:474-497): world arms
    // and cone-clamped forces.
    void contactForces(uint32_t s, f3 *F, f3 *rAw, f3 *rBw) const
    {
        const Body &a = bodies[mA[s]], &b = bodies[mB[s]];
        const int first = static_cast<int>(s) * MAXC;
        const f3 dqAL = a.pos - a.initLin, dqAA = qsub(a.q, a.initAng);
        const f3 dqBL = b.pos - b.initLin, dqBA = qsub(b.q, b.initAng);
        const m3 rotA = qmat(a.q), rotB = qmat(b.q);
        for (int i = 0; i < numC[s]; i++) {
            const f3 rA = mul(rotA, cont[first + i].rA);
            const f3 rB = mul(rotB, cont[first + i].rB);
            rAw[i] = rA; rBw[i] = rB;
            const f3 br = (dqAL - dqBL) + cross(dqAA, rA) - cross(dqBA, rB);
            const f3 C = constraintError(C0[first + i], br, basis[s], 1 - P.alpha);
            f3 f = cmul(pen[first + i], C) + lam[first + i];
            clampForce(f, mFric[s]);
            F[i] = f;
        }
    }

    static constexpr uint8_t ARMS_CACHED = 1, PRIMAL_READY = 2;   // manifold.hpp:96-97

    // manifold.cpp:465-501: the shared half of a manifold's primal, cached for the stamps
    void preCompute(uint32_t s)
    {
        const size_t first = static_cast<size_t>(s) * MAXC;
        contactForces(s, &fCache[first], &rAW[first], &rBW[first]);
        mFlags[s] |= PRIMAL_READY;
    }

    // manifold.cpp:503-568: one body's stamp from the cached forces and arms
    void stampCached(uint32_t s, int body, Block &blk) const
    {
        const bool fromA = body == mA[s];
        const float side = fromA ? 1.0f : -1.0f;
        const int first = static_cast<int>(s) * MAXC;
        const m3 &fr = basis[s];
        const f3 b0 = row(fr, 0), b1 = row(fr, 1), b2 = row(fr, 2);
        m3 accLin = zero3(), accAng = zero3(), accCross = zero3();
        f3 accRL{0, 0, 0}, accRA{0, 0, 0};
        for (int i = 0; i < numC[s]; i++) {
            const f3 f = fCache[first + i];
            const f3 r = fromA ? rAW[first + i] : rBW[first + i];
            const Hess h = basisHessian(fr, pen[first + i]);
            const m3 S{{f3{h.s00, h.s01, h.s02}, f3{h.s01, h.s11, h.s12}, f3{h.s02, h.s12, h.s22}}};
            const m3 RS{{cross(r, S[0]), cross(r, S[1]), cross(r, S[2])}};
            const float rx2 = r.x * r.x, ry2 = r.y * r.y, rz2 = r.z * r.z;
            const float rxy = r.x * r.y, rxz = r.x * r.z, ryz = r.y * r.z;
            const float h00 = h.s11 * rz2 - 2.0f * h.s12 * ryz + h.s22 * ry2;
            const float h11 = h.s22 * rx2 - 2.0f * h.s02 * rxz + h.s00 * rz2;
            const float h22 = h.s00 * ry2 - 2.0f * h.s01 * rxy + h.s11 * rx2;
            const float h01 = h.s12 * rxz + h.s02 * ryz - h.s22 * rxy - h.s01 * rz2;
            const float h02 = h.s12 * rxy + h.s01 * ryz - h.s11 * rxz - h.s02 * ry2;
            const float h12 = h.s02 * rxy + h.s01 * rxz - h.s00 * ryz - h.s12 * rx2;
            const f3 fw = (b0 * f.x + b1 * f.y) + b2 * f.z;
            const f3 fwAng = cross(r, fw);
            accLin += S;
            accAng += m3{{f3{h00, h01, h02}, f3{h01, h11, h12}, f3{h02, h12, h22}}};
            accCross += RS;
            accRL += fw;
            accRA += fwAng;
        }
        blk.lhsLin += accLin;
        blk.lhsAng += accAng;
        blk.lhsCross += accCross;
        blk.rhsLin += accRL * side;
        blk.rhsAng += accRA * side;
    }

    // manifold.cpp:570-581: the first touch of a round consumes the previous dual's cache (no
    // pose moved since); the second recomputes at the poses the first body's move left.
    void updatePrimal(uint32_t s, int body, Block &blk)
    {
        if (!(mFlags[s] & PRIMAL_READY) || g_plant == 26) preCompute(s);
        else cachedTouches++;
        stampCached(s, body, blk);
        mFlags[s] = static_cast<uint8_t>(mFlags[s] & ~PRIMAL_READY);
    }

    // manifold.cpp:583-671: lambda <- clamped force, penalty ramps; with arm caching
    // (jacobianRebuildDistance > 0) the world arms computed last are reused while neither body's
    // first rotation column has moved farther than the distance since they were captured.  The
    // next primal's force cache is left ready.
    void updateDual(uint32_t s)
    {
        const Body &a = bodies[mA[s]], &b = bodies[mB[s]];
        const int first = static_cast<int>(s) * MAXC;
        const m3 &fr = basis[s];
        const float mu = mFric[s];
        const f3 dqAL = a.pos - a.initLin, dqAA = qsub(a.q, a.initAng);
        const f3 dqBL = b.pos - b.initLin, dqBA = qsub(b.q, b.initAng);
        const m3 rotA = qmat(a.q), rotB = qmat(b.q);
        const bool cacheable = P.jacobian_rebuild_distance > 0.0f;
        const float rebuild2 = P.jacobian_rebuild_distance * P.jacobian_rebuild_distance;
        const f3 dA = rotA[0] - capA[s];
        const f3 dB = rotB[0] - capB[s];
        const bool usable = cacheable && (mFlags[s] & ARMS_CACHED) != 0 &&
                (g_plant == 23 ? length2(dA) >= rebuild2 && length2(dB) >= rebuild2
                               : length2(dA) <= rebuild2 && length2(dB) <= rebuild2);
        for (int i = 0; i < numC[s]; i++) {
            f3 rA, rB;
            if (usable) {
                rA = g_plant == 22 ? rBW[first + i] : rAW[first + i];
                rB = g_plant == 22 ? rAW[first + i] : rBW[first + i];
            } else {
                rA = mul(rotA, cont[first + i].rA);
                rB = mul(rotB, cont[first + i].rB);
                rAW[first + i] = rA;
                rBW[first + i] = rB;
            }
            const f3 br = (dqAL - dqBL) + cross(dqAA, rA) - cross(dqBA, rB);
            const f3 C = constraintError(C0[first + i], br, fr, 1 - P.alpha);
            resC[first + i] = g_plant == 28 ? C0[first + i][0] : C[0];
            f3 F = cmul(pen[first + i], C) + lam[first + i];
            const Clamp cone = clampForce(F, mu);
            lam[first + i] = F;
            const float beta = g_plant == 8 ? P.beta_lin : P.beta_lin * betaScale[first + i];
            if (F[0] < 0) pen[first + i][0] = smin(pen[first + i][0] + beta * std::fabs(C[0]), pmax());
            if (cone.scale <= cone.bounds) {
                pen[first + i][1] = smin(pen[first + i][1] + beta * std::fabs(C[1]), pmax());
                pen[first + i][2] = smin(pen[first + i][2] + beta * std::fabs(C[2]), pmax());
                stick[first + i] = length2d(C[1], C[2]) < STICK_THRESH;
            }
            f3 fNext = cmul(pen[first + i], C) + lam[first + i];
            clampForce(fNext, mu);
            fCache[first + i] = fNext;
        }
        if (usable) armHits++;
        if (!usable && cacheable && !(g_plant == 25 && (mFlags[s] & ARMS_CACHED))) {
            capA[s] = rotA[0];
            capB[s] = rotB[0];
            mFlags[s] |= ARMS_CACHED;
        }
        mFlags[s] |= PRIMAL_READY;
    }

    void decayPenalty(uint32_t s, float decay)
    {
        const int first = static_cast<int>(s) * MAXC;
        for (int i = 0; i < numC[s]; i++)
            pen[first + i] = lclamp3(pen[first + i] * decay, f3{PENALTY_MIN, PENALTY_MIN, PENALTY_MIN},
                                     f3{pmax(), pmax(), pmax()});
    }

    void rebase(uint32_t s)
    {
        const Body &a = bodies[mA[s]], &b = bodies[mB[s]];
        const int first = static_cast<int>(s) * MAXC;
        for (int i = 0; i < numC[s]; i++) {
            const f3 xA = a.pos + qrot(a.q, cont[first + i].rA);
            const f3 xB = b.pos + qrot(b.q, cont[first + i].rB);
            C0[first + i] = mul(basis[s], xA - xB) + f3{COLLISION_MARGIN, 0, 0};
            if (C0[first + i].x < -P.max_penetration) C0[first + i].x = -P.max_penetration;
        }
        mFlags[s] = static_cast<uint8_t>(mFlags[s] & ~PRIMAL_READY);   // manifold.cpp:460
    }

    // ------------------------------------------------------------ joints ----
    f3 jointLin(const Joint &j) const
    {
        const f3 pa = j.a >= 0 ? bodies[j.a].pos + qrot(bodies[j.a].q, j.rA) : j.rA;
        const Body &B = bodies[j.b];
        return pa - (B.pos + qrot(B.q, j.rB));
    }
    f3 jointAng(const Joint &j) const
    {
        const quat qa = j.a >= 0 ? bodies[j.a].q : quat{0, 0, 0, 1};
        return qsub(qa, bodies[j.b].q) * j.torqueArm;
    }
    static f3 warmPenalty(f3 p, float gamma, float stiff, float pmaxv)
    {
        return {smin(sclamp(p.x * gamma, PENALTY_MIN, pmaxv), stiff),
                smin(sclamp(p.y * gamma, PENALTY_MIN, pmaxv), stiff),
                smin(sclamp(p.z * gamma, PENALTY_MIN, pmaxv), stiff)};
    }
    // ---------------------------------------------------------------- generic joint ----
    // generic_joint.cpp
    static f3 frameAxis(int i) { return f3{i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f}; }
    quat frameOf(const Joint &j) const { return j.a >= 0 ? bodies[j.a].q : quat{0, 0, 0, 1}; }
    // :51-67
    static void removeTwist(quat &dev, int i, float &angle)
    {
        const f3 n = frameAxis(i);
        const float proj = dot(f3{dev.x, dev.y, dev.z}, n);
        const quat twist{dev.x * proj, dev.y * proj, dev.z * proj, dev.w};
        if (std::sqrt(qlength2(twist)) <= 1.0e-12f) { angle = 0.0f; return; }
        angle = 2.0f * atan2Poly(proj, dev.w);
        const float l2 = qlength2(twist);
        const quat inv{-twist.x / l2, -twist.y / l2, -twist.z / l2, twist.w / l2};
        dev = qmul(qnormalize(dev), inv);
    }
    // :69-90
    void angularValues(const Joint &j, float out[3]) const
    {
        const quat qA = frameOf(j);
        const quat relNow = qmul(qconj(qA), qnormalize(bodies[j.b].q));
        quat dev = qmul(qnormalize(j.restQ), qconj(relNow));
        for (int i = 0; i < 3; i++)
            if (j.axis[3 + i].mode != AVBD_AXIS_LOCKED) removeTwist(dev, i, out[i]);
        const float comp[3] = {dev.x, dev.y, dev.z};
        for (int i = 0; i < 3; i++)
            if (j.axis[3 + i].mode == AVBD_AXIS_LOCKED) out[i] = comp[i] * 2.0f;
    }
    // :92-96
    f3 linearValues(const Joint &j) const
    {
        const quat qA = frameOf(j);
        const f3 anchorA = j.a >= 0 ? bodies[j.a].pos + qrot(qA, j.rA) : j.rA;
        const f3 anchorB = bodies[j.b].pos + qrot(bodies[j.b].q, j.rB);
        return qrot(qconj(qA), anchorA - anchorB);
    }
    // :104-155
    static bool axisForce(const JointAxis &ax, float value, float alpha, float &C, float &F)
    {
        if (ax.mode == AVBD_AXIS_LOCKED) {
            C = value;
        } else if (ax.mode == AVBD_AXIS_LIMITED) {
            if (value > ax.upper) C = value - ax.upper;
            else if (value < ax.lower) C = value - ax.lower;
            else return false;
        } else {
            return false;
        }
        if (ax.mode == AVBD_AXIS_LOCKED) C -= ax.initial * alpha;
        float f = ax.penalty * C + ax.lambda;
        if (ax.mode == AVBD_AXIS_LIMITED) f = C > 0.0f ? smax(f, 0.0f) : smin(f, 0.0f);
        F = f;
        return true;
    }
    static m3 outerScaled(f3 a, f3 b, float s) { return m3{{a * b.x, a * b.y, a * b.z}} * s; }
    // :157-187
    static void stampAxis(const JointAxis &ax, Block &blk, float value, float alpha, f3 jLin, f3 jAng, float rate)
    {
        float C, F;
        if (axisForce(ax, value, alpha, C, F)) {
            blk.lhsLin += outerScaled(jLin, jLin, ax.penalty);
            blk.lhsAng += outerScaled(jAng, jAng, ax.penalty);
            blk.lhsCross += outerScaled(jAng, jLin, ax.penalty);
            blk.rhsLin += jLin * F;
            blk.rhsAng += jAng * F;
        }
        if (ax.spring > 0.0f) {
            const float Cs = value - ax.equilibrium;
            const float Fs = ax.spring * Cs - ax.damping * rate;
            blk.lhsLin += outerScaled(jLin, jLin, ax.spring);
            blk.lhsAng += outerScaled(jAng, jAng, ax.spring);
            blk.lhsCross += outerScaled(jAng, jLin, ax.spring);
            blk.rhsLin += jLin * Fs;
            blk.rhsAng += jAng * Fs;
        }
    }
    // :234-277
    void genericPrimal(const Joint &j, int body, Block &blk) const
    {
        const quat qA = frameOf(j);
        const f3 rAWorld = qrot(qA, j.rA);
        const Body &B = bodies[j.b];
        const f3 rBWorld = qrot(B.q, j.rB);
        const f3 offset = linearValues(j);
        float ang[3];
        angularValues(j, ang);
        const bool isA = body == j.a;
        const float sign = isA ? 1.0f : -1.0f;
        const f3 arm = isA ? rAWorld : rBWorld;
        const f3 vA = j.a >= 0 ? bodies[j.a].vel + cross(bodies[j.a].angVel, rAWorld) : f3{0, 0, 0};
        const f3 vB = B.vel + cross(B.angVel, rBWorld);
        const f3 relVel = vA - vB;
        const f3 wA = j.a >= 0 ? bodies[j.a].angVel : f3{0, 0, 0};
        const f3 wB = B.angVel;
        for (int i = 0; i < 3; i++) {
            const f3 axis = qrot(qA, frameAxis(i));
            const f3 jLin = axis * sign;
            const f3 jAng = cross(arm, jLin);
            stampAxis(j.axis[i], blk, offset[i], P.alpha, jLin, jAng, dot(relVel, axis));
            stampAxis(j.axis[3 + i], blk, ang[i], P.alpha, f3{0, 0, 0}, axis * sign, dot(wA - wB, axis));
        }
    }
    // :191-210
    float penaltyLimit(int body, f3 axis, bool angular) const
    {
        const Body &b = bodies[body];
        float effective;
        if (angular) {
            const f3 local = qrot(qconj(b.q), axis);
            effective = (local.x * local.x * b.moment.x + local.y * local.y * b.moment.y) + local.z * local.z * b.moment.z;
        } else {
            effective = b.mass;
        }
        if (effective <= 0.0f) return pmax();
        const float limit = effective * 0.25f * 0.25f / (P.dt * P.dt);
        return smin(limit, pmax());
    }
    // :279-319
    void genericDual(Joint &j) const
    {
        float ang[3];
        angularValues(j, ang);
        const f3 offset = linearValues(j);
        const quat qA = frameOf(j);
        for (int i = 0; i < 3; i++) {
            for (int k = 0; k < 2; k++) {
                JointAxis &ax = j.axis[3 * k + i];
                const float value = k == 0 ? offset[i] : ang[i];
                const float beta = k == 0 ? P.beta_lin : P.beta_ang;
                float C, F;
                if (!axisForce(ax, value, P.alpha, C, F)) {
                    if (ax.mode == AVBD_AXIS_LIMITED) ax.lambda = 0.0f;
                    continue;
                }
                ax.lambda = F;
                const f3 dir = qrot(qA, frameAxis(i));
                float cap = penaltyLimit(j.b, dir, k == 1);
                if (j.a >= 0) cap = smin(cap, penaltyLimit(j.a, dir, k == 1));
                ax.penalty = sclamp(ax.penalty + std::fabs(C) * beta, 0.0f, cap);
            }
        }
    }
    // :212-232
    void genericInit(Joint &j) const
    {
        float ang[3];
        angularValues(j, ang);
        const f3 offset = linearValues(j);
        for (int i = 0; i < 3; i++) {
            j.axis[i].initial = offset[i];
            j.axis[3 + i].initial = ang[i];
            for (int k = 0; k < 2; k++) {
                JointAxis &ax = j.axis[3 * k + i];
                ax.lambda *= P.alpha * P.gamma;
                ax.penalty = sclamp(ax.penalty * P.gamma, PENALTY_MIN, pmax());
            }
        }
    }

    bool jointInit(Joint &j)
    {
        if (j.kind == AVBD_FORCE_GENERIC) { genericInit(j); return true; }   // never breaks
        if (j.kind != AVBD_FORCE_JOINT) return true;   // Spring/IgnoreCollision::initialize
        j.C0Lin = jointLin(j);
        j.C0Ang = jointAng(j);
        j.lamLin = j.lamLin * P.alpha * P.gamma;
        j.lamAng = j.lamAng * P.alpha * P.gamma;
        j.penLin = warmPenalty(j.penLin, P.gamma, j.stiffLin, pmax());
        j.penAng = warmPenalty(j.penAng, P.gamma, j.stiffAng, pmax());
        return !j.broken;
    }
    static m3 gsBall(int k, f3 v)
    {
        m3 m = diagonal(-v[k], -v[k], -v[k]);
        m[k][0] += v[0]; m[k][1] += v[1]; m[k][2] += v[2];
        return m;
    }
    // spring.cpp:31-68
    void springPrimal(const Joint &j, int body, Block &blk) const
    {
        const Body &A = bodies[j.a], &B = bodies[j.b];
        const f3 pA = A.pos + qrot(A.q, j.rA);
        const f3 pB = B.pos + qrot(B.q, j.rB);
        const f3 d = pA - pB;
        const float dLen = length(d);
        if (dLen <= 1.0e-6f) return;
        const f3 n = d / dLen;
        const float C = dLen - j.rest;
        const float f = j.stiffLin * C;
        f3 jLin, jAng;
        if (body == j.a) {
            const f3 rWorld = qrot(A.q, j.rA);
            jLin = n;
            jAng = cross(rWorld, n);
        } else {
            const f3 rWorld = qrot(B.q, j.rB);
            jLin = -n;
            jAng = -cross(rWorld, n);
        }
        const f3 F = jLin * f, Tau = jAng * f;
        // linalg outerprod(a, b): column j is a * b[j]
        const m3 Kll = m3{{jLin * jLin.x, jLin * jLin.y, jLin * jLin.z}} * j.stiffLin;
        const m3 Kla = m3{{jAng * jLin.x, jAng * jLin.y, jAng * jLin.z}} * j.stiffLin;
        const m3 Kaa = m3{{jAng * jAng.x, jAng * jAng.y, jAng * jAng.z}} * j.stiffLin;
        blk.lhsLin += Kll;
        blk.lhsAng += Kaa;
        blk.lhsCross += Kla;
        blk.rhsLin += F;
        blk.rhsAng += Tau;
    }
    void forcePrimal(const Joint &j, int body, Block &blk) const
    {
        if (j.kind == AVBD_FORCE_JOINT) jointPrimal(j, body, blk);
        else if (j.kind == AVBD_FORCE_SPRING) springPrimal(j, body, blk);
        else if (j.kind == AVBD_FORCE_GENERIC) genericPrimal(j, body, blk);
    }
    // joint.cpp:80-160
    void jointPrimal(const Joint &j, int body, Block &blk) const
    {
        const bool isA = body == j.a;
        if (length2(j.penLin) > 0) {
            const m3 K = diagonal(j.penLin.x, j.penLin.y, j.penLin.z);
            f3 C = jointLin(j);
            if (std::isinf(j.stiffLin)) C -= j.C0Lin * (g_plant == 13 ? -P.alpha : P.alpha);
            const f3 F = mul(K, C) + j.lamLin;
            const m3 jLin = isA ? identity3() : -identity3();
            m3 jAng;
            if (isA) {
                const f3 r = -qrot(bodies[j.a].q, j.rA);
                jAng = m3{{f3{0, r.z, -r.y}, f3{-r.z, 0, r.x}, f3{r.y, -r.x, 0}}};
            } else {
                const f3 r = qrot(bodies[j.b].q, j.rB);
                jAng = m3{{f3{0, r.z, -r.y}, f3{-r.z, 0, r.x}, f3{r.y, -r.x, 0}}};
            }
            const m3 jLinT = transpose(jLin), jAngT = transpose(jAng);
            const m3 jAngTk = mul(jAngT, K);
            blk.lhsLin += mul(jLinT, K, jLin);
            blk.lhsAng += mul(jAngTk, jAng);
            blk.lhsCross += mul(jAngTk, jLin);
            const f3 r = isA ? qrot(bodies[j.a].q, j.rA) : -qrot(bodies[j.b].q, j.rB);
            const m3 H = (gsBall(0, r) * F[0] + gsBall(1, r) * F[1]) + gsBall(2, r) * F[2];
            blk.lhsAng += diagonal(H[0][0], H[1][1], H[2][2]);
            blk.rhsLin += mul(jLinT, F);
            blk.rhsAng += mul(jAngT, F);
        }
        if (length2(j.penAng) > 0) {
            const m3 K = diagonal(j.penAng.x, j.penAng.y, j.penAng.z);
            f3 C = jointAng(j);
            if (std::isinf(j.stiffAng)) C -= j.C0Ang * P.alpha;
            const f3 F = mul(K, C) + j.lamAng;
            const m3 jA = (isA ? identity3() : m3{{f3{0, 0, 0}, f3{0, -1, 0}, f3{0, 0, -1}}}) * j.torqueArm;
            blk.lhsAng += mul(transpose(jA), K, jA);
            blk.rhsAng += mul(transpose(jA), F);
        }
    }
    // joint.cpp:171-230
    void jointDual(Joint &j) const
    {
        if (j.kind == AVBD_FORCE_GENERIC) { genericDual(j); return; }
        if (j.kind != AVBD_FORCE_JOINT) return;   // Spring/IgnoreCollision::updateDual do nothing
        j.resLin = f3{0, 0, 0};
        if (length2(j.penLin) > 0) {
            const m3 K = diagonal(j.penLin.x, j.penLin.y, j.penLin.z);
            f3 C = jointLin(j);
            if (std::isinf(j.stiffLin)) {
                C -= j.C0Lin * (g_plant == 13 ? -P.alpha : P.alpha);
                if (g_plant != 12) j.lamLin = mul(K, C) + j.lamLin;
            }
            j.resLin = g_plant == 27 ? jointLin(j) : C;
            const float cap = g_plant == 14 ? 1e10f : pmax();
            j.penLin = lmin3(j.penLin + fabs3(C) * P.beta_lin,
                             lmin3(f3{j.stiffLin, j.stiffLin, j.stiffLin}, f3{cap, cap, cap}));
        }
        if (length2(j.penAng) > 0) {
            const m3 K = diagonal(j.penAng.x, j.penAng.y, j.penAng.z);
            f3 C = jointAng(j);
            if (std::isinf(j.stiffAng)) {
                C -= j.C0Ang * P.alpha;
                j.lamAng = mul(K, C) + j.lamAng;
            }
            const float cap = g_plant == 14 ? 1e10f : pmax();
            j.penAng = lmin3(j.penAng + fabs3(C) * P.beta_ang,
                             lmin3(f3{j.stiffAng, j.stiffAng, j.stiffAng}, f3{cap, cap, cap}));
        }
        if (length2(g_plant == 15 ? j.lamLin : j.lamAng) > (g_plant == 16 ? j.fracture : j.fracture * j.fracture)) {
            j.penLin = j.penAng = j.lamLin = j.lamAng = f3{0, 0, 0};
            j.broken = true;
        }
    }

    // ------------------------------------------------------------ step phases ----
    void warmstartForces()
    {
        std::vector<uint8_t> act(joints.size());
        for (size_t k = 0; k < joints.size(); k++) act[k] = jointInit(joints[k]);
        size_t w = 0;
        for (size_t k = 0; k < joints.size(); k++)
            if (act[k]) joints[w++] = joints[k];
        joints.resize(w);
        std::vector<uint8_t> mact(live.size());
        for (size_t i = 0; i < live.size(); i++) mact[i] = frozen(live[i]) || initialize(live[i]);
        size_t lw = 0;
        for (size_t i = 0; i < live.size(); i++) {
            // solver_solve.cpp:66-86: a parked manifold keeps its place in the live list
            if (!mact[i] && !(P.contact_grace_ticks > 0 && dormant[live[i]] > 0)) { freeList.push_back(live[i]); continue; }
            live[lw++] = live[i];
        }
        live.resize(lw);
        buildAdjacency();
    }

    void colourGraph()
    {
        const int n = static_cast<int>(updateOrder.size());
        colour.assign(n, kOverflow);
        colourOrder.clear();
        overflowOrder.clear();
        for (int i = 0; i < n; i++) {
            const int body = updateOrder[i];
            uint8_t used[kMaxColours] = {};
            for (uint32_t e = adjStart[body]; e < adjStart[body + 1]; e++) {
                const int o = otherOf(adj[e], body);
                if (o < 0 || bodies[o].updateIndex < 0 || bodies[o].updateIndex >= i) continue;
                const uint8_t c = colour[bodies[o].updateIndex];
                if (c < kMaxColours) used[c] = 1;
            }
            uint8_t c = kOverflow;
            for (int k = 0; k < kMaxColours; k++)
                if (!used[k]) { c = static_cast<uint8_t>(k); break; }
            colour[i] = c;
        }
        colourBegins.assign(kMaxColours + 1, 0);
        for (int i = 0; i < n; i++) {
            if ((P.contact_only_overflow || g_plant == 19) && bodies[updateOrder[i]].contactOnly) {
                colour[i] = kOverflow;
                overflowOrder.push_back(i);
            }
            if (colour[i] != kOverflow) colourBegins[colour[i] + 1]++;
        }
        for (int c = 0; c < kMaxColours; c++) colourBegins[c + 1] += colourBegins[c];
        colourOrder.resize(n - overflowOrder.size());
        int cursor[kMaxColours];
        for (int c = 0; c < kMaxColours; c++) cursor[c] = colourBegins[c];
        for (int i = 0; i < n; i++)
            if (colour[i] < kMaxColours) colourOrder[cursor[colour[i]]++] = i;
        colourCount = 0;
        for (int c = 0; c < kMaxColours; c++)
            if (colourBegins[c + 1] > colourBegins[c]) colourCount = c + 1;
    }

    // solver_solve.cpp:264-339
    void selectUpdateBodies()
    {
        updateOrder.clear();
        if (P.sleep_frames > 0) {
            // a sleeper sharing a force with a moving, awake movable neighbour wakes (list order);
            // the island pass that follows in the reference (solver_solve.cpp:177-258) only ever
            // re-sleeps islands whose members all sleep already, so it changes nothing
            const float lin2 = P.sleep_linear_threshold * P.sleep_linear_threshold;
            const float ang2 = P.sleep_angular_threshold * P.sleep_angular_threshold;
            for (size_t i = 0; i < bodies.size(); i++) {
                Body &b = bodies[i];
                if (!b.sleeping || b.sleepMode != AVBD_SLEEP_IDLE) continue;
                for (uint32_t e = adjStart[i]; e < adjStart[i + 1]; e++) {
                    const int o = otherOf(adj[e], static_cast<int>(i));
                    if (o < 0) continue;
                    const Body &x = bodies[o];
                    if (x.mass > 0 && !x.sleeping && (length2(x.vel) > lin2 || length2(x.angVel) > ang2)) {
                        b.sleeping = false;
                        b.stillFrames = 0;
                        break;
                    }
                }
            }
        }
        for (size_t i = 0; i < bodies.size(); i++) {
            Body &b = bodies[i];
            b.inUpdate = false; b.updateIndex = -1; b.contactOnly = true;
            if (b.mass > 0) {
                bool update = true;
                if (b.sleepMode == AVBD_SLEEP_IDLE) {
                    if (b.sleeping) update = false;
                } else if (b.sleepMode == AVBD_SLEEP_START_ASLEEP) {
                    if (!b.sleeping && length2(b.vel) == 0.0f && adjStart[i + 1] == adjStart[i]) update = false;
                    else b.sleeping = false;
                }
                if (update) {
                    b.updateIndex = static_cast<int>(updateOrder.size());
                    updateOrder.push_back(static_cast<int>(i));
                    b.inUpdate = true;
                }
            }
        }
        dLin.assign(updateOrder.size(), f3{0, 0, 0});
        dAng.assign(updateOrder.size(), f3{0, 0, 0});
        for (const Joint &j : joints) {
            if (j.a >= 0) bodies[j.a].contactOnly = false;
            bodies[j.b].contactOnly = false;
        }
        colourGraph();
    }

    static void maskLocked(uint8_t locks, f3 &v)
    {
        if (locks)
            for (int k = 0; k < 3; k++)
                if (locks & (1 << k)) v[k] = 0.0f;
    }

    void warmstartBody(Body &b)
    {
        const float dt = P.dt;
        maskLocked(b.lockLin, b.vel);
        if (g_plant != 5) maskLocked(b.lockAng, b.angVel);
        const float g = (b.lockLin & 0x4) && g_plant != 3 ? 0.0f : b.gravity;
        b.inertLin = b.pos + b.vel * dt;
        if (b.mass > 0) b.inertLin += f3{0, 0, g} * (dt * dt);
        b.inertAng = qadd(b.q, b.angVel * dt);
        const f3 accel = (b.vel - b.prevVel) / dt;
        const float accelExt = accel.z * std::copysign(1.0f, g);
        float w = sclamp(accelExt / std::fabs(g), 0.0f, 1.0f);
        if (!std::isfinite(w)) w = 0.0f;
        b.initLin = b.pos;
        b.initAng = b.q;
        if (b.mass > 0) {
            b.pos = (b.pos + b.vel * dt) + f3{0, 0, g} * (w * dt * dt);
            b.q = qadd(b.q, b.angVel * dt);
        }
    }

    // solver_solve.cpp:143-176: BDF1 velocity (not for sleepers), then idle detection on
    // the tick's last substep
    void finishBody(Body &b, bool idle)
    {
        const float dt = P.dt;
        b.prevVel = b.vel;
        if (b.mass > 0 && !b.sleeping) {
            b.vel = (b.pos - b.initLin) / dt;
            b.angVel = qsub(b.q, b.initAng) / dt;
        }
        if (g_plant != 4) maskLocked(b.lockLin, b.vel);
        if (g_plant != 4 && g_plant != 5) maskLocked(b.lockAng, b.angVel);
        if (idle && P.sleep_frames > 0 && b.mass > 0 && !b.sleeping && b.sleepMode == AVBD_SLEEP_IDLE) {
            const float lin2 = P.sleep_linear_threshold * P.sleep_linear_threshold;
            const float ang2 = P.sleep_angular_threshold * P.sleep_angular_threshold;
            if (length2(b.vel) < lin2 && length2(b.angVel) < ang2) ++b.stillFrames;
            else b.stillFrames = 0;
            if (b.stillFrames >= P.sleep_frames) {
                b.sleeping = true;
                b.vel = f3{0, 0, 0};
                b.angVel = f3{0, 0, 0};
            }
        }
    }

    // A body's block: inertia plus its forces in adjacency order; the colour sweep goes through
    // updatePrimal, the overflow solve stamps from the phase-A cache (solver_step.cpp:215-226,
    // 252-262).
    Block assemble(int body, bool fromCache)
    {
        const Body &b = bodies[body];
        Block blk = makeBlock(b, P.dt * P.dt);
        for (uint32_t e = adjStart[body]; e < adjStart[body + 1]; e++) {
            const int ref = adj[e];
            if (ref >= 0) {
                if (fromCache) stampCached(static_cast<uint32_t>(ref), body, blk);
                else updatePrimal(static_cast<uint32_t>(ref), body, blk);
            } else {
                forcePrimal(joints[-ref - 1], body, blk);
            }
        }
        return blk;
    }

    // manifold.cpp:672-692: the convergence probe's sum over one manifold's contacts (normal
    // rows with penetration; speculative gaps skipped)
    void accumulatePenetration(uint32_t s, float &total, int &count) const
    {
        const bool speculative = P.speculative_margin > 0.0f;
        const int first = static_cast<int>(s) * MAXC;
        for (int i = 0; i < numC[s]; i++) {
            if (speculative && C0[first + i].x > 0.0f) continue;
            const float penetration = -C0[first + i].x;
            if (penetration > 0.0f) { total += penetration; count++; }
        }
    }

    void iterate(int target)
    {
        const float rounds = P.newton_ratio * static_cast<float>(target);
        const int newtonRounds = std::clamp(
                static_cast<int>(g_plant == 9 ? std::floor(rounds) : std::ceil(rounds)), 1, target);
        float prevResidual = 0.0f;   // solver_step.cpp:88-89, per substep
        bool converged = false;
        for (int it = 0; it < target; it++) {
            const bool newton = it < newtonRounds;
            for (int c = 0; newton && c < colourCount; c++) {
                for (int k = colourBegins[c]; k < colourBegins[c + 1]; k++) {
                    const int body = updateOrder[colourOrder[k]];
                    const Block blk = assemble(body, false);
                    f3 dl, da;
                    solveDelta(blk, dl, da);
                    bodies[body].pos = bodies[body].pos + dl;
                    bodies[body].q = qadd(bodies[body].q, da);
                }
            }
            const int ov = newton ? static_cast<int>(overflowOrder.size()) : 0;
            if (ov > 0) {
                // solver_step.cpp:235-246: phase A caches the manifolds touching an overflow body
                for (uint32_t s : live) {
                    const Body &A = bodies[mA[s]], &B = bodies[mB[s]];
                    const bool aIn = A.inUpdate && colour[A.updateIndex] == kOverflow;
                    const bool bIn = B.inUpdate && colour[B.updateIndex] == kOverflow;
                    if (aIn || bIn) preCompute(s);
                }
                for (int k = 0; k < ov; k++) {
                    const int i = overflowOrder[k];
                    const Block blk = assemble(updateOrder[i], true);
                    solveDelta(blk, dLin[i], dAng[i]);
                }
                for (int k = 0; k < ov; k++) {
                    const int i = overflowOrder[k];
                    Body &b = bodies[updateOrder[i]];
                    b.pos = b.pos + dLin[i] * P.jacobi_relaxation;
                    b.q = qadd(b.q, dAng[i] * P.jacobi_relaxation);
                }
            }
            const float decay = newton ? 1.0f : P.stiffness_decay;
            for (uint32_t s : live) {
                if (frozen(s)) {
                    for (int i = 0; i < numC[s]; i++) resC[static_cast<size_t>(s) * MAXC + i] = 0.0f;
                    continue;
                }
                if (decay < 1.0f && g_plant != 10 && g_plant != 11) decayPenalty(s, decay);
                updateDual(s);
                if (decay < 1.0f && g_plant == 10) decayPenalty(s, decay);
            }
            for (Joint &j : joints) jointDual(j);
            // solver_step.cpp:299-314: after a round (not the last), the mean penetration over
            // the live manifolds against the previous round's; a change below the threshold
            // (or a worsening) ends the iterations
            if (P.convergence_threshold > 0.0f && it + 1 < target) {
                float total = 0.0f;
                int count = 0;
                for (uint32_t s : live) accumulatePenetration(s, total, count);
                const float residual = count > 0 ? total / static_cast<float>(count) : 0.0f;
                converged = (g_plant == 21 ? residual - prevResidual : prevResidual - residual) < P.convergence_threshold;
                prevResidual = residual;
            }
            if (converged) break;
        }
    }

    // solver_projection.cpp:33-53
    static float angCompliance(const Body &b, f3 t)
    {
        float sum = 0.0f;
        for (int k = 0; k < 3; k++)
            if (b.moment[k] > 0.0f) sum += t[k] * t[k] / b.moment[k];
        return sum;
    }
    static f3 angResponse(const Body &b, f3 t)
    {
        f3 o{0, 0, 0};
        for (int k = 0; k < 3; k++)
            if (b.moment[k] > 0.0f) o[k] = t[k] / b.moment[k];
        return qrot(b.q, o);
    }

    // solver_projection.cpp:67-281
    void projectContacts(bool positions, float brakeWindow)
    {
        if (live.empty()) return;
        const float window = brakeWindow > 0.0f ? brakeWindow : collisionDt;
        std::vector<uint32_t> active;
        for (uint32_t s : live) {
            if (numC[s] == 0) continue;
            const Body &A = bodies[mA[s]], &B = bodies[mB[s]];
            if (!(A.mass > 0.0f && A.inUpdate) && !(B.mass > 0.0f && B.inUpdate)) continue;
            active.push_back(s);
        }
        if (active.empty()) return;
        if (P.speculative_margin > 0.0f) {
            for (uint32_t s : active) {
                Body &A = bodies[mA[s]], &B = bodies[mB[s]];
                const int first = static_cast<int>(s) * MAXC;
                const bool moveA = A.mass > 0.0f && A.inUpdate, moveB = B.mass > 0.0f && B.inUpdate;
                const float imA = moveA ? 1.0f / A.mass : 0.0f, imB = moveB ? 1.0f / B.mass : 0.0f;
                const f3 n = row(basis[s], 0);
                f3 pRA{0, 0, 0}, pRB{0, 0, 0};
                float verr = 0.0f;
                int braked = 0;
                for (int k = 0; k < numC[s]; k++) {
                    const f3 rA = qrot(A.q, cont[first + k].rA);
                    const f3 rB = qrot(B.q, cont[first + k].rB);
                    const float gap = dot(n, (A.pos + rA) - (B.pos + rB));
                    if (!(gap > 0.0f)) continue;
                    const f3 vA = A.vel + cross(A.angVel, rA), vB = B.vel + cross(B.angVel, rB);
                    const float vn = dot(vA - vB, n);
                    const float target = -gap / window;
                    if (vn < target) { pRA += rA; pRB += rB; verr += target - vn; braked++; }
                }
                if (braked == 0) continue;
                const float invN = 1.0f / braked;
                const f3 tA = moveA ? qrot(qconj(A.q), cross(pRA * invN, n)) : f3{0, 0, 0};
                const f3 tB = moveB ? qrot(qconj(B.q), cross(pRB * invN, n)) : f3{0, 0, 0};
                const float K = imA + imB + (moveA ? angCompliance(A, tA) : 0.0f) + (moveB ? angCompliance(B, tB) : 0.0f);
                const float j = verr * invN / K;
                if (moveA) {
                    f3 dv = n * (imA * j), dw = angResponse(A, tA) * j;
                    maskLocked(A.lockLin, dv); maskLocked(A.lockAng, dw);
                    A.vel += dv; A.angVel += dw;
                }
                if (moveB) {
                    f3 dv = n * (imB * j), dw = angResponse(B, tB) * j;
                    maskLocked(B.lockLin, dv); maskLocked(B.lockAng, dw);
                    B.vel -= dv; B.angVel -= dw;
                }
            }
        }
        if (!positions || P.projection_iterations <= 0) return;
        const size_t nu = updateOrder.size();
        std::vector<f3> pl(nu), pa(nu);
        std::vector<int> cnt(nu);
        const float tol = smax(P.projection_slop * 0.01f, 1.0e-6f);
        for (int iter = 0; iter < P.projection_iterations; iter++) {
            std::fill(cnt.begin(), cnt.end(), 0);
            std::fill(pl.begin(), pl.end(), f3{0, 0, 0});
            std::fill(pa.begin(), pa.end(), f3{0, 0, 0});
            float maxDepth = 0.0f;
            for (uint32_t s : active) {
                const Body &A = bodies[mA[s]], &B = bodies[mB[s]];
                const int first = static_cast<int>(s) * MAXC;
                const bool moveA = A.mass > 0.0f && A.inUpdate, moveB = B.mass > 0.0f && B.inUpdate;
                const float imA = moveA ? 1.0f / A.mass : 0.0f, imB = moveB ? 1.0f / B.mass : 0.0f;
                const f3 n = row(basis[s], 0);
                f3 pRA{0, 0, 0}, pRB{0, 0, 0};
                float depthSum = 0.0f;
                int corr = 0;
                const m3 rotA = qmat(A.q), rotB = qmat(B.q);
                for (int k = 0; k < numC[s]; k++) {
                    const f3 rA = mul(rotA, cont[first + k].rA), rB = mul(rotB, cont[first + k].rB);
                    const float raw = -dot(n, (A.pos + rA) - (B.pos + rB));
                    const float depth = raw - P.projection_slop;
                    if (depth > 0.0f) { pRA += rA; pRB += rB; depthSum += depth; corr++; }
                }
                if (corr == 0) continue;
                const float invN = 1.0f / corr;
                maxDepth = smax(maxDepth, depthSum * invN);
                const f3 tA = moveA ? qrot(qconj(A.q), cross(pRA * invN, n)) : f3{0, 0, 0};
                const f3 tB = moveB ? qrot(qconj(B.q), cross(pRB * invN, n)) : f3{0, 0, 0};
                const float K = imA + imB + (moveA ? angCompliance(A, tA) : 0.0f) + (moveB ? angCompliance(B, tB) : 0.0f);
                const float dp = P.projection_beta * depthSum * invN / K;
                if (moveA) {
                    const int i = A.updateIndex;
                    cnt[i]++; pl[i] += n * (imA * dp); pa[i] += angResponse(A, tA) * dp;
                }
                if (moveB) {
                    const int i = B.updateIndex;
                    cnt[i]++; pl[i] -= n * (imB * dp); pa[i] -= angResponse(B, tB) * dp;
                }
            }
            for (size_t i = 0; i < nu; i++) {
                Body &b = bodies[updateOrder[i]];
                const float inv = 1.0f / std::max(cnt[i], 1);
                f3 dl = pl[i] * inv, da = pa[i] * inv;
                if (g_plant == 18) { dl = pl[i]; da = pa[i]; }
                maskLocked(b.lockLin, dl); maskLocked(b.lockAng, da);
                const float cap = P.projection_max_correction;
                if (cap > 0.0f && g_plant != 17) {
                    const float moved = length(dl);
                    if (moved > cap) dl = dl * (cap / moved);
                    const float turned = length(da);
                    if (turned > cap) da = da * (cap / turned);
                }
                b.pos = b.pos + dl;
                b.q = qadd(b.q, da);
            }
            if (maxDepth <= tol) break;
        }
    }

    void rebaseAll()
    {
        for (uint32_t s : live) rebase(s);
    }

    void runSubstep(int target, bool idle)
    {
        for (Body &b : bodies) warmstartBody(b);
        iterate(target);
        for (Body &b : bodies) finishBody(b, idle);
    }

    void step()
    {
        collisionDt = P.dt;
        broadPhase();
        warmstartForces();
        selectUpdateBodies();
        projectContacts(false, 0.0f);
        const int target = P.iterations;
        roundsThisStep = std::max(1, target);
        runSubstep(target, true);
        projectContacts(P.position_projection != 0, 0.0f);
        rebaseAll();
    }

    void stepInterleaved(int substeps, int roundsPer)
    {
        if (substeps <= 1) { step(); return; }
        collisionDt = P.dt * substeps;
        broadPhase();
        warmstartForces();
        selectUpdateBodies();
        const int rounds = roundsPer > 0 ? std::max(1, roundsPer)
                                         : std::max(2, (P.iterations + substeps - 1) / substeps);
        roundsThisStep = rounds;
        for (int sub = 0; sub < substeps; ++sub) {
            if (sub > 0) rebaseAll();
            projectContacts(false, P.dt * (substeps - sub));
            runSubstep(rounds, sub == substeps - 1);
        }
        projectContacts(P.position_projection != 0, collisionDt);
        rebaseAll();
    }
};

int checkParams(const avbd_params *p)
{
    if (!p) return AVBD_ERR_INVALID;
    return AVBD_OK;
}

f3 load3(const float *p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
void store3(float *p, int i, f3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

} // namespace

struct oracle_world {
    World w;
};

extern "C" {

oracle_world *oracle_create(const avbd_params *params)
{
    if (checkParams(params) != AVBD_OK) return nullptr;
    auto *o = new oracle_world();
    o->w.P = *params;
    o->w.collisionDt = params->dt;
    return o;
}

void oracle_destroy(oracle_world *o) { delete o; }

int oracle_set_params(oracle_world *o, const avbd_params *params)
{
    const int st = checkParams(params);
    if (st != AVBD_OK) return st;
    o->w.P = *params;
    return AVBD_OK;
}

int oracle_upload(oracle_world *o, const avbd_bodies *b, const avbd_joints *j)
{
    if (!o || !b || b->count < 0) return AVBD_ERR_INVALID;
    World &w = o->w;
    w.bodies.assign(b->count, Body{});
    for (int i = 0; i < b->count; i++) {
        Body &x = w.bodies[i];
        x.size = load3(b->size, i);
        x.mass = b->mass[i];
        x.moment = load3(b->moment, i);
        x.friction = b->friction[i];
        x.radius = b->radius[i];
        x.gravity = b->gravity ? b->gravity[i] : w.P.gravity;
        x.layer = b->layer ? b->layer[i] : 1u;
        x.mask = b->mask ? b->mask[i] : 1u;
        x.lockLin = b->lock_linear ? b->lock_linear[i] : 0;
        x.lockAng = b->lock_angular ? b->lock_angular[i] : 0;
        x.shape = b->shape ? b->shape[i] : static_cast<uint8_t>(AVBD_SHAPE_BOX);
        if (x.shape > AVBD_SHAPE_CYLINDER) return AVBD_ERR_INVALID;
        x.sleepMode = b->sleep_mode ? b->sleep_mode[i] : static_cast<uint8_t>(AVBD_SLEEP_NEVER);
        if (x.sleepMode > AVBD_SLEEP_START_ASLEEP) return AVBD_ERR_INVALID;
        x.sleeping = b->sleeping ? b->sleeping[i] != 0 : false;
        x.stillFrames = b->still_frames ? b->still_frames[i] : 0;
        x.pos = load3(b->position, i);
        x.q = quat{b->rotation[4 * i], b->rotation[4 * i + 1], b->rotation[4 * i + 2], b->rotation[4 * i + 3]};
        x.vel = load3(b->velocity, i);
        x.angVel = load3(b->angular_velocity, i);
        x.prevVel = b->prev_velocity ? load3(b->prev_velocity, i) : x.vel;
        x.initLin = x.inertLin = x.pos;
        x.initAng = x.inertAng = x.q;
    }
    w.joints.clear();
    if (j && j->count > 0) {
        w.joints.resize(j->count);
        for (int k = 0; k < j->count; k++) {
            Joint &x = w.joints[k];
            x.a = j->body_a[k]; x.b = j->body_b[k];
            x.rA = load3(j->r_a, k); x.rB = load3(j->r_b, k);
            x.stiffLin = j->stiffness_lin[k]; x.stiffAng = j->stiffness_ang[k];
            x.fracture = j->fracture[k]; x.torqueArm = j->torque_arm[k];
            x.C0Lin = j->c0_lin ? load3(j->c0_lin, k) : f3{0, 0, 0};
            x.C0Ang = j->c0_ang ? load3(j->c0_ang, k) : f3{0, 0, 0};
            x.penLin = j->penalty_lin ? load3(j->penalty_lin, k) : f3{0, 0, 0};
            x.penAng = j->penalty_ang ? load3(j->penalty_ang, k) : f3{0, 0, 0};
            x.lamLin = j->lambda_lin ? load3(j->lambda_lin, k) : f3{0, 0, 0};
            x.lamAng = j->lambda_ang ? load3(j->lambda_ang, k) : f3{0, 0, 0};
            x.broken = j->broken ? j->broken[k] != 0 : false;
            x.kind = j->kind ? j->kind[k] : AVBD_FORCE_JOINT;
            if (x.kind > AVBD_FORCE_GENERIC) return AVBD_ERR_INVALID;
            if (x.kind == AVBD_FORCE_GENERIC) {
                if (!j->rest_rotation || !j->axis_mode || !j->axis_lower || !j->axis_upper || !j->axis_spring ||
                    !j->axis_equilibrium || !j->axis_damping)
                    return AVBD_ERR_INVALID;
                x.restQ = quat{j->rest_rotation[4 * k], j->rest_rotation[4 * k + 1], j->rest_rotation[4 * k + 2],
                               j->rest_rotation[4 * k + 3]};
                for (int a = 0; a < 6; a++) {
                    JointAxis &ax = x.axis[a];
                    const int i = 6 * k + a;
                    ax.mode = j->axis_mode[i];
                    if (ax.mode > AVBD_AXIS_LIMITED) return AVBD_ERR_INVALID;
                    ax.lower = j->axis_lower[i]; ax.upper = j->axis_upper[i]; ax.spring = j->axis_spring[i];
                    ax.equilibrium = j->axis_equilibrium[i]; ax.damping = j->axis_damping[i];
                    ax.lambda = j->axis_lambda ? j->axis_lambda[i] : 0.0f;
                    ax.penalty = j->axis_penalty ? j->axis_penalty[i] : 0.0f;
                    ax.initial = j->axis_initial ? j->axis_initial[i] : 0.0f;
                }
            }
            if (x.kind == AVBD_FORCE_SPRING) {
                if (x.a < 0 || !j->rest) return AVBD_ERR_INVALID;
                x.rest = j->rest[k];
            }
            if (x.kind != AVBD_FORCE_JOINT) x.broken = false;
            if (x.b < 0 || x.b >= b->count || x.a >= b->count) return AVBD_ERR_INVALID;
        }
    }
    w.mA.clear(); w.mB.clear(); w.basis.clear(); w.numC.clear(); w.mFric.clear(); w.cont.clear();
    w.dormant.clear(); w.saved.clear();
    w.C0.clear(); w.pen.clear(); w.lam.clear(); w.stick.clear(); w.betaScale.clear();
    w.mFlags.clear(); w.capA.clear(); w.capB.clear(); w.rAW.clear(); w.rBW.clear(); w.fCache.clear(); w.resC.clear();
    w.live.clear(); w.freeList.clear();
    return AVBD_OK;
}

int oracle_write_bodies(oracle_world *o, const avbd_bodies *b)
{
    if (!o || !b || b->count != static_cast<int32_t>(o->w.bodies.size())) return AVBD_ERR_INVALID;
    for (int i = 0; i < b->count; i++) {
        Body &x = o->w.bodies[i];
        if (b->position) x.pos = load3(b->position, i);
        if (b->rotation)
            x.q = quat{b->rotation[4 * i], b->rotation[4 * i + 1], b->rotation[4 * i + 2], b->rotation[4 * i + 3]};
        if (b->velocity) x.vel = load3(b->velocity, i);
        if (b->angular_velocity) x.angVel = load3(b->angular_velocity, i);
        if (b->prev_velocity) x.prevVel = load3(b->prev_velocity, i);
        if (b->sleeping) x.sleeping = b->sleeping[i] != 0;
        if (b->still_frames) x.stillFrames = b->still_frames[i];
    }
    return AVBD_OK;
}

int oracle_step(oracle_world *o)
{
    o->w.step();
    return AVBD_OK;
}

int oracle_step_interleaved(oracle_world *o, int substeps, int rounds)
{
    o->w.stepInterleaved(substeps, rounds);
    return AVBD_OK;
}

int oracle_download(oracle_world *o, avbd_bodies *b, avbd_joints *j)
{
    const World &w = o->w;
    if (b) {
        b->count = static_cast<int32_t>(w.bodies.size());
        for (size_t i = 0; i < w.bodies.size(); i++) {
            const Body &x = w.bodies[i];
            const int k = static_cast<int>(i);
            if (b->position) store3(b->position, k, x.pos);
            if (b->rotation) {
                b->rotation[4 * k] = x.q.x; b->rotation[4 * k + 1] = x.q.y;
                b->rotation[4 * k + 2] = x.q.z; b->rotation[4 * k + 3] = x.q.w;
            }
            if (b->velocity) store3(b->velocity, k, x.vel);
            if (b->angular_velocity) store3(b->angular_velocity, k, x.angVel);
            if (b->prev_velocity) store3(b->prev_velocity, k, x.prevVel);
            if (b->sleeping) b->sleeping[k] = x.sleeping ? 1 : 0;
            if (b->still_frames) b->still_frames[k] = x.stillFrames;
        }
    }
    if (j) {
        j->count = static_cast<int32_t>(w.joints.size());
        for (size_t k = 0; k < w.joints.size(); k++) {
            const Joint &x = w.joints[k];
            const int i = static_cast<int>(k);
            if (j->body_a) j->body_a[i] = x.a;
            if (j->body_b) j->body_b[i] = x.b;
            if (j->c0_lin) store3(j->c0_lin, i, x.C0Lin);
            if (j->c0_ang) store3(j->c0_ang, i, x.C0Ang);
            if (j->penalty_lin) store3(j->penalty_lin, i, x.penLin);
            if (j->penalty_ang) store3(j->penalty_ang, i, x.penAng);
            if (j->lambda_lin) store3(j->lambda_lin, i, x.lamLin);
            if (j->lambda_ang) store3(j->lambda_ang, i, x.lamAng);
            if (j->broken) j->broken[i] = x.broken ? 1 : 0;
            for (int a = 0; a < 6; a++) {
                if (j->axis_lambda) j->axis_lambda[6 * i + a] = x.axis[a].lambda;
                if (j->axis_penalty) j->axis_penalty[6 * i + a] = x.axis[a].penalty;
                if (j->axis_initial) j->axis_initial[6 * i + a] = x.axis[a].initial;
            }
        }
    }
    return AVBD_OK;
}

int oracle_manifold_count(oracle_world *o) { return static_cast<int>(o->w.live.size()); }

// query_shape.cpp:258-274 on one shape built from a body's size (makeShape's fields).
int oracle_intersect_ray(int type, const float pos[3], const float rot[4], const float size[3], const float origin[3],
                         const float dir[3], float maxT, float *t, float point[3], float normal[3], uint8_t *inside)
{
    const RShape s = makeRShape(type, f3{pos[0], pos[1], pos[2]}, quat{rot[0], rot[1], rot[2], rot[3]},
                                f3{size[0], size[1], size[2]});
    RayHit h{};
    if (!intersectRay(s, f3{origin[0], origin[1], origin[2]}, f3{dir[0], dir[1], dir[2]}, maxT, h)) return 0;
    *t = h.t;
    store3(point, 0, h.point);
    store3(normal, 0, h.normal);
    *inside = h.inside ? 1 : 0;
    return 1;
}

// Solver::collectAABB (solver_query.cpp:158-171): list indices, in list order, of the bodies
// whose tight bounds (rigid.cpp:83-99) overlap [min, max] and whose layer passes the mask.
int oracle_collect_aabb(oracle_world *o, const float mn[3], const float mx[3], uint32_t mask, int32_t *out, int cap)
{
    const World &w = o->w;
    int n = 0;
    for (size_t i = 0; i < w.bodies.size(); i++) {
        const Body &b = w.bodies[i];
        if ((b.layer & mask) == 0) continue;
        const f3 wx = w.tightHalf(b);
        const f3 lo = b.pos - wx, hi = b.pos + wx;
        if (!(lo.x > mx[0] || lo.y > mx[1] || lo.z > mx[2] || mn[0] > hi.x || mn[1] > hi.y || mn[2] > hi.z)) {
            if (n < cap) out[n] = static_cast<int32_t>(i);
            n++;
        }
    }
    return n;
}

// The bodies part of AVBDDirectSpaceState3D::_intersect_ray (avbd_direct_space_state3d.cpp:
// 103-195) for n segments from -> to: candidates from collectAABB over the segment's box padded
// by 1e-4, in list order; the nearest hit (the earlier candidate on a tie); a hit from inside
// (only when hit_from_inside) reports t = 0 at the origin with normal -dir and ends the search.
int oracle_raycast(oracle_world *o, int n, const float *from, const float *to, const uint32_t *masks,
                   int hitFromInside, int32_t *body, float *tOut, float *point, float *normal, uint8_t *inside)
{
    World &w = o->w;
    std::vector<int32_t> cand(w.bodies.size());
    for (int r = 0; r < n; r++) {
        body[r] = -1; tOut[r] = 0.0f; inside[r] = 0;
        store3(point, r, f3{0, 0, 0}); store3(normal, r, f3{0, 0, 0});
        const f3 origin = load3(from, r), end = load3(to, r);
        f3 dir = end - origin;
        const float maxD = length(dir);
        if (maxD <= 1.0e-6f) continue;
        dir = dir * (1.0f / maxD);
        const float pad = 1.0e-4f;
        const float mn[3] = {smin(origin.x, end.x) - pad, smin(origin.y, end.y) - pad, smin(origin.z, end.z) - pad};
        const float mx[3] = {smax(origin.x, end.x) + pad, smax(origin.y, end.y) + pad, smax(origin.z, end.z) + pad};
        const int nc = oracle_collect_aabb(o, mn, mx, masks ? masks[r] : 0xFFFFFFFFu, cand.data(),
                                           static_cast<int>(cand.size()));
        bool valid = false;
        for (int k = 0; k < nc; k++) {
            const Body &b = w.bodies[cand[k]];
            const RShape s = makeRShape(b.shape, b.pos, b.q, b.size);
            RayHit h{};
            if (!intersectRay(s, origin, dir, maxD, h)) continue;
            if (h.inside && !hitFromInside) continue;
            if (h.inside) {
                body[r] = cand[k]; tOut[r] = 0.0f; store3(point, r, origin); store3(normal, r, dir * -1.0f);
                inside[r] = 1;
                valid = true;
                break;
            }
            if (h.t > maxD || (valid && h.t >= tOut[r])) continue;
            body[r] = cand[k]; tOut[r] = h.t; store3(point, r, h.point); store3(normal, r, h.normal);
            inside[r] = 0;
            valid = true;
        }
    }
    return AVBD_OK;
}

// query::containsPoint on one shape built from a body's size (makeShape's fields).
int oracle_contains_point(int type, const float pos[3], const float rot[4], const float size[3], const float point[3],
                          float margin)
{
    const RShape s = makeRShape(type, f3{pos[0], pos[1], pos[2]}, quat{rot[0], rot[1], rot[2], rot[3]},
                                f3{size[0], size[1], size[2]});
    return containsPoint(s, f3{point[0], point[1], point[2]}, margin) ? 1 : 0;
}

// query_shape.cpp:280-323: the closest point of a shape's volume to a point (the point itself
// inside a box or a cylinder side's slab; the surface otherwise)
f3 closestPoint(const RShape &s, f3 point)
{
    const f3 p = qrot(qconj(s.q), point - s.c);
    f3 local = p;
    if (s.type == AVBD_SHAPE_SPHERE) {
        const float len = length(p);
        local = len > 1.0e-6f ? p * (s.r / len) : f3{s.r, 0, 0};
    } else if (s.type == AVBD_SHAPE_CYLINDER) {
        const float hh = s.hh;
        const float radial = std::sqrt(p.x * p.x + p.y * p.y);
        const float clampedZ = sclamp(p.z, -hh, hh);
        if (radial > s.r) {
            const float scale = s.r / radial;
            local = f3{p.x * scale, p.y * scale, clampedZ};
        } else if (radial <= 1.0e-6f) {
            const float capDistance = hh - std::fabs(p.z);
            if (capDistance < s.r) local = f3{0, 0, p.z < 0.0f ? -hh : hh};
            else local = f3{s.r, 0, p.z};
        } else {
            const float sideDistance = s.r - radial;
            const float capDistance = hh - std::fabs(p.z);
            if (capDistance < sideDistance) {
                local = f3{p.x, p.y, p.z < 0.0f ? -hh : hh};
            } else {
                const float scale = s.r / radial;
                local = f3{p.x * scale, p.y * scale, p.z};
            }
        }
    } else {
        const f3 h = s.half;
        local = f3{sclamp(p.x, -h.x, h.x), sclamp(p.y, -h.y, h.y), sclamp(p.z, -h.z, h.z)};
    }
    return s.c + qrot(s.q, local);
}

void oracle_closest_point_shape(int type, const float pos[3], const float rot[4], const float size[3],
                                const float point[3], float out[3])
{
    const RShape s = makeRShape(type, f3{pos[0], pos[1], pos[2]}, quat{rot[0], rot[1], rot[2], rot[3]},
                                f3{size[0], size[1], size[2]});
    const f3 c = closestPoint(s, f3{point[0], point[1], point[2]});
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

// The server's get_closest_point_to_object_volume (avbd_direct_space_state3d.cpp:515-545) for
// bodies: the closest point of body[r]'s shape at its current pose to points[r].
int oracle_closest_point(oracle_world *o, int n, const int32_t *body, const float *points, float *out)
{
    const World &w = o->w;
    for (int r = 0; r < n; r++) {
        if (body[r] < 0 || body[r] >= static_cast<int>(w.bodies.size())) return AVBD_ERR_INVALID;
        const Body &b = w.bodies[body[r]];
        const f3 c = closestPoint(makeRShape(b.shape, b.pos, b.q, b.size), load3(points, r));
        store3(out, r, c);
    }
    return AVBD_OK;
}

// The bodies part of the server's intersect_point (avbd_direct_space_state3d.cpp:225-259) for n
// points: candidates from collectAABB over the point's box padded by 1e-4, in list order; the
// bodies whose shape contains the point, up to maxResults per point.
int oracle_intersect_point(oracle_world *o, int n, const float *points, const uint32_t *masks, int maxResults,
                           int32_t *count, int32_t *out)
{
    World &w = o->w;
    std::vector<int32_t> cand(w.bodies.size() + 1);
    for (int r = 0; r < n; r++) {
        count[r] = 0;
        const f3 p = load3(points, r);
        const float pad = 1.0e-4f;
        const float mn[3] = {p.x - pad, p.y - pad, p.z - pad}, mx[3] = {p.x + pad, p.y + pad, p.z + pad};
        const int nc = oracle_collect_aabb(o, mn, mx, masks ? masks[r] : 0xFFFFFFFFu, cand.data(),
                                           static_cast<int>(cand.size()));
        for (int k = 0; k < nc && count[r] < maxResults; k++) {
            const Body &b = w.bodies[cand[k]];
            if (!containsPoint(makeRShape(b.shape, b.pos, b.q, b.size), p, 0.0f)) continue;
            out[static_cast<size_t>(r) * maxResults + count[r]++] = cand[k];
        }
    }
    return AVBD_OK;
}

// The bodies part of the server's intersect_shape (avbd_direct_space_state3d.cpp:261-306) for
// n query shapes (type, centre, rotation, size as a body's, margin >= 0): candidates from
// collectAABB over the shape's bounds (query::shapeAabb, query_shape.cpp:371-394) grown by the
// margin, in list order; the bodies the narrow phase (collideShapes, query first) finds in
// contact within the margin, up to maxResults per query.
int oracle_intersect_shape(oracle_world *o, int n, const int32_t *type, const float *center, const float *rot,
                           const float *size, const float *margin, const uint32_t *masks, int maxResults,
                           int32_t *count, int32_t *out)
{
    World &w = o->w;
    std::vector<int32_t> cand(w.bodies.size() + 1);
    for (int r = 0; r < n; r++) {
        count[r] = 0;
        Body q{};
        q.shape = static_cast<uint8_t>(type[r]);
        q.pos = load3(center, r);
        q.q = quat{rot[4 * r], rot[4 * r + 1], rot[4 * r + 2], rot[4 * r + 3]};
        q.size = load3(size, r);
        const float m = smax(0.0f, margin[r]);
        const f3 wx = World::tightHalf(q);
        f3 lo = q.pos - wx, hi = q.pos + wx;
        if (m > 0.0f) { lo = lo - f3{m, m, m}; hi = hi + f3{m, m, m}; }
        const float mn[3] = {lo.x, lo.y, lo.z}, mx[3] = {hi.x, hi.y, hi.z};
        const int nc = oracle_collect_aabb(o, mn, mx, masks ? masks[r] : 0xFFFFFFFFu, cand.data(),
                                           static_cast<int>(cand.size()));
        const float qp[3] = {q.pos.x, q.pos.y, q.pos.z}, qq[4] = {q.q.x, q.q.y, q.q.z, q.q.w},
                    qs[3] = {q.size.x, q.size.y, q.size.z};
        for (int k = 0; k < nc && count[r] < maxResults; k++) {
            const Body &b = w.bodies[cand[k]];
            const float bp[3] = {b.pos.x, b.pos.y, b.pos.z}, bq[4] = {b.q.x, b.q.y, b.q.z, b.q.w},
                        bs[3] = {b.size.x, b.size.y, b.size.z};
            float bas[9]; int32_t keys[8]; float ra[24], rb[24];
            if (oracle_collide_shapes(q.shape, qp, qq, qs, b.shape, bp, bq, bs, m, bas, keys, ra, rb) <= 0) continue;
            out[static_cast<size_t>(r) * maxResults + count[r]++] = cand[k];
        }
    }
    return AVBD_OK;
}

namespace {
f3 shapeCentre(const float *center, int r) { return f3{center[3 * r], center[3 * r + 1], center[3 * r + 2]}; }
}  // namespace

// Candidates of query shape r exactly as intersect_shape takes them.
static int queryCandidates(oracle_world *o, int r, const int32_t *type, const float *center, const float *rot,
                           const float *size, float m, const uint32_t *masks, std::vector<int32_t> &cand, Body &q)
{
    q = Body{};
    q.shape = static_cast<uint8_t>(type[r]);
    q.pos = shapeCentre(center, r);
    q.q = quat{rot[4 * r], rot[4 * r + 1], rot[4 * r + 2], rot[4 * r + 3]};
    q.size = load3(size, r);
    const f3 wx = World::tightHalf(q);
    f3 lo = q.pos - wx, hi = q.pos + wx;
    if (m > 0.0f) { lo = lo - f3{m, m, m}; hi = hi + f3{m, m, m}; }
    const float mn[3] = {lo.x, lo.y, lo.z}, mx[3] = {hi.x, hi.y, hi.z};
    return oracle_collect_aabb(o, mn, mx, masks ? masks[r] : 0xFFFFFFFFu, cand.data(), static_cast<int>(cand.size()));
}

// The bodies part of the server's collide_shape (avbd_direct_space_state3d.cpp:386-447): for
// each query shape, its candidates in list order (as intersect_shape), the narrow phase query
// first within the margin, and every contact as a pair of world points (on the query, on the
// body) until maxPairs pairs are out.  any[r]: some candidate gave contacts.
int oracle_collide_shape(oracle_world *o, int n, const int32_t *type, const float *center, const float *rot,
                         const float *size, const float *margin, const uint32_t *masks, int maxPairs, int32_t *count,
                         uint8_t *any, float *points)
{
    World &w = o->w;
    std::vector<int32_t> cand(w.bodies.size() + 1);
    for (int r = 0; r < n; r++) {
        count[r] = 0;
        any[r] = 0;
        const float m = smax(0.0f, margin[r]);
        Body q;
        const int nc = queryCandidates(o, r, type, center, rot, size, m, masks, cand, q);
        const float qp[3] = {q.pos.x, q.pos.y, q.pos.z}, qq[4] = {q.q.x, q.q.y, q.q.z, q.q.w},
                    qs[3] = {q.size.x, q.size.y, q.size.z};
        int pairs = 0;
        for (int k = 0; k < nc; k++) {
            const Body &b = w.bodies[cand[k]];
            const float bp[3] = {b.pos.x, b.pos.y, b.pos.z}, bq[4] = {b.q.x, b.q.y, b.q.z, b.q.w},
                        bs[3] = {b.size.x, b.size.y, b.size.z};
            float bas[9]; int32_t keys[8]; float ra[24], rb[24];
            const int cnt = oracle_collide_shapes(q.shape, qp, qq, qs, b.shape, bp, bq, bs, m, bas, keys, ra, rb);
            if (cnt <= 0) continue;
            any[r] = 1;
            for (int i = 0; i < cnt && pairs < maxPairs; i++) {
                const f3 wa = q.pos + qrot(q.q, load3(ra, i));
                const f3 wb = b.pos + qrot(b.q, load3(rb, i));
                float *pt = points + (static_cast<size_t>(r) * maxPairs + pairs) * 6;
                pt[0] = wa.x; pt[1] = wa.y; pt[2] = wa.z; pt[3] = wb.x; pt[4] = wb.y; pt[5] = wb.z;
                pairs++;
            }
            if (pairs >= maxPairs) break;
        }
        count[r] = pairs;
    }
    return AVBD_OK;
}

static float queryPenetration(const Body &a, const Body &b, f3 normal);

// The bodies part of the server's rest_info (avbd_direct_space_state3d.cpp:449-513): over the
// candidates in list order and their contacts, the deepest by query::penetrationDepth along the
// contact normal (a later contact must be strictly deeper); body[r] = -1 when none.  point =
// the contact on the query shape in world space, normal = the basis' first row.
int oracle_rest_info(oracle_world *o, int n, const int32_t *type, const float *center, const float *rot,
                     const float *size, const float *margin, const uint32_t *masks, int32_t *body, float *point,
                     float *normal, float *depth)
{
    World &w = o->w;
    std::vector<int32_t> cand(w.bodies.size() + 1);
    for (int r = 0; r < n; r++) {
        const float m = smax(0.0f, margin[r]);
        Body q;
        const int nc = queryCandidates(o, r, type, center, rot, size, m, masks, cand, q);
        const float qp[3] = {q.pos.x, q.pos.y, q.pos.z}, qq[4] = {q.q.x, q.q.y, q.q.z, q.q.w},
                    qs[3] = {q.size.x, q.size.y, q.size.z};
        float best = -1.0f;
        f3 bestPoint{0, 0, 0}, bestNormal{0, 0, 0};
        int bestBody = -1;
        for (int k = 0; k < nc; k++) {
            const Body &b = w.bodies[cand[k]];
            const float bp[3] = {b.pos.x, b.pos.y, b.pos.z}, bq[4] = {b.q.x, b.q.y, b.q.z, b.q.w},
                        bs[3] = {b.size.x, b.size.y, b.size.z};
            float bas[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}; int32_t keys[8]; float ra[24], rb[24];
            const int cnt = oracle_collide_shapes(q.shape, qp, qq, qs, b.shape, bp, bq, bs, m, bas, keys, ra, rb);
            for (int i = 0; i < cnt; i++) {
                const f3 nrm{bas[0], bas[1], bas[2]};
                const float d = queryPenetration(q, b, nrm);
                if (d <= best) continue;
                best = d;
                bestNormal = nrm;
                bestPoint = q.pos + qrot(q.q, load3(ra, i));
                bestBody = cand[k];
            }
        }
        body[r] = bestBody;
        store3(point, r, bestPoint);
        store3(normal, r, bestNormal);
        depth[r] = best;
    }
    return AVBD_OK;
}

// ---- shape sweeps (query_shape.cpp:213-258, 340-369, 402-498; the server's cast_motion) ----
namespace {
constexpr float kQueryEps = 1.0e-6f;

struct QShape { int type; f3 c; quat q; f3 size; };

// sweptAabbInterval (query_shape.cpp:213-241): the t in [0, 1] where the moving box overlaps.
bool sweptInterval(f3 aMin, f3 aMax, f3 bMin, f3 bMax, f3 motion, float &rEnter, float &rExit)
{
    float enter = 0.0f, exit = 1.0f;
    const float amn[3] = {aMin.x, aMin.y, aMin.z}, amx[3] = {aMax.x, aMax.y, aMax.z};
    const float bmn[3] = {bMin.x, bMin.y, bMin.z}, bmx[3] = {bMax.x, bMax.y, bMax.z};
    const float mv[3] = {motion.x, motion.y, motion.z};
    for (int i = 0; i < 3; i++) {
        if (std::fabs(mv[i]) <= kQueryEps) {
            if (amx[i] < bmn[i] || amn[i] > bmx[i]) return false;
            continue;
        }
        float t0 = (bmn[i] - amx[i]) / mv[i], t1 = (bmx[i] - amn[i]) / mv[i];
        if (t0 > t1) std::swap(t0, t1);
        enter = smax(enter, t0);
        exit = smin(exit, t1);
        if (enter > exit) return false;
    }
    if (exit < 0.0f || enter > 1.0f) return false;
    rEnter = smax(0.0f, enter);
    rExit = smin(1.0f, exit);
    return rEnter <= rExit;
}

// contactAt (query_shape.cpp:243-258): the narrow phase at base + motion * t; the contact with
// the largest |rA - rB| (the first of equals) and the basis normal.
bool contactAt(const QShape &a, f3 motion, float t, const QShape &b, float margin, f3 &normal, f3 &rA)
{
    const f3 c = a.c + motion * t;
    const float pa[3] = {c.x, c.y, c.z}, qa[4] = {a.q.x, a.q.y, a.q.z, a.q.w}, sa[3] = {a.size.x, a.size.y, a.size.z};
    const float pb[3] = {b.c.x, b.c.y, b.c.z}, qb[4] = {b.q.x, b.q.y, b.q.z, b.q.w}, sb[3] = {b.size.x, b.size.y, b.size.z};
    float basis[9], ra[24], rb[24];
    int32_t keys[8];
    const int n = oracle_collide_shapes(a.type, pa, qa, sa, b.type, pb, qb, sb, margin, basis, keys, ra, rb);
    if (n <= 0) return false;
    int best = 0;
    float bestL = -1.0f;
    for (int k = 0; k < n; k++) {
        const float l = length2(f3{ra[3 * k], ra[3 * k + 1], ra[3 * k + 2]} - f3{rb[3 * k], rb[3 * k + 1], rb[3 * k + 2]});
        if (k == 0 || bestL < l) { best = k; bestL = l; }
    }
    normal = f3{basis[0], basis[1], basis[2]};
    rA = f3{ra[3 * best], ra[3 * best + 1], ra[3 * best + 2]};
    return true;
}

// supportRadius (query_shape.cpp:340-360)
float supportRadius(const QShape &s, f3 axis)
{
    const f3 n = normalize(axis);
    if (s.type == AVBD_SHAPE_SPHERE) return s.size.x;
    if (s.type == AVBD_SHAPE_CYLINDER) {
        const float axial = std::fabs(dot(qrot(s.q, f3{0, 0, 1}), n));
        const float radial = std::sqrt(smax(0.0f, 1.0f - axial * axial));
        return s.size.x * radial + (s.size.z * 0.5f) * axial;
    }
    const m3 R = qmat(s.q);
    const f3 h = s.size * 0.5f;
    return (h.x * std::fabs(dot(f3{R[0][0], R[1][0], R[2][0]}, n)) + h.y * std::fabs(dot(f3{R[0][1], R[1][1], R[2][1]}, n))) +
           h.z * std::fabs(dot(f3{R[0][2], R[1][2], R[2][2]}, n));
}

// penetrationDepth (query_shape.cpp:362-369)
float penetrationDepth(const QShape &a, const QShape &b, f3 normal)
{
    const float len = length(normal);
    if (len <= kQueryEps) return 0.0f;
    const f3 n = normal * (1.0f / len);
    return (supportRadius(a, n) + supportRadius(b, n)) - std::fabs(dot(b.c - a.c, n));
}

Body asBody(const QShape &s)
{
    Body b{};
    b.shape = static_cast<uint8_t>(s.type);
    b.pos = s.c;
    b.q = s.q;
    b.size = s.size;
    return b;
}

// sweepShape (query_shape.cpp:402-498): out = safe, unsafe, depth, point, normal.
bool sweepShape(const QShape &a, f3 motion, const QShape &b, float margin, float out[9])
{
    f3 normal{}, rA{};
    auto emit = [&](float safe, float unsafe, float t) {
        QShape posed = a;
        posed.c = a.c + motion * t;
        const f3 p = posed.c + qrot(posed.q, rA);
        const float v[9] = {safe, unsafe, penetrationDepth(posed, b, normal), p.x, p.y, p.z, normal.x, normal.y, normal.z};
        for (int k = 0; k < 9; k++) out[k] = v[k];
        return true;
    };
    if (length2(motion) <= kQueryEps * kQueryEps) {
        if (!contactAt(a, f3{0, 0, 0}, 0.0f, b, margin, normal, rA)) return false;
        QShape posed = a;
        const f3 p = posed.c + qrot(posed.q, rA);
        const float v[9] = {0.0f, 0.0f, penetrationDepth(a, b, normal), p.x, p.y, p.z, normal.x, normal.y, normal.z};
        for (int k = 0; k < 9; k++) out[k] = v[k];
        return true;
    }
    const f3 ha = World::tightHalf(asBody(a)), hb = World::tightHalf(asBody(b));
    f3 aMin = a.c - ha, aMax = a.c + ha;
    if (margin > 0.0f) { aMin = aMin - f3{margin, margin, margin}; aMax = aMax + f3{margin, margin, margin}; }
    float enter = 0.0f, exit = 1.0f;
    if (!sweptInterval(aMin, aMax, b.c - hb, b.c + hb, motion, enter, exit)) return false;
    if (contactAt(a, motion, enter, b, margin, normal, rA)) return emit(enter, enter, enter);
    constexpr int kSamples = 32;
    float low = enter, high = exit, previous = enter;
    bool found = false;
    for (int i = 1; i <= kSamples; i++) {
        const float t = enter + (exit - enter) * (static_cast<float>(i) / static_cast<float>(kSamples));
        if (contactAt(a, motion, t, b, margin, normal, rA)) { low = previous; high = t; found = true; break; }
        previous = t;
    }
    if (!found && contactAt(a, motion, exit, b, margin, normal, rA)) { low = previous; high = exit; found = true; }
    if (!found) return false;
    for (int i = 0; i < 12; i++) {
        const float mid = (low + high) * 0.5f;
        if (contactAt(a, motion, mid, b, margin, normal, rA)) high = mid;
        else low = mid;
    }
    if (!contactAt(a, motion, high, b, margin, normal, rA)) return false;
    return emit(low, high, high);
}
} // namespace

static float queryPenetration(const Body &a, const Body &b, f3 normal)
{
    return penetrationDepth(QShape{a.shape, a.pos, a.q, a.size}, QShape{b.shape, b.pos, b.q, b.size}, normal);
}

extern "C" int oracle_sweep_shape(int typeA, const float posA[3], const float rotA[4], const float sizeA[3],
                                  const float motion[3], int typeB, const float posB[3], const float rotB[4],
                                  const float sizeB[3], float margin, float out[9])
{
    const QShape a{typeA, load3(posA, 0), quat{rotA[0], rotA[1], rotA[2], rotA[3]}, load3(sizeA, 0)};
    const QShape b{typeB, load3(posB, 0), quat{rotB[0], rotB[1], rotB[2], rotB[3]}, load3(sizeB, 0)};
    return sweepShape(a, load3(motion, 0), b, margin, out) ? 1 : 0;
}

// The bodies part of the server's cast_motion (avbd_direct_space_state3d.cpp:310-385) for n
// query shapes: candidates from collectAABB over the shape's bounds swept by the motion and
// grown by the margin, in list order; the sweep with the smallest safe fraction (the first of
// equals) wins.  Per query: body[r] (-1: none), frac[2r..] = (safe, unsafe) (1, 1 when none),
// point[3r..], normal[3r..].
int oracle_cast_motion(oracle_world *o, int n, const int32_t *type, const float *center, const float *rot,
                       const float *size, const float *motion, const float *margin, const uint32_t *masks,
                       int32_t *body, float *frac, float *point, float *normal)
{
    World &w = o->w;
    std::vector<int32_t> cand(w.bodies.size() + 1);
    for (int r = 0; r < n; r++) {
        body[r] = -1;
        frac[2 * r] = frac[2 * r + 1] = 1.0f;
        const QShape q{type[r], load3(center, r), quat{rot[4 * r], rot[4 * r + 1], rot[4 * r + 2], rot[4 * r + 3]},
                       load3(size, r)};
        const f3 mv = load3(motion, r);
        const float m = smax(0.0f, margin[r]);
        const f3 wx = World::tightHalf(asBody(q));
        f3 lo = q.c - wx, hi = q.c + wx;
        lo = f3{smin(lo.x, lo.x + mv.x), smin(lo.y, lo.y + mv.y), smin(lo.z, lo.z + mv.z)};
        hi = f3{smax(hi.x, hi.x + mv.x), smax(hi.y, hi.y + mv.y), smax(hi.z, hi.z + mv.z)};
        if (m > 0.0f) { lo = lo - f3{m, m, m}; hi = hi + f3{m, m, m}; }
        const float mn[3] = {lo.x, lo.y, lo.z}, mx[3] = {hi.x, hi.y, hi.z};
        const int nc = oracle_collect_aabb(o, mn, mx, masks ? masks[r] : 0xFFFFFFFFu, cand.data(),
                                           static_cast<int>(cand.size()));
        float best[9];
        bool found = false;
        for (int k = 0; k < nc; k++) {
            const Body &b = w.bodies[cand[k]];
            const QShape t{b.shape, b.pos, b.q, b.size};
            float hit[9];
            if (!sweepShape(q, mv, t, m, hit)) continue;
            if (!found || hit[0] < best[0]) {
                found = true;
                body[r] = cand[k];
                for (int e = 0; e < 9; e++) best[e] = hit[e];
            }
        }
        if (!found) continue;
        frac[2 * r] = best[0];
        frac[2 * r + 1] = best[1];
        for (int e = 0; e < 3; e++) { point[3 * r + e] = best[3 + e]; normal[3 * r + e] = best[6 + e]; }
    }
    return AVBD_OK;
}

// Solver::pick (solver_query.cpp:29-155) for n rays: every body whose layer passes the mask, an
// exact sphere or the box bounding the shape (cylinder: its local box), nearest t in units of
// dir (the earlier body on a tie); body -1 when nothing is hit, else the hit point in the
// body's local frame.
int oracle_pick(oracle_world *o, int n, const float *origin, const float *dir, const uint32_t *masks, int32_t *body,
                float *local)
{
    const World &w = o->w;
    for (int r = 0; r < n; r++) {
        const f3 org = load3(origin, r), dr = load3(dir, r);
        const uint32_t mask = masks ? masks[r] : 0xFFFFFFFFu;
        float bestT = std::numeric_limits<float>::infinity();
        int best = -1;
        f3 bestLocal{0, 0, 0};
        for (size_t i = 0; i < w.bodies.size(); i++) {
            const Body &b = w.bodies[i];
            if ((b.layer & mask) == 0) continue;
            const quat inv = qconj(b.q);
            const f3 ol = qrot(inv, org - b.pos), d = qrot(inv, dr);
            if (b.shape == AVBD_SHAPE_SPHERE) {
                const float rr = b.size.x;
                const float bb = dot(ol, d);
                const float c = length2(ol) - rr * rr;
                const float disc = bb * bb - c;
                if (disc < 0.0f) continue;
                const float sq = std::sqrt(disc);
                float t = -bb - sq;
                if (t < 0.0f) t = -bb + sq;
                if (t < 0.0f || t >= bestT) continue;
                bestT = t; best = static_cast<int>(i); bestLocal = ol + d * t;
                continue;
            }
            const f3 half = b.shape == AVBD_SHAPE_CYLINDER ? f3{b.size.x, b.size.x, b.size.z * 0.5f} : b.size * 0.5f;
            float tEnter = 0.0f, tExit = std::numeric_limits<float>::infinity();
            bool hit = true;
            for (int k = 0; k < 3; k++) {
                if (std::fabs(d[k]) < 1.0e-6f) {
                    if (ol[k] < -half[k] || ol[k] > half[k]) { hit = false; break; }
                    continue;
                }
                const float invD = 1.0f / d[k];
                float t0 = (-half[k] - ol[k]) * invD, t1 = (half[k] - ol[k]) * invD;
                if (t0 > t1) std::swap(t0, t1);
                tEnter = smax(tEnter, t0);
                tExit = smin(tExit, t1);
                if (tEnter > tExit) { hit = false; break; }
            }
            if (!hit) continue;
            const float tHit = tEnter >= 0.0f ? tEnter : tExit;
            if (tHit < 0.0f) continue;
            if (tHit < bestT) { bestT = tHit; best = static_cast<int>(i); bestLocal = ol + d * tHit; }
        }
        body[r] = best;
        store3(local, r, best >= 0 ? bestLocal : f3{0, 0, 0});
    }
    return AVBD_OK;
}

float oracle_atan2(float y, float x) { return atan2Poly(y, x); }

// The server's per-body contact report (avbd_physics_server3d.cpp:1593-1652): live manifolds
// in order, each side whose body asks for reports (maxRep > 0) takes the manifold's contacts in
// order until it holds maxRep: the world point of its own arm at its current pose, the
// manifold normal (first basis row) signed towards the body (+ for A, - for B), the other body.
int oracle_contact_report(oracle_world *o, const int32_t *maxRep, int stride, int32_t *count, float *pos, float *normal,
                          int32_t *other)
{
    const World &w = o->w;
    const int n = static_cast<int>(w.bodies.size());
    for (int i = 0; i < n; i++) count[i] = 0;
    for (uint32_t s : w.live) {
        const int nc = w.numC[s];
        if (nc <= 0) continue;
        const f3 nrm{w.basis[s][0][0], w.basis[s][1][0], w.basis[s][2][0]};
        for (int side = 0; side < 2; side++) {
            const int b = side == 0 ? w.mA[s] : w.mB[s];
            const int lim = maxRep[b] < stride ? maxRep[b] : stride;
            if (lim <= 0) continue;
            const Body &me = w.bodies[b];
            for (int c = 0; c < nc && count[b] < lim; c++) {
                const Contact &ct = w.cont[static_cast<size_t>(s) * MAXC + c];
                const f3 world = me.pos + qrot(me.q, side == 0 ? ct.rA : ct.rB);
                const int k = b * stride + count[b];
                store3(pos, k, world);
                store3(normal, k, nrm * (side == 0 ? 1.0f : -1.0f));
                other[k] = side == 0 ? w.mB[s] : w.mA[s];
                count[b]++;
            }
        }
    }
    return AVBD_OK;
}

int oracle_get_manifolds(oracle_world *o, avbd_manifolds *m)
{
    const World &w = o->w;
    const int n = static_cast<int>(w.live.size());
    if (m->count < n) return AVBD_ERR_CAPACITY;
    m->count = n;
    for (int i = 0; i < n; i++) {
        const uint32_t s = w.live[i];
        if (m->body_a) m->body_a[i] = w.mA[s];
        if (m->body_b) m->body_b[i] = w.mB[s];
        if (m->basis)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) m->basis[9 * i + 3 * r + c] = w.basis[s][c][r];
        if (m->friction) m->friction[i] = w.mFric[s];
        if (m->num_contacts) m->num_contacts[i] = w.numC[s];
        for (int k = 0; k < MAXC; k++) {
            const int src = static_cast<int>(s) * MAXC + k, dst = i * MAXC + k;
            if (k >= w.numC[s]) {
                // slots beyond the count hold stale data the solver never reads
                const f3 z{0, 0, 0};
                if (m->feature) m->feature[dst] = 0;
                if (m->r_a) store3(m->r_a, dst, z);
                if (m->r_b) store3(m->r_b, dst, z);
                if (m->c0) store3(m->c0, dst, z);
                if (m->penalty) store3(m->penalty, dst, z);
                if (m->lambda) store3(m->lambda, dst, z);
                if (m->stick) m->stick[dst] = 0;
                if (m->beta_scale) m->beta_scale[dst] = 1.0f;
                continue;
            }
            if (m->feature) m->feature[dst] = w.cont[src].key;
            if (m->r_a) store3(m->r_a, dst, w.cont[src].rA);
            if (m->r_b) store3(m->r_b, dst, w.cont[src].rB);
            if (m->c0) store3(m->c0, dst, w.C0[src]);
            if (m->penalty) store3(m->penalty, dst, w.pen[src]);
            if (m->lambda) store3(m->lambda, dst, w.lam[src]);
            if (m->stick) m->stick[dst] = w.stick[src];
            if (m->beta_scale) m->beta_scale[dst] = w.betaScale[src];
        }
    }
    return AVBD_OK;
}

int oracle_set_manifolds(oracle_world *o, const avbd_manifolds *m)
{
    World &w = o->w;
    w.mA.clear(); w.mB.clear(); w.basis.clear(); w.numC.clear(); w.mFric.clear(); w.cont.clear();
    w.dormant.clear(); w.saved.clear();
    w.C0.clear(); w.pen.clear(); w.lam.clear(); w.stick.clear(); w.betaScale.clear();
    w.live.clear(); w.freeList.clear();
    for (int i = 0; i < m->count; i++) {
        const uint32_t s = w.acquire(m->body_a[i], m->body_b[i]);
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) w.basis[s][c][r] = m->basis[9 * i + 3 * r + c];
        w.mFric[s] = m->friction[i];
        w.numC[s] = m->num_contacts[i];
        for (int k = 0; k < MAXC; k++) {
            const int dst = static_cast<int>(s) * MAXC + k, src = i * MAXC + k;
            w.cont[dst] = Contact{m->feature[src], load3(m->r_a, src), load3(m->r_b, src)};
            w.C0[dst] = load3(m->c0, src);
            w.pen[dst] = load3(m->penalty, src);
            w.lam[dst] = load3(m->lambda, src);
            w.stick[dst] = m->stick[src];
            w.betaScale[dst] = m->beta_scale ? m->beta_scale[src] : 1.0f;
        }
    }
    return AVBD_OK;
}

void oracle_plant(int id) { g_plant = id; }

// The last dual's constraint values: per live manifold (live order) MAXC normal values C.x
// (entries past its contact count unused), per joint (force-list order) the linear C.
int oracle_get_residuals(oracle_world *o, float *contact_cn, float *joint_clin)
{
    const World &w = o->w;
    for (size_t i = 0; i < w.live.size(); i++)
        for (int k = 0; k < MAXC; k++)
            contact_cn[i * MAXC + k] = k < w.numC[w.live[i]] ? w.resC[static_cast<size_t>(w.live[i]) * MAXC + k] : 0.0f;
    for (size_t q = 0; q < w.joints.size(); q++) store3(joint_clin, static_cast<int>(q), w.joints[q].resLin);
    return AVBD_OK;
}

// Test hook: since the world was created, dual evaluations that reused cached arms and primal
// touches served by the force cache.
void oracle_cache_counters(oracle_world *o, long long *out)
{
    out[0] = o->w.armHits;
    out[1] = o->w.cachedTouches;
}

// avbd::Manifold::collide (collide.cpp:1100-1120) for bodies a (as A) and b at their current
// pose: the step's narrow phase with the speculative margin grown over hz.
int oracle_collide_pair(oracle_world *o, int a, int b, float hz, int32_t *keys, float *rA, float *rB, float *basis)
{
    World &w = o->w;
    const int n = static_cast<int>(w.bodies.size());
    if (a < 0 || b < 0 || a >= n || b >= n || a == b) return -1;
    const float keep = w.collisionDt;
    w.collisionDt = hz;
    const float margin = w.marginFor(w.bodies[a], w.bodies[b]);
    w.collisionDt = keep;
    const Body &A = w.bodies[a], &B = w.bodies[b];
    const float pa[3] = {A.pos.x, A.pos.y, A.pos.z}, qa[4] = {A.q.x, A.q.y, A.q.z, A.q.w},
                sa[3] = {A.size.x, A.size.y, A.size.z};
    const float pb[3] = {B.pos.x, B.pos.y, B.pos.z}, qb[4] = {B.q.x, B.q.y, B.q.z, B.q.w},
                sb[3] = {B.size.x, B.size.y, B.size.z};
    return oracle_collide_shapes(A.shape, pa, qa, sa, B.shape, pb, qb, sb, margin, basis, keys, rA, rB);
}

uint64_t oracle_digest(oracle_world *o)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ull; }
    };
    for (const Body &b : o->w.bodies) {
        mix(&b.pos, 12); mix(&b.q, 16); mix(&b.vel, 12); mix(&b.angVel, 12);
    }
    return h;
}

void oracle_stats(oracle_world *o, int32_t out[6])
{
    const World &w = o->w;
    int contacts = 0;
    for (uint32_t s : w.live) contacts += w.numC[s];
    out[0] = static_cast<int32_t>(w.live.size());
    out[1] = contacts;
    out[2] = w.lastPairs;
    out[3] = w.lastNew;
    out[4] = w.colourCount;
    out[5] = static_cast<int32_t>(w.overflowOrder.size());
}

void oracle_colours(oracle_world *o, uint8_t *out)
{
    const World &w = o->w;
    std::vector<uint8_t> jacobi(w.colour.size(), 0);
    for (int i : w.overflowOrder) jacobi[i] = 1;
    for (size_t b = 0; b < w.bodies.size(); b++) {
        const int u = w.bodies[b].updateIndex;
        if (u < 0) { out[b] = 255; continue; }
        const uint8_t c = w.colour[u];
        out[b] = c < kMaxColours ? c : (jacobi[u] ? 8 : 9);
    }
}

// Swept AABBs of the current state (solver_broadphase.cpp:101-122).
void oracle_swept_aabbs(oracle_world *o, float *centers, float *halves)
{
    std::vector<f3> c, h;
    o->w.collisionDt = o->w.P.dt;
    o->w.sweptAabbs(c, h);
    for (size_t i = 0; i < c.size(); i++) {
        store3(centers, static_cast<int>(i), c[i]);
        store3(halves, static_cast<int>(i), h[i]);
    }
}

// Candidate pairs (sorted unique) of the current state, before the filter.
int oracle_candidate_pairs(oracle_world *o, uint64_t *pairs, int capacity)
{
    std::vector<f3> c, h;
    std::vector<uint64_t> p;
    o->w.collisionDt = o->w.P.dt;
    o->w.sweptAabbs(c, h);
    o->w.candidatePairs(c, h, p);
    const int n = static_cast<int>(p.size());
    for (int i = 0; i < n && i < capacity; i++) pairs[i] = p[i];
    return n;
}

// bvh_broadphase.cpp:19-36, 69-96: Morton codes over the centres (every axis offset by
// lo.x, as the reference does) and the (code, index) order.
void oracle_morton(const float *centers, int n, uint64_t *codes, int32_t *order)
{
    auto expand = [](uint32_t v) {
        uint64_t x = v & 0x1fffffu;
        x = (x | (x << 32)) & 0x1f00000000ffffull;
        x = (x | (x << 16)) & 0x1f0000ff0000ffull;
        x = (x | (x << 8)) & 0x100f00f00f00f00full;
        x = (x | (x << 4)) & 0x10cfcfcfcfcfcfcfull;
        x = (x | (x << 2)) & 0x1249249249249249ull;
        return x;
    };
    f3 lo{FLT_MAX, FLT_MAX, FLT_MAX}, hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = 0; i < n; i++) {
        const f3 c = load3(centers, i);
        lo = f3{smin(lo.x, c.x), smin(lo.y, c.y), smin(lo.z, c.z)};
        hi = f3{smax(hi.x, c.x), smax(hi.y, c.y), smax(hi.z, c.z)};
    }
    const f3 ext{hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
    const float scale = 2097151.0f / std::max({ext.x, ext.y, ext.z, 1.0e-6f});
    auto q = [&](float v) { return static_cast<uint32_t>(sclamp((v - lo.x) * scale, 0.0f, 2097151.0f)); };
    for (int i = 0; i < n; i++) {
        const f3 c = load3(centers, i);
        codes[i] = expand(q(c.x)) | (expand(q(c.y)) << 1) | (expand(q(c.z)) << 2);
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; i++) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return codes[a] != codes[b] ? codes[a] < codes[b] : a < b; });
    for (int i = 0; i < n; i++) order[i] = ord[i];
}

// One force's contribution to one body's primal block at the current state (the per-force
// stamp of solver_solve.cpp's primal loop): lhsLin, lhsAng, lhsCross row-major, then rhsLin, rhsAng.
int oracle_force_stamp(oracle_world *o, int joint, int body, float *out33)
{
    if (!o || joint < 0 || joint >= static_cast<int>(o->w.joints.size())) return AVBD_ERR_INVALID;
    const Joint &j = o->w.joints[joint];
    if (body != j.a && body != j.b) return AVBD_ERR_INVALID;
    Block blk{};
    o->w.forcePrimal(j, body, blk);
    const m3 *ms[3] = {&blk.lhsLin, &blk.lhsAng, &blk.lhsCross};
    for (int m = 0; m < 3; m++)
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) out33[9 * m + 3 * r + c] = (*ms[m])[c][r];
    for (int r = 0; r < 3; r++) { out33[27 + r] = blk.rhsLin[r]; out33[30 + r] = blk.rhsAng[r]; }
    return AVBD_OK;
}

// One force's warm start (what = 0: Force::initialize, solver_solve.cpp:39-51) or dual
// update (what = 1: Force::updateDual, solver_solve.cpp's dual loop) applied alone at the
// current state.
int oracle_force_update(oracle_world *o, int joint, int what)
{
    if (!o || joint < 0 || joint >= static_cast<int>(o->w.joints.size()) || what < 0 || what > 1)
        return AVBD_ERR_INVALID;
    Joint &j = o->w.joints[joint];
    if (what == 0) o->w.jointInit(j);
    else o->w.jointDual(j);
    return AVBD_OK;
}

// ldlt6_solve / pivoted / solve chain on one row-major 6x6 system (maths.h:154-343).
int oracle_solve6(const float *a36, const float *b6, float *x6)
{
    float a[6][6];
    for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) a[r][c] = a36[6 * r + c];
    float x[6] = {};
    int path = 0;
    if (ldlt6(a, b6, x)) path = 1;
    else if (ldlt6Pivoted(a, b6, x)) path = 2;
    else for (int i = 0; i < 6; i++) x[i] = 0.0f;
    for (int i = 0; i < 6; i++) x6[i] = x[i];
    return path;
}

} // extern "C"

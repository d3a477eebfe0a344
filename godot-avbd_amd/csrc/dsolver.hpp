// Per-body, per-manifold and per-joint device routines of the AVBD step.
// Each routine cites the reference function whose results it reproduces; the evaluation
// order of every expression follows it (see dmath.hpp for the conventions).
#pragma once

#include "dmath.hpp"
#include "state.hpp"

namespace avbd_dev {

constexpr float PENALTY_MIN = 1.0f;
constexpr float COLLISION_MARGIN = 0.001f;
constexpr float STICK_THRESH = 0.00001f;
constexpr float CONTACT_RELEASE_GAP = 2.0e-3f;
constexpr float ADAPTIVE_BETA_MAX = 64.0f;
constexpr int MAXC = 8;
// Contact storage is sliced: the live manifolds of a step, in view order, form slices of
// CSLICE (one wavefront); contact i of the manifold in lane l of slice s is at
// sliceStart[s] + i * CSLICE + l, a slice being as wide as its largest contact count.  A
// wavefront's load of contact i then reads one contiguous run.
constexpr int CSLICE = 64;

// Penalty and lambda of dense contact c: {pen.xyz, lam.x} and {lam.y, lam.z} (24 bytes instead
// of two 16-byte rows); the stick flag is a byte of its own, so the dual writes no C0 row.
AD void loadPenLam(const State &S, uint32_t c, v3 &pen, v3 &lam)
{
    const float4 a = S.cPL[c];
    const float2 b = S.cLyz[c];
    pen = v3{a.x, a.y, a.z};
    lam = v3{a.w, b.x, b.y};
}
AD void storePenLam(const State &S, uint32_t c, v3 pen, v3 lam)
{
    S.cPL[c] = make_float4(pen.x, pen.y, pen.z, lam.x);
    S.cLyz[c] = make_float2(lam.y, lam.z);
}
constexpr float FLT_MAXV = 3.402823466e+38f;

struct Block {
    m3 lhsLin, lhsAng, lhsCross;
    v3 rhsLin, rhsAng;
};

// ------------------------------------------------------------- body loads ----
AD v3 bPos(const State &S, int b) { return ld3(S.pos[b]); }
AD qt bRot(const State &S, int b) { return ldq(S.rot[b]); }
AD float bMass(const State &S, int b) { return S.massI[b].x; }
AD v3 bMoment(const State &S, int b) { const float4 m = S.massI[b]; return {m.y, m.z, m.w}; }

// A body's packed pose record (see State::poseRec): the step-relative displacement dq of the
// primal/dual bracket, the rotation and the mass.
struct PoseRec { v3 dqL, dqA; qt q; float mass; };
AD PoseRec loadPose(const State &S, int b)
{
    const float4 r0 = S.poseRec[4 * b], r1 = S.poseRec[4 * b + 1], r2 = S.poseRec[4 * b + 2];
    return {ld3(r0), ld3(r2), ldq(r1), r0.w};
}

// types.hpp:157-172: M/dt^2 block with world inertia R diag(I) R^T / dt^2.
AD void makeBlock(const State &S, int b, Block &k)
{
    const float dtSq = S.P.dtSq;
    const float4 mi = S.massI[b];
    const v3 mo = v3{mi.x, mi.x, mi.x} / dtSq;
    k.lhsLin = diag(mo.x, mo.y, mo.z);
    k.lhsCross = zero3();
    k.rhsLin = cmul(mo, bPos(S, b) - ld3(S.inertLin[b]));
    const qt q = bRot(S, b);
    const m3 rot = qmat(q);
    const m3 iw = mul(rot, mul(diag(mi.y, mi.z, mi.w), transpose(rot)));
    k.lhsAng = mdiv(iw, dtSq);
    k.rhsAng = mul(k.lhsAng, qsub(q, ldq(S.inertAng[b])));
}

// maths.h:156-198: unpivoted LDL^T, accepted when every pivot > 1e-12 and max|x| < 10
// (the max drops NaN entries exactly as std::max does).  Fully unrolled: a, l, d and x
// live in VGPRs.
AD bool ldlt6(const float (&a)[6][6], const float (&b)[6], float (&x)[6])
{
    float l[6][6];
    float d[6];
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) l[i][j] = 0.0f;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        float dj = a[j][j];
#pragma unroll
        for (int s = 0; s < j; s++) dj -= l[j][s] * l[j][s] * d[s];
        ok = ok && (dj > 1.0e-12f);
        d[j] = dj;
        l[j][j] = 1.0f;
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            float lij = a[i][j];
#pragma unroll
            for (int s = 0; s < j; s++) lij -= l[i][s] * l[j][s] * d[s];
            l[i][j] = lij / d[j];
        }
    }
    if (!ok) return false;
    float z[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        float v = b[i];
#pragma unroll
        for (int s = 0; s < i; s++) v -= l[i][s] * z[s];
        z[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) x[i] = z[i] / d[i];
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        float v = x[i];
#pragma unroll
        for (int s = i + 1; s < 6; s++) v -= l[s][i] * x[s];
        x[i] = v;
    }
    float worst = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; i++) worst = smax(worst, fabsf(x[i]));
    return __builtin_isfinite(worst) && worst < 1.0e1f;
}

// maths.h:202-307: Eigen-order pivoted LDL^T, the fallback for indefinite blocks.  Fully
// unrolled so that every index is a constant: a row or column swap with the pivot row is a
// select over the candidate rows, and the permutation of x swaps through the same selects, so
// the matrix stays in registers (no scratch memory) and the arithmetic is the serial order.
AD bool ldlt6Pivoted(const float (&ain)[6][6], const float (&b)[6], float (&x)[6])
{
    float a[6][6];
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) a[i][j] = ain[i][j];
    float d[6] = {0, 0, 0, 0, 0, 0};
    int t[6] = {0, 1, 2, 3, 4, 5};
    bool zeroPivot = false;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        int j = k;
        float best = fabsf(a[k][k]);
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const float c = fabsf(a[i][i]);
            if (c > best) { best = c; j = i; }
        }
        t[k] = j;
#pragma unroll
        for (int r = k + 1; r < 6; r++) {
            if (j == r) {
#pragma unroll
                for (int c = 0; c < 6; c++) { const float tmp = a[k][c]; a[k][c] = a[r][c]; a[r][c] = tmp; }
#pragma unroll
                for (int q = 0; q < 6; q++) { const float tmp = a[q][k]; a[q][k] = a[q][r]; a[q][r] = tmp; }
            }
        }
        if (k > 0) {
            float temp[6], prod[6];
#pragma unroll
            for (int i = 0; i < k; i++) temp[i] = d[i] * a[k][i];
#pragma unroll
            for (int i = 0; i < k; i++) prod[i] = a[k][i] * temp[i];
            float dt;
            if (k == 1) dt = prod[0];
            else if (k == 2) dt = prod[0] + prod[1];
            else if (k == 3) dt = (prod[0] + prod[1]) + prod[2];
            else if (k == 4) dt = (prod[0] + prod[2]) + (prod[1] + prod[3]);
            else dt = ((prod[0] + prod[2]) + (prod[1] + prod[3])) + prod[4];
            a[k][k] -= dt;
#pragma unroll
            for (int i = 0; i < k; i++) {
                const float ti = temp[i];
#pragma unroll
                for (int r = k + 1; r < 6; r++) a[r][k] -= a[r][i] * ti;
            }
        }
        const float pivot = a[k][k];
        const bool valid = fabsf(pivot) > 0.0f;
        if (!valid) {
#pragma unroll
            for (int r = k + 1; r < 6; r++)
                if (a[r][k] != 0.0f) return false;
            zeroPivot = true;
        } else if (zeroPivot) {
            return false;
        }
        if (valid)
#pragma unroll
            for (int r = k + 1; r < 6; r++) a[r][k] /= pivot;
        d[k] = pivot;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) x[i] = b[i];
#pragma unroll
    for (int k = 0; k < 6; k++)
#pragma unroll
        for (int r = k + 1; r < 6; r++)
            if (t[k] == r) { const float tmp = x[k]; x[k] = x[r]; x[r] = tmp; }
#pragma unroll
    for (int jj = 0; jj < 5; jj++) {
        const float v = x[jj];
        if (v != 0.0f)
#pragma unroll
            for (int i = jj + 1; i < 6; i++) x[i] -= a[i][jj] * v;
    }
    const float tol = 1.17549435e-38f;
#pragma unroll
    for (int i = 0; i < 6; i++) x[i] = fabsf(d[i]) > tol ? x[i] / d[i] : 0.0f;
#pragma unroll
    for (int jj = 5; jj >= 1; jj--) {
        const float v = x[jj];
        if (v != 0.0f)
#pragma unroll
            for (int i = jj - 1; i >= 0; i--) x[i] -= a[jj][i] * v;
    }
#pragma unroll
    for (int k = 5; k >= 0; k--)
#pragma unroll
        for (int r = k + 1; r < 6; r++)
            if (t[k] == r) { const float tmp = x[k]; x[k] = x[r]; x[r] = tmp; }
    float worst = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; i++) worst = smax(worst, fabsf(x[i]));
    return __builtin_isfinite(worst) && worst < 1.0e1f;
}

AD void packA(const Block &k, float (&a)[6][6], float (&b)[6])
{
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            a[r][c] = k.lhsLin.c[c][r];
            a[r + 3][c + 3] = k.lhsAng.c[c][r];
            a[r + 3][c] = k.lhsCross.c[c][r];
            a[r][c + 3] = k.lhsCross.c[r][c];
        }
    b[0] = -k.rhsLin.x; b[1] = -k.rhsLin.y; b[2] = -k.rhsLin.z;
    b[3] = -k.rhsAng.x; b[4] = -k.rhsAng.y; b[5] = -k.rhsAng.z;
}

// First half of types.hpp:176-197: the unpivoted LDL^T only.  Returns false when the
// reference would fall through to the pivoted solve (queued by the caller).
AD bool solveDeltaFast(const Block &k, v3 &dLin, v3 &dAng)
{
    float a[6][6], b[6], x[6];
    packA(k, a, b);
    if (!ldlt6(a, b, x)) return false;
    dLin = v3{x[0], x[1], x[2]};
    dAng = v3{x[3], x[4], x[5]};
    return true;
}

// The rest of the chain for a queued block: pivoted LDL^T, else a zero update.
AD void solveDeltaSlow(const Block &k, v3 &dLin, v3 &dAng)
{
    float a[6][6], b[6], x[6];
    packA(k, a, b);
    if (ldlt6Pivoted(a, b, x)) {
        dLin = v3{x[0], x[1], x[2]};
        dAng = v3{x[3], x[4], x[5]};
        return;
    }
    dLin = v3{0, 0, 0};
    dAng = v3{0, 0, 0};
}

AD void storeBlock(float *p, const Block &k)
{
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            p[3 * c + r] = k.lhsLin.c[c][r];
            p[9 + 3 * c + r] = k.lhsAng.c[c][r];
            p[18 + 3 * c + r] = k.lhsCross.c[c][r];
        }
    p[27] = k.rhsLin.x; p[28] = k.rhsLin.y; p[29] = k.rhsLin.z;
    p[30] = k.rhsAng.x; p[31] = k.rhsAng.y; p[32] = k.rhsAng.z;
}
AD Block loadBlock(const float *p)
{
    Block k;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) {
            k.lhsLin.c[c][r] = p[3 * c + r];
            k.lhsAng.c[c][r] = p[9 + 3 * c + r];
            k.lhsCross.c[c][r] = p[18 + 3 * c + r];
        }
    k.rhsLin = v3{p[27], p[28], p[29]};
    k.rhsAng = v3{p[30], p[31], p[32]};
    return k;
}

// types.hpp:176-197 with maths.h:315-343: LDL^T, else pivoted LDL^T, else a zero update.
AD void solveDelta(const Block &k, v3 &dLin, v3 &dAng)
{
    float a[6][6];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            a[r][c] = k.lhsLin.c[c][r];
            a[r + 3][c + 3] = k.lhsAng.c[c][r];
            a[r + 3][c] = k.lhsCross.c[c][r];
            a[r][c + 3] = k.lhsCross.c[r][c];
        }
    const float b[6] = {-k.rhsLin.x, -k.rhsLin.y, -k.rhsLin.z, -k.rhsAng.x, -k.rhsAng.y, -k.rhsAng.z};
    float x[6];
    if (ldlt6(a, b, x) || ldlt6Pivoted(a, b, x)) {
        dLin = v3{x[0], x[1], x[2]};
        dAng = v3{x[3], x[4], x[5]};
        return;
    }
    dLin = v3{0, 0, 0};
    dAng = v3{0, 0, 0};
}

// ----------------------------------------------------------- manifolds ----
AD m3 loadBasisK(const State &S, int k)
{
    const v3 r0 = ld3(S.mBasisK[3 * k]), r1 = ld3(S.mBasisK[3 * k + 1]), r2 = ld3(S.mBasisK[3 * k + 2]);
    return transpose(m3{{r0, r1, r2}});
}
AD m3 loadBasis(const State &S, int slot)
{
    const float4 *r = S.mRec + 4 * static_cast<size_t>(slot);
    const v3 r0 = ld3(r[0]), r1 = ld3(r[1]), r2 = ld3(r[2]);
    return transpose(m3{{r0, r1, r2}});
}
// The slot's whole record (see State::mRec), one aligned line.
AD void storeRec(const State &S, int slot, const m3 &B, int a, int b, int item, float mu, int num)
{
    float4 *r = S.mRec + 4 * static_cast<size_t>(slot);
    r[0] = st3(row(B, 0), __int_as_float(a));
    r[1] = st3(row(B, 1), __int_as_float(b));
    r[2] = st3(row(B, 2), __int_as_float(item));
    r[3] = make_float4(mu, __int_as_float(num), 0.0f, 0.0f);
}
AD float recFric(const State &S, int slot) { return S.mRec[4 * static_cast<size_t>(slot) + 3].x; }
AD int recItem(const State &S, int slot) { return __float_as_int(S.mRec[4 * static_cast<size_t>(slot) + 2].w); }

// manifold.cpp:383-388
AD v3 constraintError(const State &S, v3 c0, v3 bracket, const m3 &B)
{
    const bool spec = S.P.specMargin > 0.0f;
    const bool separated = spec && c0.x > COLLISION_MARGIN + CONTACT_RELEASE_GAP;
    return c0 * (separated ? 1.0f : S.P.oneMinusAlpha) + mul(B, bracket);
}

// manifold.cpp:407-419: push-only normal, tangential clamped to the friction cone.
// Returns the pre-clamp tangential magnitude; `bounds` receives mu*|Fn|.
AD float clampForce(v3 &F, float mu, float &bounds)
{
    F.x = smin(F.x, 0.0f);
    bounds = fabsf(F.x) * mu;
    const float fs = length2d(F.y, F.z);
    if (fs > bounds && fs > 0) {
        F.y *= bounds / fs;
        F.z *= bounds / fs;
    }
    return fs;
}

struct Hess { float s00, s01, s02, s11, s12, s22; };
// manifold.cpp:362-379: S = B^T K B.
AD Hess basisHessian(const m3 &B, v3 k)
{
    const v3 b0 = row(B, 0), b1 = row(B, 1), b2 = row(B, 2);
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

// Angular and cross terms of one contact's stamp on the body with arm r (manifold.cpp:688-755).
AD void stampSide(const Hess &h, const m3 &Sm, v3 r, v3 fw, m3 &accAng, m3 &accCross, v3 &accRA)
{
    const m3 RS{{cross(r, Sm.c[0]), cross(r, Sm.c[1]), cross(r, Sm.c[2])}};
    const float rx2 = r.x * r.x, ry2 = r.y * r.y, rz2 = r.z * r.z;
    const float rxy = r.x * r.y, rxz = r.x * r.z, ryz = r.y * r.z;
    const float h00 = h.s11 * rz2 - 2.0f * h.s12 * ryz + h.s22 * ry2;
    const float h11 = h.s22 * rx2 - 2.0f * h.s02 * rxz + h.s00 * rz2;
    const float h22 = h.s00 * ry2 - 2.0f * h.s01 * rxy + h.s11 * rx2;
    const float h01 = h.s12 * rxz + h.s02 * ryz - h.s22 * rxy - h.s01 * rz2;
    const float h02 = h.s12 * rxy + h.s01 * ryz - h.s11 * rxz - h.s02 * ry2;
    const float h12 = h.s02 * rxy + h.s01 * rxz - h.s00 * ryz - h.s12 * rx2;
    accAng = madd(accAng, m3{{v3{h00, h01, h02}, v3{h01, h11, h12}, v3{h02, h12, h22}}});
    accCross = madd(accCross, RS);
    accRA += cross(r, fw);
}

// What one adjacency entry adds to its body's block, kept apart per add: the serial assembly
// adds a joint's terms to lhsAng up to three times and to rhsAng twice, and float sums depend
// on that order (adding a zero is not a no-op on a zero's sign), so a term is applied exactly
// when the serial code adds it.  flags: 1 = lin, ang, cross, rl, ra; 2 = the diagonal
// geometric-stiffness add to lhsAng; 4 = ang2, ra2 (a joint's angular rows); 8 = a force the
// block's owner stamps itself (spring, generic joint).
struct EntryTerms {
    m3 lin, ang, cross, ang2;
    v3 rl, ra, angDiag, ra2;
    int flags;
};
AD void applyTerms(Block &blk, const EntryTerms &t)
{
    if (t.flags & 1) {
        blk.lhsLin = madd(blk.lhsLin, t.lin);
        blk.lhsAng = madd(blk.lhsAng, t.ang);
        blk.lhsCross = madd(blk.lhsCross, t.cross);
        blk.rhsLin += t.rl;
        blk.rhsAng += t.ra;
    }
    if (t.flags & 2) blk.lhsAng = madd(blk.lhsAng, diag(t.angDiag.x, t.angDiag.y, t.angDiag.z));
    if (t.flags & 4) {
        blk.lhsAng = madd(blk.lhsAng, t.ang2);
        blk.rhsAng += t.ra2;
    }
}

// One manifold's contribution to one body's block (manifold.cpp:474-497 force + :503-568
// stamp, fused).  Forces are recomputed at the current poses; the reference's force cache
// holds the same floats because nothing it reads moves between fill and use.  ARMS (arm reuse
// on): a round's first touch after the first round (`cachedRound`, the other body not solved
// earlier in this round) takes the world arms the last dual left, as the reference's cached
// stamp does (manifold.cpp:570-581); every other touch computes them and stores them.
template <bool ARMS>
AD void manifoldTerms(const State &S, int k, int body, EntryTerms &t, bool cachedRound = false, int colour = 0)
{
    const int4 hd = S.mHdr[k];
    const int a = hd.x, b = hd.y, n = hd.w;
    const uint32_t cb = static_cast<uint32_t>(hd.z);
    const m3 B = loadBasisK(S, k);
    const float mu = S.mFricK[k];
    const qt qa = bRot(S, a), qb = bRot(S, b);
    const v3 dqAL = bPos(S, a) - ld3(S.initLin[a]), dqAA = qsub(qa, ldq(S.initAng[a]));
    const v3 dqBL = bPos(S, b) - ld3(S.initLin[b]), dqBA = qsub(qb, ldq(S.initAng[b]));
    const m3 rotA = qmat(qa), rotB = qmat(qb);
    const bool fromA = body == a;
    const float side = fromA ? 1.0f : -1.0f;
    const v3 b0 = row(B, 0), b1 = row(B, 1), b2 = row(B, 2);
    m3 accLin = zero3(), accAng = zero3(), accCross = zero3();
    v3 accRL{0, 0, 0}, accRA{0, 0, 0};
    bool cached = false;
    size_t arm0 = 0;
    if (ARMS) {
        const int o = fromA ? b : a;
        cached = cachedRound && !(S.updateIndex[o] >= 0 && S.colour[o] < colour);
        arm0 = static_cast<size_t>(S.kSlot[k]) * MAXC;
    }
    // contact i + 1 is loaded while contact i is computed
    float4 nPL{}, nRA{}, nRB{}, nC0{};
    float2 nLyz{};
    if (n > 0) { nPL = S.cPL[cb]; nRA = S.cRA[cb]; nRB = S.cRB[cb]; nC0 = S.cC0[cb]; nLyz = S.cLyz[cb]; }
    for (int i = 0; i < n; i++) {
        const uint32_t c = cb + i * CSLICE;
        const float4 pl4 = nPL, ra4 = nRA, rb4 = nRB, c04 = nC0;
        const float2 lyz = nLyz;
        if (i + 1 < n) { const uint32_t d = c + CSLICE; nPL = S.cPL[d]; nRA = S.cRA[d]; nRB = S.cRB[d]; nC0 = S.cC0[d]; nLyz = S.cLyz[d]; }
        v3 rA, rB;
        if (ARMS && cached) {
            rA = ld3(S.armA[arm0 + i]);
            rB = ld3(S.armB[arm0 + i]);
        } else {
            rA = mul(rotA, ld3(ra4));
            rB = mul(rotB, ld3(rb4));
            if (ARMS) { S.armA[arm0 + i] = st3(rA); S.armB[arm0 + i] = st3(rB); }
        }
        const v3 br = (dqAL - dqBL) + cross(dqAA, rA) - cross(dqBA, rB);
        const v3 C = constraintError(S, ld3(c04), br, B);
        const v3 pen = ld3(pl4), lam{pl4.w, lyz.x, lyz.y};
        v3 f = cmul(pen, C) + lam;
        float bounds;
        clampForce(f, mu, bounds);
        const Hess h = basisHessian(B, pen);
        const m3 Sm{{v3{h.s00, h.s01, h.s02}, v3{h.s01, h.s11, h.s12}, v3{h.s02, h.s12, h.s22}}};
        const v3 fw = (b0 * f.x + b1 * f.y) + b2 * f.z;
        stampSide(h, Sm, fromA ? rA : rB, fw, accAng, accCross, accRA);
        accLin = madd(accLin, Sm);
        accRL += fw;
    }
    t.lin = accLin;
    t.ang = accAng;
    t.cross = accCross;
    t.rl = accRL * side;
    t.ra = accRA * side;
    t.flags = 1;
}
template <bool ARMS>
AD void stampManifold(const State &S, int k, int body, Block &blk, bool cachedRound = false, int colour = 0)
{
    EntryTerms t;
    manifoldTerms<ARMS>(S, k, body, t, cachedRound, colour);
    applyTerms(blk, t);
}

// ---- per-manifold stamps, gathered by the Jacobi solve ----
// A manifold's summed stamp is the same for every body that gathers it except for the arm
// terms, so it is built once per iteration for both bodies (7 float4 per side in a record of
// STAMP_ROWS) and each body's solve adds its side in adjacency order: the floats are those
// stampManifold produces.

struct StampAcc {
    m3 lin, angA, crossA, angB, crossB;
    v3 rl, raA, raB;
    bool movA, movB;   // a side whose body is static is never gathered: its stamp is not stored
};
AD void stampInit(StampAcc &s, bool movA, bool movB)
{
    s.lin = s.angA = s.crossA = s.angB = s.crossB = zero3();
    s.rl = s.raA = s.raB = v3{0, 0, 0};
    s.movA = movA;
    s.movB = movB;
}
AD void stampContact(StampAcc &s, const m3 &B, v3 pen, v3 f, v3 rA, v3 rB)
{
    const Hess h = basisHessian(B, pen);
    const m3 Sm{{v3{h.s00, h.s01, h.s02}, v3{h.s01, h.s11, h.s12}, v3{h.s02, h.s12, h.s22}}};
    const v3 fw = (row(B, 0) * f.x + row(B, 1) * f.y) + row(B, 2) * f.z;
    stampSide(h, Sm, rA, fw, s.angA, s.crossA, s.raA);
    stampSide(h, Sm, rB, fw, s.angB, s.crossB, s.raB);
    s.lin = madd(s.lin, Sm);
    s.rl += fw;
}
AD void stampRows(float4 (&p)[7], const m3 &L, v3 rl, const m3 &A, const m3 &X, v3 ra)
{
    p[0] = make_float4(L.c[0][0], L.c[0][1], L.c[0][2], L.c[1][1]);
    p[1] = make_float4(L.c[1][2], L.c[2][2], rl.x, rl.y);
    p[2] = make_float4(rl.z, A.c[0][0], A.c[0][1], A.c[0][2]);
    p[3] = make_float4(A.c[1][1], A.c[1][2], A.c[2][2], X.c[0][0]);
    p[4] = make_float4(X.c[0][1], X.c[0][2], X.c[1][0], X.c[1][1]);
    p[5] = make_float4(X.c[1][2], X.c[2][0], X.c[2][1], X.c[2][2]);
    p[6] = make_float4(ra.x, ra.y, ra.z, 0.0f);
}

// A side's stamp record is STAMP_ROWS float4 (the 7 of stampRows and an empty eighth), so each
// record is one aligned 128-byte line and the primal's gather of it reads exactly two 64-byte
// sectors (seven rows at 112-byte strides straddle three sectors in half the cases).
constexpr int STAMP_ROWS = 8;
constexpr int STAMP_LDS = STAMP_ROWS * 65;   // float4 of LDS per wavefront (row stride 65)

// Wavefront-wide store of the active lanes' records (lane l's at dst + STAMP_ROWS l): the rows
// go through LDS (row q of lane l at q * 65 + l, conflict free both ways) so that every store
// instruction writes one contiguous run instead of 64 scattered 16-byte pieces.  The active
// lanes must be a prefix of the wavefront (the live-list walks guarantee it); `lds` holds
// STAMP_LDS float4 for this wavefront.  Lanes with `keep` false leave their record unwritten.
AD void waveStoreRows(float4 *dst, const float4 (&rows)[7], float4 *lds, int lane, bool keep)
{
    const int r = __popcll(__ballot(1));
    const unsigned long long kept = __ballot(keep);
#pragma unroll
    for (int q = 0; q < 7; q++) lds[q * 65 + lane] = rows[q];
    lds[7 * 65 + lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int q = 0; q < STAMP_ROWS; q++) {
        const int idx = q * r + lane, owner = idx / STAMP_ROWS;
        if ((kept >> owner) & 1ull) dst[idx] = lds[(idx % STAMP_ROWS) * 65 + owner];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Both sides' stamps of manifold k, lane = k's lane in a wavefront whose active lanes hold
// consecutive k.  Side B's force terms carry the reference's sign flip (side = -1).
AD void storeStampsWave(const State &S, int k, const StampAcc &s, float4 *lds, int lane)
{
    const size_t k0 = static_cast<size_t>(k - lane);
    float4 rows[7];
    stampRows(rows, s.lin, s.rl, s.angA, s.crossA, s.raA);
    waveStoreRows(S.stampA + STAMP_ROWS * k0, rows, lds, lane, s.movA);
    stampRows(rows, s.lin, s.rl * -1.0f, s.angB, s.crossB, s.raB * -1.0f);
    waveStoreRows(S.stampB + STAMP_ROWS * k0, rows, lds, lane, s.movB);
}
// Symmetric 3x3 from its upper triangle (the sums are symmetric bit for bit).
AD m3 sym3(float a00, float a01, float a02, float a11, float a12, float a22)
{
    return m3{{v3{a00, a01, a02}, v3{a01, a11, a12}, v3{a02, a12, a22}}};
}
AD void addStampSide(Block &blk, const float4 *p)
{
    const float4 p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3], p4 = p[4], p5 = p[5], p6 = p[6];
    blk.lhsLin = madd(blk.lhsLin, sym3(p0.x, p0.y, p0.z, p0.w, p1.x, p1.y));
    blk.lhsAng = madd(blk.lhsAng, sym3(p2.y, p2.z, p2.w, p3.x, p3.y, p3.z));
    blk.lhsCross = madd(blk.lhsCross, m3{{v3{p3.w, p4.x, p4.y}, v3{p4.z, p4.w, p5.x}, v3{p5.y, p5.z, p5.w}}});
    blk.rhsLin += v3{p1.z, p1.w, p2.x};
    blk.rhsAng += v3{p6.x, p6.y, p6.z};
}

// A manifold the reference's overflow phase A precomputes (solver_step.cpp:235-246): one of
// its bodies is in the update set and in the overflow bucket.
AD bool phaseA(const State &S, int a, int b)
{
    return (S.updateIndex[a] >= 0 && S.colour[a] == 8) || (S.updateIndex[b] >= 0 && S.colour[b] == 8);
}

// Both sides' stamps of one manifold at the current poses (zero for a frozen one, which no
// body gathers).  ARMS: the arms are stored for the manifolds the reference precomputes.
template <bool ARMS>
AD void manifoldStamps(const State &S, int k, StampAcc &acc)
{
    const int4 hd = S.mHdr[k];
    const int a = hd.x, b = hd.y, n = hd.w;
    const uint32_t cb = static_cast<uint32_t>(hd.z);
    const PoseRec A = loadPose(S, a), Bd = loadPose(S, b);
    const m3 B = loadBasisK(S, k);
    const float mu = S.mFricK[k];
    stampInit(acc, A.mass > 0.0f, Bd.mass > 0.0f);
    if (!(A.mass > 0.0f) && !(Bd.mass > 0.0f)) return;
    const v3 dqAL = A.dqL, dqAA = A.dqA, dqBL = Bd.dqL, dqBA = Bd.dqA;
    const m3 rotA = qmat(A.q), rotB = qmat(Bd.q);
    const bool store = ARMS && phaseA(S, a, b);
    const size_t arm0 = ARMS ? static_cast<size_t>(S.kSlot[k]) * MAXC : 0;
    for (int i = 0; i < n; i++) {
        const uint32_t c = cb + i * CSLICE;
        const v3 rA = mul(rotA, ld3(S.cRA[c]));
        const v3 rB = mul(rotB, ld3(S.cRB[c]));
        if (store) { S.armA[arm0 + i] = st3(rA); S.armB[arm0 + i] = st3(rB); }
        const v3 br = (dqAL - dqBL) + cross(dqAA, rA) - cross(dqBA, rB);
        const v3 C = constraintError(S, ld3(S.cC0[c]), br, B);
        v3 pen, lam;
        loadPenLam(S, c, pen, lam);
        v3 f = cmul(pen, C) + lam;
        float bounds;
        clampForce(f, mu, bounds);
        stampContact(acc, B, pen, f, rA, rB);
    }
}

// manifold.cpp:583-671: lambda <- clamped force, penalty ramps.  With STAMP, the stamps of
// the next primal iteration are built from the updated lambda and penalty: the poses do not
// move between this dual and that primal, so its constraint error is the one computed here.
// ARMS (arm reuse on): while neither body's first rotation column has moved farther than the
// rebuild distance since the arms were captured, the world arms stored last are used (the
// stamps still take arms at the current pose, as the reference's phase A recomputes them);
// otherwise the arms are computed, stored and captured.
template <bool STAMP, bool ARMS, bool RESID = false>
AD void manifoldDual(const State &S, int k, float decay, StampAcc &acc)
{
    // every load that does not depend on another is issued before the first wait; the
    // frozen test (no movable body) comes after them
    const int4 hd = S.mHdr[k];
    const int a = hd.x, b = hd.y, n = hd.w;
    const uint32_t cb = static_cast<uint32_t>(hd.z);
    const PoseRec A = loadPose(S, a), Bd = loadPose(S, b);
    const m3 B = loadBasisK(S, k);
    const float mu = S.mFricK[k];
    // contact i + 1 is read before contact i is written (distinct rows)
    float4 nPL{}, nRA{}, nRB{}, nC0{};
    float2 nLyz{};
    if (n > 0) { nPL = S.cPL[cb]; nRA = S.cRA[cb]; nRB = S.cRB[cb]; nC0 = S.cC0[cb]; nLyz = S.cLyz[cb]; }
    if (STAMP) stampInit(acc, A.mass > 0.0f, Bd.mass > 0.0f);
    if (!(A.mass > 0.0f) && !(Bd.mass > 0.0f)) {
        if (RESID)   // a frozen manifold's dual does not run: no constraint value
            for (int i = 0; i < n; i++) S.resC[static_cast<size_t>(S.kSlot[k]) * MAXC + i] = 0.0f;
        return;
    }
    const v3 dqAL = A.dqL, dqAA = A.dqA, dqBL = Bd.dqL, dqBA = Bd.dqA;
    const m3 rotA = qmat(A.q), rotB = qmat(Bd.q);
    const float pmax = S.P.penaltyMax;
    bool usable = false, store = false;
    size_t arm0 = 0;
    if (ARMS) {
        const int slot = S.kSlot[k];
        arm0 = static_cast<size_t>(slot) * MAXC;
        const float4 capA = S.armCap[2 * static_cast<size_t>(slot)], capB = S.armCap[2 * static_cast<size_t>(slot) + 1];
        const v3 dA = rotA.c[0] - ld3(capA), dB = rotB.c[0] - ld3(capB);
        usable = capA.w != 0.0f && length2(dA) <= S.P.armRebuild2 && length2(dB) <= S.P.armRebuild2;
        store = STAMP && phaseA(S, a, b);
        if (!usable) {
            S.armCap[2 * static_cast<size_t>(slot)] = st3(rotA.c[0], 1.0f);
            S.armCap[2 * static_cast<size_t>(slot) + 1] = st3(rotB.c[0], 0.0f);
        }
    }
    for (int i = 0; i < n; i++) {
        const uint32_t c = cb + i * CSLICE;
        const float4 pl4 = nPL, ra4 = nRA, rb4 = nRB, c0in = nC0;
        const float2 lyz = nLyz;
        if (i + 1 < n) { const uint32_t d = c + CSLICE; nPL = S.cPL[d]; nRA = S.cRA[d]; nRB = S.cRB[d]; nC0 = S.cC0[d]; nLyz = S.cLyz[d]; }
        v3 pen = ld3(pl4);
        if (decay < 1.0f) pen = lclamp3(pen * decay, PENALTY_MIN, pmax);
        v3 rA, rB;
        if (ARMS && usable) {
            rA = ld3(S.armA[arm0 + i]);
            rB = ld3(S.armB[arm0 + i]);
        } else {
            rA = mul(rotA, ld3(ra4));
            rB = mul(rotB, ld3(rb4));
            if (ARMS) { S.armA[arm0 + i] = st3(rA); S.armB[arm0 + i] = st3(rB); }
        }
        const v3 br = (dqAL - dqBL) + cross(dqAA, rA) - cross(dqBA, rB);
        const v3 C = constraintError(S, ld3(c0in), br, B);
        if (RESID) S.resC[static_cast<size_t>(S.kSlot[k]) * MAXC + i] = C.x;
        v3 F = cmul(pen, C) + v3{pl4.w, lyz.x, lyz.y};
        float bounds;
        const float scale = clampForce(F, mu, bounds);
        const float beta = S.P.betaLin * ra4.w;
        if (F.x < 0) pen.x = smin(pen.x + beta * fabsf(C.x), pmax);
        if (scale <= bounds) {
            pen.y = smin(pen.y + beta * fabsf(C.y), pmax);
            pen.z = smin(pen.z + beta * fabsf(C.z), pmax);
            S.cStick[c] = length2d(C.y, C.z) < STICK_THRESH ? 1 : 0;
        }
        storePenLam(S, c, pen, F);
        if (STAMP) {
            if (ARMS && usable) {
                // the next round's phase A at the current pose
                const v3 fA = mul(rotA, ld3(ra4)), fB = mul(rotB, ld3(rb4));
                const v3 fbr = (dqAL - dqBL) + cross(dqAA, fA) - cross(dqBA, fB);
                v3 f = cmul(pen, constraintError(S, ld3(c0in), fbr, B)) + F;
                float bd;
                clampForce(f, mu, bd);
                stampContact(acc, B, pen, f, fA, fB);
                if (store) { S.armA[arm0 + i] = st3(fA); S.armB[arm0 + i] = st3(fB); }
            } else {
                v3 f = cmul(pen, C) + F;
                float bd;
                clampForce(f, mu, bd);
                stampContact(acc, B, pen, f, rA, rB);
            }
        }
    }
}

// manifold.cpp:633-646
AD void manifoldRebase(const State &S, int k)
{
    const int4 hd = S.mHdr[k];
    const int a = hd.x, b = hd.y, n = hd.w;
    const uint32_t cb = static_cast<uint32_t>(hd.z);
    const v3 pa = bPos(S, a), pb = bPos(S, b);
    const qt qa = bRot(S, a), qb = bRot(S, b);
    const m3 B = loadBasisK(S, k);
    float4 nRA{}, nRB{}, nC0{};
    if (n > 0) { nRA = S.cRA[cb]; nRB = S.cRB[cb]; nC0 = S.cC0[cb]; }
    for (int i = 0; i < n; i++) {
        const uint32_t c = cb + i * CSLICE;
        const float4 ra4 = nRA, rb4 = nRB, old = nC0;
        if (i + 1 < n) { const uint32_t d = c + CSLICE; nRA = S.cRA[d]; nRB = S.cRB[d]; nC0 = S.cC0[d]; }
        const v3 xA = pa + qrot(qa, ld3(ra4));
        const v3 xB = pb + qrot(qb, ld3(rb4));
        v3 c0 = mul(B, xA - xB) + v3{COLLISION_MARGIN, 0, 0};
        if (c0.x < -S.P.maxPen) c0.x = -S.P.maxPen;
        S.cC0[c] = st3(c0, old.w);
    }
}

// ------------------------------------------------------------- joints ----
AD v3 jointLin(const State &S, int j)
{
    const int a = S.jA[j], b = S.jB[j];
    const v3 rA = ld3(S.jRA[j]);
    const v3 pa = a >= 0 ? bPos(S, a) + qrot(bRot(S, a), rA) : rA;
    return pa - (bPos(S, b) + qrot(bRot(S, b), ld3(S.jRB[j])));
}
AD v3 jointAng(const State &S, int j)
{
    const int a = S.jA[j];
    const qt qa = a >= 0 ? bRot(S, a) : qt{0, 0, 0, 1};
    return qsub(qa, bRot(S, S.jB[j])) * S.jArm[j];
}
AD m3 gsBall(int k, v3 v)
{
    m3 m = diag(-v[k], -v[k], -v[k]);
    m.c[k][0] += v[0];
    m.c[k][1] += v[1];
    m.c[k][2] += v[2];
    return m;
}
// joint.cpp:80-160
// spring.cpp:31-68: a spring has no constraint stabilisation and no dual state.
AD void springPrimal(const State &S, int j, int body, Block &blk)
{
    const int a = S.jA[j], b = S.jB[j];
    const v3 rA = ld3(S.jRA[j]), rB = ld3(S.jRB[j]);
    const qt qa = bRot(S, a), qb = bRot(S, b);
    const v3 pA = bPos(S, a) + qrot(qa, rA);
    const v3 pB = bPos(S, b) + qrot(qb, rB);
    const v3 d = pA - pB;
    const float dLen = length(d);
    if (dLen <= 1.0e-6f) return;
    const v3 n = d / dLen;
    const float k = S.jStiffL[j];
    const float C = dLen - S.jRest[j];
    const float f = k * C;
    v3 jLin, jAng;
    if (body == a) {
        jLin = n;
        jAng = cross(qrot(qa, rA), n);
    } else {
        jLin = -n;
        jAng = -cross(qrot(qb, rB), n);
    }
    const v3 F = jLin * f, Tau = jAng * f;
    // outer products a b^T (column j = a * b[j]) scaled by the stiffness
    const m3 Kll = mscale(m3{{jLin * jLin.x, jLin * jLin.y, jLin * jLin.z}}, k);
    const m3 Kla = mscale(m3{{jAng * jLin.x, jAng * jLin.y, jAng * jLin.z}}, k);
    const m3 Kaa = mscale(m3{{jAng * jAng.x, jAng * jAng.y, jAng * jAng.z}}, k);
    blk.lhsLin = madd(blk.lhsLin, Kll);
    blk.lhsAng = madd(blk.lhsAng, Kaa);
    blk.lhsCross = madd(blk.lhsCross, Kla);
    blk.rhsLin += F;
    blk.rhsAng += Tau;
}

// ------------------------------------------------------ generic 6-DOF joint ----
// generic_joint.cpp.  Axis a of joint j (0-2 linear x, y, z; 3-5 angular) is at 6 j + a:
// mode (0 free, 1 locked, 2 limited), bounds, spring, and the dual state (lambda, penalty,
// value at the step's start).
constexpr uint8_t AXIS_FREE = 0, AXIS_LOCKED = 1, AXIS_LIMITED = 2;

AD v3 frameAxis(int i) { return v3{i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f}; }
AD qt genericFrame(const State &S, int j) { return S.jA[j] >= 0 ? bRot(S, S.jA[j]) : qt{0.0f, 0.0f, 0.0f, 1.0f}; }

// generic_joint.cpp:51-67: take the twist about frame axis i out of dev; r_angle its angle.
AD void removeTwist(qt &dev, int i, float &angle)
{
    const v3 n = frameAxis(i);
    const float proj = dot(v3{dev.x, dev.y, dev.z}, n);
    const qt twist{dev.x * proj, dev.y * proj, dev.z * proj, dev.w};
    if (sqrtf(qlength2(twist)) <= 1.0e-12f) { angle = 0.0f; return; }
    angle = 2.0f * atan2Poly(proj, dev.w);
    const float l2 = qlength2(twist);
    const qt inv{-twist.x / l2, -twist.y / l2, -twist.z / l2, twist.w / l2};
    dev = qmul(qnormalize(dev), inv);
}

// generic_joint.cpp:69-90: the three angular measurements (twist of the non-locked axes,
// then swing components of the locked ones).
AD void genericAngular(const State &S, int j, float (&out)[3])
{
    const qt qA = genericFrame(S, j);
    const qt relNow = qmul(qconj(qA), qnormalize(bRot(S, S.jB[j])));
    qt dev = qmul(qnormalize(ldq(S.jGRest[j])), qconj(relNow));
    uint8_t mode[3];
#pragma unroll
    for (int i = 0; i < 3; i++) mode[i] = S.jGMode[6 * j + 3 + i];
    // unrolled so that dev and out stay in registers
    if (mode[0] != AXIS_LOCKED) removeTwist(dev, 0, out[0]);
    if (mode[1] != AXIS_LOCKED) removeTwist(dev, 1, out[1]);
    if (mode[2] != AXIS_LOCKED) removeTwist(dev, 2, out[2]);
    if (mode[0] == AXIS_LOCKED) out[0] = dev.x * 2.0f;
    if (mode[1] == AXIS_LOCKED) out[1] = dev.y * 2.0f;
    if (mode[2] == AXIS_LOCKED) out[2] = dev.z * 2.0f;
}

// generic_joint.cpp:92-96: anchor A minus anchor B in the joint frame.
AD v3 genericOffset(const State &S, int j, qt qA)
{
    const int a = S.jA[j], b = S.jB[j];
    const v3 rA = ld3(S.jRA[j]);
    const v3 anchorA = a >= 0 ? bPos(S, a) + qrot(qA, rA) : rA;
    const v3 anchorB = bPos(S, b) + qrot(bRot(S, b), ld3(S.jRB[j]));
    return qrot(qconj(qA), anchorA - anchorB);
}

// generic_joint.cpp:104-155: constraint value and force of axis x, false when it acts not.
AD bool axisForce(const State &S, int x, float value, float alpha, float &C, float &F)
{
    const uint8_t mode = S.jGMode[x];
    if (mode == AXIS_LOCKED) {
        C = value;
    } else if (mode == AXIS_LIMITED) {
        if (value > S.jGHi[x]) C = value - S.jGHi[x];
        else if (value < S.jGLo[x]) C = value - S.jGLo[x];
        else return false;
    } else {
        return false;
    }
    if (mode == AXIS_LOCKED) C -= S.jGInit[x] * alpha;
    float f = S.jGPen[x] * C + S.jGLam[x];
    if (mode == AXIS_LIMITED) f = C > 0.0f ? smax(f, 0.0f) : smin(f, 0.0f);
    F = f;
    return true;
}

AD m3 outerScaled(v3 a, v3 b, float s) { return mscale(m3{{a * b.x, a * b.y, a * b.z}}, s); }

// generic_joint.cpp:157-187
AD void stampAxis(const State &S, int x, Block &blk, float value, float alpha, v3 jLin, v3 jAng, float rate)
{
    float C, F;
    if (axisForce(S, x, value, alpha, C, F)) {
        const float k = S.jGPen[x];
        blk.lhsLin = madd(blk.lhsLin, outerScaled(jLin, jLin, k));
        blk.lhsAng = madd(blk.lhsAng, outerScaled(jAng, jAng, k));
        blk.lhsCross = madd(blk.lhsCross, outerScaled(jAng, jLin, k));
        blk.rhsLin += jLin * F;
        blk.rhsAng += jAng * F;
    }
    const float ks = S.jGK[x];
    if (ks > 0.0f) {
        const float Cs = value - S.jGEq[x];
        const float Fs = ks * Cs - S.jGDamp[x] * rate;
        blk.lhsLin = madd(blk.lhsLin, outerScaled(jLin, jLin, ks));
        blk.lhsAng = madd(blk.lhsAng, outerScaled(jAng, jAng, ks));
        blk.lhsCross = madd(blk.lhsCross, outerScaled(jAng, jLin, ks));
        blk.rhsLin += jLin * Fs;
        blk.rhsAng += jAng * Fs;
    }
}

// generic_joint.cpp:234-277
AD void genericPrimal(const State &S, int j, int body, Block &blk)
{
    const int a = S.jA[j], b = S.jB[j];
    const qt qA = genericFrame(S, j);
    const v3 rAWorld = qrot(qA, ld3(S.jRA[j]));
    const v3 rBWorld = qrot(bRot(S, b), ld3(S.jRB[j]));
    const v3 offset = genericOffset(S, j, qA);
    float ang[3];
    genericAngular(S, j, ang);
    const bool isA = body == a;
    const float sign = isA ? 1.0f : -1.0f;
    const v3 arm = isA ? rAWorld : rBWorld;
    const v3 vA = a >= 0 ? ld3(S.vel[a]) + cross(ld3(S.angVel[a]), rAWorld) : v3{0, 0, 0};
    const v3 vB = ld3(S.vel[b]) + cross(ld3(S.angVel[b]), rBWorld);
    const v3 relVel = vA - vB;
    const v3 wA = a >= 0 ? ld3(S.angVel[a]) : v3{0, 0, 0};
    const v3 wB = ld3(S.angVel[b]);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const v3 axis = qrot(qA, frameAxis(i));
        const v3 jLin = axis * sign;
        const v3 jAng = cross(arm, jLin);
        stampAxis(S, 6 * j + i, blk, offset[i], S.P.alpha, jLin, jAng, dot(relVel, axis));
        stampAxis(S, 6 * j + 3 + i, blk, ang[i], S.P.alpha, v3{0, 0, 0}, axis * sign, dot(wA - wB, axis));
    }
}

// generic_joint.cpp:191-210: the largest penalty this body resolves along the axis.
AD float penaltyLimit(const State &S, int body, v3 axis, bool angular)
{
    float effective;
    if (angular) {
        const v3 local = qrot(qconj(bRot(S, body)), axis);
        const v3 m = bMoment(S, body);
        effective = (local.x * local.x * m.x + local.y * local.y * m.y) + local.z * local.z * m.z;
    } else {
        effective = bMass(S, body);
    }
    const float pmax = S.P.penaltyMax;
    if (effective <= 0.0f) return pmax;
    const float limit = effective * 0.25f * 0.25f / (S.P.dt * S.P.dt);
    return smin(limit, pmax);
}

// generic_joint.cpp:279-319
AD void genericDual(const State &S, int j)
{
    const qt qA = genericFrame(S, j);
    float ang[3];
    genericAngular(S, j, ang);
    const v3 offset = genericOffset(S, j, qA);
    const int a = S.jA[j], b = S.jB[j];
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int x = 6 * j + 3 * k + i;
            const float value = k == 0 ? offset[i] : ang[i];
            const float beta = k == 0 ? S.P.betaLin : S.P.betaAng;
            float C, F;
            if (!axisForce(S, x, value, S.P.alpha, C, F)) {
                if (S.jGMode[x] == AXIS_LIMITED) S.jGLam[x] = 0.0f;
                continue;
            }
            S.jGLam[x] = F;
            const v3 dir = qrot(qA, frameAxis(i));
            float cap = penaltyLimit(S, b, dir, k == 1);
            if (a >= 0) cap = smin(cap, penaltyLimit(S, a, dir, k == 1));
            S.jGPen[x] = sclamp(S.jGPen[x] + fabsf(C) * beta, 0.0f, cap);
        }
    }
}

// generic_joint.cpp:212-232: values at the step's start, warm-started dual state.
AD void genericInit(const State &S, int j)
{
    float ang[3];
    genericAngular(S, j, ang);
    const v3 offset = genericOffset(S, j, genericFrame(S, j));
    const float ag = S.P.alpha * S.P.gamma, pmax = S.P.penaltyMax;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        S.jGInit[6 * j + i] = offset[i];
        S.jGInit[6 * j + 3 + i] = ang[i];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int x = 6 * j + 3 * k + i;
            S.jGLam[x] = S.jGLam[x] * ag;
            S.jGPen[x] = sclamp(S.jGPen[x] * S.P.gamma, PENALTY_MIN, pmax);
        }
    }
}

// A force of the body's list: joint, spring, generic joint, or a collision filter (which
// stamps nothing).
AD void jointPrimal(const State &S, int j, int body, Block &blk);
AD void forcePrimal(const State &S, int j, int body, Block &blk)
{
    const uint8_t kind = S.jKind[j];
    if (kind == FORCE_JOINT) jointPrimal(S, j, body, blk);
    else if (kind == FORCE_SPRING) springPrimal(S, j, body, blk);
    else if (kind == FORCE_GENERIC) genericPrimal(S, j, body, blk);
}

AD void jointTerms(const State &S, int j, int body, EntryTerms &t);
AD void jointPrimal(const State &S, int j, int body, Block &blk)
{
    EntryTerms t;
    jointTerms(S, j, body, t);
    applyTerms(blk, t);
}
// joint.cpp:80-160: a ball / finite joint's linear rows (with the geometric stiffness of the
// rotating arm) and angular rows, as terms of the block.
AD void jointTerms(const State &S, int j, int body, EntryTerms &t)
{
    t.flags = 0;
    // Every load is issued before the first branch (one round of waits for the joint's fields,
    // one for the two poses); the arithmetic is jointLin / jointAng's, in their order.
    const int ja = S.jA[j], jb = S.jB[j];
    const v3 penL = ld3(S.jPenL[j]), penA = ld3(S.jPenA[j]);
    const v3 lamL = ld3(S.jLamL[j]), lamA = ld3(S.jLamA[j]);
    const v3 c0L = ld3(S.jC0L[j]), c0A = ld3(S.jC0A[j]);
    const float stL = S.jStiffL[j], stA = S.jStiffA[j], armA = S.jArm[j];
    const v3 rA = ld3(S.jRA[j]), rB = ld3(S.jRB[j]);
    const int aa = ja >= 0 ? ja : jb;   // a world anchor (ja < 0) reads no body
    const v3 posA = bPos(S, aa), posB = bPos(S, jb);
    const qt rotA = bRot(S, aa), rotB = bRot(S, jb);
    const bool isA = body == ja;
    if (length2(penL) > 0) {
        const m3 K = diag(penL.x, penL.y, penL.z);
        const v3 pa = ja >= 0 ? posA + qrot(rotA, rA) : rA;
        v3 C = pa - (posB + qrot(rotB, rB));
        if (__builtin_isinf(stL)) C -= c0L * S.P.alpha;
        const v3 F = mul(K, C) + lamL;
        const m3 jLin = isA ? ident() : mneg(ident());
        m3 jAng;
        if (isA) {
            const v3 r = -qrot(rotA, rA);
            jAng = m3{{v3{0, r.z, -r.y}, v3{-r.z, 0, r.x}, v3{r.y, -r.x, 0}}};
        } else {
            const v3 r = qrot(rotB, rB);
            jAng = m3{{v3{0, r.z, -r.y}, v3{-r.z, 0, r.x}, v3{r.y, -r.x, 0}}};
        }
        const m3 jLinT = transpose(jLin), jAngT = transpose(jAng);
        const m3 jAngTk = mul(jAngT, K);
        t.lin = mul(mul(jLinT, K), jLin);
        t.ang = mul(jAngTk, jAng);
        t.cross = mul(jAngTk, jLin);
        const v3 r = isA ? qrot(rotA, rA) : -qrot(rotB, rB);
        const m3 H = madd(madd(mscale(gsBall(0, r), F.x), mscale(gsBall(1, r), F.y)), mscale(gsBall(2, r), F.z));
        t.angDiag = v3{H.c[0][0], H.c[1][1], H.c[2][2]};
        t.rl = mul(jLinT, F);
        t.ra = mul(jAngT, F);
        t.flags = 1 | 2;
    }
    if (length2(penA) > 0) {
        const m3 K = diag(penA.x, penA.y, penA.z);
        const qt qa = ja >= 0 ? rotA : qt{0, 0, 0, 1};
        v3 C = qsub(qa, rotB) * armA;
        if (__builtin_isinf(stA)) C -= c0A * S.P.alpha;
        const v3 F = mul(K, C) + lamA;
        const float arm = armA;
        const m3 jA = mscale(isA ? ident() : m3{{v3{0, 0, 0}, v3{0, -1, 0}, v3{0, 0, -1}}}, arm);
        t.ang2 = mul(mul(transpose(jA), K), jA);
        t.ra2 = mul(transpose(jA), F);
        t.flags |= 4;
    }
}

// joint.cpp:171-230
// `rec`: where to keep the linear rows' constraint value C (joint.cpp:178; 0 when the joint
// has no linear rows or is not a ball / finite joint), or null.
AD void jointDual(const State &S, int j, float4 *rec = nullptr)
{
    if (rec) *rec = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (S.jKind[j] == FORCE_GENERIC) { genericDual(S, j); return; }
    if (S.jKind[j] != FORCE_JOINT) return;   // springs and collision filters keep no dual state
    const float pmax = S.P.penaltyMax;
    v3 penL = ld3(S.jPenL[j]), penA = ld3(S.jPenA[j]);
    v3 lamL = ld3(S.jLamL[j]), lamA = ld3(S.jLamA[j]);
    if (length2(penL) > 0) {
        const m3 K = diag(penL.x, penL.y, penL.z);
        v3 C = jointLin(S, j);
        const float st = S.jStiffL[j];
        if (__builtin_isinf(st)) {
            C -= ld3(S.jC0L[j]) * S.P.alpha;
            lamL = mul(K, C) + lamL;
        }
        if (rec) *rec = st3(C);
        penL = lmin3(penL + fabs3(C) * S.P.betaLin, lmin3(v3{st, st, st}, v3{pmax, pmax, pmax}));
    }
    if (length2(penA) > 0) {
        const m3 K = diag(penA.x, penA.y, penA.z);
        v3 C = jointAng(S, j);
        const float st = S.jStiffA[j];
        if (__builtin_isinf(st)) {
            C -= ld3(S.jC0A[j]) * S.P.alpha;
            lamA = mul(K, C) + lamA;
        }
        penA = lmin3(penA + fabs3(C) * S.P.betaAng, lmin3(v3{st, st, st}, v3{pmax, pmax, pmax}));
    }
    const float fr = S.jFrac[j];
    if (length2(lamA) > fr * fr) {
        penL = penA = lamL = lamA = v3{0, 0, 0};
        S.jBroken[j] = 1;
    }
    S.jPenL[j] = st3(penL);
    S.jPenA[j] = st3(penA);
    S.jLamL[j] = st3(lamL);
    S.jLamA[j] = st3(lamA);
}

// Adjacency entries: e < 0 is joint -(e + 1); otherwise manifold view index k, with
// ADJ_SIDE_B set in the entry of the manifold's body b.
constexpr int ADJ_SIDE_B = 1 << 30;
AD int adjManifold(int e) { return e & (ADJ_SIDE_B - 1); }
AD bool adjSideB(int e) { return (e & ADJ_SIDE_B) != 0; }

// Block of one body: inertia plus every incident constraint, in adjacency order (joints in
// list order, then manifolds in live order; manifold.cpp:123-166, solver_step.cpp:215-226).
template <bool ARMS>
AD void assemble(const State &S, int body, Block &blk, bool cachedRound = false, int colour = 0)
{
    makeBlock(S, body, blk);
    const int e0 = S.adjStart[body], e1 = S.adjStart[body + 1];
    for (int e = e0; e < e1; e++) {
        const int ref = S.adj[e];
        if (ref >= 0) stampManifold<ARMS>(S, adjManifold(ref), body, blk, cachedRound, colour);
        else forcePrimal(S, -ref - 1, body, blk);
    }
}

// The same block from the per-manifold stamps (joints are stamped here, as in assemble).
AD void assembleStamped(const State &S, int body, Block &blk)
{
    makeBlock(S, body, blk);
    const int e0 = S.adjStart[body], e1 = S.adjStart[body + 1];
    for (int e = e0; e < e1; e++) {
        const int ref = S.adj[e];
        if (ref >= 0)
            addStampSide(blk, (adjSideB(ref) ? S.stampB : S.stampA) + STAMP_ROWS * static_cast<size_t>(adjManifold(ref)));
        else forcePrimal(S, -ref - 1, body, blk);
    }
}

// ------------------------------------------------------- narrow phase ----
// collide.cpp:71-490, 1034-1097: box-box SAT with face/edge selection, reference-face
// clipping and feature keys.
struct OBB { v3 c; qt q; v3 h; v3 ax[3]; };

AD OBB makeOBB(v3 c, qt q, v3 h)
{
    OBB o;
    o.c = c; o.q = q; o.h = h;
    o.ax[0] = qrot(q, v3{1, 0, 0});
    o.ax[1] = qrot(q, v3{0, 1, 0});
    o.ax[2] = qrot(q, v3{0, 0, 1});
    return o;
}
// Selections are made per float: a conditional between two aggregates selects their
// addresses, which keeps both boxes in scratch memory instead of registers.
AD v3 sel3(bool f, v3 a, v3 b)
{
    const float ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return v3{f ? ax : bx, f ? ay : by, f ? az : bz};
}
AD OBB pickOBB(bool first, const OBB &x, const OBB &y)
{
    OBB o;
    o.c = sel3(first, x.c, y.c);
    const qt xq = x.q, yq = y.q;
    o.q = qt{first ? xq.x : yq.x, first ? xq.y : yq.y, first ? xq.z : yq.z, first ? xq.w : yq.w};
    o.h = sel3(first, x.h, y.h);
    o.ax[0] = sel3(first, x.ax[0], y.ax[0]);
    o.ax[1] = sel3(first, x.ax[1], y.ax[1]);
    o.ax[2] = sel3(first, x.ax[2], y.ax[2]);
    return o;
}
// Axis i of a box (a run-time index into b.ax would put the box in scratch).
AD v3 axisOf(const OBB &b, int i) { return sel3(i == 0, b.ax[0], sel3(i == 1, b.ax[1], b.ax[2])); }
AD v3 supportPoint(const OBB &b, v3 dir)
{
    const float sx = dot(dir, b.ax[0]) >= 0.0f ? 1.0f : -1.0f;
    const float sy = dot(dir, b.ax[1]) >= 0.0f ? 1.0f : -1.0f;
    const float sz = dot(dir, b.ax[2]) >= 0.0f ? 1.0f : -1.0f;
    return ((b.c + b.ax[0] * (b.h.x * sx)) + b.ax[1] * (b.h.y * sy)) + b.ax[2] * (b.h.z * sz);
}
// The two in-face axes of face i (i = 0: y, z; 1: x, z; 2: x, y) with their half extents.
AD void faceAxes(const OBB &b, int i, v3 &u, v3 &v, float &eu, float &ev)
{
    const float hx = b.h.x, hy = b.h.y, hz = b.h.z;
    u = sel3(i == 0, b.ax[1], b.ax[0]);
    v = sel3(i == 2, b.ax[1], b.ax[2]);
    eu = i == 0 ? hy : hx;
    ev = i == 2 ? hy : hz;
}
constexpr int MAX_POLY = 16;
// A clip polygon of capacity MAX_POLY.  Clipping a quad by the four side planes adds at most
// one vertex per plane (eight in all, more only through rounding).  The first POLY_LDS sit in
// LDS, one lane's column (vertex v at x[v * NARROW_TPB], the caller's pointers offset by the
// lane, so a wavefront's accesses are consecutive words); the rest (the seventh and eighth of
// an octagon, and any past them) go to `spill` (3 floats each), global memory the caller owns.
// Six LDS vertices (9 KiB per 64 lanes) let a CU hold as many wavefronts as the kernel's
// registers allow; eight held it to 13 (measured 1.48 -> 1.38 ms on the settled C4 batch).
constexpr int NARROW_TPB = 64;
constexpr int POLY_LDS = 6;
struct Poly {
    float *x, *y, *z;
    float4 *spill;   // float e of the overflow is component e % 4 of row spill[(e / 4) * CSLICE]
    AD float &spilled(int e) const { return reinterpret_cast<float *>(spill + (e >> 2) * CSLICE)[e & 3]; }
    AD v3 get(int v) const
    {
        if (v < POLY_LDS) { const int i = v * NARROW_TPB; return {x[i], y[i], z[i]}; }
        const int e = 3 * (v - POLY_LDS);
        return {spilled(e), spilled(e + 1), spilled(e + 2)};
    }
    AD void put(int v, v3 a) const
    {
        if (v < POLY_LDS) { const int i = v * NARROW_TPB; x[i] = a.x; y[i] = a.y; z[i] = a.z; return; }
        const int e = 3 * (v - POLY_LDS);
        spilled(e) = a.x;
        spilled(e + 1) = a.y;
        spilled(e + 2) = a.z;
    }
};
static_assert(3 * (MAX_POLY - POLY_LDS) <= 4 * MAXC, "polygon overflow must fit one manifold's staging rows");
// collide.cpp:170-203 (Sutherland-Hodgman against one plane, capacity MAX_POLY)
AD int clipPoly(const Poly &in, int n, v3 pn, float off, const Poly &out)
{
    if (n <= 0) return 0;
    int m = 0;
    v3 a = in.get(n - 1);
    float da = dot(pn, a) - off;
    for (int i = 0; i < n; i++) {
        const v3 b = in.get(i);
        const float db = dot(pn, b) - off;
        const bool ai = da <= 1.0e-5f, bi = db <= 1.0e-5f;
        if (ai != bi) {
            float t = 0.0f;
            const float den = da - db;
            if (fabsf(den) > 1.0e-6f) t = sclamp(da / den, 0.0f, 1.0f);
            if (m < MAX_POLY) out.put(m++, a + (b - a) * t);
        }
        if (bi && m < MAX_POLY) out.put(m++, b);
        a = b;
        da = db;
    }
    return m;
}

// The contact list under construction lives in the manifold's staging rows (row i at
// i * CSLICE from the pointers): arms in ra/rb (feature key in rb.w), and the world
// midpoint, used only for the merge test, in the C0 row until the final C0 overwrites it.
struct Contacts {
    int n;
    float4 *ra, *rb, *mid;
};
AD size_t crow(int i) { return static_cast<size_t>(i) * CSLICE; }

// collide.cpp:208-233
AD bool addContact(qt qA, v3 cA, qt qB, v3 cB, Contacts &cs, v3 xA, v3 xB, int key)
{
    const v3 mid = (xA + xB) * 0.5f;
    for (int i = 0; i < cs.n; i++)
        if (length2(mid - ld3(cs.mid[crow(i)])) < 1.0e-6f) return false;
    if (cs.n >= MAXC) return false;
    cs.ra[crow(cs.n)] = st3(qrot(qconj(qA), xA - cA));
    cs.rb[crow(cs.n)] = st3(qrot(qconj(qB), xB - cB), __int_as_float(key));
    cs.mid[crow(cs.n)] = st3(mid);
    cs.n++;
    return true;
}

AD void faceManifold(const OBB &bA, const OBB &bB, bool refIsA, int refAxis, v3 nAB, float margin, Contacts &cs,
                     const Poly &c0, const Poly &c1)
{
    // field-wise selects keep both boxes in registers (a selected reference would put them in
    // scratch memory)
    const OBB ref = pickOBB(refIsA, bA, bB);
    const OBB inc = pickOBB(refIsA, bB, bA);
    const v3 refOut = refIsA ? nAB : -nAB;
    const float sgn = dot(refOut, axisOf(ref, refAxis)) >= 0.0f ? 1.0f : -1.0f;
    const v3 fn = axisOf(ref, refAxis) * sgn;
    const v3 fc = ref.c + fn * ref.h[refAxis];
    v3 fu, fv;
    float feu, fev;
    faceAxes(ref, refAxis, fu, fv, feu, fev);
    int incAxis = 0;
    {
        float best = -FLT_MAXV;
        for (int i = 0; i < 3; i++) {
            const float d = fabsf(dot(axisOf(inc, i), fn));
            if (d > best) { best = d; incAxis = i; }
        }
    }
    {
        const float s2 = dot(axisOf(inc, incAxis), fn) > 0.0f ? -1.0f : 1.0f;
        const v3 in = axisOf(inc, incAxis) * s2;
        const v3 ic = inc.c + in * inc.h[incAxis];
        v3 u, v;
        float eu, ev;
        faceAxes(inc, incAxis, u, v, eu, ev);
        c0.put(0, (ic + u * eu) + v * ev);
        c0.put(1, (ic - u * eu) + v * ev);
        c0.put(2, (ic - u * eu) - v * ev);
        c0.put(3, (ic + u * eu) - v * ev);
    }
    int n = 4;
    v3 p;
    p = fu;  n = clipPoly(c0, n, p, dot(p, fc) + feu, c1); if (!n) return;
    p = -fu; n = clipPoly(c1, n, p, dot(p, fc) + feu, c0); if (!n) return;
    p = fv;  n = clipPoly(c0, n, p, dot(p, fc) + fev, c1); if (!n) return;
    p = -fv; n = clipPoly(c1, n, p, dot(p, fc) + fev, c0); if (!n) return;
    int prefix = (refIsA ? 0 : 1) << 24;
    prefix |= (refAxis & 0xFF) << 16;
    prefix |= (incAxis & 0xFF) << 8;
    const float accept = smax(margin, 1.0e-5f);
    for (int i = 0; i < n && cs.n < MAXC; i++) {
        const v3 pi = c0.get(i);
        const float dist = dot(pi - fc, fn);
        if (dist > accept) continue;
        const v3 pr = pi - fn * dist;
        addContact(bA.q, bA.c, bB.q, bB.c, cs, refIsA ? pr : pi, refIsA ? pi : pr, prefix | (i & 0xFF));
    }
    if (!cs.n) {
        const v3 xA = supportPoint(bA, nAB);
        v3 xB = supportPoint(bB, -nAB);
        xB = xA - nAB * dot(xA - xB, nAB);
        addContact(bA.q, bA.c, bB.q, bB.c, cs, xA, xB, prefix);
    }
}

AD void supportEdge(const OBB &b, int ai, v3 dir, v3 &ea, v3 &eb)
{
    const int a1 = (ai + 1) % 3, a2 = (ai + 2) % 3;
    const float s1 = dot(dir, axisOf(b, a1)) >= 0.0f ? 1.0f : -1.0f;
    const float s2 = dot(dir, axisOf(b, a2)) >= 0.0f ? 1.0f : -1.0f;
    const v3 ec = (b.c + axisOf(b, a1) * (b.h[a1] * s1)) + axisOf(b, a2) * (b.h[a2] * s2);
    ea = ec - axisOf(b, ai) * b.h[ai];
    eb = ec + axisOf(b, ai) * b.h[ai];
}

AD void closestSeg(v3 p0, v3 p1, v3 q0, v3 q1, v3 &c0, v3 &c1)
{
    const v3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
    const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    float s = 0.0f, t = 0.0f;
    if (a <= 1.0e-6f && e <= 1.0e-6f) { c0 = p0; c1 = q0; return; }
    if (a <= 1.0e-6f) {
        t = sclamp(f / e, 0.0f, 1.0f);
    } else {
        const float c = dot(d1, r);
        if (e <= 1.0e-6f) {
            s = sclamp(-c / a, 0.0f, 1.0f);
        } else {
            const float bb = dot(d1, d2);
            const float den = a * e - bb * bb;
            if (fabsf(den) > 1.0e-6f) s = sclamp((bb * f - c * e) / den, 0.0f, 1.0f);
            t = (bb * s + f) / e;
            if (t < 0.0f) { t = 0.0f; s = sclamp(-c / a, 0.0f, 1.0f); }
            else if (t > 1.0f) { t = 1.0f; s = sclamp((bb - c) / a, 0.0f, 1.0f); }
        }
    }
    c0 = p0 + d1 * s;
    c1 = q0 + d2 * t;
}

// The separating-axis part of collideBoxes (collide.cpp:238-312): the best face and edge axes,
// or false as soon as an axis separates the boxes by more than the margin (collideBoxes then
// makes no contact).  The pair screen calls it too, so both decide with the same floats.
struct BoxSat {
    bool fValid, eValid;
    int fType, fIdx, eI, eJ;
    float fSep, eSep;
    v3 fN, eN;
};
AD bool boxSat(const OBB &bA, const OBB &bB, float margin, BoxSat &s)
{
    const v3 delta = bB.c - bA.c;
    float M[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) M[i][j] = dot(bA.ax[i], bB.ax[j]);
    const v3 dA{dot(delta, bA.ax[0]), dot(delta, bA.ax[1]), dot(delta, bA.ax[2])};
    const v3 dB{dot(delta, bB.ax[0]), dot(delta, bB.ax[1]), dot(delta, bB.ax[2])};
    // best face axis: type 0 = face of A, 1 = face of B; best edge separately.
    s.fValid = false;
    s.fType = 0;
    s.fIdx = 0;
    s.fSep = -FLT_MAXV;
    s.fN = v3{0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float sep = fabsf(dA[i]) -
                (((bA.h[i] + bB.h.x * fabsf(M[i][0])) + bB.h.y * fabsf(M[i][1])) + bB.h.z * fabsf(M[i][2]));
        if (sep > margin) return false;
        if (!s.fValid || sep > s.fSep) {
            s.fValid = true; s.fType = 0; s.fIdx = i; s.fSep = sep;
            s.fN = bA.ax[i] * (dA[i] < 0.0f ? -1.0f : 1.0f);
        }
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const float sep = fabsf(dB[j]) -
                (((bB.h[j] + bA.h.x * fabsf(M[0][j])) + bA.h.y * fabsf(M[1][j])) + bA.h.z * fabsf(M[2][j]));
        if (sep > margin) return false;
        if (!s.fValid || sep > s.fSep) {
            s.fValid = true; s.fType = 1; s.fIdx = j; s.fSep = sep;
            s.fN = bB.ax[j] * (dB[j] < 0.0f ? -1.0f : 1.0f);
        }
    }
    s.eValid = false;
    s.eI = 0;
    s.eJ = 0;
    s.eSep = -FLT_MAXV;
    s.eN = v3{0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const v3 axis = cross(bA.ax[i], bB.ax[j]);
            const float lenSq = length2(axis);
            if (lenSq < 1.0e-6f) continue;
            const float invLen = 1.0f / sqrtf(lenSq);
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            const float ra = (bA.h[i1] * fabsf(M[i2][j]) + bA.h[i2] * fabsf(M[i1][j])) * invLen;
            const float rb = (bB.h[j1] * fabsf(M[i][j2]) + bB.h[j2] * fabsf(M[i][j1])) * invLen;
            const float sd = dot(delta, axis) * invLen;
            const float sep = fabsf(sd) - (ra + rb);
            if (sep > margin) return false;
            if (!s.eValid || sep > s.eSep) {
                s.eValid = true; s.eI = i; s.eJ = j; s.eSep = sep;
                s.eN = axis * (sd < 0.0f ? -invLen : invLen);
            }
        }
    return true;
}

// Returns the contact count (0 when separated beyond the margin); `basisOut` is written
// exactly when the reference writes it.
AD int collideBoxes(v3 pA, qt qA, v3 hA, v3 pB, qt qB, v3 hB, float margin, m3 &basisOut, bool &wroteBasis,
                    Contacts &cs, const Poly &c0, const Poly &c1)
{
    cs.n = 0;
    wroteBasis = false;
    const OBB bA = makeOBB(pA, qA, hA);
    const OBB bB = makeOBB(pB, qB, hB);
    BoxSat sat;
    if (!boxSat(bA, bB, margin, sat)) return 0;
    const bool fValid = sat.fValid, eValid = sat.eValid;
    const int fType = sat.fType, fIdx = sat.fIdx, eI = sat.eI, eJ = sat.eJ;
    const float fSep = sat.fSep, eSep = sat.eSep;
    const v3 fN = sat.fN, eN = sat.eN;
    if (!fValid) return 0;
    const bool useEdge = eValid && 0.95f * eSep > fSep + 0.01f;
    const v3 nAB = useEdge ? eN : fN;
    basisOut = orthonormal(-nAB);
    wroteBasis = true;
    if (useEdge) {
        v3 a0, a1, b0, b1;
        supportEdge(bA, eI, nAB, a0, a1);
        supportEdge(bB, eJ, -nAB, b0, b1);
        v3 xA, xB;
        closestSeg(a0, a1, b0, b1, xA, xB);
        const int key = (2 << 24) | ((eI & 0xFF) << 8) | (eJ & 0xFF);
        addContact(qA, pA, qB, pB, cs, xA, xB, key);
        if (!cs.n) {
            xA = supportPoint(bA, nAB);
            xB = supportPoint(bB, -nAB);
            addContact(qA, pA, qB, pB, cs, xA, xB, key);
        }
        return cs.n;
    }
    faceManifold(bA, bB, fType == 0, fIdx, nAB, margin, cs, c0, c1);
    return cs.n;
}

// ------------------------------------------------------------ round shapes ----
// collide.cpp:505-1032: spheres and capped cylinders.  A shape holds what makeShape derives from
// a body (collide.cpp:505-516): half extents, radius = size.x, half height = size.z / 2 and the
// cylinder axis (local +Z) in world space.
struct RShp {
    v3 c; qt q; v3 half; float r, hh; v3 axis;
};
AD RShp makeRShp(v3 p, qt q, float4 size)
{
    return RShp{p, q, v3{size.x, size.y, size.z} * 0.5f, size.x, size.z * 0.5f, qrot(q, v3{0, 0, 1})};
}
AD v3 sel3v(bool c, v3 a, v3 b) { return v3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }
// field-wise select: both shapes stay in registers (a selected reference would put them in
// scratch memory)
AD RShp selR(bool c, const RShp &a, const RShp &b)
{
    return RShp{sel3v(c, a.c, b.c),
                qt{c ? a.q.x : b.q.x, c ? a.q.y : b.q.y, c ? a.q.z : b.q.z, c ? a.q.w : b.q.w},
                sel3v(c, a.half, b.half), c ? a.r : b.r, c ? a.hh : b.hh, sel3v(c, a.axis, b.axis)};
}
AD v3 clampBox(v3 l, v3 h)
{
    return v3{sclamp(l.x, -h.x, h.x), sclamp(l.y, -h.y, h.y), sclamp(l.z, -h.z, h.z)};
}

// collide.cpp:520-546: inside a box, the way out along the axis of least penetration
AD bool pointInBox(const RShp &b, v3 p, v3 &push, float &depth)
{
    const v3 l = qrot(qconj(b.q), p - b.c);
    float best = FLT_MAXV, sign = 1.0f;
    int axis = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float d = b.half[i] - fabsf(l[i]);
        if (d <= 0.0f) return false;
        if (d < best) { best = d; axis = i; sign = l[i] >= 0.0f ? 1.0f : -1.0f; }
    }
    push = qrot(b.q, v3{axis == 0 ? sign : 0.0f, axis == 1 ? sign : 0.0f, axis == 2 ? sign : 0.0f});
    depth = best;
    return true;
}
// collide.cpp:549-577: outside but within the margin: negative depth, outward push
AD bool pointNearBox(const RShp &b, v3 p, float margin, v3 &push, float &depth)
{
    const v3 l = qrot(qconj(b.q), p - b.c);
    if (!(fabsf(l.x) >= b.half.x || fabsf(l.y) >= b.half.y || fabsf(l.z) >= b.half.z)) return false;
    const v3 d = l - clampBox(l, b.half);
    const float gap = length(d);
    if (gap > margin || gap <= 1.0e-9f) return false;
    push = qrot(b.q, d / gap);
    depth = -gap;
    return true;
}
// collide.cpp:580-614
AD v3 closestOnCylinder(const RShp &cy, v3 p, v3 &outward, float &distance)
{
    const v3 l = qrot(qconj(cy.q), p - cy.c);
    const float t = sclamp(l.z, -cy.hh, cy.hh);
    const float lenSq = (0.0f + l.x * l.x) + l.y * l.y;
    float cx, cy2;
    if (lenSq > cy.r * cy.r) {
        const float scale = cy.r / sqrtf(lenSq);
        cx = l.x * scale; cy2 = l.y * scale;
    } else if (lenSq > 1.0e-12f) {
        cx = l.x; cy2 = l.y;
    } else {
        cx = cy.r; cy2 = 0.0f;
    }
    const v3 w = qrot(cy.q, v3{cx, cy2, t}) + cy.c;
    const v3 d = p - w;
    distance = length(d);
    outward = distance > 1.0e-9f ? d / distance : qrot(cy.q, v3{1, 0, 0});
    return w;
}
// collide.cpp:617-651
AD bool pointInCylinder(const RShp &cy, v3 p, v3 &push, float &surface)
{
    const v3 l = qrot(qconj(cy.q), p - cy.c);
    const float rs = l.x * l.x + l.y * l.y;
    if (rs >= cy.r * cy.r || fabsf(l.z) >= cy.hh) return false;
    const float radialDepth = cy.r - sqrtf(rs);
    const float capDepth = cy.hh - fabsf(l.z);
    v3 pl{0, 0, 0};
    if (capDepth < radialDepth) {
        pl.z = l.z >= 0.0f ? 1.0f : -1.0f;
        surface = capDepth;
    } else {
        const float len = sqrtf(rs);
        if (len > 1.0e-6f) { pl.x = l.x / len; pl.y = l.y / len; }
        else pl.x = 1.0f;
        surface = radialDepth;
    }
    push = qrot(cy.q, pl);
    return true;
}
// collide.cpp:654-671
AD bool pointNearCylinder(const RShp &cy, v3 p, float margin, v3 &push, float &depth)
{
    const v3 l = qrot(qconj(cy.q), p - cy.c);
    const float rs = l.x * l.x + l.y * l.y;
    if (rs < cy.r * cy.r && fabsf(l.z) < cy.hh) return false;
    v3 out;
    float dist;
    closestOnCylinder(cy, p, out, dist);
    if (dist > margin) return false;
    push = out;
    depth = -dist;
    return true;
}
// collide.cpp:674-684: the exact inside test first, the speculative one only when it fails
AD bool insideBox(const RShp &b, v3 p, float margin, v3 &push, float &depth)
{
    return pointInBox(b, p, push, depth) || (margin > 0.0f && pointNearBox(b, p, margin, push, depth));
}
AD bool insideCylinder(const RShp &cy, v3 p, float margin, v3 &push, float &depth)
{
    return pointInCylinder(cy, p, push, depth) || (margin > 0.0f && pointNearCylinder(cy, p, margin, push, depth));
}

// Surface sample k (0..33) of a cylinder (collide.cpp:755-784): caps (centre, then 8 rim
// points) at -h/2 and +h/2, then two inner rings.  x, y: the body's local X and Y in world.
AD v3 cylSample(const RShp &cy, v3 x, v3 y, const float *rim, int k)
{
    v3 base;
    int i;
    if (k < 18) {
        const int end = k < 9 ? -1 : 1;
        const int j = k < 9 ? k : k - 9;
        base = cy.c + cy.axis * (cy.hh * static_cast<float>(end));
        if (j == 0) return base;
        i = j - 1;
    } else {
        const int ring = k < 26 ? 1 : 2;
        i = (k - 18) & 7;
        const float t = -cy.hh + 2.0f * cy.hh * static_cast<float>(ring) / static_cast<float>(3);
        base = cy.c + cy.axis * t;
    }
    return base + (x * rim[i] + y * rim[8 + i]) * cy.r;
}

constexpr int CYL_SAMPLES = 34;
// Candidate k of a sampled pair, in the order the reference generates them (collide.cpp:894-998).
// cylBox: samples of the cylinder inside the box (k < 34), then box corners inside the cylinder;
// otherwise (two cylinders) samples of A inside B, then samples of B inside A.  Returns false
// when the point is no candidate.
AD bool roundCandidate(bool cylBox, bool cylIsA, const RShp &a, const RShp &b, const float *rim, float margin, int k,
                       v3 &xA, v3 &xB, v3 &n, float &depth, int &key)
{
    const bool cyIsA = cylBox ? cylIsA : k < CYL_SAMPLES;
    const RShp cy = selR(cyIsA, a, b);
    const RShp into = selR(!cyIsA, a, b);
    const int i = k < CYL_SAMPLES ? k : k - CYL_SAMPLES;
    v3 push;
    float d;
    if (cylBox && k >= CYL_SAMPLES) {
        // box corner i inside the cylinder
        const v3 cl{(i & 1) ? into.half.x : -into.half.x, (i & 2) ? into.half.y : -into.half.y,
                    (i & 4) ? into.half.z : -into.half.z};
        const v3 cw = qrot(into.q, cl) + into.c;
        if (!insideCylinder(cy, cw, margin, push, d)) return false;
        xA = cylIsA ? cw + push * d : cw;
        xB = cylIsA ? cw : cw + push * d;
        n = cylIsA ? push : -push;
        depth = d;
        key = static_cast<int>((4u << 24) | ((cylIsA ? 0u : 1u) << 20) | static_cast<uint32_t>(i));
        return true;
    }
    const v3 x = qrot(cy.q, v3{1, 0, 0}), y = qrot(cy.q, v3{0, 1, 0});
    const v3 smp = cylSample(cy, x, y, rim, i);
    const bool fromIsA = cylBox ? cylIsA : k < CYL_SAMPLES;
    if (!(cylBox ? insideBox(into, smp, margin, push, d) : insideCylinder(into, smp, margin, push, d))) return false;
    xA = fromIsA ? smp : smp + push * d;
    xB = fromIsA ? smp + push * d : smp;
    n = fromIsA ? -push : push;
    depth = d;
    key = static_cast<int>((3u << 24) | ((fromIsA ? 1u : 0u) << 20) | static_cast<uint32_t>(i));
    return true;
}

// Sampled pairs (cylinder-box, cylinder-cylinder): every candidate is offered in generation
// order; the deepest MAXC are kept in a register list, stable (an equal depth stays behind the
// earlier candidate), which is the head of the reference's stable insertion ranking
// (collide.cpp:702-730).  The kept ones are then rebuilt and offered to the manifold in rank
// order, stopping at the first it refuses.
AD int collideSampled(bool cylBox, bool cylIsA, const RShp &a, const RShp &b, const float *rim, float margin,
                      m3 &basisOut, Contacts &cs)
{
    const int total = cylBox ? CYL_SAMPLES + 8 : 2 * CYL_SAMPLES;
    float td[MAXC] = {};
    int tk[MAXC] = {};
    int kept = 0, nc = 0;
    v3 nAB{0, 0, 0};
    for (int k = 0; k < total; k++) {
        v3 xA, xB, n;
        float d;
        int key;
        if (!roundCandidate(cylBox, cylIsA, a, b, rim, margin, k, xA, xB, n, d, key)) continue;
        nc++;
        nAB += n;
        // insert before the first strictly shallower entry; everything after it moves down one
        float cd = d;
        int ck = k;
        bool carry = true, shifting = false;
#pragma unroll
        for (int j = 0; j < MAXC; j++) {
            const bool empty = j >= kept;
            if (carry && (shifting || empty || td[j] < cd)) {
                const float od = td[j];
                const int ok = tk[j];
                td[j] = cd; tk[j] = ck;
                cd = od; ck = ok;
                shifting = true;
                if (empty) carry = false;
            }
        }
        kept = kept < MAXC ? kept + 1 : MAXC;
    }
    if (!nc) return 0;
    const v3 axis = sel3v(!cylBox || cylIsA, a.axis, b.axis);
    basisOut = orthonormal(-(length2(nAB) > 1.0e-12f ? normalize(nAB) : axis));
#pragma unroll
    for (int j = 0; j < MAXC; j++) {
        if (j >= kept) break;
        v3 xA, xB, n;
        float d;
        int key;
        roundCandidate(cylBox, cylIsA, a, b, rim, margin, tk[j], xA, xB, n, d, key);
        if (!addContact(a.q, a.c, b.q, b.c, cs, xA, xB, key)) break;
    }
    return cs.n;
}

// collide.cpp:1001-1032 for pairs with a sphere or a cylinder: contact count, basis written
// exactly when the reference writes it (when there is a contact).
AD int collideRound(int typeA, int typeB, const RShp &a, const RShp &b, const float *rim, float margin,
                    m3 &basisOut, bool &wroteBasis, Contacts &cs)
{
    cs.n = 0;
    wroteBasis = false;
    const int S = AVBD_SHAPE_SPHERE, C = AVBD_SHAPE_CYLINDER;
    if (typeA == S || typeB == S) {
        const bool sphereIsA = typeA == S;
        const RShp sp = selR(sphereIsA, a, b);
        const RShp other = selR(!sphereIsA, a, b);
        const int otherType = sphereIsA ? typeB : typeA;
        v3 n, onOther;
        if (otherType == S) {
            // collide.cpp:786-806 (both spheres; n runs from A to B)
            const v3 d = b.c - a.c;
            const float distSq = length2(d);
            const float sumPlus = a.r + b.r + margin;
            if (distSq >= sumPlus * sumPlus) return 0;
            const float dist = sqrtf(distSq);
            const v3 nAB = dist > 1.0e-6f ? d / dist : v3{0, 0, 1};
            basisOut = orthonormal(-nAB);
            wroteBasis = true;
            addContact(a.q, a.c, b.q, b.c, cs, a.c + nAB * a.r, b.c - nAB * b.r, 1);
            return cs.n;
        }
        if (otherType == C) {
            // collide.cpp:865-891; n leaves the cylinder towards the sphere centre
            float dist;
            onOther = closestOnCylinder(other, sp.c, n, dist);
            if (dist >= sp.r + margin) return 0;
        } else {
            // collide.cpp:809-862; n leaves the box towards the sphere centre
            const v3 l = qrot(qconj(other.q), sp.c - other.c);
            onOther = qrot(other.q, clampBox(l, other.half)) + other.c;
            const v3 d = sp.c - onOther;
            const float distSq = length2(d);
            const float radPlus = sp.r + margin;
            if (distSq >= radPlus * radPlus) return 0;
            const float dist = sqrtf(distSq);
            if (dist > 1.0e-6f) {
                n = d / dist;
            } else {
                float depth;
                if (!pointInBox(other, sp.c, n, depth)) return 0;
            }
        }
        basisOut = orthonormal(-(sphereIsA ? -n : n));
        wroteBasis = true;
        const v3 onSphere = sp.c - n * sp.r;
        addContact(a.q, a.c, b.q, b.c, cs, sphereIsA ? onSphere : onOther, sphereIsA ? onOther : onSphere, 1);
        return cs.n;
    }
    const bool cylBox = !(typeA == C && typeB == C);
    const int n = collideSampled(cylBox, typeA == C, a, b, rim, margin, basisOut, cs);
    wroteBasis = n > 0;
    return n;
}

// ------------------------------------------------------ projection bits ----
// solver_projection.cpp:33-53
AD float angCompliance(v3 I, v3 t)
{
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (I[k] > 0.0f) sum += t[k] * t[k] / I[k];
    return sum;
}
AD v3 angResponse(v3 I, qt q, v3 t)
{
    v3 o{0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (I[k] > 0.0f) o[k] = t[k] / I[k];
    return qrot(q, o);
}
AD void maskLocked(uint8_t locks, v3 &v)
{
    if (locks) {
#pragma unroll
        for (int k = 0; k < 3; k++)
            if (locks & (1 << k)) v[k] = 0.0f;
    }
}

} // namespace avbd_dev

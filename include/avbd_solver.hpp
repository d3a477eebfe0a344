// avbd_solver.hpp - C++ facade with the reference solver's interface over the C ABI.
//
// A drop-in for the reference's in-process API (src/avbd/solver.h:50-374,
// src/avbd/types.hpp:75-139, src/avbd/forces.hpp:19-113): bodies and forces are created with
// `new Rigid(&solver, size[, shape], density, friction, position[, velocity])`,
// `new Joint(&solver, a, b, rA, rB[, kLin, kAng, fracture])`,
// `new Spring(&solver, a, b, rA, rB, stiffness[, rest])`, `new IgnoreCollision(&solver, a, b)`
// and `new GenericJoint(&solver, a, b, rA, rB)` (then its `linear[i]` / `angular[i]` axes set),
// which prepend themselves to the solver's intrusive lists exactly as the reference's do; the
// caller sets the same public fields, calls step() / stepInterleaved(), and reads the same
// `Rigid` / `Joint` fields.
// The state lives on the GPU: each step writes the bodies' dynamic fields to the device,
// steps there (avbd_world_step), and reads them back.  A topology change (a body or joint
// created or deleted) re-uploads the lists and carries every surviving manifold over, so
// warm starting continues as in the reference.  The solver deletes a broken joint during the
// step after it broke, as the reference does (solver_solve.cpp:39-51).
//
// Errors: the reference has no error path; here a failed device call leaves the state
// unchanged and sets lastStatus / lastError().  Header-only; link libavbd_hip.so.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>

#include "avbd_hip.h"

namespace avbd_hip {

struct float3 {
    float x = 0, y = 0, z = 0;
    float &operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
};
struct quat {
    float x = 0, y = 0, z = 0, w = 1;
};

// reference types.hpp:31-33
constexpr float COLLISION_MARGIN = 0.001f;

// The few vector and matrix operations host code written against the reference uses on
// solver output (the server's contact loops): linalg's component-wise operators and folds.
inline float3 operator+(float3 a, float3 b) { return float3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline float3 operator-(float3 a, float3 b) { return float3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float3 operator-(float3 a) { return float3{-a.x, -a.y, -a.z}; }
inline float3 operator*(float3 a, float s) { return float3{a.x * s, a.y * s, a.z * s}; }
inline float3 operator*(float s, float3 a) { return float3{s * a.x, s * a.y, s * a.z}; }
inline float dot(float3 a, float3 b) { return ((0.0f + a.x * b.x) + a.y * b.y) + a.z * b.z; }
inline float length(float3 a) { return std::sqrt(dot(a, a)); }

// 3x3 in the reference's layout (maths.h:12: m[col][row])
struct float3x3 {
    float3 x, y, z;   // columns
    float3 &operator[](int c) { return (&x)[c]; }
    const float3 &operator[](int c) const { return (&x)[c]; }
};
inline float3 row(const float3x3 &m, int r) { return float3{m.x[r], m.y[r], m.z[r]}; }
inline float3 mul(const float3x3 &m, float3 v) { return (m.x * v.x + m.y * v.y) + m.z * v.z; }

struct Solver;

// reference types.hpp:52-59: `size` is the full extents of a box, (radius, -, -) of a sphere,
// (radius, -, height) of a cylinder whose axis is the body's local +Z
enum class ShapeType : int { Box = AVBD_SHAPE_BOX, Sphere = AVBD_SHAPE_SPHERE, Cylinder = AVBD_SHAPE_CYLINDER };

// reference types.hpp:75-139
struct Rigid {
    Solver *solver;
    Rigid *next = nullptr;
    float3 positionLin;
    quat positionAng;
    float3 velocityLin;
    float3 velocityAng;
    float3 prevVelocityLin;
    float3 size;
    ShapeType shape = ShapeType::Box;
    float mass = 0;
    float3 moment;
    float friction = 0;
    float radius = 0;
    float gravity = 0;
    uint32_t collisionLayer = 1;
    uint32_t collisionMask = 1;
    uint8_t axisLockLinear = 0;
    uint8_t axisLockAngular = 0;
    bool sleeping = false;   // types.hpp:97-100, 138: asleep; steps below the thresholds; 0 NEVER,
    int stillFrames = 0;     // 1 SLEEP, 2 START_IN_SLEEP (set before the first step)
    int sleep_mode = 0;

    inline Rigid(Solver *p_solver, float3 p_size, float p_density, float p_friction, float3 p_position,
                 float3 p_velocity = float3{0, 0, 0});
    inline Rigid(Solver *p_solver, float3 p_size, ShapeType p_shape, float p_density, float p_friction,
                 float3 p_position, float3 p_velocity = float3{0, 0, 0});
    inline ~Rigid();
    Rigid(const Rigid &) = delete;
    Rigid &operator=(const Rigid &) = delete;
};

// reference forces.hpp:19-55
struct Force {
    Solver *solver;
    Rigid *bodyA;
    Rigid *bodyB;
    Force *next = nullptr;
    void *owner = nullptr;
    uint8_t kind = AVBD_FORCE_JOINT;   // which of the force types below this is
    inline Force(Solver *p_solver, Rigid *p_bodyA, Rigid *p_bodyB);
    inline virtual ~Force();
    virtual bool isBroken() const { return false; }
    Force(const Force &) = delete;
    Force &operator=(const Force &) = delete;
};

// reference forces.hpp:58-87 (joint.cpp): linear ball socket + optional angular lock
struct Joint : Force {
    float3 rA, rB;
    float3 C0Lin, C0Ang;
    float3 penaltyLin, penaltyAng;
    float3 lambdaLin, lambdaAng;
    float stiffnessLin, stiffnessAng, fracture;
    float torqueArm;
    bool broken = false;

    inline Joint(Solver *p_solver, Rigid *p_bodyA, Rigid *p_bodyB, float3 p_rA, float3 p_rB,
                 float p_stiffnessLin = std::numeric_limits<float>::infinity(), float p_stiffnessAng = 0.0f,
                 float p_fracture = std::numeric_limits<float>::infinity());
    // Break now, discarding dual state (joint.cpp:162-169).
    void breakNow()
    {
        penaltyLin = penaltyAng = lambdaLin = lambdaAng = float3{};
        broken = true;
        dirty = true;
    }
    bool isBroken() const override { return broken; }
    bool dirty = false;   // host-side edit to push before the next step
};

// reference forces.hpp:89-101 (spring.cpp): a spring between two bodies; a negative rest
// length takes the anchors' distance at creation.
struct Spring : Force {
    float3 rA, rB;
    float rest;
    float stiffness;
    inline Spring(Solver *p_solver, Rigid *p_bodyA, Rigid *p_bodyB, float3 p_rA, float3 p_rB, float p_stiffness,
                  float p_rest = -1);
};

// reference forces.hpp:103-113: no physical effect; the two bodies skip each other in the
// broad phase.
struct IgnoreCollision : Force {
    inline IgnoreCollision(Solver *p_solver, Rigid *p_bodyA, Rigid *p_bodyB);
};

// reference forces.hpp:115-121
enum class AxisMode : int { Free = AVBD_AXIS_FREE, Locked = AVBD_AXIS_LOCKED, Limited = AVBD_AXIS_LIMITED };

// reference forces.hpp:123-139: one degree of freedom of a GenericJoint and its dual state
struct JointAxis {
    AxisMode mode = AxisMode::Free;
    float lower = 0.0f;
    float upper = 0.0f;
    float springStiffness = 0.0f;
    float springEquilibrium = 0.0f;
    float springDamping = 0.0f;
    float lambda = 0.0f;
    float penalty = 0.0f;
    float initial = 0.0f;
};

// reference forces.hpp:141-190 (generic_joint.cpp): six degrees of freedom in A's frame (the
// world's if A is null), each free, locked or limited, with an optional spring; the zero is
// the arrangement at creation.
struct GenericJoint : Force {
    float3 rA, rB;
    quat rest;
    JointAxis linear[3];
    JointAxis angular[3];
    inline GenericJoint(Solver *p_solver, Rigid *p_bodyA, Rigid *p_bodyB, float3 p_rA, float3 p_rB);
};

// reference maths.h:65-71 (Eigen's order) and linalg's folds, for the host-side rest length
inline float3 hostCross(float3 a, float3 b)
{
    return float3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float3 hostQrot(quat q, float3 v)
{
    const float3 z{q.x, q.y, q.z};
    float3 uv = hostCross(z, v);
    uv = float3{uv.x + uv.x, uv.y + uv.y, uv.z + uv.z};
    const float3 c = hostCross(z, uv);
    return float3{(v.x + uv.x * q.w) + c.x, (v.y + uv.y * q.w) + c.y, (v.z + uv.z * q.w) + c.z};
}

// reference maths.h:87-95: Hamilton product, the reference's term order
inline quat hostQmul(quat a, quat b)
{
    return quat{((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y, ((a.w * b.y + a.y * b.w) + a.z * b.x) - a.x * b.z,
                ((a.w * b.z + a.z * b.w) + a.x * b.y) - a.y * b.x, ((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z};
}

inline float3 qrot(quat q, float3 v) { return hostQrot(q, v); }

// reference manifold.hpp:28-63: a contact as the manifold accessors hand it out, and the
// narrow phase for two bodies at their current pose.
struct Manifold {
    union FeaturePair {
        struct {
            char inR;
            char outR;
            char inI;
            char outI;
        };
        int key;
    };
    struct Contact {
        FeaturePair feature;
        float3 rA;   // contact offset in A's local space
        float3 rB;   // contact offset in B's local space
    };
    // collide.cpp:1100-1120: the solver's narrow phase (speculative margin grown by the pair's
    // predicted relative motion over the solver's collisionDt), run on the device through
    // avbd_world_collide_pair.  `contacts` is any contiguous range of Contact (std::span,
    // std::array, ...); returns the count written.
    template <class Range> static int collide(Rigid *bodyA, Rigid *bodyB, Range &&contacts, float3x3 &basis);
};

// reference solver.h:50-144
struct Solver {
    float dt = 1.0f / 60.0f;
    float collisionDt = 1.0f / 60.0f;
    float gravity = -10.0f;
    int iterations = 8;
    float alpha = 0.99f;
    float betaLin = 10000.0f;
    float betaAng = 100.0f;
    float gamma = 0.999f;
    int roundsThisStep = 8;
    float newtonRatio = 1.0f;
    float stiffnessDecay = 1.0f;
    bool adaptiveBeta = false;
    float adaptiveBetaErrorFloor = 1e-3f;
    float convergenceThreshold = 0.0f;
    float jacobiRelaxation = 0.7f;
    bool contactOnlyOverflow = true;
    float stickDriftHeal = 0.002f;
    float maxPenetration = 0.02f;
    float jacobianRebuildDistance = 0.0f;
    float speculativeMargin = 0.01f;
    bool positionProjection = true;
    int projectionIterations = 4;
    float projectionBeta = 1.0f;
    float projectionSlop = 0.008f;
    float projectionMaxCorrection = 0.0f;
    int contactGraceTicks = 0;
    int threads = 0;
    float sleepLinearThreshold = 0.1f;
    float sleepAngularThreshold = 0.139626f;
    int sleepFrames = 0;
    // The reference's compile-time PENALTY_MAX (types.hpp:26) as a field.
    float penaltyMax = 1e10f;

    Rigid *bodies = nullptr;
    Force *forces = nullptr;

    avbd_status lastStatus = AVBD_OK;
    const char *lastError() const { return avbd_last_error(); }

    Solver() = default;
    ~Solver()
    {
        clear();
        if (world) avbd_world_destroy(world);
    }
    Solver(const Solver &) = delete;
    Solver &operator=(const Solver &) = delete;

    // Delete every body and force (solver_step.cpp:29-40).
    void clear()
    {
        while (forces) delete forces;
        while (bodies) delete bodies;
        topologyChanged = true;
        uploaded.clear();
        uploadedJoints.clear();
    }

    void step() { advance(1, 0); }
    void stepInterleaved(int substeps, int roundsPerSubstep) { advance(substeps, roundsPerSubstep); }

    // Manifold accessors (solver.h:168-231).  `index` is a position in the live list; the live
    // manifolds are read from the device once per step, on the first accessor call.
    int manifoldCount() const { return loadManifolds() ? static_cast<int>(mfA.size()) : 0; }
    int manifoldContactPoints() const
    {
        if (!loadManifolds()) return 0;
        int total = 0;
        for (int32_t c : mfNum) total += c;
        return total;
    }
    Rigid *manifoldBodyA(int index) const { return loadManifolds() ? uploaded[static_cast<size_t>(mfA[index])] : nullptr; }
    Rigid *manifoldBodyB(int index) const { return loadManifolds() ? uploaded[static_cast<size_t>(mfB[index])] : nullptr; }
    int manifoldContactCount(int index) const { return loadManifolds() ? mfNum[static_cast<size_t>(index)] : 0; }
    const Manifold::Contact &manifoldContact(int index, int k) const
    {
        loadManifolds();
        return mfContacts[8 * static_cast<size_t>(index) + static_cast<size_t>(k)];
    }
    // Constraint value at the step-start pose: x is the normal error (positive = overlap).
    float3 manifoldContactError(int index, int k) const
    {
        loadManifolds();
        const float *c = &mfC0[24 * static_cast<size_t>(index) + 3 * static_cast<size_t>(k)];
        return float3{c[0], c[1], c[2]};
    }
    // First basis row: the contact normal, pointing from B to A.
    float3 manifoldNormal(int index) const
    {
        loadManifolds();
        const float *b = &mfBasis[9 * static_cast<size_t>(index)];
        return float3{b[0], b[1], b[2]};
    }
    // Diagnostics (solver.h:171-174, 219-220): colour groups in use, bodies in the Jacobi
    // bucket, candidate pairs of the last broad phase.
    int colourGroups() const { return stat().colours; }
    int overflowBodies() const { return stat().overflow_bodies; }
    int broadPhasePairCount() const { return stat().candidate_pairs; }
    // solver.h:242-243: host threads driving the step (one; the step itself runs on the GPU).
    int threadCount() const { return 1; }

    // solver.h:245-275: per-phase device time of the last step (ms, HIP events).  "graph" is the
    // warm start of the forces, the narrow phase and the constraint graph with its colouring.
    struct PhaseTimes {
        double broadPhase = 0.0;
        double broadIndex = 0.0;
        double broadTree = 0.0;
        double broadEmit = 0.0;
        double graph = 0.0;
        double solve = 0.0;
        double finish = 0.0;
    };
    PhaseTimes lastStepTimes;
    double get_broadPhase_time() const { return lastStepTimes.broadPhase; }
    double get_broadTree_time() const { return lastStepTimes.broadTree; }
    double get_broadEmit_time() const { return lastStepTimes.broadEmit; }
    double get_colourGraph_time() const { return lastStepTimes.graph; }
    double get_solve_time() const { return lastStepTimes.solve; }
    double get_finish_time() const { return lastStepTimes.finish; }
    // Solver::pick (solver.h:161-164, solver_query.cpp:29-155): the nearest body along the ray
    // (null if none) and the hit point in its frame; every body is tested, static ones too.
    Rigid *pick(float3 origin, float3 dir, float3 &local) { return pick(origin, dir, local, 0xFFFFFFFFu); }
    Rigid *pick(float3 origin, float3 dir, float3 &local, uint32_t mask)
    {
        std::vector<Rigid *> list;
        std::vector<Force *> jl;
        if (!bodies || !pushState(list, jl)) return nullptr;
        const float o[3] = {origin.x, origin.y, origin.z}, d[3] = {dir.x, dir.y, dir.z};
        int32_t body = -1;
        float loc[3];
        lastStatus = avbd_world_pick(world, 1, o, d, &mask, &body, loc);
        if (lastStatus != AVBD_OK || body < 0) return nullptr;
        local = float3{loc[0], loc[1], loc[2]};
        return list[static_cast<size_t>(body)];
    }
    // Solver::collectAABB (solver.h:165-167, solver_query.cpp:158-171): appends, in list order,
    // the bodies whose layer passes the mask and whose bounds overlap [min, max].
    void collectAABB(float3 min, float3 max, uint32_t mask, std::vector<Rigid *> &out)
    {
        std::vector<Rigid *> list;
        std::vector<Force *> jl;
        if (!bodies || !pushState(list, jl)) return;
        std::vector<int32_t> idx(list.size());
        int32_t count = 0;
        const float lo[3] = {min.x, min.y, min.z}, hi[3] = {max.x, max.y, max.z};
        lastStatus = avbd_world_collect_aabb(world, lo, hi, mask, idx.data(), static_cast<int32_t>(idx.size()), &count);
        if (lastStatus != AVBD_OK) return;
        for (int32_t k = 0; k < count; k++) out.push_back(list[static_cast<size_t>(idx[k])]);
    }

    avbd_step_stats lastStepStats()
    {
        avbd_step_stats s{};
        if (world) avbd_world_stats(world, &s);
        return s;
    }

    // --- internal ---------------------------------------------------------------
    bool topologyChanged = true;

    // Manifold::collide for two of this solver's bodies (the device's narrow phase).
    int collidePair(Rigid *a, Rigid *b, Manifold::Contact *out, size_t capacity, float3x3 &basis)
    {
        std::vector<Rigid *> list;
        std::vector<Force *> jl;
        if (!pushState(list, jl)) return 0;
        int ia = -1, ib = -1;
        for (size_t i = 0; i < list.size(); i++) {
            if (list[i] == a) ia = static_cast<int>(i);
            if (list[i] == b) ib = static_cast<int>(i);
        }
        if (ia < 0 || ib < 0) return 0;
        int32_t count = 0, feat[8];
        float ra[24], rb[24], bas[9];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) bas[3 * r + c] = basis[c][r];
        lastStatus = avbd_world_collide_pair(world, ia, ib, collisionDt, &count, feat, ra, rb, bas);
        if (lastStatus != AVBD_OK) return 0;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) basis[c][r] = bas[3 * r + c];
        const int n = static_cast<int>(std::min<size_t>(static_cast<size_t>(count), capacity));
        for (int k = 0; k < n; k++) {
            out[k].feature.key = feat[k];
            out[k].rA = float3{ra[3 * k], ra[3 * k + 1], ra[3 * k + 2]};
            out[k].rB = float3{rb[3 * k], rb[3 * k + 1], rb[3 * k + 2]};
        }
        return n;
    }

  private:
    avbd_world *world = nullptr;
    // live manifolds of the last step, read on the first accessor call after it
    mutable bool mfValid = false;
    mutable std::vector<int32_t> mfA, mfB, mfNum;
    mutable std::vector<float> mfBasis, mfC0;
    mutable std::vector<Manifold::Contact> mfContacts;

    avbd_step_stats stat() const
    {
        avbd_step_stats s{};
        if (world) avbd_world_stats(world, &s);
        return s;
    }

    bool loadManifolds() const
    {
        if (mfValid) return true;
        if (!world) return false;
        const int m = stat().manifolds;
        mfA.assign(m, 0); mfB.assign(m, 0); mfNum.assign(m, 0);
        mfBasis.assign(9 * static_cast<size_t>(m), 0.0f);
        mfC0.assign(24 * static_cast<size_t>(m), 0.0f);
        mfContacts.assign(8 * static_cast<size_t>(m), Manifold::Contact{});
        std::vector<float> fr(m), ra(24 * static_cast<size_t>(m)), rb(24 * static_cast<size_t>(m)),
                pen(24 * static_cast<size_t>(m)), lam(24 * static_cast<size_t>(m)), beta(8 * static_cast<size_t>(m));
        std::vector<int32_t> feat(8 * static_cast<size_t>(m));
        std::vector<uint8_t> stick(8 * static_cast<size_t>(m));
        avbd_manifolds mf{m, mfA.data(), mfB.data(), mfBasis.data(), fr.data(), mfNum.data(), feat.data(), ra.data(),
                          rb.data(), mfC0.data(), pen.data(), lam.data(), stick.data(), beta.data()};
        if (m > 0 && avbd_world_get_manifolds(world, &mf) != AVBD_OK) return false;
        for (size_t k = 0; k < 8 * static_cast<size_t>(m); k++) {
            mfContacts[k].feature.key = feat[k];
            mfContacts[k].rA = float3{ra[3 * k], ra[3 * k + 1], ra[3 * k + 2]};
            mfContacts[k].rB = float3{rb[3 * k], rb[3 * k + 1], rb[3 * k + 2]};
        }
        mfValid = true;
        return true;
    }
    int capBodies = 0, capJoints = 0;
    std::vector<Rigid *> uploaded;
    std::vector<Force *> uploadedJoints;

    avbd_params params() const
    {
        avbd_params p;
        avbd_params_default(&p);
        p.dt = dt;
        p.gravity = gravity;
        p.iterations = iterations;
        p.alpha = alpha;
        p.beta_lin = betaLin;
        p.beta_ang = betaAng;
        p.gamma = gamma;
        p.newton_ratio = newtonRatio;
        p.stiffness_decay = stiffnessDecay;
        p.adaptive_beta = adaptiveBeta ? 1 : 0;
        p.adaptive_beta_error_floor = adaptiveBetaErrorFloor;
        p.convergence_threshold = convergenceThreshold;
        p.jacobi_relaxation = jacobiRelaxation;
        p.contact_only_overflow = contactOnlyOverflow ? 1 : 0;
        p.stick_drift_heal = stickDriftHeal;
        p.max_penetration = maxPenetration;
        p.jacobian_rebuild_distance = jacobianRebuildDistance;
        p.speculative_margin = speculativeMargin;
        p.position_projection = positionProjection ? 1 : 0;
        p.projection_iterations = projectionIterations;
        p.projection_beta = projectionBeta;
        p.projection_slop = projectionSlop;
        p.projection_max_correction = projectionMaxCorrection;
        p.contact_grace_ticks = contactGraceTicks;
        p.threads = threads;
        p.sleep_linear_threshold = sleepLinearThreshold;
        p.sleep_angular_threshold = sleepAngularThreshold;
        p.sleep_frames = sleepFrames;
        p.penalty_max = penaltyMax;
        return p;
    }

    struct Arrays {
        std::vector<float> size, mass, moment, friction, radius, gravity, pos, rot, vel, ang, prev;
        std::vector<uint32_t> layer, mask;
        std::vector<uint8_t> ll, la, shape, smode, asleep;
        std::vector<int32_t> still;
        avbd_bodies view(int n)
        {
            avbd_bodies b{};
            b.count = n;
            b.size = size.data(); b.mass = mass.data(); b.moment = moment.data(); b.friction = friction.data();
            b.radius = radius.data(); b.gravity = gravity.data(); b.layer = layer.data(); b.mask = mask.data();
            b.lock_linear = ll.data(); b.lock_angular = la.data(); b.position = pos.data(); b.rotation = rot.data();
            b.velocity = vel.data(); b.angular_velocity = ang.data(); b.prev_velocity = prev.data();
            b.shape = shape.data();
            b.sleep_mode = smode.data(); b.sleeping = asleep.data(); b.still_frames = still.data();
            return b;
        }
    };

    void gather(const std::vector<Rigid *> &list, Arrays &a)
    {
        const size_t n = list.size();
        a.size.resize(3 * n); a.mass.resize(n); a.moment.resize(3 * n); a.friction.resize(n); a.radius.resize(n);
        a.gravity.resize(n); a.pos.resize(3 * n); a.rot.resize(4 * n); a.vel.resize(3 * n); a.ang.resize(3 * n);
        a.prev.resize(3 * n); a.layer.resize(n); a.mask.resize(n); a.ll.resize(n); a.la.resize(n);
        a.shape.resize(n); a.smode.resize(n); a.asleep.resize(n); a.still.resize(n);
        for (size_t i = 0; i < n; i++) {
            const Rigid *r = list[i];
            a.shape[i] = static_cast<uint8_t>(r->shape);
            a.smode[i] = static_cast<uint8_t>(r->sleep_mode);
            a.asleep[i] = r->sleeping ? 1 : 0;
            a.still[i] = r->stillFrames;
            for (int c = 0; c < 3; c++) {
                a.size[3 * i + c] = r->size[c];
                a.moment[3 * i + c] = r->moment[c];
                a.pos[3 * i + c] = r->positionLin[c];
                a.vel[3 * i + c] = r->velocityLin[c];
                a.ang[3 * i + c] = r->velocityAng[c];
                a.prev[3 * i + c] = r->prevVelocityLin[c];
            }
            a.rot[4 * i] = r->positionAng.x; a.rot[4 * i + 1] = r->positionAng.y;
            a.rot[4 * i + 2] = r->positionAng.z; a.rot[4 * i + 3] = r->positionAng.w;
            a.mass[i] = r->mass; a.friction[i] = r->friction; a.radius[i] = r->radius; a.gravity[i] = r->gravity;
            a.layer[i] = r->collisionLayer; a.mask[i] = r->collisionMask;
            a.ll[i] = r->axisLockLinear; a.la[i] = r->axisLockAngular;
        }
    }

    // Re-upload the lists, carrying the manifolds of surviving body pairs over.
    bool rebuild(const std::vector<Rigid *> &list, const std::vector<Force *> &jl)
    {
        const int n = static_cast<int>(list.size()), nj = static_cast<int>(jl.size());
        std::vector<int32_t> ma, mb, num, feat;
        std::vector<float> basis, fr, ra, rb, c0, pen, lam, beta;
        std::vector<uint8_t> stick;
        int keepCount = 0;
        if (world && !uploaded.empty()) {
            avbd_step_stats st{};
            avbd_world_stats(world, &st);
            const int m = st.manifolds > 0 ? st.manifolds : 0;
            ma.resize(m); mb.resize(m); num.resize(m); feat.resize(8 * m); basis.resize(9 * m); fr.resize(m);
            ra.resize(24 * m); rb.resize(24 * m); c0.resize(24 * m); pen.resize(24 * m); lam.resize(24 * m);
            beta.resize(8 * m); stick.resize(8 * m);
            avbd_manifolds mf{m, ma.data(), mb.data(), basis.data(), fr.data(), num.data(), feat.data(), ra.data(),
                              rb.data(), c0.data(), pen.data(), lam.data(), stick.data(), beta.data()};
            if (m > 0 && avbd_world_get_manifolds(world, &mf) != AVBD_OK) mf.count = 0;
            std::unordered_map<const Rigid *, int> now;
            for (int i = 0; i < n; i++) now[list[i]] = i;
            for (int k = 0; k < mf.count; k++) {
                auto ia = now.find(uploaded[ma[k]]), ib = now.find(uploaded[mb[k]]);
                if (ia == now.end() || ib == now.end()) continue;
                int a = ia->second, b = ib->second;
                if (a > b) continue;   // reordered pair: the broad phase recreates it
                ma[keepCount] = a; mb[keepCount] = b; num[keepCount] = num[k]; fr[keepCount] = fr[k];
                std::memmove(&basis[9 * keepCount], &basis[9 * k], 9 * sizeof(float));
                std::memmove(&feat[8 * keepCount], &feat[8 * k], 8 * sizeof(int32_t));
                std::memmove(&stick[8 * keepCount], &stick[8 * k], 8);
                std::memmove(&beta[8 * keepCount], &beta[8 * k], 8 * sizeof(float));
                for (auto *v : {&ra, &rb, &c0, &pen, &lam})
                    std::memmove(&(*v)[24 * keepCount], &(*v)[24 * k], 24 * sizeof(float));
                keepCount++;
            }
        }
        if (!world || n > capBodies || nj > capJoints) {
            if (world) avbd_world_destroy(world);
            world = nullptr;
            capBodies = n * 2 + 64;
            capJoints = nj * 2 + 16;
            const avbd_params p = params();
            lastStatus = avbd_world_create(&p, capBodies, capJoints, 0, &world);
            if (lastStatus != AVBD_OK) return false;
        }
        Arrays a;
        gather(list, a);
        avbd_bodies bv = a.view(n);
        std::vector<int32_t> ja(nj), jb(nj);
        std::vector<float> jra(3 * nj), jrb(3 * nj), ksl(nj), ksa(nj), jfr(nj), arm(nj), c0l(3 * nj), c0a(3 * nj),
                pl(3 * nj), pa(3 * nj), ll(3 * nj), la(3 * nj);
        std::vector<uint8_t> br(nj), kind(nj);
        std::vector<float> rest(nj, 0.0f), restQ(4 * nj, 0.0f), axLo(6 * nj, 0.0f), axHi(6 * nj, 0.0f),
                axK(6 * nj, 0.0f), axEq(6 * nj, 0.0f), axDamp(6 * nj, 0.0f), axLam(6 * nj, 0.0f), axPen(6 * nj, 0.0f),
                axInit(6 * nj, 0.0f);
        std::vector<uint8_t> axMode(6 * nj, 0);
        std::unordered_map<const Rigid *, int> idx;
        for (int i = 0; i < n; i++) idx[list[i]] = i;
        for (int k = 0; k < nj; k++) {
            const Force *f = jl[k];
            ja[k] = f->bodyA ? idx[f->bodyA] : -1;
            jb[k] = idx[f->bodyB];
            kind[k] = f->kind;
            jfr[k] = std::numeric_limits<float>::infinity();
            if (f->kind == AVBD_FORCE_JOINT) {
                const Joint *j = static_cast<const Joint *>(f);
                for (int c = 0; c < 3; c++) {
                    jra[3 * k + c] = j->rA[c]; jrb[3 * k + c] = j->rB[c];
                    c0l[3 * k + c] = j->C0Lin[c]; c0a[3 * k + c] = j->C0Ang[c];
                    pl[3 * k + c] = j->penaltyLin[c]; pa[3 * k + c] = j->penaltyAng[c];
                    ll[3 * k + c] = j->lambdaLin[c]; la[3 * k + c] = j->lambdaAng[c];
                }
                ksl[k] = j->stiffnessLin; ksa[k] = j->stiffnessAng; jfr[k] = j->fracture; arm[k] = j->torqueArm;
                br[k] = j->broken ? 1 : 0;
            } else if (f->kind == AVBD_FORCE_SPRING) {
                const Spring *sp = static_cast<const Spring *>(f);
                for (int c = 0; c < 3; c++) { jra[3 * k + c] = sp->rA[c]; jrb[3 * k + c] = sp->rB[c]; }
                ksl[k] = sp->stiffness;
                rest[k] = sp->rest;
            } else if (f->kind == AVBD_FORCE_GENERIC) {
                const GenericJoint *g = static_cast<const GenericJoint *>(f);
                for (int c = 0; c < 3; c++) { jra[3 * k + c] = g->rA[c]; jrb[3 * k + c] = g->rB[c]; }
                const float q[4] = {g->rest.x, g->rest.y, g->rest.z, g->rest.w};
                for (int c = 0; c < 4; c++) restQ[4 * k + c] = q[c];
                for (int a = 0; a < 6; a++) {
                    const JointAxis &x = a < 3 ? g->linear[a] : g->angular[a - 3];
                    const int i = 6 * k + a;
                    axMode[i] = static_cast<uint8_t>(x.mode);
                    axLo[i] = x.lower; axHi[i] = x.upper;
                    axK[i] = x.springStiffness; axEq[i] = x.springEquilibrium; axDamp[i] = x.springDamping;
                    axLam[i] = x.lambda; axPen[i] = x.penalty; axInit[i] = x.initial;
                }
            }
        }
        avbd_joints jv{nj, ja.data(), jb.data(), jra.data(), jrb.data(), ksl.data(), ksa.data(), jfr.data(),
                       arm.data(), c0l.data(), c0a.data(), pl.data(), pa.data(), ll.data(), la.data(), br.data(),
                       kind.data(), rest.data(), restQ.data(), axMode.data(), axLo.data(), axHi.data(), axK.data(),
                       axEq.data(), axDamp.data(), axLam.data(), axPen.data(), axInit.data()};
        lastStatus = avbd_world_upload(world, &bv, &jv);
        if (lastStatus != AVBD_OK) return false;
        if (keepCount > 0) {
            avbd_manifolds mf{keepCount, ma.data(), mb.data(), basis.data(), fr.data(), num.data(), feat.data(),
                              ra.data(), rb.data(), c0.data(), pen.data(), lam.data(), stick.data(), beta.data()};
            lastStatus = avbd_world_set_manifolds(world, &mf);
            if (lastStatus != AVBD_OK) return false;
        }
        uploaded = list;
        uploadedJoints = jl;
        mfValid = false;
        return true;
    }

    // Bring the device up to the host: parameters, then a re-upload after a topology change or
    // the bodies' dynamic fields otherwise.
    bool pushState(std::vector<Rigid *> &list, std::vector<Force *> &jl)
    {
        for (Rigid *r = bodies; r; r = r->next) list.push_back(r);
        for (Force *f = forces; f; f = f->next) {
            if (f->kind == AVBD_FORCE_JOINT) {
                Joint *j = static_cast<Joint *>(f);
                if (j->dirty) { topologyChanged = true; j->dirty = false; }
            }
            jl.push_back(f);
        }
        const avbd_params p = params();
        if (world) {
            lastStatus = avbd_world_set_params(world, &p);
            if (lastStatus != AVBD_OK) return false;
        }
        if (topologyChanged || !world) {
            if (!rebuild(list, jl)) return false;
            topologyChanged = false;
        } else {
            Arrays a;
            gather(list, a);
            avbd_bodies bv = a.view(static_cast<int>(list.size()));
            lastStatus = avbd_world_write_bodies(world, &bv);
            if (lastStatus != AVBD_OK) return false;
        }
        return true;
    }

    void advance(int substeps, int rounds)
    {
        std::vector<Rigid *> list;
        std::vector<Force *> jl;
        if (!pushState(list, jl)) return;
        // joints that were broken before this step are reaped by it (solver_solve.cpp:39-51)
        std::vector<Force *> reaped;   // only joints break
        for (Force *f : jl)
            if (f->kind == AVBD_FORCE_JOINT && static_cast<Joint *>(f)->broken) reaped.push_back(f);
        collisionDt = substeps > 1 ? dt * static_cast<float>(substeps) : dt;
        mfValid = false;
        lastStatus = substeps > 1 ? avbd_world_step_interleaved(world, substeps, rounds) : avbd_world_step(world);
        if (lastStatus != AVBD_OK) return;
        const avbd_step_stats st = stat();
        lastStepTimes.broadPhase = st.ms_broad;
        lastStepTimes.broadTree = st.ms_broad_tree;
        lastStepTimes.broadEmit = st.ms_broad_emit;
        lastStepTimes.broadIndex = static_cast<double>(st.ms_broad_tree) + st.ms_broad_emit;
        lastStepTimes.graph = st.ms_narrow;
        lastStepTimes.solve = st.ms_solve;
        lastStepTimes.finish = st.ms_finish;
        pull(list, jl, reaped);
    }

    void pull(const std::vector<Rigid *> &list, std::vector<Force *> jl, const std::vector<Force *> &reaped)
    {
        const int n = static_cast<int>(list.size());
        std::vector<float> pos(3 * n), rot(4 * n), vel(3 * n), ang(3 * n), prev(3 * n);
        avbd_bodies b{};
        b.count = n;
        b.position = pos.data(); b.rotation = rot.data(); b.velocity = vel.data();
        b.angular_velocity = ang.data(); b.prev_velocity = prev.data();
        std::vector<uint8_t> asleep(n);
        std::vector<int32_t> still(n);
        b.sleeping = asleep.data(); b.still_frames = still.data();
        const int nj = static_cast<int>(jl.size());
        std::vector<int32_t> ja(nj), jb(nj);
        std::vector<float> c0l(3 * nj), c0a(3 * nj), pl(3 * nj), pa(3 * nj), ll(3 * nj), la(3 * nj);
        std::vector<uint8_t> br(nj);
        std::vector<float> axLam(6 * nj), axPen(6 * nj), axInit(6 * nj);
        avbd_joints j{nj, ja.data(), jb.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c0l.data(),
                      c0a.data(), pl.data(), pa.data(), ll.data(), la.data(), br.data()};
        j.axis_lambda = axLam.data(); j.axis_penalty = axPen.data(); j.axis_initial = axInit.data();
        lastStatus = avbd_world_download(world, &b, nj ? &j : nullptr);
        if (lastStatus != AVBD_OK) return;
        for (int i = 0; i < n; i++) {
            Rigid *r = list[i];
            for (int c = 0; c < 3; c++) {
                r->positionLin[c] = pos[3 * i + c];
                r->velocityLin[c] = vel[3 * i + c];
                r->velocityAng[c] = ang[3 * i + c];
                r->prevVelocityLin[c] = prev[3 * i + c];
            }
            r->positionAng = quat{rot[4 * i], rot[4 * i + 1], rot[4 * i + 2], rot[4 * i + 3]};
            r->sleeping = asleep[i] != 0;
            r->stillFrames = still[i];
        }
        // the device compacted the reaped joints away; drop them from the host list too
        for (Force *dead : reaped) {
            for (size_t k = 0; k < jl.size(); k++)
                if (jl[k] == dead) { jl.erase(jl.begin() + static_cast<long>(k)); break; }
            static_cast<Joint *>(dead)->dirty = false;
            delete dead;
        }
        for (int k = 0; k < j.count && k < static_cast<int>(jl.size()); k++) {
            if (jl[k]->kind == AVBD_FORCE_GENERIC) {
                GenericJoint *g = static_cast<GenericJoint *>(jl[k]);
                for (int a = 0; a < 6; a++) {
                    JointAxis &x = a < 3 ? g->linear[a] : g->angular[a - 3];
                    x.lambda = axLam[6 * k + a]; x.penalty = axPen[6 * k + a]; x.initial = axInit[6 * k + a];
                }
                continue;
            }
            if (jl[k]->kind != AVBD_FORCE_JOINT) continue;   // springs and filters keep no dual state
            Joint *x = static_cast<Joint *>(jl[k]);
            for (int c = 0; c < 3; c++) {
                x->C0Lin[c] = c0l[3 * k + c]; x->C0Ang[c] = c0a[3 * k + c];
                x->penaltyLin[c] = pl[3 * k + c]; x->penaltyAng[c] = pa[3 * k + c];
                x->lambdaLin[c] = ll[3 * k + c]; x->lambdaAng[c] = la[3 * k + c];
            }
            x->broken = br[k] != 0;
        }
        if (!reaped.empty()) {
            // the device list changed with them; keep the upload mapping in step
            uploadedJoints = jl;
            topologyChanged = false;
        }
    }
};

inline Rigid::Rigid(Solver *p_solver, float3 p_size, float p_density, float p_friction, float3 p_position,
                    float3 p_velocity)
    : Rigid(p_solver, p_size, ShapeType::Box, p_density, p_friction, p_position, p_velocity)
{
}

inline Rigid::Rigid(Solver *p_solver, float3 p_size, ShapeType p_shape, float p_density, float p_friction,
                    float3 p_position, float3 p_velocity)
    : solver(p_solver), positionLin(p_position), velocityLin(p_velocity), prevVelocityLin(p_velocity),
      size(p_size), shape(p_shape), friction(p_friction), gravity(p_solver->gravity)
{
    next = p_solver->bodies;
    p_solver->bodies = this;
    const float s[3] = {p_size.x, p_size.y, p_size.z};
    float mom[3];
    avbd_shape_properties(static_cast<int>(p_shape), s, p_density, &mass, mom, &radius);   // rigid.cpp:34-70
    moment = float3{mom[0], mom[1], mom[2]};
    p_solver->topologyChanged = true;
}

inline Rigid::~Rigid()
{
    Rigid **p = &solver->bodies;
    while (*p != this) p = &(*p)->next;
    *p = next;
    solver->topologyChanged = true;
}

inline Force::Force(Solver *p_solver, Rigid *p_bodyA, Rigid *p_bodyB)
    : solver(p_solver), bodyA(p_bodyA), bodyB(p_bodyB)
{
    next = p_solver->forces;
    p_solver->forces = this;
}

inline Force::~Force()
{
    Force **p = &solver->forces;
    while (*p != this) p = &(*p)->next;
    *p = next;
    solver->topologyChanged = true;
}

inline Joint::Joint(Solver *p_solver, Rigid *p_bodyA, Rigid *p_bodyB, float3 p_rA, float3 p_rB, float p_stiffnessLin,
                    float p_stiffnessAng, float p_fracture)
    : Force(p_solver, p_bodyA, p_bodyB), rA(p_rA), rB(p_rB), stiffnessLin(p_stiffnessLin),
      stiffnessAng(p_stiffnessAng), fracture(p_fracture)
{
    // joint.cpp:28-42: |sizeA + sizeB|^2 with linalg's ((0 + x^2) + y^2) + z^2
    const float3 sa = p_bodyA ? p_bodyA->size : float3{0, 0, 0};
    const float sx = sa.x + p_bodyB->size.x, sy = sa.y + p_bodyB->size.y, sz = sa.z + p_bodyB->size.z;
    torqueArm = ((0.0f + sx * sx) + sy * sy) + sz * sz;
    p_solver->topologyChanged = true;
}

inline Spring::Spring(Solver *p_solver, Rigid *p_bodyA, Rigid *p_bodyB, float3 p_rA, float3 p_rB, float p_stiffness,
                      float p_rest)
    : Force(p_solver, p_bodyA, p_bodyB), rA(p_rA), rB(p_rB), rest(p_rest), stiffness(p_stiffness)
{
    kind = AVBD_FORCE_SPRING;
    if (rest < 0.0f) {   // spring.cpp:21-27
        const float3 qa = hostQrot(p_bodyA->positionAng, rA), qb = hostQrot(p_bodyB->positionAng, rB);
        const float dx = (p_bodyA->positionLin.x + qa.x) - (p_bodyB->positionLin.x + qb.x);
        const float dy = (p_bodyA->positionLin.y + qa.y) - (p_bodyB->positionLin.y + qb.y);
        const float dz = (p_bodyA->positionLin.z + qa.z) - (p_bodyB->positionLin.z + qb.z);
        rest = std::sqrt(((0.0f + dx * dx) + dy * dy) + dz * dz);
    }
    p_solver->topologyChanged = true;
}

inline GenericJoint::GenericJoint(Solver *p_solver, Rigid *p_bodyA, Rigid *p_bodyB, float3 p_rA, float3 p_rB)
    : Force(p_solver, p_bodyA, p_bodyB), rA(p_rA), rB(p_rB)
{
    kind = AVBD_FORCE_GENERIC;
    // generic_joint.cpp:24-30: rest = conj(frame) * rotation of B, frame = A's rotation or identity
    const quat f = p_bodyA ? p_bodyA->positionAng : quat{};
    rest = hostQmul(quat{-f.x, -f.y, -f.z, f.w}, p_bodyB->positionAng);
    p_solver->topologyChanged = true;
}

template <class Range>
int Manifold::collide(Rigid *bodyA, Rigid *bodyB, Range &&contacts, float3x3 &basis)
{
    return bodyA->solver->collidePair(bodyA, bodyB, contacts.data(), contacts.size(), basis);
}

inline IgnoreCollision::IgnoreCollision(Solver *p_solver, Rigid *p_bodyA, Rigid *p_bodyB)
    : Force(p_solver, p_bodyA, p_bodyB)
{
    kind = AVBD_FORCE_IGNORE_COLLISION;
    p_solver->topologyChanged = true;
}

} // namespace avbd_hip

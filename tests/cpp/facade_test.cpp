// Parity checks of the HIP step through the C++ facade (include/avbd_solver.hpp), written
// the way the reference's own core tests build their scenes (test/core_test.cpp:391-414,
// 612-631, 271-300): `new Rigid(&s, ...)`, `new Joint(&s, ...)`, `s.step()`.  The expected
// digests are the reference's recorded ones (tools/baseline/digests.txt).
#include <array>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "avbd_solver.hpp"

using namespace avbd_hip;

static int g_checks = 0, g_failures = 0;

static void report(bool ok, const char *name, const char *detail)
{
    g_checks++;
    if (!ok) g_failures++;
    std::printf("%s %-34s %s\n", ok ? "  ok " : "FAIL ", name, detail);
}

// FNV-1a over pos/quat/vel/omega in list order (test/core_test.cpp:133-149).
static uint64_t state_digest(const Solver &s)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ull; }
    };
    for (Rigid *b = s.bodies; b != nullptr; b = b->next) {
        mix(&b->positionLin, 12);
        mix(&b->positionAng, 16);
        mix(&b->velocityLin, 12);
        mix(&b->velocityAng, 12);
    }
    return h;
}

static uint64_t stack_digest()
{
    Solver s;
    new Rigid(&s, {100, 100, 1}, 0.0f, 0.5f, {0, 0, 0});
    for (int i = 0; i < 10; i++) new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {0, 0, i * 1.5f + 1.0f});
    for (int i = 0; i < 300; i++) s.step();
    return s.lastStatus == AVBD_OK ? state_digest(s) : 0;
}

static uint64_t interleaved_digest()
{
    Solver s;
    s.dt = 1.0f / 120.0f;
    new Rigid(&s, {100, 100, 1}, 0.0f, 0.5f, {0, 0, 0});
    for (int i = 0; i < 10; i++) new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {0, 0, i * 1.5f + 1.0f});
    for (int i = 0; i < 300; i++) s.stepInterleaved(2, 3);
    return s.lastStatus == AVBD_OK ? state_digest(s) : 0;
}

static void build_mixed(Solver &s)
{
    new Rigid(&s, {100, 100, 1}, 0.0f, 0.5f, {0, 0, 0});
    Rigid *prev = nullptr;
    for (int i = 0; i < 30; i++) {
        Rigid *link = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {0, 0, 20.0f - i * 1.5f});
        if (prev != nullptr) new Joint(&s, prev, link, {0, 0, -0.5f}, {0, 0, 0.5f}, 1000.0f, 250.0f);
        prev = link;
    }
    for (int i = 0; i < 40; i++) new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {5.0f + (float)i * 1.2f, 0, 3.0f + (float)(i % 5)});
}

static uint64_t mixed_digest()
{
    Solver s;
    build_mixed(s);
    for (int i = 0; i < 240; i++) s.step();
    return s.lastStatus == AVBD_OK ? state_digest(s) : 0;
}

// A body created mid-run keeps every existing manifold warm (the reference prepends it and
// leaves the manifold store alone).  Added far from a resting stack, it must leave the
// stack's trajectory bit-identical to a run without it; a lost warm start (penalty and
// lambda restarting from the floor) would make the stack sink and differ at once.
static uint64_t digest_except(const Solver &s, const Rigid *skip)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ull; }
    };
    for (Rigid *b = s.bodies; b != nullptr; b = b->next) {
        if (b == skip) continue;
        mix(&b->positionLin, 12);
        mix(&b->positionAng, 16);
        mix(&b->velocityLin, 12);
        mix(&b->velocityAng, 12);
    }
    return h;
}

static void test_topology_change_keeps_warm_start()
{
    Solver a, b;
    for (Solver *s : {&a, &b}) {
        new Rigid(s, {100, 100, 1}, 0.0f, 0.5f, {0, 0, 0});
        for (int i = 0; i < 5; i++) new Rigid(s, {1, 1, 1}, 1.0f, 0.5f, {0, 0, 1.0f + i * 1.01f});
    }
    for (int i = 0; i < 60; i++) { a.step(); b.step(); }
    Rigid *late = new Rigid(&a, {1, 1, 1}, 1.0f, 0.5f, {6, 0, 3});
    const int before = a.manifoldCount();
    for (int i = 0; i < 60; i++) { a.step(); b.step(); }
    const uint64_t da = digest_except(a, late), db = digest_except(b, nullptr);
    char d[160];
    std::snprintf(d, sizeof(d), "manifolds kept %d, stack digests %016" PRIx64 " / %016" PRIx64, before, da, db);
    report(before == 5 && da == db && a.lastStatus == AVBD_OK && late->positionLin.z < 1.1f,
           "topology change keeps warm start", d);
}

// A joint broken by the host is reaped by the next step, as in the reference.
static void test_break_now_reaps()
{
    Solver s;
    Rigid *anchor = new Rigid(&s, {1, 1, 1}, 0.0f, 0.5f, {0, 0, 10});
    Rigid *link = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {0, 0, 9});
    Joint *j = new Joint(&s, anchor, link, {0, 0, -0.5f}, {0, 0, 0.5f});
    for (int i = 0; i < 20; i++) s.step();
    const float zHeld = link->positionLin.z;
    j->breakNow();
    s.step();
    int joints = 0;
    for (Force *f = s.forces; f; f = f->next) joints++;
    for (int i = 0; i < 30; i++) s.step();
    char d[128];
    std::snprintf(d, sizeof(d), "joints after reap %d, z held %.3f -> %.3f", joints, zHeld, link->positionLin.z);
    report(joints == 0 && link->positionLin.z < zHeld - 0.5f && s.lastStatus == AVBD_OK, "breakNow reaps joint", d);
}

// Springs and collision filters through the facade (forces.hpp:89-113): a 1 kg box on a
// k = 100 spring of rest length 1 hangs m g / k = 0.1 below its rest length; a box filtered
// against the ground falls through it while an unfiltered one rests on it.
static void test_spring_and_filter()
{
    Solver s;
    Rigid *ground = new Rigid(&s, {40, 40, 1}, 0.0f, 0.5f, {0, 0, 0});
    Rigid *anchor = new Rigid(&s, {0.5f, 0.5f, 0.5f}, 0.0f, 0.5f, {10, 0, 10});
    Rigid *bob = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {10, 0, 8.5f});
    new Spring(&s, anchor, bob, {0, 0, 0}, {0, 0, 0}, 100.0f, 1.0f);
    Rigid *resting = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {-5, 0, 1.0f});
    Rigid *faller = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {5, 5, 1.0f});
    new IgnoreCollision(&s, ground, faller);
    for (int i = 0; i < 600; i++) s.step();
    char d[160];
    std::snprintf(d, sizeof(d), "spring z %.5f (closed form 8.9), resting z %.4f, filtered z %.2f", bob->positionLin.z,
                  resting->positionLin.z, faller->positionLin.z);
    report(std::fabs(bob->positionLin.z - 8.9f) < 1e-3f && std::fabs(resting->positionLin.z - 1.0f) < 5e-3f &&
               faller->positionLin.z < -10.0f && s.lastStatus == AVBD_OK,
           "spring equilibrium and collision filter", d);
}

// The reference's generic-joint rigs (test/core_test.cpp:742-790, 880-908) built with its own
// constructor and axis fields: a spinning hinge holds its locked axes and hub, a sprung strut
// hangs m g / k below its anchor, and the axes' dual state comes back to the host.
static void test_generic_joint()
{
    Solver s;
    Rigid *anchor = new Rigid(&s, {1, 1, 1}, 0.0f, 0.5f, {0, 0, 0});
    Rigid *wheel = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {0, 0, 0});
    GenericJoint *hinge = new GenericJoint(&s, anchor, wheel, {0, 0, 0}, {0, 0, 0});
    for (int i = 0; i < 3; i++) hinge->linear[i].mode = AxisMode::Locked;
    hinge->angular[0].mode = AxisMode::Locked;
    hinge->angular[2].mode = AxisMode::Locked;
    wheel->velocityAng = {0, 30.0f, 0};
    Rigid *anchor2 = new Rigid(&s, {1, 1, 1}, 0.0f, 0.5f, {10, 0, 0});
    Rigid *strut = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {10, 0, 0});
    GenericJoint *slide = new GenericJoint(&s, anchor2, strut, {0, 0, 0}, {0, 0, 0});
    for (int i = 0; i < 3; i++) slide->angular[i].mode = AxisMode::Locked;
    slide->linear[0].mode = slide->linear[1].mode = AxisMode::Locked;
    slide->linear[2] = JointAxis{AxisMode::Limited, -2.0f, 2.0f, 500.0f, 0.0f, 0.0f};
    float worstTilt = 0.0f, worstHub = 0.0f;
    for (int i = 0; i < 600; i++) {
        s.step();
        if (i < 300) {
            worstTilt = std::fmax(worstTilt, std::fmax(std::fabs(wheel->positionAng.x), std::fabs(wheel->positionAng.z)));
            const float3 p = wheel->positionLin;
            worstHub = std::fmax(worstHub, std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z));
        }
    }
    const float travel = -strut->positionLin.z, expected = 1.0f * 10.0f / 500.0f;
    char d[200];
    std::snprintf(d, sizeof(d), "hinge |q.x|,|q.z| <= %.5f, hub %.5f m; strut travel %.5f m (m g / k = %.3f)",
                  worstTilt, worstHub, travel, expected);
    report(worstTilt <= 0.005f && worstHub <= 0.01f && std::fabs(travel - expected) <= 0.2f * expected + 0.002f &&
               s.lastStatus == AVBD_OK,
           "generic joint hinge and sprung strut", d);
    // at rest the locked vertical axis carries the wheel's weight in its multiplier, the
    // strut's limited axis (inside its range) holds none, and `initial` is the strut's travel
    std::snprintf(d, sizeof(d), "hinge Z lambda %.4f (m g = 10), strut Z lambda %.3g, strut Z initial %.5f",
                  hinge->linear[2].lambda, slide->linear[2].lambda, slide->linear[2].initial);
    report(std::fabs(std::fabs(hinge->linear[2].lambda) - 10.0f) <= 0.2f && slide->linear[2].lambda == 0.0f &&
               std::fabs(slide->linear[2].initial - expected) <= 0.2f * expected + 0.002f,
           "generic joint dual state comes back", d);
}

// The reference's recorded round-shape runs (test/core_test.cpp:1061-1090 and 1409-1430;
// tools/baseline/core_test.txt:85-87 and 124): a spun sphere rolls 3.9950 m keeping 12.546
// rad/s, and a cylinder thrown at a slab at 15 m/s ends at z = 1.4665.
static void test_round_shapes()
{
    Solver s;
    new Rigid(&s, {200, 200, 1}, 0.0f, 0.9f, {0, 0, -0.5f});
    Rigid *wheel = new Rigid(&s, {0.25f, 0, 0}, ShapeType::Sphere, 1000.0f, 1.0f, {0, 0, 0.25f});
    for (int i = 0; i < 120; i++) s.step();
    wheel->velocityAng = {20.0f, 0, 0};
    const float startY = wheel->positionLin.y;
    for (int i = 0; i < 120; i++) s.step();
    const float travelled = wheel->positionLin.y - startY, spin = wheel->velocityAng.x;

    Solver t;
    t.speculativeMargin = 0.01f;
    new Rigid(&t, {100, 100, 1}, 0.0f, 0.5f, {0, 0, 0});
    Rigid *cyl = new Rigid(&t, {0.5f, 0.5f, 1.0f}, ShapeType::Cylinder, 1.0f, 0.5f, {0, 0, 1.008f}, {0, 0, -15.0f});
    cyl->gravity = 0.0f;
    float lowest = 1e9f;
    for (int i = 0; i < 60; i++) {
        t.step();
        lowest = std::fmin(lowest, cyl->positionLin.z);
    }
    char d[200];
    std::snprintf(d, sizeof(d), "rolled %.4f m at %.3f rad/s (recorded -3.9950, 12.546); cylinder z %.4f, lowest %.4f "
                  "(recorded 1.4665)", travelled, spin, cyl->positionLin.z, lowest);
    report(std::fabs(travelled + 3.9950f) <= 5.01e-5f && std::fabs(spin - 12.546f) <= 5.01e-4f &&
               std::fabs(cyl->positionLin.z - 1.4665f) <= 5.01e-5f && lowest - 0.5f >= 0.49f &&
               s.lastStatus == AVBD_OK && t.lastStatus == AVBD_OK,
           "sphere rolls and cylinder lands as recorded", d);
}

// Solver::pick and Solver::collectAABB through the facade on a small scene: a ray straight down
// finds the top box of a stack (hit 0.5 above its centre), a mask skips its layer, a body moved
// on the host is found where it was moved to, and the AABB query returns list order.
static void test_queries()
{
    Solver s;
    Rigid *ground = new Rigid(&s, {20, 20, 1}, 0.0f, 0.5f, {0, 0, -0.5f});
    Rigid *lower = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {0, 0, 0.5f});
    Rigid *upper = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {0, 0, 1.5f});
    Rigid *ball = new Rigid(&s, {0.5f, 0, 0}, ShapeType::Sphere, 1.0f, 0.5f, {3, 0, 0.5f});
    upper->collisionLayer = 2;
    s.step();
    float3 local{};
    Rigid *hitTop = s.pick({0, 0, 5}, {0, 0, -1}, local);
    const float topLocal = local.z;
    Rigid *hitMasked = s.pick({0, 0, 5}, {0, 0, -1}, local, 1u);
    ball->positionLin = float3{-4, 2, 0.5f};   // moved on the host: queries see it there
    Rigid *hitBall = s.pick({-4, 2, 5}, {0, 0, -1}, local);
    std::vector<Rigid *> found;
    s.collectAABB({-0.2f, -0.2f, -0.2f}, {0.2f, 0.2f, 1.2f}, 0xFFFFFFFFu, found);
    const bool listOrder = found.size() == 3 && found[0] == upper && found[1] == lower && found[2] == ground;
    char d[200];
    std::snprintf(d, sizeof(d), "top %s (local z %.4f), masked %s, moved ball %s, aabb %zu bodies in list order %d",
                  hitTop == upper ? "upper" : "?", topLocal, hitMasked == lower ? "lower" : "?",
                  hitBall == ball ? "ball" : "?", found.size(), listOrder ? 1 : 0);
    report(hitTop == upper && std::fabs(topLocal - 0.5f) < 1e-3f && hitMasked == lower && hitBall == ball && listOrder &&
               s.lastStatus == AVBD_OK,
           "pick and collectAABB", d);
}

// core_test.cpp:1645-1659: a two-box stack with sleep_mode SLEEP and sleepFrames 5 is asleep
// after 120 steps and stays finite; the sleepers keep zero velocity.
static void test_idle_sleep()
{
    Solver s;
    s.sleepFrames = 5;
    new Rigid(&s, {100, 100, 1}, 0.0f, 0.5f, {0, 0, -0.5f});
    Rigid *a = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {0, 0, 0.5f});
    Rigid *b = new Rigid(&s, {1, 1, 1}, 1.0f, 0.5f, {0, 0, 1.5f});
    a->sleep_mode = 1;
    b->sleep_mode = 1;
    for (int i = 0; i < 120; i++) s.step();
    char d[160];
    std::snprintf(d, sizeof(d), "a=%d b=%d still %d %d, b z %.4f vz %.1f", a->sleeping, b->sleeping, a->stillFrames,
                  b->stillFrames, b->positionLin.z, b->velocityLin.z);
    report(a->sleeping && b->sleeping && std::isfinite(b->positionLin.z) && b->velocityLin.z == 0.0f &&
               s.lastStatus == AVBD_OK,
           "idle bodies fall asleep", d);
}

// The Godot server's two per-tick loops over the solver's output, as it writes them
// (avbd_physics_server3d.cpp:1547-1585: the deepest stored contact's pair re-measured by
// Manifold::collide at the current pose; :1620-1652: each reporting body's contacts, in live
// order, up to its max_contacts_reported), run on a scene read from `scene` after `steps`
// steps.  Everything the loops read is written to `out` (floats as C99 hex) for the caller to
// compare with the oracle.  Scene format: n steps, then per body in creation order: shape,
// size xyz, mass, moment xyz, friction, radius, position xyz, rotation xyzw, velocity xyz,
// angular velocity xyz, max_contacts_reported.
namespace avbd = avbd_hip;   // what a host written against the reference switches by

static int server_loops(const char *scenePath, const char *outPath)
{
    FILE *in = std::fopen(scenePath, "r");
    if (!in) return 2;
    int n = 0, steps = 0;
    if (std::fscanf(in, "%d %d", &n, &steps) != 2) return 2;
    avbd::Solver solver;
    std::vector<avbd::Rigid *> created;
    std::vector<int> maxReported;
    for (int i = 0; i < n; i++) {
        int shape = 0, rep = 0;
        float v[22];
        if (std::fscanf(in, "%d", &shape) != 1) return 2;
        for (float &x : v)
            if (std::fscanf(in, "%f", &x) != 1) return 2;
        if (std::fscanf(in, "%d", &rep) != 1) return 2;
        auto *r = new avbd::Rigid(&solver, {v[0], v[1], v[2]}, static_cast<avbd::ShapeType>(shape), 1.0f, v[7],
                                  {v[9], v[10], v[11]}, {v[16], v[17], v[18]});
        r->mass = v[3];
        r->moment = avbd::float3{v[4], v[5], v[6]};
        r->radius = v[8];
        r->positionAng = avbd::quat{v[12], v[13], v[14], v[15]};
        r->velocityAng = avbd::float3{v[19], v[20], v[21]};
        created.push_back(r);
        maxReported.push_back(rep);
    }
    std::fclose(in);
    for (int k = 0; k < steps; k++) solver.step();
    if (solver.lastStatus != AVBD_OK) return 3;
    std::vector<avbd::Rigid *> list;
    for (avbd::Rigid *b = solver.bodies; b; b = b->next) list.push_back(b);
    auto indexOf = [&](const avbd::Rigid *r) {
        for (size_t i = 0; i < list.size(); i++)
            if (list[i] == r) return static_cast<int>(i);
        return -1;
    };
    FILE *out = std::fopen(outPath, "w");
    if (!out) return 2;
    auto f3 = [&](avbd::float3 v) { std::fprintf(out, " %a %a %a", v.x, v.y, v.z); };
    // every accessor, as the loops call them
    std::fprintf(out, "M %d\n", solver.manifoldCount());
    for (int m = 0; m < solver.manifoldCount(); m++) {
        const int count = solver.manifoldContactCount(m);
        std::fprintf(out, "%d %d %d", indexOf(solver.manifoldBodyA(m)), indexOf(solver.manifoldBodyB(m)), count);
        f3(solver.manifoldNormal(m));
        std::fprintf(out, "\n");
        for (int i = 0; i < count; i++) {
            const avbd::Manifold::Contact &c = solver.manifoldContact(m, i);
            std::fprintf(out, "%d", c.feature.key);
            f3(solver.manifoldContactError(m, i));
            f3(c.rA);
            f3(c.rB);
            std::fprintf(out, "\n");
        }
    }
    // avbd_physics_server3d.cpp:1547-1585
    avbd::Rigid *deepest_a = nullptr;
    avbd::Rigid *deepest_b = nullptr;
    double deepest_stored = 0.0;
    for (int m = 0; m < solver.manifoldCount(); m++) {
        const int contacts = solver.manifoldContactCount(m);
        for (int i = 0; i < contacts; i++) {
            const double depth = static_cast<double>(avbd::COLLISION_MARGIN - solver.manifoldContactError(m, i).x);
            if (depth > deepest_stored) {
                deepest_stored = depth;
                deepest_a = solver.manifoldBodyA(m);
                deepest_b = solver.manifoldBodyB(m);
            }
        }
    }
    double diag_max_penetration = 0.0;
    if (deepest_a != nullptr) {
        std::array<avbd::Manifold::Contact, 8> fresh{};
        avbd::float3x3 freshBasis{};
        const int freshCount = avbd::Manifold::collide(deepest_a, deepest_b, fresh, freshBasis);
        std::fprintf(out, "D %d %d %d", indexOf(deepest_a), indexOf(deepest_b), freshCount);
        f3(avbd::row(freshBasis, 0));
        f3(avbd::row(freshBasis, 1));
        f3(avbd::row(freshBasis, 2));
        std::fprintf(out, "\n");
        for (int i = 0; i < freshCount; i++) {
            const avbd::float3 xA = deepest_a->positionLin + avbd::qrot(deepest_a->positionAng, fresh[i].rA);
            const avbd::float3 xB = deepest_b->positionLin + avbd::qrot(deepest_b->positionAng, fresh[i].rB);
            const avbd::float3 c0 = avbd::mul(freshBasis, xA - xB) + avbd::float3{avbd::COLLISION_MARGIN, 0, 0};
            const double depth = static_cast<double>(avbd::COLLISION_MARGIN - c0.x);
            if (depth > diag_max_penetration) diag_max_penetration = depth;
            std::fprintf(out, "%d", fresh[i].feature.key);
            f3(fresh[i].rA);
            f3(fresh[i].rB);
            std::fprintf(out, "\n");
        }
    } else {
        std::fprintf(out, "D -1 -1 0\n");
    }
    std::fprintf(out, "P %a\n", diag_max_penetration);
    // avbd_physics_server3d.cpp:1620-1652, with frame_contacts a vector per body
    struct FrameContact { avbd::float3 position, normal; int other; };
    std::vector<std::vector<FrameContact>> frame(list.size());
    for (int m = 0; m < solver.manifoldCount(); m++) {
        const int count = solver.manifoldContactCount(m);
        if (count <= 0) continue;
        avbd::Rigid *body_a = solver.manifoldBodyA(m);
        avbd::Rigid *body_b = solver.manifoldBodyB(m);
        const int data[2] = {indexOf(body_a), indexOf(body_b)};
        const avbd::float3 normal = solver.manifoldNormal(m);
        for (int side = 0; side < 2; side++) {
            const int fill = data[side];
            const int maxc = maxReported[static_cast<size_t>(n - 1 - fill)];   // list index -> creation index
            if (fill < 0 || maxc <= 0) continue;
            const int other = data[side == 0 ? 1 : 0];
            const avbd::Rigid *self_rigid = side == 0 ? body_a : body_b;
            for (int c = 0; c < count && static_cast<int32_t>(frame[fill].size()) < maxc; c++) {
                const avbd::Manifold::Contact &contact = solver.manifoldContact(m, c);
                const avbd::float3 world = self_rigid->positionLin +
                                           avbd::qrot(self_rigid->positionAng, side == 0 ? contact.rA : contact.rB);
                frame[fill].push_back(FrameContact{world, normal * (side == 0 ? 1.0f : -1.0f), other});
            }
        }
    }
    for (size_t i = 0; i < frame.size(); i++) {
        std::fprintf(out, "R %zu %zu\n", i, frame[i].size());
        for (const FrameContact &fc : frame[i]) {
            f3(fc.position);
            f3(fc.normal);
            std::fprintf(out, " %d\n", fc.other);
        }
    }
    const avbd::Solver::PhaseTimes &t = solver.lastStepTimes;
    std::fprintf(out, "T %d %.6f %.6f %.6f %.6f %.6f %.6f %.6f\n", solver.threadCount(), t.broadPhase, t.broadIndex,
                 t.broadTree, t.broadEmit, t.graph, t.solve, t.finish);
    std::fclose(out);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 4 && std::string(argv[1]) == "--server-loops") return server_loops(argv[2], argv[3]);
    const uint64_t want[3] = {0xfbac4f08cb1c977cull, 0xb03adeb37742850aull, 0xfb3ca7a3162035c2ull};
    const uint64_t got[3] = {stack_digest(), interleaved_digest(), mixed_digest()};
    const char *names[3] = {"digest stack_300", "digest interleaved_300", "digest mixed_240"};
    for (int k = 0; k < 3; k++) {
        char d[96];
        std::snprintf(d, sizeof(d), "0x%016" PRIx64 " (reference 0x%016" PRIx64 ")", got[k], want[k]);
        report(got[k] == want[k], names[k], d);
    }
    test_topology_change_keeps_warm_start();
    test_break_now_reaps();
    test_spring_and_filter();
    test_generic_joint();
    test_round_shapes();
    test_queries();
    test_idle_sleep();
    std::printf("%d checks, %d failures\n", g_checks, g_failures);
    return g_failures == 0 ? 0 : 1;
}

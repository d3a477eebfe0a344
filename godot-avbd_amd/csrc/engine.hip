// Host side of the MI355X AVBD step: device memory, world layout, the per-step launch
// sequence, HIP-graph capture for batches, and the C ABI declared in include/avbd_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/avbd_hip.h"
#include "kernels.hpp"
#include "queries.hpp"

using namespace avbd_dev;

namespace {

thread_local std::string g_lastError;

avbd_status fail(avbd_status code, const std::string &msg)
{
    g_lastError = msg;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(AVBD_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

// A first-fit allocator of [0, capacity) ranges (world bodies, joints, manifold slots).
struct RangeAlloc {
    int capacity = 0;
    std::map<int, int> freeRanges;  // start -> length
    void init(int cap)
    {
        capacity = cap;
        freeRanges.clear();
        if (cap > 0) freeRanges[0] = cap;
    }
    int take(int n)
    {
        if (n == 0) return 0;
        for (auto it = freeRanges.begin(); it != freeRanges.end(); ++it) {
            if (it->second >= n) {
                const int s = it->first, len = it->second;
                freeRanges.erase(it);
                if (len > n) freeRanges[s + n] = len - n;
                return s;
            }
        }
        return -1;
    }
    bool fits(int n) const
    {
        if (n == 0) return true;
        for (const auto &r : freeRanges)
            if (r.second >= n) return true;
        return false;
    }
    void give(int s, int n)
    {
        if (n <= 0) return;
        auto it = freeRanges.emplace(s, n).first;
        auto nx = std::next(it);
        if (nx != freeRanges.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            freeRanges.erase(nx);
        }
        if (it != freeRanges.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                freeRanges.erase(it);
            }
        }
    }
};

// An int's bits as a float (record fields that hold ints).
float intBits(int v)
{
    float f;
    std::memcpy(&f, &v, sizeof(f));
    return f;
}

template <typename T> hipError_t dalloc(T *&p, size_t n)
{
    return hipMalloc(reinterpret_cast<void **>(&p), std::max<size_t>(n, 1) * sizeof(T));
}

void setParams(Params &P, const avbd_params &p, float collisionDt)
{
    P.dt = p.dt;
    P.dtSq = p.dt * p.dt;
    P.collisionDt = collisionDt;
    P.alpha = p.alpha;
    P.oneMinusAlpha = 1 - p.alpha;
    P.betaLin = p.beta_lin;
    P.betaAng = p.beta_ang;
    P.gamma = p.gamma;
    P.jacobiRelax = p.jacobi_relaxation;
    P.stickDriftHeal = p.stick_drift_heal;
    P.maxPen = p.max_penetration;
    P.specMargin = p.speculative_margin;
    P.projBeta = p.projection_beta;
    P.projSlop = p.projection_slop;
    P.projMaxCorr = p.projection_max_correction;
    P.penaltyMax = p.penalty_max;
    P.adaptiveFloor = p.adaptive_beta_error_floor;
    P.stiffnessDecay = p.stiffness_decay;
    P.adaptiveBeta = p.adaptive_beta;
    P.contactOnlyOverflow = p.contact_only_overflow;
    P.projIters = p.projection_iterations;
    P.sleepFrames = p.sleep_frames;
    P.sleepLin = p.sleep_linear_threshold;
    P.sleepAng = p.sleep_angular_threshold;
    P.grace = p.contact_grace_ticks > 0 ? p.contact_grace_ticks : 0;
    P.armCache = p.jacobian_rebuild_distance > 0.0f ? 1 : 0;
    P.armRebuild2 = p.jacobian_rebuild_distance * p.jacobian_rebuild_distance;
}

avbd_status checkParams(const avbd_params *p)
{
    if (!p) return fail(AVBD_ERR_INVALID, "params is null");
    if (!(p->dt > 0.0f)) return fail(AVBD_ERR_INVALID, "dt must be positive");
    if (p->iterations < 1) return fail(AVBD_ERR_INVALID, "iterations must be >= 1");
    return AVBD_OK;
}

struct WorldInfo {
    bool active = false;
    bool pending = false;    // uploaded by avbd_batch_add_world_async, active from the next step
    int bodyBase = 0, bodyCount = 0, jointBase = 0, jointCount = 0, slotBase = 0, slotCap = 0;
    bool hasJoints = false;
    bool hasRound = false;   // a sphere or cylinder among its bodies
    bool hasSleep = false;   // a body that can sleep (sleep mode) or sleeps
};

} // namespace

struct Engine {
    avbd_params prm{};
    int Wcap = 0, Ncap = 0, Jcap = 0, Scap = 0, Pcap = 0;
    State S{};
    float *rimDev = nullptr;   // S.rim
    hipStream_t stream = nullptr;
    int device = 0;             // the device current when the engine was created
    std::vector<WorldInfo> worlds;
    RangeAlloc bodyAlloc, jointAlloc, slotAlloc;
    // scratch
    uint32_t *tileSums = nullptr;
    int tileSumsCap = 0;
    uint32_t *radixHist = nullptr;
    int radixBlocks = 0;
    int *errHost = nullptr;  // pinned
    int *countsHost = nullptr;
    bool csrValid = false;
    int treeAge = 1 << 20;
    int refitBudget = 16;
    int movingWorlds = 0;   // busy or exact-walking worlds in the last checked step
    // Steps between tree rebuilds: the reference's 16, or 4 while some world moves fast (its
    // refit boxes bloat and walks lengthen; the topology never changes the pairs).
    int rebuildInterval() const { return movingWorlds > 0 ? 4 : refitBudget; }
    bool batchMode = false;
    int colourRounds = 64;   // Jones-Plassmann rounds before the finisher (measured best on the chains world)
    hipGraphExec_t graphs[2] = {nullptr, nullptr};  // [rebuild, refit]
    bool graphJoints = false;
    bool graphRound = false;
    bool graphSleep = false;
    bool graphsValid = false;
    hipEvent_t ev[8] = {};   // step phases: 0 start, 1 broad, 2 narrow, 3 solve, 4 finish; 5-7 inside broad
    avbd_step_stats stats{};
    // per-launch profiling (avbd_*_profile): event pairs around launches of each class
    bool profiling = false;
    std::vector<hipEvent_t> profEvents;
    std::vector<std::pair<int, int>> profMarks;  // (class, first event index)
    int profUsed = 0;
    double profMs[AVBD_K_CLASSES] = {};
    long profLaunches[AVBD_K_CLASSES] = {};
    int profCls = -1;

    void pbeg(int cls)
    {
        if (!profiling) return;
        if (profUsed + 2 > static_cast<int>(profEvents.size())) {
            for (int k = 0; k < 256; k++) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return;
                profEvents.push_back(e);
            }
        }
        (void)hipEventRecord(profEvents[profUsed], stream);
        profMarks.emplace_back(cls, profUsed);
        profUsed += 2;
    }
    void pend()
    {
        if (!profiling || profMarks.empty()) return;
        (void)hipEventRecord(profEvents[profMarks.back().second + 1], stream);
    }
    void profCollect()
    {
        (void)hipStreamSynchronize(stream);
        for (const auto &m : profMarks) {
            float t = 0.0f;
            if (hipEventElapsedTime(&t, profEvents[m.second], profEvents[m.second + 1]) == hipSuccess) {
                profMs[m.first] += t;
                profLaunches[m.first] += 1;
            }
        }
        profMarks.clear();
        profUsed = 0;
    }

    ~Engine() { release(); }

    void release()
    {
        for (auto &g : graphs)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
        dropPlainGraphs();
        freeAll();
        if (errHost) (void)hipHostFree(errHost);
        if (countsHost) (void)hipHostFree(countsHost);
        if (phaseTicksHost) (void)hipHostFree(phaseTicksHost);
        errHost = countsHost = nullptr;
        phaseTicksHost = nullptr;
        for (auto &e : ev)
            if (e) { (void)hipEventDestroy(e); e = nullptr; }
        for (auto &e : profEvents) (void)hipEventDestroy(e);
        profEvents.clear();
        if (stream) { (void)hipStreamDestroy(stream); stream = nullptr; }
        if (upStream) { (void)hipStreamSynchronize(upStream); (void)hipStreamDestroy(upStream); upStream = nullptr; }
        if (upEvent) { (void)hipEventDestroy(upEvent); upEvent = nullptr; }
        if (arena) { (void)hipHostFree(arena); arena = nullptr; }
        for (auto *p : arenaOld) (void)hipHostFree(p);
        arenaOld.clear();
    }

    std::vector<void *> allocs;
    bool sizing = false;        // create() only adds up the device bytes it would allocate
    size_t deviceBytes = 0;
    template <typename T> avbd_status A(T *&p, size_t n)
    {
        deviceBytes += std::max<size_t>(n, 1) * sizeof(T);
        if (sizing) return AVBD_OK;
        hipError_t e = dalloc(p, n);
        if (e != hipSuccess) return fail(AVBD_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(p);
        e = hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return fail(AVBD_ERR_DEVICE, std::string("hipMemset: ") + hipGetErrorString(e));
        return AVBD_OK;
    }
    // Constraint values of each step's last dual, kept while on (arrays allocated the first time).
    bool recordRes = false;
    avbd_status recordResiduals(bool on)
    {
        if (on && !S.resC) {
            avbd_status st;
            if ((st = A(S.resC, static_cast<size_t>(Scap) * MAXC)) != AVBD_OK) return st;
            if ((st = A(S.resJ, Jcap)) != AVBD_OK) return st;
        }
        if (on != recordRes) graphsValid = false;
        recordRes = on;
        return AVBD_OK;
    }
    // Arm-cache arrays, allocated the first time arm reuse is switched on (then kept).
    avbd_status ensureArmCache(const avbd_params &p)
    {
        if (S.armCap || !(p.jacobian_rebuild_distance > 0.0f)) return AVBD_OK;
        const size_t M = Scap;
        avbd_status st;
        if ((st = A(S.armCap, 2 * M)) != AVBD_OK) return st;
        if ((st = A(S.armA, M * MAXC)) != AVBD_OK) return st;
        return A(S.armB, M * MAXC);
    }
    void fill(int *p, int v, int n)
    {
        if (n > 0) k_fill_i32<<<nblk(n), TPB, 0, stream>>>(p, v, n);
    }
    void freeAll()
    {
        for (void *p : allocs) (void)hipFree(p);
        allocs.clear();
        qCap = 0;
    }

    // ---- space queries (solver_query.cpp, query/query_shape.cpp, the server's intersect_ray) ----
    int qCap = 0;   // rays the query buffers hold
    float4 *qRays = nullptr, *qA = nullptr, *qB = nullptr, *qBounds = nullptr;
    int *qBody = nullptr;
    uint8_t *qFlag = nullptr;

    avbd_status ensureQuery(int n)
    {
        if (n <= qCap) return AVBD_OK;
        const int cap = std::max(n, std::max(1024, 2 * qCap));
        const int bodies = static_cast<int>(Ncap);
        auto grow = [&](auto *&p, size_t count) -> avbd_status {
            if (p) {
                allocs.erase(std::remove(allocs.begin(), allocs.end(), static_cast<void *>(p)), allocs.end());
                (void)hipFree(p);
                p = nullptr;
            }
            hipError_t e = dalloc(p, count);
            if (e != hipSuccess) return fail(AVBD_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
            allocs.push_back(p);
            return AVBD_OK;
        };
        avbd_status st;
        if ((st = grow(qRays, 2 * static_cast<size_t>(cap))) != AVBD_OK) return st;
        if ((st = grow(qA, cap)) != AVBD_OK) return st;
        if ((st = grow(qB, cap)) != AVBD_OK) return st;
        if ((st = grow(qBody, cap)) != AVBD_OK) return st;
        if ((st = grow(qFlag, std::max(cap, bodies))) != AVBD_OK) return st;
        if ((st = grow(qBounds, 2 * static_cast<size_t>(bodies))) != AVBD_OK) return st;
        qCap = cap;
        return AVBD_OK;
    }

    // Query tree: the last step's LBVH topology refit over the current tight bounds in qBounds
    // (k_query_bounds of world w first).  Valid while the bodies of every world are those the
    // tree was built over: from the first step after an upload, a removal or a world change on.
    // Without one (t.box null) the query kernels scan the world's list.
    float4 *qtC = nullptr, *qtH = nullptr, *qtBox = nullptr;
    int *qtFlag = nullptr;
    size_t qtCap = 0;
    bool queryTreeOn = true;

    avbd_status queryTree(int w, QTree &t)
    {
        const WorldInfo &wi = worlds[w];
        t = QTree{nullptr, Ncap, w};
        if (!queryTreeOn || treeAge > refitBudget || wi.bodyCount < 2) return AVBD_OK;
        if (static_cast<size_t>(Ncap) > qtCap) {
            avbd_status st;
            if ((st = regrow(qtC, Ncap)) != AVBD_OK || (st = regrow(qtH, Ncap)) != AVBD_OK ||
                (st = regrow(qtBox, 4 * static_cast<size_t>(Ncap))) != AVBD_OK || (st = regrow(qtFlag, Ncap)) != AVBD_OK)
                return st;
            qtCap = Ncap;
        }
        k_qtree_leaves<<<nblk(Ncap), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, qBounds, qtC, qtH);
        State Q = S;
        Q.aabbC = qtC;
        Q.aabbH = qtH;
        Q.nodeBox = qtBox;
        Q.nodeFlag = qtFlag;
        k_refit_clear<<<nblk(Ncap), TPB, 0, stream>>>(Q, Ncap);
        k_refit_tasks<<<refitGrid(), REFIT_TPB, 0, stream>>>(Q, Ncap, 0);
        HIPCHK(hipGetLastError());
        t.box = qtBox;
        return AVBD_OK;
    }

    avbd_status raycast(int w, int n, const float *from, const float *to, const uint32_t *masks, bool fromInside,
                        int32_t *body, float *t, float *point, float *normal, uint8_t *inside)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (n < 0 || (n > 0 && (!from || !to || !body))) return fail(AVBD_ERR_INVALID, "bad ray arrays");
        if (n == 0) return AVBD_OK;
        avbd_status st = ensureQuery(n);
        if (st != AVBD_OK) return st;
        std::vector<float4> rays(2 * static_cast<size_t>(n));
        for (int r = 0; r < n; r++) {
            const uint32_t m = masks ? masks[r] : 0xFFFFFFFFu;
            float mf;
            std::memcpy(&mf, &m, sizeof(mf));
            rays[2 * r] = make_float4(from[3 * r], from[3 * r + 1], from[3 * r + 2], mf);
            rays[2 * r + 1] = make_float4(to[3 * r], to[3 * r + 1], to[3 * r + 2], 0.0f);
        }
        HIPCHK(hipMemcpyAsync(qRays, rays.data(), sizeof(float4) * rays.size(), hipMemcpyHostToDevice, stream));
        if (wi.bodyCount > 0) k_query_bounds<<<nblk(wi.bodyCount), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, qBounds);
        QTree tree;
        if ((st = queryTree(w, tree)) != AVBD_OK) return st;
        k_raycast<<<nblk(n), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, n, qRays, qBounds, fromInside ? 1 : 0, qBody,
                                               qA, qB, tree);
        HIPCHK(hipGetLastError());
        std::vector<float4> a(n), b2(n);
        HIPCHK(hipMemcpyAsync(body, qBody, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(a.data(), qA, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(b2.data(), qB, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int r = 0; r < n; r++) {
            if (t) t[r] = a[r].w;
            if (point) { point[3 * r] = a[r].x; point[3 * r + 1] = a[r].y; point[3 * r + 2] = a[r].z; }
            if (normal) { normal[3 * r] = b2[r].x; normal[3 * r + 1] = b2[r].y; normal[3 * r + 2] = b2[r].z; }
            if (inside) inside[r] = b2[r].w != 0.0f ? 1 : 0;
        }
        return AVBD_OK;
    }

    // Average device time of the bounds and ray kernels over `reps` launches of n segments
    // (every mask bit set, no hit from inside), HIP events on the library's stream.
    avbd_status profileRaycast(int w, int n, const float *from, const float *to, int reps, avbd_query_profile *out)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (n <= 0 || reps <= 0 || !from || !to || !out) return fail(AVBD_ERR_INVALID, "bad arguments");
        avbd_status st = ensureQuery(n);
        if (st != AVBD_OK) return st;
        std::vector<float4> rays(2 * static_cast<size_t>(n));
        const uint32_t all = 0xFFFFFFFFu;
        float mf;
        std::memcpy(&mf, &all, sizeof(mf));
        for (int r = 0; r < n; r++) {
            rays[2 * r] = make_float4(from[3 * r], from[3 * r + 1], from[3 * r + 2], mf);
            rays[2 * r + 1] = make_float4(to[3 * r], to[3 * r + 1], to[3 * r + 2], 0.0f);
        }
        HIPCHK(hipMemcpyAsync(qRays, rays.data(), sizeof(float4) * rays.size(), hipMemcpyHostToDevice, stream));
        hipEvent_t e[3];
        for (auto &x : e) HIPCHK(hipEventCreate(&x));
        float bounds = 0.0f, kern = 0.0f;
        for (int k = 0; k < reps; k++) {
            HIPCHK(hipEventRecord(e[0], stream));
            k_query_bounds<<<nblk(std::max(wi.bodyCount, 1)), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, qBounds);
            QTree tree;
            if ((st = queryTree(w, tree)) != AVBD_OK) return st;
            HIPCHK(hipEventRecord(e[1], stream));
            k_raycast<<<nblk(n), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, n, qRays, qBounds, 0, qBody, qA, qB, tree);
            HIPCHK(hipEventRecord(e[2], stream));
            HIPCHK(hipEventSynchronize(e[2]));
            float a = 0.0f, b = 0.0f;
            HIPCHK(hipEventElapsedTime(&a, e[0], e[1]));
            HIPCHK(hipEventElapsedTime(&b, e[1], e[2]));
            bounds += a;
            kern += b;
        }
        for (auto &x : e) (void)hipEventDestroy(x);
        out->bounds_ms = bounds / reps;
        out->kernel_ms = kern / reps;
        return AVBD_OK;
    }

    int *qOut = nullptr;
    size_t qOutCap = 0;

    avbd_status intersectPoint(int w, int n, const float *points, const uint32_t *masks, int maxRes, int32_t *count,
                               int32_t *out)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (n < 0 || maxRes < 1 || (n > 0 && (!points || !count || !out))) return fail(AVBD_ERR_INVALID, "bad arguments");
        if (n == 0) return AVBD_OK;
        avbd_status st = ensureQuery(n);
        if (st != AVBD_OK) return st;
        const size_t outN = static_cast<size_t>(n) * maxRes;
        if (outN > qOutCap) {
            qOutCap = std::max(outN, 2 * qOutCap);
            if ((st = regrow(qOut, qOutCap)) != AVBD_OK) return st;
        }
        std::vector<float4> pts(n);
        for (int r = 0; r < n; r++) {
            const uint32_t m = masks ? masks[r] : 0xFFFFFFFFu;
            float mf;
            std::memcpy(&mf, &m, sizeof(mf));
            pts[r] = make_float4(points[3 * r], points[3 * r + 1], points[3 * r + 2], mf);
        }
        HIPCHK(hipMemcpyAsync(qRays, pts.data(), sizeof(float4) * n, hipMemcpyHostToDevice, stream));
        if (wi.bodyCount > 0) k_query_bounds<<<nblk(wi.bodyCount), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, qBounds);
        QTree tree;
        if ((st = queryTree(w, tree)) != AVBD_OK) return st;
        k_intersect_point<<<nblk(n), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, n, qRays, qBounds, maxRes, qBody, qOut,
                                                       tree);
        HIPCHK(hipGetLastError());
        std::vector<int> o(outN);
        HIPCHK(hipMemcpyAsync(count, qBody, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(o.data(), qOut, sizeof(int) * outN, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int r = 0; r < n; r++)
            for (int c = 0; c < maxRes; c++) out[static_cast<size_t>(r) * maxRes + c] = c < count[r] ? o[static_cast<size_t>(r) * maxRes + c] : -1;
        return AVBD_OK;
    }

    float4 *qPts = nullptr;   // collide_shape point pairs
    size_t qPtsCap = 0;
    float4 *qShapes = nullptr, *qCRA = nullptr, *qCRB = nullptr, *qCMid = nullptr, *qSpA = nullptr, *qSpB = nullptr;
    size_t qShapeCap = 0;

    // Query shapes to the device (layout of k_intersect_shape), bounds of world w's bodies too.
    avbd_status uploadShapes(int w, int n, const int32_t *type, const float *center, const float *rot,
                             const float *size, const float *margin, const uint32_t *masks, QTree &tree)
    {
        for (int r = 0; r < n; r++)
            if (type[r] < AVBD_SHAPE_BOX || type[r] > AVBD_SHAPE_CYLINDER) return fail(AVBD_ERR_INVALID, "unknown shape");
        avbd_status st = ensureQuery(n);
        if (st != AVBD_OK || (st = ensureShapeQuery(n)) != AVBD_OK) return st;
        std::vector<float4> in(4 * static_cast<size_t>(n));
        for (int r = 0; r < n; r++) {
            float tf, mf;
            const int32_t t = type[r];
            const uint32_t m = masks ? masks[r] : 0xFFFFFFFFu;
            std::memcpy(&tf, &t, sizeof(tf));
            std::memcpy(&mf, &m, sizeof(mf));
            in[4 * r] = make_float4(center[3 * r], center[3 * r + 1], center[3 * r + 2], tf);
            in[4 * r + 1] = make_float4(rot[4 * r], rot[4 * r + 1], rot[4 * r + 2], rot[4 * r + 3]);
            in[4 * r + 2] = make_float4(size[3 * r], size[3 * r + 1], size[3 * r + 2], margin[r]);
            in[4 * r + 3] = make_float4(mf, 0.0f, 0.0f, 0.0f);
        }
        HIPCHK(hipMemcpyAsync(qShapes, in.data(), sizeof(float4) * in.size(), hipMemcpyHostToDevice, stream));
        const WorldInfo &wi = worlds[w];
        if (wi.bodyCount > 0) k_query_bounds<<<nblk(wi.bodyCount), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, qBounds);
        if ((st = queryTree(w, tree)) != AVBD_OK) return st;
        HIPCHK(hipStreamSynchronize(stream));   // `in` is host memory the copy reads
        return AVBD_OK;
    }

    avbd_status collideShape(int w, int n, const int32_t *type, const float *center, const float *rot,
                             const float *size, const float *margin, const uint32_t *masks, int maxPairs,
                             int32_t *count, uint8_t *any, float *points)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (n < 0 || maxPairs < 0 || (n > 0 && (!type || !center || !rot || !size || !margin || !count || !any ||
                                                (maxPairs > 0 && !points))))
            return fail(AVBD_ERR_INVALID, "bad arguments");
        if (n == 0) return AVBD_OK;
        QTree tree;
        avbd_status st = uploadShapes(w, n, type, center, rot, size, margin, masks, tree);
        if (st != AVBD_OK) return st;
        const size_t outN = 2 * static_cast<size_t>(n) * static_cast<size_t>(std::max(maxPairs, 1));
        if (outN > qPtsCap) {
            qPtsCap = std::max(outN, 2 * qPtsCap);
            if ((st = regrow(qPts, qPtsCap)) != AVBD_OK) return st;
        }
        k_collide_shape<<<nblk(n, NARROW_TPB), NARROW_TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, n, qShapes, qBounds,
                                                                        maxPairs, qBody, qFlag, qPts, qCRA, qCRB, qCMid,
                                                                        qSpA, qSpB, tree);
        HIPCHK(hipGetLastError());
        std::vector<float4> pts(static_cast<size_t>(n) * 2 * maxPairs);
        HIPCHK(hipMemcpyAsync(count, qBody, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(any, qFlag, n, hipMemcpyDeviceToHost, stream));
        if (maxPairs > 0)
            HIPCHK(hipMemcpyAsync(pts.data(), qPts, sizeof(float4) * pts.size(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (size_t k = 0; k < pts.size(); k++) {
            const size_t r = k / (2 * static_cast<size_t>(maxPairs)), p = (k % (2 * static_cast<size_t>(maxPairs))) / 2;
            if (static_cast<int>(p) >= count[r]) continue;
            points[3 * k] = pts[k].x; points[3 * k + 1] = pts[k].y; points[3 * k + 2] = pts[k].z;
        }
        return AVBD_OK;
    }

    avbd_status restInfo(int w, int n, const int32_t *type, const float *center, const float *rot, const float *size,
                         const float *margin, const uint32_t *masks, int32_t *body, float *point, float *normal,
                         float *depth)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (n < 0 || (n > 0 && (!type || !center || !rot || !size || !margin || !body)))
            return fail(AVBD_ERR_INVALID, "bad arguments");
        if (n == 0) return AVBD_OK;
        QTree tree;
        avbd_status st = uploadShapes(w, n, type, center, rot, size, margin, masks, tree);
        if (st != AVBD_OK) return st;
        k_rest_info<<<nblk(n, NARROW_TPB), NARROW_TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, n, qShapes, qBounds, qBody,
                                                                    qA, qB, qCRA, qCRB, qCMid, qSpA, qSpB, tree);
        HIPCHK(hipGetLastError());
        std::vector<float4> a(n), b(n);
        HIPCHK(hipMemcpyAsync(body, qBody, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(a.data(), qA, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(b.data(), qB, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int r = 0; r < n; r++) {
            if (point) { point[3 * r] = a[r].x; point[3 * r + 1] = a[r].y; point[3 * r + 2] = a[r].z; }
            if (normal) { normal[3 * r] = b[r].x; normal[3 * r + 1] = b[r].y; normal[3 * r + 2] = b[r].z; }
            if (depth) depth[r] = a[r].w;
        }
        return AVBD_OK;
    }

    avbd_status closestPointQuery(int w, int n, const int32_t *idx, const float *points, float *out)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (n < 0 || (n > 0 && (!idx || !points || !out))) return fail(AVBD_ERR_INVALID, "bad arguments");
        for (int r = 0; r < n; r++)
            if (idx[r] < 0 || idx[r] >= wi.bodyCount) return fail(AVBD_ERR_INVALID, "body index out of range");
        if (n == 0) return AVBD_OK;
        avbd_status st = ensureQuery(n);
        if (st != AVBD_OK) return st;
        std::vector<float4> in(n);
        for (int r = 0; r < n; r++) in[r] = make_float4(points[3 * r], points[3 * r + 1], points[3 * r + 2], 0.0f);
        HIPCHK(hipMemcpyAsync(qBody, idx, sizeof(int) * n, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(qA, in.data(), sizeof(float4) * n, hipMemcpyHostToDevice, stream));
        k_closest_point<<<nblk(n), TPB, 0, stream>>>(S, wi.bodyBase, n, qBody, qA, qB);
        HIPCHK(hipGetLastError());
        std::vector<float4> res(n);
        HIPCHK(hipMemcpyAsync(res.data(), qB, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int r = 0; r < n; r++) { out[3 * r] = res[r].x; out[3 * r + 1] = res[r].y; out[3 * r + 2] = res[r].z; }
        return AVBD_OK;
    }

    avbd_status intersectShape(int w, int n, const int32_t *type, const float *center, const float *rot,
                               const float *size, const float *margin, const uint32_t *masks, int maxRes,
                               int32_t *count, int32_t *out)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (n < 0 || maxRes < 1 || (n > 0 && (!type || !center || !rot || !size || !margin || !count || !out)))
            return fail(AVBD_ERR_INVALID, "bad arguments");
        for (int r = 0; r < n; r++)
            if (type[r] < AVBD_SHAPE_BOX || type[r] > AVBD_SHAPE_CYLINDER) return fail(AVBD_ERR_INVALID, "unknown shape");
        if (n == 0) return AVBD_OK;
        avbd_status st = ensureQuery(n);
        if (st != AVBD_OK) return st;
        const size_t outN = static_cast<size_t>(n) * maxRes;
        if (outN > qOutCap) {
            qOutCap = std::max(outN, 2 * qOutCap);
            if ((st = regrow(qOut, qOutCap)) != AVBD_OK) return st;
        }
        if ((st = ensureShapeQuery(n)) != AVBD_OK) return st;
        std::vector<float4> in(4 * static_cast<size_t>(n));
        for (int r = 0; r < n; r++) {
            float tf, mf;
            const int32_t t = type[r];
            const uint32_t m = masks ? masks[r] : 0xFFFFFFFFu;
            std::memcpy(&tf, &t, sizeof(tf));
            std::memcpy(&mf, &m, sizeof(mf));
            in[4 * r] = make_float4(center[3 * r], center[3 * r + 1], center[3 * r + 2], tf);
            in[4 * r + 1] = make_float4(rot[4 * r], rot[4 * r + 1], rot[4 * r + 2], rot[4 * r + 3]);
            in[4 * r + 2] = make_float4(size[3 * r], size[3 * r + 1], size[3 * r + 2], margin[r]);
            in[4 * r + 3] = make_float4(mf, 0.0f, 0.0f, 0.0f);
        }
        HIPCHK(hipMemcpyAsync(qShapes, in.data(), sizeof(float4) * in.size(), hipMemcpyHostToDevice, stream));
        if (wi.bodyCount > 0) k_query_bounds<<<nblk(wi.bodyCount), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, qBounds);
        QTree tree;
        if ((st = queryTree(w, tree)) != AVBD_OK) return st;
        k_intersect_shape<<<nblk(n, NARROW_TPB), NARROW_TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, n, qShapes, qBounds,
                                                                          maxRes, qBody, qOut, qCRA, qCRB, qCMid, qSpA, qSpB,
                                                                          tree);
        HIPCHK(hipGetLastError());
        std::vector<int> o(outN);
        HIPCHK(hipMemcpyAsync(count, qBody, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(o.data(), qOut, sizeof(int) * outN, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int r = 0; r < n; r++)
            for (int c = 0; c < maxRes; c++) out[static_cast<size_t>(r) * maxRes + c] = c < count[r] ? o[static_cast<size_t>(r) * maxRes + c] : -1;
        return AVBD_OK;
    }

    avbd_status ensureShapeQuery(int n)
    {
        if (static_cast<size_t>(n) <= qShapeCap) return AVBD_OK;
        qShapeCap = std::max<size_t>(n, 2 * qShapeCap);
        const size_t rows = (qShapeCap + CSLICE - 1) / CSLICE * CSLICE * MAXC;
        avbd_status st;
        if ((st = regrow(qShapes, 4 * qShapeCap)) != AVBD_OK || (st = regrow(qCRA, rows)) != AVBD_OK ||
            (st = regrow(qCRB, rows)) != AVBD_OK || (st = regrow(qCMid, rows)) != AVBD_OK ||
            (st = regrow(qSpA, rows)) != AVBD_OK || (st = regrow(qSpB, rows)) != AVBD_OK)
            return st;
        return AVBD_OK;
    }

    avbd_status castMotion(int w, int n, const int32_t *type, const float *center, const float *rot, const float *size,
                           const float *motion, const float *margin, const uint32_t *masks, int32_t *body, float *frac,
                           float *point, float *normal)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (n < 0 || (n > 0 && (!type || !center || !rot || !size || !motion || !margin || !body || !frac)))
            return fail(AVBD_ERR_INVALID, "bad arguments");
        for (int r = 0; r < n; r++)
            if (type[r] < AVBD_SHAPE_BOX || type[r] > AVBD_SHAPE_CYLINDER) return fail(AVBD_ERR_INVALID, "unknown shape");
        if (n == 0) return AVBD_OK;
        avbd_status st = ensureQuery(n);
        if (st != AVBD_OK || (st = ensureShapeQuery(n)) != AVBD_OK) return st;
        std::vector<float4> in(4 * static_cast<size_t>(n));
        for (int r = 0; r < n; r++) {
            float tf, mf;
            const int32_t t = type[r];
            const uint32_t m = masks ? masks[r] : 0xFFFFFFFFu;
            std::memcpy(&tf, &t, sizeof(tf));
            std::memcpy(&mf, &m, sizeof(mf));
            in[4 * r] = make_float4(center[3 * r], center[3 * r + 1], center[3 * r + 2], tf);
            in[4 * r + 1] = make_float4(rot[4 * r], rot[4 * r + 1], rot[4 * r + 2], rot[4 * r + 3]);
            in[4 * r + 2] = make_float4(size[3 * r], size[3 * r + 1], size[3 * r + 2], margin[r]);
            in[4 * r + 3] = make_float4(motion[3 * r], motion[3 * r + 1], motion[3 * r + 2], mf);
        }
        HIPCHK(hipMemcpyAsync(qShapes, in.data(), sizeof(float4) * in.size(), hipMemcpyHostToDevice, stream));
        if (wi.bodyCount > 0) k_query_bounds<<<nblk(wi.bodyCount), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, qBounds);
        QTree tree;
        if ((st = queryTree(w, tree)) != AVBD_OK) return st;
        k_cast_motion<<<nblk(n, NARROW_TPB), NARROW_TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, n, qShapes, qBounds, qBody,
                                                                      qA, qB, qCRA, qCRB, qCMid, qSpA, qSpB, tree);
        HIPCHK(hipGetLastError());
        std::vector<float4> a(n), b(n);
        HIPCHK(hipMemcpyAsync(body, qBody, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(a.data(), qA, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(b.data(), qB, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int r = 0; r < n; r++) {
            frac[2 * r] = a[r].w;
            frac[2 * r + 1] = b[r].w;
            if (point) { point[3 * r] = a[r].x; point[3 * r + 1] = a[r].y; point[3 * r + 2] = a[r].z; }
            if (normal) { normal[3 * r] = b[r].x; normal[3 * r + 1] = b[r].y; normal[3 * r + 2] = b[r].z; }
        }
        return AVBD_OK;
    }

    // avbd::Manifold::collide for two bodies of world 0 at their current pose (one narrow-phase
    // work item; staging rows of query 0).
    avbd_status collidePair(int a, int b, float hz, int32_t *count, int32_t *features, float *rA, float *rB,
                            float *basis)
    {
        const WorldInfo &wi = worlds[0];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (a < 0 || b < 0 || a >= wi.bodyCount || b >= wi.bodyCount || a == b)
            return fail(AVBD_ERR_INVALID, "bad body index");
        avbd_status st = ensureQuery(1);
        if (st != AVBD_OK || (st = ensureShapeQuery(1)) != AVBD_OK) return st;
        k_collide_pair<<<1, NARROW_TPB, 0, stream>>>(S, wi.bodyBase + a, wi.bodyBase + b, hz, qA, qCRA, qCRB, qCMid,
                                                     qSpA, qSpB);
        HIPCHK(hipGetLastError());
        float4 out[4 + 2 * 8];
        HIPCHK(hipMemcpyAsync(out, qA, sizeof(out), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        const int n = static_cast<int>(out[0].x);
        *count = n;
        if (basis && out[0].y != 0.0f)
            for (int r = 0; r < 3; r++) {
                basis[3 * r] = out[1 + r].x; basis[3 * r + 1] = out[1 + r].y; basis[3 * r + 2] = out[1 + r].z;
            }
        for (int k = 0; k < n; k++) {
            const float4 x = out[4 + 2 * k], y = out[5 + 2 * k];
            if (features) std::memcpy(&features[k], &x.w, sizeof(int32_t));
            if (rA) { rA[3 * k] = x.x; rA[3 * k + 1] = x.y; rA[3 * k + 2] = x.z; }
            if (rB) { rB[3 * k] = y.x; rB[3 * k + 1] = y.y; rB[3 * k + 2] = y.z; }
        }
        return AVBD_OK;
    }

    avbd_status pick(int w, int n, const float *origin, const float *dir, const uint32_t *masks, int32_t *body,
                     float *local)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (n < 0 || (n > 0 && (!origin || !dir || !body))) return fail(AVBD_ERR_INVALID, "bad ray arrays");
        if (n == 0) return AVBD_OK;
        avbd_status st = ensureQuery(n);
        if (st != AVBD_OK) return st;
        std::vector<float4> rays(2 * static_cast<size_t>(n));
        for (int r = 0; r < n; r++) {
            const uint32_t m = masks ? masks[r] : 0xFFFFFFFFu;
            float mf;
            std::memcpy(&mf, &m, sizeof(mf));
            rays[2 * r] = make_float4(origin[3 * r], origin[3 * r + 1], origin[3 * r + 2], mf);
            rays[2 * r + 1] = make_float4(dir[3 * r], dir[3 * r + 1], dir[3 * r + 2], 0.0f);
        }
        HIPCHK(hipMemcpyAsync(qRays, rays.data(), sizeof(float4) * rays.size(), hipMemcpyHostToDevice, stream));
        k_pick<<<nblk(n), TPB, 0, stream>>>(S, wi.bodyBase, wi.bodyCount, n, qRays, qBody, qA);
        HIPCHK(hipGetLastError());
        std::vector<float4> a(n);
        HIPCHK(hipMemcpyAsync(body, qBody, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(a.data(), qA, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (local)
            for (int r = 0; r < n; r++) { local[3 * r] = a[r].x; local[3 * r + 1] = a[r].y; local[3 * r + 2] = a[r].z; }
        return AVBD_OK;
    }

    // per-body contact reports (scratch sized to the world's bodies and live manifolds)
    int *rMax = nullptr, *rEntries = nullptr, *rCount = nullptr;
    uint32_t *rCnt = nullptr, *rOff = nullptr, *rCursor = nullptr;
    float4 *rPos = nullptr, *rNormal = nullptr;
    size_t rEntryCap = 0, rOutCap = 0;

    template <typename T> avbd_status regrow(T *&p, size_t count)
    {
        if (p) {
            allocs.erase(std::remove(allocs.begin(), allocs.end(), static_cast<void *>(p)), allocs.end());
            (void)hipFree(p);
            p = nullptr;
        }
        hipError_t e = dalloc(p, count);
        if (e != hipSuccess) return fail(AVBD_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(p);
        return AVBD_OK;
    }

    avbd_status contactReport(int w, const int32_t *maxRep, int32_t stride, int32_t *count, float *position, float *normal,
                              int32_t *other)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        const int n = wi.bodyCount;
        if (!maxRep || !count || stride < 1) return fail(AVBD_ERR_INVALID, "bad arguments");
        int need = 0;
        for (int j = 0; j < n; j++) need = std::max(need, static_cast<int>(maxRep[j]));
        if (need > stride) return fail(AVBD_ERR_INVALID, "stride is smaller than the largest max_reported");
        if (n == 0) return AVBD_OK;
        avbd_status st;
        HIPCHK(hipStreamSynchronize(stream));
        int nLive = 0;
        HIPCHK(hipMemcpy(&nLive, S.wLive + w, sizeof(int), hipMemcpyDeviceToHost));
        if (!rMax) {
            if ((st = regrow(rMax, Ncap)) != AVBD_OK || (st = regrow(rCnt, Ncap + 1)) != AVBD_OK ||
                (st = regrow(rOff, Ncap + 1)) != AVBD_OK || (st = regrow(rCursor, Ncap)) != AVBD_OK ||
                (st = regrow(rCount, Ncap)) != AVBD_OK)
                return st;
        }
        if (static_cast<size_t>(2 * nLive + 1) > rEntryCap) {
            rEntryCap = std::max<size_t>(2 * nLive + 1, 2 * rEntryCap);
            if ((st = regrow(rEntries, rEntryCap)) != AVBD_OK) return st;
        }
        const size_t outN = static_cast<size_t>(n) * stride;
        if (outN > rOutCap) {
            rOutCap = std::max(outN, 2 * rOutCap);
            if ((st = regrow(rPos, rOutCap)) != AVBD_OK || (st = regrow(rNormal, rOutCap)) != AVBD_OK) return st;
        }
        HIPCHK(hipMemcpyAsync(rMax, maxRep, sizeof(int) * n, hipMemcpyHostToDevice, stream));
        k_fill_u32<<<nblk(n + 1), TPB, 0, stream>>>(rCnt, 0u, n + 1);
        k_fill_u32<<<nblk(n), TPB, 0, stream>>>(rCursor, 0u, n);
        if (nLive > 0) {
            buildLiveList();
            k_contact_unpack<<<liveGrid(), TPB, 0, stream>>>(S);
            k_report_count<<<nblk(nLive), TPB, 0, stream>>>(S, w, rMax, rCnt);
            scan(rCnt, rOff, n);   // exclusive: rOff[0..n], rOff[n] = total
            k_report_fill<<<nblk(nLive), TPB, 0, stream>>>(S, w, rMax, rOff, rCursor, rEntries);
        } else {
            k_fill_u32<<<nblk(n + 1), TPB, 0, stream>>>(rOff, 0u, n + 1);
        }
        k_report_emit<<<nblk(n), TPB, 0, stream>>>(S, w, n, rMax, stride, rOff, rEntries, rCount, rPos, rNormal);
        HIPCHK(hipGetLastError());
        std::vector<float4> p4(outN), n4(outN);
        HIPCHK(hipMemcpyAsync(count, rCount, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(p4.data(), rPos, sizeof(float4) * outN, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(n4.data(), rNormal, sizeof(float4) * outN, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int j = 0; j < n; j++)
            for (int c = 0; c < count[j]; c++) {
                const size_t o = static_cast<size_t>(j) * stride + c;
                if (position) { position[3 * o] = p4[o].x; position[3 * o + 1] = p4[o].y; position[3 * o + 2] = p4[o].z; }
                if (normal) { normal[3 * o] = n4[o].x; normal[3 * o + 1] = n4[o].y; normal[3 * o + 2] = n4[o].z; }
                if (other) {
                    int32_t ob;
                    std::memcpy(&ob, &p4[o].w, sizeof(ob));
                    other[o] = ob;
                }
            }
        return AVBD_OK;
    }

    avbd_status collectAabb(int w, const float mn[3], const float mx[3], uint32_t mask, int32_t *out, int32_t cap,
                            int32_t *count)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        if (!mn || !mx || !count || cap < 0 || (cap > 0 && !out)) return fail(AVBD_ERR_INVALID, "bad arguments");
        avbd_status st = ensureQuery(1);
        if (st != AVBD_OK) return st;
        const int n = wi.bodyCount;
        *count = 0;
        if (n == 0) return AVBD_OK;
        k_query_bounds<<<nblk(n), TPB, 0, stream>>>(S, wi.bodyBase, n, qBounds);
        k_collect_aabb<<<nblk(n), TPB, 0, stream>>>(S, wi.bodyBase, n, qBounds, make_float4(mn[0], mn[1], mn[2], 0),
                                                   make_float4(mx[0], mx[1], mx[2], 0), mask, qFlag);
        HIPCHK(hipGetLastError());
        std::vector<uint8_t> flag(n);
        HIPCHK(hipMemcpyAsync(flag.data(), qFlag, n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        int k = 0;
        for (int j = 0; j < n; j++)
            if (flag[j]) {
                if (k < cap) out[k] = j;
                k++;
            }
        *count = k;
        return k > cap ? fail(AVBD_ERR_CAPACITY, "more bodies than the output holds (count has the total)") : AVBD_OK;
    }

    avbd_status create(const avbd_params &p, int worldsCap, int bodies, int joints, int slots, bool batch)
    {
        prm = p;
        batchMode = batch;
        Wcap = std::max(worldsCap, 1);
        Ncap = std::max(bodies, 2);
        Jcap = std::max(joints, 1);
        Scap = std::max(slots, 8);
        Pcap = 12 * Ncap + 4096;
        worlds.assign(Wcap, WorldInfo{});
        bodyAlloc.init(Ncap);
        jointAlloc.init(Jcap);
        slotAlloc.init(Scap);
        if (!sizing) {
            int ndev = 0;
            if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
                return fail(AVBD_ERR_NO_DEVICE, "no HIP device visible");
            HIPCHK(hipGetDevice(&device));
            HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            HIPCHK(hipStreamCreateWithFlags(&upStream, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&upEvent, hipEventDisableTiming));
            for (auto &e : ev) HIPCHK(hipEventCreate(&e));
            HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&errHost), sizeof(int) * Wcap, hipHostMallocDefault));
            HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&countsHost), sizeof(int) * 64, hipHostMallocDefault));
            HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&phaseTicksHost), 8 * sizeof(uint64_t), hipHostMallocDefault));
            int khz = 0;
            if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess) wallClockKHz = khz;
        }
        avbd_status st;
#define AL(ptr, n) if ((st = A(ptr, n)) != AVBD_OK) return st
        const size_t N = Ncap, J = Jcap, M = Scap, W = Wcap, P = Pcap;
        S.numWorlds = Wcap; S.numBodies = Ncap; S.slotCap = Scap; S.jointCap = Jcap; S.pairCap = Pcap;
        AL(S.pos, N); AL(S.rot, N); AL(S.initLin, N); AL(S.initAng, N); AL(S.inertLin, N); AL(S.inertAng, N);
        AL(S.vel, N); AL(S.angVel, N); AL(S.prevVel, N); AL(S.sizeR, N); AL(S.massI, N); AL(S.poseRec, 4 * N); AL(S.narrowRec, NARROW_REC * N);
        AL(S.fric, N); AL(S.grav, N); AL(S.layer, N); AL(S.mask, N); AL(S.lockLin, N); AL(S.lockAng, N); AL(S.shape, N);
        AL(S.sleepMode, N); AL(S.sleeping, N); AL(S.stillFrames, N); AL(S.updFlag, N + 1); AL(S.updScan, N + 1);
        AL(S.mDormant, M); AL(S.mSaved, M); AL(S.dRA, M * MAXC); AL(S.dRB, M * MAXC); AL(S.dC0, M * MAXC);
        AL(S.dPen, M * MAXC); AL(S.pairParked, P); AL(S.pairApart, P); AL(rimDev, 16); S.rim = rimDev;
        S.armCap = S.armA = S.armB = nullptr;
        S.resC = nullptr;
        S.resJ = nullptr;
        if (p.jacobian_rebuild_distance > 0.0f) {
            AL(S.armCap, 2 * M); AL(S.armA, M * MAXC); AL(S.armB, M * MAXC);
        }
        if (!sizing) {
            // collide.cpp:733-753: rim directions at angle steps i*5 mod 8, from the C library's
            // cosf / sinf at run time (the values the reference's table holds)
            float rim[16];
            for (int i = 0; i < 8; i++) {
                volatile float a = 6.28318530718f * static_cast<float>(i * 5 % 8) / static_cast<float>(8);
                rim[i] = std::cos(static_cast<float>(a));
                rim[8 + i] = std::sin(static_cast<float>(a));
            }
            HIPCHK(hipMemcpy(rimDev, rim, sizeof(rim), hipMemcpyHostToDevice));
        }
        AL(S.bodyWorld, N); AL(S.updateIndex, N); AL(S.colour, N); AL(S.contactOnly, N);
        AL(S.dLin, N); AL(S.dAng, N); AL(S.aabbC, N); AL(S.aabbH, N);
        AL(S.fbCount, 1); AL(S.fbBody, N); AL(S.fbBlock, 33 * N);
        AL(S.adjCount, N + 1); AL(S.adjStart, N + 1); AL(S.adj, 2 * (M + J)); AL(S.adjTmp, 2 * (M + J)); AL(S.adjPos, 2 * (M + J));
        AL(S.jA, J); AL(S.jB, J); AL(S.jWorld, J); AL(S.jRA, J); AL(S.jRB, J); AL(S.jC0L, J); AL(S.jC0A, J);
        AL(S.jPenL, J); AL(S.jPenA, J); AL(S.jLamL, J); AL(S.jLamA, J); AL(S.jStiffL, J); AL(S.jStiffA, J);
        AL(S.jFrac, J); AL(S.jArm, J); AL(S.jBroken, J); AL(S.jDead, J); AL(S.jKind, J); AL(S.jRest, J);
        AL(S.jGRest, J); AL(S.jGMode, 6 * J); AL(S.jGLo, 6 * J); AL(S.jGHi, 6 * J); AL(S.jGK, 6 * J);
        AL(S.jGEq, 6 * J); AL(S.jGDamp, 6 * J); AL(S.jGLam, 6 * J); AL(S.jGPen, 6 * J); AL(S.jGInit, 6 * J);
        AL(S.mA, M); AL(S.mB, M); AL(S.mRec, 4 * M); AL(S.mNum, M);
        // sliced contacts: at most MAXC rows per lane of every slice, the last slice padded
        const size_t CD = (static_cast<size_t>(Scap) + CSLICE) * MAXC;
        AL(S.cRA, CD); AL(S.cRB, CD); AL(S.cC0, CD); AL(S.cPL, CD); AL(S.cLyz, CD); AL(S.cStick, CD);
        // staging: sliced by narrow-phase work item (at most MAXC rows per lane of every slice)
        AL(S.sRA, CD); AL(S.sRB, CD); AL(S.sC0, CD); AL(S.sPen, CD); AL(S.sLam, CD);
        AL(S.cOff, M); AL(S.cCount, M + 1); AL(S.cScan, M + 1); AL(S.sliceCount, 1);
        AL(S.stampA, STAMP_ROWS * M); AL(S.stampB, STAMP_ROWS * M); AL(S.slotWorld, M); AL(S.gLive, M); AL(S.wLiveOff, W + 1); AL(S.live, M); AL(S.liveTmp, M); AL(S.freeList, M);
        AL(S.keep, M + 1);
        AL(S.succA, M); AL(S.succB, M); AL(S.indeg, M); AL(S.levelOrder, M);
        AL(S.brakeLevels, W); AL(S.levelStart, M); AL(S.wLevelEnd, W);
        AL(S.projLA, M); AL(S.projAA, M); AL(S.projLB, M); AL(S.projAB, M); AL(S.projOn, M);
        AL(S.wBodyBase, W); AL(S.wBodyCount, W); AL(S.wJointBase, W); AL(S.wJointCount, W);
        AL(S.wSlotBase, W); AL(S.wSlotCap, W); AL(S.wLive, W); AL(S.wFree, W); AL(S.wTop, W); AL(S.wNew, W);
        AL(S.wHasJoints, W); AL(S.wUpdateCount, W); AL(S.wErr, W); AL(S.wProjDone, W); AL(S.wMaxDepth, W);
        AL(S.wColourLeft, W); AL(S.wColourTail, W);
        AL(S.mortonCode, N); AL(S.sortKey, N); AL(S.sortKeyAlt, N); AL(S.sortVal, N); AL(S.sortValAlt, N);
        AL(S.colKey, N); AL(S.colKeyAlt, N); AL(S.colVal, N); AL(S.colValAlt, N); AL(S.colStart, 16);
        AL(phaseTicks, 8);
        AL(S.nodeLeft, N); AL(S.nodeRight, N); AL(S.nodeParent, 2 * N); AL(S.nodeFirst, N); AL(S.nodeLast, N);
        AL(S.nodeFlag, N); AL(S.nodeBox, 4 * N); AL(S.nodeRope, 2 * N); AL(S.taskFlag, 2 * N + 1); AL(S.taskScan, 2 * N + 1); AL(S.taskNode, 2 * N); AL(S.colourCnt, N); AL(S.colourQueue, N); AL(S.leafBody, N); AL(S.bodyLeaf, N);
        AL(S.pairCount, N + 1); AL(S.pairOff, N + 1); AL(S.pairI, P); AL(S.pairJ, P);
        AL(S.pairValid, P + 1); AL(S.pairValidScan, P + 1);
        AL(S.pairScratch, N * PAIR_BUF); AL(S.wRoot, W);
        AL(S.listC, N); AL(S.listH, N); AL(S.listJ, N * LIST_K); AL(S.listCnt, N); AL(S.listAdd, N); AL(S.listFlag, N);
        AL(S.wRewalk, W); AL(S.wRefit, W); AL(S.wMoved, W); AL(S.wStale, W); AL(S.listStats, 4);
        AL(S.wBoundsLo, W); AL(S.wBoundsHi, W);
        wideWords = nblk(Ncap) * (TPB / 64);
        AL(S.wWide, W); AL(S.wideMask, wideWords + 1); AL(S.wideWord, wideWords + 1); AL(S.wideScan, wideWords + 1);
        AL(S.gOfK, M); AL(S.kSlot, M); AL(S.kT, M); AL(S.newT, P); AL(S.mHdr, M); AL(S.mBasisK, 3 * M); AL(S.mFricK, M); AL(S.brakeOn, M); AL(S.wOrder, W);
        radixBlocks = (Ncap + RADIX_TILE - 1) / RADIX_TILE;
        AL(radixHist, static_cast<size_t>(256) * radixBlocks + 1);
        const int biggest = std::max({Ncap + 1, Pcap + 1, Scap + 1, 256 * radixBlocks + 1, 2 * (Scap + Jcap)});
        tileSumsCap = (biggest + SCAN_TILE - 1) / SCAN_TILE + 1;
        AL(tileSums, tileSumsCap);
#undef AL
        if (sizing) return AVBD_OK;
        // all bodies / slots / joints start outside any world
        k_fill_i32<<<nblk(Ncap), TPB, 0, stream>>>(S.bodyWorld, -1, Ncap);
        k_fill_i32<<<nblk(Ncap), TPB, 0, stream>>>(S.updateIndex, -1, Ncap);
        k_fill_i32<<<nblk(Scap), TPB, 0, stream>>>(S.slotWorld, -1, Scap);
        k_fill_i32<<<nblk(Scap), TPB, 0, stream>>>(reinterpret_cast<int *>(S.cOff), 0, Scap);
        k_fill_i32<<<nblk(Scap), TPB, 0, stream>>>(S.mNum, 0, Scap);
        k_fill_i32<<<nblk(Jcap), TPB, 0, stream>>>(S.jWorld, -1, Jcap);
        k_fill_i32<<<nblk(Wcap), TPB, 0, stream>>>(S.wWide, NO_WIDE, Wcap);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
        setParams(S.P, prm, prm.dt);
        return refreshWorldOrder();
    }

    // ---- primitives -------------------------------------------------------------
    // `dlen`: a device-side length below n (only that prefix is scanned).
    void scan(const uint32_t *in, uint32_t *out, int n, const uint32_t *dlen = nullptr)
    {
        const int nt = (n + SCAN_TILE - 1) / SCAN_TILE;
        k_scan_tiles<<<std::max(nt, 1), TPB, 0, stream>>>(in, out, tileSums, n, dlen);
        k_scan_sums<<<1, SCAN_SUMS_TPB, 0, stream>>>(tileSums, std::max(nt, 1), out, n, dlen);
        k_scan_add<<<std::max(nt, 1), TPB, 0, stream>>>(out, tileSums, n, dlen);
    }

    // Wide bodies (State::wWide): chosen at a tree rebuild, after the world bounds ...
    int wideWords = 0;
    void widePick()
    {
        k_fill_i32<<<nblk(Wcap), TPB, 0, stream>>>(S.wWide, NO_WIDE, Wcap);
        k_wide_pick<<<nblk(Ncap), TPB, 0, stream>>>(S);
    }
    // ... and their pairs marked every step before the walks (wideMask bits and their ranks).
    void wideFlags()
    {
        k_wide_flags<<<nblk(Ncap), TPB, 0, stream>>>(S);
        scan(S.wideWord, S.wideScan, wideWords);
    }

    // Bodies sorted by colour, one stable 8-bit radix pass (the colour passes then walk their
    // own run instead of every body).
    void colourRuns()
    {
        const int n = Ncap, nb = radixBlocks;
        k_colour_keys<<<nblk(n), TPB, 0, stream>>>(S);
        k_radix_hist<<<nb, TPB, 0, stream>>>(S.colKey, n, 0, radixHist, nb);
        scan(radixHist, radixHist, 256 * nb);
        k_radix_scatter<<<nb, TPB, 0, stream>>>(S.colKey, S.colVal, S.colKeyAlt, S.colValAlt, n, 0, radixHist, nb);
        k_colour_starts<<<1, 64, 0, stream>>>(S, radixHist, nb);
    }
    // one lane per body of the largest possible run, in one-wavefront workgroups so that a run
    // spreads over every CU (the passes are bound by the latency of their dependent loads)
    static constexpr int COLOUR_PASS_TPB = 64;
    int colourGrid() const { return nblk(Ncap, COLOUR_PASS_TPB); }
    // four lanes per body; the grid walks longer colour runs in strides (512 workgroups, 64 per
    // XCD: on the chains world 2.58 ms per step against 2.62 with 1024 or 2048 and 2.67 with 256)
    int quadGrid() const { return std::min(nblk(4 * Ncap, QUAD_TPB), 512); }

    void radixSort()
    {
        const int n = Ncap, nb = radixBlocks;
        unsigned long long *kin = S.sortKey, *kout = S.sortKeyAlt;
        int *vin = S.sortVal, *vout = S.sortValAlt;
        for (int shift = 0; shift < 64; shift += 8) {
            k_radix_hist<<<nb, TPB, 0, stream>>>(kin, n, shift, radixHist, nb);
            scan(radixHist, radixHist, 256 * nb);
            k_radix_scatter<<<nb, TPB, 0, stream>>>(kin, vin, kout, vout, n, shift, radixHist, nb);
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
        // eight passes: the result is back in sortKey / sortVal
    }

    // Grid for kernels that walk the compact live list: its length is only known on the
    // device, so they grid-stride over a fixed grid sized to keep every CU busy.
    // (a multiple of NUM_XCD: FOR_LIVE splits the list into one part per XCD)
    int liveGrid(int tpb = TPB) const
    {
        const int g = std::min(nblk(std::max(Scap, Jcap), tpb), std::max(1, tune.liveBlocks * TPB / tpb));
        return std::max(NUM_XCD, (g + NUM_XCD - 1) / NUM_XCD * NUM_XCD);
    }

    // Block sizes of the hot kernels and the live-list grid, chosen by autotune() (results do
    // not depend on them: every kernel's work per element is independent of its launch shape).
    struct Tuning {
        int dual = TPB, primal = TPB, stamp = TPB, proj = TPB, traverse = TPB, liveBlocks = LIVE_BLOCKS;
        int colourLanes = 1;   // lanes per body in the colour passes (1 or 4)
    } tune;

    // gLive = every world's live positions, world after world (wLiveOff = exclusive scan of wLive).
    void buildLiveList()
    {
        scan(reinterpret_cast<const uint32_t *>(S.wLive), S.wLiveOff, Wcap);
        k_glive_fill<<<Wcap, TPB, 0, stream>>>(S);
        k_manifold_order<<<Wcap, ORDER_TPB, 0, stream>>>(S);
    }

    // Dense contact layout for the current live list (staging -> cRA...).
    void packContacts(int hostLayout)
    {
        // slice widths from the counts k_manifold_order wrote, their starts, then the copy
        k_slice_widths<<<nblk(Scap), TPB, 0, stream>>>(S);
        scan(S.cScan, S.cScan, Scap / CSLICE + 1, S.sliceCount);
        k_contact_pack<<<liveGrid(), TPB, 0, stream>>>(S, hostLayout);
    }

    // Live list and dense manifold view after host edits (contacts stay where they are).
    void refreshView()
    {
        buildLiveList();
        k_manifold_view<<<liveGrid(), TPB, 0, stream>>>(S);
    }

    void buildAdjacency()
    {
        const int n = Ncap;
        k_fill_i32<<<nblk(n + 1), TPB, 0, stream>>>(S.adjCount, 0, n + 1);
        k_adj_count<<<liveGrid(), TPB, 0, stream>>>(S);
        scan(reinterpret_cast<uint32_t *>(S.adjCount), reinterpret_cast<uint32_t *>(S.adjStart), n);
        k_adj_fill<<<liveGrid(), TPB, 0, stream>>>(S);
        k_adj_sort<<<nblk(n), TPB, 0, stream>>>(S);
    }

    bool anySleep() const
    {
        for (const auto &w : worlds)
            if (w.active && w.hasSleep) return true;
        return false;
    }

    bool anyRound() const
    {
        for (const auto &w : worlds)
            if (w.active && w.hasRound) return true;
        return false;
    }

    bool anyJoints() const
    {
        if (!prm.contact_only_overflow) return true;
        for (const auto &w : worlds)
            if (w.active && w.hasJoints) return true;
        return false;
    }

    // Refit task roots of the tree just built (subtrees of at most REFIT_LEAVES leaves).
    void buildRefitTasks()
    {
        k_task_flags<<<nblk(2 * Ncap - 1), TPB, 0, stream>>>(S, Ncap);
        scan(S.taskFlag, S.taskScan, 2 * Ncap - 1);
        k_task_list<<<nblk(2 * Ncap - 1), TPB, 0, stream>>>(S, Ncap);
    }
    int refitGrid() const { return std::max(1, std::min(Wcap + 2 * Ncap / REFIT_LEAVES + 8, 8192)); }
    avbd_status listStats(int32_t *out)
    {
        HIPCHK(hipMemcpyAsync(out, S.listStats, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return AVBD_OK;
    }

    // The state as the step's refit sees it: leaves are the candidate lists' grown boxes.
    State listTree() const
    {
        State T = S;
        T.aabbC = S.listC;
        T.aabbH = S.listH;
        return T;
    }

    // One velocity brake at the current poses: flags, dependency links, level-by-level
    // brake per world (solver_projection.cpp:96-167).
    void velocityBrake(float window, int countActive)
    {
        k_brake_flags<<<liveGrid(), TPB, 0, stream>>>(S, countActive);
        k_brake_links<<<nblk(Ncap), TPB, 0, stream>>>(S);
        k_brake_levels<<<Wcap, BRAKE_TPB, 0, stream>>>(S, window, 1);
    }

    // Per-world kernels take worlds in this order, most bodies first, so the longest
    // workgroups start first.
    avbd_status refreshWorldOrder()
    {
        std::vector<int> order(Wcap);
        for (int w = 0; w < Wcap; w++) order[w] = w;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            const int na = worlds[a].active ? worlds[a].bodyCount : -1;
            const int nb = worlds[b].active ? worlds[b].bodyCount : -1;
            return na > nb;
        });
        HIPCHK(hipMemcpyAsync(S.wOrder, order.data(), sizeof(int) * Wcap, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return AVBD_OK;
    }

    // ---- the step -------------------------------------------------------------------
    // Enqueue one reference step (substeps <= 1) or interleaved tick on `stream`.
    void enqueueStep(int substeps, int roundsPer, bool rebuildTree, bool withColours, bool timed)
    {
        const bool inter = substeps > 1;
        const float collisionDt = inter ? prm.dt * static_cast<float>(substeps) : prm.dt;
        setParams(S.P, prm, collisionDt);
        const bool sleepOn = anySleep();
        S.P.sleepOn = sleepOn ? 1 : 0;
        const int nb = nblk(Ncap);
        if (timed) phaseMark(0);
        if (!csrValid) {
            refreshView();
            buildAdjacency();
        }
        // ---- broad phase (solver_broadphase.cpp:85-186)
        // every world lists again at a tree rebuild, dropping entries that no longer meet
        if (rebuildTree) k_fill_i32<<<nblk(Wcap), TPB, 0, stream>>>(S.wRewalk, 1, Wcap);
        k_swept_aabb<<<nb, TPB, 0, stream>>>(S);
        if (timed) phaseMark(5);
        if (rebuildTree) {
            k_world_bounds<<<Wcap, TPB, 0, stream>>>(S);
            widePick();
            k_morton<<<nb, TPB, 0, stream>>>(S, batchMode ? 0 : 1, 0);
            pbeg(AVBD_K_SORT);
            radixSort();
            pend();
            k_karras<<<nblk(Ncap - 1), TPB, 0, stream>>>(S, Ncap);
            k_world_roots_init<<<nblk(Wcap), TPB, 0, stream>>>(S, Ncap);
            k_world_roots<<<nblk(Ncap - 1), TPB, 0, stream>>>(S, Ncap);
            k_world_roots_check<<<nblk(Wcap), TPB, 0, stream>>>(S);
            k_ropes<<<nblk(2 * Ncap - 1), TPB, 0, stream>>>(S, Ncap);
            buildRefitTasks();
        } else {
            k_refit_clear<<<nblk(Ncap), TPB, 0, stream>>>(S, Ncap);
        }
        pbeg(AVBD_K_REFIT);
        k_refit_tasks<<<refitGrid(), REFIT_TPB, 0, stream>>>(listTree(), Ncap, rebuildTree ? 0 : 1);
        pend();
        if (timed) phaseMark(6);
        pbeg(AVBD_K_TRAVERSE);
        wideFlags();
        k_list_walk<<<nblk(Ncap, tune.traverse), tune.traverse, 0, stream>>>(S, Ncap);
        k_traverse<<<nblk(Ncap, tune.traverse), tune.traverse, 0, stream>>>(S, Ncap);
        pend();
        k_fill_i32<<<1, 64, 0, stream>>>(S.listStats, 0, 4);
        k_list_world_end<<<nblk(Wcap), TPB, 0, stream>>>(S);
        scan(S.pairCount, S.pairOff, Ncap);
        k_pair_emit<<<nb, TPB, 0, stream>>>(S, Ncap);
        if (timed) phaseMark(7);
        k_pair_filter<<<liveGrid(), TPB, 0, stream>>>(S);
        if (sleepOn) k_wake_pairs<<<nblk(Wcap), TPB, 0, stream>>>(S);
        scan(S.pairValid, S.pairValidScan, Pcap, S.pairOff + Ncap);
        k_alloc_new<<<liveGrid(), TPB, 0, stream>>>(S);
        k_alloc_counts<<<nblk(Wcap), TPB, 0, stream>>>(S);
        if (timed) phaseMark(1);
        // ---- warm start forces / narrow phase (solver_solve.cpp:31-105)
        k_joint_init<<<nblk(Jcap), TPB, 0, stream>>>(S);
        pbeg(AVBD_K_NARROW);
        // work items: the previous step's view (wLiveOff, kSlot, kT, mHdr still hold it), then
        // this step's new manifolds
        k_fill_u32<<<nblk(Scap), TPB, 0, stream>>>(S.keep, 0u, Scap);
        if (prm.contact_grace_ticks > 0) {
            // contact grace: one pass for every pair with the parking rules
            k_manifold_init_grace<<<liveGrid(NARROW_TPB), NARROW_TPB, 0, stream>>>(S);
        } else {
            k_manifold_init<<<liveGrid(NARROW_TPB), NARROW_TPB, 0, stream>>>(S);
            // pairs with a sphere or a cylinder take their own pass (only launched when such
            // bodies exist, so box scenes keep the box kernel's register budget and time)
            if (anyRound()) k_manifold_init_round<<<liveGrid(NARROW_TPB), NARROW_TPB, 0, stream>>>(S);
        }
        pend();
        k_compact_world<<<Wcap, TPB, 0, stream>>>(S);
        buildLiveList();
        packContacts(0);
        pbeg(AVBD_K_ADJACENCY);
        buildAdjacency();
        pend();
        if (sleepOn) {
            // solver_solve.cpp:264-339: wake pass, then this step's update set and its ranks
            k_wake_neighbours<<<nblk(Wcap), TPB, 0, stream>>>(S);
            k_update_flags<<<nb, TPB, 0, stream>>>(S);
            scan(S.updFlag, S.updScan, Ncap);
            k_update_index<<<nb, TPB, 0, stream>>>(S);
        }
        // ---- update set and colouring (solver_solve.cpp:263-424)
        k_colour_prepare<<<nblk(std::max(Ncap, Wcap)), TPB, 0, stream>>>(S);
        k_colour_joints<<<nblk(Jcap), TPB, 0, stream>>>(S);
        k_colour_init<<<nb, TPB, 0, stream>>>(S);
        if (withColours) {
            for (int r = 0; r + 1 < colourRounds; r++) k_colour_round<<<nb, TPB, 0, stream>>>(S, 0);
            k_fill_i32<<<nblk(Wcap), TPB, 0, stream>>>(S.wColourLeft, 0, Wcap);
            k_colour_round<<<nb, TPB, 0, stream>>>(S, 1);
            k_fill_i32<<<nblk(Wcap), TPB, 0, stream>>>(S.wColourTail, 0, Wcap);
            k_colour_counts<<<nb, TPB, 0, stream>>>(S);
            k_colour_finish<<<Wcap, COLOUR_TPB, 0, stream>>>(S);
        }
        k_colour_override<<<nb, TPB, 0, stream>>>(S);
        if (withColours) colourRuns();
        // ---- brake schedule and first brake (solver_projection.cpp:96-167)
        k_fill_i32<<<nblk(Wcap), TPB, 0, stream>>>(S.wUpdateCount, 0, Wcap);
        const bool brake = prm.speculative_margin > 0.0f;
        pbeg(AVBD_K_BRAKE_LEVELS);
        if (brake) {
            velocityBrake(collisionDt, 1);
        } else {
            k_brake_flags<<<liveGrid(), TPB, 0, stream>>>(S, 1);
            k_fill_i32<<<nblk(Wcap), TPB, 0, stream>>>(S.brakeLevels, 0, Wcap);
        }
        pend();
        if (timed) phaseMark(2);
        // ---- substeps (solver_step.cpp:79-105, 142-150)
        const int rounds = !inter ? std::max(1, prm.iterations)
                                  : (roundsPer > 0 ? std::max(1, roundsPer)
                                                   : std::max(2, (prm.iterations + substeps - 1) / substeps));
        const int newtonRounds =
                std::clamp(static_cast<int>(std::ceil(prm.newton_ratio * static_cast<float>(rounds))), 1, rounds);
        const int subs = inter ? substeps : 1;
        // solver_step.cpp:299-314: the convergence probe compares a round's mean penetration with
        // the previous round's, starting from 0 each substep (:88-89).  It reads the step-start C0,
        // which no round changes, so residual >= 0 and 0 - residual < threshold always holds: with
        // a positive threshold the reference stops after the first round of every substep.
        const int runRounds = prm.convergence_threshold > 0.0f && rounds > 1 ? 1 : rounds;
        // manifold.cpp:600-607: drift-gated arm reuse
        const bool arms = prm.jacobian_rebuild_distance > 0.0f;
        for (int sub = 0; sub < subs; sub++) {
            if (sub > 0) {
                k_rebase<<<liveGrid(), TPB, 0, stream>>>(S);
                if (brake) {
                    pbeg(AVBD_K_BRAKE_REPLAY);
                    velocityBrake(prm.dt * static_cast<float>(subs - sub), 0);
                    pend();
                }
            }
            pbeg(AVBD_K_BODY);
            k_warmstart<<<nb, TPB, 0, stream>>>(S);
            pend();
            for (int it = 0; it < runRounds; it++) {
                const bool newton = it < newtonRounds;
                // Jacobi stamps come from the previous dual unless colour passes moved bodies
                // since (or this is the first iteration of the substep).
                const bool stampHere = withColours || it == 0;
                const bool stampInDual = !withColours && it + 1 < std::min(newtonRounds, runRounds);
                if (newton) {
                    if (withColours)
                        for (int c = 0; c < 8; c++) {
                            pbeg(AVBD_K_COLOUR_PASS);
                            if (tune.colourLanes == 4) {
                                if (arms)
                                    k_colour_pass_quad<true><<<quadGrid(), QUAD_TPB, 0, stream>>>(S, c, it > 0 ? 1 : 0);
                                else
                                    k_colour_pass_quad<false><<<quadGrid(), QUAD_TPB, 0, stream>>>(S, c, 0);
                            } else if (arms) {
                                k_colour_pass<true><<<colourGrid(), COLOUR_PASS_TPB, 0, stream>>>(S, c, it > 0 ? 1 : 0);
                            } else {
                                k_colour_pass<false><<<colourGrid(), COLOUR_PASS_TPB, 0, stream>>>(S, c, 0);
                            }
                            pend();
                        }
                    if (stampHere) {
                        // pose records: written by the warm start and by every move since
                        pbeg(AVBD_K_STAMP);
                        if (arms)
                            k_stamp<true><<<liveGrid(tune.stamp), tune.stamp, 0, stream>>>(S);
                        else
                            k_stamp<false><<<liveGrid(tune.stamp), tune.stamp, 0, stream>>>(S);
                        pend();
                    }
                    pbeg(AVBD_K_PRIMAL);
                    if (withColours) {
                        k_overflow_solve<false><<<nblk(Ncap, tune.primal), tune.primal, 0, stream>>>(S);
                        pend();
                        k_solve_fallback<<<1, TPB, 0, stream>>>(S, 0);
                        pbeg(AVBD_K_BODY);
                        k_overflow_apply<<<nb, TPB, 0, stream>>>(S);
                        pend();
                    } else {
                        // no joints: bodies move inside the solve and refresh their pose records
                        k_overflow_solve<true><<<nblk(Ncap, tune.primal), tune.primal, 0, stream>>>(S);
                        pend();
                        k_solve_fallback<<<1, TPB, 0, stream>>>(S, 2);
                    }
                }
                pbeg(AVBD_K_DUAL);
                const float decay = newton ? 1.0f : prm.stiffness_decay;
                if (recordRes && it + 1 == runRounds) {
                    // the step's last round keeps its constraint values (never a stamping dual)
                    if (arms) k_dual<false, true, true><<<liveGrid(tune.dual), tune.dual, 0, stream>>>(S, decay);
                    else k_dual<false, false, true><<<liveGrid(tune.dual), tune.dual, 0, stream>>>(S, decay);
                } else if (arms) {
                    if (stampInDual) k_dual<true, true><<<liveGrid(tune.dual), tune.dual, 0, stream>>>(S, decay);
                    else k_dual<false, true><<<liveGrid(tune.dual), tune.dual, 0, stream>>>(S, decay);
                } else {
                    if (stampInDual) k_dual<true, false><<<liveGrid(tune.dual), tune.dual, 0, stream>>>(S, decay);
                    else k_dual<false, false><<<liveGrid(tune.dual), tune.dual, 0, stream>>>(S, decay);
                }
                pend();
            }
            pbeg(AVBD_K_BODY);
            k_finish<<<nb, TPB, 0, stream>>>(S, sub + 1 == subs ? 1 : 0);
            pend();
        }
        if (timed) phaseMark(3);
        // ---- final brake, position projection, rebase (solver_step.cpp:69-70)
        if (brake) {
            pbeg(AVBD_K_BRAKE_REPLAY);
            velocityBrake(collisionDt, 0);
            pend();
        }
        if (prm.position_projection && prm.projection_iterations > 0) {
            k_proj_reset<<<nblk(Wcap), TPB, 0, stream>>>(S);
            for (int pass = 0; pass < prm.projection_iterations; pass++) {
                pbeg(AVBD_K_PROJ_MANIFOLD);
                k_proj_manifold<<<liveGrid(tune.proj), tune.proj, 0, stream>>>(S);
                pend();
                pbeg(AVBD_K_PROJ_BODY);
                k_proj_body<<<nb, TPB, 0, stream>>>(S);
                pend();
                k_proj_flags<<<nblk(Wcap), TPB, 0, stream>>>(S);
            }
        }
        k_rebase<<<liveGrid(), TPB, 0, stream>>>(S);
        if (timed) phaseMark(4);
        csrValid = true;
    }

    // Worlds whose last step failed, with their reasons (AVBD_WORLD_* bits).
    std::vector<int32_t> failedIds, failedWhy;

    // Per-world error words of the step just enqueued.  The kernels never write outside a
    // world's ranges when its capacity runs out (they skip the work and set a bit), so one
    // world's failure leaves the others' results intact: a batch reports the failed worlds
    // (AVBD_ERR_WORLD, listed by avbd_batch_failed_worlds) and its caller evicts them.  A plain
    // world keeps the reference's behaviour for non-finite state (no check) and fails its step
    // on the other bits.  Overflowing the shared candidate-pair capacity fails the whole step.
    avbd_status checkErrors()
    {
        avbd_status st = queueErrorCopies();
        if (st != AVBD_OK) return st;
        return readErrors();
    }
    // The step's error words and pair total to pinned host memory (after the step's kernels).
    avbd_status queueErrorCopies()
    {
        HIPCHK(hipMemcpyAsync(errHost, S.wErr, sizeof(int) * Wcap, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(countsHost, S.pairOff + Ncap, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(countsHost + 4, S.listStats, 4 * sizeof(int), hipMemcpyDeviceToHost, stream));
        return AVBD_OK;
    }
    avbd_status readErrors()
    {
        uint32_t total = 0;
        HIPCHK(hipStreamSynchronize(stream));
        std::memcpy(&total, countsHost, sizeof(uint32_t));
        movingWorlds = countsHost[4 + 1] + countsHost[4 + 2];
        failedIds.clear();
        failedWhy.clear();
        for (int w = 0; w < Wcap; w++) {
            const int e = batchMode ? errHost[w] : (errHost[w] & ~AVBD_WORLD_NONFINITE);
            if (e) { failedIds.push_back(w); failedWhy.push_back(e); }
        }
        bool any = false;
        for (int w = 0; w < Wcap; w++) any = any || errHost[w] != 0;
        if (any) {
            HIPCHK(hipMemsetAsync(S.wErr, 0, sizeof(int) * Wcap, stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
        if (total > static_cast<uint32_t>(Pcap)) return fail(AVBD_ERR_CAPACITY, "candidate pair capacity exceeded");
        if (failedIds.empty()) return AVBD_OK;
        const int w = failedIds[0], e = failedWhy[0];
        const std::string what = (e & AVBD_WORLD_SLOTS)      ? "manifold slot capacity exceeded"
                                 : (e & AVBD_WORLD_TREE)     ? "broad-phase tree walk failed"
                                                             : "non-finite body state";
        if (batchMode)
            return fail(AVBD_ERR_WORLD, std::to_string(failedIds.size()) + " world(s) failed this step (first: world " +
                                            std::to_string(w) + ", " + what + "); the others stepped");
        return fail((e & AVBD_WORLD_SLOTS) ? AVBD_ERR_CAPACITY : AVBD_ERR_DEVICE, what);
    }

    // Phase boundaries of a timed step: a one-thread kernel stores the device's constant-rate
    // clock (wall_clock64) at each mark, so the marks time the step inside a captured graph too.
    uint64_t *phaseTicks = nullptr;    // device, 8 marks
    uint64_t *phaseTicksHost = nullptr;
    double wallClockKHz = 0.0;
    void phaseMark(int k) { k_phase_mark<<<1, 64, 0, stream>>>(phaseTicks, k); }
    float phaseMs(int a, int b) const
    {
        const uint64_t t0 = phaseTicksHost[a], t1 = phaseTicksHost[b];
        return wallClockKHz > 0.0 && t1 >= t0 ? static_cast<float>(static_cast<double>(t1 - t0) / wallClockKHz) : 0.0f;
    }

    // Captured plain steps, one per (substeps, rounds, tree rebuilt, joints, round shapes,
    // sleeping): a plain world's step is one graph launch, like a batch's.  The phase events are
    // captured with it (event record nodes); a step whose launch shapes or parameters changed
    // is captured again.
    std::map<std::array<int, 6>, hipGraphExec_t> plainGraphs;
    avbd_params plainGraphPrm{};
    bool plainGraphUnusable = false;   // capture failed once: step kernel by kernel
    void dropPlainGraphs()
    {
        for (auto &kv : plainGraphs)
            if (kv.second) (void)hipGraphExecDestroy(kv.second);
        plainGraphs.clear();
    }
    hipGraphExec_t plainGraph(int substeps, int rounds, bool rebuild)
    {
        if (std::memcmp(&plainGraphPrm, &prm, sizeof(prm)) != 0 || !graphsValid) {
            dropPlainGraphs();
            plainGraphPrm = prm;
            graphsValid = true;   // the batch graphs share the validity flag; a plain world has none
        }
        const std::array<int, 6> key{substeps, rounds, rebuild ? 1 : 0, anyJoints() ? 1 : 0, anyRound() ? 1 : 0,
                                     anySleep() ? 1 : 0};
        auto it = plainGraphs.find(key);
        if (it != plainGraphs.end()) return it->second;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
        enqueueStep(substeps, rounds, rebuild, anyJoints(), true);
        if (hipStreamEndCapture(stream, &graph) != hipSuccess || !graph) return nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return nullptr;
        plainGraphs[key] = exec;
        return exec;
    }

    avbd_status stepPlain(int substeps, int rounds)
    {
        const bool rebuild = treeAge >= rebuildInterval();
        treeAge = rebuild ? 0 : treeAge + 1;
        hipGraphExec_t g = nullptr;
        if (useGraph && !plainGraphUnusable) {
            if (!csrValid) {
                refreshView();
                buildAdjacency();
                csrValid = true;
            }
            g = plainGraph(substeps, rounds, rebuild);
            if (!g) {
                (void)hipGetLastError();
                plainGraphUnusable = true;
            }
        }
        if (g) HIPCHK(hipGraphLaunch(g, stream));
        else enqueueStep(substeps, rounds, rebuild, anyJoints(), true);
        HIPCHK(hipGetLastError());
        avbd_status st = checkErrors();
        if (st != AVBD_OK) return st;
        HIPCHK(hipMemcpy(phaseTicksHost, phaseTicks, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost));
        const float t01 = phaseMs(0, 1), t12 = phaseMs(1, 2), t23 = phaseMs(2, 3), t34 = phaseMs(3, 4);
        stats.ms_broad_tree = phaseMs(5, 6);
        stats.ms_broad_emit = phaseMs(6, 7);
        stats.ms_broad = t01;
        stats.ms_narrow = t12;
        stats.ms_solve = t23;
        stats.ms_finish = t34;
        stats.ms_total = t01 + t12 + t23 + t34;
        return AVBD_OK;
    }

    avbd_status profile(int steps, int substeps, int rounds, avbd_kernel_profile *out)
    {
        if (steps < 1) return fail(AVBD_ERR_INVALID, "steps must be >= 1");
        for (int k = 0; k < AVBD_K_CLASSES; k++) { profMs[k] = 0; profLaunches[k] = 0; }
        profiling = true;
        double stepMs = 0.0;
        for (int s2 = 0; s2 < steps; s2++) {
            const bool rebuild = treeAge >= rebuildInterval();
            treeAge = rebuild ? 0 : treeAge + 1;
            (void)hipEventRecord(ev[0], stream);
            enqueueStep(substeps, rounds, rebuild, anyJoints(), false);
            (void)hipEventRecord(ev[4], stream);
            profCollect();
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, ev[0], ev[4]);
            stepMs += t;
            avbd_status st = checkErrors();
            if (st != AVBD_OK) { profiling = false; return st; }
        }
        profiling = false;
        HIPCHK(hipGetLastError());
        std::memset(out, 0, sizeof(*out));
        out->steps = steps;
        out->step_ms = static_cast<float>(stepMs / steps);
        for (int k = 0; k < AVBD_K_CLASSES; k++) {
            out->launches[k] = static_cast<int32_t>(profLaunches[k]);
            out->avg_ms[k] = profLaunches[k] ? static_cast<float>(profMs[k] / profLaunches[k]) : 0.0f;
        }
        return AVBD_OK;
    }

    // Measure candidate launch shapes on real steps (the steps are part of the simulation:
    // results do not depend on the shapes) and keep the fastest per kernel class.
    // Launch shapes given by the caller (for instance ones an earlier autotune() chose).
    avbd_status setTuning(const avbd_tuning &t)
    {
        const int v[5] = {t.block_dual, t.block_primal, t.block_stamp, t.block_projection, t.block_traverse};
        for (int x : v)
            if (x != 64 && x != 128 && x != 256) return fail(AVBD_ERR_INVALID, "block sizes must be 64, 128 or 256");
        if (t.live_blocks < NUM_XCD || t.live_blocks > (1 << 20))
            return fail(AVBD_ERR_INVALID, "live_blocks must be in [8, 2^20]");
        if (t.colour_lanes != 0 && t.colour_lanes != 1 && t.colour_lanes != 4)
            return fail(AVBD_ERR_INVALID, "colour_lanes must be 1 or 4 (0: 1)");
        tune.dual = t.block_dual;
        tune.primal = t.block_primal;
        tune.stamp = t.block_stamp;
        tune.proj = t.block_projection;
        tune.traverse = t.block_traverse;
        tune.liveBlocks = t.live_blocks;
        tune.colourLanes = t.colour_lanes == 4 ? 4 : 1;
        graphsValid = false;
        return AVBD_OK;
    }

    avbd_status autotune(int stepsPerTrial, avbd_tuning *out)
    {
        if (stepsPerTrial < 1) return fail(AVBD_ERR_INVALID, "steps_per_trial must be >= 1");
        const int sizes[3] = {64, 128, 256};
        float t[3][AVBD_K_CLASSES];
        avbd_kernel_profile prof;
        for (int c = 0; c < 3; c++) {
            tune.dual = tune.primal = tune.stamp = tune.proj = tune.traverse = sizes[c];
            avbd_status st = profile(stepsPerTrial, 1, 0, &prof);
            if (st != AVBD_OK) return st;
            for (int k = 0; k < AVBD_K_CLASSES; k++) t[c][k] = prof.avg_ms[k];
        }
        auto best = [&](int cls) {
            int bi = 2;
            for (int c = 0; c < 3; c++)
                if (t[c][cls] > 0.0f && t[c][cls] < t[bi][cls]) bi = c;
            return sizes[bi];
        };
        tune.dual = best(AVBD_K_DUAL);
        tune.primal = best(AVBD_K_PRIMAL);
        tune.stamp = best(AVBD_K_STAMP);
        tune.proj = best(AVBD_K_PROJ_MANIFOLD);
        tune.traverse = best(AVBD_K_TRAVERSE);
        const int grids[5] = {64, 256, 1024, 2048, 4096};   // few live manifolds want a small grid
        float bestT = 0.0f;
        int bestG = LIVE_BLOCKS;
        for (int g : grids) {
            tune.liveBlocks = g;
            avbd_status st = profile(stepsPerTrial, 1, 0, &prof);
            if (st != AVBD_OK) return st;
            const float tt = prof.avg_ms[AVBD_K_DUAL] * prof.launches[AVBD_K_DUAL] +
                             prof.avg_ms[AVBD_K_STAMP] * prof.launches[AVBD_K_STAMP] +
                             prof.avg_ms[AVBD_K_PROJ_MANIFOLD] * prof.launches[AVBD_K_PROJ_MANIFOLD];
            if (bestT == 0.0f || tt < bestT) { bestT = tt; bestG = g; }
        }
        tune.liveBlocks = bestG;
        // colour passes: one lane per body, or four (a body's adjacency entries in parallel)
        if (prof.launches[AVBD_K_COLOUR_PASS] > 0) {
            float lanesT[2];
            const int lanes[2] = {1, 4};
            for (int c = 0; c < 2; c++) {
                tune.colourLanes = lanes[c];
                avbd_status st = profile(stepsPerTrial, 1, 0, &prof);
                if (st != AVBD_OK) return st;
                lanesT[c] = prof.avg_ms[AVBD_K_COLOUR_PASS];
            }
            tune.colourLanes = lanesT[1] < lanesT[0] ? 4 : 1;
        }
        graphsValid = false;
        if (out) {
            out->block_dual = tune.dual;
            out->block_primal = tune.primal;
            out->block_stamp = tune.stamp;
            out->block_projection = tune.proj;
            out->block_traverse = tune.traverse;
            out->live_blocks = tune.liveBlocks;
            out->colour_lanes = tune.colourLanes;
        }
        return AVBD_OK;
    }

    avbd_status captureGraphs()
    {
        for (auto &g : graphs)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
        const bool joints = anyJoints();
        for (int k = 0; k < 2; k++) {
            hipGraph_t graph;
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            const bool savedCsr = csrValid;
            csrValid = true;
            enqueueStep(1, 0, k == 0, joints, false);
            csrValid = savedCsr;
            HIPCHK(hipStreamEndCapture(stream, &graph));
            HIPCHK(hipGraphInstantiate(&graphs[k], graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
        }
        graphJoints = joints;
        graphRound = anyRound();
        graphSleep = anySleep();
        graphsValid = true;
        return AVBD_OK;
    }

    bool useGraph = true;

    avbd_status stepBatch()
    {
        avbd_status st = stepBatchLaunch();
        if (st != AVBD_OK) return st;
        return stepBatchWait();
    }
    // The step's work on the stream, returning at once: worlds uploaded since the last step are
    // activated first; until stepBatchWait only asynchronous admissions may be made.
    avbd_status stepBatchLaunch()
    {
        if (stepInFlight) return fail(AVBD_ERR_INVALID, "a step is in flight");
        avbd_status st = activatePending();
        if (st != AVBD_OK) return st;
        if (!csrValid) {
            refreshView();
            buildAdjacency();
            csrValid = true;
        }
        if (!useGraph) {
            const bool rebuild = treeAge >= rebuildInterval();
            treeAge = rebuild ? 0 : treeAge + 1;
            enqueueStep(1, 0, rebuild, anyJoints(), false);
            HIPCHK(hipGetLastError());
        } else {
            if (!graphsValid || graphJoints != anyJoints() || graphRound != anyRound() || graphSleep != anySleep()) {
                st = captureGraphs();
                if (st != AVBD_OK) return st;
            }
            const bool rebuild = treeAge >= rebuildInterval();
            treeAge = rebuild ? 0 : treeAge + 1;
            HIPCHK(hipGraphLaunch(graphs[rebuild ? 0 : 1], stream));
        }
        if ((st = queueErrorCopies()) != AVBD_OK) return st;
        stepInFlight = true;
        return AVBD_OK;
    }
    avbd_status stepBatchWait()
    {
        if (!stepInFlight) return fail(AVBD_ERR_INVALID, "no step in flight");
        stepInFlight = false;
        avbd_status st = readErrors();
        arenaQuiesce();
        return st;
    }
    // Entry points other than the step and asynchronous admission: no step in flight, and
    // worlds admitted asynchronously made active first.
    avbd_status idle()
    {
        if (stepInFlight) return fail(AVBD_ERR_INVALID, "a step is in flight: call avbd_batch_step_wait first");
        avbd_status st = activatePending();
        if (st != AVBD_OK) return st;
        arenaQuiesce();
        return AVBD_OK;
    }

    // ---- upload / download --------------------------------------------------------
    // A world's arrays are packed into pinned host memory (the arena) and copied from there:
    // its data arrays at once (on the main stream, or on the upload stream for an admission
    // that overlaps a running step), its identity arrays (bodyWorld, updateIndex, jWorld,
    // slotWorld, the world table), which every kernel checks first, on the main stream when
    // the world becomes active.  The arena is reused once the streams have drained.
    struct Copy { void *dst; const void *src; size_t bytes; };
    struct Pending { int w; std::vector<Copy> activate; };
    unsigned char *arena = nullptr;
    size_t arenaCap = 0, arenaUsed = 0;
    std::vector<unsigned char *> arenaOld;
    hipStream_t upStream = nullptr;
    hipEvent_t upEvent = nullptr;
    bool upRecorded = false;
    std::vector<Pending> pending;
    bool stepInFlight = false;

    avbd_status ensureArena(size_t bytes)
    {
        if (arenaUsed + bytes <= arenaCap) return AVBD_OK;
        const size_t cap = std::max<size_t>({bytes, 2 * arenaCap, 1 << 20});
        unsigned char *p = nullptr;
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&p), cap, hipHostMallocDefault));
        if (arena) {
            if (arenaUsed == 0 && pending.empty() && !stepInFlight) (void)hipHostFree(arena);
            else arenaOld.push_back(arena);   // copies from it may still be queued
        }
        arena = p;
        arenaCap = cap;
        arenaUsed = 0;
        return AVBD_OK;
    }
    const void *arenaPut(const void *src, size_t bytes)
    {
        unsigned char *p = arena + arenaUsed;
        std::memcpy(p, src, bytes);
        arenaUsed += (bytes + 255) & ~static_cast<size_t>(255);
        return p;
    }
    // Both streams drained: the arena's copies are done.
    void arenaQuiesce()
    {
        if (!pending.empty() || stepInFlight) return;
        arenaUsed = 0;
        for (auto *p : arenaOld) (void)hipHostFree(p);
        arenaOld.clear();
    }
    // Worlds uploaded since the last step become active (main stream, after their copies).
    avbd_status activatePending()
    {
        if (pending.empty()) return AVBD_OK;
        if (upRecorded) HIPCHK(hipStreamWaitEvent(stream, upEvent, 0));
        for (auto &p : pending) {
            for (auto &c : p.activate) HIPCHK(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, stream));
            worlds[p.w].pending = false;
            worlds[p.w].active = true;
        }
        pending.clear();
        csrValid = false;
        treeAge = 1 << 20;
        return refreshWorldOrder();   // waits for the copies
    }

    // async: the data copies go on the upload stream and the world is activated by the next
    // step (avbd_batch_add_world_async); otherwise it is active on return.
    avbd_status uploadWorld(int w, const avbd_bodies *b, const avbd_joints *j, int slotCap, bool async = false)
    {
        const int nj = j ? j->count : 0;
        avbd_status st = ensureArena(209 * static_cast<size_t>(b->count) + 377 * static_cast<size_t>(nj) +
                                     4 * static_cast<size_t>(slotCap) + 64 * 256 + 4096);
        if (st != AVBD_OK) return st;
        const size_t mark = arenaUsed;
        st = stageWorld(w, b, j, slotCap, async);
        if (st != AVBD_OK && !worlds[w].pending) arenaUsed = mark;   // rejected input: nothing was queued
        return st;
    }

    avbd_status stageWorld(int w, const avbd_bodies *b, const avbd_joints *j, int slotCap, bool async)
    {
        const int n = b->count;
        const int nj = j ? j->count : 0;
        hipStream_t cs = async ? upStream : stream;
        std::vector<Copy> data;
        Pending act{w, {}};
        WorldInfo &wi = worlds[w];
        wi.bodyCount = n;
        wi.jointCount = nj;
        wi.slotCap = slotCap;
        wi.hasJoints = nj > 0;
        wi.hasRound = false;
        wi.hasSleep = false;
        std::vector<float4> pos(n), rot(n), vel(n), ang(n), prev(n), sizeR(n), massI(n);
        std::vector<float> fric(n), grav(n);
        std::vector<uint32_t> layer(n), mask(n);
        std::vector<uint8_t> ll(n), la(n), shp(n, AVBD_SHAPE_BOX), smode(n, AVBD_SLEEP_NEVER), slp(n, 0);
        std::vector<int> still(n, 0);
        std::vector<int> bw(n, w), ui(n, -1);
        int movable = 0;
        for (int i = 0; i < n; i++) {
            if (b->shape) {
                if (b->shape[i] > AVBD_SHAPE_CYLINDER) return fail(AVBD_ERR_INVALID, "unknown body shape");
                shp[i] = b->shape[i];
                if (shp[i] != AVBD_SHAPE_BOX) wi.hasRound = true;
            }
            if (b->sleep_mode) {
                if (b->sleep_mode[i] > AVBD_SLEEP_START_ASLEEP) return fail(AVBD_ERR_INVALID, "unknown sleep mode");
                smode[i] = b->sleep_mode[i];
            }
            if (b->sleeping) slp[i] = b->sleeping[i] ? 1 : 0;
            if (b->still_frames) still[i] = b->still_frames[i];
            if (smode[i] != AVBD_SLEEP_NEVER || slp[i]) wi.hasSleep = true;
            pos[i] = make_float4(b->position[3 * i], b->position[3 * i + 1], b->position[3 * i + 2], 0);
            rot[i] = make_float4(b->rotation[4 * i], b->rotation[4 * i + 1], b->rotation[4 * i + 2], b->rotation[4 * i + 3]);
            vel[i] = make_float4(b->velocity[3 * i], b->velocity[3 * i + 1], b->velocity[3 * i + 2], 0);
            ang[i] = make_float4(b->angular_velocity[3 * i], b->angular_velocity[3 * i + 1], b->angular_velocity[3 * i + 2], 0);
            prev[i] = b->prev_velocity ? make_float4(b->prev_velocity[3 * i], b->prev_velocity[3 * i + 1], b->prev_velocity[3 * i + 2], 0)
                                       : vel[i];
            sizeR[i] = make_float4(b->size[3 * i], b->size[3 * i + 1], b->size[3 * i + 2], b->radius[i]);
            massI[i] = make_float4(b->mass[i], b->moment[3 * i], b->moment[3 * i + 1], b->moment[3 * i + 2]);
            fric[i] = b->friction[i];
            grav[i] = b->gravity ? b->gravity[i] : prm.gravity;
            layer[i] = b->layer ? b->layer[i] : 1u;
            mask[i] = b->mask ? b->mask[i] : 1u;
            ll[i] = b->lock_linear ? b->lock_linear[i] : 0;
            la[i] = b->lock_angular ? b->lock_angular[i] : 0;
            if (b->mass[i] > 0.0f) ui[i] = movable++;
        }
        const int o = wi.bodyBase;
        auto put = [&](auto *dst, const auto &src) {
            const size_t bytes = src.size() * sizeof(src[0]);
            data.push_back({dst + o, arenaPut(src.data(), bytes), bytes});
            return hipSuccess;
        };
        auto idn = [&](auto *dst, const auto &src) {   // identity arrays: copied on activation
            const size_t bytes = src.size() * sizeof(src[0]);
            act.activate.push_back({dst, arenaPut(src.data(), bytes), bytes});
        };
        HIPCHK(put(S.pos, pos)); HIPCHK(put(S.rot, rot)); HIPCHK(put(S.vel, vel)); HIPCHK(put(S.angVel, ang));
        HIPCHK(put(S.prevVel, prev)); HIPCHK(put(S.initLin, pos)); HIPCHK(put(S.initAng, rot));
        HIPCHK(put(S.inertLin, pos)); HIPCHK(put(S.inertAng, rot));
        HIPCHK(put(S.sizeR, sizeR)); HIPCHK(put(S.massI, massI)); HIPCHK(put(S.fric, fric)); HIPCHK(put(S.grav, grav));
        HIPCHK(put(S.layer, layer)); HIPCHK(put(S.mask, mask)); HIPCHK(put(S.lockLin, ll)); HIPCHK(put(S.lockAng, la));
        HIPCHK(put(S.shape, shp));
        HIPCHK(put(S.sleepMode, smode)); HIPCHK(put(S.sleeping, slp)); HIPCHK(put(S.stillFrames, still));
        idn(S.bodyWorld + o, bw);
        idn(S.updateIndex + o, ui);
        if (nj > 0) {
            std::vector<int> ja(nj), jb(nj), jw(nj, w);
            std::vector<float4> ra(nj), rb(nj), c0l(nj), c0a(nj), pl(nj), pa(nj), ll2(nj), la2(nj);
            std::vector<float> sl(nj), sa(nj), fr(nj), arm(nj), rest(nj, 0.0f);
            std::vector<uint8_t> br(nj), dead(nj, 0), kind(nj, AVBD_FORCE_JOINT), gmode(6 * nj, 0);
            std::vector<float4> grest(nj, make_float4(0, 0, 0, 1));
            std::vector<float> glo(6 * nj, 0.0f), ghi(6 * nj, 0.0f), gk(6 * nj, 0.0f), geq(6 * nj, 0.0f),
                    gd(6 * nj, 0.0f), glam(6 * nj, 0.0f), gpen(6 * nj, 0.0f), ginit(6 * nj, 0.0f);
            bool anyGeneric = false;
            auto f4 = [](const float *p, int k) { return p ? make_float4(p[3 * k], p[3 * k + 1], p[3 * k + 2], 0) : make_float4(0, 0, 0, 0); };
            for (int k = 0; k < nj; k++) {
                if (j->body_b[k] < 0 || j->body_b[k] >= n || j->body_a[k] >= n)
                    return fail(AVBD_ERR_INVALID, "joint body index out of range");
                ja[k] = j->body_a[k] >= 0 ? j->body_a[k] + o : -1;
                jb[k] = j->body_b[k] + o;
                ra[k] = f4(j->r_a, k); rb[k] = f4(j->r_b, k);
                c0l[k] = f4(j->c0_lin, k); c0a[k] = f4(j->c0_ang, k);
                pl[k] = f4(j->penalty_lin, k); pa[k] = f4(j->penalty_ang, k);
                ll2[k] = f4(j->lambda_lin, k); la2[k] = f4(j->lambda_ang, k);
                sl[k] = j->stiffness_lin[k]; sa[k] = j->stiffness_ang[k];
                fr[k] = j->fracture[k]; arm[k] = j->torque_arm[k];
                br[k] = j->broken ? j->broken[k] : 0;
                if (j->kind) kind[k] = j->kind[k];
                if (kind[k] > AVBD_FORCE_GENERIC) return fail(AVBD_ERR_INVALID, "unknown force kind");
                if (kind[k] == AVBD_FORCE_GENERIC) {
                    if (!j->rest_rotation || !j->axis_mode || !j->axis_lower || !j->axis_upper || !j->axis_spring ||
                        !j->axis_equilibrium || !j->axis_damping)
                        return fail(AVBD_ERR_INVALID, "generic joint arrays are null");
                    anyGeneric = true;
                    grest[k] = make_float4(j->rest_rotation[4 * k], j->rest_rotation[4 * k + 1],
                                           j->rest_rotation[4 * k + 2], j->rest_rotation[4 * k + 3]);
                    for (int a = 0; a < 6; a++) {
                        const int x = 6 * k + a;
                        if (j->axis_mode[x] > AVBD_AXIS_LIMITED) return fail(AVBD_ERR_INVALID, "unknown axis mode");
                        gmode[x] = j->axis_mode[x];
                        glo[x] = j->axis_lower[x]; ghi[x] = j->axis_upper[x]; gk[x] = j->axis_spring[x];
                        geq[x] = j->axis_equilibrium[x]; gd[x] = j->axis_damping[x];
                        glam[x] = j->axis_lambda ? j->axis_lambda[x] : 0.0f;
                        gpen[x] = j->axis_penalty ? j->axis_penalty[x] : 0.0f;
                        ginit[x] = j->axis_initial ? j->axis_initial[x] : 0.0f;
                    }
                }
                if (kind[k] == AVBD_FORCE_SPRING) {
                    // spring.cpp:19-28 reads both bodies' poses
                    if (j->body_a[k] < 0) return fail(AVBD_ERR_INVALID, "a spring needs two bodies");
                    if (!j->rest) return fail(AVBD_ERR_INVALID, "spring rest lengths are null");
                    rest[k] = j->rest[k];
                }
                if (kind[k] != AVBD_FORCE_JOINT) br[k] = 0;   // only joints break
            }
            const int jo = wi.jointBase;
            auto putj = [&](auto *dst, const auto &src) {
                const size_t bytes = src.size() * sizeof(src[0]);
                data.push_back({dst + jo, arenaPut(src.data(), bytes), bytes});
                return hipSuccess;
            };
            idn(S.jWorld + jo, jw);
            HIPCHK(putj(S.jA, ja)); HIPCHK(putj(S.jB, jb)); HIPCHK(putj(S.jRA, ra));
            HIPCHK(putj(S.jRB, rb)); HIPCHK(putj(S.jC0L, c0l)); HIPCHK(putj(S.jC0A, c0a)); HIPCHK(putj(S.jPenL, pl));
            HIPCHK(putj(S.jPenA, pa)); HIPCHK(putj(S.jLamL, ll2)); HIPCHK(putj(S.jLamA, la2));
            HIPCHK(putj(S.jStiffL, sl)); HIPCHK(putj(S.jStiffA, sa)); HIPCHK(putj(S.jFrac, fr)); HIPCHK(putj(S.jArm, arm));
            HIPCHK(putj(S.jBroken, br)); HIPCHK(putj(S.jDead, dead)); HIPCHK(putj(S.jKind, kind));
            HIPCHK(putj(S.jRest, rest));
            if (anyGeneric) {
                HIPCHK(putj(S.jGRest, grest));
                auto put6 = [&](auto *dst, const auto &src) {
                    const size_t bytes = src.size() * sizeof(src[0]);
                    data.push_back({dst + 6 * jo, arenaPut(src.data(), bytes), bytes});
                    return hipSuccess;
                };
                HIPCHK(put6(S.jGMode, gmode)); HIPCHK(put6(S.jGLo, glo)); HIPCHK(put6(S.jGHi, ghi));
                HIPCHK(put6(S.jGK, gk)); HIPCHK(put6(S.jGEq, geq)); HIPCHK(put6(S.jGDamp, gd));
                HIPCHK(put6(S.jGLam, glam)); HIPCHK(put6(S.jGPen, gpen)); HIPCHK(put6(S.jGInit, ginit));
            }
        }
        // slot range ownership and empty manifold store, the world table
        idn(S.slotWorld + wi.slotBase, std::vector<int>(slotCap, w));
        const std::vector<int> tab{wi.bodyBase, n, wi.jointBase, nj, wi.slotBase, slotCap, 0, 1};
        const int *tp = static_cast<const int *>(arenaPut(tab.data(), sizeof(int) * tab.size()));
        int *dst[6] = {S.wBodyBase, S.wBodyCount, S.wJointBase, S.wJointCount, S.wSlotBase, S.wSlotCap};
        for (int k = 0; k < 6; k++) act.activate.push_back({dst[k] + w, tp + k, sizeof(int)});
        int *zs[6] = {S.wLive, S.wFree, S.wTop, S.wNew, S.wErr, S.wStale};
        for (int k = 0; k < 6; k++) act.activate.push_back({zs[k] + w, tp + 6, sizeof(int)});
        act.activate.push_back({S.wRewalk + w, tp + 7, sizeof(int)});
        if (arenaUsed > arenaCap) return fail(AVBD_ERR_DEVICE, "upload staging overflow");
        for (const auto &c : data) HIPCHK(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, cs));
        pending.push_back(std::move(act));
        wi.pending = true;
        if (async) {
            HIPCHK(hipEventRecord(upEvent, upStream));
            upRecorded = true;
            return AVBD_OK;
        }
        avbd_status st = activatePending();
        if (st != AVBD_OK) return st;
        HIPCHK(hipStreamSynchronize(stream));
        arenaQuiesce();
        return AVBD_OK;
    }

    avbd_status writeBodies(int w, const avbd_bodies *b)
    {
        const WorldInfo &wi = worlds[w];
        if (b->count != wi.bodyCount) return fail(AVBD_ERR_INVALID, "body count differs from the uploaded world");
        const int n = wi.bodyCount, o = wi.bodyBase;
        std::vector<float4> tmp(n);
        auto put = [&](float4 *dst, const float *src, int comps) -> hipError_t {
            if (!src) return hipSuccess;
            hipError_t e = hipMemcpy(tmp.data(), dst + o, sizeof(float4) * n, hipMemcpyDeviceToHost);
            if (e != hipSuccess) return e;
            for (int i = 0; i < n; i++) {
                float v[4] = {tmp[i].x, tmp[i].y, tmp[i].z, tmp[i].w};
                for (int c = 0; c < comps; c++) v[c] = src[comps * i + c];
                tmp[i] = make_float4(v[0], v[1], v[2], v[3]);
            }
            return hipMemcpy(dst + o, tmp.data(), sizeof(float4) * n, hipMemcpyHostToDevice);
        };
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(put(S.pos, b->position, 3));
        HIPCHK(put(S.rot, b->rotation, 4));
        HIPCHK(put(S.vel, b->velocity, 3));
        HIPCHK(put(S.angVel, b->angular_velocity, 3));
        HIPCHK(put(S.prevVel, b->prev_velocity, 3));
        if (b->sleeping) {
            std::vector<uint8_t> slp(n);
            for (int i = 0; i < n; i++) {
                slp[i] = b->sleeping[i] ? 1 : 0;
                if (slp[i]) worlds[w].hasSleep = true;
            }
            HIPCHK(hipMemcpy(S.sleeping + o, slp.data(), n, hipMemcpyHostToDevice));
        }
        if (b->still_frames) HIPCHK(hipMemcpy(S.stillFrames + o, b->still_frames, sizeof(int) * n, hipMemcpyHostToDevice));
        return AVBD_OK;
    }

    avbd_status clearWorld(int w)
    {
        WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        fill(S.bodyWorld + wi.bodyBase, -1, wi.bodyCount);
        fill(S.updateIndex + wi.bodyBase, -1, wi.bodyCount);
        fill(S.jWorld + wi.jointBase, -1, wi.jointCount);
        fill(S.slotWorld + wi.slotBase, -1, wi.slotCap);
        const int zero = 0;
        int *zs[8] = {S.wBodyCount, S.wJointCount, S.wSlotCap, S.wLive, S.wFree, S.wTop, S.wNew, S.wErr};
        for (int k = 0; k < 8; k++) HIPCHK(hipMemcpyAsync(zs[k] + w, &zero, sizeof(int), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        bodyAlloc.give(wi.bodyBase, wi.bodyCount);
        jointAlloc.give(wi.jointBase, wi.jointCount);
        slotAlloc.give(wi.slotBase, wi.slotCap);
        wi = WorldInfo{};
        csrValid = false;
        treeAge = 1 << 20;
        return refreshWorldOrder();
    }

    avbd_status downloadWorld(int w, avbd_bodies *b, avbd_joints *j)
    {
        const WorldInfo &wi = worlds[w];
        if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
        const int n = wi.bodyCount, o = wi.bodyBase;
        HIPCHK(hipStreamSynchronize(stream));
        if (b) {
            if (b->count < n) return fail(AVBD_ERR_CAPACITY, "body arrays are too small");
            b->count = n;
            std::vector<float4> tmp(n);
            auto get = [&](const float4 *src, float *dst, int comps) -> hipError_t {
                if (!dst) return hipSuccess;
                hipError_t e = hipMemcpy(tmp.data(), src + o, sizeof(float4) * n, hipMemcpyDeviceToHost);
                if (e != hipSuccess) return e;
                for (int i = 0; i < n; i++) {
                    const float v[4] = {tmp[i].x, tmp[i].y, tmp[i].z, tmp[i].w};
                    for (int c = 0; c < comps; c++) dst[comps * i + c] = v[c];
                }
                return hipSuccess;
            };
            HIPCHK(get(S.pos, b->position, 3));
            HIPCHK(get(S.rot, b->rotation, 4));
            HIPCHK(get(S.vel, b->velocity, 3));
            HIPCHK(get(S.angVel, b->angular_velocity, 3));
            HIPCHK(get(S.prevVel, b->prev_velocity, 3));
            if (b->sleeping) HIPCHK(hipMemcpy(b->sleeping, S.sleeping + o, n, hipMemcpyDeviceToHost));
            if (b->still_frames) HIPCHK(hipMemcpy(b->still_frames, S.stillFrames + o, sizeof(int) * n, hipMemcpyDeviceToHost));
        }
        if (j) {
            const int nj = wi.jointCount, jo = wi.jointBase;
            std::vector<uint8_t> dead(std::max(nj, 1)), broken(std::max(nj, 1));
            std::vector<int> ja(std::max(nj, 1)), jb(std::max(nj, 1));
            std::vector<float4> c0l(nj), c0a(nj), pl(nj), pa(nj), ll(nj), la(nj);
            std::vector<float> glam(6 * nj), gpen(6 * nj), ginit(6 * nj);
            const bool wantGeneric = j->axis_lambda || j->axis_penalty || j->axis_initial;
            if (nj > 0 && wantGeneric) {
                HIPCHK(hipMemcpy(glam.data(), S.jGLam + 6 * jo, sizeof(float) * 6 * nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(gpen.data(), S.jGPen + 6 * jo, sizeof(float) * 6 * nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(ginit.data(), S.jGInit + 6 * jo, sizeof(float) * 6 * nj, hipMemcpyDeviceToHost));
            }
            if (nj > 0) {
                HIPCHK(hipMemcpy(dead.data(), S.jDead + jo, nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(broken.data(), S.jBroken + jo, nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(ja.data(), S.jA + jo, sizeof(int) * nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(jb.data(), S.jB + jo, sizeof(int) * nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(c0l.data(), S.jC0L + jo, sizeof(float4) * nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(c0a.data(), S.jC0A + jo, sizeof(float4) * nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(pl.data(), S.jPenL + jo, sizeof(float4) * nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(pa.data(), S.jPenA + jo, sizeof(float4) * nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(ll.data(), S.jLamL + jo, sizeof(float4) * nj, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(la.data(), S.jLamA + jo, sizeof(float4) * nj, hipMemcpyDeviceToHost));
            }
            // a dead joint was reaped by the reference: compact the list as it does
            int k = 0;
            auto put3 = [](float *dst, int i, float4 v) { if (dst) { dst[3 * i] = v.x; dst[3 * i + 1] = v.y; dst[3 * i + 2] = v.z; } };
            for (int q = 0; q < nj; q++) {
                if (dead[q]) continue;
                if (k >= j->count) return fail(AVBD_ERR_CAPACITY, "joint arrays are too small");
                if (j->body_a) j->body_a[k] = ja[q] >= 0 ? ja[q] - o : -1;
                if (j->body_b) j->body_b[k] = jb[q] - o;
                put3(j->c0_lin, k, c0l[q]); put3(j->c0_ang, k, c0a[q]);
                put3(j->penalty_lin, k, pl[q]); put3(j->penalty_ang, k, pa[q]);
                put3(j->lambda_lin, k, ll[q]); put3(j->lambda_ang, k, la[q]);
                if (j->broken) j->broken[k] = broken[q];
                for (int a = 0; a < 6; a++) {
                    if (j->axis_lambda) j->axis_lambda[6 * k + a] = glam[6 * q + a];
                    if (j->axis_penalty) j->axis_penalty[6 * k + a] = gpen[6 * q + a];
                    if (j->axis_initial) j->axis_initial[6 * k + a] = ginit[6 * q + a];
                }
                k++;
            }
            j->count = k;
        }
        return AVBD_OK;
    }

    avbd_status getManifolds(int w, avbd_manifolds *m)
    {
        const WorldInfo &wi = worlds[w];
        HIPCHK(hipStreamSynchronize(stream));
        int nLive = 0;
        HIPCHK(hipMemcpy(&nLive, S.wLive + w, sizeof(int), hipMemcpyDeviceToHost));
        if (m->count < nLive) {
            m->count = nLive;
            return fail(AVBD_ERR_CAPACITY, "manifold arrays are too small");
        }
        m->count = nLive;
        if (nLive == 0) return AVBD_OK;
        buildLiveList();
        k_contact_unpack<<<liveGrid(), TPB, 0, stream>>>(S);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
        std::vector<int> live(nLive);
        HIPCHK(hipMemcpy(live.data(), S.live + wi.slotBase, sizeof(int) * nLive, hipMemcpyDeviceToHost));
        const int cap = wi.slotCap, sb = wi.slotBase;
        std::vector<int> mA(cap), mB(cap), num(cap);
        std::vector<float4> rec(4 * cap), cra(cap * MAXC), crb(cap * MAXC), cc0(cap * MAXC), cp(cap * MAXC), cl(cap * MAXC);
        HIPCHK(hipMemcpy(mA.data(), S.mA + sb, sizeof(int) * cap, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(mB.data(), S.mB + sb, sizeof(int) * cap, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(num.data(), S.mNum + sb, sizeof(int) * cap, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(rec.data(), S.mRec + 4 * static_cast<size_t>(sb), sizeof(float4) * 4 * cap, hipMemcpyDeviceToHost));
        const size_t cb = static_cast<size_t>(sb) * MAXC, cn = static_cast<size_t>(cap) * MAXC;
        HIPCHK(hipMemcpy(cra.data(), S.sRA + cb, sizeof(float4) * cn, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(crb.data(), S.sRB + cb, sizeof(float4) * cn, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(cc0.data(), S.sC0 + cb, sizeof(float4) * cn, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(cp.data(), S.sPen + cb, sizeof(float4) * cn, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(cl.data(), S.sLam + cb, sizeof(float4) * cn, hipMemcpyDeviceToHost));
        const int o = wi.bodyBase;
        auto put3 = [](float *dst, int i, float4 v) { if (dst) { dst[3 * i] = v.x; dst[3 * i + 1] = v.y; dst[3 * i + 2] = v.z; } };
        for (int i = 0; i < nLive; i++) {
            const int s = live[i] - sb;
            if (m->body_a) m->body_a[i] = mA[s] - o;
            if (m->body_b) m->body_b[i] = mB[s] - o;
            if (m->basis)
                for (int r = 0; r < 3; r++) {
                    const float4 v = rec[4 * s + r];
                    m->basis[9 * i + 3 * r] = v.x; m->basis[9 * i + 3 * r + 1] = v.y; m->basis[9 * i + 3 * r + 2] = v.z;
                }
            if (m->friction) m->friction[i] = rec[4 * s + 3].x;
            if (m->num_contacts) m->num_contacts[i] = num[s];
            for (int k = 0; k < MAXC; k++) {
                const int src = s * MAXC + k, dst = i * MAXC + k;
                const bool used = k < num[s];
                const float4 z = make_float4(0, 0, 0, 0);
                int key = 0;
                std::memcpy(&key, &crb[src].w, sizeof(int));
                if (m->feature) m->feature[dst] = used ? key : 0;
                put3(m->r_a, dst, used ? cra[src] : z);
                put3(m->r_b, dst, used ? crb[src] : z);
                put3(m->c0, dst, used ? cc0[src] : z);
                put3(m->penalty, dst, used ? cp[src] : z);
                put3(m->lambda, dst, used ? cl[src] : z);
                if (m->stick) m->stick[dst] = used && cc0[src].w != 0.0f ? 1 : 0;
                if (m->beta_scale) m->beta_scale[dst] = used ? cra[src].w : 1.0f;
            }
        }
        return AVBD_OK;
    }

    avbd_status getResiduals(int w, int32_t *numManifolds, float *cn, int32_t *numJoints, float *cl)
    {
        if (!S.resC) return fail(AVBD_ERR_INVALID, "residuals are not being kept (avbd_world_record_residuals)");
        const WorldInfo &wi = worlds[w];
        HIPCHK(hipStreamSynchronize(stream));
        int nLive = 0;
        HIPCHK(hipMemcpy(&nLive, S.wLive + w, sizeof(int), hipMemcpyDeviceToHost));
        const int nj = wi.jointCount, jo = wi.jointBase;
        std::vector<uint8_t> dead(std::max(nj, 1));
        if (nj > 0) HIPCHK(hipMemcpy(dead.data(), S.jDead + jo, nj, hipMemcpyDeviceToHost));
        int alive = 0;
        for (int q = 0; q < nj; q++) alive += dead[q] ? 0 : 1;
        const bool fits = *numManifolds >= nLive && *numJoints >= alive;
        *numManifolds = nLive;
        *numJoints = alive;
        if (!fits) return fail(AVBD_ERR_CAPACITY, "residual arrays are too small");
        if (nLive > 0 && cn) {
            std::vector<int> live(nLive), num(wi.slotCap);
            std::vector<float> res(static_cast<size_t>(wi.slotCap) * MAXC);
            HIPCHK(hipMemcpy(live.data(), S.live + wi.slotBase, sizeof(int) * nLive, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(num.data(), S.mNum + wi.slotBase, sizeof(int) * wi.slotCap, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(res.data(), S.resC + static_cast<size_t>(wi.slotBase) * MAXC, sizeof(float) * res.size(),
                             hipMemcpyDeviceToHost));
            for (int i = 0; i < nLive; i++) {
                const int s = live[i] - wi.slotBase;
                for (int k = 0; k < MAXC; k++) cn[MAXC * i + k] = k < num[s] ? res[static_cast<size_t>(s) * MAXC + k] : 0.0f;
            }
        }
        if (nj > 0 && cl) {
            std::vector<float4> res(nj);
            HIPCHK(hipMemcpy(res.data(), S.resJ + jo, sizeof(float4) * nj, hipMemcpyDeviceToHost));
            int k = 0;
            for (int q = 0; q < nj; q++) {
                if (dead[q]) continue;
                cl[3 * k] = res[q].x; cl[3 * k + 1] = res[q].y; cl[3 * k + 2] = res[q].z;
                k++;
            }
        }
        return AVBD_OK;
    }

    avbd_status setManifolds(int w, const avbd_manifolds *m)
    {
        const WorldInfo &wi = worlds[w];
        const int n = m->count;
        if (n > wi.slotCap) return fail(AVBD_ERR_CAPACITY, "more manifolds than the world's slot capacity");
        const int sb = wi.slotBase, o = wi.bodyBase;
        fill(S.mDormant + sb, 0, wi.slotCap);   // restored manifolds are not parked
        fill(S.mSaved + sb, 0, wi.slotCap);
        if (S.armCap) fill(reinterpret_cast<int *>(S.armCap + 2 * static_cast<size_t>(sb)), 0, 8 * wi.slotCap);   // no captured arms
        std::vector<int> live(n), mA(n), mB(n), num(n);
        std::vector<float4> rec(4 * n), cra(n * MAXC), crb(n * MAXC), cc0(n * MAXC), cp(n * MAXC), cl(n * MAXC);
        for (int i = 0; i < n; i++) {
            live[i] = sb + i;
            if (m->body_a[i] < 0 || m->body_a[i] >= wi.bodyCount || m->body_b[i] < 0 || m->body_b[i] >= wi.bodyCount)
                return fail(AVBD_ERR_INVALID, "manifold body index out of range");
            mA[i] = m->body_a[i] + o;
            mB[i] = m->body_b[i] + o;
            num[i] = m->num_contacts[i];
            if (num[i] < 0 || num[i] > MAXC) return fail(AVBD_ERR_INVALID, "manifold contact count out of range");
            rec[4 * i + 3] = make_float4(m->friction[i], intBits(num[i]), 0.0f, 0.0f);
            for (int r = 0; r < 3; r++)
                rec[4 * i + r] = make_float4(m->basis[9 * i + 3 * r], m->basis[9 * i + 3 * r + 1], m->basis[9 * i + 3 * r + 2],
                                             intBits(r == 0 ? mA[i] : r == 1 ? mB[i] : 0));
            for (int k = 0; k < MAXC; k++) {
                const int q = i * MAXC + k;
                float keyf;
                std::memcpy(&keyf, &m->feature[q], sizeof(float));
                cra[q] = make_float4(m->r_a[3 * q], m->r_a[3 * q + 1], m->r_a[3 * q + 2], m->beta_scale ? m->beta_scale[q] : 1.0f);
                crb[q] = make_float4(m->r_b[3 * q], m->r_b[3 * q + 1], m->r_b[3 * q + 2], keyf);
                cc0[q] = make_float4(m->c0[3 * q], m->c0[3 * q + 1], m->c0[3 * q + 2], m->stick[q] ? 1.0f : 0.0f);
                cp[q] = make_float4(m->penalty[3 * q], m->penalty[3 * q + 1], m->penalty[3 * q + 2], 0);
                cl[q] = make_float4(m->lambda[3 * q], m->lambda[3 * q + 1], m->lambda[3 * q + 2], 0);
            }
        }
        // every other live manifold's contacts go to staging, this world's are written there,
        // and the dense layout is rebuilt from staging
        HIPCHK(hipStreamSynchronize(stream));
        buildLiveList();
        k_contact_unpack<<<liveGrid(), TPB, 0, stream>>>(S);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
        if (n > 0) {
            HIPCHK(hipMemcpy(S.live + sb, live.data(), sizeof(int) * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(S.mA + sb, mA.data(), sizeof(int) * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(S.mB + sb, mB.data(), sizeof(int) * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(S.mNum + sb, num.data(), sizeof(int) * n, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(S.mRec + 4 * static_cast<size_t>(sb), rec.data(), sizeof(float4) * 4 * n, hipMemcpyHostToDevice));
            const size_t cb = static_cast<size_t>(sb) * MAXC, cn = static_cast<size_t>(n) * MAXC;
            HIPCHK(hipMemcpy(S.sRA + cb, cra.data(), sizeof(float4) * cn, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(S.sRB + cb, crb.data(), sizeof(float4) * cn, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(S.sC0 + cb, cc0.data(), sizeof(float4) * cn, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(S.sPen + cb, cp.data(), sizeof(float4) * cn, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(S.sLam + cb, cl.data(), sizeof(float4) * cn, hipMemcpyHostToDevice));
        }
        const int vals[3] = {n, 0, n};
        int *dst[3] = {S.wLive, S.wFree, S.wTop};
        for (int k = 0; k < 3; k++) HIPCHK(hipMemcpy(dst[k] + w, &vals[k], sizeof(int), hipMemcpyHostToDevice));
        buildLiveList();
        packContacts(1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
        csrValid = false;
        return AVBD_OK;
    }

    avbd_status probeBroadphase(uint64_t *codes, int32_t *order, uint64_t *pairs, int32_t *pairCount)
    {
        const WorldInfo &wi = worlds[0];
        setParams(S.P, prm, prm.dt);
        if (!csrValid) {
            refreshView();
            buildAdjacency();
        }
        csrValid = true;
        const int nb = nblk(Ncap);
        k_fill_i32<<<nblk(Wcap), TPB, 0, stream>>>(S.wRewalk, 1, Wcap);
        k_fill_i32<<<nblk(Wcap), TPB, 0, stream>>>(S.wStale, 0, Wcap);
        k_swept_aabb<<<nb, TPB, 0, stream>>>(S);
        k_world_bounds<<<Wcap, TPB, 0, stream>>>(S);
        widePick();
        k_morton<<<nb, TPB, 0, stream>>>(S, 1, 1);
        radixSort();
        k_karras<<<nblk(Ncap - 1), TPB, 0, stream>>>(S, Ncap);
        k_world_roots_init<<<nblk(Wcap), TPB, 0, stream>>>(S, Ncap);
        k_world_roots<<<nblk(Ncap - 1), TPB, 0, stream>>>(S, Ncap);
        k_world_roots_check<<<nblk(Wcap), TPB, 0, stream>>>(S);
        k_ropes<<<nblk(2 * Ncap - 1), TPB, 0, stream>>>(S, Ncap);
        buildRefitTasks();
        k_refit_tasks<<<refitGrid(), REFIT_TPB, 0, stream>>>(listTree(), Ncap, 0);
        wideFlags();
        k_list_walk<<<nb, TPB, 0, stream>>>(S, Ncap);
        k_traverse<<<nb, TPB, 0, stream>>>(S, Ncap);
        k_fill_i32<<<1, 64, 0, stream>>>(S.listStats, 0, 4);
        k_list_world_end<<<nblk(Wcap), TPB, 0, stream>>>(S);
        scan(S.pairCount, S.pairOff, Ncap);
        k_pair_emit<<<nb, TPB, 0, stream>>>(S, Ncap);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
        treeAge = 1 << 20;
        const int n = wi.bodyCount;
        if (codes) {
            std::vector<unsigned long long> c(n);
            HIPCHK(hipMemcpy(c.data(), S.mortonCode, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost));
            for (int i = 0; i < n; i++) codes[i] = c[i];
        }
        if (order) {
            // the world occupies positions [0, n) in sorted order (bodies outside it sort later)
            std::vector<int> v(n);
            HIPCHK(hipMemcpy(v.data(), S.sortVal, sizeof(int) * n, hipMemcpyDeviceToHost));
            for (int i = 0; i < n; i++) order[i] = v[i];
        }
        uint32_t total = 0;
        HIPCHK(hipMemcpy(&total, S.pairOff + Ncap, sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (pairCount) {
            const int cap = *pairCount;
            *pairCount = static_cast<int32_t>(total);
            if (pairs && total > 0) {
                if (static_cast<int>(total) > cap) return fail(AVBD_ERR_CAPACITY, "pair array too small");
                if (total > static_cast<uint32_t>(Pcap)) return fail(AVBD_ERR_CAPACITY, "candidate pair capacity exceeded");
                std::vector<int> pi(total), pj(total);
                HIPCHK(hipMemcpy(pi.data(), S.pairI, sizeof(int) * total, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(pj.data(), S.pairJ, sizeof(int) * total, hipMemcpyDeviceToHost));
                for (uint32_t k = 0; k < total; k++)
                    pairs[k] = (static_cast<uint64_t>(static_cast<uint32_t>(pi[k])) << 32) | static_cast<uint32_t>(pj[k]);
            }
        }
        return AVBD_OK;
    }

    avbd_status totals(int64_t *movable, int64_t *manifolds, int64_t *contacts, int64_t *joints)
    {
        HIPCHK(hipStreamSynchronize(stream));
        std::vector<int> bw(Ncap), ui(Ncap), live(Wcap), slotWorld(Scap), liveList(Scap), num(Scap), jw(Jcap);
        std::vector<uint8_t> dead(Jcap);
        HIPCHK(hipMemcpy(ui.data(), S.updateIndex, sizeof(int) * Ncap, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(bw.data(), S.bodyWorld, sizeof(int) * Ncap, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(live.data(), S.wLive, sizeof(int) * Wcap, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(liveList.data(), S.live, sizeof(int) * Scap, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(num.data(), S.mNum, sizeof(int) * Scap, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(jw.data(), S.jWorld, sizeof(int) * Jcap, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(dead.data(), S.jDead, Jcap, hipMemcpyDeviceToHost));
        int64_t nb = 0, nm = 0, nc = 0, nj = 0;
        for (int i = 0; i < Ncap; i++) nb += (bw[i] >= 0 && ui[i] >= 0) ? 1 : 0;
        for (int w = 0; w < Wcap; w++) {
            if (!worlds[w].active) continue;
            nm += live[w];
            for (int k = 0; k < live[w]; k++) nc += num[liveList[worlds[w].slotBase + k]];
        }
        for (int j = 0; j < Jcap; j++) nj += (jw[j] >= 0 && !dead[j]) ? 1 : 0;
        if (movable) *movable = nb;
        if (manifolds) *manifolds = nm;
        if (contacts) *contacts = nc;
        if (joints) *joints = nj;
        return AVBD_OK;
    }

    // Pairs of the world's last step that the pair filter found apart: the reference makes and
    // releases a manifold for each, so they count as new manifolds.
    avbd_status apartPairs(const WorldInfo &wi, int &n)
    {
        n = 0;
        uint32_t r[2] = {0, 0};
        HIPCHK(hipMemcpy(&r[0], S.pairOff + wi.bodyBase, sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&r[1], S.pairOff + wi.bodyBase + wi.bodyCount, sizeof(uint32_t), hipMemcpyDeviceToHost));
        const uint32_t a = std::min<uint32_t>(r[0], static_cast<uint32_t>(Pcap));
        const uint32_t b = std::min<uint32_t>(r[1], static_cast<uint32_t>(Pcap));
        if (b <= a) return AVBD_OK;
        std::vector<uint8_t> f(b - a);
        HIPCHK(hipMemcpy(f.data(), S.pairApart + a, b - a, hipMemcpyDeviceToHost));
        for (uint8_t x : f) n += x;
        return AVBD_OK;
    }

    avbd_status fillStats(int w, avbd_step_stats *out)
    {
        HIPCHK(hipStreamSynchronize(stream));
        avbd_step_stats s = stats;
        const WorldInfo &wi = worlds[w];
        s.bodies = wi.bodyCount;
        s.joints = wi.jointCount;
        int v[3] = {0, 0, 0};
        HIPCHK(hipMemcpy(&v[0], S.wLive + w, sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&v[1], S.wNew + w, sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&v[2], S.brakeLevels + w, sizeof(int), hipMemcpyDeviceToHost));
        int apart = 0;
        if (avbd_status st = apartPairs(wi, apart); st != AVBD_OK) return st;
        s.manifolds = v[0];
        s.new_manifolds = v[1] + apart;
        s.brake_levels = v[2];
        if (v[0] > 0) {
            std::vector<int> live(v[0]);
            HIPCHK(hipMemcpy(live.data(), S.live + wi.slotBase, sizeof(int) * v[0], hipMemcpyDeviceToHost));
            std::vector<int> num(wi.slotCap);
            HIPCHK(hipMemcpy(num.data(), S.mNum + wi.slotBase, sizeof(int) * wi.slotCap, hipMemcpyDeviceToHost));
            int c = 0;
            for (int s2 : live) c += num[s2 - wi.slotBase];
            s.contacts = c;
        }
        uint32_t a = 0, b = 0;
        HIPCHK(hipMemcpy(&a, S.pairOff + wi.bodyBase, sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&b, S.pairOff + wi.bodyBase + wi.bodyCount, sizeof(uint32_t), hipMemcpyDeviceToHost));
        s.candidate_pairs = static_cast<int32_t>(b - a);
        std::vector<uint8_t> col(wi.bodyCount);
        HIPCHK(hipMemcpy(col.data(), S.colour + wi.bodyBase, wi.bodyCount, hipMemcpyDeviceToHost));
        std::vector<int> ui(wi.bodyCount);
        HIPCHK(hipMemcpy(ui.data(), S.updateIndex + wi.bodyBase, sizeof(int) * wi.bodyCount, hipMemcpyDeviceToHost));
        int maxc = -1, ov = 0;
        for (int i = 0; i < wi.bodyCount; i++) {
            if (ui[i] < 0) continue;
            if (col[i] < 8) maxc = std::max<int>(maxc, col[i]);
            else if (col[i] == 8) ov++;
        }
        s.colours = maxc + 1;
        s.overflow_bodies = ov;
        *out = s;
        return AVBD_OK;
    }
};

// ---------------------------------------------------------------------- C ABI ----
struct avbd_world {
    Engine eng;
};
struct avbd_batch {
    Engine eng;
};

extern "C" {

void avbd_params_default(avbd_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->dt = 1.0f / 60.0f;
    p->gravity = -10.0f;
    p->iterations = 8;
    p->alpha = 0.99f;
    p->beta_lin = 10000.0f;
    p->beta_ang = 100.0f;
    p->gamma = 0.999f;
    p->newton_ratio = 1.0f;
    p->stiffness_decay = 1.0f;
    p->adaptive_beta = 0;
    p->adaptive_beta_error_floor = 1e-3f;
    p->convergence_threshold = 0.0f;
    p->jacobi_relaxation = 0.7f;
    p->contact_only_overflow = 1;
    p->stick_drift_heal = 0.002f;
    p->max_penetration = 0.02f;
    p->jacobian_rebuild_distance = 0.0f;
    p->speculative_margin = 0.01f;
    p->position_projection = 1;
    p->projection_iterations = 4;
    p->projection_beta = 1.0f;
    p->projection_slop = 0.008f;
    p->projection_max_correction = 0.0f;
    p->contact_grace_ticks = 0;
    p->threads = 0;
    p->sleep_linear_threshold = 0.1f;
    p->sleep_angular_threshold = 0.139626f;
    p->sleep_frames = 0;
    p->penalty_max = 1e10f;
}

void avbd_box_properties(const float size[3], float density, float *mass, float moment[3], float *radius)
{
    const float m = size[0] * size[1] * size[2] * density;
    if (mass) *mass = m;
    if (moment) {
        moment[0] = (size[1] * size[1] + size[2] * size[2]) / 12.0f * m;
        moment[1] = (size[0] * size[0] + size[2] * size[2]) / 12.0f * m;
        moment[2] = (size[0] * size[0] + size[1] * size[1]) / 12.0f * m;
    }
    if (radius) {
        const float hx = size[0] * 0.5f, hy = size[1] * 0.5f, hz = size[2] * 0.5f;
        *radius = std::sqrt(((0.0f + hx * hx) + hy * hy) + hz * hz);
    }
}

int avbd_shape_properties(int shape, const float size[3], float density, float *mass, float moment[3], float *radius)
{
    float m, mo[3], rad;
    const float r = size[0];
    if (shape == AVBD_SHAPE_SPHERE) {
        m = (4.0f / 3.0f) * 3.14159265358979f * r * r * r * density;
        const float i = 0.4f * m * r * r;
        mo[0] = mo[1] = mo[2] = i;
        rad = r;
    } else if (shape == AVBD_SHAPE_CYLINDER) {
        const float h = size[2];
        m = 3.14159265358979f * r * r * h * density;
        const float axial = 0.5f * m * r * r;
        const float radial = m * (3.0f * r * r + h * h) / 12.0f;
        mo[0] = mo[1] = radial; mo[2] = axial;
        const float hh = h * 0.5f;
        rad = std::sqrt(((0.0f + r * r) + hh * hh) + r * r);
    } else if (shape == AVBD_SHAPE_BOX) {
        avbd_box_properties(size, density, &m, mo, &rad);
    } else {
        return fail(AVBD_ERR_INVALID, "unknown shape");
    }
    if (mass) *mass = m;
    if (moment) { moment[0] = mo[0]; moment[1] = mo[1]; moment[2] = mo[2]; }
    if (radius) *radius = rad;
    return AVBD_OK;
}

uint64_t avbd_state_digest(const avbd_bodies *b)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; i++) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    if (!b) return h;
    for (int i = 0; i < b->count; i++) {
        mix(b->position + 3 * i, 12);
        mix(b->rotation + 4 * i, 16);
        mix(b->velocity + 3 * i, 12);
        mix(b->angular_velocity + 3 * i, 12);
    }
    return h;
}

const char *avbd_last_error(void) { return g_lastError.c_str(); }

int32_t avbd_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

static avbd_status validBodies(const avbd_bodies *b)
{
    if (!b) return fail(AVBD_ERR_INVALID, "bodies is null");
    if (b->count < 0) return fail(AVBD_ERR_INVALID, "negative body count");
    if (b->count > 0 && (!b->size || !b->mass || !b->moment || !b->friction || !b->radius || !b->position ||
                         !b->rotation || !b->velocity || !b->angular_velocity))
        return fail(AVBD_ERR_INVALID, "a required body array is null");
    return AVBD_OK;
}
static avbd_status validJoints(const avbd_joints *j)
{
    if (!j || j->count == 0) return AVBD_OK;
    if (j->count < 0) return fail(AVBD_ERR_INVALID, "negative joint count");
    if (!j->body_a || !j->body_b || !j->r_a || !j->r_b || !j->stiffness_lin || !j->stiffness_ang || !j->fracture ||
        !j->torque_arm)
        return fail(AVBD_ERR_INVALID, "a required joint array is null");
    return AVBD_OK;
}

avbd_status avbd_world_create(const avbd_params *params, int32_t max_bodies, int32_t max_joints,
                              int32_t max_manifolds, avbd_world **out)
{
    if (!out) return fail(AVBD_ERR_INVALID, "out is null");
    *out = nullptr;
    avbd_status st = checkParams(params);
    if (st != AVBD_OK) return st;
    if (max_bodies < 0 || max_joints < 0) return fail(AVBD_ERR_INVALID, "negative capacity");
    auto *w = new avbd_world();
    const int slots = max_manifolds > 0 ? max_manifolds : 8 * std::max(max_bodies, 1) + 64;
    st = w->eng.create(*params, 1, max_bodies, max_joints, slots, false);
    if (st != AVBD_OK) {
        delete w;
        return st;
    }
    *out = w;
    return AVBD_OK;
}

avbd_status avbd_world_destroy(avbd_world *world)
{
    delete world;
    return AVBD_OK;
}

avbd_status avbd_world_set_params(avbd_world *world, const avbd_params *params)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    avbd_status st = checkParams(params);
    if (st != AVBD_OK) return st;
    if ((st = world->eng.ensureArmCache(*params)) != AVBD_OK) return st;
    world->eng.prm = *params;
    return AVBD_OK;
}

avbd_status avbd_world_upload(avbd_world *world, const avbd_bodies *bodies, const avbd_joints *joints)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    avbd_status st = validBodies(bodies);
    if (st != AVBD_OK) return st;
    if ((st = validJoints(joints)) != AVBD_OK) return st;
    Engine &e = world->eng;
    if (bodies->count > e.Ncap) return fail(AVBD_ERR_CAPACITY, "more bodies than the world's capacity");
    if (joints && joints->count > e.Jcap) return fail(AVBD_ERR_CAPACITY, "more joints than the world's capacity");
    if (e.worlds[0].active) {
        st = e.clearWorld(0);
        if (st != AVBD_OK) return st;
    }
    WorldInfo &wi = e.worlds[0];
    wi.bodyBase = e.bodyAlloc.take(bodies->count);
    wi.jointBase = e.jointAlloc.take(joints ? joints->count : 0);
    wi.slotBase = e.slotAlloc.take(e.Scap);
    return e.uploadWorld(0, bodies, joints, e.Scap);
}

avbd_status avbd_world_write_bodies(avbd_world *world, const avbd_bodies *bodies)
{
    if (!world || !bodies) return fail(AVBD_ERR_INVALID, "null argument");
    if (!world->eng.worlds[0].active) return fail(AVBD_ERR_INVALID, "nothing uploaded");
    return world->eng.writeBodies(0, bodies);
}

avbd_status avbd_world_step(avbd_world *world)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    if (!world->eng.worlds[0].active) return fail(AVBD_ERR_INVALID, "nothing uploaded");
    return world->eng.stepPlain(1, 0);
}

avbd_status avbd_world_step_interleaved(avbd_world *world, int32_t substeps, int32_t rounds)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    if (!world->eng.worlds[0].active) return fail(AVBD_ERR_INVALID, "nothing uploaded");
    return world->eng.stepPlain(substeps <= 1 ? 1 : substeps, rounds);
}

avbd_status avbd_world_download(avbd_world *world, avbd_bodies *bodies, avbd_joints *joints)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.downloadWorld(0, bodies, joints);
}

avbd_status avbd_world_get_manifolds(avbd_world *world, avbd_manifolds *out)
{
    if (!world || !out) return fail(AVBD_ERR_INVALID, "null argument");
    return world->eng.getManifolds(0, out);
}

avbd_status avbd_world_record_residuals(avbd_world *world, int32_t on)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.recordResiduals(on != 0);
}

avbd_status avbd_world_get_residuals(avbd_world *world, int32_t *num_manifolds, float *contact_cn,
                                     int32_t *num_joints, float *joint_clin)
{
    if (!world || !num_manifolds || !num_joints) return fail(AVBD_ERR_INVALID, "null argument");
    return world->eng.getResiduals(0, num_manifolds, contact_cn, num_joints, joint_clin);
}

avbd_status avbd_world_set_manifolds(avbd_world *world, const avbd_manifolds *in)
{
    if (!world || !in) return fail(AVBD_ERR_INVALID, "null argument");
    if (in->count > 0 && (!in->body_a || !in->body_b || !in->basis || !in->friction || !in->num_contacts ||
                          !in->feature || !in->r_a || !in->r_b || !in->c0 || !in->penalty || !in->lambda || !in->stick))
        return fail(AVBD_ERR_INVALID, "a required manifold array is null");
    return world->eng.setManifolds(0, in);
}

avbd_status avbd_world_stats(avbd_world *world, avbd_step_stats *out)
{
    if (!world || !out) return fail(AVBD_ERR_INVALID, "null argument");
    return world->eng.fillStats(0, out);
}

avbd_status avbd_world_colours(avbd_world *world, uint8_t *out)
{
    if (!world || !out) return fail(AVBD_ERR_INVALID, "null argument");
    Engine &e = world->eng;
    if (!e.worlds[0].active) return fail(AVBD_ERR_INVALID, "nothing uploaded");
    const WorldInfo &wi = e.worlds[0];
    std::vector<uint8_t> col(wi.bodyCount);
    std::vector<int> ui(wi.bodyCount);
    HIPCHK(hipStreamSynchronize(e.stream));
    HIPCHK(hipMemcpy(col.data(), e.S.colour + wi.bodyBase, wi.bodyCount, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ui.data(), e.S.updateIndex + wi.bodyBase, sizeof(int) * wi.bodyCount, hipMemcpyDeviceToHost));
    for (int i = 0; i < wi.bodyCount; i++) out[i] = ui[i] < 0 ? 255 : col[i];
    return AVBD_OK;
}

avbd_status avbd_world_profile(avbd_world *world, int32_t steps, avbd_kernel_profile *out)
{
    if (!world || !out) return fail(AVBD_ERR_INVALID, "null argument");
    if (!world->eng.worlds[0].active) return fail(AVBD_ERR_INVALID, "nothing uploaded");
    return world->eng.profile(steps, 1, 0, out);
}

avbd_status avbd_batch_profile(avbd_batch *batch, int32_t steps, avbd_kernel_profile *out)
{
    if (!batch || !out) return fail(AVBD_ERR_INVALID, "null argument");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    if (!batch->eng.csrValid) {
        batch->eng.buildAdjacency();
        batch->eng.csrValid = true;
    }
    return batch->eng.profile(steps, 1, 0, out);
}

avbd_status avbd_world_probe_broadphase(avbd_world *world, uint64_t *codes, int32_t *order, uint64_t *pairs,
                                        int32_t *pair_count)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    if (!world->eng.worlds[0].active) return fail(AVBD_ERR_INVALID, "nothing uploaded");
    return world->eng.probeBroadphase(codes, order, pairs, pair_count);
}

avbd_status avbd_batch_create(const avbd_params *params, int32_t max_worlds, int32_t max_bodies, int32_t max_joints,
                              int32_t max_manifolds, avbd_batch **out)
{
    if (!out) return fail(AVBD_ERR_INVALID, "out is null");
    *out = nullptr;
    avbd_status st = checkParams(params);
    if (st != AVBD_OK) return st;
    if (max_worlds < 1 || max_bodies < 0 || max_joints < 0) return fail(AVBD_ERR_INVALID, "bad capacity");
    auto *b = new avbd_batch();
    const int slots = max_manifolds > 0 ? max_manifolds : 8 * std::max(max_bodies, 1) + 64 * max_worlds;
    st = b->eng.create(*params, max_worlds, max_bodies, max_joints, slots, true);
    if (st != AVBD_OK) {
        delete b;
        return st;
    }
    *out = b;
    return AVBD_OK;
}

avbd_status avbd_batch_autotune(avbd_batch *batch, int32_t steps_per_trial, avbd_tuning *out)
{
    if (!batch) return fail(AVBD_ERR_INVALID, "null argument");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    return batch->eng.autotune(steps_per_trial, out);
}

avbd_status avbd_batch_set_tuning(avbd_batch *batch, const avbd_tuning *tuning)
{
    if (!batch || !tuning) return fail(AVBD_ERR_INVALID, "null argument");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    return batch->eng.setTuning(*tuning);
}

avbd_status avbd_world_autotune(avbd_world *world, int32_t steps_per_trial, avbd_tuning *out)
{
    if (!world) return fail(AVBD_ERR_INVALID, "null argument");
    if (!world->eng.worlds[0].active) return fail(AVBD_ERR_INVALID, "nothing uploaded");
    return world->eng.autotune(steps_per_trial, out);
}

avbd_status avbd_batch_footprint(const avbd_params *params, int32_t max_worlds, int32_t max_bodies,
                                 int32_t max_joints, int32_t max_manifolds, int64_t *bytes)
{
    if (!bytes) return fail(AVBD_ERR_INVALID, "bytes is null");
    avbd_status st = checkParams(params);
    if (st != AVBD_OK) return st;
    if (max_worlds < 1 || max_bodies < 0 || max_joints < 0) return fail(AVBD_ERR_INVALID, "bad capacity");
    Engine e;
    e.sizing = true;
    const int slots = max_manifolds > 0 ? max_manifolds : 8 * std::max(max_bodies, 1) + 64 * max_worlds;
    st = e.create(*params, max_worlds, max_bodies, max_joints, slots, true);
    if (st != AVBD_OK) return st;
    *bytes = static_cast<int64_t>(e.deviceBytes);
    return AVBD_OK;
}

avbd_status avbd_batch_device_bytes(avbd_batch *batch, int64_t *bytes)
{
    if (!batch || !bytes) return fail(AVBD_ERR_INVALID, "null argument");
    *bytes = static_cast<int64_t>(batch->eng.deviceBytes);
    return AVBD_OK;
}

avbd_status avbd_batch_destroy(avbd_batch *batch)
{
    if (batch && batch->eng.stepInFlight) (void)batch->eng.stepBatchWait();
    delete batch;
    return AVBD_OK;
}

// Manifold slots a world of nb bodies is given in a batch.
static int worldSlots(int nb) { return std::max(8 * nb + 16, 32); }

avbd_status avbd_batch_fits(avbd_batch *batch, int32_t n_bodies, int32_t n_joints, int32_t *fits_now,
                            int32_t *fits_ever)
{
    if (!batch || n_bodies < 0 || n_joints < 0) return fail(AVBD_ERR_INVALID, "bad argument");
    const Engine &e = batch->eng;
    const int slots = worldSlots(n_bodies);
    bool freeWorld = false;
    for (int k = 0; k < e.Wcap; k++) freeWorld = freeWorld || (!e.worlds[k].active && !e.worlds[k].pending);
    if (fits_ever) *fits_ever = n_bodies <= e.Ncap && n_joints <= e.Jcap && slots <= e.Scap;
    if (fits_now)
        *fits_now = freeWorld && e.bodyAlloc.fits(n_bodies) && e.jointAlloc.fits(n_joints) && e.slotAlloc.fits(slots);
    return AVBD_OK;
}

static avbd_status addWorld(avbd_batch *batch, const avbd_bodies *bodies, const avbd_joints *joints,
                            int32_t *world_id, bool async)
{
    avbd_status st = validBodies(bodies);
    if (st != AVBD_OK) return st;
    if ((st = validJoints(joints)) != AVBD_OK) return st;
    Engine &e = batch->eng;
    int w = -1;
    for (int k = 0; k < e.Wcap; k++)
        if (!e.worlds[k].active && !e.worlds[k].pending) { w = k; break; }
    if (w < 0) return fail(AVBD_ERR_CAPACITY, "no free world slot");
    const int nb = bodies->count, nj = joints ? joints->count : 0;
    const int slots = worldSlots(nb);
    const int bb = e.bodyAlloc.take(nb);
    if (bb < 0) return fail(AVBD_ERR_CAPACITY, "not enough body capacity");
    const int jb = e.jointAlloc.take(nj);
    if (jb < 0) { e.bodyAlloc.give(bb, nb); return fail(AVBD_ERR_CAPACITY, "not enough joint capacity"); }
    const int sb = e.slotAlloc.take(slots);
    if (sb < 0) {
        e.bodyAlloc.give(bb, nb);
        e.jointAlloc.give(jb, nj);
        return fail(AVBD_ERR_CAPACITY, "not enough manifold capacity");
    }
    WorldInfo &wi = e.worlds[w];
    wi.bodyBase = bb;
    wi.jointBase = jb;
    wi.slotBase = sb;
    st = e.uploadWorld(w, bodies, joints, slots, async);
    if (st != AVBD_OK) {   // rejected input: the world stays inactive and its ranges go back
        e.bodyAlloc.give(bb, nb);
        e.jointAlloc.give(jb, nj);
        e.slotAlloc.give(sb, slots);
        return st;
    }
    *world_id = w;
    return AVBD_OK;
}

avbd_status avbd_batch_add_world(avbd_batch *batch, const avbd_bodies *bodies, const avbd_joints *joints,
                                 int32_t *world_id)
{
    if (!batch || !world_id) return fail(AVBD_ERR_INVALID, "null argument");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    return addWorld(batch, bodies, joints, world_id, false);
}

avbd_status avbd_batch_add_world_async(avbd_batch *batch, const avbd_bodies *bodies, const avbd_joints *joints,
                                       int32_t *world_id)
{
    if (!batch || !world_id) return fail(AVBD_ERR_INVALID, "null argument");
    return addWorld(batch, bodies, joints, world_id, true);
}

avbd_status avbd_batch_step_launch(avbd_batch *batch)
{
    if (!batch) return fail(AVBD_ERR_INVALID, "batch is null");
    return batch->eng.stepBatchLaunch();
}

avbd_status avbd_batch_step_wait(avbd_batch *batch)
{
    if (!batch) return fail(AVBD_ERR_INVALID, "batch is null");
    return batch->eng.stepBatchWait();
}

avbd_status avbd_batch_remove_world(avbd_batch *batch, int32_t world_id)
{
    if (!batch) return fail(AVBD_ERR_INVALID, "batch is null");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    if (world_id < 0 || world_id >= batch->eng.Wcap) return fail(AVBD_ERR_INVALID, "bad world id");
    return batch->eng.clearWorld(world_id);
}

avbd_status avbd_world_use_graph(avbd_world *world, int32_t enable)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    world->eng.useGraph = enable != 0;
    return AVBD_OK;
}

avbd_status avbd_batch_use_graph(avbd_batch *batch, int32_t enable)
{
    if (!batch) return fail(AVBD_ERR_INVALID, "batch is null");
    batch->eng.useGraph = enable != 0;
    return AVBD_OK;
}

avbd_status avbd_world_use_query_tree(avbd_world *world, int32_t enable)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    world->eng.queryTreeOn = enable != 0;
    return AVBD_OK;
}

avbd_status avbd_batch_use_query_tree(avbd_batch *batch, int32_t enable)
{
    if (!batch) return fail(AVBD_ERR_INVALID, "batch is null");
    batch->eng.queryTreeOn = enable != 0;
    return AVBD_OK;
}

avbd_status avbd_batch_step(avbd_batch *batch)
{
    if (!batch) return fail(AVBD_ERR_INVALID, "batch is null");
    return batch->eng.stepBatch();
}

avbd_status avbd_batch_failed_worlds(avbd_batch *batch, int32_t capacity, int32_t *world_ids, int32_t *reasons,
                                     int32_t *count)
{
    if (!batch || !count || capacity < 0 || (capacity > 0 && !world_ids)) return fail(AVBD_ERR_INVALID, "null argument");
    const Engine &e = batch->eng;
    const int n = static_cast<int>(e.failedIds.size());
    for (int k = 0; k < n && k < capacity; k++) {
        world_ids[k] = e.failedIds[k];
        if (reasons) reasons[k] = e.failedWhy[k];
    }
    *count = n;
    return AVBD_OK;
}

avbd_status avbd_world_raycast(avbd_world *world, int32_t n, const float *from, const float *to,
                               const uint32_t *masks, int32_t hit_from_inside, avbd_ray_hits *out)
{
    if (!world || !out) return fail(AVBD_ERR_INVALID, "null argument");
    return world->eng.raycast(0, n, from, to, masks, hit_from_inside != 0, out->body, out->t, out->point, out->normal,
                              out->inside);
}

avbd_status avbd_world_collide_shape(avbd_world *world, int32_t n, const int32_t *type, const float *center,
                                     const float *rotation, const float *size, const float *margin,
                                     const uint32_t *masks, int32_t max_pairs, int32_t *count, uint8_t *any,
                                     float *points)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.collideShape(0, n, type, center, rotation, size, margin, masks, max_pairs, count, any, points);
}

avbd_status avbd_world_rest_info(avbd_world *world, int32_t n, const int32_t *type, const float *center,
                                 const float *rotation, const float *size, const float *margin, const uint32_t *masks,
                                 int32_t *body, float *point, float *normal, float *depth)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.restInfo(0, n, type, center, rotation, size, margin, masks, body, point, normal, depth);
}

avbd_status avbd_world_closest_point(avbd_world *world, int32_t n, const int32_t *body, const float *points,
                                     float *out)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.closestPointQuery(0, n, body, points, out);
}

avbd_status avbd_world_collide_pair(avbd_world *world, int32_t body_a, int32_t body_b, float collision_dt,
                                    int32_t *count, int32_t *features, float *r_a, float *r_b, float *basis)
{
    if (!world || !count) return fail(AVBD_ERR_INVALID, "null argument");
    return world->eng.collidePair(body_a, body_b, collision_dt, count, features, r_a, r_b, basis);
}

// ---- avbd_world_raycast without allocation ------------------------------------------------
// All device memory the call needs beyond the world's own state lives in `scratch`, one
// fixed array per device (this library's code object is loaded once per device, each with its
// own copy at its own address).  Layout, in 256-byte aligned parts: the caller's segments and
// masks as copied (from, to, masks), the packed segments, the bodies' bounds, the kernel's
// hits, and the hits unpacked into the caller's layout (t, point, normal, inside).
namespace {
constexpr size_t NA_RAYS = AVBD_NOALLOC_MAX_RAYS, NA_BODIES = AVBD_NOALLOC_MAX_BODIES;
constexpr size_t naAlign(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }
constexpr size_t NA_FROM = 0;
constexpr size_t NA_TO = NA_FROM + naAlign(12 * NA_RAYS);
constexpr size_t NA_MASK = NA_TO + naAlign(12 * NA_RAYS);
constexpr size_t NA_RAYS4 = NA_MASK + naAlign(4 * NA_RAYS);
constexpr size_t NA_BOUNDS = NA_RAYS4 + naAlign(32 * NA_RAYS);
constexpr size_t NA_BODY = NA_BOUNDS + naAlign(32 * NA_BODIES);
constexpr size_t NA_HITA = NA_BODY + naAlign(4 * NA_RAYS);
constexpr size_t NA_HITB = NA_HITA + naAlign(16 * NA_RAYS);
constexpr size_t NA_T = NA_HITB + naAlign(16 * NA_RAYS);
constexpr size_t NA_POINT = NA_T + naAlign(4 * NA_RAYS);
constexpr size_t NA_NORMAL = NA_POINT + naAlign(12 * NA_RAYS);
constexpr size_t NA_INSIDE = NA_NORMAL + naAlign(12 * NA_RAYS);
constexpr size_t NA_BYTES = NA_INSIDE + naAlign(NA_RAYS);
static_assert(NA_BYTES == AVBD_NOALLOC_SCRATCH_BYTES, "scratch size in avbd_hip.h is out of date");
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
static unsigned char *g_scratchPtr[NA_MAX_DEVICES];
static std::once_flag g_scratchOnce[NA_MAX_DEVICES];
static hipError_t g_scratchErr[NA_MAX_DEVICES];

// The device address of `scratch` on `dev`, resolved once per device with that device current
// (host code never uses the array's own name: that is a host-side stand-in).
static avbd_status scratchOn(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(AVBD_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::call_once(g_scratchOnce[dev], [dev] {
        int cur = 0;
        g_scratchErr[dev] = hipGetDevice(&cur);
        if (g_scratchErr[dev] != hipSuccess) return;
        if (cur != dev && (g_scratchErr[dev] = hipSetDevice(dev)) != hipSuccess) return;
        void *p = nullptr;
        g_scratchErr[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratchPtr[dev] = static_cast<unsigned char *>(p);
        if (cur != dev) (void)hipSetDevice(cur);
    });
    if (g_scratchErr[dev] != hipSuccess || !g_scratchPtr[dev])
        return fail(AVBD_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_scratchErr[dev]));
    *out = g_scratchPtr[dev];
    return AVBD_OK;
}

__global__ void k_na_pack_rays(int n, const float *from, const float *to, const uint32_t *masks, int haveMasks,
                               float4 *rays)
{
    const int r = GID;
    if (r >= n) return;
    const uint32_t m = haveMasks ? masks[r] : 0xFFFFFFFFu;
    rays[2 * r] = make_float4(from[3 * r], from[3 * r + 1], from[3 * r + 2], __uint_as_float(m));
    rays[2 * r + 1] = make_float4(to[3 * r], to[3 * r + 1], to[3 * r + 2], 0.0f);
}

__global__ void k_na_unpack_hits(int n, const float4 *a, const float4 *b, float *t, float *point, float *normal,
                                 uint8_t *inside)
{
    const int r = GID;
    if (r >= n) return;
    const float4 x = a[r], y = b[r];
    t[r] = x.w;
    point[3 * r] = x.x; point[3 * r + 1] = x.y; point[3 * r + 2] = x.z;
    normal[3 * r] = y.x; normal[3 * r + 1] = y.y; normal[3 * r + 2] = y.z;
    inside[r] = y.w != 0.0f ? 1 : 0;
}

avbd_status avbd_world_raycast_noalloc(avbd_world *world, int32_t n, const float *from, const float *to,
                                       const uint32_t *masks, int32_t hit_from_inside, avbd_ray_hits *out)
{
    if (!world || !out) return fail(AVBD_ERR_INVALID, "null argument");
    Engine &e = world->eng;
    const WorldInfo &wi = e.worlds[0];
    if (!wi.active) return fail(AVBD_ERR_INVALID, "world is not active");
    if (n < 0 || (n > 0 && (!from || !to || !out->body))) return fail(AVBD_ERR_INVALID, "bad ray arrays");
    if (n > AVBD_NOALLOC_MAX_RAYS)
        return fail(AVBD_ERR_CAPACITY, "more than AVBD_NOALLOC_MAX_RAYS segments in one no-allocation call");
    if (wi.bodyCount > AVBD_NOALLOC_MAX_BODIES)
        return fail(AVBD_ERR_CAPACITY, "world has more than AVBD_NOALLOC_MAX_BODIES bodies for a no-allocation call");
    if (n == 0) return AVBD_OK;
    unsigned char *base = nullptr;
    avbd_status st = scratchOn(e.device, &base);
    if (st != AVBD_OK) return st;
    auto *dFrom = reinterpret_cast<float *>(base + NA_FROM);
    auto *dTo = reinterpret_cast<float *>(base + NA_TO);
    auto *dMask = reinterpret_cast<uint32_t *>(base + NA_MASK);
    auto *rays = reinterpret_cast<float4 *>(base + NA_RAYS4);
    auto *bounds = reinterpret_cast<float4 *>(base + NA_BOUNDS);
    auto *body = reinterpret_cast<int *>(base + NA_BODY);
    auto *hitA = reinterpret_cast<float4 *>(base + NA_HITA);
    auto *hitB = reinterpret_cast<float4 *>(base + NA_HITB);
    auto *t = reinterpret_cast<float *>(base + NA_T);
    auto *point = reinterpret_cast<float *>(base + NA_POINT);
    auto *normal = reinterpret_cast<float *>(base + NA_NORMAL);
    auto *inside = reinterpret_cast<uint8_t *>(base + NA_INSIDE);
    hipStream_t s = e.stream;
    HIPCHK(hipMemcpyAsync(dFrom, from, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dTo, to, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
    if (masks) HIPCHK(hipMemcpyAsync(dMask, masks, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    k_na_pack_rays<<<nblk(n), TPB, 0, s>>>(n, dFrom, dTo, dMask, masks ? 1 : 0, rays);
    if (wi.bodyCount > 0) k_query_bounds<<<nblk(wi.bodyCount), TPB, 0, s>>>(e.S, wi.bodyBase, wi.bodyCount, bounds);
    // no query tree here (its refit needs buffers this call may not allocate): the list scan,
    // whose results are the tree walk's
    k_raycast<<<nblk(n), TPB, 0, s>>>(e.S, wi.bodyBase, wi.bodyCount, n, rays, bounds, hit_from_inside ? 1 : 0, body,
                                      hitA, hitB, QTree{nullptr, e.Ncap, 0});
    k_na_unpack_hits<<<nblk(n), TPB, 0, s>>>(n, hitA, hitB, t, point, normal, inside);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out->body, body, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    if (out->t) HIPCHK(hipMemcpyAsync(out->t, t, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    if (out->point) HIPCHK(hipMemcpyAsync(out->point, point, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
    if (out->normal) HIPCHK(hipMemcpyAsync(out->normal, normal, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
    if (out->inside) HIPCHK(hipMemcpyAsync(out->inside, inside, n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return AVBD_OK;
}

avbd_status avbd_batch_raycast(avbd_batch *batch, int32_t world_id, int32_t n, const float *from, const float *to,
                               const uint32_t *masks, int32_t hit_from_inside, avbd_ray_hits *out)
{
    if (!batch || !out) return fail(AVBD_ERR_INVALID, "null argument");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    if (world_id < 0 || world_id >= batch->eng.Wcap) return fail(AVBD_ERR_INVALID, "bad world id");
    return batch->eng.raycast(world_id, n, from, to, masks, hit_from_inside != 0, out->body, out->t, out->point,
                              out->normal, out->inside);
}

avbd_status avbd_world_profile_raycast(avbd_world *world, int32_t n, const float *from, const float *to, int32_t reps,
                                       avbd_query_profile *out)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.profileRaycast(0, n, from, to, reps, out);
}

avbd_status avbd_world_intersect_point(avbd_world *world, int32_t n, const float *points, const uint32_t *masks,
                                       int32_t max_results, int32_t *count, int32_t *out)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.intersectPoint(0, n, points, masks, max_results, count, out);
}

avbd_status avbd_world_intersect_shape(avbd_world *world, int32_t n, const int32_t *type, const float *center,
                                       const float *rotation, const float *size, const float *margin,
                                       const uint32_t *masks, int32_t max_results, int32_t *count, int32_t *out)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.intersectShape(0, n, type, center, rotation, size, margin, masks, max_results, count, out);
}

avbd_status avbd_world_cast_motion(avbd_world *world, int32_t n, const int32_t *type, const float *center,
                                   const float *rotation, const float *size, const float *motion, const float *margin,
                                   const uint32_t *masks, int32_t *body, float *fraction, float *point, float *normal)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.castMotion(0, n, type, center, rotation, size, motion, margin, masks, body, fraction, point, normal);
}

avbd_status avbd_world_contact_report(avbd_world *world, const int32_t *max_reported, int32_t stride, int32_t *count,
                                      float *position, float *normal, int32_t *other)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.contactReport(0, max_reported, stride, count, position, normal, other);
}

avbd_status avbd_batch_contact_report(avbd_batch *batch, int32_t world_id, const int32_t *max_reported, int32_t stride,
                                      int32_t *count, float *position, float *normal, int32_t *other)
{
    if (!batch) return fail(AVBD_ERR_INVALID, "batch is null");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    if (world_id < 0 || world_id >= batch->eng.Wcap) return fail(AVBD_ERR_INVALID, "bad world id");
    return batch->eng.contactReport(world_id, max_reported, stride, count, position, normal, other);
}

avbd_status avbd_world_pick(avbd_world *world, int32_t n, const float *origin, const float *dir,
                            const uint32_t *masks, int32_t *body, float *local)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.pick(0, n, origin, dir, masks, body, local);
}

avbd_status avbd_world_collect_aabb(avbd_world *world, const float min[3], const float max[3], uint32_t mask,
                                    int32_t *out, int32_t capacity, int32_t *count)
{
    if (!world) return fail(AVBD_ERR_INVALID, "world is null");
    return world->eng.collectAabb(0, min, max, mask, out, capacity, count);
}

avbd_status avbd_batch_download_world(avbd_batch *batch, int32_t world_id, avbd_bodies *bodies, avbd_joints *joints)
{
    if (!batch) return fail(AVBD_ERR_INVALID, "batch is null");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    if (world_id < 0 || world_id >= batch->eng.Wcap) return fail(AVBD_ERR_INVALID, "bad world id");
    return batch->eng.downloadWorld(world_id, bodies, joints);
}

avbd_status avbd_batch_world_count(avbd_batch *batch, int32_t *count)
{
    if (!batch || !count) return fail(AVBD_ERR_INVALID, "null argument");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    int c = 0;
    for (const auto &w : batch->eng.worlds) c += w.active ? 1 : 0;
    *count = c;
    return AVBD_OK;
}

avbd_status avbd_batch_get_manifolds(avbd_batch *batch, int32_t world_id, avbd_manifolds *out)
{
    if (!batch || !out) return fail(AVBD_ERR_INVALID, "null argument");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    if (world_id < 0 || world_id >= batch->eng.Wcap || !batch->eng.worlds[world_id].active)
        return fail(AVBD_ERR_INVALID, "bad world id");
    return batch->eng.getManifolds(world_id, out);
}

avbd_status avbd_batch_set_manifolds(avbd_batch *batch, int32_t world_id, const avbd_manifolds *in)
{
    if (!batch || !in) return fail(AVBD_ERR_INVALID, "null argument");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    if (world_id < 0 || world_id >= batch->eng.Wcap || !batch->eng.worlds[world_id].active)
        return fail(AVBD_ERR_INVALID, "bad world id");
    if (in->count > 0 && (!in->body_a || !in->body_b || !in->basis || !in->friction || !in->num_contacts ||
                          !in->feature || !in->r_a || !in->r_b || !in->c0 || !in->penalty || !in->lambda || !in->stick))
        return fail(AVBD_ERR_INVALID, "a required manifold array is null");
    return batch->eng.setManifolds(world_id, in);
}

avbd_status avbd_batch_totals(avbd_batch *batch, int64_t *movable_bodies, int64_t *manifolds, int64_t *contacts,
                              int64_t *joints)
{
    if (!batch) return fail(AVBD_ERR_INVALID, "batch is null");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    return batch->eng.totals(movable_bodies, manifolds, contacts, joints);
}

avbd_status avbd_batch_list_stats(avbd_batch *batch, int32_t out[4])
{
    if (!batch || !out) return fail(AVBD_ERR_INVALID, "null argument");
    if (avbd_status busy = batch->eng.idle(); busy != AVBD_OK) return busy;
    return batch->eng.listStats(out);
}

avbd_status avbd_world_list_stats(avbd_world *world, int32_t out[4])
{
    if (!world || !out) return fail(AVBD_ERR_INVALID, "null argument");
    return world->eng.listStats(out);
}

} // extern "C"

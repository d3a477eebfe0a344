// Kernels of the AVBD step.  One thread per body, per manifold slot, per joint or per
// candidate pair, with per-world workgroups for the two passes the reference runs in
// a data-dependent order (greedy colouring, velocity brake).  Grids are sized by the
// batch capacities, so a launch sequence never depends on data and can be captured in a
// HIP graph; per-world counts on the device say which threads have work.
#pragma once

#include "dsolver.hpp"

namespace avbd_dev {

constexpr int TPB = 256;
__host__ __device__ inline int nblk(int n, int tpb = TPB) { return (n + tpb - 1) / tpb; }
// Workgroups of a grid-stride launch over the live manifold list (8 per CU on 256 CUs).
constexpr int LIVE_BLOCKS = 2048;

#define GID (static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x))

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8), and each
// XCD has its own L2.  xcdBlock() renumbers workgroups so that every XCD gets one contiguous
// run of logical blocks: consecutive bodies / manifolds, which share bodies, then meet in one
// L2.  It is a bijection on [0, gridDim.x) for any grid size.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcdBlock()
{
    const int nb = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
    const int x = b % NUM_XCD, j = b / NUM_XCD;
    const int q = nb / NUM_XCD, r = nb % NUM_XCD;
    return x * q + (x < r ? x : r) + j;
}
#define XGID (xcdBlock() * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x))

// ------------------------------------------------------------------ utilities ----
__global__ void k_fill_i32(int *p, int v, int n)
{
    const int i = GID;
    if (i < n) p[i] = v;
}
// A phase boundary of a timed step: the device's constant-rate clock at this point of the stream.
__global__ void k_phase_mark(uint64_t *ticks, int k)
{
    if (threadIdx.x == 0) ticks[k] = wall_clock64();
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int n)
{
    const int i = GID;
    if (i < n) p[i] = v;
}

// Exclusive scan of n u32 values, SCAN_TILE per block; out has n+1 entries (out[n] = total).
// With `dlen`, only the first min(n, *dlen) values are scanned (a length known on the device
// only): later tiles do nothing and out[min(n, *dlen)] = out[n] = total.
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = TPB * SCAN_ITEMS;
constexpr int SCAN_SUMS_TPB = 1024;

// Block-wide exclusive scan: wavefront shuffles, then one wavefront scans the wavefront sums.
// `sm` holds NT / 64 entries.
template <int NT>
__device__ __forceinline__ uint32_t blockExclusiveScan(uint32_t v, uint32_t *sm, uint32_t &total)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sm[wave] = x;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < NT / 64 ? sm[lane] : 0u;
#pragma unroll
        for (int off = 1; off < NT / 64; off <<= 1) {
            const uint32_t y = __shfl_up(w, off, 64);
            if (lane >= off) w += y;
        }
        if (lane < NT / 64) sm[lane] = w;
    }
    __syncthreads();
    const uint32_t below = wave > 0 ? sm[wave - 1] : 0u;
    total = sm[NT / 64 - 1];
    __syncthreads();
    return below + x - v;
}

__device__ __forceinline__ int scanLength(int n, const uint32_t *dlen)
{
    if (!dlen) return n;
    const uint32_t d = *dlen;
    return d < static_cast<uint32_t>(n) ? static_cast<int>(d) : n;
}

__global__ void __launch_bounds__(TPB) k_scan_tiles(const uint32_t *in, uint32_t *out, uint32_t *tileSums, int n,
                                                    const uint32_t *dlen)
{
    __shared__ uint32_t sm[TPB / 64];
    n = scanLength(n, dlen);
    if (static_cast<int>(blockIdx.x) * SCAN_TILE >= n) {
        if (threadIdx.x == 0) tileSums[blockIdx.x] = 0;
        return;
    }
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = blockExclusiveScan<TPB>(s, sm, total);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) tileSums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_SUMS_TPB) k_scan_sums(uint32_t *tileSums, int nt, uint32_t *out, int n,
                                                            const uint32_t *dlen)
{
    __shared__ uint32_t sm[SCAN_SUMS_TPB / 64];
    uint32_t carry = 0;
    for (int start = 0; start < nt; start += SCAN_SUMS_TPB) {
        const int i = start + threadIdx.x;
        const uint32_t v = i < nt ? tileSums[i] : 0u;
        uint32_t total;
        const uint32_t ex = blockExclusiveScan<SCAN_SUMS_TPB>(v, sm, total);
        if (i < nt) tileSums[i] = ex + carry;
        carry += total;
    }
    if (threadIdx.x == 0) {
        out[scanLength(n, dlen)] = carry;
        out[n] = carry;
    }
}

__global__ void __launch_bounds__(TPB) k_scan_add(uint32_t *out, const uint32_t *tileSums, int n, const uint32_t *dlen)
{
    n = scanLength(n, dlen);
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    if (static_cast<int>(blockIdx.x) * SCAN_TILE >= n) return;
    const uint32_t add = tileSums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
        if (base + k < n) out[base + k] += add;
}

// LSD radix sort pass over 8 key bits: per-tile digit histograms, then a stable scatter
// that ranks equal digits within each wavefront with ballots.
constexpr int RADIX_ITEMS = 8;
constexpr int RADIX_TILE = TPB * RADIX_ITEMS;

__global__ void k_radix_hist(const unsigned long long *keys, int n, int shift, uint32_t *hist, int nb)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int start = blockIdx.x * RADIX_TILE;
    for (int r = 0; r < RADIX_ITEMS; r++) {
        const int i = start + r * TPB + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ void k_radix_scatter(const unsigned long long *kin, const int *vin, unsigned long long *kout, int *vout,
                                int n, int shift, const uint32_t *offs, int nb)
{
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[TPB / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    base[tid] = offs[tid * nb + blockIdx.x];
#pragma unroll
    for (int w = 0; w < TPB / 64; w++) wcnt[w][tid] = 0;
    __syncthreads();
    const int start = blockIdx.x * RADIX_TILE;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int r = 0; r < RADIX_ITEMS; r++) {
        const int i = start + r * TPB + tid;
        const bool valid = i < n;
        const unsigned long long key = valid ? kin[i] : 0ull;
        const int val = valid ? vin[i] : 0;
        const uint32_t d = valid ? static_cast<uint32_t>((key >> shift) & 255u) : 0u;
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const unsigned long long b = __ballot((d >> bit) & 1u);
            m &= ((d >> bit) & 1u) ? b : ~b;
        }
        const uint32_t rank = __popcll(m & below);
        if (valid && rank == 0) wcnt[wave][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t off = base[d];
            for (int w = 0; w < wave; w++) off += wcnt[w][d];
            kout[off + rank] = key;
            vout[off + rank] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < TPB / 64; w++) { add += wcnt[w][tid]; wcnt[w][tid] = 0; }
        base[tid] += add;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ broad phase ----
// solver_broadphase.cpp:101-122 + rigid.cpp:83-106: swept AABB of every body.  Bodies
// outside any world get an empty box so they never overlap.
constexpr float LIST_GROW = 0.1f;
__global__ void k_swept_aabb(State S)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    const int w = S.bodyWorld[i];
    if (w < 0) {
        S.aabbC[i] = make_float4(0, 0, 0, 0);
        S.aabbH[i] = make_float4(-1, -1, -1, 0);
        S.listC[i] = make_float4(0, 0, 0, 0);
        S.listH[i] = make_float4(-1, -1, -1, 0);
        S.listFlag[i] = 0;
        return;
    }
    const float4 sr = S.sizeR[i];
    const int shape = S.shape[i];
    // rigid.cpp:83-99: the shape's local half extents
    const v3 h = shape == AVBD_SHAPE_SPHERE     ? v3{sr.x, sr.x, sr.x}
                 : shape == AVBD_SHAPE_CYLINDER ? v3{sr.x, sr.x, sr.z * 0.5f}
                                                : v3{sr.x, sr.y, sr.z} * 0.5f;
    const m3 R = qmat(bRot(S, i));
    const v3 tight{(fabsf(R.c[0][0]) * h.x + fabsf(R.c[1][0]) * h.y) + fabsf(R.c[2][0]) * h.z,
                   (fabsf(R.c[0][1]) * h.x + fabsf(R.c[1][1]) * h.y) + fabsf(R.c[2][1]) * h.z,
                   (fabsf(R.c[0][2]) * h.x + fabsf(R.c[1][2]) * h.y) + fabsf(R.c[2][2]) * h.z};
    const float cdt = S.P.collisionDt;
    v3 disp = ld3(S.vel[i]) * cdt;
    if (bMass(S, i) > 0) disp += v3{0, 0, S.grav[i] * cdt * cdt};
    const float4 w4 = S.angVel[i];
    const float rotSweep = sr.w * length(ld3(w4)) * cdt;
    const float grow = rotSweep + S.P.specMargin;
    const float4 p4 = S.pos[i], v4 = S.vel[i];
    const v3 ac = ld3(p4) + disp * 0.5f, ah = (tight + fabs3(disp) * 0.5f) + v3{grow, grow, grow};
    S.aabbC[i] = st3(ac);
    S.aabbH[i] = st3(ah);
    // A box that left the grown box of its candidate list is grown again (by LIST_GROW of its
    // smallest half extent plus twice this step's displacement) and walks.  While every box
    // stays inside its grown box, two boxes can only overlap if their grown boxes do (the same
    // float comparisons), and those pairs are what the lists hold.  A stale world's boxes are
    // not grown (its tree holds the exact boxes).  Where the last boxes were not grown ones
    // (a stale world, or one listing again), a body counts as moved when this step's
    // displacement by its velocity exceeds half the base growth, LIST_GROW of its smallest half
    // extent.
    const float g0 = LIST_GROW * smin(smin(ah.x, ah.y), ah.z);
    const v3 vdisp = ld3(S.vel[i]) * cdt;   // the velocity's part of the displacement
    const float step = smax(smax(fabsf(vdisp.x), fabsf(vdisp.y)), fabsf(vdisp.z));
    const float reach = smax(smax(fabsf(disp.x), fabsf(disp.y)), fabsf(disp.z));
    const bool stale = S.wStale[w] != 0, listing = stale || S.wRewalk[w];
    const v3 ec = ld3(S.listC[i]), eh = ld3(S.listH[i]);
    const v3 alo = ac - ah, ahi = ac + ah, elo = ec - eh, ehi = ec + eh;
    const bool inside = alo.x >= elo.x && alo.y >= elo.y && alo.z >= elo.z && ahi.x <= ehi.x && ahi.y <= ehi.y &&
                        ahi.z <= ehi.z;
    const bool moved = listing ? step > 0.5f * g0 : !inside;
    const bool regrow = listing || !inside;
    if (regrow) {
        const float g = stale ? 0.0f : g0 + 2.0f * reach;
        S.listC[i] = st3(ac);
        S.listH[i] = st3(ah + v3{g, g, g});
        S.wRefit[w] = 1;
    }
    S.listFlag[i] = regrow ? 2 : 0;
    // moved bodies per world: one atomic per world present among the wavefront's moved lanes
    unsigned long long pending = __ballot(moved);
    while (pending) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const int wl = __shfl(w, leader);
        const unsigned long long same = __ballot(moved && w == wl);
        if (static_cast<int>(__lane_id()) == leader) atomicAdd(S.wMoved + wl, static_cast<int>(__popcll(same)));
        pending &= ~same;
    }
    // the narrow phase's view of the body, one aligned line instead of eight arrays
    float4 *r = S.narrowRec + NARROW_REC * static_cast<size_t>(i);
    r[0] = make_float4(p4.x, p4.y, p4.z, bMass(S, i));
    r[1] = S.rot[i];
    r[2] = sr;
    r[3] = make_float4(v4.x, v4.y, v4.z, S.grav[i]);
    r[4] = make_float4(w4.x, w4.y, w4.z, S.fric[i]);
    // w: the body's awake mass (0 while it sleeps; k_wake_pairs refreshes it for bodies it wakes)
    r[5] = make_float4(__int_as_float(shape), 0.0f, 0.0f, S.P.sleepOn && S.sleeping[i] ? 0.0f : bMass(S, i));
}

// Scene bounds of each world's AABB centres (bvh_broadphase.cpp:332-342); one block per
// world.  min/max are order-independent, so the block reduction is exact.
__global__ void k_world_bounds(State S)
{
    const int w = blockIdx.x;
    __shared__ float lo[3][TPB], hi[3][TPB];
    const int base = S.wBodyBase[w], cnt = S.wBodyCount[w];
    float l0 = FLT_MAXV, l1 = FLT_MAXV, l2 = FLT_MAXV, h0 = -FLT_MAXV, h1 = -FLT_MAXV, h2 = -FLT_MAXV;
    for (int k = threadIdx.x; k < cnt; k += TPB) {
        const float4 c = S.aabbC[base + k];
        l0 = smin(l0, c.x); l1 = smin(l1, c.y); l2 = smin(l2, c.z);
        h0 = smax(h0, c.x); h1 = smax(h1, c.y); h2 = smax(h2, c.z);
    }
    lo[0][threadIdx.x] = l0; lo[1][threadIdx.x] = l1; lo[2][threadIdx.x] = l2;
    hi[0][threadIdx.x] = h0; hi[1][threadIdx.x] = h1; hi[2][threadIdx.x] = h2;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                lo[a][threadIdx.x] = smin(lo[a][threadIdx.x], lo[a][threadIdx.x + s]);
                hi[a][threadIdx.x] = smax(hi[a][threadIdx.x], hi[a][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        S.wBoundsLo[w] = make_float4(lo[0][0], lo[1][0], lo[2][0], 0);
        S.wBoundsHi[w] = make_float4(hi[0][0], hi[1][0], hi[2][0], 0);
    }
}

__device__ __forceinline__ unsigned long long expandBits21(uint32_t v)
{
    unsigned long long x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10cfcfcfcfcfcfcfull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// The standard 3-D bit spreading (two zero bits after each bit of v).
__device__ __forceinline__ unsigned long long spreadBits21(uint32_t v)
{
    unsigned long long x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// bvh_broadphase.cpp:19-36, 343-348: 63-bit Morton code of the centre, every axis offset
// by the scene's lo.x (the reference's quantisation and bit spreading, kept for bit-exact
// codes in mortonCode).  The reference's fourth spreading mask keeps bits that belong to the
// neighbouring axes, so its codes interleave the axes only roughly: sorted by them, leaves that
// are far apart share subtrees (on the chains scene a walk visits ≈6 times the nodes).  The
// tree the step walks is therefore sorted on the standard interleave of the same quantised
// coordinates; the tree only decides how long the walks are, never which pairs they find
// (refOrder 1 sorts on the reference's codes instead, for the broad-phase probe).  The sort key
// is the full code for a single world; for a batch it is (segment << 43 | code >> 20) so each
// world's bodies sort within their own index range.
__global__ void k_morton(State S, int fullKey, int refOrder)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    const int w = S.bodyWorld[i];
    unsigned long long code = 0, tree = 0;
    if (w >= 0) {
        const float4 lo = S.wBoundsLo[w], hi = S.wBoundsHi[w];
        const float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
        float m = ex;
        m = smax(m, ey);
        m = smax(m, ez);
        m = smax(m, 1.0e-6f);
        const float scale = 2097151.0f / m;
        const float4 c = S.aabbC[i];
        const uint32_t qx = static_cast<uint32_t>(sclamp((c.x - lo.x) * scale, 0.0f, 2097151.0f));
        const uint32_t qy = static_cast<uint32_t>(sclamp((c.y - lo.x) * scale, 0.0f, 2097151.0f));
        const uint32_t qz = static_cast<uint32_t>(sclamp((c.z - lo.x) * scale, 0.0f, 2097151.0f));
        code = expandBits21(qx) | (expandBits21(qy) << 1) | (expandBits21(qz) << 2);
        tree = refOrder ? code : (spreadBits21(qx) | (spreadBits21(qy) << 1) | (spreadBits21(qz) << 2));
    }
    S.mortonCode[i] = code;
    const unsigned long long seg = w >= 0 ? static_cast<unsigned long long>(S.wBodyBase[w]) : static_cast<unsigned long long>(i);
    // bodies outside every world sort after the world (single-world key) or into their
    // own index position (segment key)
    S.sortKey[i] = fullKey ? (w >= 0 ? tree : ~0ull) : ((seg << 43) | (tree >> 20));
    S.sortVal[i] = i;
}

// Common-prefix length of sorted keys i and j, ties broken by position (Karras 2012).
__device__ __forceinline__ int kdelta(const unsigned long long *k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const unsigned long long a = k[i], b = k[j];
    if (a == b) return 64 + __clz(static_cast<uint32_t>(i ^ j));
    return __clzll(a ^ b);
}

// Karras radix-tree build: internal node i of n-1, children encoded as node ids
// (internal [0, n-1), leaf p at n-1+p).
__global__ void k_karras(State S, int n)
{
    const int i = GID;
    if (i >= n - 1) return;
    const unsigned long long *k = S.sortKey;
    const int d = (kdelta(k, n, i, i + 1) - kdelta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = kdelta(k, n, i, i - d);
    int lmax = 2;
    while (kdelta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(k, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = kdelta(k, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (kdelta(k, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int left = (lo == gamma) ? (n - 1 + gamma) : gamma;
    const int right = (hi == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    S.nodeLeft[i] = left;
    S.nodeRight[i] = right;
    S.nodeFirst[i] = lo;
    S.nodeLast[i] = hi;
    S.nodeParent[left] = i;
    S.nodeParent[right] = i;
    S.nodeFlag[i] = 0;
}

// Refit tasks: a world's tree is cut into subtrees of at most REFIT_LEAVES leaves (the task
// roots: nodes of that size whose parent is larger, or the world's root), chosen at each tree
// build.  The step's tree holds the grown boxes of the candidate lists (listC/listH).  One workgroup refits a task's subtree bottom up: leaf boxes (centre -/+ half, the
// reference's bounds), and the second child to arrive at a parent writes the union and climbs
// on; arrivals are counted with workgroup-scope atomics and ordered with workgroup-scope
// fences (a device-scope fence per thread would write back the L2 of every XCD).  Above the
// task roots (worlds larger than REFIT_LEAVES only) one thread per task climbs on with
// device-scope arrivals, so those few nodes pay the device-scope ordering.  Queries start at
// world roots, so the levels above them are never read.
constexpr int REFIT_TPB = 1024;
constexpr int REFIT_LEAVES = 2048;

__device__ __forceinline__ int nodeLeaves(const State &S, int n, int x)
{
    return x >= n - 1 ? 1 : S.nodeLast[x] - S.nodeFirst[x] + 1;
}

__global__ void k_task_flags(State S, int n)
{
    const int x = GID;
    if (x >= 2 * n - 1) return;
    const int first = x >= n - 1 ? x - (n - 1) : S.nodeFirst[x];
    const int last = x >= n - 1 ? first : S.nodeLast[x];
    const int w = S.bodyWorld[S.sortVal[first]];
    uint32_t flag = 0;
    if (w >= 0 && S.wRoot[w] >= 0 && first >= S.wBodyBase[w] && last <= S.wBodyBase[w] + S.wBodyCount[w] - 1 &&
        last - first + 1 <= REFIT_LEAVES)
        flag = (x == S.wRoot[w] || nodeLeaves(S, n, S.nodeParent[x]) > REFIT_LEAVES) ? 1u : 0u;
    S.taskFlag[x] = flag;
}

__global__ void k_task_list(State S, int n)
{
    const int x = GID;
    if (x >= 2 * n - 1) return;
    if (S.taskFlag[x]) S.taskNode[S.taskScan[x]] = x;
}

__device__ __forceinline__ void unionInto(const State &S, int parent)
{
    const int l = S.nodeLeft[parent], r = S.nodeRight[parent];
    const float4 a0 = S.nodeBox[2 * l], a1 = S.nodeBox[2 * l + 1];
    const float4 b0 = S.nodeBox[2 * r], b1 = S.nodeBox[2 * r + 1];
    S.nodeBox[2 * parent] = make_float4(smin(a0.x, b0.x), smin(a0.y, b0.y), smin(a0.z, b0.z), __int_as_float(l));
    S.nodeBox[2 * parent + 1] = make_float4(smax(a1.x, b1.x), smax(a1.y, b1.y), smax(a1.z, b1.z),
                                            __int_as_float(S.nodeRope[parent]));
}

__global__ void __launch_bounds__(REFIT_TPB) k_refit_tasks(State S, int n, int onlyRewalked)
{
    const int nTasks = static_cast<int>(S.taskScan[2 * n - 1]);
    for (int t = blockIdx.x; t < nTasks; t += gridDim.x) {
        const int root = S.taskNode[t];
        const int lo = root >= n - 1 ? root - (n - 1) : S.nodeFirst[root];
        // a world whose grown boxes did not change keeps its node boxes (same topology)
        if (onlyRewalked && !S.wRefit[S.bodyWorld[S.sortVal[lo]]]) continue;
        const int cnt = nodeLeaves(S, n, root);
        for (int p = lo + threadIdx.x; p < lo + cnt; p += REFIT_TPB) {
            const int body = S.sortVal[p];
            S.leafBody[p] = body;
            S.bodyLeaf[body] = p;
            const v3 c = ld3(S.aabbC[body]), h = ld3(S.aabbH[body]);
            const int node = n - 1 + p;
            S.nodeBox[2 * node] = st3(c - h, __int_as_float(body));
            S.nodeBox[2 * node + 1] = st3(c + h, __int_as_float(S.nodeRope[node]));
            if (node == root) continue;
            __threadfence_block();
            int parent = S.nodeParent[node];
            while (true) {
                if (__hip_atomic_fetch_add(&S.nodeFlag[parent], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
                    break;
                __threadfence_block();
                unionInto(S, parent);
                __threadfence_block();
                if (parent == root) break;
                parent = S.nodeParent[parent];
            }
        }
        __syncthreads();
        const int w = S.bodyWorld[S.sortVal[lo]];
        if (threadIdx.x == 0 && root != S.wRoot[w]) {
            int node = root;
            __threadfence();
            while (node != S.wRoot[w]) {
                const int parent = S.nodeParent[node];
                if (__hip_atomic_fetch_add(&S.nodeFlag[parent], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 0)
                    break;
                __threadfence();
                unionInto(S, parent);
                __threadfence();
                node = parent;
            }
        }
        __syncthreads();
    }
}

// Arrival counters of the internal nodes, cleared before each refit.
__global__ void k_refit_clear(State S, int n)
{
    const int i = GID;
    if (i < n - 1) S.nodeFlag[i] = 0;
}

// Closed-interval AABB overlap exactly as the reference tests it
// (solver_broadphase.cpp:35-40, bvh_broadphase.cpp:494-500).
__device__ __forceinline__ bool aabbOverlap(v3 aMin, v3 aMax, v3 bMin, v3 bMax)
{
    return !(bMin.x > aMax.x || bMin.y > aMax.y || bMin.z > aMax.z || aMin.x > bMax.x || aMin.y > bMax.y ||
             aMin.z > bMax.z);
}

constexpr int PAIR_BUF = 16;

// Every world's sort keys share a prefix that no other key has, so its leaves form exactly
// one subtree of the Karras tree; queries start there (a one-body world is its leaf).
__global__ void k_world_roots_init(State S, int n)
{
    const int w = GID;
    if (w >= S.numWorlds) return;
    const int cnt = S.wBodyCount[w];
    S.wRoot[w] = cnt == 1 ? n - 1 + S.wBodyBase[w] : -1;
}
__global__ void k_world_roots_check(State S)
{
    const int w = GID;
    if (w >= S.numWorlds) return;
    if (S.wBodyCount[w] >= 1 && S.wRoot[w] < 0) atomicOr(S.wErr + w, 8);
}
__global__ void k_world_roots(State S, int n)
{
    const int i = GID;
    if (i >= n - 1) return;
    const int first = S.nodeFirst[i], last = S.nodeLast[i];
    const int w = S.bodyWorld[S.sortVal[first]];
    if (w < 0) return;
    if (first == S.wBodyBase[w] && last == S.wBodyBase[w] + S.wBodyCount[w] - 1) S.wRoot[w] = i;
}

// Next node after the subtree whose leaves end at position `last`, in left-first
// depth-first order of the world's subtree (-1 when the world's last leaf is done).  That
// node is the right child of the split at `last`: internal node last+1 when its range
// starts there (an internal node always has its own index as one end of its range), else
// leaf last+1.  Computed once per tree build for every node.
__global__ void k_ropes(State S, int n)
{
    const int node = GID;
    if (node >= 2 * n - 1) return;
    const int last = node >= n - 1 ? node - (n - 1) : S.nodeLast[node];
    const int w = S.bodyWorld[S.sortVal[last]];
    int rope = -1;
    if (w >= 0 && last < S.wBodyBase[w] + S.wBodyCount[w] - 1) {
        const int q = last + 1;
        rope = (q < n - 1 && S.nodeFirst[q] == q) ? q : n - 1 + q;
    }
    S.nodeRope[node] = rope;
}

// Walk body i's world subtree without a stack (each node carries its rope); `visit(j)` for
// every j > i (Dir 1), j < i (Dir -1) or j != i (Dir 0) whose leaf box overlaps [qMin, qMax],
// in left-first depth-first order.  A walk visits each node at most once; one that does not
// end within 2n steps (a corrupt tree) stops and returns false.
template <int Dir = 1, typename F>
__device__ __forceinline__ bool walkWorld(const State &S, int n, int i, int w, v3 qMin, v3 qMax, F &&visit)
{
    const int root = S.wRoot[w];
    if (root < 0) return true;   // flagged by k_world_roots_check
    int node = root;
    for (int guard = 2 * n; node >= 0; guard--) {
        if (guard == 0) return false;
        const float4 lo = S.nodeBox[2 * node], hi = S.nodeBox[2 * node + 1];
        const bool ov = aabbOverlap(qMin, qMax, ld3(lo), ld3(hi));
        const int down = __float_as_int(lo.w), rope = __float_as_int(hi.w);
        if (node >= n - 1) {
            if (ov && (Dir > 0 ? down > i : Dir < 0 ? down < i : down != i)) visit(down);
            node = rope;
        } else {
            node = ov ? down : rope;
        }
    }
    return true;
}

// The reference's pair test on the swept boxes (closed intervals, solver_broadphase.cpp:35-40).
__device__ __forceinline__ bool sweptOverlap(const State &S, v3 qMin, v3 qMax, int j)
{
    const v3 c = ld3(S.aabbC[j]), h = ld3(S.aabbH[j]);
    return aabbOverlap(qMin, qMax, c - h, c + h);
}

constexpr int LIST_K = PAIR_BUF;

// Over half of a world's bodies moved out of their grown boxes: it is busy.  This step every
// body of it lists again (appends would cost more); from the next step on, until it calms
// down, it is stale.
__device__ __forceinline__ bool worldBusy(const State &S, int w)
{
    return S.wMoved[w] * 2 > S.wBodyCount[w];
}

// ---- wide bodies (State::wWide) ----
constexpr int NO_WIDE = 0x7fffffff;

// At a tree rebuild, after k_world_bounds (wWide filled with NO_WIDE): a body whose swept box
// spans at least a quarter of its world's centre bounds on two axes is a candidate; the
// world's lowest-indexed candidate is its wide body.  Which body is wide changes how its
// pairs are found, never which pairs are found.
__global__ void k_wide_pick(State S)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    const int w = S.bodyWorld[i];
    if (w < 0 || S.wBodyCount[w] < 2) return;
    const float4 lo = S.wBoundsLo[w], hi = S.wBoundsHi[w];
    const v3 h = ld3(S.aabbH[i]);
    const int axes = (8.0f * h.x >= hi.x - lo.x ? 1 : 0) + (8.0f * h.y >= hi.y - lo.y ? 1 : 0) +
                     (8.0f * h.z >= hi.z - lo.z ? 1 : 0);
    if (axes >= 2) atomicMin(S.wWide + w, i);
}

// Every step, after the swept boxes and before the walks: bit j of wideMask is set when body j
// lies above its world's wide body and their swept boxes meet (the pair test the wide body's
// own walk would make); wideWord counts each word's bits for the scan.  The wide body's list
// is kept closed, so that no walk appends to it.  Launched with a multiple of 64 threads per
// workgroup over every word of wideMask.
__global__ void __launch_bounds__(TPB) k_wide_flags(State S)
{
    const int j = GID;
    bool on = false;
    if (j < S.numBodies) {
        const int w = S.bodyWorld[j];
        const int wi = w >= 0 ? S.wWide[w] : NO_WIDE;
        if (wi == j) {
            S.listCnt[j] = LIST_K + 1;
        } else if (wi != NO_WIDE && j > wi) {
            const v3 c = ld3(S.aabbC[wi]), h = ld3(S.aabbH[wi]);
            on = sweptOverlap(S, c - h, c + h, j);
        }
    }
    const unsigned long long m = __ballot(on);
    if ((threadIdx.x & 63) == 0) {
        S.wideMask[j >> 6] = m;
        S.wideWord[j >> 6] = static_cast<uint32_t>(__popcll(m));
    }
}

// Number of bodies below index x with their wideMask bit set (x <= numBodies).
__device__ __forceinline__ uint32_t wideRank(const State &S, int x)
{
    const int word = x >> 6, bit = x & 63;
    const unsigned long long below = bit ? (S.wideMask[word] & ((1ull << bit) - 1ull)) : 0ull;
    return S.wideScan[word] + static_cast<uint32_t>(__popcll(below));
}

// Body i's swept-box pairs (j > i) by walking the tree; the count and the first PAIR_BUF
// to pairScratch.  `exactLeaves`: the leaves are the swept boxes themselves.
__device__ __forceinline__ void walkPairs(const State &S, int n, int i, int w, v3 qMin, v3 qMax, bool exactLeaves)
{
    uint32_t count = 0;
    int *buf = S.pairScratch + static_cast<size_t>(i) * PAIR_BUF;
    const bool ok = walkWorld(S, n, i, w, qMin, qMax, [&](int j) {
        if (!exactLeaves && !sweptOverlap(S, qMin, qMax, j)) return;
        if (count < PAIR_BUF) buf[count] = j;
        count++;
    });
    if (!ok) atomicOr(S.wErr + w, 4);
    S.pairCount[i] = count;
}

// The walks of a step, one kernel so that they wait on each other's latency only once:
//  - a stale world's bodies walk the exact boxes for their pairs;
//  - in a world that lists again (new, calmed down, busy, or at a tree rebuild) every body lists
//    the j > i whose grown boxes meet its own;
//  - elsewhere a body whose grown box changed lists again, and adds itself to the list of each
//    lower body whose grown box meets its new one and that keeps its list (unless already there:
//    the lists' first listCnt entries do not change in this kernel, appends go after them);
//  - a body with more than LIST_K candidates walks for its pairs.
// Bodies whose pairs are done here get listFlag 3.
__global__ void __launch_bounds__(TPB) k_list_walk(State S, int n)
{
    const int p = XGID;
    if (p >= S.numBodies || n < 2) return;
    const int i = S.sortVal[p];
    const int w = S.bodyWorld[i];
    if (w < 0) return;
    if (S.wWide[w] == i) {   // its pairs are the bodies above it with their wideMask bit set
        S.pairCount[i] = wideRank(S, S.wBodyBase[w] + S.wBodyCount[w]) - wideRank(S, i + 1);
        S.listFlag[i] = 3;
        return;
    }
    const v3 c = ld3(S.aabbC[i]), h = ld3(S.aabbH[i]);
    if (S.wStale[w]) {
        walkPairs(S, n, i, w, c - h, c + h, true);
        S.listFlag[i] = 3;
        return;
    }
    const bool all = S.wRewalk[w] || worldBusy(S, w);
    int lc;
    if (all || S.listFlag[i] == 2) {
        const v3 ec = ld3(S.listC[i]), eh = ld3(S.listH[i]);
        int v[LIST_K];
        lc = 0;
        auto own = [&](int j) {
            if (lc < LIST_K) {
                int b = lc;
                while (b > 0 && v[b - 1] > j) { v[b] = v[b - 1]; b--; }
                v[b] = j;
            }
            lc++;
        };
        bool ok;
        if (all) {
            ok = walkWorld(S, n, i, w, ec - eh, ec + eh, own);
        } else {
            ok = walkWorld<0>(S, n, i, w, ec - eh, ec + eh, [&](int j) {
                if (j > i) { own(j); return; }
                if (S.listFlag[j] != 0) return;   // lists again itself
                const int base = S.listCnt[j];
                if (base > LIST_K) return;          // walks for its pairs
                const int *lj = S.listJ + static_cast<size_t>(j) * LIST_K;
                for (int k = 0; k < base; k++)
                    if (lj[k] == i) return;
                const int slot = base + atomicAdd(S.listAdd + j, 1);
                if (slot < LIST_K) S.listJ[static_cast<size_t>(j) * LIST_K + slot] = i;
            });
        }
        if (!ok) atomicOr(S.wErr + w, 4);
        int *list = S.listJ + static_cast<size_t>(i) * LIST_K;
        if (lc <= LIST_K)
            for (int k = 0; k < lc; k++) list[k] = v[k];
        S.listCnt[i] = lc;
    } else {
        lc = S.listCnt[i];
    }
    if (lc > LIST_K) {
        walkPairs(S, n, i, w, c - h, c + h, false);
        S.listFlag[i] = 3;
    }
}

// Candidate pairs of body i: every j > i of the same world whose swept box overlaps i's, from
// its list (with this step's appends; a list that overflowed with them walks), unless the walk
// kernel made them.  The count and the first PAIR_BUF pairs go to pairScratch; the emit pass
// copies those (or, for a body with more, walks again) into pairJ at pairOff[i], sorted by j.
__global__ void __launch_bounds__(TPB) k_traverse(State S, int n)
{
    // body order (not the tree's): a lane's list and outputs sit next to its neighbours'
    const int i = XGID;
    if (i >= S.numBodies) return;
    const int w = S.bodyWorld[i];
    if (w < 0 || n < 2) { S.pairCount[i] = 0; return; }
    if (S.listFlag[i] == 3) return;
    int lc = S.listCnt[i];
    const int add = S.listAdd[i];
    if (add) {
        lc += add;
        S.listAdd[i] = 0;
        S.listCnt[i] = lc;
    }
    const v3 c = ld3(S.aabbC[i]), h = ld3(S.aabbH[i]);
    const v3 qMin = c - h, qMax = c + h;
    if (lc > LIST_K) {
        walkPairs(S, n, i, w, qMin, qMax, false);
        return;
    }
    uint32_t count = 0;
    int *buf = S.pairScratch + static_cast<size_t>(i) * PAIR_BUF;
    const int *list = S.listJ + static_cast<size_t>(i) * LIST_K;
    for (int k = 0; k < lc; k++) {
        const int j = list[k];
        if (sweptOverlap(S, qMin, qMax, j)) buf[count++] = j;
    }
    S.pairCount[i] = count;
}

// Per world after the walks: busy now = stale next step; a stale world that calmed down lists
// again; counters for the next step cleared.  listStats (cleared by the host's fill) gathers
// the step's figures.
__global__ void k_list_world_end(State S)
{
    const int w = GID;
    if (w >= S.numWorlds) return;
    const bool busy = worldBusy(S, w), wasStale = S.wStale[w] != 0;
    const int moved = S.wMoved[w];
    if (moved) atomicAdd(S.listStats, moved);
    if (busy) atomicAdd(S.listStats + 1, 1);
    if (wasStale) atomicAdd(S.listStats + 2, 1);
    if (S.wRewalk[w] && !wasStale) atomicAdd(S.listStats + 3, 1);
    S.wRewalk[w] = wasStale && !busy ? 1 : 0;
    S.wStale[w] = busy ? 1 : 0;
    S.wMoved[w] = 0;
    S.wRefit[w] = 0;
}

__global__ void __launch_bounds__(TPB) k_pair_emit(State S, int n)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    const int w = S.bodyWorld[i];
    if (w >= 0) {
        // the wide body's range is written by the bodies that meet it, each at its rank
        const int wi = S.wWide[w];
        if (wi == i) return;
        if (wi != NO_WIDE && i > wi && ((S.wideMask[i >> 6] >> (i & 63)) & 1ull) &&
            S.pairOff[wi + 1] <= static_cast<uint32_t>(S.pairCap)) {
            const uint32_t at = S.pairOff[wi] + (wideRank(S, i) - wideRank(S, wi + 1));
            S.pairJ[at] = i;
            S.pairI[at] = wi;
        }
    }
    const uint32_t begin = S.pairOff[i], end = S.pairOff[i + 1];
    const uint32_t cnt = end - begin;
    if (cnt == 0) return;
    if (end > static_cast<uint32_t>(S.pairCap)) return;   // reported as a capacity error
    if (cnt <= PAIR_BUF) {
        const int *buf = S.pairScratch + static_cast<size_t>(i) * PAIR_BUF;
        int v[PAIR_BUF];
        for (uint32_t k = 0; k < cnt; k++) {
            const int x = buf[k];
            uint32_t b = k;
            while (b > 0 && v[b - 1] > x) { v[b] = v[b - 1]; b--; }
            v[b] = x;
        }
        for (uint32_t k = 0; k < cnt; k++) {
            S.pairJ[begin + k] = v[k];
            S.pairI[begin + k] = i;
        }
        return;
    }
    const v3 c = ld3(S.aabbC[i]), h = ld3(S.aabbH[i]);
    const v3 qMin = c - h, qMax = c + h;
    uint32_t out = begin;
    walkWorld(S, n, i, S.bodyWorld[i], qMin, qMax, [&](int j) {
        if (out < end && sweptOverlap(S, qMin, qMax, j)) S.pairJ[out++] = j;
    });
    for (uint32_t a = begin + 1; a < out; a++) {
        const int v = S.pairJ[a];
        uint32_t b = a;
        while (b > begin && S.pairJ[b - 1] > v) { S.pairJ[b] = S.pairJ[b - 1]; b--; }
        S.pairJ[b] = v;
    }
    for (uint32_t a = begin; a < out; a++) S.pairI[a] = i;
}

// The body on the other side of an adjacency entry (solver_internal.hpp:364-384).
__device__ __forceinline__ int otherOf(const State &S, int ref, int body)
{
    int a, b;
    if (ref >= 0) { const int4 h = S.mHdr[adjManifold(ref)]; a = h.x; b = h.y; }
    else { const int j = -ref - 1; a = S.jA[j]; b = S.jB[j]; }
    if (a == body) return b;
    if (b == body) return a;
    return -1;
}

// Start of XCD x's part of [0, total): about x / NUM_XCD of it, rounded up to a multiple of
// 64 so that (with blocks of whole wavefronts) each wavefront covers one contact slice.
__device__ __forceinline__ int xcdPartStart(int total, int x)
{
    if (x >= NUM_XCD) return total;
    const int b = static_cast<int>((static_cast<long long>(total) * x) / NUM_XCD);
    const int r = (b + 63) & ~63;
    return r < total ? r : total;
}

// Grid-stride walk over every world's live manifolds (the compact list gLive).
// Each XCD takes one contiguous eighth of [0, total) (its workgroups grid-stride inside it).
#define FOR_LIVE(k) FOR_RANGE(k, static_cast<int>(S.wLiveOff[S.numWorlds]))
#define FOR_RANGE(k, total)                                                                                 \
    for (int k##_t = (total), k##_x = static_cast<int>(blockIdx.x) % NUM_XCD,                              \
             k##_nb = static_cast<int>(gridDim.x) / NUM_XCD, k##_hi = xcdPartStart(k##_t, k##_x + 1),      \
             k = xcdPartStart(k##_t, k##_x) + static_cast<int>(blockIdx.x) / NUM_XCD * static_cast<int>(blockDim.x) + \
                 static_cast<int>(threadIdx.x);                                                             \
         k < k##_hi; k += k##_nb * static_cast<int>(blockDim.x))

// One workgroup per world: the view order, a stable counting sort of the world's live
// manifolds by contact count (ballot ranks inside a wavefront, LDS counts across them).
constexpr int ORDER_TPB = 256;
__global__ void __launch_bounds__(ORDER_TPB) k_manifold_order(State S)
{
    const int w = S.wOrder[blockIdx.x];
    const int base = static_cast<int>(S.wLiveOff[w]), n = S.wLive[w];
    if (n <= 0) return;
    constexpr int NB = 16, NW = ORDER_TPB / 64;
    __shared__ uint32_t start[NB], wcnt[NW][NB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    if (tid < NB) start[tid] = 0;
    if (tid < NW * NB) wcnt[tid / NB][tid % NB] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += ORDER_TPB) atomicAdd(&start[S.mNum[S.live[S.gLive[base + i]]]], 1u);
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int v = 0; v < NB; v++) { const uint32_t c = start[v]; start[v] = run; run += c; }
    }
    __syncthreads();
    for (int c0 = 0; c0 < n; c0 += ORDER_TPB) {
        const int i = c0 + tid;
        const bool valid = i < n;
        const int slot = valid ? S.live[S.gLive[base + i]] : 0;
        const uint32_t v = valid ? static_cast<uint32_t>(S.mNum[slot]) : 0u;
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 4; bit++) {
            const unsigned long long b = __ballot((v >> bit) & 1u);
            m &= ((v >> bit) & 1u) ? b : ~b;
        }
        const uint32_t rank = __popcll(m & below);
        if (valid && rank == 0) wcnt[wave][v] = __popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t off = start[v];
            for (int q = 0; q < wave; q++) off += wcnt[q][v];
            const int k = base + static_cast<int>(off + rank);
            S.gOfK[k] = base + i;
            S.kSlot[k] = slot;
            S.kT[k] = S.gLive[base + i];
            S.cCount[k] = v;   // contact count in view order (scanned into dense offsets)
        }
        __syncthreads();
        if (tid < NB) {
            uint32_t add = 0;
            for (int q = 0; q < NW; q++) { add += wcnt[q][tid]; wcnt[q][tid] = 0; }
            start[tid] += add;
        }
        __syncthreads();
    }
}

__global__ void k_glive_fill(State S)
{
    const int w = blockIdx.x;
    const int off = static_cast<int>(S.wLiveOff[w]), n = S.wLive[w], base = S.wSlotBase[w];
    for (int i = threadIdx.x; i < n; i += blockDim.x) S.gLive[off + i] = base + i;
}

// solver_broadphase.cpp:30-65: bounds, layers, no force or manifold between the pair,
// not both static.  `constrained` scans the shorter adjacency list (same answer).
// Number of candidate pairs of this step (pairs past the capacity are reported, not kept).
__device__ __forceinline__ int pairTotal(const State &S)
{
    const uint32_t t = S.pairOff[S.numBodies];
    return t < static_cast<uint32_t>(S.pairCap) ? static_cast<int>(t) : S.pairCap;
}

// collide.cpp:1107-1117: the speculative margin grown by the pair's predicted relative motion,
// from the two bodies' narrow-phase records ({pos, mass}, rot, {size, radius}, {vel, gravity},
// {angVel, friction}).
AD float pairMargin(const State &S, float ma, float4 a2, float4 a3, float4 a4, float mb, float4 b2, float4 b3,
                    float4 b4)
{
    float margin = S.P.specMargin;
    if (margin > 0.0f) {
        const float hz = S.P.collisionDt;
        const float gA = ma > 0.0f ? a3.w : 0.0f;
        const float gB = mb > 0.0f ? b3.w : 0.0f;
        const v3 disp = (ld3(a3) - ld3(b3)) * hz + v3{0, 0, (gA - gB) * hz * hz};
        margin += length(disp) + (a2.w * length(ld3(a4)) + b2.w * length(ld3(b4))) * hz;
    }
    return margin;
}

// A new pair of boxes that a separating axis parts by more than the speculative margin: the
// narrow phase would make its manifold with no contact (collide.cpp:238-312) and release it in
// the same step (solver_solve.cpp:66-87), so no manifold is made.  Same inputs and floats as
// the narrow phase's own test.  Pairs with no awake movable side are left to the narrow phase,
// which keeps their manifolds; a side asleep now may still be woken before it runs, which only
// turns such a pair from frozen to awake, so an awake pair here is awake there.
AD bool boxesApart(const State &S, int i, int j)
{
    const float4 *ra = S.narrowRec + NARROW_REC * static_cast<size_t>(i);
    const float4 *rb = S.narrowRec + NARROW_REC * static_cast<size_t>(j);
    const float4 a5 = ra[5], b5 = rb[5];
    if (__float_as_int(a5.x) != AVBD_SHAPE_BOX || __float_as_int(b5.x) != AVBD_SHAPE_BOX) return false;
    if (!(a5.w > 0.0f) && !(b5.w > 0.0f)) return false;
    const float4 a0 = ra[0], a1 = ra[1], a2 = ra[2], a3 = ra[3], a4 = ra[4];
    const float4 b0 = rb[0], b1 = rb[1], b2 = rb[2], b3 = rb[3], b4 = rb[4];
    const float margin = pairMargin(S, a0.w, a2, a3, a4, b0.w, b2, b3, b4);
    const OBB bA = makeOBB(ld3(a0), ldq(a1), v3{a2.x, a2.y, a2.z} * 0.5f);
    const OBB bB = makeOBB(ld3(b0), ldq(b1), v3{b2.x, b2.y, b2.z} * 0.5f);
    BoxSat sat;
    return !boxSat(bA, bB, margin, sat);
}

__device__ void pairFilter(const State &S, int p);
__global__ void k_pair_filter(State S)
{
    FOR_RANGE(p, pairTotal(S)) pairFilter(S, p);
}
__device__ void pairFilter(const State &S, int p)
{
    const int i = S.pairI[p], j = S.pairJ[p];
    const v3 ci = ld3(S.aabbC[i]), hi = ld3(S.aabbH[i]), cj = ld3(S.aabbC[j]), hj = ld3(S.aabbH[j]);
    const bool ov = aabbOverlap(ci - hi, ci + hi, cj - hj, cj + hj);
    const bool layers = ((S.layer[i] & S.mask[j]) != 0) || ((S.layer[j] & S.mask[i]) != 0);
    const bool staticI = bMass(S, i) <= 0.0f, staticJ = bMass(S, j) <= 0.0f;
    const bool bothStatic = staticI && staticJ;
    bool constrained = false, parked = false;
    if (ov && layers && !bothStatic) {
        // A static body's list holds its joints only (manifold entries are kept for movable
        // bodies), so a pair with a static body scans the movable body's list: every force
        // between the two is in it.
        const int ni = S.adjStart[i + 1] - S.adjStart[i], nj = S.adjStart[j + 1] - S.adjStart[j];
        const bool useI = staticJ || (!staticI && ni <= nj);
        const int a = useI ? i : j, b = useI ? j : i;
        for (int e = S.adjStart[a]; e < S.adjStart[a + 1]; e++) {
            if (otherOf(S, S.adj[e], a) != b) continue;
            // solver_internal.hpp:67-84: a manifold parked by contact grace does not join the
            // pair (it can still wake sleepers), but it keeps a second manifold from being made
            if (S.P.grace > 0 && S.adj[e] >= 0) {
                const int k = adjManifold(S.adj[e]);
                if (S.mHdr[k].w == 0 && S.mDormant[S.kSlot[k]] > 0) { parked = true; continue; }
            }
            constrained = true;
            break;
        }
    }
    const bool pass = ov && layers && !bothStatic && !constrained;
    // contact grace keeps a new manifold with no contact (parked), so only without it can a
    // separated pair skip the narrow phase
    const bool apart = pass && !parked && S.P.grace == 0 && boxesApart(S, i, j);
    S.pairValid[p] = (pass && !parked && !apart) ? 1u : 0u;
    S.pairApart[p] = apart ? 1 : 0;
    if (S.P.grace > 0) S.pairParked[p] = pass && parked ? 1 : 0;
}

// solver_broadphase.cpp:153-184 + manifold.cpp:205-268: the r-th new pair of a world
// takes the r-th slot popped from its free list (LIFO), else a fresh slot, and is
// appended to the live list in sorted pair order.
__device__ void allocNew(const State &S, int p);
__global__ void k_alloc_new(State S)
{
    FOR_RANGE(p, pairTotal(S)) allocNew(S, p);
}
__device__ void allocNew(const State &S, int p)
{
    if (!S.pairValid[p]) return;
    const int i = S.pairI[p], j = S.pairJ[p];
    const int w = S.bodyWorld[i];
    const uint32_t start = S.pairOff[S.wBodyBase[w]];
    const int r = static_cast<int>(S.pairValidScan[p] - S.pairValidScan[start]);
    const int base = S.wSlotBase[w], cap = S.wSlotCap[w];
    const int fr = S.wFree[w], lv = S.wLive[w], top = S.wTop[w];
    // the narrow phase takes new manifolds in pair order, after the previous step's view
    int *newT = S.newT + S.pairValidScan[p];
    int slot;
    if (r < fr) {
        slot = S.freeList[base + fr - 1 - r];
    } else {
        const int fresh = top + (r - fr);
        if (fresh >= cap) { atomicOr(S.wErr + w, 1); *newT = -1; return; }
        slot = base + fresh;
    }
    if (lv + r >= cap) { atomicOr(S.wErr + w, 1); *newT = -1; return; }
    *newT = base + lv + r;
    S.live[base + lv + r] = slot;
    S.mA[slot] = i;
    S.mB[slot] = j;
    S.mNum[slot] = 0;
    S.mDormant[slot] = 0;
    S.mSaved[slot] = 0;
    storeRec(S, slot, zero3(), i, j, 0, 0.0f, 0);   // manifold.cpp:55-66: basis, friction, count reset
}

__global__ void k_alloc_counts(State S)
{
    const int w = GID;
    if (w >= S.numWorlds) return;
    const int cnt = S.wBodyCount[w];
    int newCount = 0;
    if (cnt > 0) {
        const uint32_t a = S.pairOff[S.wBodyBase[w]], b = S.pairOff[S.wBodyBase[w] + cnt];
        newCount = static_cast<int>(S.pairValidScan[b] - S.pairValidScan[a]);
    }
    const int fr = S.wFree[w];
    const int used = newCount < fr ? newCount : fr;
    int top = S.wTop[w] + (newCount - used);
    int lv = S.wLive[w] + newCount;
    if (top > S.wSlotCap[w]) top = S.wSlotCap[w];
    if (lv > S.wSlotCap[w]) lv = S.wSlotCap[w];
    S.wFree[w] = fr - used;
    S.wTop[w] = top;
    S.wLive[w] = lv;
    S.wNew[w] = newCount;
}

// ------------------------------------------------------- contact storage ----
// Copy n contact rows of the five contact fields (source and destination never overlap, so
// every load of a row is issued before its stores).
// Staging row of contact i of narrow-phase work item q: sliced like the dense contacts, so
// the narrow phase (one lane per work item) writes contact i of a wavefront as one run.
__device__ __forceinline__ size_t stageRow(int q, int i)
{
    return static_cast<size_t>(q / CSLICE) * (CSLICE * MAXC) + static_cast<size_t>(i) * CSLICE + q % CSLICE;
}

// Staging rows (float4 rA|beta, rB|key, C0|stick, penalty, lambda; row i of a manifold at
// x + i * sx) to dense contacts (row i at y + i * sy, penalty and lambda packed, stick a byte),
// and back.
__device__ __forceinline__ void stageToDense(const State &S, size_t y, size_t sy, size_t x, size_t sx, int n)
{
    for (int i = 0; i < n; i++) {
        const size_t xi = x + i * sx, yi = y + i * sy;
        const float4 a = S.sRA[xi], b = S.sRB[xi], c = S.sC0[xi], p = S.sPen[xi], l = S.sLam[xi];
        S.cRA[yi] = a;
        S.cRB[yi] = b;
        S.cC0[yi] = c;
        S.cPL[yi] = make_float4(p.x, p.y, p.z, l.x);
        S.cLyz[yi] = make_float2(l.y, l.z);
        S.cStick[yi] = c.w != 0.0f ? 1 : 0;
    }
}
__device__ __forceinline__ void denseToStage(const State &S, size_t x, size_t sx, size_t y, size_t sy, int n)
{
    for (int i = 0; i < n; i++) {
        const size_t xi = x + i * sx, yi = y + i * sy;
        const float4 a = S.cRA[yi], b = S.cRB[yi], c = S.cC0[yi], pl = S.cPL[yi];
        const float2 lyz = S.cLyz[yi];
        const uint8_t st = S.cStick[yi];
        S.sRA[xi] = a;
        S.sRB[xi] = b;
        S.sC0[xi] = make_float4(c.x, c.y, c.z, st ? 1.0f : 0.0f);
        S.sPen[xi] = make_float4(pl.x, pl.y, pl.z, 0.0f);
        S.sLam[xi] = make_float4(pl.w, lyz.x, lyz.y, 0.0f);
    }
}

// Copy a live manifold's dense contacts to its host-layout staging rows (slot * MAXC + i).
__device__ __forceinline__ void stageContacts(const State &S, int slot)
{
    denseToStage(S, static_cast<size_t>(slot) * MAXC, 1, S.cOff[slot], CSLICE, S.mNum[slot]);
}

// Dense per-step view of live manifold k (stored in `slot`, contacts from `o`).
__device__ __forceinline__ void manifoldView(const State &S, int k, int slot, uint32_t o, int n)
{
    const float4 *r = S.mRec + 4 * static_cast<size_t>(slot);
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    S.mHdr[k] = make_int4(__float_as_int(r0.w), __float_as_int(r1.w), static_cast<int>(o), n);
    S.mBasisK[3 * k] = make_float4(r0.x, r0.y, r0.z, 0.0f);
    S.mBasisK[3 * k + 1] = make_float4(r1.x, r1.y, r1.z, 0.0f);
    S.mBasisK[3 * k + 2] = make_float4(r2.x, r2.y, r2.z, 0.0f);
    S.mFricK[k] = r3.x;
}

// The view alone, contacts left where they are (after host edits of the live lists).
__global__ void k_manifold_view(State S)
{
    FOR_LIVE(k) {
        const int slot = S.kSlot[k];
        manifoldView(S, k, slot, S.cOff[slot], S.mNum[slot]);
    }
}

// Width of each contact slice (CSLICE consecutive view indices, one wavefront of this
// launch): CSLICE times its largest contact count.  sliceCount = the number of slices.
__global__ void __launch_bounds__(TPB) k_slice_widths(State S)
{
    const int total = static_cast<int>(S.wLiveOff[S.numWorlds]);
    const int k = GID;
    uint32_t w = k < total ? S.cCount[k] : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) w = max(w, static_cast<uint32_t>(__shfl_xor(static_cast<int>(w), off)));
    // cScan has slotCap + 1 entries; slices past slotCap / CSLICE hold no live manifold
    if ((threadIdx.x & 63) == 0 && k / CSLICE <= S.slotCap / CSLICE) S.cScan[k / CSLICE] = w * CSLICE;
    if (k == 0) *S.sliceCount = static_cast<uint32_t>((total + CSLICE - 1) / CSLICE);
}

// Staging -> sliced dense arrays (cScan holds each slice's start), one lane per manifold.
// After the narrow phase the staging rows are those of the manifold's work item (its record's
// row 2 w), which
// mostly runs in the same order as the view, so reads and writes are both runs; after host
// edits they are the host layout (slot * MAXC + i).
__global__ void k_contact_pack(State S, int hostLayout)
{
    FOR_LIVE(k) {
        const int slot = S.kSlot[k];
        const int n = static_cast<int>(S.cCount[k]);
        const uint32_t o = S.cScan[k / CSLICE] + static_cast<uint32_t>(k % CSLICE);
        if (hostLayout) stageToDense(S, o, CSLICE, static_cast<size_t>(slot) * MAXC, 1, n);
        else stageToDense(S, o, CSLICE, stageRow(recItem(S, slot), 0), CSLICE, n);
        S.cOff[slot] = o;
        manifoldView(S, k, slot, o, n);
    }
}

// Dense -> staging for every live manifold (host reads and writes go through staging).
__global__ void k_contact_unpack(State S)
{
    FOR_LIVE(k) stageContacts(S, S.live[S.gLive[k]]);
}

// ------------------------------------------------------------ narrow phase ----
// manifold.cpp:237-435 with collide.cpp:1100-1120: collide at the step-start pose,
// carry hot state across matching feature keys, C0, release, warm-start decay.
// One thread per live manifold.  Work item q < P is view index q of the previous step (the
// manifolds that were live then, in that order, so that their old contacts, headers and the
// staging rows are read and written as runs); items past P are this step's new manifolds in
// pair order.  The clip polygons live in LDS, two per lane.  `keep` of positions that are not
// live is cleared beforehand.
// Box pairs run in k_manifold_init; pairs with a sphere or a cylinder in k_manifold_init_round
// (launched only when such bodies exist).  Each work item is handled by exactly one of the two:
// the box pass also parks frozen pairs and reports capacity errors.
typedef float NarrowPolys[2][3][POLY_LDS * NARROW_TPB];
// MODE 0: box pairs, MODE 1: pairs with a round shape, MODE 2 (contact grace on): every pair,
// with the grace rules (manifold.cpp:249-378).
template <int MODE> AD void manifoldInit(const State &S, int q, int P, NarrowPolys *poly);
// At most 128 registers (four wavefronts per SIMD; it fits without spilling vector registers).
__global__ void __launch_bounds__(NARROW_TPB) __attribute__((amdgpu_waves_per_eu(4))) k_manifold_init(State S)
{
    __shared__ NarrowPolys poly;   // 12 KiB
    const int P = static_cast<int>(S.wLiveOff[S.numWorlds]);
    const int fresh = static_cast<int>(S.pairValidScan[pairTotal(S)]);
    FOR_RANGE(q, P + fresh) manifoldInit<0>(S, q, P, &poly);
}
__global__ void __launch_bounds__(NARROW_TPB) k_manifold_init_grace(State S)
{
    __shared__ NarrowPolys poly;
    const int P = static_cast<int>(S.wLiveOff[S.numWorlds]);
    const int fresh = static_cast<int>(S.pairValidScan[pairTotal(S)]);
    FOR_RANGE(q, P + fresh) manifoldInit<2>(S, q, P, &poly);
}
__global__ void __launch_bounds__(NARROW_TPB) k_manifold_init_round(State S)
{
    const int P = static_cast<int>(S.wLiveOff[S.numWorlds]);
    const int fresh = static_cast<int>(S.pairValidScan[pairTotal(S)]);
    FOR_RANGE(q, P + fresh) manifoldInit<1>(S, q, P, nullptr);
}
template <int MODE> AD void manifoldInit(const State &S, int q, int P, NarrowPolys *poly)
{
    constexpr bool ROUND = MODE == 1;
    int t, slot, a, b, oldN;
    uint32_t ob;
    if (q < P) {
        slot = S.kSlot[q];
        t = S.kT[q];
        const int4 hd = S.mHdr[q];
        a = hd.x; b = hd.y; ob = static_cast<uint32_t>(hd.z); oldN = hd.w;
    } else {
        t = S.newT[q - P];
        if (t < 0) return;   // not allocated (capacity error reported)
        slot = S.live[t];
        a = S.mA[slot]; b = S.mB[slot]; ob = 0u; oldN = 0;
        // past the staging capacity only when allocations failed this step (items of failed
        // pairs hold places in the list): report it for this manifold's world too
        if (q >= S.slotCap) {
            if (!ROUND) atomicOr(S.wErr + S.bodyWorld[a], 1);
            return;
        }
    }
    const float4 *ra = S.narrowRec + NARROW_REC * static_cast<size_t>(a);
    const float4 *rb = S.narrowRec + NARROW_REC * static_cast<size_t>(b);
    const float4 a0 = ra[0], a1 = ra[1], a2 = ra[2], a3 = ra[3], a4 = ra[4], a5 = ra[5];
    const float4 b0 = rb[0], b1 = rb[1], b2 = rb[2], b3 = rb[3], b4 = rb[4], b5 = rb[5];
    const float ma = a0.w, mb = b0.w;
    // manifold.cpp:168-172: parked when neither side is an awake movable body
    const bool frozen = !(a5.w > 0.0f) && !(b5.w > 0.0f);
    const int shA = __float_as_int(a5.x), shB = __float_as_int(b5.x);
    const bool round = (shA | shB) != AVBD_SHAPE_BOX;
    if (ROUND && (frozen || !round)) return;
    if (!ROUND && frozen) {   // parked verbatim (under grace with its kept contacts too)
        denseToStage(S, stageRow(q, 0), CSLICE, ob, CSLICE, oldN);
        S.keep[t] = 1;
        S.mRec[4 * static_cast<size_t>(slot) + 2].w = __int_as_float(q);
        return;
    }
    if (MODE == 0 && round) return;   // the round pass takes it
    // manifold.cpp:246-247: a new contact set drops the captured arms
    if (S.armCap) S.armCap[2 * static_cast<size_t>(slot)].w = 0.0f;
    const float fa = a4.w, fb = b4.w;
    const float mu = sqrtf(fa * fb);
    const v3 pa = ld3(a0), pb = ld3(b0);
    const qt qa = ldq(a1), qb = ldq(b1);
    const float4 sa = a2, sb = b2;
    // speculative margin grown by the pair's predicted relative motion
    const float margin = pairMargin(S, ma, a2, a3, a4, mb, b2, b3, b4);
    m3 B;
    bool wrote;
    const size_t d = stageRow(q, 0);
    Contacts cs{0, S.sRA + d, S.sRB + d, S.sC0 + d};
    int nn;
    if (ROUND || (MODE == 2 && round)) {
        nn = collideRound(shA, shB, makeRShp(pa, qa, sa), makeRShp(pb, qb, sb), S.rim, margin, B, wrote, cs);
    } else {
        const int l = threadIdx.x;
        // polygon vertices past the LDS part use this manifold's penalty and lambda staging rows
        // (32 floats each), which are written only after the contacts are built
        const Poly c0{&(*poly)[0][0][l], &(*poly)[0][1][l], &(*poly)[0][2][l], S.sPen + d};
        const Poly c1{&(*poly)[1][0][l], &(*poly)[1][1][l], &(*poly)[1][2][l], S.sLam + d};
        nn = collideBoxes(pa, qa, v3{sa.x, sa.y, sa.z} * 0.5f, pb, qb, v3{sb.x, sb.y, sb.z} * 0.5f, margin, B, wrote,
                          cs, c0, c1);
    }
    if (!wrote) B = loadBasis(S, slot);
    // old contacts stay in the dense arrays; the new ones go to staging
    const float pmax = S.P.penaltyMax;
    // contact grace (manifold.cpp:249-253): a parked manifold (no contacts in the view) matches
    // against the contacts it kept, from the slot's keep rows
    bool revive = false;
    int liveCount = oldN;
    const size_t ds = static_cast<size_t>(slot) * MAXC;
    if constexpr (MODE == 2) {
        if (oldN == 0 && S.mSaved[slot] > 0) { revive = true; liveCount = S.mSaved[slot]; }
    }
    for (int i = 0; i < nn; i++) {
        const float4 nra = cs.ra[crow(i)], nrb = cs.rb[crow(i)];
        const int key = __float_as_int(nrb.w);
        v3 rA = ld3(nra), rB = ld3(nrb);
        v3 pen{0, 0, 0}, lam{0, 0, 0};
        uint8_t stick = 0;
        float beta = 1.0f, prevErr = -1.0f;
        for (int j = 0; j < liveCount; j++) {
            const uint32_t oc = ob + j * CSLICE;
            float4 orb, ora, oc0;
            if (MODE == 2 && revive) {
                orb = S.dRB[ds + j];
                if (key != __float_as_int(orb.w)) continue;
                ora = S.dRA[ds + j];
                oc0 = S.dC0[ds + j];
                pen = ld3(S.dPen[ds + j]);   // kept with no force: lambda 0, not stuck
                lam = v3{0, 0, 0};
                stick = 0;
            } else {
                orb = S.cRB[oc];
                if (key != __float_as_int(orb.w)) continue;
                ora = S.cRA[oc];
                oc0 = S.cC0[oc];
                loadPenLam(S, oc, pen, lam);
                stick = S.cStick[oc];
            }
            prevErr = smax(0.0f, COLLISION_MARGIN - oc0.x);
            beta = ora.w;
            if (stick) {
                const v3 oldA = pa + qrot(qa, ld3(ora));
                const v3 oldB = pb + qrot(qb, ld3(orb));
                const v3 newA = pa + qrot(qa, rA);
                const v3 newB = pb + qrot(qb, rB);
                const float drift = smax(length(oldA - newA), length(oldB - newB));
                if (drift > S.P.stickDriftHeal) {
                    stick = 0;
                    lam = v3{lam.x, 0.0f, 0.0f};
                } else {
                    rA = ld3(ora); rB = ld3(orb);
                }
            }
            break;
        }
        const v3 xA = pa + qrot(qa, rA);
        const v3 xB = pb + qrot(qb, rB);
        v3 c0 = mul(B, xA - xB) + v3{COLLISION_MARGIN, 0, 0};
        if (c0.x < -S.P.maxPen) c0.x = -S.P.maxPen;
        float scale = 1.0f;
        if (S.P.adaptiveBeta && prevErr >= 0.0f) {
            const float err = smax(0.0f, COLLISION_MARGIN - c0.x);
            const float floor = S.P.adaptiveFloor;
            scale = beta;
            if (err <= floor) scale = smax(scale * 0.5f, 1.0f);
            else if (prevErr <= floor) scale = smin(scale * 2.0f, ADAPTIVE_BETA_MAX);
            else if (err > 0.5f * prevErr) scale = smin(scale * 2.0f, ADAPTIVE_BETA_MAX);
            else if (err < 0.1f * prevErr) scale = smax(scale * 0.5f, 1.0f);
        }
        if (c0.x > COLLISION_MARGIN + CONTACT_RELEASE_GAP) {
            lam = v3{0, 0, 0};
            stick = 0;
        }
        if constexpr (MODE == 2) {
            // manifold.cpp:338-347: an unmatched contact starts at 5 % of beta_lin
            if (prevErr < 0.0f) {
                const float seed = smax(PENALTY_MIN, 0.05f * S.P.betaLin);
                pen = v3{seed, seed, seed};
            }
        }
        lam = lam * S.P.alpha * S.P.gamma;
        pen = lclamp3(pen * S.P.gamma, PENALTY_MIN, pmax);
        const size_t c = d + crow(i);
        S.sRA[c] = st3(rA, scale);
        S.sRB[c] = st3(rB, __int_as_float(key));
        S.sC0[c] = st3(c0, stick ? 1.0f : 0.0f);
        S.sPen[c] = st3(pen);
        S.sLam[c] = st3(lam);
    }
    if constexpr (MODE == 2) {
        // manifold.cpp:359-378: a separated pair is parked for up to `grace` ticks, keeping its
        // contacts (in the slot's keep rows) with no force
        if (nn == 0) {
            if (S.mDormant[slot] < S.P.grace) {
                if (!revive)
                    for (int j = 0; j < oldN; j++) {
                        const uint32_t oc = ob + j * CSLICE;
                        v3 kp, kl;
                        loadPenLam(S, oc, kp, kl);
                        S.dRA[ds + j] = S.cRA[oc];
                        S.dRB[ds + j] = S.cRB[oc];
                        S.dC0[ds + j] = S.cC0[oc];
                        S.dPen[ds + j] = st3(kp);
                    }
                S.mDormant[slot] += 1;
                S.mSaved[slot] = liveCount;
                S.mNum[slot] = 0;
                storeRec(S, slot, B, a, b, q, mu, 0);
                S.keep[t] = 1u;
                return;
            }
            S.mSaved[slot] = 0;
        }
        S.mDormant[slot] = 0;
    }
    S.mNum[slot] = nn;
    storeRec(S, slot, B, a, b, q, mu, nn);
    S.keep[t] = nn > 0 ? 1u : 0u;
}

// joint.cpp:62-78: C0 at the step-start pose, warm-start decay; a broken joint is reaped
// (marked dead: it leaves the adjacency exactly as a deleted force leaves the list).
__global__ void k_joint_init(State S)
{
    const int j = GID;
    if (j >= S.jointCap) return;
    if (S.jWorld[j] < 0 || S.jDead[j]) return;
    if (S.jKind[j] == FORCE_GENERIC) { genericInit(S, j); return; }   // never breaks
    if (S.jKind[j] != FORCE_JOINT) return;   // springs and collision filters: nothing to warm
    if (S.jBroken[j]) { S.jDead[j] = 1; return; }
    S.jC0L[j] = st3(jointLin(S, j));
    S.jC0A[j] = st3(jointAng(S, j));
    S.jLamL[j] = st3(ld3(S.jLamL[j]) * S.P.alpha * S.P.gamma);
    S.jLamA[j] = st3(ld3(S.jLamA[j]) * S.P.alpha * S.P.gamma);
    const float pmax = S.P.penaltyMax;
    const v3 pl = ld3(S.jPenL[j]), pa = ld3(S.jPenA[j]);
    const float sl = S.jStiffL[j], sa = S.jStiffA[j];
    const float g = S.P.gamma;
    S.jPenL[j] = st3(v3{smin(sclamp(pl.x * g, PENALTY_MIN, pmax), sl), smin(sclamp(pl.y * g, PENALTY_MIN, pmax), sl),
                        smin(sclamp(pl.z * g, PENALTY_MIN, pmax), sl)});
    S.jPenA[j] = st3(v3{smin(sclamp(pa.x * g, PENALTY_MIN, pmax), sa), smin(sclamp(pa.y * g, PENALTY_MIN, pmax), sa),
                        smin(sclamp(pa.z * g, PENALTY_MIN, pmax), sa)});
}

// solver_solve.cpp:66-87: stable compaction of each world's live list, releasing the
// slots of separated pairs to the free list.  One workgroup per world walks its live prefix
// in chunks (the narrow phase set keep only there): a block scan of the kept and released
// flags (two 16-bit counts in one word) places each slot, kept ones in order in liveTmp and
// released ones after the world's free entries, then the kept run is copied back.  The
// world's slot range is this workgroup's alone.
__global__ void __launch_bounds__(TPB) k_compact_world(State S)
{
    const int w = blockIdx.x;
    if (S.wSlotCap[w] <= 0) return;
    const int base = S.wSlotBase[w], live = S.wLive[w], freeBase = S.wFree[w];
    __shared__ uint32_t sm[TPB / 64];
    int kept = 0, released = 0;
    for (int c = 0; c < live; c += TPB) {
        const int at = c + static_cast<int>(threadIdx.x);
        const bool in = at < live;
        const uint32_t k = in && S.keep[base + at] ? 1u : 0u;
        const uint32_t r = in && !k ? 1u : 0u;
        uint32_t total;
        const uint32_t ex = blockExclusiveScan<TPB>(k | (r << 16), sm, total);
        if (in) {
            const int slot = S.live[base + at];
            if (k) S.liveTmp[base + kept + static_cast<int>(ex & 0xFFFFu)] = slot;
            else S.freeList[base + freeBase + released + static_cast<int>(ex >> 16)] = slot;
        }
        kept += static_cast<int>(total & 0xFFFFu);
        released += static_cast<int>(total >> 16);
    }
    __threadfence_block();
    __syncthreads();
    for (int c = threadIdx.x; c < kept; c += TPB) S.live[base + c] = S.liveTmp[base + c];
    if (threadIdx.x == 0) {
        S.wLive[w] = kept;
        S.wFree[w] = freeBase + released;
    }
}

// ------------------------------------------------------------ adjacency ----
// manifold.cpp:123-166: CSR per body, joints in list order then manifolds in live order.
// Counts and fills use atomics; each segment is then sorted by its sequence number.
// Manifold entries are made for movable bodies only: no pass reads a static body's manifold
// entries (the solves, colouring, brake and projection gather movable bodies; the pair filter
// scans the movable side), and a ground body touched by hundreds of manifolds would otherwise
// serialise their atomics on one counter.
// Each entry's place in its body's list is the count its atomic took (adjPos: two per
// manifold view index, then two per joint after 2 * slotCap), so the fill writes without
// atomics; k_adj_sort then orders every list.
__global__ void k_adj_count(State S)
{
    int *jp = S.adjPos + 2 * static_cast<size_t>(S.slotCap);
    for (int t = GID; t < S.jointCap; t += gridDim.x * blockDim.x) {
        if (S.jWorld[t] >= 0 && !S.jDead[t]) {
            if (S.jA[t] >= 0) jp[2 * t] = atomicAdd(&S.adjCount[S.jA[t]], 1);
            jp[2 * t + 1] = atomicAdd(&S.adjCount[S.jB[t]], 1);
        }
    }
    FOR_LIVE(k) {
        const int4 h = S.mHdr[k];
        if (bMass(S, h.x) > 0.0f) S.adjPos[2 * k] = atomicAdd(&S.adjCount[h.x], 1);
        if (bMass(S, h.y) > 0.0f) S.adjPos[2 * k + 1] = atomicAdd(&S.adjCount[h.y], 1);
    }
}

__global__ void k_adj_fill(State S)
{
    const int *jp = S.adjPos + 2 * static_cast<size_t>(S.slotCap);
    for (int t = GID; t < S.jointCap; t += gridDim.x * blockDim.x) {
        const int w = S.jWorld[t];
        if (w >= 0 && !S.jDead[t]) {
            const unsigned long long seq = static_cast<unsigned long long>(t - S.wJointBase[w]);
            const unsigned long long e = (seq << 32) | static_cast<uint32_t>(-(t + 1));
            if (S.jA[t] >= 0) S.adjTmp[S.adjStart[S.jA[t]] + jp[2 * t]] = e;
            S.adjTmp[S.adjStart[S.jB[t]] + jp[2 * t + 1]] = e;
        }
    }
    FOR_LIVE(k) {
        // the live position orders a body's manifolds (a body's entries are all of one world)
        const unsigned long long seq = 0x40000000ull + static_cast<unsigned long long>(S.gOfK[k]);
        const unsigned long long e = (seq << 32) | static_cast<uint32_t>(k);
        const int4 h = S.mHdr[k];
        const int a = h.x, b = h.y;
        if (bMass(S, a) > 0.0f) S.adjTmp[S.adjStart[a] + S.adjPos[2 * k]] = e;
        if (bMass(S, b) > 0.0f) S.adjTmp[S.adjStart[b] + S.adjPos[2 * k + 1]] = e | static_cast<uint32_t>(ADJ_SIDE_B);
    }
}

// Sorts up to ADJ_REG keys in registers (odd-even transposition; unused places hold the
// largest key and stay last); the keys of one body are distinct.
constexpr int ADJ_REG = 8;
__global__ void k_adj_sort(State S)
{
    const int b = GID;
    if (b >= S.numBodies) return;
    const int e0 = S.adjStart[b], e1 = S.adjStart[b + 1];
    const bool ordered = bMass(S, b) > 0.0f;  // only movable bodies sum in this order
    if (ordered && e1 - e0 <= ADJ_REG) {
        unsigned long long v[ADJ_REG];
#pragma unroll
        for (int q = 0; q < ADJ_REG; q++) v[q] = e0 + q < e1 ? S.adjTmp[e0 + q] : ~0ull;
#pragma unroll
        for (int phase = 0; phase < ADJ_REG; phase++) {
#pragma unroll
            for (int q = phase & 1; q + 1 < ADJ_REG; q += 2) {
                const unsigned long long lo = v[q] < v[q + 1] ? v[q] : v[q + 1];
                const unsigned long long hi = v[q] < v[q + 1] ? v[q + 1] : v[q];
                v[q] = lo;
                v[q + 1] = hi;
            }
        }
#pragma unroll
        for (int q = 0; q < ADJ_REG; q++)
            if (e0 + q < e1) S.adj[e0 + q] = static_cast<int>(static_cast<uint32_t>(v[q] & 0xFFFFFFFFull));
        return;
    }
    if (ordered) {
        for (int a = e0 + 1; a < e1; a++) {
            const unsigned long long v = S.adjTmp[a];
            int c = a;
            while (c > e0 && S.adjTmp[c - 1] > v) { S.adjTmp[c] = S.adjTmp[c - 1]; c--; }
            S.adjTmp[c] = v;
        }
    }
    for (int a = e0; a < e1; a++) S.adj[a] = static_cast<int>(static_cast<uint32_t>(S.adjTmp[a] & 0xFFFFFFFFull));
}

// ------------------------------------------------------ update set, colours ----
// solver_solve.cpp:296-358: contactOnly = no force touches the body; worlds with a
// live joint (or with contact-only bodies kept out of the overflow) need colours.
__global__ void k_colour_prepare(State S)
{
    const int i = GID;
    if (i < S.numBodies) S.contactOnly[i] = 1;
    if (i < S.numWorlds) S.wHasJoints[i] = S.P.contactOnlyOverflow ? 0 : 1;
}
__global__ void k_colour_joints(State S)
{
    const int j = GID;
    if (j >= S.jointCap) return;
    if (S.jWorld[j] < 0 || S.jDead[j]) return;
    if (S.jA[j] >= 0) S.contactOnly[S.jA[j]] = 0;
    S.contactOnly[S.jB[j]] = 0;
    S.wHasJoints[S.jWorld[j]] = 1;
}
__global__ void k_colour_init(State S)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    const int w = S.bodyWorld[i];
    if (w < 0) return;
    S.colour[i] = (S.wHasJoints[w] && S.updateIndex[i] >= 0) ? 0xFF : 8;
}

// Jones-Plassmann round with priority = lower updateIndex: a body takes the smallest
// colour its lower-index neighbours do not use once all of them are coloured.  This is
// exactly the reference's sequential greedy pass (solver_solve.cpp:365-397).
__device__ __forceinline__ bool tryColour(const State &S, int i)
{
    const int ui = S.updateIndex[i];
    uint32_t used = 0;
    for (int e = S.adjStart[i]; e < S.adjStart[i + 1]; e++) {
        const int o = otherOf(S, S.adj[e], i);
        if (o < 0) continue;
        const int uo = S.updateIndex[o];
        if (uo < 0 || uo >= ui) continue;
        const uint8_t c = __hip_atomic_load(&S.colour[o], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c == 0xFF) return false;
        if (c < 8) used |= 1u << c;
    }
    uint8_t c = 8;
    for (int k = 0; k < 8; k++)
        if (!(used & (1u << k))) { c = static_cast<uint8_t>(k); break; }
    __hip_atomic_store(&S.colour[i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}
// `count`: the last round also flags the worlds that still have uncoloured bodies (a plain
// store; many bodies of one world may be left).
__global__ void k_colour_round(State S, int count)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    if (S.colour[i] != 0xFF) return;
    if (!tryColour(S, i) && count) S.wColourLeft[S.bodyWorld[i]] = 1;
}
// Bodies left after the bounded rounds (worlds flagged by the last round) are finished by
// dependency counting: a body is ready once every lower-ordered neighbour has its colour, and
// ready bodies are coloured level by level (Kahn) in one workgroup per world.  Every body then
// takes the colour the sequential greedy pass gives it (bodies of one level are never
// neighbours, so the order inside a level does not matter).  The counts are per adjacency entry
// (entries are symmetric between the two bodies), so each entry of a newly coloured body
// releases one count of its higher-ordered neighbour.  The bodies ready from the start are
// queued here (wColourTail cleared before), so the finisher does not scan its world.
__global__ void k_colour_counts(State S)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    const int w = S.bodyWorld[i];
    if (w < 0 || !S.wColourLeft[w] || S.colour[i] != 0xFF) return;
    const int ui = S.updateIndex[i];
    int cnt = 0;
    for (int e = S.adjStart[i]; e < S.adjStart[i + 1]; e++) {
        const int o = otherOf(S, S.adj[e], i);
        if (o < 0) continue;
        const int uo = S.updateIndex[o];
        if (uo < 0 || uo >= ui) continue;
        if (S.colour[o] == 0xFF) cnt++;
    }
    S.colourCnt[i] = cnt;
    if (cnt == 0) S.colourQueue[S.wBodyBase[w] + atomicAdd(S.wColourTail + w, 1)] = i;
}

constexpr int COLOUR_TPB = 256;
__global__ void __launch_bounds__(COLOUR_TPB) k_colour_finish(State S)
{
    const int w = S.wOrder[blockIdx.x];
    if (!S.wHasJoints[w] || !S.wColourLeft[w]) return;
    int *queue = S.colourQueue + S.wBodyBase[w];
    __shared__ int tail, head;
    if (threadIdx.x == 0) { tail = S.wColourTail[w]; head = 0; }
    __syncthreads();
    while (true) {
        const int h = head, e = tail;
        __syncthreads();   // every thread holds this level's bounds before the next is appended
        if (h == e) break;
        for (int k = h + threadIdx.x; k < e; k += COLOUR_TPB) {
            const int i = queue[k];
            tryColour(S, i);
            const int ui = S.updateIndex[i];
            for (int a = S.adjStart[i]; a < S.adjStart[i + 1]; a++) {
                const int o = otherOf(S, S.adj[a], i);
                if (o < 0 || S.updateIndex[o] <= ui) continue;
                if (__hip_atomic_load(&S.colour[o], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0xFF) continue;
                if (atomicSub(&S.colourCnt[o], 1) == 1) queue[atomicAdd(&tail, 1)] = o;
            }
        }
        __threadfence_block();
        __syncthreads();
        if (threadIdx.x == 0) head = e;
        __syncthreads();
    }
}
// solver_solve.cpp:399-418: contact-only bodies form the Jacobi bucket (colour 8).  A
// body whose greedy colour overflowed but touches a joint lands in neither the colour runs
// nor the bucket, so the reference never solves it; it is marked 9 here and skipped too.
__global__ void k_colour_override(State S)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    if (S.bodyWorld[i] < 0) return;
    if (S.P.contactOnlyOverflow && S.contactOnly[i]) S.colour[i] = 8;
    else if (S.colour[i] == 8) S.colour[i] = 9;
}

// ------------------------------------------------------------ velocity brake ----
// solver_projection.cpp:96-167 walks the active manifolds in live order and changes body
// velocities in place.  A manifold depends only on the earlier active manifolds that
// share a movable body with it, so running it after exactly those (level by level) gives
// the reference's result.  succ* link each manifold to the next active manifold of each
// movable body (that body's adjacency is in live order).
__device__ __forceinline__ bool bodyMoves(const State &S, int b) { return bMass(S, b) > 0.0f && S.updateIndex[b] >= 0; }

__device__ __forceinline__ bool manifoldActive(const State &S, int k)
{
    const int4 h = S.mHdr[k];
    return h.w > 0 && (bodyMoves(S, h.x) || bodyMoves(S, h.y));
}

// Separation of one contact along the normal at the current poses
// (solver_projection.cpp:120-123).
__device__ __forceinline__ float contactGap(const State &S, v3 n, v3 pa, qt qa, v3 pb, qt qb, uint32_t c, v3 &rA, v3 &rB)
{
    rA = qrot(qa, ld3(S.cRA[c]));
    rB = qrot(qb, ld3(S.cRB[c]));
    return dot(n, (pa + rA) - (pb + rB));
}

// A manifold can change a velocity in the brake only if it is active and one of its
// contacts has a positive gap; gaps depend on poses alone, which the brake leaves alone,
// so the others are left out of the dependency links altogether.  `count` also counts the
// world's active manifolds (used by the projection).
__global__ void k_brake_flags(State S, int count)
{
    FOR_LIVE(k) {
        const int4 h = S.mHdr[k];
        const float ma = bMass(S, h.x), mb = bMass(S, h.y);
        const int ua = S.updateIndex[h.x], ub = S.updateIndex[h.y];
        const v3 n = ld3(S.mBasisK[3 * k]);
        const v3 pa = bPos(S, h.x), pb = bPos(S, h.y);
        const qt qa = bRot(S, h.x), qb = bRot(S, h.y);
        const bool active = h.w > 0 && ((ma > 0.0f && ua >= 0) || (mb > 0.0f && ub >= 0));
        bool on = false;
        if (active) {
            if (count) S.wUpdateCount[S.bodyWorld[h.x]] = 1;   // the world has an active manifold
            for (int i = 0; i < h.w && !on; i++) {
                v3 rA, rB;
                on = contactGap(S, n, pa, qa, pb, qb, static_cast<uint32_t>(h.z) + i * CSLICE, rA, rB) > 0.0f;
            }
        }
        S.brakeOn[k] = on ? 1 : 0;
        S.succA[k] = -1;
        S.succB[k] = -1;
        S.indeg[k] = on ? 0 : -1;
    }
}

// Each movable body walks its adjacency once (manifolds in live order after its joints) and
// links each braking manifold to the next one of that body: succA / succB by the side the body
// is on, one count of the later manifold's in-degree per link (both bodies of a manifold may
// add one).  k_brake_flags has set succ* to -1 and indeg to 0 (braking) or -1 (not).
__global__ void k_brake_links(State S)
{
    const int i = XGID;
    if (i >= S.numBodies || S.bodyWorld[i] < 0 || !bodyMoves(S, i)) return;
    int prev = -1;
    bool prevB = false;
    const int e1 = S.adjStart[i + 1];
    for (int e = S.adjStart[i]; e < e1; e++) {
        const int r = S.adj[e];
        if (r < 0) continue;
        const int k = adjManifold(r);
        if (!S.brakeOn[k]) continue;
        if (prev >= 0) {
            if (prevB) S.succB[prev] = k;
            else S.succA[prev] = k;
            atomicAdd(S.indeg + k, 1);
        }
        prev = k;
        prevB = adjSideB(r);
    }
}

__device__ void brakeManifold(const State &S, int k, float window)
{
    const int4 hd = S.mHdr[k];
    const int a = hd.x, b = hd.y;
    const bool moveA = bodyMoves(S, a), moveB = bodyMoves(S, b);
    const float ma = bMass(S, a), mb = bMass(S, b);
    const float imA = moveA ? 1.0f / ma : 0.0f, imB = moveB ? 1.0f / mb : 0.0f;
    const v3 n = ld3(S.mBasisK[3 * k]);
    const v3 pa = bPos(S, a), pb = bPos(S, b);
    const qt qa = bRot(S, a), qb = bRot(S, b);
    const v3 va = ld3(S.vel[a]), wa = ld3(S.angVel[a]), vb = ld3(S.vel[b]), wb = ld3(S.angVel[b]);
    v3 pRA{0, 0, 0}, pRB{0, 0, 0};
    float verr = 0.0f;
    int braked = 0;
    const int nc = hd.w;
    for (int i = 0; i < nc; i++) {
        v3 rA, rB;
        const float gap = contactGap(S, n, pa, qa, pb, qb, static_cast<uint32_t>(hd.z) + i * CSLICE, rA, rB);
        if (!(gap > 0.0f)) continue;
        const v3 vAp = va + cross(wa, rA), vBp = vb + cross(wb, rB);
        const float vn = dot(vAp - vBp, n);
        const float target = -gap / window;
        if (vn < target) {
            pRA += rA;
            pRB += rB;
            verr += target - vn;
            braked++;
        }
    }
    if (braked == 0) return;
    const float invN = 1.0f / static_cast<float>(braked);
    const v3 IA = bMoment(S, a), IB = bMoment(S, b);
    const v3 tA = moveA ? qrot(qconj(qa), cross(pRA * invN, n)) : v3{0, 0, 0};
    const v3 tB = moveB ? qrot(qconj(qb), cross(pRB * invN, n)) : v3{0, 0, 0};
    const float K = imA + imB + (moveA ? angCompliance(IA, tA) : 0.0f) + (moveB ? angCompliance(IB, tB) : 0.0f);
    const float j = verr * invN / K;
    if (moveA) {
        v3 dv = n * (imA * j), dw = angResponse(IA, qa, tA) * j;
        maskLocked(S.lockLin[a], dv);
        maskLocked(S.lockAng[a], dw);
        S.vel[a] = st3(va + dv);
        S.angVel[a] = st3(wa + dw);
    }
    if (moveB) {
        v3 dv = n * (imB * j), dw = angResponse(IB, qb, tB) * j;
        maskLocked(S.lockLin[b], dv);
        maskLocked(S.lockAng[b], dw);
        S.vel[b] = st3(vb - dv);
        S.angVel[b] = st3(wb - dw);
    }
}

// One workgroup per world: Kahn's algorithm by levels over the dependency links,
// braking each level's manifolds in parallel; the level order is kept for the replay.
// levelStart[base + l] is the first entry of level l; wLevelEnd[w] ends the last level.
constexpr int BRAKE_TPB = 256;
__global__ void __launch_bounds__(BRAKE_TPB) k_brake_levels(State S, float window, int doBrake)
{
    const int w = S.wOrder[blockIdx.x];
    const int base = static_cast<int>(S.wLiveOff[w]);
    const int nLive = S.wLive[w];
    __shared__ int tail, head, levels;
    if (threadIdx.x == 0) { tail = 0; head = 0; levels = 0; }
    __syncthreads();
    if (nLive <= 0 || S.wSlotCap[w] <= 0) {
        if (threadIdx.x == 0) { S.brakeLevels[w] = 0; S.wLevelEnd[w] = 0; }
        return;
    }
    for (int k = threadIdx.x; k < nLive; k += BRAKE_TPB)
        if (S.indeg[base + k] == 0) S.levelOrder[base + atomicAdd(&tail, 1)] = base + k;
    __syncthreads();
    while (true) {
        const int h = head, e = tail;
        // every thread holds this level's bounds before any thread appends to the next
        // level, so all threads see the same h == e and leave the loop together
        __syncthreads();
        if (h == e) break;
        if (threadIdx.x == 0) S.levelStart[base + levels] = h;
        for (int k = h + threadIdx.x; k < e; k += BRAKE_TPB) {
            const int t = S.levelOrder[base + k];
            if (doBrake) brakeManifold(S, t, window);
            const int sa = S.succA[t], sb = S.succB[t];
            if (sa >= 0 && atomicSub(&S.indeg[sa], 1) == 1) S.levelOrder[base + atomicAdd(&tail, 1)] = sa;
            if (sb >= 0 && atomicSub(&S.indeg[sb], 1) == 1) S.levelOrder[base + atomicAdd(&tail, 1)] = sb;
        }
        __threadfence_block();
        __syncthreads();
        if (threadIdx.x == 0) { head = e; levels++; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        S.brakeLevels[w] = levels;
        S.wLevelEnd[w] = tail;
    }
}

// ------------------------------------------------------------ body phases ----
// A body's pose record (see State::poseRec) at pose (p, q): its displacement from the step's
// initial pose, its awake mass, the rotation and the rotation difference.
__device__ __forceinline__ void writePose(const State &S, int i, v3 p, qt q)
{
    S.poseRec[4 * i] = st3(p - ld3(S.initLin[i]), S.P.sleepOn && S.sleeping[i] ? 0.0f : bMass(S, i));
    S.poseRec[4 * i + 1] = stq(q);
    S.poseRec[4 * i + 2] = st3(qsub(q, ldq(S.initAng[i])));
}

// solver_solve.cpp:109-142
__global__ void k_warmstart(State S)
{
    const int i = XGID;
    if (i >= S.numBodies || S.bodyWorld[i] < 0) return;
    const float dt = S.P.dt;
    v3 vel = ld3(S.vel[i]), ang = ld3(S.angVel[i]);
    const uint8_t ll = S.lockLin[i];
    maskLocked(ll, vel);
    maskLocked(S.lockAng[i], ang);
    S.vel[i] = st3(vel);
    S.angVel[i] = st3(ang);
    const float g = (ll & 0x4) ? 0.0f : S.grav[i];
    const float m = bMass(S, i);
    const v3 p = bPos(S, i);
    const qt q = bRot(S, i);
    v3 inl = p + vel * dt;
    if (m > 0) inl += v3{0, 0, g} * (dt * dt);
    S.inertLin[i] = st3(inl);
    S.inertAng[i] = stq(qadd(q, ang * dt));
    const v3 accel = (vel - ld3(S.prevVel[i])) / dt;
    const float accelExt = accel.z * __builtin_copysignf(1.0f, g);
    float wgt = sclamp(accelExt / fabsf(g), 0.0f, 1.0f);
    if (!__builtin_isfinite(wgt)) wgt = 0.0f;
    S.initLin[i] = st3(p);
    S.initAng[i] = stq(q);
    v3 np = p;
    qt nq = q;
    if (m > 0) {
        np = (p + vel * dt) + v3{0, 0, g} * (wgt * dt * dt);
        nq = qadd(q, ang * dt);
        S.pos[i] = st3(np);
        S.rot[i] = stq(nq);
    }
    // the substep's first pose records (moves in the rounds refresh their bodies')
    S.poseRec[4 * i] = st3(np - p, S.P.sleepOn && S.sleeping[i] ? 0.0f : m);
    S.poseRec[4 * i + 1] = stq(nq);
    S.poseRec[4 * i + 2] = st3(qsub(nq, q));
}

// Pose records for the per-manifold passes (the displacement and rotation difference are the
// primal/dual bracket terms, manifold.cpp:465-501, computed once per body instead of once per
// manifold side; same floats).
// mass slot: the body's mass while it is awake, 0 while it sleeps (the manifold passes treat a
// sleeper as they treat a static body: manifold.cpp:168-172)

// solver_step.cpp:268-273 for one body: move by the relaxed Jacobi delta.
__device__ __forceinline__ void jacobiApply(const State &S, int i, v3 dl, v3 da, bool pose)
{
    const float r = S.P.jacobiRelax;
    const v3 p = bPos(S, i) + dl * r;
    const qt q = qadd(bRot(S, i), da * r);
    S.pos[i] = st3(p);
    S.rot[i] = stq(q);
    if (pose) writePose(S, i, p, q);
}

// solver_solve.cpp:146-176 (sleep off)
// solver_solve.cpp:143-176: BDF1 velocity (a sleeper keeps its own), locks, then on the tick's
// last substep (`idle`) the idle count and falling asleep.
__global__ void k_finish(State S, int idle)
{
    const int i = XGID;
    if (i >= S.numBodies || S.bodyWorld[i] < 0) return;
    const float dt = S.P.dt;
    v3 vel = ld3(S.vel[i]), ang = ld3(S.angVel[i]);
    S.prevVel[i] = st3(vel);
    const bool asleep = S.P.sleepOn && S.sleeping[i];
    const bool movable = bMass(S, i) > 0;
    if (movable && !asleep) {
        vel = (bPos(S, i) - ld3(S.initLin[i])) / dt;
        ang = qsub(bRot(S, i), ldq(S.initAng[i])) / dt;
    }
    maskLocked(S.lockLin[i], vel);
    maskLocked(S.lockAng[i], ang);
    if (idle && S.P.sleepOn && S.P.sleepFrames > 0 && movable && !asleep && S.sleepMode[i] == AVBD_SLEEP_IDLE) {
        const float lin2 = S.P.sleepLin * S.P.sleepLin, ang2 = S.P.sleepAng * S.P.sleepAng;
        int still = S.stillFrames[i];
        if (length2(vel) < lin2 && length2(ang) < ang2) ++still;
        else still = 0;
        if (still >= S.P.sleepFrames) {
            S.sleeping[i] = 1;
            vel = v3{0, 0, 0};
            ang = v3{0, 0, 0};
        }
        S.stillFrames[i] = still;
    }
    S.vel[i] = st3(vel);
    S.angVel[i] = st3(ang);
    if (movable) {
        const v3 p = bPos(S, i);
        const float4 q = S.rot[i];
        const float v[13] = {p.x, p.y, p.z, q.x, q.y, q.z, q.w, vel.x, vel.y, vel.z, ang.x, ang.y, ang.z};
        bool finite = true;
        for (int c = 0; c < 13; c++) finite = finite && isfinite(v[c]);
        if (!finite) atomicOr(S.wErr + S.bodyWorld[i], AVBD_WORLD_NONFINITE);
    }
}

// ------------------------------------------------------------ sleeping ----
// solver_broadphase.cpp:67-80 with :153-161: for each pair about to get its first manifold, in
// pair order, a sleeper whose bounds touch an awake movable body wakes.  Serial per world (one
// thread each): a body woken by one pair can wake the next.
__device__ __forceinline__ void wakeBody(const State &S, int a)
{
    S.sleeping[a] = 0;
    S.stillFrames[a] = 0;
    S.narrowRec[NARROW_REC * static_cast<size_t>(a) + 5].w = bMass(S, a);   // awake mass for the narrow phase
}
__device__ __forceinline__ void wakePair(const State &S, int a, int b)
{
    if (S.sleeping[a] && !S.sleeping[b] && bMass(S, b) > 0.0f) wakeBody(S, a);
    else if (S.sleeping[b] && !S.sleeping[a] && bMass(S, a) > 0.0f) wakeBody(S, b);
}
__global__ void k_wake_pairs(State S)
{
    const int w = GID;
    if (w >= S.numWorlds || S.wBodyCount[w] <= 0) return;
    const int base = S.wBodyBase[w];
    const uint32_t p0 = S.pairOff[base], p1 = S.pairOff[base + S.wBodyCount[w]];
    for (uint32_t p = p0; p < p1; p++)
        if (S.pairValid[p] || S.pairApart[p] || (S.P.grace > 0 && S.pairParked[p]))
            wakePair(S, S.pairI[p], S.pairJ[p]);
}
// solver_solve.cpp:264-289: with idle detection on, a sleeper (mode IDLE) sharing a force with
// an awake movable body that moves faster than the thresholds wakes; bodies in list order,
// serially per world.  (The island pass that follows in the reference only ever re-sleeps
// islands whose members all sleep already, so it changes nothing and is not run.)
__global__ void k_wake_neighbours(State S)
{
    const int w = GID;
    if (w >= S.numWorlds || S.P.sleepFrames <= 0) return;
    const int base = S.wBodyBase[w], cnt = S.wBodyCount[w];
    const float lin2 = S.P.sleepLin * S.P.sleepLin, ang2 = S.P.sleepAng * S.P.sleepAng;
    for (int i = base; i < base + cnt; i++) {
        if (!S.sleeping[i] || S.sleepMode[i] != AVBD_SLEEP_IDLE) continue;
        for (int e = S.adjStart[i]; e < S.adjStart[i + 1]; e++) {
            const int o = otherOf(S, S.adj[e], i);
            if (o < 0) continue;
            if (bMass(S, o) > 0.0f && !S.sleeping[o] &&
                (length2(ld3(S.vel[o])) > lin2 || length2(ld3(S.angVel[o])) > ang2)) {
                S.sleeping[i] = 0;
                S.stillFrames[i] = 0;
                break;
            }
        }
    }
}
// solver_solve.cpp:296-339: the update set (movable bodies that are awake; a START_ASLEEP body
// at rest with no force stays out, else it wakes) ...
__global__ void k_update_flags(State S)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    uint32_t f = 0;
    if (S.bodyWorld[i] >= 0 && bMass(S, i) > 0.0f) {
        f = 1;
        const uint8_t mode = S.sleepMode[i];
        if (mode == AVBD_SLEEP_IDLE) {
            if (S.sleeping[i]) f = 0;
        } else if (mode == AVBD_SLEEP_START_ASLEEP) {
            if (!S.sleeping[i] && length2(ld3(S.vel[i])) == 0.0f && S.adjStart[i + 1] == S.adjStart[i]) f = 0;
            else S.sleeping[i] = 0;
        }
    }
    S.updFlag[i] = f;
}
// ... ranked in list order within each world
__global__ void k_update_index(State S)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    const int w = S.bodyWorld[i];
    if (w < 0) return;
    S.updateIndex[i] = S.updFlag[i] ? static_cast<int>(S.updScan[i] - S.updScan[S.wBodyBase[w]]) : -1;
}

__device__ __forceinline__ void queueFallback(const State &S, int body, const Block &blk)
{
    const int k = atomicAdd(S.fbCount, 1);
    S.fbBody[k] = body;
    storeBlock(S.fbBlock + 33 * k, blk);
}

// Sort keys of the colour runs: a body's colour (0..9), 255 outside the update set.
__global__ void k_colour_keys(State S)
{
    const int i = GID;
    if (i >= S.numBodies) return;
    const bool in = S.bodyWorld[i] >= 0 && S.updateIndex[i] >= 0;
    S.colKey[i] = in ? static_cast<unsigned long long>(S.colour[i]) : 255ull;
    S.colVal[i] = i;
}
// Run starts from the pass's exclusive digit offsets (digit d of tile 0 = keys below d).
__global__ void k_colour_starts(State S, const uint32_t *offs, int nb)
{
    const int d = threadIdx.x;
    if (d < 16) S.colStart[d] = offs[d * nb];
}

// solver_step.cpp:209-229: one colour of the Gauss-Seidel sweep, over that colour's run.
// The pivoted solve (rare) runs in place: the next colour reads this one's poses.  ARMS: arm
// reuse is on; `cachedRound` = a round after the substep's first.
template <bool ARMS>
__global__ void __launch_bounds__(TPB) k_colour_pass(State S, int colour, int cachedRound)
{
    const uint32_t begin = S.colStart[colour], end = S.colStart[colour + 1];
    for (uint32_t r = begin + GID; r < end; r += gridDim.x * blockDim.x) {
        const int i = S.colValAlt[r];
        Block blk;
        assemble<ARMS>(S, i, blk, cachedRound != 0, colour);
        v3 dl, da;
        if (!solveDeltaFast(blk, dl, da)) solveDeltaSlow(blk, dl, da);
        const v3 np = bPos(S, i) + dl;
        const qt nq = qadd(bRot(S, i), da);
        S.pos[i] = st3(np);
        S.rot[i] = stq(nq);
        writePose(S, i, np, nq);
    }
}

// The colour pass with four lanes per body: lane s of a body's quad takes adjacency entries
// s, s + 4, ... (their loads are in flight together instead of one entry after another), and
// the quad's first lane adds the entries' terms to the block in adjacency order, so the block
// holds the floats the one-lane assembly makes.  Springs and generic joints (flag 8) are
// stamped by the first lane itself at their place in the order.
AD float shflF(float v, int src) { return __shfl(v, src); }
AD v3 shfl3(v3 v, int src) { return v3{shflF(v.x, src), shflF(v.y, src), shflF(v.z, src)}; }
AD m3 shflM(const m3 &m, int src) { return m3{{shfl3(m.c[0], src), shfl3(m.c[1], src), shfl3(m.c[2], src)}}; }
AD EntryTerms shflTerms(const EntryTerms &t, int src)
{
    EntryTerms u;
    u.flags = __shfl(t.flags, src);
    u.lin = shflM(t.lin, src);
    u.ang = shflM(t.ang, src);
    u.cross = shflM(t.cross, src);
    u.ang2 = shflM(t.ang2, src);
    u.rl = shfl3(t.rl, src);
    u.ra = shfl3(t.ra, src);
    u.angDiag = shfl3(t.angDiag, src);
    u.ra2 = shfl3(t.ra2, src);
    return u;
}
constexpr int QUAD_TPB = 256;
template <bool ARMS>
__global__ void __launch_bounds__(QUAD_TPB) k_colour_pass_quad(State S, int colour, int cachedRound)
{
    const uint32_t begin = S.colStart[colour], end = S.colStart[colour + 1];
    const int sub = threadIdx.x & 3;
    const int lead = (threadIdx.x & 63) & ~3;
    const uint32_t quads = (gridDim.x * blockDim.x) >> 2;
    for (uint32_t r = begin + (GID >> 2); r < end; r += quads) {
        const int i = S.colValAlt[r];
        const int e0 = S.adjStart[i], e1 = S.adjStart[i + 1];
        Block blk;
        if (sub == 0) makeBlock(S, i, blk);
        for (int c = e0; c < e1; c += 4) {
            EntryTerms t;
            t.flags = 0;
            const int e = c + sub;
            if (e < e1) {
                const int ref = S.adj[e];
                if (ref >= 0) {
                    manifoldTerms<ARMS>(S, adjManifold(ref), i, t, cachedRound != 0, colour);
                } else {
                    const uint8_t kind = S.jKind[-ref - 1];
                    if (kind == FORCE_JOINT) jointTerms(S, -ref - 1, i, t);
                    else if (kind != FORCE_IGNORE) t.flags = 8;
                }
            }
            const int n = min(4, e1 - c);
            for (int q = 0; q < n; q++) {
                const EntryTerms u = shflTerms(t, lead + q);
                if (sub == 0) {
                    if (u.flags & 8) forcePrimal(S, -S.adj[c + q] - 1, i, blk);
                    else applyTerms(blk, u);
                }
            }
        }
        if (sub == 0) {
            v3 dl, da;
            if (!solveDeltaFast(blk, dl, da)) solveDeltaSlow(blk, dl, da);
            const v3 np = bPos(S, i) + dl;
            const qt nq = qadd(bRot(S, i), da);
            S.pos[i] = st3(np);
            S.rot[i] = stq(nq);
            writePose(S, i, np, nq);
        }
    }
}

// solver_step.cpp:248-264: overflow bodies solve against one pose snapshot.  APPLY (worlds
// without joints): a body's block reads no other body's pose, so each body moves at once
// (solver_step.cpp:268-273) and refreshes its pose record; otherwise the delta is kept for
// k_overflow_apply.
template <bool APPLY>
__global__ void __launch_bounds__(TPB) k_overflow_solve(State S)
{
    const int i = XGID;
    if (i >= S.numBodies || S.colour[i] != 8 || S.bodyWorld[i] < 0) return;
    if (S.updateIndex[i] < 0) return;
    Block blk;
    assembleStamped(S, i, blk);
    v3 dl, da;
    if (!solveDeltaFast(blk, dl, da)) { queueFallback(S, i, blk); return; }
    if (APPLY) {
        jacobiApply(S, i, dl, da, true);
    } else {
        S.dLin[i] = st3(dl);
        S.dAng[i] = st3(da);
    }
}

// Pivoted solves of the queued blocks (maths.h:202-343), then the queue is emptied.  One
// workgroup; `apply`: 1 = colour pass (move the body now), 0 = Jacobi (store the delta),
// 2 = Jacobi moving the body now (with its pose record).
__global__ void k_solve_fallback(State S, int apply)
{
    const int n = *S.fbCount;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int i = S.fbBody[k];
        const Block blk = loadBlock(S.fbBlock + 33 * k);
        v3 dl, da;
        solveDeltaSlow(blk, dl, da);
        if (apply == 2) {
            jacobiApply(S, i, dl, da, true);
        } else if (apply) {
            S.pos[i] = st3(bPos(S, i) + dl);
            S.rot[i] = stq(qadd(bRot(S, i), da));
        } else {
            S.dLin[i] = st3(dl);
            S.dAng[i] = st3(da);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *S.fbCount = 0;
}

// solver_step.cpp:268-273
__global__ void k_overflow_apply(State S)
{
    const int i = XGID;
    if (i >= S.numBodies || S.colour[i] != 8 || S.bodyWorld[i] < 0) return;
    if (S.updateIndex[i] < 0) return;
    jacobiApply(S, i, ld3(S.dLin[i]), ld3(S.dAng[i]), true);
}

// Stamps of every live manifold that touches a movable body, at the current poses.
template <bool ARMS>
__global__ void __launch_bounds__(TPB) k_stamp(State S)
{
    __shared__ float4 lds[TPB / 64][STAMP_LDS];
    const int lane = threadIdx.x & 63;
    FOR_LIVE(k) {
        StampAcc acc;
        manifoldStamps<ARMS>(S, k, acc);
        storeStampsWave(S, k, acc, lds[threadIdx.x >> 6], lane);
    }
}

// solver_step.cpp:277-296: manifolds (not frozen), then joints.  STAMP: also build the next
// Jacobi iteration's stamps.
// RESID: keep each constraint value the dual computes (the step's last round; host reads).
template <bool STAMP, bool ARMS, bool RESID = false>
__global__ void __launch_bounds__(TPB) k_dual(State S, float decay)
{
    __shared__ float4 lds[STAMP ? TPB / 64 : 1][STAMP ? STAMP_LDS : 1];
    const int lane = threadIdx.x & 63;
    FOR_LIVE(k) {
        StampAcc acc;
        manifoldDual<STAMP, ARMS, RESID>(S, k, decay, acc);
        if (STAMP) storeStampsWave(S, k, acc, lds[threadIdx.x >> 6], lane);
    }
    for (int t = GID; t < S.jointCap; t += gridDim.x * blockDim.x)
        if (S.jWorld[t] >= 0 && !S.jDead[t]) jointDual(S, t, RESID ? S.resJ + t : nullptr);
}

// ------------------------------------------------------------ projection ----
__global__ void k_proj_reset(State S)
{
    const int w = GID;
    if (w >= S.numWorlds) return;
    S.wMaxDepth[w] = 0u;
    S.wProjDone[w] = (S.wUpdateCount[w] > 0) ? 0 : 1;  // no active manifold: no pass
}

// solver_projection.cpp:642-709: per active manifold, patch depth beyond the slop and the
// per-side corrections (applied by the body gather below).  The world's largest average
// depth is a max over its manifolds: each wavefront reduces its lanes per world (the view is
// world after world, so a world's lanes are one run) and one lane per run takes the atomic,
// instead of every manifold contending for the world's counter.
__device__ float projManifold(const State &S, int k, int &w);
__global__ void __launch_bounds__(TPB) k_proj_manifold(State S)
{
    FOR_LIVE(k) {
        int w;
        float d = projManifold(S, k, w);
        // FOR_RANGE's active lanes are a prefix of the wavefront, so every lane read below is
        // an active one
        const int lane = static_cast<int>(threadIdx.x) & 63;
        const int active = __popcll(__ballot(1));
        for (int off = 1; off < 64; off <<= 1) {
            const float od = __shfl_up(d, off);
            const int ow = __shfl_up(w, off);
            if (lane >= off && ow == w) d = fmaxf(d, od);
        }
        const int nw = __shfl_down(w, 1);
        if ((lane == active - 1 || nw != w) && d > 0.0f) atomicMax(&S.wMaxDepth[w], __float_as_uint(d));
    }
}
__device__ float projManifold(const State &S, int k, int &w)
{
    const int4 hd = S.mHdr[k];
    const int a = hd.x, b = hd.y;
    w = S.bodyWorld[a];
    const float ma = bMass(S, a), mb = bMass(S, b);
    const int ua = S.updateIndex[a], ub = S.updateIndex[b];
    const v3 n = ld3(S.mBasisK[3 * k]);
    const v3 pa = bPos(S, a), pb = bPos(S, b);
    const qt qa = bRot(S, a), qb = bRot(S, b);
    if (S.wProjDone[w]) return 0.0f;
    S.projOn[k] = 0;
    const bool moveA = ma > 0.0f && ua >= 0, moveB = mb > 0.0f && ub >= 0;
    if (!(hd.w > 0 && (moveA || moveB))) return 0.0f;
    const float imA = moveA ? 1.0f / ma : 0.0f, imB = moveB ? 1.0f / mb : 0.0f;
    const m3 rotA = qmat(qa), rotB = qmat(qb);
    v3 pRA{0, 0, 0}, pRB{0, 0, 0};
    float depthSum = 0.0f;
    int corr = 0;
    const int nc = hd.w;
    for (int i = 0; i < nc; i++) {
        const uint32_t c = static_cast<uint32_t>(hd.z) + i * CSLICE;
        const v3 rA = mul(rotA, ld3(S.cRA[c])), rB = mul(rotB, ld3(S.cRB[c]));
        const float raw = -dot(n, (pa + rA) - (pb + rB));
        const float depth = raw - S.P.projSlop;
        if (depth > 0.0f) {
            pRA += rA;
            pRB += rB;
            depthSum += depth;
            corr++;
        }
    }
    if (corr == 0) return 0.0f;
    const float invN = 1.0f / static_cast<float>(corr);
    const float avg = depthSum * invN;
    const v3 IA = bMoment(S, a), IB = bMoment(S, b);
    const v3 tA = moveA ? qrot(qconj(qa), cross(pRA * invN, n)) : v3{0, 0, 0};
    const v3 tB = moveB ? qrot(qconj(qb), cross(pRB * invN, n)) : v3{0, 0, 0};
    const float K = imA + imB + (moveA ? angCompliance(IA, tA) : 0.0f) + (moveB ? angCompliance(IB, tB) : 0.0f);
    const float dp = S.P.projBeta * depthSum * invN / K;
    if (moveA) {
        S.projLA[k] = st3(n * (imA * dp));
        S.projAA[k] = st3(angResponse(IA, qa, tA) * dp);
    }
    if (moveB) {
        S.projLB[k] = st3(n * (imB * dp));
        S.projAB[k] = st3(angResponse(IB, qb, tB) * dp);
    }
    S.projOn[k] = 1;
    return avg > 0.0f ? avg : 0.0f;
}

// solver_projection.cpp:711-730: each movable body averages its corrections, summed in
// live order, and moves; every body of a world with active manifolds is renormalised.
__global__ void k_proj_body(State S)
{
    const int i = XGID;
    if (i >= S.numBodies) return;
    const int w = S.bodyWorld[i];
    if (w < 0 || S.updateIndex[i] < 0) return;
    if (S.wProjDone[w]) return;
    v3 sl{0, 0, 0}, sa{0, 0, 0};
    int cnt = 0;
    for (int e = S.adjStart[i]; e < S.adjStart[i + 1]; e++) {
        const int ref = S.adj[e];
        if (ref < 0) continue;
        const int m = adjManifold(ref);
        if (!S.projOn[m]) continue;
        cnt++;
        if (!adjSideB(ref)) {
            sl += ld3(S.projLA[m]);
            sa += ld3(S.projAA[m]);
        } else {
            sl -= ld3(S.projLB[m]);
            sa -= ld3(S.projAB[m]);
        }
    }
    const float inv = 1.0f / static_cast<float>(cnt > 1 ? cnt : 1);
    v3 dl = sl * inv, da = sa * inv;
    maskLocked(S.lockLin[i], dl);
    maskLocked(S.lockAng[i], da);
    const float cap = S.P.projMaxCorr;
    if (cap > 0.0f) {
        const float moved = length(dl);
        if (moved > cap) dl = dl * (cap / moved);
        const float turned = length(da);
        if (turned > cap) da = da * (cap / turned);
    }
    S.pos[i] = st3(bPos(S, i) + dl);
    S.rot[i] = stq(qadd(bRot(S, i), da));
}

__global__ void k_proj_flags(State S)
{
    const int w = GID;
    if (w >= S.numWorlds) return;
    if (S.wProjDone[w]) return;
    const float tol = smax(S.P.projSlop * 0.01f, 1.0e-6f);
    if (__uint_as_float(S.wMaxDepth[w]) <= tol) S.wProjDone[w] = 1;
    S.wMaxDepth[w] = 0u;
}

// solver_step.cpp:164-170
__global__ void k_rebase(State S)
{
    FOR_LIVE(k) manifoldRebase(S, k);
}

} // namespace avbd_dev

// Device-resident state of a batch of worlds, passed by value to every kernel.
//
// HBM layout (structure of arrays, 16-byte vectors so one thread issues one dwordx4 per
// field):
//   bodies    float4 pos/rot/init*/inertial*/vel/angVel/prevVel, float4 size+radius,
//             float4 mass+moment; index = world body base + reference list position
//   joints    float4 anchors, C0, penalty, lambda (lin and ang); index = world joint base
//             + reference force-list position
//   manifolds stable slots in per-world ranges; per slot: bodies, 3 float4 basis rows,
//             friction, count; staging rows per contact (slot*8 + i) written by the
//             narrow phase
//   contacts  float4 rA|betaScale, rB|featureKey, C0|stick, penalty, lambda per contact,
//             in view order, sliced by wavefront (contact i of a manifold at its first
//             row + 64 i)
//   live      per-world slot lists in creation order (the walk order of every
//             order-dependent pass)
//   adjacency CSR per body: joints in list order, then manifolds in live order
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace avbd_dev {

struct Params {
    float dt, dtSq, collisionDt, alpha, oneMinusAlpha, betaLin, betaAng, gamma;
    float jacobiRelax, stickDriftHeal, maxPen, specMargin;
    float projBeta, projSlop, projMaxCorr, penaltyMax, adaptiveFloor, stiffnessDecay;
    int adaptiveBeta, contactOnlyOverflow, projIters;
    // sleeping (solver_solve.cpp:143-339): sleepOn = some body can sleep or sleeps; sleepFrames
    // = idle steps before sleeping (0: no idle detection); thresholds in m/s and rad/s
    int sleepOn, sleepFrames;
    int grace;   // contact grace ticks (0: a separated pair's manifold is released at once)
    float sleepLin, sleepAng;
    // drift-gated arm reuse (manifold.cpp:600-607): on when the rebuild distance is positive;
    // armRebuild2 = distance squared
    int armCache;
    float armRebuild2;
};

constexpr int NARROW_REC = 8;
constexpr uint8_t FORCE_JOINT = 0, FORCE_SPRING = 1, FORCE_IGNORE = 2, FORCE_GENERIC = 3;

struct State {
    Params P;
    int numWorlds;
    int numBodies;   // total over worlds (capacity in use)
    int slotCap;     // total manifold slot capacity
    int jointCap;
    int pairCap;

    // ---- bodies
    float4 *pos, *rot, *initLin, *initAng, *inertLin, *inertAng, *vel, *angVel, *prevVel;
    float4 *sizeR, *massI;
    // Per body, 64 B (written by the substep's warm start and by every move in the rounds, so
    // it is current whenever a pass gathers body poses per manifold): {pos - initial, awake
    // mass}, rot, {rot (-) initial rotation, 0}, unused.  One aligned half line per gather
    // instead of five separate rows.
    float4 *poseRec;
    // Per body, NARROW_REC float4 (one 128-byte line) written with the swept AABB for the
    // narrow phase: {pos, mass}, rot, {size, radius}, {vel, gravity}, {angVel, friction}.
    float4 *narrowRec;
    float *fric, *grav;
    uint32_t *layer, *mask;
    uint8_t *lockLin, *lockAng;
    uint8_t *shape;       // AVBD_SHAPE_* per body
    uint8_t *sleepMode, *sleeping;   // AVBD_SLEEP_* per body; 1 while asleep
    int *stillFrames;                // steps below the sleep thresholds
    uint32_t *updFlag, *updScan;     // per step: in the update set; its per-world ranks
    // contact grace: per slot, ticks parked and contacts kept; the kept contacts (slot * MAXC
    // rows: arm A + beta scale, arm B + key, C0, penalty); per pair, "has a parked manifold"
    int *mDormant, *mSaved;
    float4 *dRA, *dRB, *dC0, *dPen;
    // arm cache (allocated once arm reuse is first switched on, else null): per slot the first
    // rotation columns of both bodies when the arms were captured (armCap[2 slot] = {column of
    // A, 1 when captured}, armCap[2 slot + 1] = {column of B, 0}); per contact (slot * MAXC + i)
    // the world arms computed last (armA, armB)
    float4 *armCap, *armA, *armB;
    // constraint values of the step's last dual, kept while the host asks for them (else null):
    // per contact (slot * MAXC + i) the normal row, per joint the linear rows
    float *resC;
    float4 *resJ;
    uint8_t *pairParked;
    const float *rim;     // cylinder rim directions: 8 cosines then 8 sines (host libm values)
    int *bodyWorld;
    int *updateIndex;  // rank among movable bodies of the world (list order), -1 static
    uint8_t *colour;   // 0..7, 8 = Jacobi overflow, 0xFF = not yet coloured
    uint8_t *contactOnly;
    float4 *dLin, *dAng;
    float4 *aabbC, *aabbH;
    // bodies whose block failed the unpivoted LDL^T this pass: (body, block) queued for the
    // pivoted solve, which runs in its own kernel to keep the hot kernels call- and scratch-free
    int *fbCount;
    int *fbBody;
    float *fbBlock;  // 33 floats per entry (lhsLin, lhsAng, lhsCross column-major, rhsLin, rhsAng)
    int *adjCount, *adjStart, *adj;
    unsigned long long *adjTmp;
    int *adjPos;     // place of each manifold side / joint end in its body's list (count pass)

    // ---- joints
    int *jA, *jB, *jWorld;
    float4 *jRA, *jRB, *jC0L, *jC0A, *jPenL, *jPenA, *jLamL, *jLamA;
    float *jStiffL, *jStiffA, *jFrac, *jArm;
    uint8_t *jKind;    // FORCE_JOINT, FORCE_SPRING, FORCE_IGNORE, FORCE_GENERIC
    float *jRest;      // spring rest length
    // generic joints: rest orientation per joint; per axis (6 j + a, a = 0-2 linear, 3-5
    // angular) mode, bounds, spring stiffness / equilibrium / damping, lambda, penalty, and
    // the value at the step's start
    float4 *jGRest;
    uint8_t *jGMode;
    float *jGLo, *jGHi, *jGK, *jGEq, *jGDamp, *jGLam, *jGPen, *jGInit;
    uint8_t *jBroken, *jDead;

    // ---- manifold slots
    int *mA, *mB;
    // Per slot one aligned 64-byte record (mRec[4 slot + r]), written whole by the narrow phase
    // for every manifold it keeps (one full line instead of scattered partial lines): rows 0-2
    // the basis rows (xyz) with body a, body b and this step's narrow-phase work item (its
    // staging rows) in w as int bits; row 3 {friction, contact count (int bits), 0, 0}.
    float4 *mRec;
    int *mNum;          // contact count per slot (also in the record)
    // Contacts in view order (all worlds, world after world), sliced by wavefront: contact i
    // of the manifold in `slot` is at cOff[slot] + i * CSLICE.  Fields: rA | beta scale,
    // rB | feature key, C0, penalty and lambda (packed in 24 bytes), stick flag.
    float4 *cRA, *cRB, *cC0;
    float4 *cPL;                // {penalty.xyz, lambda.x}
    float2 *cLyz;               // {lambda.y, lambda.z}
    uint8_t *cStick;            // stick flag
    uint32_t *cOff;
    // Staging: the narrow phase writes the new contacts here (sliced by its work items) and
    // k_contact_pack lays them out in view order once the live list is final; host reads and
    // edits use it with a fixed stride (slot * MAXC + i).
    float4 *sRA, *sRB, *sC0, *sPen, *sLam;
    uint32_t *cCount;           // per view index k: contact count
    uint32_t *cScan;            // per contact slice: its width, then (scanned) its start
    uint32_t *sliceCount;       // number of contact slices this step
    // Dense view of the live manifolds for one step, index k = gLive position (world after
    // world, live order inside a world).  Adjacency entries, stamps, brake links and
    // projection records use k; slots are only the persistent storage across steps.
    // Inside each world's range the view is ordered by contact count (stable), so that a
    // wavefront's per-manifold contact loops run about as often as its manifolds have
    // contacts; gOfK[k] is the manifold's live position (its place in the reference's live
    // order, which orders adjacency and therefore every sum).
    int *gOfK;
    int *kSlot;                 // storage slot of view index k
    int *kT;                    // slot-range position (live list entry) of view index k
    int *newT;                  // per new manifold of this step (pair order): its live entry
    int4 *mHdr;                 // {body a, body b, first contact, contact count}
    float4 *mBasisK;            // 3 rows per k
    float *mFricK;
    uint8_t *brakeOn;           // per k: active and some contact has a positive gap now
    int *wOrder;                // world ids, most bodies first (per-world launches)
    int *slotWorld;     // static: world owning slot-range position t
    int *gLive;         // all worlds' live positions t, world after world (compact)
    uint32_t *wLiveOff; // per world: offset of its live entries in gLive; [numWorlds] = total
    float4 *stampA, *stampB;    // per view index k, STAMP_ROWS float4: the Jacobi stamp of each side
    int *live;          // per world range: slots in live order
    int *liveTmp;
    int *freeList;      // per world range
    uint32_t *keep;     // per live-list position: the manifold stays live this step
    // per-manifold brake schedule (indexed by slot-range position of the live entry)
    int *succA, *succB, *indeg, *levelOrder;
    int *brakeLevels;   // per world: number of levels
    int *levelStart;    // per world slot range: first entry of each level
    int *wLevelEnd;     // per world: end of the last level
    float4 *projLA, *projAA, *projLB, *projAB; // per live position: per-side corrections
    uint8_t *projOn;    // per live position: manifold corrected this pass

    // ---- per-world tables
    int *wBodyBase, *wBodyCount, *wJointBase, *wJointCount, *wSlotBase, *wSlotCap;
    int *wLive, *wFree, *wTop, *wNew, *wHasJoints, *wUpdateCount;
    int *wErr;
    int *wProjDone;
    unsigned int *wMaxDepth;   // float bits (non-negative) for atomicMax
    int *wColourLeft;
    int *wColourTail;            // per world: bodies k_colour_counts queued as ready (colourQueue)

    // ---- broad phase
    unsigned long long *mortonCode, *sortKey, *sortKeyAlt;
    int *sortVal, *sortValAlt;
    int *nodeLeft, *nodeRight, *nodeParent, *nodeFirst, *nodeLast, *nodeFlag;
    // Two float4 per node (internal nodes [0, n-1), leaves [n-1, 2n-1)):
    // {min, left child id | leaf: body}, {max, rope = next node after this subtree, -1 = end}
    float4 *nodeBox;
    int *nodeRope;
    uint32_t *taskFlag, *taskScan;   // refit task roots (see k_task_flags), per node
    int *taskNode;
    int *colourCnt, *colourQueue;    // colour finisher: uncoloured lower neighbours, ready queue
    // bodies sorted by colour (stable in body order; key 255 = outside the update set) and the
    // start of each colour's run: run c is colVal[colStart[c] .. colStart[c + 1])
    unsigned long long *colKey, *colKeyAlt;
    int *colVal, *colValAlt;
    uint32_t *colStart;
    int *leafBody;               // sorted position -> body
    int *bodyLeaf;               // body -> sorted position
    uint32_t *pairCount, *pairOff;
    int *pairI, *pairJ;
    int *pairScratch;            // PAIR_BUF candidates per body from the single traversal
    // Candidate lists: each body's box grown by a margin (listC/listH, centre and half) is what
    // the tree's leaves hold; listJ keeps the LIST_K lowest j > i of the same world whose grown
    // box meets i's (sorted), listCnt their number (more than LIST_K: walk the tree each step).
    // listFlag per body this step: 2 = its box left its grown box (grown again, walks its list
    // and marks), 1 = marked by such a body (walks its list), 0 = keeps its list.
    // Per world: wRewalk = every body walks and lists again (a new world, or one that calmed
    // down); wStale = it was busy last step, so its boxes are not grown and its bodies walk the
    // exact boxes without lists; wMoved = bodies that moved out of their grown boxes this step
    // (busy: over half); wRefit = some grown box changed.  listStats (last step): bodies moved,
    // busy worlds, stale worlds, worlds that listed again.
    // listAdd: entries appended to a list this step (after its first listCnt).
    float4 *listC, *listH;
    int *listJ, *listCnt, *listAdd, *listFlag;
    int *wRewalk, *wRefit, *wMoved, *wStale;
    int *listStats;
    int *wRoot;                  // per world: node whose leaf range is exactly the world
    uint32_t *pairValid, *pairValidScan;
    // per pair: a new pair of boxes parted beyond the margin (no manifold is made for it, though
    // the reference makes one and releases it in the same step; it still wakes a sleeper)
    uint8_t *pairApart;
    float4 *wBoundsLo, *wBoundsHi;
    // A world's wide body (its lowest-indexed body whose box spans at least a quarter of the
    // world's extent on two axes, chosen at each tree rebuild; NO_WIDE if none) does not walk
    // the tree: its rope walk would visit every node that reaches into its box (a ground box
    // under hanging chains: thousands of dependent node loads on one lane).  Instead every
    // body j above it tests its box against the wide body's (wideMask: one bit per body,
    // wideScan: exclusive scan of the per-word bit counts), and writes itself into the wide
    // body's pair range at its rank, so the range comes out sorted by j.
    int *wWide;
    unsigned long long *wideMask;
    uint32_t *wideWord, *wideScan;
};

} // namespace avbd_dev

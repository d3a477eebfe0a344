/*
 * avbd_hip.h - C ABI of the MI355X AVBD rigid-body step.
 *
 * One step of the augmented-vertex-block-descent solver for box bodies and ball/angular
 * joints, run on one AMD Instinct MI355X (gfx950).  The ABI is plain C: POD structs of
 * host pointers and sizes, integer status codes, no exceptions, no torch types.
 *
 * Index conventions (they decide results, so they are part of the contract):
 *  - Body index i is the reference solver's list position: the reference prepends every
 *    new body to an intrusive list, so index 0 is the body created LAST
 *    (reference src/avbd/rigid.cpp:28-30, solver_broadphase.cpp:87-92).
 *  - Joint index j is likewise the force-list position (reference src/avbd/force.cpp:19-21).
 *  - Manifold index m is the position in the solver's live list (creation order,
 *    reference src/avbd/manifold.hpp:92, solver_broadphase.cpp:153-184).
 *
 * Ownership: every pointer passed in belongs to the caller and is only read/written for
 * the duration of the call.  A world or batch handle owns its device memory.  A handle is
 * used by one host thread at a time; distinct handles may be used from distinct threads.
 */
#ifndef AVBD_HIP_H
#define AVBD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t avbd_status;
#define AVBD_OK 0
#define AVBD_ERR_INVALID (-1)      /* bad argument (null pointer, negative count, ...)   */
#define AVBD_ERR_UNSUPPORTED (-2)  /* parameter or feature outside this build's path     */
#define AVBD_ERR_CAPACITY (-3)     /* a device buffer would overflow its capacity        */
#define AVBD_ERR_DEVICE (-4)       /* a HIP runtime call failed                          */
#define AVBD_ERR_NO_DEVICE (-5)    /* no usable gfx950 device                            */
#define AVBD_ERR_WORLD (-6)        /* some worlds of a batch failed their step; the others
                                      stepped (avbd_batch_failed_worlds lists them)      */

/* Why a world of a batch failed its step (bits of avbd_batch_failed_worlds' reasons).  The
 * reference has no such check (it never detects non-finite state, reference
 * test/core_test.cpp:105-120 checks it after the fact); a batch needs it so that one diverged
 * or overfull world can be evicted without stopping the others. */
#define AVBD_WORLD_SLOTS 1         /* more live manifolds than the world's slot range     */
#define AVBD_WORLD_TREE 12         /* its broad-phase tree walk failed (4) or was missing (8) */
#define AVBD_WORLD_NONFINITE 16    /* a movable body's pose or velocity is not finite     */

/* Solver parameters.  Field meaning and defaults follow the reference's public Solver
 * fields (reference src/avbd/solver.h:54-144) and constants (src/avbd/types.hpp:23-45).
 * avbd_params_default() fills the reference defaults.  penalty_max is the reference's
 * compile-time PENALTY_MAX (1e10) made a parameter. */
typedef struct avbd_params {
    float dt;                        /* 1/60 */
    float gravity;                   /* -10, along -Z; copied into bodies that pass no gravity */
    int32_t iterations;              /* 8 */
    float alpha;                     /* 0.99 */
    float beta_lin;                  /* 1e4 */
    float beta_ang;                  /* 100 */
    float gamma;                     /* 0.999 */
    float newton_ratio;              /* 1 */
    float stiffness_decay;           /* 1 */
    int32_t adaptive_beta;           /* 0 */
    float adaptive_beta_error_floor; /* 1e-3 */
    float convergence_threshold;     /* 0; > 0 ends each substep's rounds as the reference's
                                        probe does (solver_step.cpp:299-314) */
    float jacobi_relaxation;         /* 0.7 */
    int32_t contact_only_overflow;   /* 1 */
    float stick_drift_heal;          /* 0.002 */
    float max_penetration;           /* 0.02 */
    float jacobian_rebuild_distance; /* 0 = off; > 0: duals reuse world arms until a rotation drifts this far */
    float speculative_margin;        /* 0.01 */
    int32_t position_projection;     /* 1 */
    int32_t projection_iterations;   /* 4 */
    float projection_beta;           /* 1 */
    float projection_slop;           /* 0.008 */
    float projection_max_correction; /* 0 */
    int32_t contact_grace_ticks;     /* 0 (only 0 is supported) */
    int32_t threads;                 /* ignored on the GPU */
    float sleep_linear_threshold;    /* 0.1 */
    float sleep_angular_threshold;   /* 0.139626 */
    int32_t sleep_frames;            /* 0 (only 0 is supported) */
    float penalty_max;               /* 1e10 */
} avbd_params;

/* Body arrays, all of length `count`, list order.  Vectors are packed xyz (3 floats),
 * quaternions xyzw (4 floats).  On upload the constant fields and the dynamic state are
 * read; on download only non-null pointers are written. */
typedef struct avbd_bodies {
    int32_t count;
    float *size;             /* 3n full box extents                                  */
    float *mass;             /* n, 0 = static                                        */
    float *moment;           /* 3n principal moments of inertia                      */
    float *friction;         /* n                                                    */
    float *radius;           /* n bounding radius |size/2|                           */
    float *gravity;          /* n signed gravity along -Z (null: params.gravity)     */
    uint32_t *layer;         /* n collision layer (null: 1)                          */
    uint32_t *mask;          /* n collision mask (null: 1)                           */
    uint8_t *lock_linear;    /* n axis-lock bits 0..2 (null: 0)                      */
    uint8_t *lock_angular;   /* n (null: 0)                                          */
    float *position;         /* 3n                                                   */
    float *rotation;         /* 4n                                                   */
    float *velocity;         /* 3n                                                   */
    float *angular_velocity; /* 3n                                                   */
    float *prev_velocity;    /* 3n previous step's linear velocity (null: velocity)  */
    uint8_t *shape;          /* n AVBD_SHAPE_* (null: every body a box)              */
    uint8_t *sleep_mode;     /* n AVBD_SLEEP_* (null: never)                         */
    uint8_t *sleeping;       /* n in/out: 1 asleep (null: awake; download fills it)  */
    int32_t *still_frames;   /* n in/out: steps below the thresholds (null: 0)       */
} avbd_bodies;

/* Sleep modes (reference `Rigid::sleep_mode`, src/avbd/types.hpp:138): never; sleep after
 * params.sleep_frames idle steps and wake on contact or a moving neighbour; start asleep and
 * wake on motion or forces. */
#define AVBD_SLEEP_NEVER 0
#define AVBD_SLEEP_IDLE 1
#define AVBD_SLEEP_START_ASLEEP 2

/* Body shapes (reference `ShapeType`, src/avbd/types.hpp:52-59).  `size` per shape: box =
 * full extents; sphere = (radius, unused, unused); cylinder = (radius, unused, height), its
 * axis the body's local +Z. */
#define AVBD_SHAPE_BOX 0
#define AVBD_SHAPE_SPHERE 1
#define AVBD_SHAPE_CYLINDER 2

/* Joint arrays (reference `Joint`, src/avbd/forces.hpp:58-87), list order.  body_a = -1
 * anchors body_b to the world point r_a.  A ball joint is stiffness_lin = +inf,
 * stiffness_ang = 0.  The dual state fields are read on upload and written on download. */
typedef struct avbd_joints {
    int32_t count;
    int32_t *body_a;
    int32_t *body_b;
    float *r_a;           /* 3n local anchor on A (world point if body_a < 0) */
    float *r_b;           /* 3n local anchor on B                             */
    float *stiffness_lin; /* n */
    float *stiffness_ang; /* n */
    float *fracture;      /* n (+inf: unbreakable)                            */
    float *torque_arm;    /* n |size_a + size_b|^2                            */
    float *c0_lin;        /* 3n */
    float *c0_ang;        /* 3n */
    float *penalty_lin;   /* 3n */
    float *penalty_ang;   /* 3n */
    float *lambda_lin;    /* 3n */
    float *lambda_ang;    /* 3n */
    uint8_t *broken;      /* n */
    /* Force kind per entry (NULL: every entry is a joint).  The forces share one list, in the
     * reference's list order (forces.hpp:19-113): AVBD_FORCE_JOINT is the joint above;
     * AVBD_FORCE_SPRING is a spring between two bodies (spring.cpp:16-75) with anchors r_a, r_b,
     * stiffness in stiffness_lin and rest length in rest; AVBD_FORCE_IGNORE_COLLISION keeps the
     * pair of bodies out of the broad phase and acts on nothing else (forces.hpp:104-113). */
    uint8_t *kind;        /* n, or NULL */
    float *rest;          /* n spring rest length (read for springs; NULL when there are none) */
    /* AVBD_FORCE_GENERIC: the 6-DOF joint (generic_joint.cpp) with anchors r_a, r_b.  Axis a of
     * entry k is at 6 k + a: 0-2 linear x, y, z and 3-5 angular x, y, z of A's frame (the
     * world's when body_a < 0).  Read only for generic entries; NULL when there are none. */
    float *rest_rotation;    /* 4n B's orientation relative to A at creation, xyzw          */
    uint8_t *axis_mode;      /* 6n AVBD_AXIS_FREE / _LOCKED / _LIMITED                      */
    float *axis_lower;       /* 6n limit bounds                                            */
    float *axis_upper;       /* 6n                                                          */
    float *axis_spring;      /* 6n spring stiffness (0: none)                               */
    float *axis_equilibrium; /* 6n spring equilibrium                                       */
    float *axis_damping;     /* 6n spring damping                                           */
    float *axis_lambda;      /* 6n dual state: read on upload, written on download         */
    float *axis_penalty;     /* 6n                                                          */
    float *axis_initial;     /* 6n value at the start of the last step                     */
} avbd_joints;

#define AVBD_FORCE_JOINT 0
#define AVBD_FORCE_SPRING 1
#define AVBD_FORCE_IGNORE_COLLISION 2
#define AVBD_FORCE_GENERIC 3
#define AVBD_AXIS_FREE 0
#define AVBD_AXIS_LOCKED 1
#define AVBD_AXIS_LIMITED 2

#define AVBD_MAX_CONTACTS 8

/* Contact manifolds in live order (reference ManifoldStore, src/avbd/manifold.hpp:68-131).
 * Per-contact arrays have AVBD_MAX_CONTACTS entries per manifold.  basis is row-major with
 * row 0 the contact normal (pointing from B to A) and rows 1-2 the tangents. */
typedef struct avbd_manifolds {
    int32_t count;
    int32_t *body_a;
    int32_t *body_b;
    float *basis;        /* 9 per manifold          */
    float *friction;     /* 1 per manifold          */
    int32_t *num_contacts;
    int32_t *feature;    /* 8 per manifold          */
    float *r_a;          /* 24 per manifold (8 x 3) */
    float *r_b;          /* 24 per manifold         */
    float *c0;           /* 24 per manifold         */
    float *penalty;      /* 24 per manifold         */
    float *lambda;       /* 24 per manifold         */
    uint8_t *stick;      /* 8 per manifold          */
    float *beta_scale;   /* 8 per manifold          */
} avbd_manifolds;

/* Per-step counters and phase times of the last step (milliseconds, from HIP events). */
typedef struct avbd_step_stats {
    int32_t bodies;
    int32_t joints;
    int32_t manifolds;
    int32_t contacts;
    int32_t candidate_pairs;
    int32_t new_manifolds;
    int32_t colours;
    int32_t overflow_bodies;
    int32_t brake_levels;
    float ms_broad;
    float ms_narrow;
    float ms_solve;
    float ms_finish;
    float ms_total;
    float ms_broad_tree;             /* tree build or refit (part of ms_broad)              */
    float ms_broad_emit;             /* candidate pair walk and emission (part of ms_broad) */
} avbd_step_stats;

/* Kernel classes timed by avbd_*_profile (average duration per launch, HIP events on the
 * handle's stream around each launch). */
#define AVBD_K_PRIMAL 0        /* Jacobi primal: per-body block assembly + 6x6 solve     */
#define AVBD_K_DUAL 1          /* dual: per-manifold lambda / penalty update + joints      */
#define AVBD_K_NARROW 2        /* narrow phase + warm start, per manifold                  */
#define AVBD_K_BRAKE_LEVELS 3  /* level schedule + first velocity brake, per world         */
#define AVBD_K_BRAKE_REPLAY 4  /* second velocity brake, per world                         */
#define AVBD_K_TRAVERSE 5      /* LBVH traversal (count + emit)                            */
#define AVBD_K_COLOUR_PASS 6   /* one colour of the Gauss-Seidel sweep                     */
#define AVBD_K_PROJ_MANIFOLD 7 /* projection: per-manifold corrections                     */
#define AVBD_K_PROJ_BODY 8     /* projection: per-body gather and apply                    */
#define AVBD_K_ADJACENCY 9     /* CSR adjacency build                                      */
#define AVBD_K_SORT 10         /* Morton radix sort                                        */
#define AVBD_K_BODY 11         /* warm start + finish + Jacobi combine                     */
#define AVBD_K_STAMP 12        /* Jacobi stamps of every manifold at the current poses     */
#define AVBD_K_REFIT 13        /* LBVH refit, per world                                    */
#define AVBD_K_CLASSES 14

typedef struct avbd_kernel_profile {
    int32_t steps;
    float avg_ms[AVBD_K_CLASSES];
    int32_t launches[AVBD_K_CLASSES];
    float step_ms;   /* average wall time of a profiled step (events, stream launches) */
} avbd_kernel_profile;

/* ---------------------------------------------------------------- helpers ---- */

/* Fill `p` with the reference defaults. */
void avbd_params_default(avbd_params *p);

/* Box mass properties exactly as the reference body constructor computes them
 * (src/avbd/rigid.cpp:59-70): mass = x*y*z*density, moment_x = (y^2+z^2)/12*mass, ...,
 * radius = |size/2|. */
void avbd_box_properties(const float size[3], float density, float *mass, float moment[3],
                         float *radius);
/* Mass properties of any shape exactly as the reference body constructor computes them
 * (src/avbd/rigid.cpp:34-70): sphere m = 4/3 pi r^3 density, I = 2/5 m r^2, radius r;
 * cylinder m = pi r^2 h density, I = (m (3 r^2 + h^2) / 12, same, m r^2 / 2), radius
 * |(r, h/2, r)|; box as avbd_box_properties.  Returns AVBD_ERR_INVALID for an unknown shape. */
int avbd_shape_properties(int shape, const float size[3], float density, float *mass, float moment[3],
                          float *radius);

/* FNV-1a digest over the raw bytes of position(12), rotation(16), velocity(12) and
 * angular velocity(12) of every body in list order (reference test/core_test.cpp:133-149). */
uint64_t avbd_state_digest(const avbd_bodies *bodies);

/* Message of the last error on the calling thread ("" if none). */
const char *avbd_last_error(void);

/* Number of visible gfx950 devices (0 when none; never fails). */
int32_t avbd_device_count(void);

/* ------------------------------------------------- single world (plain path) ---- */

typedef struct avbd_world avbd_world;

/* Create a world with room for the given counts on the current HIP device.  Replaces
 * constructing a reference `avbd::Solver` (src/avbd/solver.h:150). */
avbd_status avbd_world_create(const avbd_params *params, int32_t max_bodies,
                              int32_t max_joints, int32_t max_manifolds, avbd_world **out);
avbd_status avbd_world_destroy(avbd_world *world);
/* Replaces writing the public Solver fields (src/avbd/solver.h:54-144). */
avbd_status avbd_world_set_params(avbd_world *world, const avbd_params *params);
/* Load bodies and joints (in list order) and clear every manifold, as creating the
 * reference's `Rigid`/`Joint` objects in a fresh Solver does (types.hpp:114-117,
 * forces.hpp:68). */
avbd_status avbd_world_upload(avbd_world *world, const avbd_bodies *bodies,
                              const avbd_joints *joints);
/* Overwrite the dynamic state (position, rotation, velocity, angular velocity, previous
 * velocity; null arrays are skipped) of the uploaded bodies, keeping manifolds and joints:
 * what a host does when it writes `Rigid` fields between steps.  count must equal the
 * uploaded body count. */
avbd_status avbd_world_write_bodies(avbd_world *world, const avbd_bodies *bodies);
/* One step: replaces `Solver::step()` (src/avbd/solver.h:234, solver_step.cpp:44-75). */
avbd_status avbd_world_step(avbd_world *world);
/* Replaces `Solver::stepInterleaved(substeps, rounds)` (solver.h:241, solver_step.cpp:109-160);
 * params.dt is the substep dt. */
avbd_status avbd_world_step_interleaved(avbd_world *world, int32_t substeps, int32_t rounds);
/* Copy the state back: replaces reading the `Rigid`/`Joint` fields.  Null members skip. */
avbd_status avbd_world_download(avbd_world *world, avbd_bodies *bodies, avbd_joints *joints);
/* Manifold accessors: replace Solver::manifoldCount/manifoldContact/... (solver.h:170-231).
 * get: `out` arrays must hold `out->count` manifolds on entry; count is set on return. */
/* Per-ray results of avbd_world_raycast (arrays the caller owns; null arrays are skipped,
 * except body). */
typedef struct avbd_ray_hits {
    int32_t *body;     /* n list index of the hit body, -1 for none                   */
    float *t;          /* n distance along the segment (0 for a hit from inside)      */
    float *point;      /* 3n hit point (the origin for a hit from inside)             */
    float *normal;     /* 3n outward surface normal (-direction for a hit from inside) */
    uint8_t *inside;   /* n 1 when the segment starts inside the body                 */
} avbd_ray_hits;

/* n segments from[3r..] -> to[3r..] against the bodies at their current pose: replaces the
 * bodies part of the Godot server's intersect_ray (src/server/avbd_direct_space_state3d.cpp:
 * 103-195 with query::intersectRay, src/avbd/query/query_shape.cpp:258-274).  Candidates are
 * the bodies whose layer passes masks[r] (null: all) and whose tight bounds overlap the
 * segment's box padded by 1e-4, in list order; the nearest hit wins, the earlier body on a tie;
 * with hit_from_inside, a segment starting inside a body reports it at t = 0 and stops there.
 * Exclusion lists, areas and ray_pickable are the server's and stay there. */
avbd_status avbd_world_raycast(avbd_world *world, int32_t n, const float *from, const float *to,
                               const uint32_t *masks, int32_t hit_from_inside, avbd_ray_hits *out);
/* Per-body contact reports: replaces the server's contact collection after a step
 * (src/server/avbd_physics_server3d.cpp:1593-1652, read through Solver::manifoldBodyA/B,
 * manifoldContact and manifoldNormal, solver.h:193-220).  Body j asks for max_reported[j]
 * contacts (0: none).  Live manifolds are walked in order; each side whose body asks takes the
 * manifold's contacts in order until it holds its count: position[3(j*stride+c)..] = the
 * body's own arm in world space at its current pose, normal = the manifold normal (first basis
 * row) signed towards the body (+ for A, - for B), other = the other body's list index.
 * count[j] = contacts written.  stride >= every max_reported[j]. */
avbd_status avbd_world_contact_report(avbd_world *world, const int32_t *max_reported, int32_t stride, int32_t *count,
                                      float *position, float *normal, int32_t *other);
/* The bodies part of the Godot server's collide_shape (src/server/avbd_direct_space_state3d.cpp:
 * 386-447): for each of n query shapes (laid out as for avbd_world_intersect_shape; the caller
 * applies any motion to the centre), the candidates in list order, the narrow phase (query first)
 * within the margin, and every contact as a pair of world points (on the query shape, on the
 * body) until max_pairs pairs are out: points[6 (r max_pairs + p) ..] = (query point, body
 * point).  count[r] = pairs written; any[r] = 1 when some candidate touched.  Areas stay in the
 * server. */
avbd_status avbd_world_collide_shape(avbd_world *world, int32_t n, const int32_t *type, const float *center,
                                     const float *rotation, const float *size, const float *margin,
                                     const uint32_t *masks, int32_t max_pairs, int32_t *count, uint8_t *any,
                                     float *points);
/* The bodies part of the server's rest_info (avbd_direct_space_state3d.cpp:449-513): over the
 * candidates in list order and their contacts, the deepest by query::penetrationDepth along the
 * contact normal (src/avbd/query/query_shape.cpp:340-369; a later contact must be strictly
 * deeper).  body[r] = its list index (-1: none); point = the contact on the query shape, world
 * space; normal = the basis' first row (from the body towards the query); depth.  Null point,
 * normal or depth skip. */
avbd_status avbd_world_rest_info(avbd_world *world, int32_t n, const int32_t *type, const float *center,
                                 const float *rotation, const float *size, const float *margin, const uint32_t *masks,
                                 int32_t *body, float *point, float *normal, float *depth);
/* The server's get_closest_point_to_object_volume for bodies (avbd_direct_space_state3d.cpp:
 * 515-545 with query::closestPoint, query_shape.cpp:280-323): out[3r..] = the point of body
 * body[r]'s volume closest to points[3r..] at the body's current pose. */
avbd_status avbd_world_closest_point(avbd_world *world, int32_t n, const int32_t *body, const float *points,
                                     float *out);
/* Replaces `avbd::Manifold::collide(bodyA, bodyB, contacts, basis)` (reference
 * src/avbd/manifold.hpp:61-63, collide.cpp:1100-1120): the solver's narrow phase for bodies
 * body_a (as A) and body_b (list indices) of the world at their current pose, with the
 * speculative margin grown by their predicted relative motion over collision_dt (the solver's
 * collisionDt).  Writes *count contacts (at most 8): features[k] (the feature key), r_a[3k..] /
 * r_b[3k..] (contact offsets in A's / B's local frame), and basis[9] (row-major; row 0 is the
 * normal from B to A; left as given when the pair is beyond the margin).  Null outputs skip. */
avbd_status avbd_world_collide_pair(avbd_world *world, int32_t body_a, int32_t body_b, float collision_dt,
                                    int32_t *count, int32_t *features, float *r_a, float *r_b, float *basis);
/* avbd_world_raycast for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No reference counterpart.
 * Same arguments and results, bit for bit; it calls no hipMalloc / hipFree: its staging (the
 * segments, the bodies' bounds and the hits) lives in one fixed device array per GPU,
 * AVBD_NOALLOC_SCRATCH_BYTES long, held until the process ends.  Limits: at most
 * AVBD_NOALLOC_MAX_RAYS segments per call and AVBD_NOALLOC_MAX_BODIES bodies in the world
 * (AVBD_ERR_CAPACITY beyond).  Calls on one GPU share that array: calls that overlap on one GPU
 * (from several host threads) race and must be serialised by the caller. */
#define AVBD_NOALLOC_MAX_RAYS 16384
#define AVBD_NOALLOC_MAX_BODIES 65536
#define AVBD_NOALLOC_SCRATCH_BYTES 4145152
avbd_status avbd_world_raycast_noalloc(avbd_world *world, int32_t n, const float *from, const float *to,
                                       const uint32_t *masks, int32_t hit_from_inside, avbd_ray_hits *out);
/* Measurement (no reference counterpart): average device time of the bounds kernel with the
 * query tree's refit (bounds_ms) and of the ray kernel (kernel_ms) of avbd_world_raycast over
 * `reps` launches of n segments (all layers, no hit from inside), timed with HIP events on the
 * library's stream. */
typedef struct avbd_query_profile {
    float bounds_ms;
    float kernel_ms;
} avbd_query_profile;
avbd_status avbd_world_profile_raycast(avbd_world *world, int32_t n, const float *from, const float *to,
                                       int32_t reps, avbd_query_profile *out);
/* n points against the bodies at their current pose: replaces the bodies part of the server's
 * intersect_point (src/server/avbd_direct_space_state3d.cpp:225-259 with query::containsPoint,
 * query_shape.cpp:334-337).  Candidates are the bodies whose layer passes masks[r] (null: all)
 * and whose bounds overlap the point's box padded by 1e-4, in list order; out[r*max_results + c]
 * = the list index of the c-th body containing point r (-1 past count[r]). */
avbd_status avbd_world_intersect_point(avbd_world *world, int32_t n, const float *points, const uint32_t *masks,
                                       int32_t max_results, int32_t *count, int32_t *out);
/* n query shapes against the bodies at their current pose: replaces the bodies part of the
 * server's intersect_shape (src/server/avbd_direct_space_state3d.cpp:261-306 with
 * query::shapeAabb and query::shapeContacts = collideShapes).  Query r: type[r] (AVBD_SHAPE_*),
 * center[3r..] (the pose already moved by any motion), rotation[4r..], size[3r..] (as a body's),
 * margin[r] >= 0.  Candidates are the bodies whose layer passes masks[r] (null: all) and whose
 * bounds overlap the query's bounds grown by the margin, in list order; a hit is a body the narrow
 * phase (query first) finds in contact within the margin.  out as avbd_world_intersect_point. */
avbd_status avbd_world_intersect_shape(avbd_world *world, int32_t n, const int32_t *type, const float *center,
                                       const float *rotation, const float *size, const float *margin,
                                       const uint32_t *masks, int32_t max_results, int32_t *count, int32_t *out);
/* n query shapes swept by a motion against the bodies at their current pose: replaces the
 * bodies part of the server's cast_motion (src/server/avbd_direct_space_state3d.cpp:310-385 with
 * query::sweepShape, src/avbd/query/query_shape.cpp:402-498).  Query r as in
 * avbd_world_intersect_shape plus motion[3r..].  Candidates are the bodies whose layer passes
 * masks[r] and whose bounds overlap the query's bounds swept by the motion and grown by the
 * margin, in list order; the sweep with the smallest safe fraction (the first of equals) wins:
 * body[r] (-1: none), fraction[2r..] = (safe, unsafe) ((1, 1) when none), point[3r..] and
 * normal[3r..] (optional) at the first colliding pose. */
avbd_status avbd_world_cast_motion(avbd_world *world, int32_t n, const int32_t *type, const float *center,
                                   const float *rotation, const float *size, const float *motion, const float *margin,
                                   const uint32_t *masks, int32_t *body, float *fraction, float *point, float *normal);
/* Replaces `Solver::pick(origin, dir, local, mask)` (src/avbd/solver.h:161-164,
 * solver_query.cpp:29-155) for n rays: body[r] = list index of the nearest body (-1 none),
 * local[3r..] = the hit point in its frame (zeros for none); masks null = all layers. */
avbd_status avbd_world_pick(avbd_world *world, int32_t n, const float *origin, const float *dir,
                            const uint32_t *masks, int32_t *body, float *local);
/* Replaces `Solver::collectAABB(min, max, mask, out)` (solver.h:165-167, solver_query.cpp:
 * 158-171): list indices, in list order, of the bodies whose layer passes the mask and whose
 * tight bounds overlap [min, max].  *count is the total; AVBD_ERR_CAPACITY when it exceeds
 * capacity (the first capacity indices are written). */
avbd_status avbd_world_collect_aabb(avbd_world *world, const float min[3], const float max[3], uint32_t mask,
                                    int32_t *out, int32_t capacity, int32_t *count);
avbd_status avbd_world_get_manifolds(avbd_world *world, avbd_manifolds *out);
avbd_status avbd_world_set_manifolds(avbd_world *world, const avbd_manifolds *in);
avbd_status avbd_world_stats(avbd_world *world, avbd_step_stats *out);
/* Keep (on = 1) or stop keeping the constraint values C the last dual of each step computes:
 * a contact's normal row (ManifoldStore::updateDual, src/avbd/manifold.cpp:626-630) and a ball
 * / finite joint's linear rows (Joint::updateDual, src/avbd/joint.cpp:176-182), which the
 * reference computes and discards.  Off by default; on, the step's last dual writes them.
 * For long-run residual statistics (no reference counterpart). */
avbd_status avbd_world_record_residuals(avbd_world *world, int32_t on);
/* The kept values of the last step: contact_cn[8 i + k] for live manifold i (the order of
 * avbd_world_get_manifolds) and contact k below its count (0 past it and for a manifold whose
 * bodies were frozen), joint_clin[3 q..] for joint q in the force-list order of
 * avbd_world_download (0 for joints without linear rows).  *num_manifolds / *num_joints hold
 * the arrays' capacities and receive the counts (AVBD_ERR_CAPACITY when too small). */
avbd_status avbd_world_get_residuals(avbd_world *world, int32_t *num_manifolds, float *contact_cn,
                                     int32_t *num_joints, float *joint_clin);
/* Colour of every body after the last step, in list order (solver_solve.cpp:365-424): 0-7 a
 * Gauss-Seidel colour, 8 the Jacobi (overflow) bucket, 9 an overflow body the step leaves
 * unsolved, 255 not in the update set.  `out` holds the world's body count. */
avbd_status avbd_world_colours(avbd_world *world, uint8_t *out);
/* Run `steps` steps with every launch bracketed by HIP events; fills per-class averages. */
avbd_status avbd_world_profile(avbd_world *world, int32_t steps, avbd_kernel_profile *out);

/* Broad-phase probe for parity tests: Morton codes of the current swept-AABB centres
 * (reference bvh_broadphase.cpp:19-36, 69-85) with the (code, index) sort order, and the
 * sorted unique candidate pairs (min<<32 | max) the step would emit.  `codes`/`order`
 * hold count bodies; `pairs` holds *pair_count on entry, the number found on return. */
avbd_status avbd_world_probe_broadphase(avbd_world *world, uint64_t *codes, int32_t *order,
                                        uint64_t *pairs, int32_t *pair_count);

/* ----------------------------------------------------- batched worlds ---- */

typedef struct avbd_batch avbd_batch;

/* A batch steps many independent worlds with one captured graph per step.  Capacities
 * are totals over all worlds. */
avbd_status avbd_batch_create(const avbd_params *params, int32_t max_worlds,
                              int32_t max_bodies, int32_t max_joints, int32_t max_manifolds,
                              avbd_batch **out);
avbd_status avbd_batch_destroy(avbd_batch *batch);
/* Launch shapes of the hot kernels (threads per workgroup; workgroups of the live-manifold
 * walks).  Results do not depend on them. */
typedef struct avbd_tuning {
    int32_t block_dual, block_primal, block_stamp, block_projection, block_traverse, live_blocks;
    int32_t colour_lanes;   /* lanes per body in the colour passes: 1 or 4 (0 = 1) */
} avbd_tuning;
/* Autotuner: steps the batch (or world) with candidate launch shapes, steps_per_trial steps per
 * candidate (3 block sizes, 5 live-list grids, and 2 colour-pass lane counts when colour passes
 * run: 8 or 10 x steps_per_trial steps), timing each kernel class with HIP events, and keeps the
 * fastest per class.  The steps are real steps of the simulation.  No reference counterpart
 * (the reference runs on the CPU). */
avbd_status avbd_batch_autotune(avbd_batch *batch, int32_t steps_per_trial, avbd_tuning *out);
avbd_status avbd_world_autotune(avbd_world *world, int32_t steps_per_trial, avbd_tuning *out);
/* Use these launch shapes (block sizes 64, 128 or 256; live_blocks in [8, 2^20]; colour_lanes 1
 * or 4), e.g. ones an earlier autotune chose, without timing anything.  Results do not depend on
 * them. */
avbd_status avbd_batch_set_tuning(avbd_batch *batch, const avbd_tuning *tuning);
/* Device bytes a batch of these capacities allocates (computed on the host; no device needed),
 * and the bytes an existing batch holds.  A serving loop sizes its batch from an HBM budget
 * with the first and admits worlds while their bodies, joints and manifold slots fit. */
avbd_status avbd_batch_footprint(const avbd_params *params, int32_t max_worlds, int32_t max_bodies,
                                 int32_t max_joints, int32_t max_manifolds, int64_t *bytes);
avbd_status avbd_batch_device_bytes(avbd_batch *batch, int64_t *bytes);
/* Add a world; its id is written to *world_id.  Takes effect at the next step. */
avbd_status avbd_batch_add_world(avbd_batch *batch, const avbd_bodies *bodies,
                                 const avbd_joints *joints, int32_t *world_id);
/* Serving without a reference counterpart: the same, but the world's arrays are copied to the
 * device on a second stream from pinned host memory and the call returns before they land, so
 * an admission made between avbd_batch_step_launch and avbd_batch_step_wait overlaps the
 * running step.  The world becomes active (and is stepped) from the next step launch; every
 * other call except the step, avbd_batch_fits and more asynchronous additions first makes
 * such worlds active.  The caller's arrays may be reused as soon as the call returns. */
avbd_status avbd_batch_add_world_async(avbd_batch *batch, const avbd_bodies *bodies,
                                       const avbd_joints *joints, int32_t *world_id);
/* avbd_batch_step in two halves: launch queues the step (after activating worlds added
 * asynchronously) and returns; wait blocks until it is done and returns what avbd_batch_step
 * would (AVBD_ERR_WORLD with avbd_batch_failed_worlds).  Between the two only
 * avbd_batch_add_world_async, avbd_batch_fits and avbd_batch_failed_worlds may be called;
 * anything else returns AVBD_ERR_INVALID. */
avbd_status avbd_batch_step_launch(avbd_batch *batch);
avbd_status avbd_batch_step_wait(avbd_batch *batch);
avbd_status avbd_batch_remove_world(avbd_batch *batch, int32_t world_id);
/* Whether a world of n_bodies bodies and n_joints joints can be added now (*fits_now: a free
 * world slot and free body, joint and manifold-slot ranges) and whether it could be added to
 * this batch when empty (*fits_ever); either pointer may be null.  A serving loop rejects a
 * request that never fits instead of waiting for it. */
avbd_status avbd_batch_fits(avbd_batch *batch, int32_t n_bodies, int32_t n_joints, int32_t *fits_now,
                            int32_t *fits_ever);
/* Launch batch steps as one captured HIP graph (enable = 1, the default) or as the same
 * kernels launched one by one on the batch's stream (enable = 0; for tracing tools). */
avbd_status avbd_batch_use_graph(avbd_batch *batch, int32_t enable);
/* The same choice for a plain world: its step is one captured HIP graph per (substeps, rounds,
 * tree rebuild) by default; enable = 0 launches the kernels one by one. */
avbd_status avbd_world_use_graph(avbd_world *world, int32_t enable);
/* Space queries (ray, point, shape, rest info, motion cast) walk the last step's LBVH refit to
 * the bodies' current tight bounds, from the first step after the world set last changed on;
 * enable = 0 makes them scan the body list instead.  The results are the same either way: the
 * tree only skips bodies whose bounds cannot overlap the query's, and candidates are still taken
 * in list order.  No reference counterpart (the reference's server scans the list). */
avbd_status avbd_world_use_query_tree(avbd_world *world, int32_t enable);
avbd_status avbd_batch_use_query_tree(avbd_batch *batch, int32_t enable);
/* One step of every world in the batch.  AVBD_ERR_WORLD: every world stepped except the ones
 * avbd_batch_failed_worlds lists, whose state this step is not valid (remove them; until then
 * they keep failing). */
avbd_status avbd_batch_step(avbd_batch *batch);
/* The worlds that failed the last step (ids, and AVBD_WORLD_* reason bits when reasons is not
 * null): at most `capacity` are written, *count is the total. */
avbd_status avbd_batch_failed_worlds(avbd_batch *batch, int32_t capacity, int32_t *world_ids, int32_t *reasons,
                                     int32_t *count);
/* avbd_world_contact_report for one world of a batch. */
avbd_status avbd_batch_contact_report(avbd_batch *batch, int32_t world_id, const int32_t *max_reported, int32_t stride,
                                      int32_t *count, float *position, float *normal, int32_t *other);
/* avbd_world_raycast for one world of a batch. */
avbd_status avbd_batch_raycast(avbd_batch *batch, int32_t world_id, int32_t n, const float *from, const float *to,
                               const uint32_t *masks, int32_t hit_from_inside, avbd_ray_hits *out);
avbd_status avbd_batch_download_world(avbd_batch *batch, int32_t world_id,
                                      avbd_bodies *bodies, avbd_joints *joints);
avbd_status avbd_batch_world_count(avbd_batch *batch, int32_t *count);
/* Manifolds of one world of the batch, as avbd_world_get_manifolds. */
avbd_status avbd_batch_get_manifolds(avbd_batch *batch, int32_t world_id, avbd_manifolds *out);
/* Upload a world's manifold state (e.g. a snapshot), as avbd_world_set_manifolds. */
avbd_status avbd_batch_set_manifolds(avbd_batch *batch, int32_t world_id, const avbd_manifolds *in);
/* Run `steps` batch steps as stream launches with per-launch HIP events (same kernels as
 * the captured graph); fills per-class average launch durations. */
avbd_status avbd_batch_profile(avbd_batch *batch, int32_t steps, avbd_kernel_profile *out);
/* Totals over the batch after the last step: bodies (movable), manifolds, contacts. */
avbd_status avbd_batch_totals(avbd_batch *batch, int64_t *movable_bodies, int64_t *manifolds,
                              int64_t *contacts, int64_t *joints);
/* Measurement (no reference counterpart): the broad phase's candidate lists in the last step.
 * Each body lists the bodies whose boxes, grown by a margin, meet its own, and walks the tree
 * again only when its box leaves its grown box (or a moved body's new grown box meets it);
 * a world where over half the bodies moved walks the exact boxes until it calms down.
 * out[0] = bodies that moved out of their grown boxes, out[1] = busy worlds, out[2] = worlds
 * that walked the exact boxes, out[3] = worlds that listed all their bodies again. */
avbd_status avbd_batch_list_stats(avbd_batch *batch, int32_t out[4]);
avbd_status avbd_world_list_stats(avbd_world *world, int32_t out[4]);

#ifdef __cplusplus
}
#endif

#endif /* AVBD_HIP_H */

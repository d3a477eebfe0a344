// ORACLE - test infrastructure only.  C interface of the serial CPU restatement of the
// reference AVBD step (avbd_oracle.cpp).  Used by tests/, the smoke check and the bench's
// CPU baseline as the checker of the HIP path; never by the product.
#pragma once
#include "../include/avbd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_world oracle_world;

oracle_world *oracle_create(const avbd_params *params); /* null on unsupported params */
void oracle_destroy(oracle_world *o);
int oracle_set_params(oracle_world *o, const avbd_params *params);
int oracle_upload(oracle_world *o, const avbd_bodies *b, const avbd_joints *j);
int oracle_write_bodies(oracle_world *o, const avbd_bodies *b);
int oracle_step(oracle_world *o);
int oracle_step_interleaved(oracle_world *o, int substeps, int rounds);
int oracle_download(oracle_world *o, avbd_bodies *b, avbd_joints *j);
int oracle_manifold_count(oracle_world *o);
/* The generic joint's atan2 (the same expression as the device's). */
float oracle_atan2(float y, float x);
int oracle_get_manifolds(oracle_world *o, avbd_manifolds *m);
int oracle_set_manifolds(oracle_world *o, const avbd_manifolds *m);
// Planted mistakes for the oracle's own tests (0 = none; the ids are listed in avbd_oracle.cpp).
void oracle_plant(int id);
void oracle_cache_counters(oracle_world *o, long long *out);
int oracle_get_residuals(oracle_world *o, float *contact_cn, float *joint_clin);
// query::closestPoint on one shape; the server's closest point to a body's volume for n bodies.
void oracle_closest_point_shape(int type, const float pos[3], const float rot[4], const float size[3],
                                const float point[3], float out[3]);
int oracle_closest_point(oracle_world *o, int n, const int32_t *body, const float *points, float *out);
// The bodies parts of the server's collide_shape and rest_info for n query shapes.
int oracle_collide_shape(oracle_world *o, int n, const int32_t *type, const float *center, const float *rot,
                         const float *size, const float *margin, const uint32_t *masks, int maxPairs, int32_t *count,
                         uint8_t *any, float *points);
int oracle_rest_info(oracle_world *o, int n, const int32_t *type, const float *center, const float *rot,
                     const float *size, const float *margin, const uint32_t *masks, int32_t *body, float *point,
                     float *normal, float *depth);
// avbd::Manifold::collide for bodies a and b at their current pose (margin grown over hz);
// returns the contact count (-1 for bad indices).
int oracle_collide_pair(oracle_world *o, int a, int b, float hz, int32_t *keys, float *rA, float *rB, float *basis);
uint64_t oracle_digest(oracle_world *o);
/* out: manifolds, contacts, candidate pairs, new manifolds, colours, overflow bodies */
void oracle_stats(oracle_world *o, int32_t out[6]);
/* per body in list order: 0-7 colour, 8 Jacobi bucket, 9 overflow left unsolved, 255 not updated */
void oracle_colours(oracle_world *o, uint8_t *out);
void oracle_swept_aabbs(oracle_world *o, float *centers, float *halves);
int oracle_candidate_pairs(oracle_world *o, uint64_t *pairs, int capacity);
void oracle_morton(const float *centers, int n, uint64_t *codes, int32_t *order);
int oracle_force_stamp(oracle_world *o, int joint, int body, float *out33);
int oracle_force_update(oracle_world *o, int joint, int what);
int oracle_intersect_ray(int type, const float pos[3], const float rot[4], const float size[3], const float origin[3],
                         const float dir[3], float maxT, float *t, float point[3], float normal[3], uint8_t *inside);
int oracle_collect_aabb(oracle_world *o, const float mn[3], const float mx[3], uint32_t mask, int32_t *out, int cap);
int oracle_raycast(oracle_world *o, int n, const float *from, const float *to, const uint32_t *masks,
                   int hitFromInside, int32_t *body, float *t, float *point, float *normal, uint8_t *inside);
int oracle_pick(oracle_world *o, int n, const float *origin, const float *dir, const uint32_t *masks, int32_t *body,
                float *local);
int oracle_contact_report(oracle_world *o, const int32_t *maxRep, int stride, int32_t *count, float *pos, float *normal,
                          int32_t *other);
int oracle_contains_point(int type, const float pos[3], const float rot[4], const float size[3], const float point[3],
                          float margin);
int oracle_intersect_point(oracle_world *o, int n, const float *points, const uint32_t *masks, int maxResults,
                           int32_t *count, int32_t *out);
int oracle_intersect_shape(oracle_world *o, int n, const int32_t *type, const float *center, const float *rot,
                           const float *size, const float *margin, const uint32_t *masks, int maxResults,
                           int32_t *count, int32_t *out);
int oracle_sweep_shape(int typeA, const float posA[3], const float rotA[4], const float sizeA[3],
                       const float motion[3], int typeB, const float posB[3], const float rotB[4],
                       const float sizeB[3], float margin, float out[9]);
int oracle_cast_motion(oracle_world *o, int n, const int32_t *type, const float *center, const float *rot,
                       const float *size, const float *motion, const float *margin, const uint32_t *masks,
                       int32_t *body, float *frac, float *point, float *normal);
int oracle_solve6(const float *a36, const float *b6, float *x6);
int oracle_collide_boxes(const float posA[3], const float rotA[4], const float sizeA[3],
                         const float posB[3], const float rotB[4], const float sizeB[3],
                         float margin, float basis[9], int32_t keys[8], float rA[24], float rB[24]);
int oracle_collide_shapes(int typeA, const float posA[3], const float rotA[4], const float sizeA[3],
                          int typeB, const float posB[3], const float rotB[4], const float sizeB[3],
                          float margin, float basis[9], int32_t keys[8], float rA[24], float rB[24]);

#ifdef __cplusplus
}
#endif

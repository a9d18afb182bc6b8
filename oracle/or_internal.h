/* Oracle-internal entry points shared between its translation units. */
#ifndef DEVILRAY_OR_INTERNAL_H
#define DEVILRAY_OR_INTERNAL_H

#include "oracle.h"

/* One ray through the canonical BVH (build semantics). */
void or_trace_one(const void *bvh, const float *verts, const int32_t *tris, const or_ray *ray, int any,
                  or_hit *hit, uint32_t *occluded);
/* The same, also returning its box tests (1 for the root, then 2 per
 * interior node visited, so visits = (box_tests - 1) / 2) and triangle tests. */
void or_trace_one_counted(const void *bvh, const float *verts, const int32_t *tris, const or_ray *ray, int any,
                          or_hit *hit, uint32_t *occluded, int32_t *box_tests, int32_t *tri_tests);

/* One ray through a reference-semantics BBH (the stackless bit-trail walk,
 * getIntersectionImpl, IntersectionTestsImpl.hpp:141-220): the first
 * strictly nearest hit's local triangle id, t and barycentrics (wA, wB, wC);
 * box and triangle tests are added to the counters when given. */
int or_refbbh_closest_one(const void *bbh, const float *verts, const int32_t *tris, const float p[3],
                          const float d[3], int32_t *id_out, float *t_out, float bari_out[3], int64_t *box_tests,
                          int64_t *tri_tests);

/* Philox path random numbers and samplers shared by the two integrators. */
void or_path_rng(uint32_t pixel, uint32_t sample, uint32_t vertex, uint32_t stream, uint64_t seed, uint32_t out[4]);
/* The camera ray of a pixel sample (jittered by the Philox camera stream
 * when the camera samples uniformly, else through the pixel centre). */
void or_camera_sample(const or_camera *cam, uint64_t seed, uint32_t pixel, uint32_t sample, or_ray *out);

#endif

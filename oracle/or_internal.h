/* Oracle-internal entry points shared between its translation units. */
#ifndef DEVILRAY_OR_INTERNAL_H
#define DEVILRAY_OR_INTERNAL_H

#include "oracle.h"

/* One ray through the canonical BVH (build semantics). */
void or_trace_one(const void *bvh, const float *verts, const int32_t *tris, const or_ray *ray, int any,
                  or_hit *hit, uint32_t *occluded);

#endif

// Marching cubes.
#pragma once
#include "common.h"

namespace dwg {
// Returns the triangle count; throws Error(DWG_ERR_CAPACITY, value = count)
// when it exceeds capacity (nothing written then).
int64_t marching_cubes(dwg_ctx* ctx, const float* field, const dwg_grid* g, float iso, int64_t capacity, float* tris,
                       float* cent, float* nrm);
// The same triangles' centroids and normals in one pass, in no fixed order;
// throws Error(DWG_ERR_CAPACITY, value = count) when the count exceeds
// capacity (the buffers then hold an unspecified part of the triangles).
int64_t marching_cubes_any_order(dwg_ctx* ctx, const float* field, const dwg_grid* g, float iso, int64_t capacity,
                                 float* cent, float* nrm);
}  // namespace dwg

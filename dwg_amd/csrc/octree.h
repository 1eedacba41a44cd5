// Octree construction on the GPU and the point-set helpers of the setup
// phase (normalisation, bounds, octree leaves, leaf normal smoothing).
#pragma once
#include "common.h"

struct dwg_octree;

namespace dwg {

// Octree over n device points (FP64, n×3) with optional per-point areas
// (device FP64, nullptr → 2e-6 each); octreeDWG split rule with leaf
// capacity leaf_size.  Returns an octree ready for octree_update().
dwg_octree* octree_build(dwg_ctx* ctx, const double* P, const double* A, int64_t n, int max_depth, int leaf_size);

// Leaves with more than min_num points, in breadth-first order
// (GetOctreeLeaves, DWGProcess.h:383-428): cell centres (n×3), widths and
// node ids; each output needs room for T->n_leaves entries (widths and
// nodes may be null).  Returns the count.
int64_t octree_leaves(dwg_ctx* ctx, const dwg_octree* T, int min_num, double* centres, double* widths, int* nodes);

// SmoothNormalstoOctreeLeaves (DWGProcess.h:430-472) with the whole-matrix
// Frobenius normalisation of DWGProcess.h:537.  P, N: FP64 points and
// normals (n×3) the octree was built over.
void leaf_smooth_normals(dwg_ctx* ctx, const dwg_octree* T, const int* nodes, const double* centres,
                         const double* widths, int64_t nl, const double* P, const double* N, float* out);

// ReadPointandNormals normalisation (DWGProcess.h:53-97): centre at the
// mean, divide by max_i 2|p_i − c|_∞; out64 (n×3) and optionally out32.
void normalize_points(dwg_ctx* ctx, const float* p, int64_t n, double* out64, float* out32, double center[3],
                      double* scale);

void bounds_f64(dwg_ctx* ctx, const double* P, int64_t n, double lo[3], double hi[3]);
void to_f32(dwg_ctx* ctx, const double* a, int64_t n, float* b);
void to_f64(dwg_ctx* ctx, const float* a, int64_t n, double* b);

}  // namespace dwg

// Normal diffusion, radii and normal initialisation.
#pragma once
#include "common.h"

namespace dwg {
void diffuse_normals(dwg_ctx* ctx, const dwg_knn_index* I, const float* src, int64_t n, const float* cent,
                     const float* tri_nrm, int64_t ntri, int k, int zero_k, int zero_threshold, float* nrm,
                     int64_t* h_nzero);
void radii(dwg_ctx* ctx, const dwg_knn_index* I, const float* p, int64_t n, int k, float rmin, float rmax, float* out);
void gauss_map(dwg_ctx* ctx, const float* p, int64_t n, const float* center, float* nrm);
void pca_normals(dwg_ctx* ctx, const dwg_knn_index* I, const float* p, int64_t n, int k, const float* viewpoint,
                 float* out);
void point_areas(dwg_ctx* ctx, const dwg_knn_index* I, const float* p, int64_t n, int k, const float* nrm, float* out);
void transfer_normals(dwg_ctx* ctx, const dwg_knn_index* I, const float* src_nrm, const float* q, int64_t nq, int k,
                      float* out);
void outlier_filter(dwg_ctx* ctx, float* area, int64_t n, float threshold);
// mean over i of 1 − |cos ∠(old_i, new_i)|, FP64 on the GPU (n×3 normals)
double normal_change(dwg_ctx* ctx, const float* old_nrm, const float* new_nrm, int64_t n);
}  // namespace dwg

// Winding-number sums: direct, octree build and Barnes-Hut traversal.
#pragma once
#include "common.h"
#include "scan.h"

namespace dwg {

// Octree in octreeDWG's node numbering, points sorted into contiguous node
// ranges.  P points to the caller's host doubles (not owned).
struct HostOctree {
    std::vector<int> first_child, begin, end, depth;
    std::vector<double> cm, r;  // 3 per node, 1 per node
    std::vector<double> cn, w;  // cell centre (3 per node) and cell width
    std::vector<int> perm;      // packed slot → caller index
    const double* P = nullptr;
    int64_t n_points = 0, n_leaves = 0;
    int max_level = 0;
};

HostOctree build_octree_host(const double* P, const double* A, int64_t n, int max_depth, int leaf_size);

}  // namespace dwg

struct dwg_octree {
    int64_t n_nodes = 0, n_points = 0, n_leaves = 0;
    int max_level = 0;
    bool updated = false;
    int4* info = nullptr;      // (first_child | −1, begin, end, depth)
    float4* cm_r = nullptr;    // (float(CM), R)
    float4* cm_lo = nullptr;   // (CM − float(CM), 0): residual of the rounded centre
    float4* ec = nullptr;      // (EC/4π, 0), refreshed per iteration
    uint2* ec_bf16 = nullptr;  // bf16 copy of EC/4π
    float4* sa = nullptr;      // packed sources in tree order (x, y, z, s)
    float4* sb = nullptr;      // (A·n/4π, A)
    int* perm = nullptr;       // tree slot → caller index
    dwg::Double3* an64 = nullptr;  // FP64 prefix sums of A·n (n+1)
    ~dwg_octree();
};

namespace dwg {

dwg_octree* octree_upload(dwg_ctx* ctx, const HostOctree& H);
void octree_update(dwg_ctx* ctx, dwg_octree* T, const float* nrm, const float* area, const float* radii);

void gwn_direct(dwg_ctx* ctx, const float* pts, const float* nrm, const float* area, const float* radii, int64_t n,
                const float* q, int64_t nq, float sigma, uint32_t flags, float* out_w, float* out_g);
void gwn_tree(dwg_ctx* ctx, const dwg_octree* T, const float* q, int64_t nq, float beta, float sigma, uint32_t flags,
              float* out);
void gwn_tree_grid(dwg_ctx* ctx, const dwg_octree* T, const dwg_grid* g, float beta, float sigma, uint32_t flags,
                   float* out);
void gwn_tree_terms(dwg_ctx* ctx, const dwg_octree* T, const float* q, int64_t nq, float beta, int64_t* terms);
int64_t gwn_tree_grid_terms(dwg_ctx* ctx, const dwg_octree* T, const dwg_grid* g, float beta);

}  // namespace dwg

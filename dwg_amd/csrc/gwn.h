// Winding-number sums: direct, octree build and Barnes-Hut traversal.
#pragma once
#include "common.h"
#include "scan.h"

namespace dwg {

// sources the allocation-free direct sums can take (DWG_NOALLOC_MAX_SOURCES)
constexpr int64_t kNoAllocSources = DWG_NOALLOC_MAX_SOURCES;

}  // namespace dwg

struct dwg_octree {
    int64_t n_nodes = 0, n_points = 0, n_leaves = 0;
    int max_level = 0, max_depth = 0;
    bool updated = false;
    std::vector<int64_t> level_start;  // nodes of depth d: [level_start[d], level_start[d+1])
    int4* info = nullptr;      // (first_child | −1, begin, end, depth)
    int2* kid = nullptr;       // traversal record: (first_child, 0) internal, (−count, begin) leaf
    float4* cm_r = nullptr;    // (float(CM), (β·R)² for the β of the last evaluation); one-point leaf: (point, guard)
    float* radius = nullptr;   // float(R)
    float far_beta = -1.f;     // β that cm_r.w was computed for
    float far_floor = -1.f;    // lower bound cm_r.w was clamped to (0 for the exact far test)
    float far_guard = 0.f;     // cm_r.w of one-point leaves: the point term's guard (1e-20 screened, −1 plain)
    float4* cm_lo = nullptr;   // (CM − float(CM), 0): residual of the rounded centre
    float4* ec = nullptr;      // (EC/4π, 0), refreshed per iteration

    float4* sa = nullptr;      // packed sources in tree order (x, y, z, s)
    float4* sb = nullptr;      // (A·n/4π, A)
    int* perm = nullptr;       // tree slot → caller index
    dwg::Double3* an64 = nullptr;  // FP64 prefix sums of A·n (n+1)
    double4* cell = nullptr;       // cell centre and width (breadth-first node order)
    double4* cm64 = nullptr;       // (CM, ΣA) in FP64
    // 8 words per node, read by the traversal for all 8 children of a node
    // in one vector load: internal node (EC/4π, s², CM − float(CM), 0) with
    // s the default screening radius; leaf with one point: that point's
    // (A·n/4π, s², 0, 0, 0, 0), its position in cm_r
    float4* payload = nullptr;
    bool payload_bf16 = false;  // internal nodes' EC/4π in the payload rounded to BF16
    ~dwg_octree();
};

namespace dwg {

void octree_update(dwg_ctx* ctx, dwg_octree* T, const float* nrm, const float* area, const float* radii);
// re-packs the point terms (A·n, A) with other areas, keeping the far-field
// moments of the last octree_update
void octree_set_point_areas(dwg_ctx* ctx, dwg_octree* T, const float* nrm, const float* area);

void gwn_direct(dwg_ctx* ctx, const float* pts, const float* nrm, const float* area, const float* radii, int64_t n,
                const float* q, int64_t nq, float sigma, uint32_t flags, float* out_w, float* out_g);
// device address of the fixed scratch array of `device` (must be current),
// looked up once per device; called when a context is created
void noalloc_init(int device);
// gwn_direct with the packed sources in that array: no allocation, n ≤
// kNoAllocSources; concurrent calls on one device share the array
void gwn_direct_noalloc(dwg_ctx* ctx, const float* pts, const float* nrm, const float* area, const float* radii,
                        int64_t n, const float* q, int64_t nq, float sigma, uint32_t flags, float* out_w,
                        float* out_g);
void gwn_tree(dwg_ctx* ctx, const dwg_octree* T, const float* q, int64_t nq, float beta, float sigma, uint32_t flags,
              float* out);
// sums at the tree's own points (evaluated in tree order), out in caller order
void gwn_tree_self(dwg_ctx* ctx, const dwg_octree* T, float beta, float sigma, uint32_t flags, float* out);
void gwn_tree_grid(dwg_ctx* ctx, const dwg_octree* T, const dwg_grid* g, float beta, float sigma, uint32_t flags,
                   float* out);
// value and gradient of the Barnes-Hut sums at query points (FP32 walk)
void gwn_tree_grad(dwg_ctx* ctx, const dwg_octree* T, const float* q, int64_t nq, float beta, float sigma,
                   uint32_t flags, float* out_w, float* out_g);
void gwn_tree_terms(dwg_ctx* ctx, const dwg_octree* T, const float* q, int64_t nq, float beta, int64_t* terms);
int64_t gwn_tree_grid_terms(dwg_ctx* ctx, const dwg_octree* T, const dwg_grid* g, float beta);

}  // namespace dwg

// Exact kNN on a uniform cell grid.
#pragma once
#include "common.h"

namespace dwg {
// value (3 floats per query) to add, in 2^-40 fixed point, into acc (3
// int64 per point) at every neighbour of the query
struct KnnScatter {
    const float* val;
    unsigned long long* acc;
};

struct KnnGridArgs {
    float lo[3];
    float cell, inv_cell;
    int dim[3];
};

// levels of the fine level's occupancy pyramid (up to 2^15 cells per axis)
constexpr int kMaxPyr = 16;
}  // namespace dwg

// Two cell grids over the same points: a fine one sized from the measured
// occupancy and a coarse one (8× larger cells) for queries whose fine search
// does not close within a few shells.
struct dwg_knn_index {
    int64_t n = 0;
    struct Level {
        dwg::KnnGridArgs G{};
        int64_t ncells = 0;
        int64_t* cell_start = nullptr;  // ncells + 1
        float4* pts = nullptr;          // (x, y, z, index bits) bucketed by cell
        int* empty = nullptr;           // coarse level: Chebyshev distance (cells) to the nearest occupied cell
        // fine level: occupancy pyramid.  Level l ≥ 1 has one byte per
        // block of 2^l cells per axis, bit c set when child block c (octant
        // bits x, y, z) of level l − 1 holds a point; level 0 are the cells.
        uint8_t* pyr = nullptr;
        int4* pyr_dim = nullptr;        // (blocks per axis, 0) of every level
        int64_t* pyr_off = nullptr;     // byte offset of every level in pyr
        int4 pyr_dim_h[dwg::kMaxPyr];   // host copies (source of the uploads)
        int64_t pyr_off_h[dwg::kMaxPyr];
        int pyr_levels = 0;             // 0: no pyramid (grid too elongated)
        float pyr_margin = 0.f;         // rounding allowance of cell boundaries
        // bounding box of the points of every cell (at 2·cell) and pyramid
        // block (at 2·(ncells + byte offset)): (lo, 0), (hi, 0); (∞, −∞) empty
        float4* pbox = nullptr;
    } lvl[2];
    ~dwg_knn_index();
};

namespace dwg {
void bounds(dwg_ctx* ctx, const float* p, int64_t n, float* lo, float* hi);
dwg_knn_index* knn_build(dwg_ctx* ctx, const float* p, int64_t n, int k_hint);
void knn_query(dwg_ctx* ctx, const dwg_knn_index* I, const float* q, int64_t nq, int k, const uint8_t* skip,
               const uint8_t* qmask, int32_t* out_idx, float* out_d2, const char* tag = nullptr,
               const KnnScatter* scatter = nullptr);
}  // namespace dwg

// C ABI of libdwg_hip.so: argument checks, exception → status translation,
// context lifetime.  Declarations and reference citations: include/dwg_hip.h.
#include <cstring>

#include "common.h"
#include "diffuse.h"
#include "gwn.h"
#include "knn.h"
#include "mc.h"
#include "octree.h"
#include "scan.h"

namespace {

template <class F>
int guarded(dwg_ctx* ctx, F&& fn) {
    try {
        fn();
        if (ctx) ctx->err.clear();
        return DWG_OK;
    } catch (const dwg::Error& e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return DWG_ERR_HIP;
    } catch (...) {
        if (ctx) ctx->err = "unknown error";
        return DWG_ERR_HIP;
    }
}

void set_device(dwg_ctx* ctx) { DWG_HIP_CHECK(hipSetDevice(ctx->device)); }

}  // namespace

extern "C" {

int dwg_ctx_create(int device, void* stream, dwg_ctx** out) {
    if (!out) return DWG_ERR_ARG;
    *out = nullptr;
    auto* ctx = new dwg_ctx();
    ctx->device = device;
    ctx->stream = static_cast<hipStream_t>(stream);
    int rc = guarded(ctx, [&] {
        int count = 0;
        DWG_HIP_CHECK(hipGetDeviceCount(&count));
        DWG_REQUIRE(device >= 0 && device < count, "no such HIP device");
        set_device(ctx);
        // per-device address of the allocation-free entry point's scratch array
        dwg::noalloc_init(device);
    });
    if (rc != DWG_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return DWG_OK;
}

int dwg_cache_release(int64_t* out_bytes) {
    try {
        const size_t freed = dwg::dev_cache_release();
        if (out_bytes) *out_bytes = int64_t(freed);
        return DWG_OK;
    } catch (...) {
        return DWG_ERR_HIP;
    }
}

void dwg_ctx_destroy(dwg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->times)
        for (auto& ab : kv.second.pending) {
            (void)hipEventDestroy(ab.first);
            (void)hipEventDestroy(ab.second);
        }
    ctx->scratch.clear();
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->readback) (void)hipEventDestroy(ctx->readback);
    delete ctx;
}

const char* dwg_ctx_error(const dwg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int dwg_ctx_sync(dwg_ctx* ctx) {
    if (!ctx) return DWG_ERR_ARG;
    return guarded(ctx, [&] { DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream)); });
}

int dwg_ctx_profile(dwg_ctx* ctx, int enable) {
    if (!ctx) return DWG_ERR_ARG;
    ctx->profile = enable != 0;
    return DWG_OK;
}

int dwg_ctx_profile_get(dwg_ctx* ctx, const char* name, double* total_ms, int64_t* launches) {
    if (!ctx || !name) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        auto it = ctx->times.find(name);
        double ms = 0;
        int64_t cnt = 0;
        if (it != ctx->times.end()) {
            auto& kt = it->second;
            for (auto& ab : kt.pending) {
                float t = 0;
                DWG_HIP_CHECK(hipEventElapsedTime(&t, ab.first, ab.second));
                kt.total_ms += t;
                kt.launches += 1;
                (void)hipEventDestroy(ab.first);
                (void)hipEventDestroy(ab.second);
            }
            kt.pending.clear();
            ms = kt.total_ms;
            cnt = kt.launches;
        }
        if (total_ms) *total_ms = ms;
        if (launches) *launches = cnt;
    });
}

int dwg_ctx_counter_get(dwg_ctx* ctx, const char* name, int64_t* value) {
    if (!ctx || !name || !value) return DWG_ERR_ARG;
    auto it = ctx->counters.find(name);
    *value = it == ctx->counters.end() ? 0 : it->second;
    return DWG_OK;
}

int dwg_ctx_profile_reset(dwg_ctx* ctx) {
    if (!ctx) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (auto& kv : ctx->times)
            for (auto& ab : kv.second.pending) {
                (void)hipEventDestroy(ab.first);
                (void)hipEventDestroy(ab.second);
            }
        ctx->times.clear();
        ctx->counters.clear();
    });
}

int dwg_gwn_direct(dwg_ctx* ctx, const float* d_pts, const float* d_nrm, const float* d_area, const float* d_radii,
                   int64_t n, const float* d_q, int64_t nq, float sigma, uint32_t flags, float* d_out_w,
                   float* d_out_grad) {
    if (!ctx) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(n == 0 || (d_pts && d_nrm), "null source arrays");
        DWG_REQUIRE(nq == 0 || (d_q && d_out_w), "null query/output arrays");
        set_device(ctx);
        dwg::gwn_direct(ctx, d_pts, d_nrm, d_area, d_radii, n, d_q, nq, sigma, flags, d_out_w, d_out_grad);
    });
}

int dwg_gwn_direct_noalloc(dwg_ctx* ctx, const float* d_pts, const float* d_nrm, const float* d_area,
                           const float* d_radii, int64_t n, const float* d_q, int64_t nq, float sigma, uint32_t flags,
                           float* d_out_w, float* d_out_grad) {
    if (!ctx) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(n == 0 || (d_pts && d_nrm), "null source arrays");
        DWG_REQUIRE(nq == 0 || (d_q && d_out_w), "null query/output arrays");
        set_device(ctx);
        dwg::gwn_direct_noalloc(ctx, d_pts, d_nrm, d_area, d_radii, n, d_q, nq, sigma, flags, d_out_w, d_out_grad);
    });
}

int dwg_octree_build(dwg_ctx* ctx, const double* h_pts, const double* h_area, int64_t n, int max_depth, int leaf_size,
                     dwg_octree** out) {
    if (!ctx || !out) return DWG_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        DWG_REQUIRE(h_pts, "null points");
        DWG_REQUIRE(n > 0, "octree needs at least one point");
        set_device(ctx);
        double* p = static_cast<double*>(ctx->buf("octree_upload_pts", sizeof(double) * 3 * size_t(n)));
        double* a = h_area ? static_cast<double*>(ctx->buf("octree_upload_area", sizeof(double) * size_t(n))) : nullptr;
        DWG_HIP_CHECK(hipMemcpyAsync(p, h_pts, sizeof(double) * 3 * size_t(n), hipMemcpyHostToDevice, ctx->stream));
        if (a) DWG_HIP_CHECK(hipMemcpyAsync(a, h_area, sizeof(double) * size_t(n), hipMemcpyHostToDevice, ctx->stream));
        *out = dwg::octree_build(ctx, p, a, n, max_depth, leaf_size);
    });
}

int dwg_octree_build_device(dwg_ctx* ctx, const double* d_pts, const double* d_area, int64_t n, int max_depth,
                            int leaf_size, dwg_octree** out) {
    if (!ctx || !out) return DWG_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        DWG_REQUIRE(d_pts, "null points");
        set_device(ctx);
        *out = dwg::octree_build(ctx, d_pts, d_area, n, max_depth, leaf_size);
    });
}

void dwg_octree_destroy(dwg_octree* tree) { delete tree; }

int dwg_octree_leaves(dwg_ctx* ctx, const dwg_octree* tree, int min_num, double* d_centres, double* d_widths,
                      int64_t* out_count) {
    if (!ctx || !tree || !out_count) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(d_centres, "null output");
        DWG_REQUIRE(min_num >= 0, "min_num must be ≥ 0");
        set_device(ctx);
        *out_count = dwg::octree_leaves(ctx, tree, min_num, d_centres, d_widths, nullptr);
    });
}

int dwg_octree_leaf_normals(dwg_ctx* ctx, const dwg_octree* tree, int min_num, const double* d_pts,
                            const double* d_nrm, float* d_out, int64_t* out_count) {
    if (!ctx || !tree || !out_count) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(d_pts && d_nrm && d_out, "null input or output");
        DWG_REQUIRE(min_num >= 0, "min_num must be ≥ 0");
        set_device(ctx);
        const int64_t nl_max = tree->n_leaves;
        auto* cen = static_cast<double*>(ctx->buf("leaf_normals_centres", sizeof(double) * 3 * size_t(nl_max + 1)));
        auto* wid = static_cast<double*>(ctx->buf("leaf_normals_widths", sizeof(double) * size_t(nl_max + 1)));
        auto* nodes = static_cast<int*>(ctx->buf("leaf_normals_nodes", sizeof(int) * size_t(nl_max + 1)));
        const int64_t nl = dwg::octree_leaves(ctx, tree, min_num, cen, wid, nodes);
        dwg::leaf_smooth_normals(ctx, tree, nodes, cen, wid, nl, d_pts, d_nrm, d_out);
        *out_count = nl;
    });
}

int dwg_octree_info(const dwg_octree* tree, int64_t* n_nodes, int64_t* n_leaves, int* max_level) {
    if (!tree) return DWG_ERR_ARG;
    if (n_nodes) *n_nodes = tree->n_nodes;
    if (n_leaves) *n_leaves = tree->n_leaves;
    if (max_level) *max_level = tree->max_level;
    return DWG_OK;
}

int dwg_octree_point_order(dwg_ctx* ctx, const dwg_octree* tree, int* d_order) {
    if (!ctx || !tree) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(d_order || tree->n_points == 0, "null output");
        set_device(ctx);
        DWG_HIP_CHECK(hipMemcpyAsync(d_order, tree->perm, sizeof(int) * size_t(tree->n_points),
                                     hipMemcpyDeviceToDevice, ctx->stream));
    });
}

int dwg_octree_update(dwg_ctx* ctx, dwg_octree* tree, const float* d_nrm, const float* d_area, const float* d_radii) {
    if (!ctx || !tree) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(d_nrm, "null normals");
        set_device(ctx);
        dwg::octree_update(ctx, tree, d_nrm, d_area, d_radii);
    });
}

int dwg_octree_set_point_areas(dwg_ctx* ctx, dwg_octree* tree, const float* d_nrm, const float* d_area) {
    if (!ctx || !tree) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(d_nrm, "null normals");
        set_device(ctx);
        dwg::octree_set_point_areas(ctx, tree, d_nrm, d_area);
    });
}

int dwg_gwn_tree(dwg_ctx* ctx, const dwg_octree* tree, const float* d_q, int64_t nq, float beta, float sigma,
                 uint32_t flags, float* d_out) {
    if (!ctx || !tree) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(nq == 0 || d_out, "null output");
        set_device(ctx);
        dwg::gwn_tree(ctx, tree, d_q, nq, beta, sigma, flags, d_out);
    });
}

int dwg_gwn_tree_grad(dwg_ctx* ctx, const dwg_octree* tree, const float* d_q, int64_t nq, float beta, float sigma,
                      uint32_t flags, float* d_out_w, float* d_out_grad) {
    if (!ctx || !tree) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(nq == 0 || (d_q && d_out_w), "null query or output array");
        set_device(ctx);
        dwg::gwn_tree_grad(ctx, tree, d_q, nq, beta, sigma, flags, d_out_w, d_out_grad);
    });
}

int dwg_gwn_tree_self(dwg_ctx* ctx, const dwg_octree* tree, float beta, float sigma, uint32_t flags, float* d_out) {
    if (!ctx || !tree) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(d_out, "null output");
        set_device(ctx);
        dwg::gwn_tree_self(ctx, tree, beta, sigma, flags, d_out);
    });
}

int dwg_gwn_tree_grid(dwg_ctx* ctx, const dwg_octree* tree, const dwg_grid* grid, float beta, float sigma,
                      uint32_t flags, float* d_out) {
    if (!ctx || !tree) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(d_out, "null output");
        set_device(ctx);
        dwg::gwn_tree_grid(ctx, tree, grid, beta, sigma, flags, d_out);
    });
}

int dwg_gwn_tree_terms(dwg_ctx* ctx, const dwg_octree* tree, const float* d_q, int64_t nq, float beta,
                       int64_t* d_terms) {
    if (!ctx || !tree) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(nq == 0 || d_terms, "null output");
        set_device(ctx);
        dwg::gwn_tree_terms(ctx, tree, d_q, nq, beta, d_terms);
    });
}

int dwg_gwn_tree_grid_terms(dwg_ctx* ctx, const dwg_octree* tree, const dwg_grid* grid, float beta,
                            int64_t* h_total_terms) {
    if (!ctx || !tree || !h_total_terms) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        set_device(ctx);
        *h_total_terms = dwg::gwn_tree_grid_terms(ctx, tree, grid, beta);
    });
}

int dwg_knn_build(dwg_ctx* ctx, const float* d_pts, int64_t n, int k_hint, dwg_knn_index** out) {
    if (!ctx || !out) return DWG_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        DWG_REQUIRE(d_pts, "null points");
        set_device(ctx);
        *out = dwg::knn_build(ctx, d_pts, n, k_hint);
    });
}

void dwg_knn_destroy(dwg_knn_index* index) { delete index; }

int dwg_knn_query(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_q, int64_t nq, int k, const uint8_t* d_skip,
                  const uint8_t* d_qmask, int32_t* d_idx, float* d_d2) {
    if (!ctx || !index) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(nq == 0 || (d_q && d_idx), "null query/output arrays");
        set_device(ctx);
        dwg::knn_query(ctx, index, d_q, nq, k, d_skip, d_qmask, d_idx, d_d2);
    });
}

int dwg_marching_cubes(dwg_ctx* ctx, const float* d_field, const dwg_grid* grid, float iso, int64_t capacity,
                       float* d_tris, float* d_cent, float* d_nrm, int64_t* h_count) {
    if (!ctx) return DWG_ERR_ARG;
    try {
        DWG_REQUIRE(d_field && grid, "null field/grid");
        set_device(ctx);
        int64_t t = dwg::marching_cubes(ctx, d_field, grid, iso, capacity, d_tris, d_cent, d_nrm);
        if (h_count) *h_count = t;
        ctx->err.clear();
        return DWG_OK;
    } catch (const dwg::Error& e) {
        ctx->err = e.what();
        if (e.code == DWG_ERR_CAPACITY && h_count) *h_count = e.value;
        return e.code;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return DWG_ERR_HIP;
    }
}

int dwg_marching_cubes_any_order(dwg_ctx* ctx, const float* d_field, const dwg_grid* grid, float iso,
                                 int64_t capacity, float* d_cent, float* d_nrm, int64_t* h_count) {
    if (!ctx) return DWG_ERR_ARG;
    try {
        DWG_REQUIRE(d_field && grid, "null field/grid");
        set_device(ctx);
        int64_t t = dwg::marching_cubes_any_order(ctx, d_field, grid, iso, capacity, d_cent, d_nrm);
        if (h_count) *h_count = t;
        ctx->err.clear();
        return DWG_OK;
    } catch (const dwg::Error& e) {
        ctx->err = e.what();
        if (e.code == DWG_ERR_CAPACITY && h_count) *h_count = e.value;
        return e.code;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return DWG_ERR_HIP;
    }
}

int dwg_diffuse_normals(dwg_ctx* ctx, const dwg_knn_index* src_index, const float* d_src, int64_t n_src,
                        const float* d_cent, const float* d_tri_nrm, int64_t ntri, int k, int zero_k,
                        int zero_threshold, float* d_nrm, int64_t* h_nzero) {
    if (!ctx || !src_index) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(d_src && d_nrm && (ntri == 0 || (d_cent && d_tri_nrm)), "null arrays");
        set_device(ctx);
        dwg::diffuse_normals(ctx, src_index, d_src, n_src, d_cent, d_tri_nrm, ntri, k, zero_k, zero_threshold, d_nrm,
                             h_nzero);
    });
}

int dwg_radii(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_pts, int64_t n, int k, float rmin, float rmax,
              float* d_out) {
    if (!ctx || !index) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(k >= 1 && k <= 32, "k must be in [1, 32]");
        set_device(ctx);
        dwg::radii(ctx, index, d_pts, n, k, rmin, rmax, d_out);
    });
}

int dwg_pca_normals(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_pts, int64_t n, int k,
                    const float* h_viewpoint, float* d_out) {
    if (!ctx || !index) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(n == 0 || (d_pts && d_out && h_viewpoint), "null arrays");
        set_device(ctx);
        if (n > 0) dwg::pca_normals(ctx, index, d_pts, n, k, h_viewpoint, d_out);
    });
}

int dwg_point_areas(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_pts, int64_t n, int k,
                    const float* d_nrm, float* d_out) {
    if (!ctx || !index) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(n == 0 || (d_pts && d_out), "null arrays");
        set_device(ctx);
        if (n > 0) dwg::point_areas(ctx, index, d_pts, n, k, d_nrm, d_out);
    });
}

int dwg_gauss_map(dwg_ctx* ctx, const float* d_pts, int64_t n, const float* h_center, float* d_nrm) {
    if (!ctx) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(h_center && (n == 0 || (d_pts && d_nrm)), "null arrays");
        set_device(ctx);
        dwg::gauss_map(ctx, d_pts, n, h_center, d_nrm);
    });
}

int dwg_transfer_normals(dwg_ctx* ctx, const dwg_knn_index* src_index, const float* d_src_nrm, const float* d_q,
                         int64_t nq, int k, float* d_out_nrm) {
    if (!ctx || !src_index) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(k >= 1 && k <= 32, "k must be in [1, 32]");
        set_device(ctx);
        dwg::transfer_normals(ctx, src_index, d_src_nrm, d_q, nq, k, d_out_nrm);
    });
}

int dwg_outlier_filter(dwg_ctx* ctx, float* d_area, int64_t n, float threshold) {
    if (!ctx) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        set_device(ctx);
        dwg::outlier_filter(ctx, d_area, n, threshold);
    });
}

int dwg_normal_change(dwg_ctx* ctx, const float* d_old, const float* d_new, int64_t n, double* h_residual) {
    if (!ctx || !h_residual) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        DWG_REQUIRE(n == 0 || (d_old && d_new), "null normals");
        set_device(ctx);
        *h_residual = dwg::normal_change(ctx, d_old, d_new, n);
    });
}

int dwg_reduce_sum(dwg_ctx* ctx, const float* d_vals, int64_t n, double* h_sum) {
    if (!ctx || !h_sum) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        set_device(ctx);
        *h_sum = n > 0 ? dwg::reduce_sum_f32(ctx, d_vals, n) : 0.0;
    });
}

int dwg_bounds(dwg_ctx* ctx, const float* d_pts, int64_t n, float* h_lo, float* h_hi) {
    if (!ctx || !h_lo || !h_hi) return DWG_ERR_ARG;
    return guarded(ctx, [&] {
        set_device(ctx);
        dwg::bounds(ctx, d_pts, n, h_lo, h_hi);
    });
}

}  // extern "C"

// Host-side DWG driver over the C ABI: TurboDWG (src/TurboDWG.cu) and the
// test3D orchestration (include/DWGProcess.h), without Eigen/libigl/PCL.
#include "turbo_dwg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <memory>
#include <mutex>
#include <random>
#include <sstream>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

#include "common.h"
#include "diffuse.h"
#include "gwn.h"
#include "knn.h"
#include "octree.h"
#include "scan.h"

namespace dwg_host {

using Clock = std::chrono::steady_clock;

static double seconds_since(Clock::time_point t) {
    return std::chrono::duration<double>(Clock::now() - t).count();
}

static std::mutex g_pinned_mu;
static std::unordered_map<void*, size_t> g_pinned;             // live pinned buffers and their sizes
static std::vector<std::pair<void*, size_t>> g_pinned_kept;    // released ones, kept for reuse
constexpr size_t kPinnedKept = 2;

// Pinned host memory for meshes handed to the caller.  Pinning a buffer
// costs milliseconds per hundred MB, so up to two released buffers are kept
// and reused by later allocations of up to the same size.
float* alloc_host_pinned(size_t bytes) {
    {
        std::lock_guard<std::mutex> lock(g_pinned_mu);
        size_t best = g_pinned_kept.size();
        for (size_t i = 0; i < g_pinned_kept.size(); ++i) {
            const size_t sz = g_pinned_kept[i].second;
            if (sz >= bytes && sz <= 2 * bytes + (size_t(1) << 20) &&
                (best == g_pinned_kept.size() || sz < g_pinned_kept[best].second))
                best = i;
        }
        if (best < g_pinned_kept.size()) {
            const auto kept = g_pinned_kept[best];
            g_pinned_kept.erase(g_pinned_kept.begin() + std::ptrdiff_t(best));
            g_pinned[kept.first] = kept.second;
            return static_cast<float*>(kept.first);
        }
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p)
        throw dwg::Error(DWG_ERR_HIP, "pinned host allocation failed");
    std::lock_guard<std::mutex> lock(g_pinned_mu);
    g_pinned[p] = bytes;
    return static_cast<float*>(p);
}

void free_host(void* p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lock(g_pinned_mu);
        const auto it = g_pinned.find(p);
        if (it != g_pinned.end()) {
            const size_t sz = it->second;
            g_pinned.erase(it);
            if (g_pinned_kept.size() < kPinnedKept) {
                g_pinned_kept.emplace_back(p, sz);
            } else {
                (void)hipHostFree(p);
            }
            return;
        }
    }
    std::free(p);
}

template <class T>
DevArray<T>::~DevArray() {
    dwg::dev_free(ptr);
}

template <class T>
void DevArray<T>::resize(int64_t n) {
    if (n <= count && ptr) return;
    dwg::dev_free(ptr);
    ptr = nullptr;
    count = 0;
    ptr = static_cast<T*>(dwg::dev_malloc(sizeof(T) * size_t(std::max<int64_t>(n, 1))));
    count = n;
}

template <class T>
void DevArray<T>::upload(const T* h, int64_t n, void* stream) {
    resize(n);
    if (n > 0 && hipMemcpyAsync(ptr, h, sizeof(T) * n, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)) != hipSuccess)
        throw dwg::Error(DWG_ERR_HIP, "upload failed");
    // the host buffer is pageable and often a temporary: wait for the copy
    if (n > 0 && hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
        throw dwg::Error(DWG_ERR_HIP, "sync failed");
}

template <class T>
void DevArray<T>::download(T* h, int64_t n, void* stream) const {
    if (n > 0 && hipMemcpyAsync(h, ptr, sizeof(T) * n, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)) != hipSuccess)
        throw dwg::Error(DWG_ERR_HIP, "download failed");
    if (hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess) throw dwg::Error(DWG_ERR_HIP, "sync failed");
}

template struct DevArray<float>;
template struct DevArray<double>;
template struct DevArray<int32_t>;
template struct DevArray<int64_t>;
template struct DevArray<uint16_t>;

dwg_grid Grid::abi() const {
    dwg_grid g;
    for (int d = 0; d < 3; ++d) {
        g.origin[d] = float(origin[d]);
        g.res[d] = res[d];
    }
    g.h = float(h);
    return g;
}

Grid voxel_grid(const double* V, int64_t n, double offset, int s) {
    if (n <= 0) throw dwg::Error(DWG_ERR_ARG, "voxel grid of no points");
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], V[3 * i + d]);
            hi[d] = std::max(hi[d], V[3 * i + d]);
        }
    return voxel_grid_bounds(lo, hi, offset, s);
}

Grid voxel_grid_bounds(const double lo_in[3], const double hi_in[3], double offset, int s) {
    if (s < 4) throw dwg::Error(DWG_ERR_ARG, "grid resolution must be ≥ 4");
    double lo[3] = {lo_in[0], lo_in[1], lo_in[2]}, hi[3] = {hi_in[0], hi_in[1], hi_in[2]};
    double ext[3], ctr[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] -= offset;
        hi[d] += offset;
        ext[d] = hi[d] - lo[d];
        ctr[d] = (lo[d] + hi[d]) / 2;
    }
    int si = 0;
    for (int d = 1; d < 3; ++d)
        if (ext[d] > ext[si]) si = d;
    // the longest side gets s samples, the others ceil((s−2)·ext/ext_max)+2,
    // one padding sample each side; spacing from the largest per-axis scale
    double side[3];
    for (int d = 0; d < 3; ++d) side[d] = (d == si ? double(s - 2) : std::ceil(double(s - 2) * ext[d] / ext[si])) + 2;
    double amax = -1;
    int ai = 0;
    for (int d = 0; d < 3; ++d) {
        const double a = ext[d] / (1.0 - 2.0 / (side[d] - 1.0));
        if (a > amax) {
            amax = a;
            ai = d;
        }
    }
    Grid g;
    g.h = amax / (side[ai] - 1.0);
    for (int d = 0; d < 3; ++d) {
        g.res[d] = int(side[d]);
        g.origin[d] = ctr[d] - (side[d] - 1.0) / 2.0 * g.h;
    }
    return g;
}

void save_binary_stl(const std::string& path, const float* tris, int64_t ntri) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw dwg::Error(DWG_ERR_ARG, "cannot create " + path);
    char header[80] = {0};
    f.write(header, 80);
    const uint32_t n = uint32_t(ntri);
    f.write(reinterpret_cast<const char*>(&n), 4);
    const char pad[2] = {0, 0};
    for (int64_t t = 0; t < ntri; ++t) {
        const float* v = tris + 9 * t;
        const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const float l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (float& x : c) x = l > 0 ? x / l : 0.f;
        f.write(reinterpret_cast<const char*>(c), 12);
        f.write(reinterpret_cast<const char*>(v), 36);
        f.write(pad, 2);
    }
    if (!f.good()) throw dwg::Error(DWG_ERR_ARG, "cannot write " + path);
}

void save_xyz(const std::string& path, const float* pts, const float* nrm, int64_t n) {
    std::ofstream f(path);
    if (!f) throw dwg::Error(DWG_ERR_ARG, "cannot create " + path);
    for (int64_t i = 0; i < n; ++i)
        f << pts[3 * i] << " " << pts[3 * i + 1] << " " << pts[3 * i + 2] << " " << nrm[3 * i] << " " << nrm[3 * i + 1]
          << " " << nrm[3 * i + 2] << "\n";
}

bool read_xyz(const std::string& path, std::vector<float>& pts, std::vector<float>& nrm) {
    std::ifstream in(path);
    if (!in) return false;
    pts.clear();
    nrm.clear();
    for (std::string line; std::getline(in, line);) {
        std::istringstream ls(line);
        double v[6] = {0, 0, 0, 0, 0, 0};
        if (!(ls >> v[0] >> v[1] >> v[2])) continue;
        ls >> v[3] >> v[4] >> v[5];
        for (int d = 0; d < 3; ++d) {
            pts.push_back(float(v[d]));
            nrm.push_back(float(v[3 + d]));
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
TurboDWG::TurboDWG(dwg_ctx* ctx, const dwg_options& opt) : ctx_(ctx), opt_(opt) {}

TurboDWG::~TurboDWG() {
    if (tree_) dwg_octree_destroy(tree_);
    if (knn_) dwg_knn_destroy(knn_);
}

void TurboDWG::check(int rc, const char* what) const {
    if (rc != DWG_OK) throw dwg::Error(rc, std::string(what) + ": " + dwg_ctx_error(ctx_));
}

void TurboDWG::SetSources(const double* d_pts, int64_t n, const float* d_nrm, const Grid& grid,
                          dwg_octree* prebuilt, const float* d_area) {
    std::unique_ptr<dwg_octree, void (*)(dwg_octree*)> owned(prebuilt, dwg_octree_destroy);
    n_ = n;
    hipStream_t s = ctx_->stream;
    pts_.resize(3 * n_);
    dwg::to_f32(ctx_, d_pts, 3 * n_, pts_.ptr);
    nrm_.resize(3 * n_);
    if (d_nrm) {
        if (hipMemcpyAsync(nrm_.ptr, d_nrm, sizeof(float) * 3 * n_, hipMemcpyDeviceToDevice, s) != hipSuccess)
            throw dwg::Error(DWG_ERR_HIP, "normal copy failed");
    } else {
        dwg::fill_u32(ctx_, nrm_.ptr, 0u, 3 * n_);
    }
    area_.resize(n_);
    const float a0 = 2e-6f, r0 = 0.0015f;
    uint32_t bits;
    std::memcpy(&bits, &a0, 4);
    dwg::fill_u32(ctx_, area_.ptr, bits, n_);
    radii_.resize(n_);
    std::memcpy(&bits, &r0, 4);
    dwg::fill_u32(ctx_, radii_.ptr, bits, n_);
    w_src_.resize(n_);
    grid_ = grid;
    w_grid_.resize(grid_.size());
    if (tree_) dwg_octree_destroy(tree_);
    tree_ = nullptr;
    if (knn_) dwg_knn_destroy(knn_);
    knn_ = nullptr;
    if (d_area) {
        // computed areas weight CM as well (fast_winding_number with As, DWGProcess.h:711-722)
        SetAreas(d_area);
        DevArray<double> a64(n_);
        dwg::to_f64(ctx_, area_.ptr, n_, a64.ptr);
        check(dwg_octree_build_device(ctx_, d_pts, a64.ptr, n_, opt_.max_depth, opt_.leaf_size, &tree_), "octree");
    } else if (owned && opt_.leaf_size == 1 && owned->n_points == n_) {
        tree_ = owned.release();
    } else {
        check(dwg_octree_build_device(ctx_, d_pts, nullptr, n_, opt_.max_depth, opt_.leaf_size, &tree_), "octree");
    }
}

dwg_knn_index* TurboDWG::index() {
    if (!knn_) check(dwg_knn_build(ctx_, pts_.ptr, n_, 15, &knn_), "knn index");
    return knn_;
}

void TurboDWG::SetAreas(const float* d_area) {
    if (hipMemcpyAsync(area_.ptr, d_area, sizeof(float) * n_, hipMemcpyDeviceToDevice, ctx_->stream) != hipSuccess)
        throw dwg::Error(DWG_ERR_HIP, "area copy failed");
}

void TurboDWG::UpdateRadii(int k) {
    // without --use_radii no sum reads the radii (the reference computes them
    // regardless, DWGProcess.h:283-343)
    if (!opt_.use_radii) return;
    check(dwg_radii(ctx_, index(), pts_.ptr, n_, k, 0.0015f, 0.015f, radii_.ptr), "radii");
}

void TurboDWG::GaussMapInit() {
    float lo[3], hi[3];
    check(dwg_bounds(ctx_, pts_.ptr, n_, lo, hi), "bounds");
    const float c[3] = {(lo[0] + hi[0]) / 2.0f, (lo[1] + hi[1]) / 2.0f, (lo[2] + hi[2]) / 2.0f};
    check(dwg_gauss_map(ctx_, pts_.ptr, n_, c, nrm_.ptr), "gauss map");
}

void TurboDWG::MlpInit(const dwg_mlp* mlp) {
    // the predicted normal line, oriented like the Gauss map: away from the
    // bounding-box centre (TurboNormalInit.cu:3-44)
    float lo[3], hi[3];
    check(dwg_bounds(ctx_, pts_.ptr, n_, lo, hi), "bounds");
    const float c[3] = {(lo[0] + hi[0]) / 2.0f, (lo[1] + hi[1]) / 2.0f, (lo[2] + hi[2]) / 2.0f};
    check(dwg_mlp_init(ctx_, mlp, pts_.ptr, n_, c, nrm_.ptr), "mlp init");
}

void TurboDWG::download_normals(std::vector<float>& out) const {
    out.resize(3 * n_);
    nrm_.download(out.data(), 3 * n_, ctx_->stream);
}

void TurboDWG::set_normals(const std::vector<float>& n) { nrm_.upload(n.data(), 3 * n_, ctx_->stream); }

static float sigma_schedule(int it, float sigma_in, int warm_iters) {
    // TurboDWG.cu:526-533 (warm_iters = 4 there)
    float s = sigma_in;
    if (it < warm_iters) s = 20.0f;
    if (it > 0 && it % 7 == 0) s = 10.0f;
    return s;
}

double TurboDWG::TurboScreenedDWGLaunch(int total_iters, double* w_mean, int64_t* iter_tris, int64_t* iter_zeros,
                                        double* residual, int* iters_run) {
    const auto t0 = Clock::now();
    if (opt_.As_threshold > 0) check(dwg_outlier_filter(ctx_, area_.ptr, n_, opt_.As_threshold), "outlier filter");
    const uint32_t flags = DWG_SCREENED | (opt_.use_radii ? DWG_USE_RADII : 0u) | opt_.precision;
    const dwg_grid g = grid_.abi();
    DevArray<float> grid_pts;
    if (opt_.beta <= 0) {  // direct sums need explicit grid points
        std::vector<float> gp(3 * grid_.size());
        for (int64_t k = 0; k < grid_.res[2]; ++k)
            for (int64_t j = 0; j < grid_.res[1]; ++j)
                for (int64_t i = 0; i < grid_.res[0]; ++i) {
                    const int64_t id = i + grid_.res[0] * (j + grid_.res[1] * k);
                    gp[3 * id] = std::fma(float(i), g.h, g.origin[0]);
                    gp[3 * id + 1] = std::fma(float(j), g.h, g.origin[1]);
                    gp[3 * id + 2] = std::fma(float(k), g.h, g.origin[2]);
                }
        grid_pts.upload(gp.data(), int64_t(gp.size()), ctx_->stream);
    }
    int64_t capacity = std::max<int64_t>(cent_.count / 3, 1 << 20);
    // the convergence residual compares each iteration's normals with the
    // previous ones (only when it is asked for or stops the loop)
    const bool want_residual = residual || opt_.residual_tol > 0;
    DevArray<float> prev;
    if (want_residual) prev.resize(3 * n_);
    int it = 0;
    for (; it < total_iters; ++it) {
        const float sigma = sigma_schedule(it, opt_.sigma, opt_.sigma_warm_iters);
        const float* rad = opt_.use_radii ? radii_.ptr : nullptr;
        check(dwg_octree_update(ctx_, tree_, nrm_.ptr, area_.ptr, rad), "moments");
        // the sources' mean (the iso value) is read back while the grid is
        // evaluated, so marching cubes queues right behind the grid
        if (opt_.beta > 0) {
            check(dwg_gwn_tree_self(ctx_, tree_, opt_.beta, sigma, flags, w_src_.ptr), "gwn sources");
            dwg::reduce_sum_f32_async(ctx_, w_src_.ptr, n_);
            check(dwg_gwn_tree_grid(ctx_, tree_, &g, opt_.beta, sigma, flags, w_grid_.ptr), "gwn grid");
        } else {
            check(dwg_gwn_direct(ctx_, pts_.ptr, nrm_.ptr, area_.ptr, rad, n_, pts_.ptr, n_, sigma, flags, w_src_.ptr,
                                 nullptr), "gwn sources");
            dwg::reduce_sum_f32_async(ctx_, w_src_.ptr, n_);
            check(dwg_gwn_direct(ctx_, pts_.ptr, nrm_.ptr, area_.ptr, rad, n_, grid_pts.ptr, grid_.size(), sigma, flags,
                                 w_grid_.ptr, nullptr), "gwn grid");
        }
        const double sum = dwg::reduce_sum_f32_wait(ctx_);
        float iso = float(sum / double(n_));
        if (w_mean) w_mean[it] = double(iso);
        if (it % 5 == 0 && it != 0) iso *= 0.5f;  // TurboDWG.cu:547-550
        int64_t ntri = 0;
        cent_.resize(3 * capacity);
        tri_nrm_.resize(3 * capacity);
        // the diffusion's sums do not depend on the triangles' order: one pass
        int rc = dwg_marching_cubes_any_order(ctx_, w_grid_.ptr, &g, iso, capacity, cent_.ptr, tri_nrm_.ptr, &ntri);
        if (rc == DWG_ERR_CAPACITY) {
            capacity = ntri + ntri / 4 + 1024;
            cent_.resize(3 * capacity);
            tri_nrm_.resize(3 * capacity);
            rc = dwg_marching_cubes_any_order(ctx_, w_grid_.ptr, &g, iso, capacity, cent_.ptr, tri_nrm_.ptr, &ntri);
        }
        check(rc, "marching cubes");
        if (iter_tris) iter_tris[it] = ntri;
        if (!dump_dir_.empty() && opt_.test_iter != 999999 && opt_.test_iter > 0 && it % opt_.test_iter == 0)
            dump_mesh(g, iso, ntri, it);
        int64_t nz = 0;
        if (want_residual)
            DWG_HIP_CHECK(hipMemcpyAsync(prev.ptr, nrm_.ptr, sizeof(float) * 3 * n_, hipMemcpyDeviceToDevice,
                                         ctx_->stream));
        check(dwg_diffuse_normals(ctx_, index(), pts_.ptr, n_, cent_.ptr, tri_nrm_.ptr, ntri, 15, 10, 10, nrm_.ptr, &nz),
              "diffuse");
        if (iter_zeros) iter_zeros[it] = nz;
        double res = -1.0;
        if (want_residual) {
            check(dwg_normal_change(ctx_, prev.ptr, nrm_.ptr, n_, &res), "residual");
            if (residual) residual[it] = res;
        }
        if (opt_.verbose) std::printf("iter = %d  W_mean = %.6g  triangles = %lld  zero normals = %lld  residual = %.3g\n",
                                      it, double(w_mean ? w_mean[it] : iso), (long long)ntri, (long long)nz, res);
        // an even count keeps the start's side (the sign flips every iteration)
        if (opt_.residual_tol > 0 && (it + 1) % 2 == 0 && res < double(opt_.residual_tol)) {
            ++it;
            break;
        }
    }
    if (iters_run) *iters_run = it;
    check(dwg_ctx_sync(ctx_), "sync");
    return seconds_since(t0);
}

void TurboDWG::dump_mesh(const dwg_grid& g, float iso, int64_t ntri, int it) {
    // the iteration's mesh again, now with its vertices (the iteration itself
    // only keeps centroids and face normals)
    DevArray<float> tris(9 * std::max<int64_t>(ntri, 1)), c(3 * std::max<int64_t>(ntri, 1)),
        nr(3 * std::max<int64_t>(ntri, 1));
    int64_t n2 = 0;
    check(dwg_marching_cubes(ctx_, w_grid_.ptr, &g, iso, std::max<int64_t>(ntri, 1), tris.ptr, c.ptr, nr.ptr, &n2),
          "marching cubes (dump)");
    std::vector<float> h(size_t(9 * n2));
    if (n2 > 0) tris.download(h.data(), 9 * n2, ctx_->stream);
    const std::time_t now = std::time(nullptr);
    const std::tm lt = *std::localtime(&now);
    const std::string name = dump_dir_ + std::to_string(lt.tm_mon + 1) + std::to_string(lt.tm_mday) +
                             std::to_string(lt.tm_hour) + std::to_string(lt.tm_min) + std::to_string(lt.tm_sec) + "_" +
                             std::to_string(it) + "iter.stl";
    save_binary_stl(name, h.data(), n2);
    std::printf("Mesh saved to %s (%lld triangle(s).\n", name.c_str(), (long long)n2);
}

void TurboDWG::CountTerms(int64_t* src_terms, int64_t* grid_terms) {
    *src_terms = 0;
    *grid_terms = 0;
    if (opt_.beta <= 0) {  // direct sums: every pair
        *src_terms = n_ * n_;
        *grid_terms = n_ * grid_.size();
        return;
    }
    check(dwg_octree_update(ctx_, tree_, nrm_.ptr, area_.ptr, nullptr), "moments");
    DevArray<int64_t> t(n_);
    check(dwg_gwn_tree_terms(ctx_, tree_, pts_.ptr, n_, opt_.beta, t.ptr), "terms");
    std::vector<int64_t> h(n_);
    t.download(h.data(), n_, ctx_->stream);
    for (int64_t v : h) *src_terms += v;
    const dwg_grid g = grid_.abi();
    check(dwg_gwn_tree_grid_terms(ctx_, tree_, &g, opt_.beta, grid_terms), "grid terms");
}

void TurboDWG::mesh_pass(const Grid& grid, float iso_scale, int64_t* tri_count, HostMesh* tris, bool filter_areas) {
    // one screened evaluation at the sources and on the grid with the
    // unscheduled σ, iso = mean source value, marching cubes
    struct Tag {  // kernels of the final pass are profiled as "<kernel>:final"
        dwg_ctx* c;
        explicit Tag(dwg_ctx* x) : c(x) { c->profile_suffix = ":final"; }
        ~Tag() { c->profile_suffix.clear(); }
    } tag(ctx_);
    // the mesh that is returned comes from FP32 far-field moments; BF16 is
    // confined to the iterations
    const uint32_t flags = DWG_SCREENED | (opt_.use_radii ? DWG_USE_RADII : 0u) | (opt_.precision & ~DWG_FAR_BF16);
    const dwg_grid g = grid.abi();
    // field and triangles in the context's scratch space: a second
    // reconstruction on the same context allocates nothing here
    float* wg = static_cast<float*>(ctx_->buf("final_field", sizeof(float) * size_t(grid.size())));
    const float* rad = opt_.use_radii ? radii_.ptr : nullptr;
    check(dwg_octree_update(ctx_, tree_, nrm_.ptr, area_.ptr, rad), "moments");
    if (filter_areas && opt_.As_threshold > 0) {
        // GenerateMesh filters the areas after the moment refresh
        // (TurboDWG.cu:870-875): the far-field moments keep the unfiltered
        // areas, the point terms take the filtered ones
        check(dwg_outlier_filter(ctx_, area_.ptr, n_, opt_.As_threshold), "outlier filter");
        check(dwg_octree_set_point_areas(ctx_, tree_, nrm_.ptr, area_.ptr), "point areas");
    }
    const float beta = opt_.beta > 0 ? opt_.beta : 1e30f;  // β → ∞ takes every point term
    check(dwg_gwn_tree_self(ctx_, tree_, beta, opt_.sigma, flags, w_src_.ptr), "gwn points");
    dwg::reduce_sum_f32_async(ctx_, w_src_.ptr, n_);  // read back while the grid is evaluated
    check(dwg_gwn_tree_grid(ctx_, tree_, &g, beta, opt_.sigma, flags, wg), "gwn final grid");
    const double sum = dwg::reduce_sum_f32_wait(ctx_);
    const float iso = float(sum / double(n_)) * iso_scale;
    int64_t ntri = 0;
    if (!tris) {
        const int rc = dwg_marching_cubes(ctx_, wg, &g, iso, 0, nullptr, nullptr, nullptr, &ntri);
        if (rc != DWG_OK && rc != DWG_ERR_CAPACITY) check(rc, "marching cubes");
    } else {
        // count and emit in one call when the kept triangle buffer is large
        // enough, else once more with a buffer of the counted size
        size_t cap = ctx_->buf_size("final_tris") / (9 * sizeof(float));
        float* t = static_cast<float*>(ctx_->buf("final_tris", sizeof(float) * 9 * std::max<size_t>(cap, 1)));
        int rc = dwg_marching_cubes(ctx_, wg, &g, iso, int64_t(cap), t, nullptr, nullptr, &ntri);
        if (rc == DWG_ERR_CAPACITY) {
            cap = size_t(ntri + ntri / 8 + 1024);
            t = static_cast<float*>(ctx_->buf("final_tris", sizeof(float) * 9 * cap));
            rc = dwg_marching_cubes(ctx_, wg, &g, iso, int64_t(cap), t, nullptr, nullptr, &ntri);
        }
        check(rc, "marching cubes");
        // straight into the buffer the caller receives (no second host copy)
        free_host(tris->data);
        tris->data = nullptr;
        tris->data = alloc_host_pinned(sizeof(float) * 9 * size_t(std::max<int64_t>(ntri, 1)));
        tris->ntri = ntri;
        DWG_HIP_CHECK(hipMemcpyAsync(tris->data, t, sizeof(float) * 9 * size_t(ntri), hipMemcpyDeviceToHost,
                                     ctx_->stream));
        DWG_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    }
    *tri_count = ntri;
}

void TurboDWG::GenerateMesh(const double* d_points, int64_t np, const Grid& final_grid, int64_t* tri_count,
                            HostMesh* tris, float* point_normals, dwg_octree* points_tree) {
    std::unique_ptr<dwg_octree, void (*)(dwg_octree*)> owned(points_tree, dwg_octree_destroy);
    // normals of the points from their normal_kw nearest sources (LoadNormals)
    DevArray<float> dp(3 * np), dn(3 * np);
    dwg::to_f32(ctx_, d_points, 3 * np, dp.ptr);
    check(dwg_transfer_normals(ctx_, index(), nrm_.ptr, dp.ptr, np, opt_.normal_kw, dn.ptr), "transfer normals");
    // --com_As 1: Voronoi areas of the points from their 15 nearest points and
    // the transferred normals (PreGeneration + point_areas + UpdateAs,
    // DWGProcess.h:771-792)
    DevArray<float> pa;
    if (opt_.com_As) {
        pa.resize(np);
        std::unique_ptr<dwg_knn_index> pidx(dwg::knn_build(ctx_, dp.ptr, np, 15));
        dwg::point_areas(ctx_, pidx.get(), dp.ptr, np, 15, dn.ptr, pa.ptr);
    }
    // the points become the sources (TurboDWG.cu:865-876), reusing the octree
    // over the input points (TurboDWG.cu:869-870), built with constant areas.
    // With --com_As 1, MeshGeneration refreshes CM, R and EC with the computed
    // areas (UpdateGWNInfo, TurboLibiGWN.cu:954-1017, called at
    // TurboDWG.cu:931): the octree is built again with area-weighted moments
    SetSources(d_points, np, dn.ptr, grid_, owned.release(), opt_.com_As ? pa.ptr : nullptr);
    UpdateRadii(opt_.final_kw);
    mesh_pass(final_grid, 1.0f, tri_count, tris, true);
    if (point_normals) dn.download(point_normals, 3 * np, ctx_->stream);
}

void TurboDWG::GenerateMeshFromSources(const Grid& final_grid, int64_t* tri_count, HostMesh* tris) {
    // GenerateMesh(grid) (TurboDWG.cu:768-806): the areas were filtered once
    // before the iterations and are not filtered again
    UpdateRadii(opt_.final_kw);
    mesh_pass(final_grid, 1.0f, tri_count, tris, false);
}

// ---------------------------------------------------------------------------
namespace test3D {

void NormalizePoints(dwg_ctx* ctx, const float* d_pts, int64_t n, double* d_out64, float* d_out32, double center[3],
                     double* scale) {
    dwg::normalize_points(ctx, d_pts, n, d_out64, d_out32, center, scale);
}

int64_t GetOctreeLeaves(dwg_ctx* ctx, const double* d_P, int64_t n, int max_depth, int min_num, dwg_octree** tree,
                        DevArray<double>& leaves, DevArray<double>& widths, DevArray<int32_t>& nodes) {
    dwg_octree* T = dwg::octree_build(ctx, d_P, nullptr, n, max_depth, 1);
    try {
        leaves.resize(3 * T->n_leaves);
        widths.resize(T->n_leaves);
        nodes.resize(T->n_leaves);
        const int64_t nl = dwg::octree_leaves(ctx, T, min_num, leaves.ptr, widths.ptr, nodes.ptr);
        *tree = T;
        return nl;
    } catch (...) {
        dwg_octree_destroy(T);
        throw;
    }
}

void SmoothNormalstoOctreeLeaves(dwg_ctx* ctx, const dwg_octree* tree, const DevArray<int32_t>& nodes,
                                 const DevArray<double>& leaves, const DevArray<double>& widths, int64_t nl,
                                 const double* d_P, const double* d_N, float* d_leaf_normals) {
    dwg::leaf_smooth_normals(ctx, tree, nodes.ptr, leaves.ptr, widths.ptr, nl, d_P, d_N, d_leaf_normals);
}

void GetRandomNs(int64_t n, uint64_t seed, std::vector<double>& N) {
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> dis(0.0, 2.0 * M_PI);
    N.resize(3 * n);
    for (int64_t i = 0; i < n; ++i) {
        const double u = dis(gen), v = dis(gen);
        N[3 * i] = std::sin(u) * std::cos(v);
        N[3 * i + 1] = std::sin(u) * std::sin(v);
        N[3 * i + 2] = std::cos(u);
    }
}

void DWGLaunch(dwg_ctx* ctx, const float* pts, const float* nrm, int64_t n, const dwg_options& opt,
               const char* out_base, dwg_result* res) {
    if (n <= 0 || !pts) throw dwg::Error(DWG_ERR_ARG, "no input points");
    // a run without iterations once ended in a host crash whose cause is not
    // known: refused instead of run
    if (opt.total_iter < 1) throw dwg::Error(DWG_ERR_ARG, "total_iter must be at least 1");
    if (opt.input_normal == 1 && !nrm) throw dwg::Error(DWG_ERR_ARG, "--in_normal 1 needs input normals");
    if (opt.input_normal == 4 && !opt.mlp) throw dwg::Error(DWG_ERR_ARG, "--in_normal 4 needs MLP weights");
    const auto t0 = Clock::now();
    hipStream_t stream = ctx->stream;
    // the input goes to the GPU once; everything up to the final mesh stays there
    DevArray<float> P32(3 * n);
    DevArray<double> P(3 * n);
    double center[3] = {0, 0, 0}, scale = 1;
    {
        DevArray<float> raw;
        raw.upload(pts, 3 * n, stream);
        if (opt.normalize_input) {
            NormalizePoints(ctx, raw.ptr, n, P.ptr, P32.ptr, center, &scale);
        } else {
            std::swap(P32, raw);
            dwg::to_f64(ctx, P32.ptr, 3 * n, P.ptr);
        }
    }
    // point normals (DWGProcess.h:152-210, 488-510).  Input normals for
    // --in_normal 1; PCA normals (k = 10, turned towards the bounding-box
    // centre) when --com_As 1 needs them for the area filter or for the PCA
    // start (5, an extension); random otherwise.  The Gauss-map and learned
    // starts overwrite the source normals later.
    DevArray<double> Ns;
    if (opt.input_normal == 1) {
        std::vector<double> h(nrm, nrm + 3 * n);
        Ns.upload(h.data(), 3 * n, stream);
    } else if (opt.com_As || opt.input_normal == 5) {
        float lo[3], hi[3];
        dwg::bounds(ctx, P32.ptr, n, lo, hi);
        const float vp[3] = {(lo[0] + hi[0]) / 2.0f, (lo[1] + hi[1]) / 2.0f, (lo[2] + hi[2]) / 2.0f};
        std::unique_ptr<dwg_knn_index> index(dwg::knn_build(ctx, P32.ptr, n, 10));
        DevArray<float> pn(3 * n);
        dwg::pca_normals(ctx, index.get(), P32.ptr, n, 10, vp, pn.ptr);
        Ns.resize(3 * n);
        dwg::to_f64(ctx, pn.ptr, 3 * n, Ns.ptr);
    } else if (opt.input_normal == 0 || opt.input_normal == 2) {
        std::vector<double> h;
        GetRandomNs(n, opt.seed, h);
        Ns.upload(h.data(), 3 * n, stream);
    }
    // sources
    std::unique_ptr<dwg_octree, void (*)(dwg_octree*)> points_tree(nullptr, dwg_octree_destroy);
    DevArray<double> leaves, leaf_w;
    DevArray<int32_t> leaf_node;
    DevArray<float> Sn;
    const double* S = P.ptr;
    int64_t ns = n;
    if (opt.source_mode == 0) {
        dwg_octree* T = nullptr;
        ns = GetOctreeLeaves(ctx, P.ptr, n, opt.max_depth, opt.min_num, &T, leaves, leaf_w, leaf_node);
        points_tree.reset(T);
        S = leaves.ptr;
        if (Ns.ptr) {
            Sn.resize(3 * ns);
            SmoothNormalstoOctreeLeaves(ctx, T, leaf_node, leaves, leaf_w, ns, P.ptr, Ns.ptr, Sn.ptr);
        }
    } else if (Ns.ptr) {
        Sn.resize(3 * n);
        dwg::to_f32(ctx, Ns.ptr, 3 * n, Sn.ptr);
    }
    if (ns <= 0) throw dwg::Error(DWG_ERR_ARG, "no octree leaf holds more than min_num points");
    // the iteration grid is built over the float input points (DWGProcess.h:662-672)
    double plo[3], phi[3];
    {
        float lo[3], hi[3];
        dwg::bounds(ctx, P32.ptr, n, lo, hi);
        for (int d = 0; d < 3; ++d) {
            plo[d] = lo[d];
            phi[d] = hi[d];
        }
    }
    const Grid grid = voxel_grid_bounds(plo, phi, opt.offset_1, opt.grid_reso);
    // --com_As 1: Voronoi areas of the sources from their 15 nearest sources
    // and normals (GetRadiiandAs, DWGProcess.h:345-380)
    DevArray<float> src_area;
    if (opt.com_As) {
        DevArray<float> S32(3 * ns);
        dwg::to_f32(ctx, S, 3 * ns, S32.ptr);
        src_area.resize(ns);
        std::unique_ptr<dwg_knn_index> sidx(dwg::knn_build(ctx, S32.ptr, ns, 15));
        dwg::point_areas(ctx, sidx.get(), S32.ptr, ns, 15, Sn.ptr, src_area.ptr);
        if (opt.input_normal == 0 || opt.input_normal == 2) {
            // the PCA normals only served the areas: the sources start from
            // random normals (DWGProcess.h:643-656)
            std::vector<double> h;
            GetRandomNs(ns, opt.seed + 1, h);
            std::vector<float> hf(h.begin(), h.end());
            Sn.upload(hf.data(), 3 * ns, stream);
        }
    }
    TurboDWG dwg(ctx, opt);
    if (out_base) {  // the dumps go to the output directory (main.cpp:155-156)
        const std::string ob(out_base);
        const size_t slash = ob.find_last_of('/');
        dwg.SetDumpDirectory(slash == std::string::npos ? std::string("./") : ob.substr(0, slash + 1));
    }
    dwg.SetSources(S, ns, Sn.ptr, grid, nullptr, opt.com_As ? src_area.ptr : nullptr);
    dwg.UpdateRadii(opt.kw);
    if (opt.input_normal == 3) dwg.GaussMapInit();
    if (opt.input_normal == 4) dwg.MlpInit(opt.mlp);
    if (opt.verbose)
        std::printf("sources %lld, grid %d x %d x %d\n", (long long)ns, grid.res[0], grid.res[1], grid.res[2]);
    if (dwg_ctx_sync(ctx) != DWG_OK) throw dwg::Error(DWG_ERR_HIP, dwg_ctx_error(ctx));
    const double t_setup = seconds_since(t0);
    if (res) res->grid_points = grid.size();
    if (opt.count_terms && res) {
        // outside every timed phase: the term set depends only on the tree and β
        std::vector<float> keep;
        dwg.download_normals(keep);
        dwg.CountTerms(&res->terms_src, &res->terms_grid);
        dwg.set_normals(keep);
    }
    const auto t1 = Clock::now();
    int iters_run = 0;
    const double t_iter = dwg.TurboScreenedDWGLaunch(opt.total_iter, res ? res->w_mean : nullptr,
                                                     res ? res->iter_tris : nullptr, res ? res->iter_zeros : nullptr,
                                                     res ? res->residual : nullptr, &iters_run);
    if (res) res->iterations = iters_run;
    if (res && res->source_normals) {  // the sources' normals after the iterations
        std::vector<float> sn;
        dwg.download_normals(sn);
        std::copy(sn.begin(), sn.end(), res->source_normals);
    }
    const auto t2 = Clock::now();
    int64_t ntri = 0;
    HostMesh tris;
    // the points' final normals on the host: straight into the caller's buffer
    // when it takes them (pn), else a local copy for the .xyz file
    std::vector<float> point_nrm;
    float* pn = nullptr;
    int64_t pn_count = 0;
    const bool want_xyz = out_base && (opt.output_mode == 2 || opt.output_mode == 3);
    const bool want_tris = (res && res->want_mesh) || (out_base && (opt.output_mode == 1 || opt.output_mode == 3));
    if (opt.use_leaves_generate == 1) {
        double slo[3], shi[3];
        dwg::bounds_f64(ctx, S, ns, slo, shi);
        const Grid fg = voxel_grid_bounds(slo, shi, 0.01, opt.final_reso);
        dwg.GenerateMeshFromSources(fg, &ntri, want_tris ? &tris : nullptr);
        dwg.download_normals(point_nrm);
    } else {
        const Grid fg = voxel_grid_bounds(plo, phi, opt.offset_2, opt.final_reso);
        if (opt.source_mode == 0) {
            if (res && res->normals) {
                pn = res->normals;
            } else if (want_xyz) {
                point_nrm.resize(3 * n);
                pn = point_nrm.data();
            }
            pn_count = pn ? n : 0;
            dwg.GenerateMesh(P.ptr, n, fg, &ntri, want_tris ? &tris : nullptr, pn, points_tree.release());
        } else {
            dwg.GenerateMeshFromSources(fg, &ntri, want_tris ? &tris : nullptr);
            dwg.download_normals(point_nrm);
        }
    }
    if (!pn) {
        pn = point_nrm.data();
        pn_count = int64_t(point_nrm.size() / 3);
    }
    const double t_final = seconds_since(t2);
    if (out_base) {
        if (opt.output_mode == 1 || opt.output_mode == 3) save_binary_stl(std::string(out_base) + ".stl", tris.data, ntri);
        if (want_xyz) {
            const int64_t np = pn_count;
            DevArray<float> f(3 * np);
            dwg::to_f32(ctx, (opt.use_leaves_generate == 1) ? S : P.ptr, 3 * np, f.ptr);
            std::vector<float> spf(3 * np);
            f.download(spf.data(), 3 * np, stream);
            save_xyz(std::string(out_base) + ".xyz", spf.data(), pn, np);
        }
    }
    if (res) {
        // the points' normals (with --use_leaves_generate 1 the final pass
        // keeps the sources' normals: returned when the sources are the points)
        if (res->normals && pn != res->normals && pn_count == n) std::copy(pn, pn + 3 * n, res->normals);
        res->tri_count = ntri;
        res->mesh = (res->want_mesh && ntri > 0) ? tris.release() : nullptr;
        res->n_sources = ns;
        for (int d = 0; d < 3; ++d) res->center[d] = center[d];
        res->scale = scale;
        res->time_setup_s = t_setup;
        res->time_iter_s = t_iter;
        res->time_final_s = t_final;
        res->time_total_s = seconds_since(t0) - std::chrono::duration<double>(t1 - t0).count() + t_setup;
    }
}

}  // namespace test3D
}  // namespace dwg_host

extern "C" {

void dwg_options_default(dwg_options* o) {
    if (!o) return;
    // src/main.cpp:19-39
    o->input_normal = 3;
    o->grid_reso = 400;
    o->final_reso = 500;
    o->kw = 10;
    o->final_kw = 10;
    o->normal_kw = 10;
    o->max_depth = 10;
    o->sigma = 10.0f;
    o->min_num = 0;
    o->use_radii = 0;
    o->com_As = 0;
    o->As_threshold = 0.0f;
    o->offset_1 = 0.05f;
    o->offset_2 = 0.01f;
    o->use_leaves_generate = 0;
    o->total_iter = 40;
    o->test_iter = 999999;
    o->output_mode = 1;
    o->source_mode = 0;
    o->leaf_size = 1;
    o->beta = 2.3f;
    o->precision = 0;
    o->seed = 1;
    o->normalize_input = 1;
    o->verbose = 0;
    o->count_terms = 0;
    o->mlp = nullptr;
    o->sigma_warm_iters = 4;
    o->residual_tol = 0.0f;
}

void dwg_free(void* p) { dwg_host::free_host(p); }

int dwg_launch(dwg_ctx* ctx, const float* h_pts, const float* h_nrm, int64_t n, const dwg_options* opt,
               const char* out_base, dwg_result* res) {
    if (!ctx || !opt) return DWG_ERR_ARG;
    try {
        if (hipSetDevice(ctx->device) != hipSuccess) throw dwg::Error(DWG_ERR_HIP, "hipSetDevice failed");
        dwg_host::test3D::DWGLaunch(ctx, h_pts, h_nrm, n, *opt, out_base, res);
        ctx->err.clear();
        return DWG_OK;
    } catch (const dwg::Error& e) {
        ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return DWG_ERR_HIP;
    }
}

}  // extern "C"

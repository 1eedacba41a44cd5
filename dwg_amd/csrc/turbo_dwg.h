// Host-side DWG driver: the reference's TurboDWG class and test3D entry
// points (include/TurboDWG.h:29-75, include/DWGProcess.h:32-808) over the
// libdwg_hip C ABI.  Eigen-free; point sets stay on the GPU from upload to
// the final mesh (row-major n×3 device arrays).
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/dwg_hip.h"

namespace dwg_host {

// Device array owned by the host driver.
template <class T>
struct DevArray {
    T* ptr = nullptr;
    int64_t count = 0;
    DevArray() = default;
    explicit DevArray(int64_t n) { resize(n); }
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    DevArray(DevArray&& o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
    DevArray& operator=(DevArray&& o) noexcept {
        std::swap(ptr, o.ptr);
        std::swap(count, o.count);
        return *this;
    }
    ~DevArray();
    void resize(int64_t n);
    void upload(const T* h, int64_t n, void* stream);
    void download(T* h, int64_t n, void* stream) const;
};

// Triangle soup on the host (ntri × 9 floats, std::malloc'd so that it can
// be handed to a dwg_result and released by dwg_free).
// Host buffers handed to the caller (released with dwg_free): pinned, so
// the device-to-host copy runs at full rate into pages that are already
// resident (a fresh pageable buffer costs ~25 ms of page faults per 100 MB).
float* alloc_host_pinned(size_t bytes);
void free_host(void* p);  // pinned buffers from alloc_host_pinned, else std::free

struct HostMesh {
    float* data = nullptr;
    int64_t ntri = 0;
    HostMesh() = default;
    HostMesh(const HostMesh&) = delete;
    HostMesh& operator=(const HostMesh&) = delete;
    ~HostMesh() { free_host(data); }
    float* release() {
        float* d = data;
        data = nullptr;
        return d;
    }
};

struct Grid {
    double origin[3];
    double h;
    int res[3];
    dwg_grid abi() const;
    int64_t size() const { return int64_t(res[0]) * res[1] * res[2]; }
};

// igl::voxel_grid(V, offset, s, pad = 1) in closed form (voxel_grid.cpp:15-87):
// from the points, or from their bounding box.
Grid voxel_grid(const double* V, int64_t n, double offset, int s);
Grid voxel_grid_bounds(const double lo[3], const double hi[3], double offset, int s);

// Binary STL (TurboMC.cu:343-371) and the .xyz normals writer (TurboDWG.cu:912-923).
void save_binary_stl(const std::string& path, const float* tris, int64_t ntri);
void save_xyz(const std::string& path, const float* pts, const float* nrm, int64_t n);
// .xyz reader (DWGProcess.h:53-97 without the normalisation): 6 columns per line.
bool read_xyz(const std::string& path, std::vector<float>& pts, std::vector<float>& nrm);

// The iteration driver (TurboDWG, src/TurboDWG.cu).  Sources stay resident
// on the GPU for the whole loop; every per-iteration buffer is reused.
class TurboDWG {
public:
    TurboDWG(dwg_ctx* ctx, const dwg_options& opt);
    ~TurboDWG();

    // TurboDWG::TurboDWG (TurboDWG.cu:14-83): sources (device FP64, n×3),
    // normals (device, n×3; nullptr → zeros), areas (device; nullptr → 2e-6,
    // DWGProcess.h:619), radii 0.0015, grid.  `prebuilt` (nullable, ownership
    // taken): an octree already built over exactly these points with constant
    // areas and leaf size 1.
    void SetSources(const double* d_pts, int64_t n, const float* d_nrm, const Grid& grid,
                    dwg_octree* prebuilt = nullptr, const float* d_area = nullptr);
    // per-source areas (device, n floats) for the far-field moments EC = Σ A·n
    // (TurboDWG::UpdateAs, TurboDWG.cu:693-701); CM and R keep the areas the
    // octree was built with (SetSources with d_area builds it area-weighted)
    void SetAreas(const float* d_area);
    // Radii of the sources from their k nearest neighbours (GetRadiiandAs /
    // TurboDWG::UpdateRadii); only with --use_radii, the one option under
    // which the sums read them.
    void UpdateRadii(int k);
    void GaussMapInit();
    // learned initialisation: MLP normals from the k-NN patches of the sources
    void MlpInit(const dwg_mlp* mlp);
    // TurboScreenedDWGLaunch (TurboDWG.cu:456-672); returns wall seconds.
    // residual (nullable): per-iteration convergence residual; with
    // residual_tol > 0 the loop stops after the first even iteration count
    // whose residual is below it; *iters_run receives the count run.
    double TurboScreenedDWGLaunch(int total_iters, double* w_mean, int64_t* iter_tris, int64_t* iter_zeros,
                                  double* residual = nullptr, int* iters_run = nullptr);
    // GenerateMesh (TurboDWG.cu:808-924): transfer the source normals to the
    // points, one screened evaluation at the points and on the final grid, MC.
    // d_points: device FP64 (np×3); point_normals (nullable): host buffer of
    // np×3 floats for the points' normals; points_tree (nullable, ownership
    // taken): their octree, reused as in TurboDWG.cu:869-870.
    void GenerateMesh(const double* d_points, int64_t np, const Grid& final_grid, int64_t* tri_count,
                      HostMesh* tris, float* point_normals, dwg_octree* points_tree = nullptr);
    // GenerateMesh(grid) variant of --use_leaves_generate 1 (TurboDWG.cu:768-806).
    void GenerateMeshFromSources(const Grid& final_grid, int64_t* tri_count, HostMesh* tris);

    // --test_iter dumps (TurboMC.cu:602-627): with a directory set, iteration
    // it with it % test_iter == 0 writes its marching-cubes mesh there as
    // <month><day><hour><minute><second>_<it>iter.stl
    void SetDumpDirectory(const std::string& dir) { dump_dir_ = dir; }

    // exact Barnes-Hut term counts of one source and one grid evaluation
    void CountTerms(int64_t* src_terms, int64_t* grid_terms);
    int64_t num_sources() const { return n_; }
    float* d_normals() { return nrm_.ptr; }
    const float* d_points() const { return pts_.ptr; }
    void download_normals(std::vector<float>& out) const;
    void set_normals(const std::vector<float>& n);

private:
    void check(int rc, const char* what) const;
    // filter_areas: apply the outlier filter to the point terms after the
    // moment refresh (GenerateMesh over the points)
    void mesh_pass(const Grid& g, float iso_scale, int64_t* tri_count, HostMesh* tris, bool filter_areas);
    void dump_mesh(const dwg_grid& g, float iso, int64_t ntri, int it);

    dwg_ctx* ctx_;
    dwg_options opt_;
    int64_t n_ = 0;
    DevArray<float> pts_, nrm_, area_, radii_, w_src_, w_grid_, cent_, tri_nrm_;
    Grid grid_{};
    dwg_octree* tree_ = nullptr;
    dwg_knn_index* knn_ = nullptr;  // the sources' kNN index, built on first use (index())
    std::string dump_dir_;
    dwg_knn_index* index();
};

// test3D entry points (DWGProcess.h) over in-memory points.
namespace test3D {

// ReadPointandNormals normalisation (DWGProcess.h:53-97) on the GPU: centre
// at the mean, divide by max_i 2|p_i − c|_∞.  d_pts: device floats (n×3);
// outputs device FP64 and float copies.
void NormalizePoints(dwg_ctx* ctx, const float* d_pts, int64_t n, double* d_out64, float* d_out32, double center[3],
                     double* scale);

// GetOctreeLeaves (DWGProcess.h:383-428) on the GPU: the octree over the
// points (returned in *tree, caller owns) and its breadth-first leaf cells
// with more than min_num points: centres (FP64), widths, node ids.
int64_t GetOctreeLeaves(dwg_ctx* ctx, const double* d_P, int64_t n, int max_depth, int min_num, dwg_octree** tree,
                        DevArray<double>& leaves, DevArray<double>& widths, DevArray<int32_t>& nodes);

// SmoothNormalstoOctreeLeaves (DWGProcess.h:430-472) + the whole-matrix
// Frobenius normalisation of DWGProcess.h:537, on the GPU.
void SmoothNormalstoOctreeLeaves(dwg_ctx* ctx, const dwg_octree* tree, const DevArray<int32_t>& nodes,
                                 const DevArray<double>& leaves, const DevArray<double>& widths, int64_t nl,
                                 const double* d_P, const double* d_N, float* d_leaf_normals);

// GetRandomNs (DWGProcess.h:264-281), seeded.
void GetRandomNs(int64_t n, uint64_t seed, std::vector<double>& N);

// DWGLaunch over in-memory points (src/main.cpp:5-15).
void DWGLaunch(dwg_ctx* ctx, const float* pts, const float* nrm, int64_t n, const dwg_options& opt,
               const char* out_base, dwg_result* res);

}  // namespace test3D
}  // namespace dwg_host

// Host-side DWG driver: the reference's TurboDWG class and test3D entry
// points (include/TurboDWG.h:29-75, include/DWGProcess.h:32-808) over the
// libdwg_hip C ABI.  Eigen-free: matrices are row-major std::vector<double>.
#pragma once

#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dwg_hip.h"

namespace dwg {
struct HostOctree;
}

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

struct Grid {
    double origin[3];
    double h;
    int res[3];
    dwg_grid abi() const;
    int64_t size() const { return int64_t(res[0]) * res[1] * res[2]; }
};

// igl::voxel_grid(V, offset, s, pad = 1) in closed form (voxel_grid.cpp:15-87).
Grid voxel_grid(const double* V, int64_t n, double offset, int s);

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

    // TurboDWG::TurboDWG (TurboDWG.cu:14-83): sources, normals, areas, radii, grid.
    // `prebuilt` (nullable): an octree already built over exactly these points
    // with constant areas and leaf size 1.
    void SetSources(const std::vector<double>& pts, const std::vector<float>& nrm, const std::vector<float>& areas,
                    const Grid& grid, const dwg::HostOctree* prebuilt = nullptr);
    // Radii of the sources from their k nearest neighbours (GetRadiiandAs /
    // TurboDWG::UpdateRadii).
    void UpdateRadii(int k);
    void GaussMapInit();
    // learned initialisation: MLP normals from the k-NN patches of the sources
    void MlpInit(const dwg_mlp* mlp);
    // TurboScreenedDWGLaunch (TurboDWG.cu:456-672); returns wall seconds.
    double TurboScreenedDWGLaunch(int total_iters, double* w_mean, int64_t* iter_tris, int64_t* iter_zeros);
    // GenerateMesh (TurboDWG.cu:808-924): transfer the source normals to the
    // points, one screened evaluation at the points and on the final grid, MC.
    void GenerateMesh(const std::vector<double>& points, const Grid& final_grid, int64_t* tri_count,
                      std::vector<float>* tris, std::vector<float>* point_normals,
                      const dwg::HostOctree* points_tree = nullptr);
    // GenerateMesh(grid) variant of --use_leaves_generate 1 (TurboDWG.cu:768-806).
    void GenerateMeshFromSources(const Grid& final_grid, int64_t* tri_count, std::vector<float>* tris);

    // exact Barnes-Hut term counts of one source and one grid evaluation
    void CountTerms(int64_t* src_terms, int64_t* grid_terms);
    int64_t num_sources() const { return n_; }
    float* d_normals() { return nrm_.ptr; }
    const float* d_points() const { return pts_.ptr; }
    void download_normals(std::vector<float>& out) const;
    void set_normals(const std::vector<float>& n);

private:
    void check(int rc, const char* what) const;
    void mesh_pass(const Grid& g, float iso_scale, int64_t* tri_count, std::vector<float>* tris);

    dwg_ctx* ctx_;
    dwg_options opt_;
    int64_t n_ = 0;
    DevArray<float> pts_, nrm_, area_, radii_, w_src_, w_grid_, cent_, tri_nrm_;
    Grid grid_{};
    dwg_octree* tree_ = nullptr;
    dwg_knn_index* knn_ = nullptr;
};

// test3D entry points (DWGProcess.h) over in-memory points.
namespace test3D {

// ReadPointandNormals normalisation (DWGProcess.h:53-97): centre at the mean,
// divide by max_i 2|p_i − c|_∞.
void NormalizePoints(const float* pts, int64_t n, std::vector<double>& out, double center[3], double* scale);

// GetOctreeLeaves (DWGProcess.h:383-428): breadth-first leaf cells with more
// than min_num points; returns cell centres, cell ids.
void GetOctreeLeaves(const std::vector<double>& P, int max_depth, int min_num, std::vector<double>& leaves,
                     std::vector<int>& leaf_node, std::vector<double>& leaf_width,
                     std::vector<std::vector<int>>& leaf_points, dwg::HostOctree* tree_out = nullptr);

// SmoothNormalstoOctreeLeaves (DWGProcess.h:430-472) + the whole-matrix
// Frobenius normalisation of DWGProcess.h:537.
void SmoothNormalstoOctreeLeaves(const std::vector<double>& leaves, const std::vector<double>& leaf_width,
                                 const std::vector<std::vector<int>>& leaf_points, const std::vector<double>& P,
                                 const std::vector<double>& N, std::vector<float>& leaf_normals);

// GetRandomNs (DWGProcess.h:264-281), seeded.
void GetRandomNs(int64_t n, uint64_t seed, std::vector<double>& N);

// DWGLaunch over in-memory points (src/main.cpp:5-15).
void DWGLaunch(dwg_ctx* ctx, const float* pts, const float* nrm, int64_t n, const dwg_options& opt,
               const char* out_base, dwg_result* res);

}  // namespace test3D
}  // namespace dwg_host

// Normal diffusion, radii and normal initialisation.
//
// Scatter: the reference adds every triangle normal into its 15 nearest
// sources with float atomics (TurboDWG.cu:179-204), so sums depend on
// arrival order.  Here each component is converted to 2^-40 fixed point and
// added as 64-bit integers (inside the centroid kNN, knn.hip): integer
// addition is associative, so the sum is exact up to the one rounding per
// term and bit-reproducible.
#include <algorithm>
#include <cmath>
#include <numeric>

#include <memory>

#include "diffuse.h"
#include "knn.h"
#include "scan.h"

namespace dwg {

constexpr double kInvFix = 1.0 / 1099511627776.0;  // 2^-40, the scatter's fixed point

// NormalizeNormals (TurboDWG.cu:238-266): |acc| < 1e-10 keeps the old normal
// and counts as zero; otherwise n = acc / (|acc| + 1e-8).
__global__ __launch_bounds__(256) void normalize_kernel(const unsigned long long* __restrict__ acc, int64_t n,
                                                        float* __restrict__ nrm, uint8_t* __restrict__ zero,
                                                        unsigned long long* __restrict__ nzero) {
    __shared__ int cnt[256];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int z = 0;
    if (i < n) {
        const double ax = double((long long)acc[3 * i]) * kInvFix;
        const double ay = double((long long)acc[3 * i + 1]) * kInvFix;
        const double az = double((long long)acc[3 * i + 2]) * kInvFix;
        const double l = sqrt(ax * ax + ay * ay + az * az);
        if (l < 1e-10) {
            z = 1;
        } else {
            const double s = 1.0 / (l + 1e-8);
            nrm[3 * i] = float(ax * s);
            nrm[3 * i + 1] = float(ay * s);
            nrm[3 * i + 2] = float(az * s);
        }
        zero[i] = uint8_t(z);
    }
    cnt[threadIdx.x] = z;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) cnt[threadIdx.x] += cnt[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0 && cnt[0]) atomicAdd(nzero, (unsigned long long)cnt[0]);
}

__global__ void nonzero_flags_kernel(const uint8_t* __restrict__ zero, int64_t n, int64_t* __restrict__ slot) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > n) return;
    slot[i] = (i < n && !zero[i]) ? 1 : 0;
}

// the zero sources, compacted: their indices and positions (slot[i] counts
// the non-zero sources before i, so i − slot[i] counts the zero ones)
__global__ void zero_gather_kernel(const int64_t* __restrict__ slot, const float* __restrict__ p, int64_t n,
                                   float* __restrict__ zq, int32_t* __restrict__ zlist) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || slot[i + 1] != slot[i]) return;
    const int64_t j = i - slot[i];
    for (int d = 0; d < 3; ++d) zq[3 * j + d] = p[3 * i + d];
    zlist[j] = int32_t(i);
}

// SimpleUpdateNonzeroNormals (TurboDWG.cu:322-359): a zero source takes the
// normalised sum of its nearest non-zero sources' normals.
// knn row t belongs to zero source zlist[t] and holds source indices.
__global__ void zero_fill_kernel(const int32_t* __restrict__ knn, int k, const int32_t* __restrict__ zlist,
                                 int64_t nzero, float* __restrict__ nrm) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nzero) return;
    const int64_t i = zlist[t];
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int j = 0; j < k; ++j) {
        const int32_t c = knn[t * k + j];
        if (c < 0) continue;
        const int64_t s = c;
        sx += nrm[3 * s];
        sy += nrm[3 * s + 1];
        sz += nrm[3 * s + 2];
    }
    const float sc = 1.0f / (sqrtf(sx * sx + sy * sy + sz * sz) + 1e-8f);
    nrm[3 * i] = sx * sc;
    nrm[3 * i + 1] = sy * sc;
    nrm[3 * i + 2] = sz * sc;
}

void diffuse_normals(dwg_ctx* ctx, const dwg_knn_index* I, const float* src, int64_t n, const float* cent,
                     const float* tri_nrm, int64_t ntri, int k, int zero_k, int zero_threshold, float* nrm,
                     int64_t* h_nzero) {
    DWG_REQUIRE(I && I->n == n, "kNN index does not match the source count");
    DWG_REQUIRE(k >= 1 && k <= 32 && zero_k >= 1 && zero_k <= 32, "k out of range");
    unsigned long long* acc = static_cast<unsigned long long*>(ctx->buf("diff_acc", sizeof(uint64_t) * (3 * n + 1)));
    uint8_t* zero = static_cast<uint8_t*>(ctx->buf("diff_zero", size_t(n) + 4));
    unsigned long long* nz = acc + 3 * n;
    fill_u32(ctx, acc, 0, 2 * (3 * n + 1));
    if (ntri > 0) {
        // the kNN of the centroids also scatters the face normals
        int32_t* knn = static_cast<int32_t*>(ctx->buf("diff_knn", sizeof(int32_t) * ntri * k));
        const KnnScatter sc{tri_nrm, acc};
        knn_query(ctx, I, cent, ntri, k, nullptr, nullptr, knn, nullptr, "knn_query_centroids", &sc);
    }
    {
        ProfileScope ps(ctx, "diffuse_normalize");
        hipLaunchKernelGGL(normalize_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, acc, n, nrm, zero, nz);
        check_launch();
    }
    // the zero sources' positions and indices, compacted (room for n): queued
    // before their count is read back, as the zero fill nearly always runs
    int64_t* slot = static_cast<int64_t*>(ctx->buf("diff_nz_slot", sizeof(int64_t) * (n + 1)));
    float* zq = static_cast<float*>(ctx->buf("diff_z_pts", sizeof(float) * 3 * n));
    int32_t* zlist = static_cast<int32_t*>(ctx->buf("diff_z_list", sizeof(int32_t) * n));
    {
        ProfileScope ps(ctx, "diffuse_nonzero_index");
        hipLaunchKernelGGL(nonzero_flags_kernel, dim3(grid_for(n + 1, 256)), dim3(256), 0, ctx->stream, zero, n, slot);
        check_launch();
        exclusive_scan<int64_t>(ctx, slot, n);
        hipLaunchKernelGGL(zero_gather_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, slot, src, n, zq,
                           zlist);
        check_launch();
    }
    const int64_t nzero = int64_t(read_scalar<unsigned long long>(ctx, nz));
    if (h_nzero) *h_nzero = nzero;
    if (nzero > zero_threshold && nzero < n) {
        // the zero sources query the sources' own index with the zero ones
        // skipped (knn_query then walks a pyramid of the non-zero sources
        // only, as the zero sources come in clusters); results are source
        // indices, ties broken by source index as over the non-zero set
        int32_t* zknn = static_cast<int32_t*>(ctx->buf("diff_zknn", sizeof(int32_t) * nzero * zero_k));
        knn_query(ctx, I, zq, nzero, zero_k, zero, nullptr, zknn, nullptr, "knn_query_zero_fill");
        ProfileScope ps(ctx, "diffuse_zero_fill");
        hipLaunchKernelGGL(zero_fill_kernel, dim3(grid_for(nzero, 256)), dim3(256), 0, ctx->stream, zknn, zero_k, zlist,
                           nzero, nrm);
        check_launch();
    }
}

// r = clamp(sqrt(Σ_{j∈kNN} |p_i − p_j|² / k), rmin, rmax), FP64 accumulation
// (DWGProcess.h:325-343).
__global__ void radii_kernel(const int32_t* __restrict__ knn, int k, const float* __restrict__ p, int64_t n, float rmin,
                             float rmax, float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    for (int j = 0; j < k; ++j) {
        const int32_t q = knn[i * k + j];
        if (q < 0) continue;
        for (int d = 0; d < 3; ++d) {
            const double df = double(p[3 * i + d]) - double(p[3 * int64_t(q) + d]);
            s += df * df;
        }
    }
    double r = sqrt(s / k);
    r = r < rmax ? r : rmax;
    r = r > rmin ? r : rmin;
    out[i] = float(r);
}

void radii(dwg_ctx* ctx, const dwg_knn_index* I, const float* p, int64_t n, int k, float rmin, float rmax, float* out) {
    DWG_REQUIRE(I && I->n == n, "kNN index does not match the point count");
    int32_t* knn = static_cast<int32_t*>(ctx->buf("radii_knn", sizeof(int32_t) * n * k));
    knn_query(ctx, I, p, n, k, nullptr, nullptr, knn, nullptr, "knn_query_radii");
    ProfileScope ps(ctx, "radii");
    hipLaunchKernelGGL(radii_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, knn, k, p, n, rmin, rmax, out);
    check_launch();
}

// Smallest-eigenvalue eigenvector of the symmetric 3×3 matrix
// (a00 a01 a02; a11 a12; a22): eigenvalue by the trigonometric closed form
// (Smith, CACM 4(4), 1961), vector as the longest cross product of two rows
// of A − λI.  An isotropic matrix returns (0, 0, 1).
__device__ void smallest_eigenvector(double a00, double a01, double a02, double a11, double a12, double a22,
                                     double v[3]) {
    const double m = (a00 + a11 + a22) / 3.0;
    const double b00 = a00 - m, b11 = a11 - m, b22 = a22 - m;
    const double p2 = (b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * (a01 * a01 + a02 * a02 + a12 * a12)) / 6.0;
    v[0] = 0.0;
    v[1] = 0.0;
    v[2] = 1.0;
    if (!(p2 > 0.0)) return;
    const double p = sqrt(p2);
    const double det = b00 * (b11 * b22 - a12 * a12) - a01 * (a01 * b22 - a12 * a02) + a02 * (a01 * a12 - b11 * a02);
    double c = det / (2.0 * p2 * p);
    c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
    const double phi = acos(c) / 3.0;
    const double lam = m + 2.0 * p * cos(phi + 2.0943951023931954923);  // + 2π/3: the smallest root
    const double r0[3] = {a00 - lam, a01, a02}, r1[3] = {a01, a11 - lam, a12}, r2[3] = {a02, a12, a22 - lam};
    const double c01[3] = {r0[1] * r1[2] - r0[2] * r1[1], r0[2] * r1[0] - r0[0] * r1[2], r0[0] * r1[1] - r0[1] * r1[0]};
    const double c02[3] = {r0[1] * r2[2] - r0[2] * r2[1], r0[2] * r2[0] - r0[0] * r2[2], r0[0] * r2[1] - r0[1] * r2[0]};
    const double c12[3] = {r1[1] * r2[2] - r1[2] * r2[1], r1[2] * r2[0] - r1[0] * r2[2], r1[0] * r2[1] - r1[1] * r2[0]};
    const double n01 = c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2];
    const double n02 = c02[0] * c02[0] + c02[1] * c02[1] + c02[2] * c02[2];
    const double n12 = c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2];
    const double* w = c01;
    double nw = n01;
    if (n02 > nw) {
        w = c02;
        nw = n02;
    }
    if (n12 > nw) {
        w = c12;
        nw = n12;
    }
    if (!(nw > 0.0)) return;
    const double inv = 1.0 / sqrt(nw);
    v[0] = w[0] * inv;
    v[1] = w[1] * inv;
    v[2] = w[2] * inv;
}

// PCA normal of each point from its k nearest neighbours (the point itself
// included), turned to face the viewpoint: PCL NormalEstimation with
// setKSearch(k) and setViewPoint(bbox centre), DWGProcess.h:152-200.
__global__ void pca_normals_kernel(const int32_t* __restrict__ knn, int k, const float* __restrict__ p, int64_t n,
                                   float vx, float vy, float vz, float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double mx = 0, my = 0, mz = 0;
    int m = 0;
    for (int j = 0; j < k; ++j) {
        const int32_t q = knn[i * k + j];
        if (q < 0) continue;
        mx += p[3 * int64_t(q)];
        my += p[3 * int64_t(q) + 1];
        mz += p[3 * int64_t(q) + 2];
        ++m;
    }
    double v[3] = {0.0, 0.0, 1.0};
    if (m >= 3) {
        mx /= m;
        my /= m;
        mz /= m;
        double c00 = 0, c01 = 0, c02 = 0, c11 = 0, c12 = 0, c22 = 0;
        for (int j = 0; j < k; ++j) {
            const int32_t q = knn[i * k + j];
            if (q < 0) continue;
            const double dx = p[3 * int64_t(q)] - mx, dy = p[3 * int64_t(q) + 1] - my, dz = p[3 * int64_t(q) + 2] - mz;
            c00 += dx * dx;
            c01 += dx * dy;
            c02 += dx * dz;
            c11 += dy * dy;
            c12 += dy * dz;
            c22 += dz * dz;
        }
        smallest_eigenvector(c00, c01, c02, c11, c12, c22, v);
    }
    const double tx = vx - p[3 * i], ty = vy - p[3 * i + 1], tz = vz - p[3 * i + 2];
    if (v[0] * tx + v[1] * ty + v[2] * tz < 0.0) {
        v[0] = -v[0];
        v[1] = -v[1];
        v[2] = -v[2];
    }
    out[3 * i] = float(v[0]);
    out[3 * i + 1] = float(v[1]);
    out[3 * i + 2] = float(v[2]);
}

void pca_normals(dwg_ctx* ctx, const dwg_knn_index* I, const float* p, int64_t n, int k, const float* viewpoint,
                 float* out) {
    DWG_REQUIRE(I && I->n == n, "kNN index does not match the point count");
    DWG_REQUIRE(k >= 3 && k <= 32, "PCA needs 3 ≤ k ≤ 32");
    int32_t* knn = static_cast<int32_t*>(ctx->buf("pca_knn", sizeof(int32_t) * n * k));
    knn_query(ctx, I, p, n, k, nullptr, nullptr, knn, nullptr, "knn_query_pca");
    ProfileScope ps(ctx, "pca_normals");
    hipLaunchKernelGGL(pca_normals_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, knn, k, p, n,
                       viewpoint[0], viewpoint[1], viewpoint[2], out);
    check_launch();
}

// ---------------------------------------------------------------------------
// Voronoi point areas without CGAL: igl::copyleft::cgal::point_areas
// (deps/libigl/include/igl/copyleft/cgal/point_areas.cpp:41-205).  Per
// point: its neighbours (the first is the point itself), kept where
// n_first·n_j > 0; ≤ 2 kept → 0; the PCA plane of the kept points; the
// Delaunay triangles at the first point, found directly: (0, j, l) is one
// when no other kept point lies inside its circumcircle.  The in-circle
// determinant is evaluated once per set of four points in index order and
// exact ties are broken by Edelsbrunner–Mücke simulation of simplicity (the
// lifted coordinate of point i perturbed by ε^i), so cocircular neighbours
// give one consistent triangulation.  Each triangle adds the first point's
// Voronoi share with the reference's rules.
// ---------------------------------------------------------------------------
constexpr int kAreaMax = 32;

__device__ __forceinline__ double orient2(const double* X, const double* Y, int a, int b, int c) {
    return (X[b] - X[a]) * (Y[c] - Y[a]) - (Y[b] - Y[a]) * (X[c] - X[a]);
}

// sign of the lifted determinant D(a, b, c, d) = det[x y x²+y² 1] (positive
// when d is inside the circle through a, b, c taken counter-clockwise)
__device__ int incircle_sign(const double* X, const double* Y, int a, int b, int c, int d) {
    int v[4] = {a, b, c, d};
    int parity = 1;
    for (int i = 0; i < 3; ++i)  // sort the four indices, tracking the permutation parity
        for (int j = 0; j < 3 - i; ++j)
            if (v[j] > v[j + 1]) {
                const int t = v[j];
                v[j] = v[j + 1];
                v[j + 1] = t;
                parity = -parity;
            }
    const int s0 = v[0];
    double m[3][3];
    for (int r = 0; r < 3; ++r) {
        const double x = X[v[r + 1]] - X[s0], y = Y[v[r + 1]] - Y[s0];
        m[r][0] = x;
        m[r][1] = y;
        m[r][2] = x * x + y * y;
    }
    const double det3 = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                        m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                        m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    double D = -det3;
    if (D == 0.0) {  // the largest perturbation belongs to the lowest index
        D = orient2(X, Y, v[1], v[2], v[3]);
        if (D == 0.0) D = -orient2(X, Y, v[0], v[2], v[3]);
    }
    const int sd = D > 0.0 ? 1 : (D < 0.0 ? -1 : 0);
    return parity * sd;
}

__global__ void point_areas_kernel(const int32_t* __restrict__ knn, int k, const float* __restrict__ p,
                                   const float* __restrict__ nrm, int64_t n, float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double px[kAreaMax], py[kAreaMax], pz[kAreaMax];
    int m = 0;
    const int32_t first = knn[i * k];
    for (int j = 0; j < k && j < kAreaMax; ++j) {
        const int32_t c = knn[i * k + j];
        if (c < 0) continue;
        if (nrm && first >= 0) {
            const double dot = double(nrm[3 * int64_t(first)]) * nrm[3 * int64_t(c)] +
                               double(nrm[3 * int64_t(first) + 1]) * nrm[3 * int64_t(c) + 1] +
                               double(nrm[3 * int64_t(first) + 2]) * nrm[3 * int64_t(c) + 2];
            if (!(dot > 0.0)) continue;
        }
        px[m] = p[3 * int64_t(c)];
        py[m] = p[3 * int64_t(c) + 1];
        pz[m] = p[3 * int64_t(c) + 2];
        ++m;
    }
    if (m <= 2) {
        out[i] = 0.f;
        return;
    }
    double mx = 0, my = 0, mz = 0;
    for (int j = 0; j < m; ++j) {
        mx += px[j];
        my += py[j];
        mz += pz[j];
    }
    mx /= m;
    my /= m;
    mz /= m;
    double c00 = 0, c01 = 0, c02 = 0, c11 = 0, c12 = 0, c22 = 0;
    for (int j = 0; j < m; ++j) {
        const double dx = px[j] - mx, dy = py[j] - my, dz = pz[j] - mz;
        c00 += dx * dx;
        c01 += dx * dy;
        c02 += dx * dz;
        c11 += dy * dy;
        c12 += dy * dz;
        c22 += dz * dz;
    }
    double nv[3];
    smallest_eigenvector(c00, c01, c02, c11, c12, c22, nv);
    // an orthonormal basis (u, v) of the plane: distances in it equal those
    // of the SVD scores' first two columns
    const double ax = fabs(nv[0]), ay = fabs(nv[1]), az = fabs(nv[2]);
    double e[3] = {0, 0, 0};
    e[(ax <= ay && ax <= az) ? 0 : (ay <= az ? 1 : 2)] = 1.0;
    double u[3] = {nv[1] * e[2] - nv[2] * e[1], nv[2] * e[0] - nv[0] * e[2], nv[0] * e[1] - nv[1] * e[0]};
    const double un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int d = 0; d < 3; ++d) u[d] /= un;
    const double v[3] = {nv[1] * u[2] - nv[2] * u[1], nv[2] * u[0] - nv[0] * u[2], nv[0] * u[1] - nv[1] * u[0]};
    double X[kAreaMax], Y[kAreaMax];
    for (int j = 0; j < m; ++j) {
        const double dx = px[j] - mx, dy = py[j] - my, dz = pz[j] - mz;
        X[j] = dx * u[0] + dy * u[1] + dz * u[2];
        Y[j] = dx * v[0] + dy * v[1] + dz * v[2];
    }
    double acc = 0.0;
    for (int j = 1; j < m; ++j)
        for (int l = j + 1; l < m; ++l) {
            const double o = orient2(X, Y, 0, j, l);
            if (o == 0.0) continue;
            bool empty = true;
            for (int t = 1; t < m && empty; ++t) {
                if (t == j || t == l) continue;
                const int s = incircle_sign(X, Y, 0, j, l, t);
                if ((o > 0.0 ? s : -s) > 0) empty = false;
            }
            if (!empty) continue;
            // triangle X=0, Y=j, Z=l (point_areas.cpp:160-195)
            const double x = sqrt((X[j] - X[l]) * (X[j] - X[l]) + (Y[j] - Y[l]) * (Y[j] - Y[l]));
            const double y = sqrt((X[0] - X[l]) * (X[0] - X[l]) + (Y[0] - Y[l]) * (Y[0] - Y[l]));
            const double z = sqrt((X[j] - X[0]) * (X[j] - X[0]) + (Y[j] - Y[0]) * (Y[j] - Y[0]));
            const double cx = (z * z + y * y - x * x) / (2 * y * z);
            const double cy = (z * z + x * x - y * y) / (2 * x * z);
            const double cz = (x * x + y * y - z * z) / (2 * y * x);
            double b0 = x * cx, b1 = y * cy, b2 = z * cz;
            const double bs = b0 + b1 + b2;
            b1 /= bs;
            b2 /= bs;
            const double full = 0.25 * sqrt((x + (y + z)) * (z - (x - y)) * (z + (x - y)) * (x + (y - z)));
            if (cx < 0) acc += 0.5 * full;
            else if (cy < 0 || cz < 0) acc += 0.25 * full;
            else acc += (b1 * full + b2 * full) / 2;
            (void)b0;
        }
    out[i] = isfinite(acc) ? float(acc) : 0.f;
}

void point_areas(dwg_ctx* ctx, const dwg_knn_index* I, const float* p, int64_t n, int k, const float* nrm, float* out) {
    DWG_REQUIRE(I && I->n == n, "kNN index does not match the point count");
    DWG_REQUIRE(k >= 3 && k <= kAreaMax, "point areas need 3 ≤ k ≤ 32");
    int32_t* knn = static_cast<int32_t*>(ctx->buf("areas_knn", sizeof(int32_t) * n * k));
    knn_query(ctx, I, p, n, k, nullptr, nullptr, knn, nullptr, "knn_query_areas");
    ProfileScope ps(ctx, "point_areas");
    hipLaunchKernelGGL(point_areas_kernel, dim3(grid_for(n, 128)), dim3(128), 0, ctx->stream, knn, k, p, nrm, n, out);
    check_launch();
}

__global__ void gauss_map_kernel(const float* __restrict__ p, int64_t n, float cx, float cy, float cz,
                                 float* __restrict__ nrm) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = p[3 * i] - cx, y = p[3 * i + 1] - cy, z = p[3 * i + 2] - cz;
    const float s = 1.0f / (sqrtf(x * x + y * y + z * z) + 1e-8f);
    nrm[3 * i] = x * s;
    nrm[3 * i + 1] = y * s;
    nrm[3 * i + 2] = z * s;
}

void gauss_map(dwg_ctx* ctx, const float* p, int64_t n, const float* center, float* nrm) {
    if (n == 0) return;
    ProfileScope ps(ctx, "gauss_map");
    hipLaunchKernelGGL(gauss_map_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, p, n, center[0], center[1],
                       center[2], nrm);
    check_launch();
}

// LoadNormals (TurboDWG.cu:206-236): mean of k neighbour normals, then
// n / (|n| + 1e-8).
__global__ void transfer_kernel(const int32_t* __restrict__ knn, int k, const float* __restrict__ src_nrm, int64_t nq,
                                float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int j = 0; j < k; ++j) {
        const int32_t s = knn[i * k + j];
        if (s < 0) continue;
        sx += src_nrm[3 * s];
        sy += src_nrm[3 * s + 1];
        sz += src_nrm[3 * s + 2];
    }
    sx /= float(k);
    sy /= float(k);
    sz /= float(k);
    const float sc = 1.0f / (sqrtf(sx * sx + sy * sy + sz * sz) + 1e-8f);
    out[3 * i] = sx * sc;
    out[3 * i + 1] = sy * sc;
    out[3 * i + 2] = sz * sc;
}

void transfer_normals(dwg_ctx* ctx, const dwg_knn_index* I, const float* src_nrm, const float* q, int64_t nq, int k,
                      float* out) {
    if (nq == 0) return;
    int32_t* knn = static_cast<int32_t*>(ctx->buf("transfer_knn", sizeof(int32_t) * nq * k));
    knn_query(ctx, I, q, nq, k, nullptr, nullptr, knn, nullptr, "knn_query_transfer");
    ProfileScope ps(ctx, "transfer_normals");
    hipLaunchKernelGGL(transfer_kernel, dim3(grid_for(nq, 256)), dim3(256), 0, ctx->stream, knn, k, src_nrm, nq, out);
    check_launch();
}

// Convergence residual of one diffusion step, reduced in FP64: the mean over
// the sources of 1 − |cos ∠(n_old, n_new)|.  The absolute value makes it
// blind to the global sign flip every iteration makes (the face normals
// point towards increasing winding numbers); a source whose old or new
// normal is zero counts 1.
__global__ __launch_bounds__(256) void normal_change_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                            int64_t n, double* __restrict__ partial) {
    __shared__ double lds[256];
    double s = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const double ax = a[3 * i], ay = a[3 * i + 1], az = a[3 * i + 2];
        const double bx = b[3 * i], by = b[3 * i + 1], bz = b[3 * i + 2];
        const double la = ax * ax + ay * ay + az * az, lb = bx * bx + by * by + bz * bz;
        s += (la > 0 && lb > 0) ? 1.0 - fabs(ax * bx + ay * by + az * bz) / sqrt(la * lb) : 1.0;
    }
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

__global__ void sum_partials_kernel(const double* __restrict__ partial, int m, double* __restrict__ out) {
    __shared__ double lds[256];
    double s = 0;
    for (int i = threadIdx.x; i < m; i += 256) s += partial[i];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = lds[0];
}

double normal_change(dwg_ctx* ctx, const float* old_nrm, const float* new_nrm, int64_t n) {
    if (n <= 0) return 0.0;
    constexpr int blocks = 512;
    auto* partial = static_cast<double*>(ctx->buf("normal_change", sizeof(double) * (blocks + 1)));
    {
        ProfileScope ps(ctx, "normal_change");
        hipLaunchKernelGGL(normal_change_kernel, dim3(blocks), dim3(256), 0, ctx->stream, old_nrm, new_nrm, n, partial);
        check_launch();
        hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, ctx->stream, partial, blocks, partial + blocks);
        check_launch();
    }
    return read_scalar<double>(ctx, partial + blocks) / double(n);
}

// OutlierFilter (TurboDWG.cu:991-1009): the ceil(n·thr) largest areas are
// divided by 1000.  Ties are ordered by index (stable), on the host.
// OutlierFilter (TurboDWG.cu:991-1009): the reference sorts the areas in
// descending order (a stable radix sort, so equal areas keep ascending index
// order) and divides the first ceil(n·threshold) by 1000.  Here a radix
// select finds the m-th largest key in four 8-bit passes; every area above
// it, and the first of those equal to it in index order, are divided.
__device__ __forceinline__ uint32_t float_key(float f) {  // order-preserving
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(256) void select_hist_kernel(const float* __restrict__ v, int64_t n, uint32_t prefix,
                                                          uint32_t mask, int shift, unsigned int* __restrict__ hist) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const uint32_t k = float_key(v[i]);
        if ((k & mask) == prefix) atomicAdd(&h[(k >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void equal_flags_kernel(const float* __restrict__ v, int64_t n, uint32_t key, int64_t* __restrict__ eq) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) eq[i] = float_key(v[i]) == key ? 1 : 0;
}

__global__ void outlier_divide_kernel(float* __restrict__ v, int64_t n, uint32_t key, const int64_t* __restrict__ rank,
                                      int64_t take_equal) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = float_key(v[i]);
    if (k > key || (k == key && rank[i] < take_equal)) v[i] /= 1000.0f;
}

void outlier_filter(dwg_ctx* ctx, float* area, int64_t n, float threshold) {
    const int64_t m = int64_t(std::ceil(float(n) * threshold));  // the reference's float product
    if (m <= 0 || n == 0) return;
    const int64_t take = std::min(m, n);
    auto* hist = static_cast<unsigned int*>(ctx->buf("outlier_hist", 256 * sizeof(unsigned int)));
    const unsigned blocks = unsigned(std::min<int64_t>(grid_for(n, 256), 1024));
    uint32_t prefix = 0, mask = 0;
    int64_t remaining = take;  // rank of the wanted key among those matching the prefix
    for (int shift = 24; shift >= 0; shift -= 8) {
        fill_u32(ctx, hist, 0u, 256);
        hipLaunchKernelGGL(select_hist_kernel, dim3(blocks), dim3(256), 0, ctx->stream, area, n, prefix, mask, shift,
                           hist);
        check_launch();
        unsigned int h[256];
        DWG_HIP_CHECK(hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        int d = 255;
        for (; d > 0 && int64_t(h[d]) < remaining; --d) remaining -= h[d];
        prefix |= uint32_t(d) << shift;
        mask |= 0xFFu << shift;
    }
    // `remaining` of the areas equal to the selected key are taken, lowest index first
    auto* rank = static_cast<int64_t*>(ctx->buf("outlier_rank", sizeof(int64_t) * size_t(n + 1)));
    hipLaunchKernelGGL(equal_flags_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, area, n, prefix, rank);
    check_launch();
    exclusive_scan_i64(ctx, rank, n);
    hipLaunchKernelGGL(outlier_divide_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, area, n, prefix, rank,
                       remaining);
    check_launch();
}

}  // namespace dwg

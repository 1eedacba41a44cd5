// Normal diffusion, radii and normal initialisation.
//
// Scatter: the reference adds every triangle normal into its 15 nearest
// sources with float atomics (TurboDWG.cu:179-204), so sums depend on
// arrival order.  Here each component is converted to 2^-40 fixed point and
// added with 64-bit integer atomics: integer addition is associative, so
// the sum is exact up to the one rounding per term and bit-reproducible.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "diffuse.h"
#include "knn.h"
#include "scan.h"

namespace dwg {

constexpr double kFix = 1099511627776.0;  // 2^40
constexpr double kInvFix = 1.0 / 1099511627776.0;

__global__ void scatter_fixed_kernel(const int32_t* __restrict__ knn, int k, const float* __restrict__ tri_nrm,
                                     int64_t ntri, unsigned long long* __restrict__ acc) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= ntri * k) return;
    const int64_t tri = t / k;
    const int32_t s = knn[t];
    if (s < 0) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long long v = __double2ll_rn(double(tri_nrm[3 * tri + d]) * kFix);
        atomicAdd(&acc[3 * int64_t(s) + d], (unsigned long long)v);
    }
}

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

// SimpleUpdateNonzeroNormals (TurboDWG.cu:322-359): a zero source takes the
// normalised sum of its nearest non-zero sources' normals.
__global__ void zero_fill_kernel(const int32_t* __restrict__ knn, int k, const uint8_t* __restrict__ zero, int64_t n,
                                 float* __restrict__ nrm) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !zero[i]) return;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int j = 0; j < k; ++j) {
        const int32_t s = knn[i * k + j];
        if (s < 0) continue;
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
        int32_t* knn = static_cast<int32_t*>(ctx->buf("diff_knn", sizeof(int32_t) * ntri * k));
        knn_query(ctx, I, cent, ntri, k, nullptr, nullptr, knn, nullptr, "knn_query_centroids");
        ProfileScope ps(ctx, "diffuse_scatter");
        hipLaunchKernelGGL(scatter_fixed_kernel, dim3(grid_for(ntri * k, 256)), dim3(256), 0, ctx->stream, knn, k, tri_nrm,
                           ntri, acc);
        check_launch();
    }
    {
        ProfileScope ps(ctx, "diffuse_normalize");
        hipLaunchKernelGGL(normalize_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, acc, n, nrm, zero, nz);
        check_launch();
    }
    const int64_t nzero = int64_t(read_scalar<unsigned long long>(ctx, nz));
    if (h_nzero) *h_nzero = nzero;
    if (nzero > zero_threshold && nzero < n) {
        int32_t* zknn = static_cast<int32_t*>(ctx->buf("diff_zknn", sizeof(int32_t) * n * zero_k));
        knn_query(ctx, I, src, n, zero_k, zero, zero, zknn, nullptr, "knn_query_zero_fill");
        ProfileScope ps(ctx, "diffuse_zero_fill");
        hipLaunchKernelGGL(zero_fill_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, zknn, zero_k, zero, n, nrm);
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

// OutlierFilter (TurboDWG.cu:991-1009): the ceil(n·thr) largest areas are
// divided by 1000.  Ties are ordered by index (stable), on the host.
void outlier_filter(dwg_ctx* ctx, float* area, int64_t n, float threshold) {
    const int64_t m = int64_t(std::ceil(float(n) * threshold));
    if (m <= 0 || n == 0) return;
    std::vector<float> h(n);
    DWG_HIP_CHECK(hipMemcpyAsync(h.data(), area, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return h[a] > h[b]; });
    for (int64_t j = 0; j < std::min(m, n); ++j) h[idx[j]] /= 1000.0f;
    DWG_HIP_CHECK(hipMemcpyAsync(area, h.data(), sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace dwg

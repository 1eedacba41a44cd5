// Exact k-nearest neighbours on a uniform cell grid.
//
// The reference builds a left-balanced kd-tree with log2(N) thrust sorts and
// walks it per query (KNN_GWN.cu:42-66, builder_thrust.h:486-724).  Here the
// points are bucketed once into a dense cell grid (counting sort: one
// histogram, one scan, one scatter); a query visits cells in growing
// Chebyshev shells around its own cell and stops once its k-th key is closer
// than any cell not yet visited.  Candidates are kept as 64-bit keys
// (float bits of d² << 32 | index) in a register array, so results are the
// exact k smallest (d², index) pairs — identical to a brute-force scan with
// the same float d² formula, ties to the lower index.
#include <algorithm>
#include <cmath>

#include "knn.h"
#include "scan.h"

dwg_knn_index::~dwg_knn_index() {
    for (auto& L : lvl) {
        if (L.cell_start) (void)hipFree(L.cell_start);
        if (L.pts) (void)hipFree(L.pts);
    }
}

namespace dwg {

__device__ __forceinline__ int cell_coord(float v, float lo, float inv_cell, int dim) {
    int c = int(floorf((v - lo) * inv_cell));
    return min(max(c, 0), dim - 1);
}

__global__ void knn_count_kernel(const float* __restrict__ p, int64_t n, KnnGridArgs G,
                                 unsigned long long* __restrict__ counts, int* __restrict__ cell_of) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cx = cell_coord(p[3 * i], G.lo[0], G.inv_cell, G.dim[0]);
    const int cy = cell_coord(p[3 * i + 1], G.lo[1], G.inv_cell, G.dim[1]);
    const int cz = cell_coord(p[3 * i + 2], G.lo[2], G.inv_cell, G.dim[2]);
    const int64_t c = cx + int64_t(G.dim[0]) * (cy + int64_t(G.dim[1]) * cz);
    cell_of[i] = int(c);
    atomicAdd(&counts[c], 1ull);
}

__global__ void knn_scatter_kernel(const float* __restrict__ p, int64_t n, const int* __restrict__ cell_of,
                                   unsigned long long* __restrict__ cursor, float4* __restrict__ out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long s = atomicAdd(&cursor[cell_of[i]], 1ull);
    out[s] = make_float4(p[3 * i], p[3 * i + 1], p[3 * i + 2], __int_as_float(int(i)));
}

__global__ void copy_i64_kernel(const int64_t* __restrict__ a, unsigned long long* __restrict__ b, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (unsigned long long)a[i];
}

__global__ __launch_bounds__(256) void bounds_kernel(const float* __restrict__ p, int64_t n, float* __restrict__ partial) {
    __shared__ float lds[6][256];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        for (int d = 0; d < 3; ++d) {
            mn[d] = fminf(mn[d], p[3 * i + d]);
            mx[d] = fmaxf(mx[d], p[3 * i + d]);
        }
    for (int d = 0; d < 3; ++d) {
        lds[d][threadIdx.x] = mn[d];
        lds[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int d = 0; d < 3; ++d) {
                lds[d][threadIdx.x] = fminf(lds[d][threadIdx.x], lds[d][threadIdx.x + off]);
                lds[3 + d][threadIdx.x] = fmaxf(lds[3 + d][threadIdx.x], lds[3 + d][threadIdx.x + off]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = lds[threadIdx.x][0];
}

void bounds(dwg_ctx* ctx, const float* p, int64_t n, float* lo, float* hi) {
    DWG_REQUIRE(n > 0, "bounds of an empty point set");
    const int blocks = 256;
    float* partial = static_cast<float*>(ctx->buf("bounds_partial", sizeof(float) * 6 * blocks));
    hipLaunchKernelGGL(bounds_kernel, dim3(blocks), dim3(256), 0, ctx->stream, p, n, partial);
    check_launch();
    std::vector<float> h(6 * blocks);
    DWG_HIP_CHECK(hipMemcpyAsync(h.data(), partial, sizeof(float) * 6 * blocks, hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int d = 0; d < 3; ++d) {
        lo[d] = INFINITY;
        hi[d] = -INFINITY;
    }
    for (int b = 0; b < blocks; ++b)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], h[6 * b + d]);
            hi[d] = std::max(hi[d], h[6 * b + 3 + d]);
        }
}

__global__ void count_occupied_kernel(const int64_t* __restrict__ counts, int64_t nc,
                                      unsigned long long* __restrict__ out) {
    __shared__ int lds[256];
    int c = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < nc; i += int64_t(gridDim.x) * 256) c += counts[i] != 0;
    lds[threadIdx.x] = c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)lds[0]);
}

static KnnGridArgs grid_args(const float* lo, const float* hi, double cell) {
    KnnGridArgs G{};
    G.cell = float(cell);
    G.inv_cell = float(1.0 / cell);
    for (int d = 0; d < 3; ++d) {
        G.lo[d] = lo[d];
        G.dim[d] = std::max(1, int(std::floor((hi[d] - lo[d]) / cell)) + 1);
    }
    return G;
}

static double cells_of(const float* lo, const float* hi, double cell) {
    double nc = 1;
    for (int d = 0; d < 3; ++d) nc *= std::floor((hi[d] - lo[d]) / cell) + 1;
    return nc;
}

// histogram + scan (+ scatter when out != nullptr) of the points into level L
static void bucket(dwg_ctx* ctx, const float* p, int64_t n, dwg_knn_index::Level& L, bool scatter) {
    L.ncells = int64_t(L.G.dim[0]) * L.G.dim[1] * L.G.dim[2];
    if (L.cell_start) (void)hipFree(L.cell_start);
    L.cell_start = nullptr;
    DWG_HIP_CHECK(hipMalloc(&L.cell_start, sizeof(int64_t) * (L.ncells + 1)));
    int* cell_of = static_cast<int*>(ctx->buf("knn_cell_of", sizeof(int) * n));
    fill_u32(ctx, L.cell_start, 0, 2 * (L.ncells + 1));
    hipLaunchKernelGGL(knn_count_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, p, n, L.G,
                       reinterpret_cast<unsigned long long*>(L.cell_start), cell_of);
    check_launch();
    if (!scatter) return;
    unsigned long long* cursor = static_cast<unsigned long long*>(ctx->buf("knn_cursor", sizeof(uint64_t) * L.ncells));
    exclusive_scan<int64_t>(ctx, L.cell_start, L.ncells);
    hipLaunchKernelGGL(copy_i64_kernel, dim3(grid_for(L.ncells, 256)), dim3(256), 0, ctx->stream, L.cell_start,
                       cursor, L.ncells);
    check_launch();
    DWG_HIP_CHECK(hipMalloc(&L.pts, sizeof(float4) * n));
    hipLaunchKernelGGL(knn_scatter_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, p, n, cell_of, cursor,
                       L.pts);
    check_launch();
}

dwg_knn_index* knn_build(dwg_ctx* ctx, const float* p, int64_t n, int k_hint) {
    DWG_REQUIRE(n > 0 && n < (int64_t(1) << 31) - 1, "kNN index needs 1 ≤ n < 2^31 points");
    float lo[3], hi[3];
    bounds(ctx, p, n, lo, hi);
    auto* I = new dwg_knn_index();
    try {
        I->n = n;
        ProfileScope ps(ctx, "knn_build");
        const double ext = std::max({double(hi[0] - lo[0]), double(hi[1] - lo[1]), double(hi[2] - lo[2]), 1e-6});
        const double max_cells = double(1 << 27);
        // about `target` points per occupied cell puts ~k points in the 3×3×3
        // block around a query on a surface-like cloud
        const double target = std::max(2.0, 0.3 * double(std::max(k_hint, 1)));
        double cell = ext / std::max(1.0, std::sqrt(double(n) / target));
        while (cells_of(lo, hi, cell) > max_cells) cell *= 1.25;
        // measure the occupancy at this cell size and rescale (occupancy of a
        // surface-like cloud grows with cell²)
        dwg_knn_index::Level probe;
        probe.G = grid_args(lo, hi, cell);
        bucket(ctx, p, n, probe, false);
        unsigned long long* occ = static_cast<unsigned long long*>(ctx->buf("knn_occ", sizeof(uint64_t)));
        fill_u32(ctx, occ, 0, 2);
        hipLaunchKernelGGL(count_occupied_kernel, dim3(512), dim3(256), 0, ctx->stream, probe.cell_start, probe.ncells,
                           occ);
        check_launch();
        const double occupied = double(std::max<unsigned long long>(1, read_scalar<unsigned long long>(ctx, occ)));
        (void)hipFree(probe.cell_start);
        probe.cell_start = nullptr;
        const double o = double(n) / occupied;
        double scale = std::sqrt(target / o);
        scale = std::min(16.0, std::max(0.25, scale));
        cell *= scale;
        while (cells_of(lo, hi, cell) > max_cells) cell *= 1.25;
        I->lvl[0].G = grid_args(lo, hi, cell);
        bucket(ctx, p, n, I->lvl[0], true);
        I->lvl[1].G = grid_args(lo, hi, cell * 8.0);
        bucket(ctx, p, n, I->lvl[1], true);
    } catch (...) {
        delete I;
        throw;
    }
    return I;
}

template <int K>
__device__ __forceinline__ void knn_insert(uint64_t (&best)[K], int k, uint64_t key) {
    uint64_t v = key;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        if (i < k) {
            const uint64_t b = best[i];
            const bool sw = v < b;
            best[i] = sw ? v : b;
            v = sw ? b : v;
        }
    }
}

template <int K>
__device__ __forceinline__ uint64_t kth_of(const uint64_t (&best)[K], int k) {
    uint64_t v = best[0];
#pragma unroll
    for (int j = 1; j < K; ++j)
        if (j == k - 1) v = best[j];
    return v;
}

// Visits shells r = 0..rmax of cells around q on one grid level, keeping
// the k smallest (d², index) keys.  Returns true once no unvisited cell can
// hold a closer point (the search is then exact on this level).
template <int K>
__device__ bool search_level(const KnnGridArgs& G, const int64_t* __restrict__ cell_start,
                             const float4* __restrict__ pts, float qx, float qy, float qz, int k,
                             const uint8_t* __restrict__ skip, uint64_t (&best)[K], int rmax) {
    uint64_t kth = ~0ull;
#pragma unroll
    for (int j = 0; j < K; ++j) best[j] = ~0ull;
    const int c0x = cell_coord(qx, G.lo[0], G.inv_cell, G.dim[0]);
    const int c0y = cell_coord(qy, G.lo[1], G.inv_cell, G.dim[1]);
    const int c0z = cell_coord(qz, G.lo[2], G.inv_cell, G.dim[2]);
    for (int r = 0; r <= rmax; ++r) {
        const int zlo = max(c0z - r, 0), zhi = min(c0z + r, G.dim[2] - 1);
        const int ylo = max(c0y - r, 0), yhi = min(c0y + r, G.dim[1] - 1);
        for (int z = zlo; z <= zhi; ++z) {
            const bool zedge = abs(z - c0z) == r;
            for (int y = ylo; y <= yhi; ++y) {
                const bool yedge = zedge || abs(y - c0y) == r;
                // rows on the shell's faces are scanned fully, other rows only at both ends
                const int xs = yedge ? max(c0x - r, 0) : c0x - r;
                const int step = yedge ? 1 : 2 * r;
                const int xe = min(c0x + r, G.dim[0] - 1);
                for (int x = xs; x <= xe; x += step) {
                    if (x < 0) continue;
                    const int64_t c = x + int64_t(G.dim[0]) * (y + int64_t(G.dim[1]) * z);
                    const int64_t sb = cell_start[c], se = cell_start[c + 1];
                    for (int64_t s = sb; s < se; ++s) {
                        const float4 p = pts[s];
                        const int idx = __float_as_int(p.w);
                        if (skip && skip[idx]) continue;
                        const float d2 = dist2(qx - p.x, qy - p.y, qz - p.z);
                        const uint64_t key = (uint64_t(__float_as_uint(d2)) << 32) | uint32_t(idx);
                        if (key < kth) {
                            knn_insert<K>(best, k, key);
                            kth = kth_of<K>(best, k);
                        }
                    }
                }
            }
        }
        // distance from q to the nearest face of the searched block that is
        // not on the grid boundary; no unvisited point can be closer
        float dmin = INFINITY;
        if (c0x - r > 0) dmin = fminf(dmin, qx - (G.lo[0] + float(c0x - r) * G.cell));
        if (c0x + r + 1 < G.dim[0]) dmin = fminf(dmin, (G.lo[0] + float(c0x + r + 1) * G.cell) - qx);
        if (c0y - r > 0) dmin = fminf(dmin, qy - (G.lo[1] + float(c0y - r) * G.cell));
        if (c0y + r + 1 < G.dim[1]) dmin = fminf(dmin, (G.lo[1] + float(c0y + r + 1) * G.cell) - qy);
        if (c0z - r > 0) dmin = fminf(dmin, qz - (G.lo[2] + float(c0z - r) * G.cell));
        if (c0z + r + 1 < G.dim[2]) dmin = fminf(dmin, (G.lo[2] + float(c0z + r + 1) * G.cell) - qz);
        if (dmin == INFINITY) return true;  // the block covers the whole grid
        if (kth != ~0ull) {
            const float kd2 = __uint_as_float(uint32_t(kth >> 32));
            const float safe = dmin * (1.0f - 1e-5f) - 1e-6f * G.cell;
            if (safe > 0.f && kd2 < safe * safe) return true;
        }
    }
    return false;
}

template <int K>
__global__ __launch_bounds__(256) void knn_query_kernel(KnnGridArgs G0, const int64_t* __restrict__ cs0,
                                                        const float4* __restrict__ p0, KnnGridArgs G1,
                                                        const int64_t* __restrict__ cs1,
                                                        const float4* __restrict__ p1, const float* __restrict__ q,
                                                        int64_t nq, int k, const uint8_t* __restrict__ skip,
                                                        const uint8_t* __restrict__ qmask, int32_t* __restrict__ out_idx,
                                                        float* __restrict__ out_d2) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    if (qmask && !qmask[i]) return;
    const float qx = q[3 * i], qy = q[3 * i + 1], qz = q[3 * i + 2];
    uint64_t best[K];
    if (!search_level<K>(G0, cs0, p0, qx, qy, qz, k, skip, best, 2)) {
        const int maxr = max(G1.dim[0], max(G1.dim[1], G1.dim[2]));
        (void)search_level<K>(G1, cs1, p1, qx, qy, qz, k, skip, best, maxr);
    }
    for (int j = 0; j < K; ++j) {
        if (j >= k) break;
        const uint64_t b = best[j];
        out_idx[i * k + j] = b == ~0ull ? -1 : int32_t(uint32_t(b));
        if (out_d2) out_d2[i * k + j] = b == ~0ull ? INFINITY : __uint_as_float(uint32_t(b >> 32));
    }
}

void knn_query(dwg_ctx* ctx, const dwg_knn_index* I, const float* q, int64_t nq, int k, const uint8_t* skip,
               const uint8_t* qmask, int32_t* out_idx, float* out_d2, const char* tag) {
    DWG_REQUIRE(I, "null kNN index");
    DWG_REQUIRE(k >= 1 && k <= 32, "k must be in [1, 32]");
    DWG_REQUIRE(nq >= 0, "negative query count");
    if (nq == 0) return;
    dim3 grid(grid_for(nq, 256)), block(256);
    ProfileScope ps(ctx, tag ? tag : "knn_query");
#define DWG_KNN(KK) \
    hipLaunchKernelGGL(knn_query_kernel<KK>, grid, block, 0, ctx->stream, I->lvl[0].G, I->lvl[0].cell_start,   \
                       I->lvl[0].pts, I->lvl[1].G, I->lvl[1].cell_start, I->lvl[1].pts, q, nq, k, skip, qmask, \
                       out_idx, out_d2)
    if (k <= 8) DWG_KNN(8);
    else if (k <= 10) DWG_KNN(10);
    else if (k <= 16) DWG_KNN(16);
    else DWG_KNN(32);
#undef DWG_KNN
    check_launch();
}

}  // namespace dwg

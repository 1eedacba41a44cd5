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
        dwg::dev_free(L.cell_start);
        dwg::dev_free(L.pts);
        dwg::dev_free(L.empty);
        dwg::dev_free(L.pyr);
        dwg::dev_free(L.pyr_dim);
        dwg::dev_free(L.pyr_off);
        dwg::dev_free(L.pbox);
    }
}

namespace dwg {

__device__ __forceinline__ int cell_coord(float v, float lo, float inv_cell, int dim) {
    int c = int(floorf((v - lo) * inv_cell));
    return min(max(c, 0), dim - 1);
}

// Points and queries arrive in runs that share a cell or tile (sources in
// tree order, triangle centroids in brick order): each wave adds one run
// with one atomic.  A run starts at lane 0 and wherever the key changes; its
// first lane (the head) reserves the run's slots, the others take theirs
// from the head in lane order.
__device__ __forceinline__ uint64_t run_starts(int tl) {
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(tl, 1, 64);
    return __ballot(lane == 0 || tl != prev);
}

// lanes from `lane` up to the next run start (or the end of the wave)
__device__ __forceinline__ int run_length(uint64_t starts, int lane) {
    const uint64_t after = lane == 63 ? 0ull : starts >> (lane + 1);
    return after ? __builtin_ctzll(after) + 1 : 64 - lane;
}

// the run start at or before `lane`
__device__ __forceinline__ int run_head(uint64_t starts, int lane) {
    const uint64_t upto = lane == 63 ? starts : starts & ((2ull << lane) - 1ull);
    return 63 - __builtin_clzll(upto);
}

__global__ void knn_count_kernel(const float* __restrict__ p, int64_t n, KnnGridArgs G,
                                 unsigned long long* __restrict__ counts, int* __restrict__ cell_of) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int c = -1;  // every lane takes part in the run detection
    if (i < n) {
        const int cx = cell_coord(p[3 * i], G.lo[0], G.inv_cell, G.dim[0]);
        const int cy = cell_coord(p[3 * i + 1], G.lo[1], G.inv_cell, G.dim[1]);
        const int cz = cell_coord(p[3 * i + 2], G.lo[2], G.inv_cell, G.dim[2]);
        c = int(cx + int64_t(G.dim[0]) * (cy + int64_t(G.dim[1]) * cz));
        cell_of[i] = c;
    }
    const int lane = threadIdx.x & 63;
    const uint64_t starts = run_starts(c);
    if (c >= 0 && ((starts >> lane) & 1ull)) atomicAdd(&counts[c], (unsigned long long)run_length(starts, lane));
}

__global__ void knn_scatter_kernel(const float* __restrict__ p, int64_t n, const int* __restrict__ cell_of,
                                   unsigned long long* __restrict__ cursor, float4* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int c = i < n ? cell_of[i] : -1;
    const int lane = threadIdx.x & 63;
    const uint64_t starts = run_starts(c);
    int64_t base = 0;
    if (c >= 0 && ((starts >> lane) & 1ull))
        base = int64_t(atomicAdd(&cursor[c], (unsigned long long)run_length(starts, lane)));
    const int head = run_head(starts, lane);
    const int64_t hb = (int64_t(__shfl(int(uint32_t(base >> 32)), head, 64)) << 32) |
                       int64_t(uint32_t(__shfl(int(uint32_t(base)), head, 64)));
    if (c >= 0) out[hb + (lane - head)] = make_float4(p[3 * i], p[3 * i + 1], p[3 * i + 2], __int_as_float(int(i)));
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
    dwg::dev_free(L.cell_start);
    L.cell_start = nullptr;
    L.cell_start = static_cast<decltype(L.cell_start)>(dev_malloc(sizeof(int64_t) * (L.ncells + 1)));
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
    L.pts = static_cast<decltype(L.pts)>(dev_malloc(sizeof(float4) * n));
    hipLaunchKernelGGL(knn_scatter_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, p, n, cell_of, cursor,
                       L.pts);
    check_launch();
}

// Chebyshev distance (in cells) from every cell to the nearest occupied one,
// as three separable passes: along x the distance to the nearest occupied
// cell of the row, then d(y) = min_y' max(|y − y'|, d_prev(y')) along y and z.
__global__ void empty_dist_x_kernel(const int64_t* __restrict__ cell_start, int dx, int dy, int dz,
                                    int* __restrict__ out) {
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= int64_t(dx) * dy * dz) return;
    const int x = int(c % dx);
    const int64_t row = c - x;
    int best = 1 << 28;
    for (int x2 = 0; x2 < dx; ++x2)
        if (cell_start[row + x2 + 1] > cell_start[row + x2]) best = min(best, abs(x2 - x));
    out[c] = best;
}

__global__ void empty_dist_pass_kernel(const int* __restrict__ in, int dx, int dy, int dz, int axis,
                                       int* __restrict__ out) {
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= int64_t(dx) * dy * dz) return;
    const int x = int(c % dx), y = int((c / dx) % dy), z = int(c / (int64_t(dx) * dy));
    const int len = axis == 1 ? dy : dz, pos = axis == 1 ? y : z;
    const int64_t stride = axis == 1 ? int64_t(dx) : int64_t(dx) * dy;
    const int64_t base = c - int64_t(pos) * stride;
    int best = 1 << 28;
    for (int t = 0; t < len; ++t) best = min(best, max(abs(t - pos), in[base + int64_t(t) * stride]));
    out[c] = best;
    (void)x;
}

static void empty_distance(dwg_ctx* ctx, dwg_knn_index::Level& L) {
    const int dx = L.G.dim[0], dy = L.G.dim[1], dz = L.G.dim[2];
    L.empty = static_cast<decltype(L.empty)>(dev_malloc(sizeof(int) * L.ncells));
    int* tmp = static_cast<int*>(ctx->buf("knn_empty_tmp", sizeof(int) * L.ncells));
    const unsigned g = grid_for(L.ncells, 256);
    hipLaunchKernelGGL(empty_dist_x_kernel, dim3(g), dim3(256), 0, ctx->stream, L.cell_start, dx, dy, dz, L.empty);
    check_launch();
    hipLaunchKernelGGL(empty_dist_pass_kernel, dim3(g), dim3(256), 0, ctx->stream, L.empty, dx, dy, dz, 1, tmp);
    check_launch();
    hipLaunchKernelGGL(empty_dist_pass_kernel, dim3(g), dim3(256), 0, ctx->stream, tmp, dx, dy, dz, 2, L.empty);
    check_launch();
}

// Occupancy pyramid of a level: byte b of level l ≥ 1 has bit c set when
// child c (x + 2y + 4z) of block b, at level l − 1, holds a point.
__global__ void pyr_first_kernel(const int64_t* __restrict__ cell_start, int dx, int dy, int dz, int ux, int uy, int uz,
                                 uint8_t* __restrict__ out) {
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= int64_t(ux) * uy * uz) return;
    const int x = int(b % ux), y = int((b / ux) % uy), z = int(b / (int64_t(ux) * uy));
    uint32_t m = 0;
    for (int c = 0; c < 8; ++c) {
        const int X = 2 * x + (c & 1), Y = 2 * y + ((c >> 1) & 1), Z = 2 * z + (c >> 2);
        if (X < dx && Y < dy && Z < dz) {
            const int64_t cell = X + int64_t(dx) * (Y + int64_t(dy) * Z);
            if (cell_start[cell + 1] > cell_start[cell]) m |= 1u << c;
        }
    }
    out[b] = uint8_t(m);
}

__global__ void pyr_up_kernel(const uint8_t* __restrict__ lower, int dx, int dy, int dz, int ux, int uy, int uz,
                              uint8_t* __restrict__ out) {
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= int64_t(ux) * uy * uz) return;
    const int x = int(b % ux), y = int((b / ux) % uy), z = int(b / (int64_t(ux) * uy));
    uint32_t m = 0;
    for (int c = 0; c < 8; ++c) {
        const int X = 2 * x + (c & 1), Y = 2 * y + ((c >> 1) & 1), Z = 2 * z + (c >> 2);
        if (X < dx && Y < dy && Z < dz && lower[X + int64_t(dx) * (Y + int64_t(dy) * Z)]) m |= 1u << c;
    }
    out[b] = uint8_t(m);
}

// Bounding boxes of the points themselves, per cell and per pyramid block:
// a query far from the cloud is as close to a box of a surface patch as to
// the patch, where the cell's own box can be a cell width nearer.  Exact
// bounds of float points need no rounding allowance: the box distance,
// computed like a point's d², never exceeds the d² of a point inside.
__global__ void cell_box_kernel(const int64_t* __restrict__ cell_start, const float4* __restrict__ pts, int64_t ncells,
                                float4* __restrict__ box) {
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
    for (int64_t s = cell_start[c], e = cell_start[c + 1]; s < e; ++s) {
        const float4 p = pts[s];
        lo.x = fminf(lo.x, p.x);
        lo.y = fminf(lo.y, p.y);
        lo.z = fminf(lo.z, p.z);
        hi.x = fmaxf(hi.x, p.x);
        hi.y = fmaxf(hi.y, p.y);
        hi.z = fmaxf(hi.z, p.z);
    }
    box[2 * c] = lo;
    box[2 * c + 1] = hi;
}

__global__ void box_up_kernel(const float4* __restrict__ lower, int dx, int dy, int dz, int ux, int uy, int uz,
                              float4* __restrict__ out) {
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= int64_t(ux) * uy * uz) return;
    const int x = int(b % ux), y = int((b / ux) % uy), z = int(b / (int64_t(ux) * uy));
    float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
    for (int c = 0; c < 8; ++c) {
        const int X = 2 * x + (c & 1), Y = 2 * y + ((c >> 1) & 1), Z = 2 * z + (c >> 2);
        if (X < dx && Y < dy && Z < dz) {
            const int64_t i = X + int64_t(dx) * (Y + int64_t(dy) * Z);
            const float4 a = lower[2 * i], h = lower[2 * i + 1];
            lo.x = fminf(lo.x, a.x);
            lo.y = fminf(lo.y, a.y);
            lo.z = fminf(lo.z, a.z);
            hi.x = fmaxf(hi.x, h.x);
            hi.y = fmaxf(hi.y, h.y);
            hi.z = fmaxf(hi.z, h.z);
        }
    }
    out[2 * b] = lo;
    out[2 * b + 1] = hi;
}

// the points' bounding boxes of every cell and pyramid block, built the
// first time a walk needs them (only the wave-shared walk does)
static void ensure_boxes(dwg_ctx* ctx, dwg_knn_index::Level& L) {
    if (L.pbox) return;
    const int levels = L.pyr_levels;
    const int4* dims = L.pyr_dim_h;
    const int64_t* off = L.pyr_off_h;
    int64_t total = 0;
    for (int l = 1; l < levels; ++l) total += int64_t(dims[l].x) * dims[l].y * dims[l].z;
    L.pbox = static_cast<decltype(L.pbox)>(dev_malloc(sizeof(float4) * 2 * size_t(L.ncells + total)));
    hipLaunchKernelGGL(cell_box_kernel, dim3(grid_for(L.ncells, 256)), dim3(256), 0, ctx->stream, L.cell_start, L.pts,
                       L.ncells, L.pbox);
    check_launch();
    for (int l = 1; l < levels; ++l) {
        const int4 lo = dims[l - 1], up = dims[l];
        const int64_t nb = int64_t(up.x) * up.y * up.z;
        const float4* lower = L.pbox + 2 * (l == 1 ? 0 : L.ncells + off[l - 1]);
        hipLaunchKernelGGL(box_up_kernel, dim3(grid_for(nb, 256)), dim3(256), 0, ctx->stream, lower, lo.x, lo.y, lo.z,
                           up.x, up.y, up.z, L.pbox + 2 * (L.ncells + off[l]));
        check_launch();
    }
}

// cell boundaries are computed as lo + b·w in float and points are
// assigned by floor((v − lo)/cell): both are within a few ulp of the exact
// boundary, so a box of cells is shrunk by 16 ulp of the largest coordinate
// plus 1e-5 cell before it may exclude a point
static float cell_margin(const KnnGridArgs& G) {
    float ext = 0.f;
    for (int a = 0; a < 3; ++a)
        ext = std::max({ext, std::fabs(G.lo[a]), std::fabs(G.lo[a] + G.cell * float(G.dim[a]))});
    return 16.f * 1.1920929e-7f * ext + 1e-5f * G.cell;
}

// the first pyramid level counting only points that `skip` keeps
__global__ void pyr_first_masked_kernel(const int64_t* __restrict__ cell_start, const float4* __restrict__ pts,
                                        const uint8_t* __restrict__ skip, int dx, int dy, int dz, int ux, int uy,
                                        int uz, uint8_t* __restrict__ out) {
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= int64_t(ux) * uy * uz) return;
    const int x = int(b % ux), y = int((b / ux) % uy), z = int(b / (int64_t(ux) * uy));
    uint32_t m = 0;
    for (int c = 0; c < 8; ++c) {
        const int X = 2 * x + (c & 1), Y = 2 * y + ((c >> 1) & 1), Z = 2 * z + (c >> 2);
        if (X < dx && Y < dy && Z < dz) {
            const int64_t cell = X + int64_t(dx) * (Y + int64_t(dy) * Z);
            for (int64_t s = cell_start[cell], e = cell_start[cell + 1]; s < e; ++s)
                if (!skip[__float_as_int(pts[s].w)]) {
                    m |= 1u << c;
                    break;
                }
        }
    }
    out[b] = uint8_t(m);
}

// An occupancy pyramid of the fine level that ignores the points `skip`
// drops (same layout as Level::pyr, in a context buffer): a walk with that
// skip mask then never opens a block of skipped points only
static const uint8_t* masked_pyramid(dwg_ctx* ctx, const dwg_knn_index::Level& L, const uint8_t* skip) {
    const int levels = L.pyr_levels;
    const int4* dims = L.pyr_dim_h;
    const int64_t* off = L.pyr_off_h;
    int64_t total = 0;
    for (int l = 1; l < levels; ++l) total += int64_t(dims[l].x) * dims[l].y * dims[l].z;
    uint8_t* pyr = static_cast<uint8_t*>(ctx->buf("knn_pyr_masked", size_t(total) + 64));
    for (int l = 1; l < levels; ++l) {
        const int4 lo = dims[l - 1], up = dims[l];
        const int64_t nb = int64_t(up.x) * up.y * up.z;
        if (l == 1)
            hipLaunchKernelGGL(pyr_first_masked_kernel, dim3(grid_for(nb, 256)), dim3(256), 0, ctx->stream,
                               L.cell_start, L.pts, skip, lo.x, lo.y, lo.z, up.x, up.y, up.z, pyr + off[1]);
        else
            hipLaunchKernelGGL(pyr_up_kernel, dim3(grid_for(nb, 256)), dim3(256), 0, ctx->stream, pyr + off[l - 1],
                               lo.x, lo.y, lo.z, up.x, up.y, up.z, pyr + off[l]);
        check_launch();
    }
    return pyr;
}

static void pyramid(dwg_ctx* ctx, dwg_knn_index::Level& L) {
    int4* dims = L.pyr_dim_h;
    int64_t* off = L.pyr_off_h;
    int d[3] = {L.G.dim[0], L.G.dim[1], L.G.dim[2]};
    dims[0] = make_int4(d[0], d[1], d[2], 0);
    off[0] = 0;
    int levels = 1;
    int64_t total = 0;
    while (d[0] > 1 || d[1] > 1 || d[2] > 1) {
        if (levels == kMaxPyr) return;  // no pyramid: queries take the shell search
        for (int a = 0; a < 3; ++a) d[a] = (d[a] + 1) / 2;
        dims[levels] = make_int4(d[0], d[1], d[2], 0);
        off[levels] = total;
        total += int64_t(d[0]) * d[1] * d[2];
        ++levels;
    }
    if (levels == 1) return;  // one cell: the shell search is already optimal
    L.pyr = static_cast<decltype(L.pyr)>(dev_malloc(size_t(total)));
    L.pyr_dim = static_cast<decltype(L.pyr_dim)>(dev_malloc(sizeof(int4) * kMaxPyr));
    L.pyr_off = static_cast<decltype(L.pyr_off)>(dev_malloc(sizeof(int64_t) * kMaxPyr));
    DWG_HIP_CHECK(hipMemcpyAsync(L.pyr_dim, dims, sizeof(int4) * levels, hipMemcpyHostToDevice, ctx->stream));
    DWG_HIP_CHECK(hipMemcpyAsync(L.pyr_off, off, sizeof(int64_t) * levels, hipMemcpyHostToDevice, ctx->stream));
    for (int l = 1; l < levels; ++l) {
        const int4 lo = dims[l - 1], up = dims[l];
        const int64_t nb = int64_t(up.x) * up.y * up.z;
        if (l == 1)
            hipLaunchKernelGGL(pyr_first_kernel, dim3(grid_for(nb, 256)), dim3(256), 0, ctx->stream, L.cell_start, lo.x,
                               lo.y, lo.z, up.x, up.y, up.z, L.pyr + off[1]);
        else
            hipLaunchKernelGGL(pyr_up_kernel, dim3(grid_for(nb, 256)), dim3(256), 0, ctx->stream, L.pyr + off[l - 1],
                               lo.x, lo.y, lo.z, up.x, up.y, up.z, L.pyr + off[l]);
        check_launch();
    }
    L.pyr_margin = cell_margin(L.G);
    L.pyr_levels = levels;
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
        dev_free(probe.cell_start);
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
        empty_distance(ctx, I->lvl[1]);
        pyramid(ctx, I->lvl[0]);
    } catch (...) {
        delete I;
        throw;
    }
    return I;
}

// The k best keys of a query live in registers as a descending list:
// best[0] is the k-th smallest key so far (the insertion threshold), the
// slots k..K−1 hold 0 and never move (every key is ≥ 0).  Every access uses
// a constant index, so the list is never demoted to scratch memory.
// An empty slot: the bits of +∞ as a double, above every key (the high
// word of a key is the float bits of a d² ≥ 0, at most 0x7F800000)
constexpr uint64_t kNoKey = 0x7FF0000000000000ull;

template <int K>
__device__ __forceinline__ void knn_reset(uint64_t (&best)[K], int k) {
#pragma unroll
    for (int j = 0; j < K; ++j) best[j] = j < k ? kNoKey : 0ull;
}

// key < best[0]: it replaces the largest key and sinks to its place.  Every
// key (and kNoKey, and the 0 of an unused slot) read as a double is a
// non-negative number, never NaN, and non-negative doubles order as their
// bit patterns do (the kernels keep FP64 denormals): each compare-exchange
// is one FP64 max and one min, which return one of their operands bit for
// bit, instead of a 64-bit compare and four selects.  Written as the two
// instructions themselves: the compiler would first quiet possible
// signalling NaNs of both operands, which these keys never are
__device__ __forceinline__ uint64_t key_max(uint64_t a, uint64_t b) {
    uint64_t r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ uint64_t key_min(uint64_t a, uint64_t b) {
    uint64_t r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

template <int K>
__device__ __forceinline__ void knn_insert(uint64_t (&best)[K], uint64_t key) {
    best[0] = key;
#pragma unroll
    for (int i = 0; i < K - 1; ++i) {
        const uint64_t a = best[i], b = best[i + 1];
        best[i] = key_max(a, b);
        best[i + 1] = key_min(a, b);
    }
}

// writes the k keys in ascending order (index −1 / d² ∞ where fewer were found)
template <int K>
__device__ __forceinline__ void knn_store(const uint64_t (&best)[K], int k, int64_t row, int32_t* __restrict__ out_idx,
                                          float* __restrict__ out_d2) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j < k) {
            const uint64_t b = best[j];
            const int64_t o = row * k + (k - 1 - j);
            out_idx[o] = b == kNoKey ? -1 : int32_t(uint32_t(b));
            if (out_d2) out_d2[o] = b == kNoKey ? INFINITY : __uint_as_float(uint32_t(b >> 32));
        }
    }
}

// Visits shells r = 0..rmax of cells around q on one grid level, keeping
// the k smallest (d², index) keys.  Returns true once no unvisited cell can
// hold a closer point (the search is then exact on this level).
template <int K>
__device__ bool search_level(const KnnGridArgs& G, const int64_t* __restrict__ cell_start,
                             const float4* __restrict__ pts, float qx, float qy, float qz, int k,
                             const uint8_t* __restrict__ skip, uint64_t (&best)[K], int rmax,
                             const int* __restrict__ empty) {
    knn_reset<K>(best, k);
    const int c0x = cell_coord(qx, G.lo[0], G.inv_cell, G.dim[0]);
    const int c0y = cell_coord(qy, G.lo[1], G.inv_cell, G.dim[1]);
    const int c0z = cell_coord(qz, G.lo[2], G.inv_cell, G.dim[2]);
    // every cell closer (Chebyshev) than empty[c0] holds no point: start there
    const int r0 = empty ? min(empty[c0x + int64_t(G.dim[0]) * (c0y + int64_t(G.dim[1]) * c0z)], rmax) : 0;
    for (int r = r0; r <= rmax; ++r) {
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
                        const float d2 = dist2(qx - p.x, qy - p.y, qz - p.z);
                        const uint64_t key = (uint64_t(__float_as_uint(d2)) << 32) | uint32_t(idx);
                        // the skip byte only for a point that would enter the list
                        if (key < best[0] && !(skip && skip[idx])) knn_insert<K>(best, key);
                    }
                }
            }
        }
        const uint64_t kth = best[0];
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
        if (kth != kNoKey) {
            const float kd2 = __uint_as_float(uint32_t(kth >> 32));
            const float safe = dmin * (1.0f - 1e-5f) - 1e-6f * G.cell;
            if (safe > 0.f && kd2 < safe * safe) return true;
        }
    }
    return false;
}

// One query per lane; `qlist` (nullable) lists the queries to answer
// (nq entries), otherwise queries 0..nq−1.
template <int K>
__global__ __launch_bounds__(256) void knn_query_kernel(KnnGridArgs G0, const int64_t* __restrict__ cs0,
                                                        const float4* __restrict__ p0, KnnGridArgs G1,
                                                        const int64_t* __restrict__ cs1,
                                                        const float4* __restrict__ p1, const int* __restrict__ empty1,
                                                        const float* __restrict__ q, int64_t nq,
                                                        const int* __restrict__ qlist, int k,
                                                        const uint8_t* __restrict__ skip,
                                                        const uint8_t* __restrict__ qmask, int32_t* __restrict__ out_idx,
                                                        float* __restrict__ out_d2) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const int64_t i = qlist ? int64_t(qlist[t]) : t;
    if (qmask && !qmask[i]) return;
    const float qx = q[3 * i], qy = q[3 * i + 1], qz = q[3 * i + 2];
    uint64_t best[K];
    // a query whose coarse cell is empty skips the fine shells
    const int c1 = cell_coord(qx, G1.lo[0], G1.inv_cell, G1.dim[0]) +
                   G1.dim[0] * (cell_coord(qy, G1.lo[1], G1.inv_cell, G1.dim[1]) +
                                G1.dim[1] * cell_coord(qz, G1.lo[2], G1.inv_cell, G1.dim[2]));
    if (empty1[c1] > 0 || !search_level<K>(G0, cs0, p0, qx, qy, qz, k, skip, best, 2, nullptr)) {
        const int maxr = max(G1.dim[0], max(G1.dim[1], G1.dim[2]));
        (void)search_level<K>(G1, cs1, p1, qx, qy, qz, k, skip, best, maxr, empty1);
    }
    knn_store<K>(best, k, i, out_idx, out_d2);
}

// ---------------------------------------------------------------------------
// Tiled queries.  Queries are bucketed by tile (kTile³ fine cells); one
// workgroup per non-empty tile stages the points of the tile plus a kHalo
// ring of cells in LDS and every lane (one query each) scans the same
// candidates in lockstep, so LDS reads are broadcasts and no lane diverges.
// A query is final when its k-th key is closer than every face of the
// staged box that is not on the grid boundary; the rest (sparse or clustered
// regions) go to the per-lane shell search above.  Both paths keep the k
// smallest (d², index) keys, so the result does not depend on the path.
// ---------------------------------------------------------------------------
constexpr double kScatterFix = 1099511627776.0;  // 2^40
constexpr int kTile = 4;
constexpr int kStage = 512;
constexpr int kTileThreads = 256;
// halos of the tiled passes: most queries close within 2 cells; the ones
// that do not (off-surface centroids) get passes with 5 and 9 cells
constexpr int kHalo1 = 2;
constexpr int kHalo2 = 5;
constexpr int kHalo3 = 9;
// the tiled passes bucket their queries by 2×2×2-cell eighth of the tile
// (8 buckets per tile), so that the 64 queries of a wave sit close together
// (first pass 2.07 → 1.97 ms, and the far walk that takes its open queries
// in this order 1.89 → 1.81 ms); the sparse walk's order keeps one bucket
// per tile (0.86 instead of 0.93 ms)
constexpr int kSubTiled = 8;

// exclusive sum over the block's threads (NT threads, a multiple of 64);
// every thread calls it; wsum holds NT / 64 ints of LDS
template <int NT>
__device__ __forceinline__ int block_exclusive_sum(int v, int* wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int before = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) before += i < w ? wsum[i] : 0;
    return before + incl - v;
}

struct TileArgs {
    int tdim[3];   // tiles per axis
    float margin;  // rounding allowance of cell boundaries (cell_margin)
    int sub;       // query buckets per tile: 1 or 8 (by eighth of the tile)
};

// bucket of a point: its tile, times 8 plus the eighth of the tile it lies
// in when T.sub is 8
__device__ __forceinline__ int bucket_of_point(const KnnGridArgs& G, const TileArgs& T, float x, float y, float z) {
    const int cx = cell_coord(x, G.lo[0], G.inv_cell, G.dim[0]);
    const int cy = cell_coord(y, G.lo[1], G.inv_cell, G.dim[1]);
    const int cz = cell_coord(z, G.lo[2], G.inv_cell, G.dim[2]);
    const int tile = cx / kTile + T.tdim[0] * (cy / kTile + T.tdim[1] * (cz / kTile));
    if (T.sub == 1) return tile;
    constexpr int h = kTile / 2;
    return tile * 8 + int(cx % kTile >= h) + 2 * int(cy % kTile >= h) + 4 * int(cz % kTile >= h);
}

// query t of the pass is qlist[t] (or t); masked queries are left out
__global__ void tile_count_kernel(const float* __restrict__ q, int64_t nq, const int* __restrict__ qlist,
                                  const uint8_t* __restrict__ qmask, KnnGridArgs G, TileArgs T,
                                  int* __restrict__ tile_of, unsigned long long* __restrict__ count) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int tl = -1;  // every lane takes part in the run detection
    if (t < nq) {
        const int64_t i = qlist ? int64_t(qlist[t]) : t;
        if (!qmask || qmask[i]) tl = bucket_of_point(G, T, q[3 * i], q[3 * i + 1], q[3 * i + 2]);
        tile_of[t] = tl;
    }
    const int lane = threadIdx.x & 63;
    const uint64_t starts = run_starts(tl);
    if (tl >= 0 && ((starts >> lane) & 1ull)) atomicAdd(&count[tl], (unsigned long long)run_length(starts, lane));
}

__global__ void tile_scatter_kernel(const int* __restrict__ tile_of, int64_t nq, const int* __restrict__ qlist,
                                    unsigned long long* __restrict__ cursor, int* __restrict__ order) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int tl = t < nq ? tile_of[t] : -1;
    const int lane = threadIdx.x & 63;
    const uint64_t starts = run_starts(tl);
    int base = 0;
    if (tl >= 0 && ((starts >> lane) & 1ull))
        base = int(atomicAdd(&cursor[tl], (unsigned long long)run_length(starts, lane)));
    const int head = run_head(starts, lane);
    const int hb = __shfl(base, head, 64);
    if (tl >= 0) order[hb + (lane - head)] = qlist ? qlist[t] : int(t);
}

__global__ void tile_flags_kernel(const int64_t* __restrict__ start, int64_t ntiles, int sub,
                                  int64_t* __restrict__ flags) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    flags[t] = (t < ntiles && start[sub * (t + 1)] > start[sub * t]) ? 1 : 0;
}

__global__ void tile_list_kernel(const int64_t* __restrict__ flags, int64_t ntiles, int* __restrict__ list) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= ntiles || flags[t + 1] == flags[t]) return;
    list[flags[t]] = int(t);
}

// Staged candidates are ordered by ring: the rows of cells that cross the
// tile first, then the halo rings outwards, so a query meets its nearest
// points early and its k-th key (the insertion threshold) tightens fast.
// SCATTER: besides its neighbours, every closed query adds its value
// (3 floats) in 2^-40 fixed point into acc[neighbour] — the normal scatter
// of the diffusion (TurboDWG.cu:179-204).  The tile gathers these sums in an
// LDS hash table keyed by point index and flushes each key once with 64-bit
// global atomics; integer sums are associative, so the result equals
// per-pair atomics bit for bit.
template <int K, int HALO, bool SCATTER>
__global__ __launch_bounds__(kTileThreads) void knn_tile_kernel(
    KnnGridArgs G, TileArgs T, const int64_t* __restrict__ cell_start, const float4* __restrict__ pts,
    const int* __restrict__ tiles, const int64_t* __restrict__ tile_start, const int* __restrict__ order,
    const float* __restrict__ q, int k, const uint8_t* __restrict__ skip, int32_t* __restrict__ out_idx,
    float* __restrict__ out_d2, int* __restrict__ redo, float* __restrict__ redo_kth, unsigned int* __restrict__ n_redo,
    int64_t* __restrict__ redo_flag, KnnScatter sc) {
    constexpr int kRegion = kTile + 2 * HALO;
    // the first pass splits each row into the tile's own cells and the halo
    // cells at x distance 1 and 2 on either side, so that the staging order
    // (by Chebyshev ring from the tile) starts with the tile's own cells
    constexpr int kParts = HALO == kHalo1 ? 5 : 1;
    constexpr int kRows = kRegion * kRegion * kParts;
    constexpr int kHash = SCATTER ? 256 : 1;
    __shared__ int s_hkey[kHash];
    __shared__ unsigned long long s_hacc[kHash][3];
    if (SCATTER) {
        for (int h = threadIdx.x; h < kHash; h += kTileThreads) {
            s_hkey[h] = -1;
            s_hacc[h][0] = s_hacc[h][1] = s_hacc[h][2] = 0ull;
        }
    }
    __shared__ float4 s_pts[kStage];
    __shared__ int64_t s_rb[kRows];
    __shared__ int s_cnt[kRows];
    __shared__ int s_pref[kRows + 1];
    __shared__ int s_ryz[kRows];  // y | z << 16 of each row
    __shared__ int s_rx[kRows];   // its cells xa | xb << 16
    __shared__ int s_wsum[kTileThreads / 64];
    __shared__ int s_ring[HALO + 2];
    const int tile = tiles[blockIdx.x];
    const int tx = tile % T.tdim[0], ty = (tile / T.tdim[0]) % T.tdim[1], tz = tile / (T.tdim[0] * T.tdim[1]);
    const int x0 = max(tx * kTile - HALO, 0), x1 = min(tx * kTile + kTile - 1 + HALO, G.dim[0] - 1);
    const int y0 = max(ty * kTile - HALO, 0), y1 = min(ty * kTile + kTile - 1 + HALO, G.dim[1] - 1);
    const int z0 = max(tz * kTile - HALO, 0), z1 = min(tz * kTile + kTile - 1 + HALO, G.dim[2] - 1);
    const int ny = y1 - y0 + 1, nrows = ny * (z1 - z0 + 1) * kParts;
    if (threadIdx.x < HALO + 2) s_ring[threadIdx.x] = 0;
    __syncthreads();
    // each (y, z) row of cells x0..x1 (or each of its parts) is one
    // contiguous range of points; a thread takes rows threadIdx.x +
    // h · kTileThreads, h < kRowIters
    constexpr int kRowIters = (kRows + kTileThreads - 1) / kTileThreads;
    int ring[kRowIters], rank[kRowIters], rc[kRowIters], rx[kRowIters];
    int64_t rb[kRowIters];
#pragma unroll
    for (int h = 0; h < kRowIters; ++h) {
        ring[h] = rank[h] = rc[h] = rx[h] = 0;
        rb[h] = 0;
        const int r = threadIdx.x + h * kTileThreads;
        if (r < nrows) {
            const int row = r / kParts, part = r % kParts;
            const int y = y0 + row % ny, z = z0 + row / ny;
            // the part's cells xa..xb and their x distance from the tile
            int xa = x0, xb = x1, dx = 0;
            if (kParts > 1) {
                const int t0 = tx * kTile, t1 = tx * kTile + kTile - 1;
                if (part == 0) { xa = t0; xb = t1; }
                else if (part <= 2) { xa = xb = t0 - part; dx = part; }
                else { xa = xb = t1 + (part - 2); dx = part - 2; }
                xa = max(xa, x0);
                xb = min(xb, x1);
            }
            const int64_t c = xa + int64_t(G.dim[0]) * (y + int64_t(G.dim[1]) * z);
            rb[h] = cell_start[c];
            rc[h] = xb >= xa ? int(cell_start[c + (xb - xa) + 1] - rb[h]) : 0;
            rx[h] = xa | (xb << 16);
            const int dy = max(max(ty * kTile - y, y - (ty * kTile + kTile - 1)), 0);
            const int dz = max(max(tz * kTile - z, z - (tz * kTile + kTile - 1)), 0);
            ring[h] = max(max(dy, dz), dx);
            rank[h] = atomicAdd(&s_ring[ring[h] + 1], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int r = 1; r <= HALO + 1; ++r) s_ring[r] += s_ring[r - 1];
    __syncthreads();
#pragma unroll
    for (int h = 0; h < kRowIters; ++h) {
        if (threadIdx.x + h * kTileThreads < nrows) {
            const int pos = s_ring[ring[h]] + rank[h];
            const int r = (threadIdx.x + h * kTileThreads) / kParts;
            s_rb[pos] = rb[h];
            s_cnt[pos] = rc[h];
            s_ryz[pos] = (y0 + r % ny) | ((z0 + r / ny) << 16);
            s_rx[pos] = rx[h];
        }
    }
    __syncthreads();
    {
        // s_pref: exclusive prefix of the rows' point counts, thread t
        // owning rows kRowIters·t .. kRowIters·t + kRowIters − 1
        int loc[kRowIters], sum = 0;
#pragma unroll
        for (int i = 0; i < kRowIters; ++i) {
            const int r = threadIdx.x * kRowIters + i;
            loc[i] = r < nrows ? s_cnt[r] : 0;
            sum += loc[i];
        }
        int run = block_exclusive_sum<kTileThreads>(sum, s_wsum);
#pragma unroll
        for (int i = 0; i < kRowIters; ++i) {
            const int r = threadIdx.x * kRowIters + i;
            if (r <= nrows) s_pref[r] = run;
            run += loc[i];
        }
        if (threadIdx.x == kTileThreads - 1) s_pref[nrows] = run;  // the total, also when nrows fills every slot
    }
    __syncthreads();
    const int total = s_pref[nrows];
    // faces of the staged box that are not on the grid boundary
    const float flx = x0 > 0 ? G.lo[0] + float(x0) * G.cell : -INFINITY;
    const float fhx = x1 + 1 < G.dim[0] ? G.lo[0] + float(x1 + 1) * G.cell : INFINITY;
    const float fly = y0 > 0 ? G.lo[1] + float(y0) * G.cell : -INFINITY;
    const float fhy = y1 + 1 < G.dim[1] ? G.lo[1] + float(y1 + 1) * G.cell : INFINITY;
    const float flz = z0 > 0 ? G.lo[2] + float(z0) * G.cell : -INFINITY;
    const float fhz = z1 + 1 < G.dim[2] ? G.lo[2] + float(z1 + 1) * G.cell : INFINITY;
    const int64_t qs = tile_start[int64_t(tile) * T.sub], qe = tile_start[int64_t(tile + 1) * T.sub];
    for (int64_t qb = qs; qb < qe; qb += kTileThreads) {
        const int64_t t = qb + threadIdx.x;
        const bool active = t < qe;
        const int qi = active ? order[t] : 0;
        const float qx = active ? q[3 * int64_t(qi)] : 0.f, qy = active ? q[3 * int64_t(qi) + 1] : 0.f,
                    qz = active ? q[3 * int64_t(qi) + 2] : 0.f;
        uint64_t best[K];
        knn_reset<K>(best, k);
        float kd2 = INFINITY;  // float d² of the k-th key: a cheap first test
        for (int cb = 0; cb < total; cb += kStage) {
            const int cnt = min(kStage, total - cb);
            __syncthreads();
            for (int j = threadIdx.x; j < cnt; j += kTileThreads) {
                const int g = cb + j;
                int lo = 0, hi = nrows - 1;  // last row with s_pref[row] <= g
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_pref[mid] <= g) lo = mid;
                    else hi = mid - 1;
                }
                float4 p = pts[s_rb[lo] + (g - s_pref[lo])];
                if (skip && skip[__float_as_int(p.w)]) p.w = __int_as_float(-1);
                s_pts[j] = p;
            }
            __syncthreads();
            auto candidate = [&](const float4 p) {
                const float d2 = dist2(qx - p.x, qy - p.y, qz - p.z);
                if (!(d2 <= kd2)) return;
                const int idx = __float_as_int(p.w);
                if (idx < 0) return;
                const uint64_t key = (uint64_t(__float_as_uint(d2)) << 32) | uint32_t(idx);
                if (key < best[0]) {
                    knn_insert<K>(best, key);
                    kd2 = best[0] == kNoKey ? INFINITY : __uint_as_float(uint32_t(best[0] >> 32));
                }
            };
            if (active && HALO == kHalo1) {
                // the first pass's small region is nearly all needed
                for (int j = 0; j < cnt; ++j) candidate(s_pts[j]);
            } else if (active) {
                // row by row in ring order: a row whose box of cells is
                // farther than the k-th distance of every lane of the wave
                // is skipped whole
                int r = 0;
                for (int hi = nrows - 1; r < hi;) {  // the row holding point cb
                    const int mid = (r + hi + 1) >> 1;
                    if (s_pref[mid] <= cb) r = mid;
                    else hi = mid - 1;
                }
                for (; r < nrows; ++r) {
                    const int a = max(s_pref[r], cb), b = min(s_pref[r + 1], cb + cnt);
                    if (a >= cb + cnt) break;
                    if (a >= b) continue;
                    const int yz = s_ryz[r], xr = s_rx[r];
                    const float rxlo = __builtin_fmaf(float(xr & 0xFFFF), G.cell, G.lo[0]);
                    const float rxhi = __builtin_fmaf(float((xr >> 16) + 1), G.cell, G.lo[0]);
                    const float ylo = __builtin_fmaf(float(yz & 0xFFFF), G.cell, G.lo[1]);
                    const float zlo = __builtin_fmaf(float(yz >> 16), G.cell, G.lo[2]);
                    const float gy = fmaxf(fmaxf(ylo - qy, qy - (ylo + G.cell)) - T.margin, 0.f);
                    const float gz = fmaxf(fmaxf(zlo - qz, qz - (zlo + G.cell)) - T.margin, 0.f);
                    const float gx = fmaxf(fmaxf(rxlo - qx, qx - rxhi) - T.margin, 0.f);
                    if (!__builtin_amdgcn_ballot_w64(dist2(gx, gy, gz) <= kd2)) continue;
                    for (int j = a - cb; j < b - cb; ++j) candidate(s_pts[j]);
                }
            }
        }
        if (!active) continue;
        const float dmin = fminf(fminf(fminf(qx - flx, fhx - qx), fminf(qy - fly, fhy - qy)), fminf(qz - flz, fhz - qz));
        bool exact = dmin == INFINITY;
        if (!exact && best[0] != kNoKey) {
            const float safe = dmin * (1.0f - 1e-5f) - 1e-6f * G.cell;
            exact = safe > 0.f && kd2 < safe * safe;
        }
        if (!exact && redo_flag) {
            // ordered: the query keeps its place in tile order (compacted
            // afterwards), so the far walk's waves get spatially coherent runs
            redo_flag[t] = 1;
            redo[t] = qi;
            if (redo_kth) redo_kth[t] = kd2;
            continue;
        }
        if (!exact) {
            // the wave's lanes that give up here take consecutive slots from
            // one atomic; the k candidates a lane has bound its k-th distance
            const uint64_t m = __ballot(1);
            const int lane = threadIdx.x & 63, leader = __builtin_ctzll(m);
            unsigned int base = 0;
            if (lane == leader) base = atomicAdd(n_redo, (unsigned int)__builtin_popcountll(m));
            base = unsigned(__shfl(int(base), leader, 64));
            const unsigned int slot = base + unsigned(__builtin_popcountll(m & ((1ull << lane) - 1ull)));
            redo[slot] = qi;
            if (redo_kth) redo_kth[slot] = kd2;
            continue;
        }
        knn_store<K>(best, k, qi, out_idx, out_d2);
        if (SCATTER) {
            long long v[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] = __double2ll_rn(double(sc.val[3 * int64_t(qi) + d]) * kScatterFix);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (j >= k || best[j] == kNoKey) continue;
                const int idx = int(uint32_t(best[j]));
                int h = int((uint32_t(idx) * 0x9E3779B1u) >> 24);  // 8 bits
                bool done = false;
                for (int probe = 0; probe < 8 && !done; ++probe) {
                    const int prev = atomicCAS(&s_hkey[h], -1, idx);
                    if (prev == -1 || prev == idx) {
#pragma unroll
                        for (int d = 0; d < 3; ++d) atomicAdd(&s_hacc[h][d], (unsigned long long)v[d]);
                        done = true;
                    }
                    h = (h + 1) & (kHash - 1);
                }
                if (!done)
#pragma unroll
                    for (int d = 0; d < 3; ++d) atomicAdd(&sc.acc[3 * int64_t(idx) + d], (unsigned long long)v[d]);
            }
        }
    }
    if (SCATTER) {
        __syncthreads();
        for (int h = threadIdx.x; h < kHash; h += kTileThreads) {
            const int key = s_hkey[h];
            if (key >= 0)
#pragma unroll
                for (int d = 0; d < 3; ++d) atomicAdd(&sc.acc[3 * int64_t(key) + d], s_hacc[h][d]);
        }
    }
}

// acc[neighbour] += value of the query, for the queries of qlist (or all)
__global__ void scatter_pairs_kernel(const int32_t* __restrict__ knn, int k, const int* __restrict__ qlist, int64_t nq,
                                     KnnScatter sc) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nq * k) return;
    const int64_t qi = qlist ? int64_t(qlist[t / k]) : t / k;
    const int32_t s = knn[qi * k + t % k];
    if (s < 0) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long long v = __double2ll_rn(double(sc.val[3 * qi + d]) * kScatterFix);
        atomicAdd(&sc.acc[3 * int64_t(s) + d], (unsigned long long)v);
    }
}

// The same for 256 consecutive queries of qlist per block (qlist is in tile
// order, so they are neighbours and share most of their neighbours): the
// pairs are summed in an LDS hash keyed by point index and each key is
// flushed once with 64-bit global atomics (integer sums: bit for bit the
// per-pair result).
__global__ __launch_bounds__(256) void scatter_pairs_hashed_kernel(const int32_t* __restrict__ knn, int k,
                                                                   const int* __restrict__ qlist, int64_t nq,
                                                                   KnnScatter sc) {
    constexpr int kHash = 1024;
    __shared__ int s_hkey[kHash];
    __shared__ unsigned long long s_hacc[kHash][3];
    for (int h = threadIdx.x; h < kHash; h += 256) {
        s_hkey[h] = -1;
        s_hacc[h][0] = s_hacc[h][1] = s_hacc[h][2] = 0ull;
    }
    __syncthreads();
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t < nq) {
        const int64_t qi = qlist ? int64_t(qlist[t]) : t;
        long long v[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] = __double2ll_rn(double(sc.val[3 * qi + d]) * kScatterFix);
        for (int j = 0; j < k; ++j) {
            const int32_t idx = knn[qi * k + j];
            if (idx < 0) continue;
            int h = int((uint32_t(idx) * 0x9E3779B1u) >> 22);  // 10 bits
            bool done = false;
            for (int probe = 0; probe < 8 && !done; ++probe) {
                const int prev = atomicCAS(&s_hkey[h], -1, idx);
                if (prev == -1 || prev == idx) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) atomicAdd(&s_hacc[h][d], (unsigned long long)v[d]);
                    done = true;
                }
                h = (h + 1) & (kHash - 1);
            }
            if (!done)
#pragma unroll
                for (int d = 0; d < 3; ++d) atomicAdd(&sc.acc[3 * int64_t(idx) + d], (unsigned long long)v[d]);
        }
    }
    __syncthreads();
    for (int h = threadIdx.x; h < kHash; h += 256) {
        const int key = s_hkey[h];
        if (key >= 0)
#pragma unroll
            for (int d = 0; d < 3; ++d) atomicAdd(&sc.acc[3 * int64_t(key) + d], s_hacc[h][d]);
    }
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const uint32_t lo = uint32_t(__shfl(int(uint32_t(v)), src, 64));
    const uint32_t hi = uint32_t(__shfl(int(uint32_t(v >> 32)), src, 64));
    return (uint64_t(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = shfl64(v, (threadIdx.x & 63) ^ off);
        v = o < v ? o : v;
    }
    return v;
}

// One wave per query, for the queries the tiled passes could not close (far
// from the points, or inside a cluster of skipped points).  On the coarse
// level the cells of each Chebyshev shell are read 64 at a time, their
// points flattened and scanned 64 per step (one candidate per lane), each
// lane keeping its own k best keys.  The smallest lane-local k-th key bounds
// the wave's k-th key from above, so the search stops once it is closer than
// every cell not yet visited; k rounds of wave minima then merge the lists
// into the exact k smallest (d², index) keys.
// Cell offset of entry id of the Chebyshev shell of radius r, which has
// 24r² + 2 cells (one for r = 0): the two z faces, then the two y faces
// without their z rims, then the two x faces without either rim.
__device__ __forceinline__ int3 shell_cell(int id, int r) {
    if (r == 0) return make_int3(0, 0, 0);
    const int s = 2 * r + 1, t = 2 * r - 1, nz = s * s, ny = s * t;
    if (id < 2 * nz) {
        const int f = id / nz, rem = id - f * nz;
        return make_int3(rem % s - r, rem / s - r, f ? r : -r);
    }
    id -= 2 * nz;
    if (id < 2 * ny) {
        const int f = id / ny, rem = id - f * ny;
        return make_int3(rem % s - r, f ? r : -r, rem / s - (r - 1));
    }
    id -= 2 * ny;
    const int tt = t * t, f = id / tt, rem = id - f * tt;
    return make_int3(f ? r : -r, rem % t - (r - 1), rem / t - (r - 1));
}

template <int K>
__global__ __launch_bounds__(256) void knn_wave_kernel(KnnGridArgs G, const int64_t* __restrict__ cell_start,
                                                       const float4* __restrict__ pts, const int* __restrict__ empty,
                                                       const float* __restrict__ q, int64_t nq,
                                                       const int* __restrict__ qlist, int k,
                                                       const uint8_t* __restrict__ skip, int32_t* __restrict__ out_idx,
                                                       float* __restrict__ out_d2) {
    const int lane = threadIdx.x & 63;
    // wave w takes query w/2 of the list if w is even, else the (w/2)-th
    // from the end: the long searches, which come in runs along the list,
    // start early from either end (zero fill 0.61 → 0.57 ms per iteration;
    // the tiled passes lose L2 reuse that way and keep their order)
    const int64_t w = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= nq) return;  // uniform per wave
    const int64_t t = (w & 1) ? nq - 1 - (w >> 1) : (w >> 1);
    const int64_t i = qlist ? int64_t(qlist[t]) : t;
    const float qx = q[3 * i], qy = q[3 * i + 1], qz = q[3 * i + 2];
    uint64_t best[K];
    knn_reset<K>(best, k);
    const int c0x = cell_coord(qx, G.lo[0], G.inv_cell, G.dim[0]);
    const int c0y = cell_coord(qy, G.lo[1], G.inv_cell, G.dim[1]);
    const int c0z = cell_coord(qz, G.lo[2], G.inv_cell, G.dim[2]);
    const int maxr = max(G.dim[0], max(G.dim[1], G.dim[2]));
    int r = empty[c0x + int64_t(G.dim[0]) * (c0y + int64_t(G.dim[1]) * c0z)];
    for (; r <= maxr; ++r) {
        // only the shell's own cells: O(r²) per shell
        const int shell = r == 0 ? 1 : 24 * r * r + 2;
        for (int base = 0; base < shell; base += 64) {
            const int id = base + lane;
            int64_t b = 0;
            int cnt = 0;
            if (id < shell) {
                const int3 o = shell_cell(id, r);
                const int x = c0x + o.x, y = c0y + o.y, z = c0z + o.z;
                if (x >= 0 && y >= 0 && z >= 0 && x < G.dim[0] && y < G.dim[1] && z < G.dim[2]) {
                    const int64_t c = x + int64_t(G.dim[0]) * (y + int64_t(G.dim[1]) * z);
                    b = cell_start[c];
                    cnt = int(cell_start[c + 1] - b);
                }
            }
            // inclusive scan of the 64 cell sizes
            int incl = cnt;
            for (int off = 1; off < 64; off <<= 1) {
                const int v = __shfl_up(incl, off, 64);
                if (lane >= off) incl += v;
            }
            const int total = __shfl(incl, 63, 64);
            for (int cb = 0; cb < total; cb += 64) {
                const int j = cb + lane;
                // owner lane: the number of lanes whose inclusive count is ≤ j
                // (six uniform steps, every lane active for the shuffles)
                int lo = 0;
                for (int step = 32; step > 0; step >>= 1)
                    if (__shfl(incl, lo + step - 1, 64) <= j) lo += step;
                lo = min(lo, 63);
                const int64_t ob = shfl64(uint64_t(b), lo);
                const int oexcl = __shfl(incl - cnt, lo, 64);
                if (j < total) {
                    const float4 p = pts[ob + (j - oexcl)];
                    const int idx = __float_as_int(p.w);
                    if (!(skip && skip[idx])) {
                        const float d2 = dist2(qx - p.x, qy - p.y, qz - p.z);
                        const uint64_t key = (uint64_t(__float_as_uint(d2)) << 32) | uint32_t(idx);
                        if (key < best[0]) knn_insert<K>(best, key);
                    }
                }
            }
        }
        // no unvisited point is closer than the nearest face of the searched
        // block that is not on the grid boundary
        float dmin = INFINITY;
        if (c0x - r > 0) dmin = fminf(dmin, qx - (G.lo[0] + float(c0x - r) * G.cell));
        if (c0x + r + 1 < G.dim[0]) dmin = fminf(dmin, (G.lo[0] + float(c0x + r + 1) * G.cell) - qx);
        if (c0y - r > 0) dmin = fminf(dmin, qy - (G.lo[1] + float(c0y - r) * G.cell));
        if (c0y + r + 1 < G.dim[1]) dmin = fminf(dmin, (G.lo[1] + float(c0y + r + 1) * G.cell) - qy);
        if (c0z - r > 0) dmin = fminf(dmin, qz - (G.lo[2] + float(c0z - r) * G.cell));
        if (c0z + r + 1 < G.dim[2]) dmin = fminf(dmin, (G.lo[2] + float(c0z + r + 1) * G.cell) - qz);
        if (dmin == INFINITY) break;
        const uint64_t bound = wave_min_u64(best[0]);
        if (bound != kNoKey) {
            const float kd2 = __uint_as_float(uint32_t(bound >> 32));
            const float safe = dmin * (1.0f - 1e-5f) - 1e-6f * G.cell;
            if (safe > 0.f && kd2 < safe * safe) break;
        }
    }
    // merge: reverse each list to ascending order, drop the K−k zero slots
    // from its front, then the j-th smallest key is the wave minimum of the
    // lanes' heads (keys are distinct: one lane pops)
    uint64_t asc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) asc[j] = best[K - 1 - j];
    for (int s = 0; s < K - k; ++s) {
#pragma unroll
        for (int c = 0; c < K - 1; ++c) asc[c] = asc[c + 1];
        asc[K - 1] = kNoKey;
    }
    for (int j = 0; j < k; ++j) {
        const uint64_t m = wave_min_u64(asc[0]);
        const uint64_t win = __ballot(asc[0] == m && m != kNoKey);
        if (win && lane == __ffsll((unsigned long long)win) - 1) {
#pragma unroll
            for (int c = 0; c < K - 1; ++c) asc[c] = asc[c + 1];
            asc[K - 1] = kNoKey;
        }
        if (lane == 0) {
            out_idx[i * k + j] = m == kNoKey ? -1 : int32_t(uint32_t(m));
            if (out_d2) out_d2[i * k + j] = m == kNoKey ? INFINITY : __uint_as_float(uint32_t(m >> 32));
        }
    }
}

static void launch_wave(dwg_ctx* ctx, const dwg_knn_index* I, const float* q, int64_t nq, const int* qlist, int k,
                        const uint8_t* skip, int32_t* out_idx, float* out_d2) {
    if (nq == 0) return;
    dim3 grid(grid_for(nq, 4)), block(256);
    const auto& L = I->lvl[1];
#define DWG_KNNW(KK)                                                                                               \
    hipLaunchKernelGGL(knn_wave_kernel<KK>, grid, block, 0, ctx->stream, L.G, L.cell_start, L.pts, L.empty, q, nq, \
                       qlist, k, skip, out_idx, out_d2)
    if (k <= 8) DWG_KNNW(8);
    else if (k <= 10) DWG_KNNW(10);
    else if (k <= 16) DWG_KNNW(16);
    else if (k <= 17) DWG_KNNW(17);  // the learned start's 16-point patches with the point itself
    else DWG_KNNW(32);
#undef DWG_KNNW
    check_launch();
}

// Exact search over the occupancy pyramid of the fine level, one query per
// lane: a depth-first walk from the top block that always descends into the
// nearest remaining child block and drops every block whose box is farther
// than the current k-th key.  A query far from every point reaches its
// neighbours in a few hundred short steps, where a shell search visits
// every cell of the shells in between.  The stack holds one entry per level
// in LDS: block coordinates and the mask of children not yet visited.
struct PyrArgs {
    const uint8_t* mask;
    const int4* dim;
    const int64_t* off;
    int levels;
    float margin;
    const float4* box;  // the points' bounding boxes (dwg_knn_index::Level::pbox)
    int64_t ncells;
};

// squared distance from q to a box (lo, hi), computed like a point's d²
// (dist2): every step is monotone, so it never exceeds the d² of a point
// inside the box
__device__ __forceinline__ float tight_box_d2(const float4 lo, const float4 hi, float qx, float qy, float qz) {
    const float gx = fmaxf(fmaxf(lo.x - qx, qx - hi.x), 0.f);
    const float gy = fmaxf(fmaxf(lo.y - qy, qy - hi.y), 0.f);
    const float gz = fmaxf(fmaxf(lo.z - qz, qz - hi.z), 0.f);
    return dist2(gx, gy, gz);
}

typedef float KnnFloat4 __attribute__((ext_vector_type(4)));
// a wave-uniform box through the scalar data cache
__device__ __forceinline__ float tight_box_d2_uniform(const float4* box, int64_t i, float qx, float qy, float qz) {
    const auto* b = (const __attribute__((address_space(4))) KnnFloat4*)box;
    const KnnFloat4 lo = b[2 * i], hi = b[2 * i + 1];
    return tight_box_d2(make_float4(lo.x, lo.y, lo.z, 0.f), make_float4(hi.x, hi.y, hi.z, 0.f), qx, qy, qz);
}

__device__ __forceinline__ uint64_t pyr_entry(int x, int y, int z, uint32_t m) {
    return uint64_t(uint32_t(x)) | (uint64_t(uint32_t(y)) << 16) | (uint64_t(uint32_t(z)) << 32) | (uint64_t(m) << 48);
}

// squared distance from q to the box of cells [b·2^l, (b+1)·2^l) along each
// axis, each axis gap shrunk by the rounding margin
__device__ __forceinline__ float pyr_box_d2(const KnnGridArgs& G, float w, float margin, float qx, float qy, float qz,
                                            int x, int y, int z) {
    auto gap = [&](float v, float lo, int b) {
        const float a = __builtin_fmaf(float(b), w, lo);
        return fmaxf(fmaxf(a - v, v - (a + w)) - margin, 0.f);
    };
    const float gx = gap(qx, G.lo[0], x), gy = gap(qy, G.lo[1], y), gz = gap(qz, G.lo[2], z);
    return __builtin_fmaf(gz, gz, __builtin_fmaf(gy, gy, gx * gx));
}

constexpr int kPyrThreads = 64;  // one wave per block: a block frees its slot when its slowest walk ends

template <int K>
__global__ __launch_bounds__(kPyrThreads) void knn_pyramid_kernel(KnnGridArgs G, PyrArgs Y,
                                                                  const int64_t* __restrict__ cell_start,
                                                                  const float4* __restrict__ pts,
                                                                  const float* __restrict__ q, int64_t nq,
                                                                  const int* __restrict__ qlist, int k,
                                                                  const uint8_t* __restrict__ skip,
                                                                  const float* __restrict__ init_kth,
                                                                  int32_t* __restrict__ out_idx,
                                                                  float* __restrict__ out_d2) {
    extern __shared__ uint64_t stk[];  // Y.levels × kPyrThreads entries
    __shared__ int4 s_dim[kMaxPyr];
    __shared__ int64_t s_off[kMaxPyr];
    if (threadIdx.x < Y.levels) {
        s_dim[threadIdx.x] = Y.dim[threadIdx.x];
        s_off[threadIdx.x] = Y.off[threadIdx.x];
    }
    __syncthreads();
    const int64_t t = int64_t(blockIdx.x) * kPyrThreads + threadIdx.x;
    if (t >= nq) return;
    const int64_t i = qlist ? int64_t(qlist[t]) : t;
    const float qx = q[3 * i], qy = q[3 * i + 1], qz = q[3 * i + 2];
    uint64_t best[K];
    knn_reset<K>(best, k);
    // an upper bound of the k-th d² known beforehand (k points at most that
    // far exist): blocks beyond it are never opened.  A point at exactly the
    // bound has a box distance ≤ it, so ties are still seen.
    const float cap = init_kth ? init_kth[t] : INFINITY;
    uint64_t* st = stk + threadIdx.x;
    const int top = Y.levels - 1;
    st[top * kPyrThreads] = pyr_entry(0, 0, 0, Y.mask[s_off[top]]);
    int l = top;
    while (l <= top) {
        const uint64_t e = st[l * kPyrThreads];
        const int bx = int(e & 0xFFFF), by = int((e >> 16) & 0xFFFF), bz = int((e >> 32) & 0xFFFF);
        uint32_t m = uint32_t(e >> 48);
        // the nearest child block not yet visited
        const float w = G.cell * float(1 << (l - 1));
        float bd = INFINITY;
        int bc = -1;
        for (uint32_t r = m; r; r &= r - 1) {
            const int c = __builtin_ctz(r);
            const float d = pyr_box_d2(G, w, Y.margin, qx, qy, qz, 2 * bx + (c & 1), 2 * by + ((c >> 1) & 1),
                                       2 * bz + (c >> 2));
            if (d < bd) {
                bd = d;
                bc = c;
            }
        }
        const float kth = best[0] == kNoKey ? cap : fminf(cap, __uint_as_float(uint32_t(best[0] >> 32)));
        if (bc < 0 || !(bd <= kth)) {  // nothing left here can hold a closer point
            ++l;
            continue;
        }
        m &= ~(1u << bc);
        st[l * kPyrThreads] = pyr_entry(bx, by, bz, m);
        const int x = 2 * bx + (bc & 1), y = 2 * by + ((bc >> 1) & 1), z = 2 * bz + (bc >> 2);
        if (l == 1) {  // a cell: its points
            const int64_t c = x + int64_t(G.dim[0]) * (y + int64_t(G.dim[1]) * z);
            const int64_t se = cell_start[c + 1];
            for (int64_t s = cell_start[c]; s < se; ++s) {
                const float4 p = pts[s];
                const int idx = __float_as_int(p.w);
                const float d2 = dist2(qx - p.x, qy - p.y, qz - p.z);
                const uint64_t key = (uint64_t(__float_as_uint(d2)) << 32) | uint32_t(idx);
                // the skip byte only for a point that would enter the list
                if (key < best[0] && !(skip && skip[idx])) knn_insert<K>(best, key);
            }
        } else {
            const int4 cd = s_dim[l - 1];
            --l;
            st[l * kPyrThreads] = pyr_entry(x, y, z, Y.mask[s_off[l] + x + int64_t(cd.x) * (y + int64_t(cd.y) * z)]);
        }
    }
    knn_store<K>(best, k, i, out_idx, out_d2);
}

// The same pyramid walk shared by a wave: 64 neighbouring queries (the far
// queries arrive in tile order) walk the union of their searches with one
// wave-uniform stack of (block, lanes that need it) in LDS, the way the
// Barnes-Hut walk shares its tree.  Every block and every point is then one
// load for the wave instead of 64 scattered ones, and the lanes stay in
// step.  Children are taken nearest-first for the wave's first active lane;
// each lane still prunes with its own k-th key, so its result is exactly
// its own k nearest points.  A child's distance is that of the bounding box
// of its points, not of its cells: a far query over a surface patch then
// opens the few cells near its foot point instead of every cell whose box
// comes within a cell width of its k-th distance.
//
// The walk is a chain of dependent loads, so each step issues its loads
// together: opening a block fetches its 8 children's boxes and their own
// occupancy bytes (or, for cells, their point ranges) in one vector load
// over a few lanes; a stack entry carries its block's occupancy byte; a
// cell's points come in one vector load (64 at a time) staged in LDS.
constexpr int kPwWaves = 2;  // small blocks free their slots sooner: the walks end at very different times
constexpr int kPwStack = 128;  // ≥ 7 · kMaxPyr + 1 entries

template <int K>
__global__ __launch_bounds__(kPwWaves * 64) void knn_pyramid_wave_kernel(KnnGridArgs G, PyrArgs Y,
                                                                        const int64_t* __restrict__ cell_start,
                                                                        const float4* __restrict__ pts,
                                                                        const float* __restrict__ q, int64_t nq,
                                                                        const int* __restrict__ qlist, int k,
                                                                        const uint8_t* __restrict__ skip,
                                                                        const float* __restrict__ init_kth,
                                                                        int32_t* __restrict__ out_idx,
                                                                        float* __restrict__ out_d2) {
    // stack entry: (x | y << 16, z | level << 16 | occupancy << 20, lanes lo, lanes hi)
    __shared__ int4 s_stk[kPwWaves][kPwStack];
    __shared__ float4 s_box[kPwWaves][16];    // box of child c: lo at 2c, hi at 2c + 1
    __shared__ int64_t s_rng[kPwWaves][16];   // level 1: point range of child cell c at 2c, 2c + 1
    __shared__ int s_cocc[kPwWaves][16];      // level ≥ 2: occupancy byte of child c at 2c
    __shared__ float4 s_pts[kPwWaves][64];    // points of the cell being scanned
    __shared__ int4 s_dim[kMaxPyr];
    __shared__ int64_t s_off[kMaxPyr];
    if (threadIdx.x < Y.levels) {
        s_dim[threadIdx.x] = Y.dim[threadIdx.x];
        s_off[threadIdx.x] = Y.off[threadIdx.x];
    }
    __syncthreads();
    // waves take the 64-query chunks alternately from the front and the back
    // of the list (wave w: chunk w/2 if w is even, else the (w/2)-th from
    // the end).  The longest walks are not known before they run and come
    // in runs along the list; started from both ends, fewer of them start
    // late and end after everything else (far walk 1.65 → 1.54 ms per
    // learned iteration)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t nchunks = (nq + 63) / 64, w = int64_t(blockIdx.x) * kPwWaves + wave;
    const int64_t chunk = (w & 1) ? nchunks - 1 - (w >> 1) : (w >> 1);
    const int64_t t = chunk * 64 + lane;
    const bool valid = w < nchunks && t < nq;
    const int64_t i = valid ? (qlist ? int64_t(qlist[t]) : t) : 0;
    const float qx = valid ? q[3 * i] : 0.f, qy = valid ? q[3 * i + 1] : 0.f, qz = valid ? q[3 * i + 2] : 0.f;
    uint64_t best[K];
    knn_reset<K>(best, k);
    const float cap = (valid && init_kth) ? init_kth[t] : INFINITY;
    const uint64_t all = __builtin_amdgcn_ballot_w64(valid);
    if (all == 0) return;
    auto kth_of = [&]() { return best[0] == kNoKey ? cap : fminf(cap, __uint_as_float(uint32_t(best[0] >> 32))); };
    int4* stk = s_stk[wave];
    float4* box = s_box[wave];
    const int top = Y.levels - 1;
    const int occ_top = uniform(int(Y.mask[s_off[top]]));
    // every lane writes the same entries (no divergent branch: the stack
    // pointer stays in a scalar register)
    stk[0] = make_int4(0, (top << 16) | (occ_top << 20), int(uint32_t(all)), int(uint32_t(all >> 32)));
    __builtin_amdgcn_wave_barrier();
    int sp = 1;
    while (sp > 0) {
        --sp;
        const int4 e = stk[sp];
        const uint32_t e0 = uint32_t(uniform(e.x)), e1 = uint32_t(uniform(e.y));
        const int bx = int(e0 & 0xFFFFu), by = int(e0 >> 16), bz = int(e1 & 0xFFFFu), l = int((e1 >> 16) & 0xFu);
        const uint32_t occ = e1 >> 20;
        const uint64_t mask = (uint64_t(uint32_t(uniform(e.w))) << 32) | uint32_t(uniform(e.z));
        const bool act = __builtin_amdgcn_inverse_ballot_w64(mask);
        // the children's loads, together: lanes 2c and 2c + 1 fetch child
        // c's box (lo, hi); lanes 16 + 2c and 17 + 2c its point range (a
        // cell) or lane 16 + 2c its occupancy byte (a block)
        {
            const int c = (lane >> 1) & 7;
            if ((occ >> c) & 1u) {
                const int4 cdm = s_dim[l - 1];
                const int x = 2 * bx + (c & 1), y = 2 * by + ((c >> 1) & 1), z = 2 * bz + (c >> 2);
                const int64_t ci = x + int64_t(cdm.x) * (y + int64_t(cdm.y) * z);
                if (lane < 16) {
                    box[lane] = Y.box[2 * ((l == 1 ? 0 : Y.ncells + s_off[l - 1]) + ci) + (lane & 1)];
                } else if (lane < 32) {
                    if (l == 1) s_rng[wave][lane - 16] = cell_start[ci + (lane & 1)];
                    else if (!(lane & 1)) s_cocc[wave][lane - 16] = int(Y.mask[s_off[l - 1] + ci]);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        const float kth = kth_of();
        // the children that some lane still needs, keyed by the first active
        // lane's box distance (float bits, positive) with the child number in
        // the low 3 bits, sorted on the scalar unit
        const int ref = __builtin_ctzll(mask);
        uint32_t key[8];
        uint64_t need[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            need[c] = 0;
            key[c] = 0xFFFFFFFFu;
            if ((occ >> c) & 1u) {
                const float d = tight_box_d2(box[2 * c], box[2 * c + 1], qx, qy, qz);
                need[c] = __builtin_amdgcn_ballot_w64(act && d <= kth);
                if (need[c]) {
                    const uint32_t dr = uint32_t(__builtin_amdgcn_readlane(int(__float_as_uint(d)), ref));
                    key[c] = (dr & ~7u) | uint32_t(c);
                }
            }
        }
        // sorting network on 8 scalar keys (19 compare-exchanges)
        auto cx = [&](int a, int b) {
            const uint32_t lo = min(key[a], key[b]), hi = max(key[a], key[b]);
            key[a] = lo;
            key[b] = hi;
        };
        cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
        cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
        cx(1, 2); cx(5, 6); cx(0, 4); cx(3, 7);
        cx(1, 5); cx(2, 6);
        cx(1, 4); cx(3, 6);
        cx(2, 4); cx(3, 5);
        cx(3, 4);
        if (l == 1) {  // children are cells: scan their points, nearest first
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (key[j] == 0xFFFFFFFFu) break;
                const int c = int(key[j] & 7u);
                // the lanes that still need this cell with their k-th key now
                const float d = tight_box_d2(box[2 * c], box[2 * c + 1], qx, qy, qz);
                const bool use = act && d <= kth_of();
                if (!__builtin_amdgcn_ballot_w64(use)) continue;
                const int64_t cb = uniform64(uint64_t(s_rng[wave][2 * c])), ce = uniform64(uint64_t(s_rng[wave][2 * c + 1]));
                for (int64_t s0 = cb; s0 < ce; s0 += 64) {
                    const int cnt = int(min(ce - s0, int64_t(64)));
                    if (lane < cnt) {
                        float4 p = pts[s0 + lane];
                        if (skip && skip[__float_as_int(p.w)]) p.w = __int_as_float(-1);
                        s_pts[wave][lane] = p;
                    }
                    __builtin_amdgcn_wave_barrier();
                    for (int jj = 0; jj < cnt; ++jj) {
                        const float4 p = s_pts[wave][jj];
                        const int idx = __float_as_int(p.w);
                        if (idx < 0) continue;
                        const float d2 = dist2(qx - p.x, qy - p.y, qz - p.z);
                        const uint64_t kk = (uint64_t(__float_as_uint(d2)) << 32) | uint32_t(idx);
                        if (use && kk < best[0]) knn_insert<K>(best, kk);
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
        } else {  // push the needed children, farthest first so the nearest pops first
#pragma unroll
            for (int j = 7; j >= 0; --j) {
                if (key[j] == 0xFFFFFFFFu) continue;
                const int c = int(key[j] & 7u);
                uint64_t m = 0;
                int co = 0;
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) {
                    m = cc == c ? need[cc] : m;
                }
                co = uniform(s_cocc[wave][2 * c]);
                const int x = 2 * bx + (c & 1), y = 2 * by + ((c >> 1) & 1), z = 2 * bz + (c >> 2);
                stk[sp] = make_int4(x | (y << 16), z | ((l - 1) << 16) | (co << 20), int(uint32_t(m)),
                                    int(uint32_t(m >> 32)));
                ++sp;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (valid) knn_store<K>(best, k, i, out_idx, out_d2);
}

// far queries: the pyramid walk when there are enough of them to fill the
// lanes, else (or without a pyramid) one wave per query on the coarse level
constexpr int64_t kPyrMinQueries = 2048;
// below this many queries (and fewer than one per 8 indexed points) a query
// set skips the tiled passes for the pyramid walk
constexpr int64_t kSparseMaxQueries = 65536;

static void launch_far(dwg_ctx* ctx, const dwg_knn_index* I, const float* q, int64_t nq, const int* qlist, int k,
                       const uint8_t* skip, int32_t* out_idx, float* out_d2, const float* init_kth = nullptr,
                       bool shared = false, const uint8_t* pyr_mask = nullptr) {
    if (nq == 0) return;
    const auto& L = I->lvl[0];
    if (L.pyr_levels < 2 || nq < kPyrMinQueries) {
        launch_wave(ctx, I, q, nq, qlist, k, skip, out_idx, out_d2);
        return;
    }
    // the boxes are a cache of the index's own points: building them
    // leaves the index unchanged
    if (shared) ensure_boxes(ctx, const_cast<dwg_knn_index::Level&>(L));
    const PyrArgs Y{pyr_mask ? pyr_mask : L.pyr, L.pyr_dim, L.pyr_off, L.pyr_levels, L.pyr_margin, L.pbox, L.ncells};
    if (shared) {
        DWG_REQUIRE(L.pyr_levels <= kMaxPyr && 7 * L.pyr_levels + 1 <= kPwStack, "pyramid deeper than the wave stack");
        dim3 grid(grid_for(nq, kPwWaves * 64)), block(kPwWaves * 64);
#define DWG_KNNPW(KK)                                                                                              \
    hipLaunchKernelGGL(knn_pyramid_wave_kernel<KK>, grid, block, 0, ctx->stream, L.G, Y, L.cell_start, L.pts, q, nq, \
                       qlist, k, skip, init_kth, out_idx, out_d2)
        if (k <= 8) DWG_KNNPW(8);
        else if (k <= 10) DWG_KNNPW(10);
        else if (k == 15) DWG_KNNPW(15);  // the diffusion's 15 neighbours: no spare slot to sink through
        else if (k <= 16) DWG_KNNPW(16);
        else if (k <= 17) DWG_KNNPW(17);  // the learned start's 16-point patches with the point itself
        else DWG_KNNPW(32);
#undef DWG_KNNPW
        check_launch();
        return;
    }
    dim3 grid(grid_for(nq, kPyrThreads)), block(kPyrThreads);
    const size_t lds = sizeof(uint64_t) * size_t(L.pyr_levels) * kPyrThreads;
#define DWG_KNNP(KK)                                                                                             \
    hipLaunchKernelGGL(knn_pyramid_kernel<KK>, grid, block, lds, ctx->stream, L.G, Y, L.cell_start, L.pts, q, nq, \
                       qlist, k, skip, init_kth, out_idx, out_d2)
    if (k <= 8) DWG_KNNP(8);
    else if (k <= 10) DWG_KNNP(10);
    else if (k <= 16) DWG_KNNP(16);
    else if (k <= 17) DWG_KNNP(17);  // the learned start's 16-point patches with the point itself
    else DWG_KNNP(32);
#undef DWG_KNNP
    check_launch();
}

static void launch_per_lane(dwg_ctx* ctx, const dwg_knn_index* I, const float* q, int64_t nq, const int* qlist, int k,
                            const uint8_t* skip, const uint8_t* qmask, int32_t* out_idx, float* out_d2) {
    if (nq == 0) return;
    dim3 grid(grid_for(nq, 256)), block(256);
#define DWG_KNN(KK)                                                                                                  \
    hipLaunchKernelGGL(knn_query_kernel<KK>, grid, block, 0, ctx->stream, I->lvl[0].G, I->lvl[0].cell_start,        \
                       I->lvl[0].pts, I->lvl[1].G, I->lvl[1].cell_start, I->lvl[1].pts, I->lvl[1].empty, q, nq, qlist, \
                       k, skip, qmask,                                                                            \
                       out_idx, out_d2)
    if (k <= 8) DWG_KNN(8);
    else if (k <= 10) DWG_KNN(10);
    else if (k <= 16) DWG_KNN(16);
    else if (k <= 17) DWG_KNN(17);  // the learned start's 16-point patches with the point itself
    else DWG_KNN(32);
#undef DWG_KNN
    check_launch();
}

// The queries qlist[0..nq) (or 0..nq−1, masked by qmask) bucketed by tile
// of kTile³ fine cells (and by eighth of it when sub is 8): order[start[sub·t]
// ..start[sub·(t+1))) are tile t's queries, list[0..nlist) the tiles that
// have any
struct TileBuckets {
    TileArgs T;
    int64_t ntiles = 0, nlist = 0;
    int64_t* start = nullptr;
    int* order = nullptr;
    int* list = nullptr;
};

static TileBuckets tile_buckets(dwg_ctx* ctx, const dwg_knn_index* I, const float* q, int64_t nq, const int* qlist,
                                const uint8_t* qmask, int sub, bool want_list = true) {
    const KnnGridArgs& G = I->lvl[0].G;
    TileBuckets B;
    TileArgs& T = B.T;
    for (int d = 0; d < 3; ++d) T.tdim[d] = (G.dim[d] + kTile - 1) / kTile;
    T.margin = cell_margin(G);
    T.sub = sub;
    const int64_t ntiles = int64_t(T.tdim[0]) * T.tdim[1] * T.tdim[2], nb = ntiles * sub;
    DWG_REQUIRE(nb < (int64_t(1) << 31), "too many query buckets");
    B.ntiles = ntiles;
    int* tile_of = static_cast<int*>(ctx->buf("knnt_tile_of", sizeof(int) * nq));
    int64_t* start = static_cast<int64_t*>(ctx->buf("knnt_start", sizeof(int64_t) * (nb + 1)));
    unsigned long long* cursor = static_cast<unsigned long long*>(ctx->buf("knnt_cursor", sizeof(uint64_t) * nb));
    int* order = static_cast<int*>(ctx->buf("knnt_order", sizeof(int) * nq));
    int64_t* flags = static_cast<int64_t*>(ctx->buf("knnt_flags", sizeof(int64_t) * (ntiles + 1)));
    int* list = static_cast<int*>(ctx->buf("knnt_list", sizeof(int) * ntiles));
    fill_u32(ctx, start, 0, 2 * (nb + 1));
    hipLaunchKernelGGL(tile_count_kernel, dim3(grid_for(nq, 256)), dim3(256), 0, ctx->stream, q, nq, qlist, qmask, G, T,
                       tile_of, reinterpret_cast<unsigned long long*>(start));
    check_launch();
    exclusive_scan<int64_t>(ctx, start, nb);
    hipLaunchKernelGGL(copy_i64_kernel, dim3(grid_for(nb, 256)), dim3(256), 0, ctx->stream, start, cursor, nb);
    check_launch();
    hipLaunchKernelGGL(tile_scatter_kernel, dim3(grid_for(nq, 256)), dim3(256), 0, ctx->stream, tile_of, nq, qlist,
                       cursor, order);
    check_launch();
    B.start = start;
    B.order = order;
    if (!want_list) {  // the caller takes the queries in tile order only: no list, no host read
        B.nlist = 0;
        B.list = nullptr;
        return B;
    }
    hipLaunchKernelGGL(tile_flags_kernel, dim3(grid_for(ntiles + 1, 256)), dim3(256), 0, ctx->stream, start, ntiles,
                       sub, flags);
    check_launch();
    exclusive_scan<int64_t>(ctx, flags, ntiles);
    hipLaunchKernelGGL(tile_list_kernel, dim3(grid_for(ntiles, 256)), dim3(256), 0, ctx->stream, flags, ntiles, list);
    check_launch();
    B.nlist = read_scalar<int64_t>(ctx, flags + ntiles);
    B.list = list;
    return B;
}

// the queries a tiled pass leaves open, gathered in tile order
__global__ void redo_compact_kernel(const int64_t* __restrict__ pos, const int* __restrict__ tmp_q,
                                    const float* __restrict__ tmp_kth, int64_t n, int* __restrict__ redo,
                                    float* __restrict__ redo_kth) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n || pos[t + 1] == pos[t]) return;
    redo[pos[t]] = tmp_q[t];
    if (redo_kth) redo_kth[pos[t]] = tmp_kth[t];
}

// One tiled pass over the queries qlist[0..nq) (or 0..nq−1, masked by
// qmask); the queries it cannot close are listed in redo (with their k-th
// bound in redo_kth), in tile order when `ordered`, else in the order the
// waves end.  Returns their count.
template <int HALO>
static int64_t tiled_pass(dwg_ctx* ctx, const dwg_knn_index* I, const float* q, int64_t nq, const int* qlist,
                          const uint8_t* qmask, int k, const uint8_t* skip, int32_t* out_idx, float* out_d2,
                          int* redo, const std::string& name, const KnnScatter* sc = nullptr,
                          float* redo_kth = nullptr, bool ordered = false) {
    const KnnGridArgs& G = I->lvl[0].G;
    const TileBuckets B = tile_buckets(ctx, I, q, nq, qlist, qmask, kSubTiled);
    const TileArgs& T = B.T;
    const int64_t ntiles = B.ntiles, nlist = B.nlist;
    const int64_t* start = B.start;
    const int* order = B.order;
    const int* list = B.list;
    unsigned int* n_redo = static_cast<unsigned int*>(ctx->buf("knnt_nredo", 64));
    fill_u32(ctx, n_redo, 0, 1);
    if (ctx->profile) ctx->count(name + ":queries", read_scalar<int64_t>(ctx, start + ntiles * T.sub));
    // ordered: each query's place t in tile order gets its flag (and, when
    // open, its index and bound) in scratch arrays, compacted below
    int64_t* rflag = nullptr;
    int* tmp_q = nullptr;
    float* tmp_kth = nullptr;
    if (ordered) {
        rflag = static_cast<int64_t*>(ctx->buf("knnt_rflag", sizeof(int64_t) * (nq + 1)));
        tmp_q = static_cast<int*>(ctx->buf("knnt_rtmp", sizeof(int) * nq + 64));
        if (redo_kth) tmp_kth = static_cast<float*>(ctx->buf("knnt_rkth", sizeof(float) * nq + 64));
        fill_u32(ctx, rflag, 0, 2 * (nq + 1));
    }
    int* redo_out = ordered ? tmp_q : redo;
    float* kth_out = ordered ? tmp_kth : redo_kth;
    if (nlist > 0) {
        ProfileScope pt(ctx, name.c_str());
        const KnnScatter scv = sc ? *sc : KnnScatter{nullptr, nullptr};
#define DWG_KNNT(KK)                                                                                                   \
    do {                                                                                                               \
        if (sc)                                                                                                        \
            hipLaunchKernelGGL((knn_tile_kernel<KK, HALO, true>), dim3(unsigned(nlist)), dim3(kTileThreads), 0,        \
                               ctx->stream, G, T, I->lvl[0].cell_start, I->lvl[0].pts, list, start, order, q, k, skip, \
                               out_idx, out_d2, redo_out, kth_out, n_redo, rflag, scv);                                \
        else                                                                                                           \
            hipLaunchKernelGGL((knn_tile_kernel<KK, HALO, false>), dim3(unsigned(nlist)), dim3(kTileThreads), 0,       \
                               ctx->stream, G, T, I->lvl[0].cell_start, I->lvl[0].pts, list, start, order, q, k, skip, \
                               out_idx, out_d2, redo_out, kth_out, n_redo, rflag, scv);                                \
    } while (0)
        if (k <= 8) DWG_KNNT(8);
        else if (k <= 10) DWG_KNNT(10);
        else if (k == 15) DWG_KNNT(15);  // the diffusion's 15 neighbours: no spare slot to sink through
        else if (k <= 16) DWG_KNNT(16);
        else if (k <= 17) DWG_KNNT(17);  // the learned start's 16-point patches with the point itself
        else DWG_KNNT(32);
#undef DWG_KNNT
        check_launch();
    }
    int64_t nredo;
    if (ordered) {
        exclusive_scan<int64_t>(ctx, rflag, nq);
        hipLaunchKernelGGL(redo_compact_kernel, dim3(grid_for(nq, 256)), dim3(256), 0, ctx->stream, rflag, tmp_q,
                           tmp_kth, nq, redo, redo_kth);
        check_launch();
        nredo = read_scalar<int64_t>(ctx, rflag + nq);
    } else {
        nredo = int64_t(read_scalar<unsigned int>(ctx, n_redo));
    }
    ctx->count(name + ":redo", nredo);
    return nredo;
}

void knn_query(dwg_ctx* ctx, const dwg_knn_index* I, const float* q, int64_t nq, int k, const uint8_t* skip,
               const uint8_t* qmask, int32_t* out_idx, float* out_d2, const char* tag, const KnnScatter* sc) {
    DWG_REQUIRE(I, "null kNN index");
    DWG_REQUIRE(k >= 1 && k <= 32, "k must be in [1, 32]");
    DWG_REQUIRE(nq >= 0, "negative query count");
    DWG_REQUIRE(nq < (int64_t(1) << 31) - 1, "too many queries");
    DWG_REQUIRE(!sc || !qmask, "a scattering query covers every point");
    if (nq == 0) return;
    ProfileScope ps(ctx, tag ? tag : "knn_query");
    const std::string base = tag ? tag : "knn_query";
    if (nq < 4096) {  // too few to pay for the bucketing
        launch_per_lane(ctx, I, q, nq, nullptr, k, skip, qmask, out_idx, out_d2);
        if (sc) {
            hipLaunchKernelGGL(scatter_pairs_kernel, dim3(grid_for(nq * k, 256)), dim3(256), 0, ctx->stream, out_idx, k,
                               nullptr, nq, *sc);
            check_launch();
        }
        return;
    }
    if (!qmask && nq < kSparseMaxQueries && nq * 8 < I->n && I->lvl[0].pyr_levels >= 2) {
        // a few tens of thousands of queries spread thin over a large index
        // (the zero fill's) leave the tiles nearly empty, so staging a tile's
        // points serves one or two queries: a search per query costs less
        // (1.2 vs 2.1 ms for 35k queries over 950k points).  One wave per
        // query, its 64 lanes scanning the points of the coarse grid's
        // shells, beats one query per lane walking the occupancy pyramid
        // (0.63 vs 0.79 ms for the zero fill's ~30k queries, skipped points
        // included): the per-lane walks are few and long, the wave searches
        // many and short
        ctx->count(base + ":sparse", nq);
        // the queries in tile order: neighbouring waves read the same cells
        const TileBuckets B = tile_buckets(ctx, I, q, nq, nullptr, nullptr, 1, false);
        launch_wave(ctx, I, q, nq, B.order, k, skip, out_idx, out_d2);
        if (sc) {
            hipLaunchKernelGGL(scatter_pairs_kernel, dim3(grid_for(nq * k, 256)), dim3(256), 0, ctx->stream, out_idx, k,
                               nullptr, nq, *sc);
            check_launch();
        }
        return;
    }
    int* redo3 = static_cast<int*>(ctx->buf("knnt_redo3", sizeof(int) * nq + 64));
    float* kth3 = static_cast<float*>(ctx->buf("knnt_kth3", sizeof(float) * nq + 64));
    // at least as many queries as indexed points (triangle centroids): the
    // far ones come in dense patches (spurious surface pieces) and in tile
    // order, so a wave's 64 share most of their search and take the
    // wave-shared walk (6.4 vs 7.2 ms per centroid query at 1M points);
    // sparser far queries are spread apart and walk per lane
    const bool shared = nq >= I->n;
    int64_t n3 = 0;
    if (shared) {
        // the queries the first pass leaves open go straight to the shared
        // walk: its time is set by its longest walks (far queries deep
        // inside large closed pieces), not by how many queries it gets, so
        // wider tiled passes before it cost more than they save (4.41
        // instead of 5.20 ms per centroid query at 1M points)
        n3 = tiled_pass<kHalo1>(ctx, I, q, nq, nullptr, qmask, k, skip, out_idx, out_d2, redo3, base + ":tiles", sc,
                                kth3, true);
    } else {
        int* redo1 = static_cast<int*>(ctx->buf("knnt_redo1", sizeof(int) * nq + 64));
        int* redo2 = static_cast<int*>(ctx->buf("knnt_redo2", sizeof(int) * nq + 64));
        const int64_t n1 = tiled_pass<kHalo1>(ctx, I, q, nq, nullptr, qmask, k, skip, out_idx, out_d2, redo1,
                                              base + ":tiles", sc);
        // every tiled pass scatters the queries it closes (summed per tile
        // in LDS first); the far queries scatter after the last search
        const int64_t n2 = n1 > 0 ? tiled_pass<kHalo2>(ctx, I, q, n1, redo1, nullptr, k, skip, out_idx, out_d2,
                                                       redo2, base + ":tiles2", sc)
                                  : 0;
        n3 = n2 > 0 ? tiled_pass<kHalo3>(ctx, I, q, n2, redo2, nullptr, k, skip, out_idx, out_d2, redo3,
                                         base + ":tiles3", sc, kth3, true)
                    : 0;
    }
    {
        const std::string name = base + ":redo";
        ProfileScope pr(ctx, name.c_str());
        if (shared) ctx->count(base + ":far_shared", n3);
        const uint8_t* pm = (skip && n3 >= kPyrMinQueries && I->lvl[0].pyr_levels >= 2)
                                ? masked_pyramid(ctx, I->lvl[0], skip)
                                : nullptr;
        launch_far(ctx, I, q, n3, redo3, k, skip, out_idx, out_d2, kth3, shared, pm);
    }
    if (sc && n3 > 0) {
        const std::string name = base + ":scatter";
        ProfileScope pr(ctx, name.c_str());
        hipLaunchKernelGGL(scatter_pairs_hashed_kernel, dim3(grid_for(n3, 256)), dim3(256), 0, ctx->stream, out_idx,
                           k, redo3, n3, *sc);
        check_launch();
    }
}

}  // namespace dwg

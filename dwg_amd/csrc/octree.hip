// Octree construction on the GPU: octreeDWG's split rule
// (deps/libigl/include/igl/octree.cpp:171-335) and the order-0 moments of
// deps/libigl/include/igl/fast_winding_number.cpp:54-119, which the
// reference computes on the host before uploading (TurboLibiGWN.cu:132-231,
// DWGProcess.h:227, 608).
//
// 1. Each point's octant path is computed once, in FP64, by the same chain of
//    centre updates a recursive split performs (c ± w/4, w/2 per level; all
//    exact power-of-two scalings), 3 bits per level, most significant bits =
//    level 1.  The cell-centre comparisons are therefore bit-identical to a
//    recursive build.
// 2. A stable radix sort by path makes every node a contiguous range of the
//    sorted points; equal paths keep the caller's order.
// 3. Nodes are created one level at a time (breadth-first numbering, the 8
//    children of a node contiguous): split flags, an exclusive scan for the
//    child slots, and 8 binary searches over the parent's range per split
//    node.
// 4. ΣA and ΣA·p are summed bottom-up (a leaf over its points in sorted
//    order, an internal node over its 8 children in octant order), so CM is
//    deterministic; R = max |p − CM| is an exact maximum over every point of
//    the node (order-free), taken by a walk from the root per point with
//    wave-aggregated 64-bit atomic max on the IEEE bits (monotone for
//    non-negative doubles).

#include <algorithm>
#include <cmath>

#include "gwn.h"
#include "octree.h"
#include "scan.h"

namespace dwg {

namespace {

constexpr double kAreaDefault = 2e-6;  // DWGProcess.h:619

__global__ __launch_bounds__(256) void bounds_f64_kernel(const double* __restrict__ p, int64_t n,
                                                         double* __restrict__ partial) {
    __shared__ double lds[6][256];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], p[3 * i + d]);
            mx[d] = fmax(mx[d], p[3 * i + d]);
        }
    for (int d = 0; d < 3; ++d) {
        lds[d][threadIdx.x] = mn[d];
        lds[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int d = 0; d < 3; ++d) {
                lds[d][threadIdx.x] = fmin(lds[d][threadIdx.x], lds[d][threadIdx.x + off]);
                lds[3 + d][threadIdx.x] = fmax(lds[3 + d][threadIdx.x], lds[3 + d][threadIdx.x + off]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = lds[threadIdx.x][0];
}

__global__ void octree_codes_kernel(const double* __restrict__ P, int64_t n, double cx, double cy, double cz,
                                    double w0, int max_depth, uint64_t* __restrict__ code, int* __restrict__ idx) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = P[3 * i], y = P[3 * i + 1], z = P[3 * i + 2];
    double c0 = cx, c1 = cy, c2 = cz, w = w0;
    uint64_t k = 0;
    for (int d = 0; d < max_depth; ++d) {
        const int o = (x >= c0 ? 1 : 0) | (y >= c1 ? 2 : 0) | (z >= c2 ? 4 : 0);
        k = (k << 3) | uint64_t(o);
        const double q = w / 4;
        c0 += (o & 1) ? q : -q;
        c1 += (o & 2) ? q : -q;
        c2 += (o & 4) ? q : -q;
        w /= 2;
    }
    code[i] = k;
    idx[i] = int(i);
}

// flags[j] = 1 when node ls+j splits; flags[count] = 0 (scan total slot)
__global__ void split_flags_kernel(const int4* __restrict__ info, int64_t ls, int64_t count, int leaf_size,
                                   int64_t* __restrict__ flags) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j > count) return;
    if (j == count) {
        flags[j] = 0;
        return;
    }
    const int4 nf = info[ls + j];
    flags[j] = (nf.z - nf.y > leaf_size) ? 1 : 0;
}

__device__ __forceinline__ int digit_at(uint64_t code, int shift) { return int((code >> shift) & 7u); }

// Children of every split node on one level: child k covers the points of
// the parent's range whose octant digit at this level is k.
__global__ void children_kernel(int4* __restrict__ info, double4* __restrict__ cell, int64_t ls, int64_t count,
                                int64_t le, const int64_t* __restrict__ slot, const uint64_t* __restrict__ code,
                                int shift, int child_depth) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= count) return;
    if (slot[j + 1] == slot[j]) return;
    const int64_t idx = ls + j;
    const int64_t fc = le + 8 * slot[j];
    int4 nf = info[idx];
    nf.x = int(fc);
    info[idx] = nf;
    const double4 pc = cell[idx];
    const double q = pc.w / 4;
    int cur = nf.y;
    for (int k = 0; k < 8; ++k) {
        int lo = cur, hi = nf.z;
        while (lo < hi) {  // first slot whose digit exceeds k
            const int mid = lo + ((hi - lo) >> 1);
            if (digit_at(code[mid], shift) <= k) lo = mid + 1;
            else hi = mid;
        }
        info[fc + k] = make_int4(-1, cur, lo, child_depth);
        cell[fc + k] = make_double4(pc.x + ((k & 1) ? q : -q), pc.y + ((k & 2) ? q : -q), pc.z + ((k & 4) ? q : -q),
                                    pc.w / 2);
        cur = lo;
    }
}

// Bottom-up ΣA·p and ΣA of the nodes of one level; cm = ΣA·p/ΣA (0 for an
// empty node; NaN for a node whose points all have zero area, as the
// reference's 0/0 gives, TurboLibiGWN.cu:986-988: such a node is never far,
// its radius comes out NaN too).
__global__ void node_sums_kernel(const int4* __restrict__ info, int64_t ls, int64_t le, const int* __restrict__ perm,
                                 const double* __restrict__ P, const double* __restrict__ A,
                                 double4* __restrict__ sums, double4* __restrict__ cm) {
    const int64_t i = ls + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= le) return;
    const int4 nf = info[i];
    double sx = 0, sy = 0, sz = 0, sa = 0;
    if (nf.x < 0) {
        for (int s = nf.y; s < nf.z; ++s) {
            const int p = perm[s];
            const double a = A ? A[p] : kAreaDefault;
            sa += a;
            sx += a * P[3 * p];
            sy += a * P[3 * p + 1];
            sz += a * P[3 * p + 2];
        }
    } else {
        for (int k = 0; k < 8; ++k) {
            const double4 c = sums[nf.x + k];
            sx += c.x;
            sy += c.y;
            sz += c.z;
            sa += c.w;
        }
    }
    sums[i] = make_double4(sx, sy, sz, sa);
    const double nan = __longlong_as_double(0x7ff8000000000000ll);
    cm[i] = sa > 0 ? make_double4(sx / sa, sy / sa, sz / sa, sa)
                   : (nf.z > nf.y ? make_double4(nan, nan, nan, 0) : make_double4(0, 0, 0, 0));
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(v)), off, 64));
        const uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(v >> 32)), off, 64));
        const uint64_t o = (uint64_t(hi) << 32) | lo;
        v = o > v ? o : v;
    }
    return v;
}

// R = max |p − CM| over the points of each node: every point walks from the
// root to its leaf.  Where all 64 lanes of a wave stand on one node (the
// upper levels, points being sorted) the wave reduces first and issues one
// atomic.
__global__ __launch_bounds__(256) void node_radius_kernel(const int4* __restrict__ info, const double4* __restrict__ cm,
                                                          const uint64_t* __restrict__ code, const int* __restrict__ perm,
                                                          const double* __restrict__ P, int64_t n, int max_depth,
                                                          unsigned long long* __restrict__ rbits) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    const double x = P[3 * p], y = P[3 * p + 1], z = P[3 * p + 2];
    const uint64_t c = code[i];
    int node = 0, shift = 3 * (max_depth - 1);
    for (;;) {
        const double4 m = cm[node];
        const double dx = x - m.x, dy = y - m.y, dz = z - m.z;
        uint64_t bits = uint64_t(__double_as_longlong(sqrt(dx * dx + dy * dy + dz * dz)));
        const uint64_t active = __ballot(1);
        const int first = __builtin_amdgcn_readfirstlane(node);
        if (active == ~0ull && __ballot(node == first) == active) {
            bits = wave_max_u64(bits);
            if ((threadIdx.x & 63) == 0) atomicMax(&rbits[node], (unsigned long long)bits);
        } else {
            atomicMax(&rbits[node], (unsigned long long)bits);
        }
        const int fc = info[node].x;
        if (fc < 0) break;
        node = fc + digit_at(c, shift);
        shift -= 3;
    }
}

// Float copies the traversal reads: (float(CM), float(R)) and the rounding
// residual CM − float(CM).
__global__ void node_pack_kernel(const double4* __restrict__ cm, const unsigned long long* __restrict__ rbits,
                                 const int4* __restrict__ info, int64_t m, float4* __restrict__ cm_r,
                                 float* __restrict__ radius, float4* __restrict__ cm_lo, float4* __restrict__ payload,
                                 int2* __restrict__ kid) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    // traversal record: (first child, 0) for an internal node (first child
    // ≥ 1), (−count, begin) for a leaf (0 when empty)
    const int4 nf = info[i];
    kid[i] = nf.x >= 0 ? make_int2(nf.x, 0) : make_int2(nf.y - nf.z, nf.y);
    const double4 c = cm[i];
    const float cx = float(c.x), cy = float(c.y), cz = float(c.z);
    const float r = float(__longlong_as_double((long long)rbits[i]));
    radius[i] = r;
    cm_r[i] = make_float4(cx, cy, cz, r);
    // the dipole of a node next to the query is (CM − q)·EC with CM − q
    // nearly tangent to EC; the residual makes CM − q accurate to one
    // rounding of the difference itself
    cm_lo[i] = make_float4(float(c.x - double(cx)), float(c.y - double(cy)), float(c.z - double(cz)), 0.f);
    payload[2 * i] = make_float4(0.f, 0.f, 0.f, 0.f);
    payload[2 * i + 1] = cm_lo[i];
}

__global__ void pack_positions_kernel(const double* __restrict__ P, const int* __restrict__ perm, int64_t n,
                                      float4* __restrict__ sa) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    sa[i] = make_float4(float(P[3 * p]), float(P[3 * p + 1]), float(P[3 * p + 2]), kScreenDefault);
}

__global__ void leaf_flags_kernel(const int4* __restrict__ info, int64_t m, int min_num, int64_t* __restrict__ flags) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > m) return;
    if (i == m) {
        flags[i] = 0;
        return;
    }
    const int4 nf = info[i];
    flags[i] = (nf.x < 0 && nf.z - nf.y > min_num) ? 1 : 0;
}

__global__ void leaf_emit_kernel(const int4* __restrict__ info, const double4* __restrict__ cell, int64_t m,
                                 const int64_t* __restrict__ slot, double* __restrict__ centres,
                                 double* __restrict__ widths, int* __restrict__ nodes) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m || slot[i + 1] == slot[i]) return;
    const int64_t j = slot[i];
    const double4 c = cell[i];
    centres[3 * j] = c.x;
    centres[3 * j + 1] = c.y;
    centres[3 * j + 2] = c.z;
    if (widths) widths[j] = c.w;
    if (nodes) nodes[j] = int(i);
}

// Σ_j exp(−((P_jk − C_k)/W)²/4)·N_jk per component over the points of each
// leaf in ascending caller index (DWGProcess.h:430-472).
__global__ void leaf_smooth_kernel(const int4* __restrict__ info, const int* __restrict__ perm,
                                   const int* __restrict__ nodes, const double* __restrict__ centres,
                                   const double* __restrict__ widths, int64_t nl, const double* __restrict__ P,
                                   const double* __restrict__ N, double* __restrict__ out) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nl) return;
    const int4 nf = info[nodes[j]];
    const double W = widths[j];
    double acc[3] = {0, 0, 0};
    for (int s = nf.y; s < nf.z; ++s) {
        const int p = perm[s];
        for (int k = 0; k < 3; ++k) {
            const double x = (P[3 * p + k] - centres[3 * j + k]) / W;
            acc[k] += exp(-x * x / 4.0) * N[3 * p + k];
        }
    }
    for (int k = 0; k < 3; ++k) out[3 * j + k] = acc[k];
}

__global__ __launch_bounds__(256) void sumsq_f64_kernel(const double* __restrict__ v, int64_t n,
                                                        double* __restrict__ partial) {
    __shared__ double lds[256];
    double s = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) s += v[i] * v[i];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

__global__ void scale_to_f32_kernel(const double* __restrict__ v, int64_t n, double f, float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = float(f > 0 ? v[i] / f : 0.0);
}

__global__ __launch_bounds__(256) void sum_f32_as_f64_kernel(const float* __restrict__ p, int64_t n,
                                                             double* __restrict__ partial) {
    __shared__ double lds[3][256];
    double s[3] = {0, 0, 0};
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        for (int d = 0; d < 3; ++d) s[d] += double(p[3 * i + d]);
    for (int d = 0; d < 3; ++d) lds[d][threadIdx.x] = s[d];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int d = 0; d < 3; ++d) lds[d][threadIdx.x] += lds[d][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 3) partial[3 * blockIdx.x + threadIdx.x] = lds[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void max_dev_kernel(const float* __restrict__ p, int64_t n, double cx, double cy,
                                                      double cz, double* __restrict__ partial) {
    __shared__ double lds[256];
    double s = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        s = fmax(s, 2.0 * fabs(double(p[3 * i]) - cx));
        s = fmax(s, 2.0 * fabs(double(p[3 * i + 1]) - cy));
        s = fmax(s, 2.0 * fabs(double(p[3 * i + 2]) - cz));
    }
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

__global__ void normalize_kernel(const float* __restrict__ p, int64_t n, double cx, double cy, double cz, double s,
                                 double* __restrict__ out64, float* __restrict__ out32) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    const int d = int(i % 3);
    const double c = d == 0 ? cx : (d == 1 ? cy : cz);
    const double v = (double(p[i]) - c) / s;
    out64[i] = v;
    if (out32) out32[i] = float(v);
}

__global__ void f64_to_f32_kernel(const double* __restrict__ a, int64_t n, float* __restrict__ b) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = float(a[i]);
}

__global__ void f32_to_f64_kernel(const float* __restrict__ a, int64_t n, double* __restrict__ b) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = double(a[i]);
}

constexpr int kReduceBlocks = 256;

template <class T>
T* dev_alloc(int64_t count) {
    return static_cast<T*>(dev_malloc(sizeof(T) * size_t(std::max<int64_t>(count, 1))));
}

}  // namespace

void bounds_f64(dwg_ctx* ctx, const double* P, int64_t n, double lo[3], double hi[3]) {
    DWG_REQUIRE(n > 0, "bounds of an empty point set");
    double* partial = static_cast<double*>(ctx->buf("bounds64_partial", sizeof(double) * 6 * kReduceBlocks));
    hipLaunchKernelGGL(bounds_f64_kernel, dim3(kReduceBlocks), dim3(256), 0, ctx->stream, P, n, partial);
    check_launch();
    std::vector<double> h(6 * kReduceBlocks);
    DWG_HIP_CHECK(hipMemcpyAsync(h.data(), partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int d = 0; d < 3; ++d) {
        lo[d] = INFINITY;
        hi[d] = -INFINITY;
    }
    for (int b = 0; b < kReduceBlocks; ++b)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], h[6 * b + d]);
            hi[d] = std::max(hi[d], h[6 * b + 3 + d]);
        }
}

dwg_octree* octree_build(dwg_ctx* ctx, const double* P, const double* A, int64_t n, int max_depth, int leaf_size) {
    DWG_REQUIRE(n > 0, "octree needs at least one point");
    DWG_REQUIRE(max_depth >= 0 && max_depth <= 21, "max_depth must be in [0, 21]");
    DWG_REQUIRE(leaf_size >= 1, "leaf_size must be ≥ 1");
    DWG_REQUIRE(n < (int64_t(1) << 31) - 1, "too many points");
    ProfileScope ps(ctx, "octree_build");
    double lo[3], hi[3];
    bounds_f64(ctx, P, n, lo, hi);
    const double c0[3] = {(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2};
    const double w0 = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});

    auto* T = new dwg_octree();
    int4* info = nullptr;
    double4* cell = nullptr;
    try {
        T->n_points = n;
        T->max_depth = max_depth;
        // 1. octant paths and the stable sort by path
        uint64_t* code_in = static_cast<uint64_t*>(ctx->buf("oct_code_in", sizeof(uint64_t) * n));
        uint64_t* code = static_cast<uint64_t*>(ctx->buf("oct_code", sizeof(uint64_t) * n));
        int* idx_in = static_cast<int*>(ctx->buf("oct_idx_in", sizeof(int) * n));
        T->perm = dev_alloc<int>(n);
        hipLaunchKernelGGL(octree_codes_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, P, n, c0[0], c0[1],
                           c0[2], w0, max_depth, code_in, idx_in);
        check_launch();
        radix_sort_pairs(ctx, code_in, idx_in, code, T->perm, n, std::max(1, 3 * max_depth));

        // 2. nodes, one level at a time
        int64_t cap = std::max<int64_t>(1024, 2 * n + 9);
        info = dev_alloc<int4>(cap);
        cell = dev_alloc<double4>(cap);
        const int4 root_info = make_int4(-1, 0, int(n), 0);
        const double4 root_cell = make_double4(c0[0], c0[1], c0[2], w0);
        DWG_HIP_CHECK(hipMemcpyAsync(info, &root_info, sizeof(int4), hipMemcpyHostToDevice, ctx->stream));
        DWG_HIP_CHECK(hipMemcpyAsync(cell, &root_cell, sizeof(double4), hipMemcpyHostToDevice, ctx->stream));
        T->level_start.assign({0, 1});
        int64_t ls = 0, le = 1;
        for (int d = 0; d < max_depth; ++d) {
            const int64_t count = le - ls;
            int64_t* slot = static_cast<int64_t*>(ctx->buf("oct_slot", sizeof(int64_t) * (count + 1)));
            hipLaunchKernelGGL(split_flags_kernel, dim3(grid_for(count + 1, 256)), dim3(256), 0, ctx->stream, info, ls,
                               count, leaf_size, slot);
            check_launch();
            exclusive_scan<int64_t>(ctx, slot, count);
            const int64_t nsplit = read_scalar<int64_t>(ctx, slot + count);
            if (nsplit == 0) break;
            const int64_t total = le + 8 * nsplit;
            DWG_REQUIRE(total < (int64_t(1) << 31), "octree too large");
            if (total > cap) {
                const int64_t ncap = std::max(total, cap + cap / 2);
                int4* ni = dev_alloc<int4>(ncap);
                double4* nc = dev_alloc<double4>(ncap);
                DWG_HIP_CHECK(hipMemcpyAsync(ni, info, sizeof(int4) * le, hipMemcpyDeviceToDevice, ctx->stream));
                DWG_HIP_CHECK(hipMemcpyAsync(nc, cell, sizeof(double4) * le, hipMemcpyDeviceToDevice, ctx->stream));
                DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
                dev_free(info);
                dev_free(cell);
                info = ni;
                cell = nc;
                cap = ncap;
            }
            hipLaunchKernelGGL(children_kernel, dim3(grid_for(count, 256)), dim3(256), 0, ctx->stream, info, cell, ls,
                               count, le, slot, code, 3 * (max_depth - 1 - d), d + 1);
            check_launch();
            ls = le;
            le = total;
            T->level_start.push_back(le);
        }
        const int64_t m = le;
        T->n_nodes = m;
        T->max_level = int(T->level_start.size()) - 2;
        T->info = info;
        T->cell = cell;
        info = nullptr;
        cell = nullptr;

        // 3. moments: bottom-up sums, then the exact radius
        double4* sums = static_cast<double4*>(ctx->buf("oct_sums", sizeof(double4) * m));
        T->cm64 = dev_alloc<double4>(m);
        for (int lv = T->max_level; lv >= 0; --lv) {
            const int64_t a = T->level_start[lv], b = T->level_start[lv + 1];
            hipLaunchKernelGGL(node_sums_kernel, dim3(grid_for(b - a, 256)), dim3(256), 0, ctx->stream, T->info, a, b,
                               T->perm, P, A, sums, T->cm64);
            check_launch();
        }
        unsigned long long* rbits = static_cast<unsigned long long*>(ctx->buf("oct_rbits", sizeof(uint64_t) * m));
        fill_u32(ctx, rbits, 0, 2 * m);
        hipLaunchKernelGGL(node_radius_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, T->info, T->cm64,
                           code, T->perm, P, n, max_depth, rbits);
        check_launch();
        T->cm_r = dev_alloc<float4>(m);
        T->radius = dev_alloc<float>(m);
        T->cm_lo = dev_alloc<float4>(m);
        T->payload = dev_alloc<float4>(2 * m);
        T->kid = dev_alloc<int2>(m);
        hipLaunchKernelGGL(node_pack_kernel, dim3(grid_for(m, 256)), dim3(256), 0, ctx->stream, T->cm64, rbits, T->info,
                           m, T->cm_r, T->radius, T->cm_lo, T->payload, T->kid);
        check_launch();

        // 4. sources in tree order and the per-iteration buffers
        T->ec = dev_alloc<float4>(m);
        T->sa = dev_alloc<float4>(n);
        T->sb = dev_alloc<float4>(n);
        T->an64 = dev_alloc<Double3>(n + 1);
        hipLaunchKernelGGL(pack_positions_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, P, T->perm, n,
                           T->sa);
        check_launch();

        // leaves (non-empty nodes without children)
        int64_t* lslot = static_cast<int64_t*>(ctx->buf("oct_lslot", sizeof(int64_t) * (m + 1)));
        hipLaunchKernelGGL(leaf_flags_kernel, dim3(grid_for(m + 1, 256)), dim3(256), 0, ctx->stream, T->info, m, 0,
                           lslot);
        check_launch();
        exclusive_scan<int64_t>(ctx, lslot, m);
        T->n_leaves = read_scalar<int64_t>(ctx, lslot + m);
    } catch (...) {
        if (info) dev_free(info);
        if (cell) dev_free(cell);
        delete T;
        throw;
    }
    return T;
}

int64_t octree_leaves(dwg_ctx* ctx, const dwg_octree* T, int min_num, double* centres, double* widths, int* nodes) {
    const int64_t m = T->n_nodes;
    int64_t* slot = static_cast<int64_t*>(ctx->buf("oct_leaf_slot", sizeof(int64_t) * (m + 1)));
    hipLaunchKernelGGL(leaf_flags_kernel, dim3(grid_for(m + 1, 256)), dim3(256), 0, ctx->stream, T->info, m, min_num,
                       slot);
    check_launch();
    exclusive_scan<int64_t>(ctx, slot, m);
    const int64_t nl = read_scalar<int64_t>(ctx, slot + m);
    DWG_REQUIRE(nl <= T->n_leaves, "leaf count exceeds the octree's leaves");
    hipLaunchKernelGGL(leaf_emit_kernel, dim3(grid_for(m, 256)), dim3(256), 0, ctx->stream, T->info, T->cell, m, slot,
                       centres, widths, nodes);
    check_launch();
    return nl;
}

void leaf_smooth_normals(dwg_ctx* ctx, const dwg_octree* T, const int* nodes, const double* centres,
                         const double* widths, int64_t nl, const double* P, const double* N, float* out) {
    if (nl == 0) return;
    double* acc = static_cast<double*>(ctx->buf("leaf_smooth_acc", sizeof(double) * 3 * nl));
    hipLaunchKernelGGL(leaf_smooth_kernel, dim3(grid_for(nl, 256)), dim3(256), 0, ctx->stream, T->info, T->perm, nodes,
                       centres, widths, nl, P, N, acc);
    check_launch();
    // whole-matrix Frobenius normalisation (DWGProcess.h:537)
    double* partial = static_cast<double*>(ctx->buf("leaf_smooth_partial", sizeof(double) * kReduceBlocks));
    hipLaunchKernelGGL(sumsq_f64_kernel, dim3(kReduceBlocks), dim3(256), 0, ctx->stream, acc, 3 * nl, partial);
    check_launch();
    std::vector<double> h(kReduceBlocks);
    DWG_HIP_CHECK(hipMemcpyAsync(h.data(), partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double f = 0;
    for (double v : h) f += v;
    f = std::sqrt(f);
    hipLaunchKernelGGL(scale_to_f32_kernel, dim3(grid_for(3 * nl, 256)), dim3(256), 0, ctx->stream, acc, 3 * nl, f, out);
    check_launch();
}

void normalize_points(dwg_ctx* ctx, const float* p, int64_t n, double* out64, float* out32, double center[3],
                      double* scale) {
    DWG_REQUIRE(n > 0, "no points to normalise");
    double* partial = static_cast<double*>(ctx->buf("normalize_partial", sizeof(double) * 3 * kReduceBlocks));
    hipLaunchKernelGGL(sum_f32_as_f64_kernel, dim3(kReduceBlocks), dim3(256), 0, ctx->stream, p, n, partial);
    check_launch();
    std::vector<double> h(3 * kReduceBlocks);
    DWG_HIP_CHECK(hipMemcpyAsync(h.data(), partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double c[3] = {0, 0, 0};
    for (int b = 0; b < kReduceBlocks; ++b)
        for (int d = 0; d < 3; ++d) c[d] += h[3 * b + d];
    for (int d = 0; d < 3; ++d) c[d] /= double(n);
    hipLaunchKernelGGL(max_dev_kernel, dim3(kReduceBlocks), dim3(256), 0, ctx->stream, p, n, c[0], c[1], c[2], partial);
    check_launch();
    DWG_HIP_CHECK(hipMemcpyAsync(h.data(), partial, sizeof(double) * kReduceBlocks, hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double s = 0;
    for (int b = 0; b < kReduceBlocks; ++b) s = std::max(s, h[b]);
    if (s <= 0) s = 1;
    hipLaunchKernelGGL(normalize_kernel, dim3(grid_for(3 * n, 256)), dim3(256), 0, ctx->stream, p, n, c[0], c[1], c[2],
                       s, out64, out32);
    check_launch();
    for (int d = 0; d < 3; ++d) center[d] = c[d];
    *scale = s;
}

void to_f32(dwg_ctx* ctx, const double* a, int64_t n, float* b) {
    if (n <= 0) return;
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, a, n, b);
    check_launch();
}

void to_f64(dwg_ctx* ctx, const float* a, int64_t n, double* b) {
    if (n <= 0) return;
    hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, a, n, b);
    check_launch();
}

}  // namespace dwg

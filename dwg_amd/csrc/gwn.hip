// Generalised-winding-number dipole sums on CDNA4.
//
//   direct:  all-pairs sums, one query per lane, sources streamed through
//            the scalar cache (wave-uniform loads → SGPR operands), per-tile
//            partial sums folded with Kahan compensation.
//   tree:    Barnes-Hut over an octree built with octreeDWG's split rule;
//            a wave walks the union of its 64 lanes' traversals with one
//            wave-uniform stack in LDS (node, lane mask), so every node and
//            source load is wave-uniform and each lane still takes exactly
//            the reference's per-query term set.
//
// Terms (reference semantics, d = p − q, r = |d|):
//   plain     A(d·n) / (4π r³ + 1e-8)                         fast_winding_number.cpp:182-184
//   screened  (σρ+1) A(d·n) / (e^{σρ} 4π ρ³ + 1e-7), ρ = max(s, r), r > 1e-10
//                                                             TurboLibiGWN.cu:544-585
// Both are evaluated with numerator and denominator divided by 4π (A·n is
// pre-scaled by 1/(4π) when sources are packed), which is the same value.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <functional>
#include <numeric>
#include <thread>

#include "gwn.h"
#include "scan.h"

namespace dwg {

// ---------------------------------------------------------------------------
// per-pair terms
// ---------------------------------------------------------------------------
// term from the offset d = p − q; s = screening radius of the source
template <bool SCREENED>
__device__ __forceinline__ float gwn_term_d(float dx, float dy, float dz, float s, const float4 b, float sigma,
                                            float sigma_l2e) {
    const float r2 = dist2(dx, dy, dz);
    const float dot = __builtin_fmaf(dz, b.z, __builtin_fmaf(dy, b.y, dx * b.x));
    const float r = fast_sqrt(r2);
    if constexpr (!SCREENED) {
        const float den = __builtin_fmaf(r2, r, kEpsPlain * kInv4Pi);
        return dot * fast_rcp(den);
    } else {
        const float rho = fmaxf(r, s);
        const float e = fast_exp2(sigma_l2e * rho);
        const float rho3 = rho * rho * rho;
        const float den = __builtin_fmaf(e, rho3, kEpsScreened * kInv4Pi);
        const float t = (__builtin_fmaf(sigma, rho, 1.0f) * dot) * fast_rcp(den);
        return r > 1e-10f ? t : 0.0f;
    }
}

template <bool SCREENED>
__device__ __forceinline__ float gwn_term(float qx, float qy, float qz, const float4 a, const float4 b,
                                          float sigma, float sigma_l2e) {
    return gwn_term_d<SCREENED>(a.x - qx, a.y - qy, a.z - qz, a.w, b, sigma, sigma_l2e);
}

// plain term and its gradient w.r.t. q (b pre-scaled by 1/4π):
//   w = (d·b)/D, D = r³ + ε/4π;  ∇w = −b/D + 3 r (d·b) d / D²
__device__ __forceinline__ void gwn_term_grad(float qx, float qy, float qz, const float4 a, const float4 b,
                                              float& w, float& gx, float& gy, float& gz) {
    const float dx = a.x - qx, dy = a.y - qy, dz = a.z - qz;
    const float r2 = dist2(dx, dy, dz);
    const float dot = __builtin_fmaf(dz, b.z, __builtin_fmaf(dy, b.y, dx * b.x));
    const float r = fast_sqrt(r2);
    const float inv = fast_rcp(__builtin_fmaf(r2, r, kEpsPlain * kInv4Pi));
    w = dot * inv;
    const float c = 3.0f * r * w * inv;
    gx = __builtin_fmaf(c, dx, -b.x * inv);
    gy = __builtin_fmaf(c, dy, -b.y * inv);
    gz = __builtin_fmaf(c, dz, -b.z * inv);
}

// ---------------------------------------------------------------------------
// source packing: (x, y, z, s) and (A·n/4π, A)
// ---------------------------------------------------------------------------
__global__ void pack_sources_kernel(const float* __restrict__ pts, const float* __restrict__ nrm,
                                    const float* __restrict__ area, const float* __restrict__ radii,
                                    const int* __restrict__ perm, int64_t n, int use_radii,
                                    float4* __restrict__ sa, float4* __restrict__ sb,
                                    Double3* __restrict__ an64) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = perm ? perm[i] : i;
    const float A = area ? area[p] : 2e-6f;
    const float s = use_radii ? radii[p] : kScreenDefault;
    if (sa) sa[i] = make_float4(pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], s);
    const float nx = nrm[3 * p], ny = nrm[3 * p + 1], nz = nrm[3 * p + 2];
    sb[i] = make_float4(A * nx * kInv4Pi, A * ny * kInv4Pi, A * nz * kInv4Pi, A);
    if (an64) an64[i] = Double3(double(A) * nx, double(A) * ny, double(A) * nz);
}

// ---------------------------------------------------------------------------
// direct all-pairs kernel
// ---------------------------------------------------------------------------
constexpr int kDirectThreads = 256;
constexpr int kDirectTile = 64;

template <bool SCREENED, bool KAHAN>
__global__ __launch_bounds__(kDirectThreads) void gwn_direct_kernel(
    const float4* __restrict__ sa, const float4* __restrict__ sb, int64_t n, const float* __restrict__ q,
    int64_t nq, float sigma, float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * kDirectThreads + threadIdx.x;
    const bool valid = i < nq;
    const float qx = valid ? q[3 * i] : 0.f, qy = valid ? q[3 * i + 1] : 0.f, qz = valid ? q[3 * i + 2] : 0.f;
    const float sl2e = sigma * kLog2e;
    float sum = 0.f, comp = 0.f;
    int64_t j0 = 0;
    for (; j0 + kDirectTile <= n; j0 += kDirectTile) {
        float part = 0.f;
#pragma unroll 8
        for (int j = 0; j < kDirectTile; ++j) part += gwn_term<SCREENED>(qx, qy, qz, sa[j0 + j], sb[j0 + j], sigma, sl2e);
        if (KAHAN) kahan_add(sum, comp, part); else sum += part;
    }
    float part = 0.f;
    for (int64_t j = j0; j < n; ++j) part += gwn_term<SCREENED>(qx, qy, qz, sa[j], sb[j], sigma, sl2e);
    if (KAHAN) kahan_add(sum, comp, part); else sum += part;
    if (valid) out[i] = sum - comp;
}

template <bool KAHAN>
__global__ __launch_bounds__(kDirectThreads) void gwn_direct_grad_kernel(
    const float4* __restrict__ sa, const float4* __restrict__ sb, int64_t n, const float* __restrict__ q,
    int64_t nq, float* __restrict__ out_w, float* __restrict__ out_g) {
    const int64_t i = int64_t(blockIdx.x) * kDirectThreads + threadIdx.x;
    const bool valid = i < nq;
    const float qx = valid ? q[3 * i] : 0.f, qy = valid ? q[3 * i + 1] : 0.f, qz = valid ? q[3 * i + 2] : 0.f;
    float s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    for (int64_t j0 = 0; j0 < n; j0 += kDirectTile) {
        const int64_t j1 = j0 + kDirectTile < n ? j0 + kDirectTile : n;
        float p[4] = {0, 0, 0, 0};
        for (int64_t j = j0; j < j1; ++j) {
            float w, gx, gy, gz;
            gwn_term_grad(qx, qy, qz, sa[j], sb[j], w, gx, gy, gz);
            p[0] += w;
            p[1] += gx;
            p[2] += gy;
            p[3] += gz;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (KAHAN) kahan_add(s[k], c[k], p[k]); else s[k] += p[k];
        }
    }
    if (valid) {
        out_w[i] = s[0] - c[0];
        if (out_g) {
            out_g[3 * i] = s[1] - c[1];
            out_g[3 * i + 1] = s[2] - c[2];
            out_g[3 * i + 2] = s[3] - c[3];
        }
    }
}

void gwn_direct(dwg_ctx* ctx, const float* pts, const float* nrm, const float* area, const float* radii, int64_t n,
                const float* q, int64_t nq, float sigma, uint32_t flags, float* out_w, float* out_g) {
    DWG_REQUIRE(n >= 0 && nq >= 0, "negative size");
    DWG_REQUIRE(!(flags & DWG_USE_RADII) || radii, "DWG_USE_RADII needs radii");
    DWG_REQUIRE(!(out_g && (flags & DWG_SCREENED)), "gradient is defined for the plain kernel only");
    if (nq == 0) return;
    float4* sa = static_cast<float4*>(ctx->buf("direct_sa", sizeof(float4) * (n + 1)));
    float4* sb = static_cast<float4*>(ctx->buf("direct_sb", sizeof(float4) * (n + 1)));
    if (n > 0) {
        ProfileScope ps(ctx, "pack_sources");
        hipLaunchKernelGGL(pack_sources_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, pts, nrm, area,
                           radii, nullptr, n, (flags & DWG_USE_RADII) ? 1 : 0, sa, sb, nullptr);
        check_launch();
    }
    const bool kahan = flags & DWG_KAHAN;
    dim3 grid(grid_for(nq, kDirectThreads)), block(kDirectThreads);
    ProfileScope ps(ctx, out_g ? "gwn_direct_grad" : ((flags & DWG_SCREENED) ? "gwn_direct_screened" : "gwn_direct_plain"));
    if (out_g) {
        if (kahan) hipLaunchKernelGGL(gwn_direct_grad_kernel<true>, grid, block, 0, ctx->stream, sa, sb, n, q, nq, out_w, out_g);
        else hipLaunchKernelGGL(gwn_direct_grad_kernel<false>, grid, block, 0, ctx->stream, sa, sb, n, q, nq, out_w, out_g);
    } else if (flags & DWG_SCREENED) {
        if (kahan) hipLaunchKernelGGL((gwn_direct_kernel<true, true>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w);
        else hipLaunchKernelGGL((gwn_direct_kernel<true, false>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w);
    } else {
        if (kahan) hipLaunchKernelGGL((gwn_direct_kernel<false, true>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w);
        else hipLaunchKernelGGL((gwn_direct_kernel<false, false>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w);
    }
    check_launch();
}

// ---------------------------------------------------------------------------
// octree build (host).  Split rule and child numbering of octreeDWG
// (octree.cpp:171-335): the root cube is centred on the AABB with the
// largest extent as width; a node with more than leaf_size points and depth
// < max_depth gets 8 children, allocated contiguously in pre-order; octant
// bit 1: x ≥ cx, 2: y ≥ cy, 4: z ≥ cz (exact double comparisons).  Points
// are stored sorted by their octant path so every node owns a contiguous
// range [begin, end) of the packed source array.
// ---------------------------------------------------------------------------
static void parallel_ranges(int64_t n, const std::function<void(int64_t, int64_t)>& fn) {
    unsigned nt = std::max(1u, std::thread::hardware_concurrency());
    if (n < 20000 || nt == 1) {
        fn(0, n);
        return;
    }
    nt = std::min(nt, 64u);
    std::vector<std::thread> th;
    int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) th.emplace_back([=, &fn]() { fn(b, e); });
    }
    for (auto& x : th) x.join();
}

HostOctree build_octree_host(const double* P, const double* A, int64_t n, int max_depth, int leaf_size) {
    DWG_REQUIRE(n > 0, "octree needs at least one point");
    DWG_REQUIRE(max_depth >= 0 && max_depth <= 21, "max_depth must be in [0, 21]");
    DWG_REQUIRE(leaf_size >= 1, "leaf_size must be ≥ 1");
    DWG_REQUIRE(n < (int64_t(1) << 31) - 1, "too many points");
    HostOctree T;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], P[3 * i + d]);
            hi[d] = std::max(hi[d], P[3 * i + d]);
        }
    const double c0[3] = {(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2};
    const double w0 = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
    // octant path codes, 3 bits per level, most significant = level 1
    std::vector<uint64_t> code(n);
    parallel_ranges(n, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            double c[3] = {c0[0], c0[1], c0[2]}, w = w0;
            uint64_t k = 0;
            for (int d = 0; d < max_depth; ++d) {
                int o = (P[3 * i] >= c[0] ? 1 : 0) | (P[3 * i + 1] >= c[1] ? 2 : 0) | (P[3 * i + 2] >= c[2] ? 4 : 0);
                k = (k << 3) | uint64_t(o);
                const double q = w / 4;
                c[0] += (o & 1) ? q : -q;
                c[1] += (o & 2) ? q : -q;
                c[2] += (o & 4) ? q : -q;
                w /= 2;
            }
            code[i] = k;
        }
    });
    // stable LSD radix sort of (code, index) by code, 11-bit digits
    T.perm.resize(n);
    std::iota(T.perm.begin(), T.perm.end(), 0);
    std::vector<uint64_t> sc(code);
    {
        std::vector<uint64_t> k2(n);
        std::vector<int> v2(n);
        const int bits = 3 * max_depth;
        for (int shift = 0; shift < bits; shift += 11) {
            std::vector<int64_t> cnt(2049, 0);
            for (int64_t i = 0; i < n; ++i) ++cnt[((sc[i] >> shift) & 2047) + 1];
            for (int d = 0; d < 2048; ++d) cnt[d + 1] += cnt[d];
            for (int64_t i = 0; i < n; ++i) {
                const int64_t dst = cnt[(sc[i] >> shift) & 2047]++;
                k2[dst] = sc[i];
                v2[dst] = T.perm[i];
            }
            sc.swap(k2);
            T.perm.swap(v2);
        }
    }
    // positions and areas in sorted order, contiguous for the moment loops
    std::vector<double> Ps(4 * n);
    parallel_ranges(n, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            const int p = T.perm[i];
            Ps[4 * i] = P[3 * p];
            Ps[4 * i + 1] = P[3 * p + 1];
            Ps[4 * i + 2] = P[3 * p + 2];
            Ps[4 * i + 3] = A ? A[p] : 2e-6;
        }
    });
    // Nodes level by level (breadth-first numbering; the 8 children of a node
    // stay contiguous, which is all the traversal relies on).  Each level is
    // one parallel pass: split flags, a prefix sum for the child slots, then
    // every split node cuts its sorted range at the octant digits.
    T.first_child.assign(1, -1);
    T.begin.assign(1, 0);
    T.end.assign(1, int(n));
    T.depth.assign(1, 0);
    T.cn.assign({c0[0], c0[1], c0[2]});
    T.w.assign(1, w0);
    int64_t ls = 0, le = 1;
    for (int d = 0; d < max_depth && ls < le; ++d) {
        std::vector<int64_t> slot(le - ls + 1, 0);
        for (int64_t i = ls; i < le; ++i) slot[i - ls + 1] = slot[i - ls] + ((T.end[i] - T.begin[i] > leaf_size) ? 1 : 0);
        const int64_t nsplit = slot[le - ls];
        if (nsplit == 0) break;
        const int64_t total = le + 8 * nsplit;
        T.first_child.resize(total, -1);
        T.begin.resize(total);
        T.end.resize(total);
        T.depth.resize(total);
        T.cn.resize(3 * total);
        T.w.resize(total);
        const int shift = 3 * (max_depth - 1 - d);
        parallel_ranges(le - ls, [&](int64_t b0, int64_t e0) {
            for (int64_t j = b0; j < e0; ++j) {
                const int64_t idx = ls + j;
                if (slot[j + 1] == slot[j]) continue;
                const int64_t fc = le + 8 * slot[j];
                T.first_child[idx] = int(fc);
                const int b = T.begin[idx], e = T.end[idx];
                const double q = T.w[idx] / 4;
                int cur = b;
                for (int k = 0; k < 8; ++k) {
                    int stop = cur;
                    while (stop < e && int((sc[stop] >> shift) & 7) == k) ++stop;
                    const int64_t c = fc + k;
                    T.begin[c] = cur;
                    T.end[c] = stop;
                    T.depth[c] = d + 1;
                    T.cn[3 * c] = T.cn[3 * idx] + ((k & 1) ? q : -q);
                    T.cn[3 * c + 1] = T.cn[3 * idx + 1] + ((k & 2) ? q : -q);
                    T.cn[3 * c + 2] = T.cn[3 * idx + 2] + ((k & 4) ? q : -q);
                    T.w[c] = T.w[idx] / 2;
                    cur = stop;
                }
            }
        });
        ls = le;
        le = total;
    }
    const int64_t m = int64_t(T.first_child.size());
    T.cm.assign(3 * m, 0.0);
    T.r.assign(m, 0.0);
    // CM = ΣA p / ΣA, R = max |p − CM| (fast_winding_number.cpp:59-110), FP64.
    // Node costs are uneven (the root spans every point), so threads take
    // blocks of nodes from a shared counter.
    std::atomic<int64_t> next{0};
    const int64_t blk = 2048;
    parallel_ranges(int64_t(std::max(1u, std::thread::hardware_concurrency())) * 20000, [&](int64_t, int64_t) {
        for (;;) {
            const int64_t b = next.fetch_add(blk);
            if (b >= m) return;
            const int64_t e = std::min(m, b + blk);
            for (int64_t i = b; i < e; ++i) {
                const int pb = T.begin[i], pe = T.end[i];
                if (pb == pe) continue;
                double at = 0, c[3] = {0, 0, 0};
                for (int s = pb; s < pe; ++s) {
                    const double a = Ps[4 * s + 3];
                    at += a;
                    for (int d = 0; d < 3; ++d) c[d] += a * Ps[4 * s + d];
                }
                for (int d = 0; d < 3; ++d) c[d] /= at;
                double rm = 0;
                for (int s = pb; s < pe; ++s) {
                    const double dx = Ps[4 * s] - c[0], dy = Ps[4 * s + 1] - c[1], dz = Ps[4 * s + 2] - c[2];
                    rm = std::max(rm, std::sqrt(dx * dx + dy * dy + dz * dz));
                }
                for (int d = 0; d < 3; ++d) T.cm[3 * i + d] = c[d];
                T.r[i] = rm;
            }
        }
    });
    T.n_points = n;
    T.max_level = 0;
    T.n_leaves = 0;
    for (int64_t i = 0; i < m; ++i) {
        T.max_level = std::max(T.max_level, T.depth[i]);
        if (T.first_child[i] < 0 && T.end[i] > T.begin[i]) ++T.n_leaves;
    }
    return T;
}

}  // namespace dwg

// ---------------------------------------------------------------------------
// device octree
// ---------------------------------------------------------------------------
dwg_octree::~dwg_octree() {
    for (void* p : {(void*)info, (void*)cm_r, (void*)cm_lo, (void*)ec, (void*)ec_bf16, (void*)sa, (void*)sb, (void*)perm, (void*)an64})
        if (p) (void)hipFree(p);
}

namespace dwg {

__global__ void node_ec_kernel(const int4* __restrict__ info, const Double3* __restrict__ scan, int64_t m,
                               float4* __restrict__ ec, uint2* __restrict__ ec_bf16) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int4 nf = info[i];
    const Double3 s = scan[nf.z] - scan[nf.y];
    const float x = float(s.x * double(kInv4Pi)), y = float(s.y * double(kInv4Pi)), z = float(s.z * double(kInv4Pi));
    ec[i] = make_float4(x, y, z, 0.f);
    if (ec_bf16) {
        // round-to-nearest-even bf16 of each component
        auto bf = [](float f) -> uint32_t {
            uint32_t u = __float_as_uint(f);
            u += 0x7FFFu + ((u >> 16) & 1u);
            return u >> 16;
        };
        ec_bf16[i] = make_uint2(bf(x) | (bf(y) << 16), bf(z));
    }
}

dwg_octree* octree_upload(dwg_ctx* ctx, const HostOctree& H) {
    auto* T = new dwg_octree();
    try {
        const int64_t m = int64_t(H.first_child.size()), n = H.n_points;
        T->n_nodes = m;
        T->n_points = n;
        T->n_leaves = H.n_leaves;
        T->max_level = H.max_level;
        std::vector<int4> info(m);
        std::vector<float4> cmr(m), cml(m);
        for (int64_t i = 0; i < m; ++i) {
            info[i] = make_int4(H.first_child[i], H.begin[i], H.end[i], H.depth[i]);
            const float cx = float(H.cm[3 * i]), cy = float(H.cm[3 * i + 1]), cz = float(H.cm[3 * i + 2]);
            cmr[i] = make_float4(cx, cy, cz, float(H.r[i]));
            // the dipole of a node next to the query is (CM − q)·EC with CM − q
            // nearly tangent to EC; keeping the rounding residual of CM makes
            // CM − q accurate to one rounding of the difference itself
            cml[i] = make_float4(float(H.cm[3 * i] - double(cx)), float(H.cm[3 * i + 1] - double(cy)),
                                 float(H.cm[3 * i + 2] - double(cz)), 0.f);
        }
        DWG_HIP_CHECK(hipMalloc(&T->info, sizeof(int4) * m));
        DWG_HIP_CHECK(hipMalloc(&T->cm_r, sizeof(float4) * m));
        DWG_HIP_CHECK(hipMalloc(&T->cm_lo, sizeof(float4) * m));
        DWG_HIP_CHECK(hipMalloc(&T->ec, sizeof(float4) * m));
        DWG_HIP_CHECK(hipMalloc(&T->ec_bf16, sizeof(uint2) * m));
        DWG_HIP_CHECK(hipMalloc(&T->sa, sizeof(float4) * n));
        DWG_HIP_CHECK(hipMalloc(&T->sb, sizeof(float4) * n));
        DWG_HIP_CHECK(hipMalloc(&T->perm, sizeof(int) * n));
        DWG_HIP_CHECK(hipMalloc(&T->an64, sizeof(Double3) * (n + 1)));
        DWG_HIP_CHECK(hipMemcpyAsync(T->info, info.data(), sizeof(int4) * m, hipMemcpyHostToDevice, ctx->stream));
        DWG_HIP_CHECK(hipMemcpyAsync(T->cm_r, cmr.data(), sizeof(float4) * m, hipMemcpyHostToDevice, ctx->stream));
        DWG_HIP_CHECK(hipMemcpyAsync(T->cm_lo, cml.data(), sizeof(float4) * m, hipMemcpyHostToDevice, ctx->stream));
        DWG_HIP_CHECK(hipMemcpyAsync(T->perm, H.perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
        // positions in tree order (x, y, z, s = 0.0015 until the first update)
        std::vector<float4> sa(n);
        for (int64_t i = 0; i < n; ++i) {
            const int p = H.perm[i];
            sa[i] = make_float4(float(H.P[3 * p]), float(H.P[3 * p + 1]), float(H.P[3 * p + 2]), kScreenDefault);
        }
        DWG_HIP_CHECK(hipMemcpyAsync(T->sa, sa.data(), sizeof(float4) * n, hipMemcpyHostToDevice, ctx->stream));
        DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        delete T;
        throw;
    }
    return T;
}

__global__ void set_radius_kernel(float4* __restrict__ sa, const float* __restrict__ radii, const int* __restrict__ perm,
                                  int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sa[i].w = radii ? radii[perm[i]] : kScreenDefault;
}

void octree_update(dwg_ctx* ctx, dwg_octree* T, const float* nrm, const float* area, const float* radii) {
    const int64_t n = T->n_points, m = T->n_nodes;
    {
        ProfileScope ps(ctx, "octree_pack");
        hipLaunchKernelGGL(set_radius_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, T->sa, radii, T->perm, n);
        check_launch();
        hipLaunchKernelGGL(pack_sources_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, nullptr, nrm, area,
                           nullptr, T->perm, n, 0, nullptr, T->sb, T->an64);
        check_launch();
    }
    {
        ProfileScope ps(ctx, "octree_ec_scan");
        exclusive_scan<Double3>(ctx, T->an64, n);
    }
    ProfileScope ps(ctx, "octree_ec");
    hipLaunchKernelGGL(node_ec_kernel, dim3(grid_for(m, 256)), dim3(256), 0, ctx->stream, T->info, T->an64, m, T->ec,
                       T->ec_bf16);
    check_launch();
    T->updated = true;
}

// ---------------------------------------------------------------------------
// Barnes-Hut traversal, one wave per 64 queries
// ---------------------------------------------------------------------------
constexpr int kTreeThreads = 256;
constexpr int kTreeWaves = kTreeThreads / 64;
constexpr int kStack = 7 * 21 + 8;

struct TreeArgs {
    const int4* __restrict__ info;
    const float4* __restrict__ cm_r;
    const float4* __restrict__ cm_lo;
    const float4* __restrict__ ec;
    const uint2* __restrict__ ec_bf16;
    const float4* __restrict__ sa;
    const float4* __restrict__ sb;
};

struct QueryArgs {
    const float* q;  // explicit points (nullptr → grid)
    int64_t nq;
    float ox, oy, oz, h;
    int rx, ry, rz;
    int bx, by, bz;  // bricks per axis (4×4×4 points per brick)
    int64_t n_bricks;
};

__device__ __forceinline__ float bf16_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t u) { return __uint_as_float(u & 0xFFFF0000u); }

// Block index remap so each XCD (blocks are dealt round-robin over 8 XCDs)
// gets a contiguous run of spatially adjacent bricks / query chunks.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
    const int64_t per = nb / 8;
    if (b >= per * 8) return b;  // tail keeps its slot
    return (b % 8) * per + b / 8;
}

template <bool SCREENED, bool COUNT, bool BF16, bool GRID>
__global__ __launch_bounds__(kTreeThreads) void gwn_tree_kernel(TreeArgs T, QueryArgs Q, float beta, float sigma,
                                                                 uint32_t flags, float* __restrict__ out,
                                                                 int64_t* __restrict__ terms_out) {
    __shared__ int s_node[kTreeWaves][kStack];
    __shared__ uint64_t s_mask[kTreeWaves][kStack];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t wid = blk * kTreeWaves + wave;
    int64_t qi;
    bool valid;
    float qx, qy, qz;
    if (!GRID) {
        qi = wid * 64 + lane;
        valid = qi < Q.nq;
        qx = valid ? Q.q[3 * qi] : 0.f;
        qy = valid ? Q.q[3 * qi + 1] : 0.f;
        qz = valid ? Q.q[3 * qi + 2] : 0.f;
    } else {
        const int64_t br = wid;
        const int bxi = int(br % Q.bx), byi = int((br / Q.bx) % Q.by), bzi = int(br / (int64_t(Q.bx) * Q.by));
        const int x = bxi * 4 + (lane & 3), y = byi * 4 + ((lane >> 2) & 3), z = bzi * 4 + (lane >> 4);
        valid = br < Q.n_bricks && x < Q.rx && y < Q.ry && z < Q.rz;
        qi = int64_t(x) + int64_t(Q.rx) * (int64_t(y) + int64_t(Q.ry) * z);
        qx = __builtin_fmaf(float(x), Q.h, Q.ox);
        qy = __builtin_fmaf(float(y), Q.h, Q.oy);
        qz = __builtin_fmaf(float(z), Q.h, Q.oz);
    }
    const uint64_t all = __ballot(valid);
    if (all == 0) return;
    const float sl2e = sigma * kLog2e;
    float sum = 0.f, comp = 0.f;
    int64_t nterms = 0;
    const bool kahan = flags & DWG_KAHAN;

    auto leaf_sum = [&](int b, int e, bool act) {
        float part = 0.f;
        for (int j = b; j < e; ++j) {
            const float t = gwn_term<SCREENED>(qx, qy, qz, T.sa[j], T.sb[j], sigma, sl2e);
            part += act ? t : 0.f;
        }
        if (COUNT && act) nterms += e - b;
        if (kahan) kahan_add(sum, comp, part); else sum += part;
    };

    int sp = 0;
    const int4 root = T.info[0];
    if (root.x < 0) {
        leaf_sum(root.y, root.z, valid);
    } else {
        if (lane == 0) {
            s_node[wave][0] = 0;
            s_mask[wave][0] = all;
        }
        sp = 1;
    }
    while (sp > 0) {
        --sp;
        const int node = uniform(s_node[wave][sp]);
        const uint64_t mask = uniform64(s_mask[wave][sp]);
        const bool act = (mask >> lane) & 1;
        const int fc = uniform(T.info[node].x);
        float part = 0.f;
        // the 8 children are contiguous: fetch their records together
        int4 nfs[8];
        float4 crs[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) nfs[c] = T.info[fc + c];
#pragma unroll
        for (int c = 0; c < 8; ++c) crs[c] = T.cm_r[fc + c];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int ch = fc + c;
            const int4 nf = nfs[c];
            if (nf.z == nf.y) continue;
            if (nf.x < 0) {  // leaf child: summed point by point whether near or far
                for (int j = nf.y; j < nf.z; ++j) {
                    const float t = gwn_term<SCREENED>(qx, qy, qz, T.sa[j], T.sb[j], sigma, sl2e);
                    part += act ? t : 0.f;
                }
                if (COUNT && act) nterms += nf.z - nf.y;
                continue;
            }
            const float4 cr = crs[c];
            const float dx = cr.x - qx, dy = cr.y - qy, dz = cr.z - qz;
            const float br = beta * cr.w;
            const bool far = dist2(dx, dy, dz) > br * br;
            const bool near = act && !far;
            const uint64_t nm = __ballot(near);
            if (__ballot(act && far)) {
                float4 e4;
                if (BF16) {
                    const uint2 u = T.ec_bf16[ch];
                    e4 = make_float4(bf16_lo(u.x), bf16_hi(u.x), bf16_lo(u.y), 0.f);
                } else {
                    e4 = T.ec[ch];
                }
                // d = (CM_f − q) + δ: the first difference is exact for nearby
                // values, so d carries a single rounding
                const float4 lo = T.cm_lo[ch];
                const float t = gwn_term_d<SCREENED>((cr.x - qx) + lo.x, (cr.y - qy) + lo.y, (cr.z - qz) + lo.z,
                                                     kScreenDefault, e4, sigma, sl2e);
                // the plain dipole has no r > 1e-10 guard in the reference either
                part += (act && far) ? t : 0.f;
                if (COUNT && act && far) nterms += 1;
            }
            if (nm) {
                if (lane == 0) {
                    s_node[wave][sp] = ch;
                    s_mask[wave][sp] = nm;
                }
                ++sp;
            }
        }
        if (kahan) kahan_add(sum, comp, part); else sum += part;
    }
    if (valid) {
        if (out) out[qi] = sum - comp;
        if (COUNT) terms_out[qi] = nterms;
    }
}

static QueryArgs make_points(const float* q, int64_t nq) {
    QueryArgs Q{};
    Q.q = q;
    Q.nq = nq;
    return Q;
}

static QueryArgs make_grid(const dwg_grid* g) {
    QueryArgs Q{};
    Q.q = nullptr;
    Q.ox = g->origin[0];
    Q.oy = g->origin[1];
    Q.oz = g->origin[2];
    Q.h = g->h;
    Q.rx = g->res[0];
    Q.ry = g->res[1];
    Q.rz = g->res[2];
    Q.bx = (Q.rx + 3) / 4;
    Q.by = (Q.ry + 3) / 4;
    Q.bz = (Q.rz + 3) / 4;
    Q.n_bricks = int64_t(Q.bx) * Q.by * Q.bz;
    Q.nq = int64_t(Q.rx) * Q.ry * Q.rz;
    return Q;
}

static void launch_tree(dwg_ctx* ctx, const dwg_octree* T, const QueryArgs& Q, float beta, float sigma, uint32_t flags,
                        float* out, int64_t* terms, bool count) {
    DWG_REQUIRE(T && T->updated, "octree has no moments yet: call dwg_octree_update first");
    DWG_REQUIRE(T->max_level <= 21, "octree deeper than the traversal stack");
    const int64_t waves = Q.q ? (Q.nq + 63) / 64 : Q.n_bricks;
    if (waves == 0) return;
    TreeArgs A{T->info, T->cm_r, T->cm_lo, T->ec, T->ec_bf16, T->sa, T->sb};
    dim3 grid(grid_for(waves, kTreeWaves)), block(kTreeThreads);
    const bool scr = flags & DWG_SCREENED, bf = flags & DWG_FAR_BF16;
    const char* name = count ? "gwn_tree_terms" : (Q.q ? (scr ? "gwn_tree_screened" : "gwn_tree_plain")
                                                       : (scr ? "gwn_tree_grid_screened" : "gwn_tree_grid_plain"));
    ProfileScope ps(ctx, name);
#define DWG_TREE_LAUNCH(S, C, B, G) \
    hipLaunchKernelGGL((gwn_tree_kernel<S, C, B, G>), grid, block, 0, ctx->stream, A, Q, beta, sigma, flags, out, terms)
#define DWG_TREE_LAUNCH_Q(S, C, B)       \
    do {                                 \
        if (Q.q)                         \
            DWG_TREE_LAUNCH(S, C, B, false); \
        else                             \
            DWG_TREE_LAUNCH(S, C, B, true);  \
    } while (0)
    if (count) DWG_TREE_LAUNCH_Q(false, true, false);
    else if (scr && bf) DWG_TREE_LAUNCH_Q(true, false, true);
    else if (scr) DWG_TREE_LAUNCH_Q(true, false, false);
    else if (bf) DWG_TREE_LAUNCH_Q(false, false, true);
    else DWG_TREE_LAUNCH_Q(false, false, false);
#undef DWG_TREE_LAUNCH_Q
#undef DWG_TREE_LAUNCH
    check_launch();
}

void gwn_tree(dwg_ctx* ctx, const dwg_octree* T, const float* q, int64_t nq, float beta, float sigma, uint32_t flags,
              float* out) {
    DWG_REQUIRE(nq >= 0 && (q || nq == 0), "bad query array");
    launch_tree(ctx, T, make_points(q, nq), beta, sigma, flags, out, nullptr, false);
}

void gwn_tree_grid(dwg_ctx* ctx, const dwg_octree* T, const dwg_grid* g, float beta, float sigma, uint32_t flags,
                   float* out) {
    DWG_REQUIRE(g && g->res[0] > 0 && g->res[1] > 0 && g->res[2] > 0 && g->h > 0, "bad grid");
    launch_tree(ctx, T, make_grid(g), beta, sigma, flags, out, nullptr, false);
}

void gwn_tree_terms(dwg_ctx* ctx, const dwg_octree* T, const float* q, int64_t nq, float beta, int64_t* terms) {
    launch_tree(ctx, T, make_points(q, nq), beta, 0.f, 0, nullptr, terms, true);
}

__global__ void sum_i64_kernel(const int64_t* __restrict__ v, int64_t n, unsigned long long* __restrict__ out) {
    __shared__ unsigned long long lds[256];
    unsigned long long s = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) s += v[i];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, lds[0]);
}

int64_t gwn_tree_grid_terms(dwg_ctx* ctx, const dwg_octree* T, const dwg_grid* g, float beta) {
    QueryArgs Q = make_grid(g);
    int64_t* terms = static_cast<int64_t*>(ctx->buf("grid_terms", sizeof(int64_t) * (Q.nq + 1)));
    unsigned long long* tot = reinterpret_cast<unsigned long long*>(terms + Q.nq);
    fill_u32(ctx, tot, 0, 2);
    launch_tree(ctx, T, Q, beta, 0.f, 0, nullptr, terms, true);
    hipLaunchKernelGGL(sum_i64_kernel, dim3(1024), dim3(256), 0, ctx->stream, terms, Q.nq, tot);
    check_launch();
    return int64_t(read_scalar<unsigned long long>(ctx, tot));
}

}  // namespace dwg

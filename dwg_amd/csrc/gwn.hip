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
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "gwn.h"
#include "scan.h"

namespace dwg {

// ---------------------------------------------------------------------------
// per-pair terms
// ---------------------------------------------------------------------------
// term from the offset d = p − q; s = screening radius of the source
// term from the offset d and a squared distance r2 (normally |d|²)
template <bool SCREENED>
__device__ __forceinline__ float gwn_term_r2(float dx, float dy, float dz, float r2, float s, const float4 b,
                                             float sigma, float sigma_l2e) {
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

// max of two non-negative floats on their bit patterns (the integer order is
// the float order there): one instruction, no NaN quieting of the operands
__device__ __forceinline__ float max_nonneg(float a, float b) {
    return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
}

// The tree walk's point and dipole terms in two factors, t = x · inv, from
// d, r² = |d|² and s²: ρ² = max(r², s²), ρ = √ρ², ρ³ = ρ²·ρ; the screened
// guard r > 1e-10 as r² > 1e-20
template <bool SCREENED>
__device__ __forceinline__ float tree_term_split(float dx, float dy, float dz, float r2, float s2, const float4 b,
                                                 float sigma, float sigma_l2e, float& inv, bool& ok) {
    const float dot = __builtin_fmaf(dz, b.z, __builtin_fmaf(dy, b.y, dx * b.x));
    if constexpr (!SCREENED) {
        const float r = fast_sqrt(r2);
        inv = fast_rcp(__builtin_fmaf(r2, r, kEpsPlain * kInv4Pi));
        ok = true;
        return dot;
    } else {
        const float rho2 = max_nonneg(r2, s2);
        const float rho = fast_sqrt(rho2);
        inv = fast_rcp(__builtin_fmaf(fast_exp2(sigma_l2e * rho), rho2 * rho, kEpsScreened * kInv4Pi));
        ok = r2 > 1e-20f;
        return __builtin_fmaf(sigma, rho, 1.0f) * dot;
    }
}

template <bool SCREENED>
__device__ __forceinline__ float gwn_term_d(float dx, float dy, float dz, float s, const float4 b, float sigma,
                                            float sigma_l2e) {
    return gwn_term_r2<SCREENED>(dx, dy, dz, dist2(dx, dy, dz), s, b, sigma, sigma_l2e);
}

template <bool SCREENED>
__device__ __forceinline__ float gwn_term(float qx, float qy, float qz, const float4 a, const float4 b,
                                          float sigma, float sigma_l2e) {
    return gwn_term_d<SCREENED>(a.x - qx, a.y - qy, a.z - qz, a.w, b, sigma, sigma_l2e);
}

// DWG_NEAR_FP64: a screened point term inside its source's screening radius
// (r < s, so ρ = s) in FP64, from the float source a = (p, s), b = A·n/4π
// and query q; 0 unless r > 1e-10 (TurboLibiGWN.cu:544-585)
__device__ __forceinline__ double near_term64(const float4 a, const float4 b, float qx, float qy, float qz,
                                              float sigma) {
    const double dx = double(a.x) - double(qx), dy = double(a.y) - double(qy), dz = double(a.z) - double(qz);
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 > 1e-20)) return 0.0;
    const double rho = a.w, sg = sigma;
    const double dot = dx * double(b.x) + dy * double(b.y) + dz * double(b.z);
    const double den = exp(sg * rho) * (rho * rho * rho) + 1e-7 / (4.0 * 3.14159265358979323846);
    return (sg * rho + 1.0) * dot / den;
}

// the split of a screened point term between the FP32 sum and the FP64 one:
// near (r < s, r > 1e-10) goes to FP64
__device__ __forceinline__ bool near_field(float r2, float s) { return r2 < s * s && r2 > 1e-20f; }

// a term and its gradient w.r.t. q from the offset d = p − q (b pre-scaled
// by 1/4π, ε' = ε/4π):
//   plain     w = (d·b)/D, D = r³ + ε';  ∇w = −b/D + 3 r (d·b) d / D²
//   screened  w = g(ρ)(d·b), g = (1 + σρ)/D, D = e^{σρ}ρ³ + ε', ρ = max(r, s),
//             0 unless r > 1e-10;  ∇w = −g b − (d·b) g'(ρ) d / r for r ≥ s
//             (∂ρ/∂q = −d/r), −g b for r < s;
//             g' = (σε' − e^{σρ}ρ²(σ²ρ² + 3σρ + 3)) / D²  (no cancellation)
template <bool SCREENED>
__device__ __forceinline__ void gwn_term_grad_d(float dx, float dy, float dz, float s, const float4 b, float sigma,
                                                float sigma_l2e, float& w, float& gx, float& gy, float& gz) {
    const float r2 = dist2(dx, dy, dz);
    const float dot = __builtin_fmaf(dz, b.z, __builtin_fmaf(dy, b.y, dx * b.x));
    const float r = fast_sqrt(r2);
    if constexpr (!SCREENED) {
        const float inv = fast_rcp(__builtin_fmaf(r2, r, kEpsPlain * kInv4Pi));
        w = dot * inv;
        const float c = 3.0f * r * w * inv;
        gx = __builtin_fmaf(c, dx, -b.x * inv);
        gy = __builtin_fmaf(c, dy, -b.y * inv);
        gz = __builtin_fmaf(c, dz, -b.z * inv);
    } else {
        const bool ok = r > 1e-10f;
        const bool inside = s > r;  // ρ = s: no dependence on q through ρ
        const float rho = inside ? s : r;
        const float er2 = fast_exp2(sigma_l2e * rho) * (rho * rho);
        const float eps = kEpsScreened * kInv4Pi;
        const float inv = fast_rcp(__builtin_fmaf(er2, rho, eps));
        const float sr = sigma * rho;
        const float g = __builtin_fmaf(sigma, rho, 1.0f) * inv;
        const float gp = (sigma * eps - er2 * __builtin_fmaf(sr, sr + 3.0f, 3.0f)) * inv * inv;
        const float c = (inside || !ok) ? 0.0f : -dot * gp * fast_rcp(r);
        w = ok ? g * dot : 0.0f;
        gx = ok ? __builtin_fmaf(c, dx, -g * b.x) : 0.0f;
        gy = ok ? __builtin_fmaf(c, dy, -g * b.y) : 0.0f;
        gz = ok ? __builtin_fmaf(c, dz, -g * b.z) : 0.0f;
    }
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

template <bool SCREENED, bool KAHAN, bool NEAR64 = false>
__global__ __launch_bounds__(kDirectThreads) void gwn_direct_kernel(
    const float4* __restrict__ sa, const float4* __restrict__ sb, int64_t n, const float* __restrict__ q,
    int64_t nq, float sigma, float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * kDirectThreads + threadIdx.x;
    const bool valid = i < nq;
    const float qx = valid ? q[3 * i] : 0.f, qy = valid ? q[3 * i + 1] : 0.f, qz = valid ? q[3 * i + 2] : 0.f;
    const float sl2e = sigma * kLog2e;
    float sum = 0.f, comp = 0.f;
    double near = 0.0;  // DWG_NEAR_FP64: the terms inside the screening radius, in FP64
    auto term = [&](const float4 a, const float4 b) {
        if constexpr (NEAR64) {
            const float dx = a.x - qx, dy = a.y - qy, dz = a.z - qz;
            const bool nr = near_field(dist2(dx, dy, dz), a.w);
            if (__builtin_amdgcn_ballot_w64(nr)) {
                if (nr) near += near_term64(a, b, qx, qy, qz, sigma);
            }
            return nr ? 0.f : gwn_term<SCREENED>(qx, qy, qz, a, b, sigma, sl2e);
        }
        return gwn_term<SCREENED>(qx, qy, qz, a, b, sigma, sl2e);
    };
    int64_t j0 = 0;
    for (; j0 + kDirectTile <= n; j0 += kDirectTile) {
        float part = 0.f;
#pragma unroll 8
        for (int j = 0; j < kDirectTile; ++j) part += term(sa[j0 + j], sb[j0 + j]);
        if (KAHAN) kahan_add(sum, comp, part); else sum += part;
    }
    float part = 0.f;
    for (int64_t j = j0; j < n; ++j) part += term(sa[j], sb[j]);
    if (KAHAN) kahan_add(sum, comp, part); else sum += part;
    if (valid) out[i] = NEAR64 ? float(double(sum - comp) + near) : sum - comp;
}

template <bool SCREENED, bool KAHAN>
__global__ __launch_bounds__(kDirectThreads) void gwn_direct_grad_kernel(
    const float4* __restrict__ sa, const float4* __restrict__ sb, int64_t n, const float* __restrict__ q,
    int64_t nq, float sigma, float* __restrict__ out_w, float* __restrict__ out_g) {
    const int64_t i = int64_t(blockIdx.x) * kDirectThreads + threadIdx.x;
    const bool valid = i < nq;
    const float qx = valid ? q[3 * i] : 0.f, qy = valid ? q[3 * i + 1] : 0.f, qz = valid ? q[3 * i + 2] : 0.f;
    const float sl2e = sigma * kLog2e;
    float s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    for (int64_t j0 = 0; j0 < n; j0 += kDirectTile) {
        const int64_t j1 = j0 + kDirectTile < n ? j0 + kDirectTile : n;
        float p[4] = {0, 0, 0, 0};
        for (int64_t j = j0; j < j1; ++j) {
            float w, gx, gy, gz;
            const float4 a = sa[j];
            gwn_term_grad_d<SCREENED>(a.x - qx, a.y - qy, a.z - qz, a.w, sb[j], sigma, sl2e, w, gx, gy, gz);
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

static void check_direct_args(int64_t n, int64_t nq, const float* radii, uint32_t flags, const float*) {
    DWG_REQUIRE(n >= 0 && nq >= 0, "negative size");
    DWG_REQUIRE(!(flags & DWG_USE_RADII) || radii, "DWG_USE_RADII needs radii");
}

// packs the n sources into sa/sb (room for n + 1 each) and runs the direct
// kernel; `profile` off records no events (the allocation-free entry point)
static void direct_sums(dwg_ctx* ctx, float4* sa, float4* sb, const float* pts, const float* nrm, const float* area,
                        const float* radii, int64_t n, const float* q, int64_t nq, float sigma, uint32_t flags,
                        float* out_w, float* out_g, bool profile) {
    if (n > 0) {
        ProfileScope ps(profile ? ctx : nullptr, "pack_sources");
        hipLaunchKernelGGL(pack_sources_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, pts, nrm, area,
                           radii, nullptr, n, (flags & DWG_USE_RADII) ? 1 : 0, sa, sb, nullptr);
        check_launch();
    }
    const bool kahan = flags & DWG_KAHAN;
    dim3 grid(grid_for(nq, kDirectThreads)), block(kDirectThreads);
    ProfileScope ps(profile ? ctx : nullptr,
                    out_g ? "gwn_direct_grad" : ((flags & DWG_SCREENED) ? "gwn_direct_screened" : "gwn_direct_plain"));
    if (out_g) {
        const bool scr = flags & DWG_SCREENED;
#define DWG_GRAD_LAUNCH(S, K) \
    hipLaunchKernelGGL((gwn_direct_grad_kernel<S, K>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w, out_g)
        if (scr && kahan) DWG_GRAD_LAUNCH(true, true);
        else if (scr) DWG_GRAD_LAUNCH(true, false);
        else if (kahan) DWG_GRAD_LAUNCH(false, true);
        else DWG_GRAD_LAUNCH(false, false);
#undef DWG_GRAD_LAUNCH
    } else if ((flags & DWG_SCREENED) && (flags & DWG_NEAR_FP64)) {
        if (kahan) hipLaunchKernelGGL((gwn_direct_kernel<true, true, true>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w);
        else hipLaunchKernelGGL((gwn_direct_kernel<true, false, true>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w);
    } else if (flags & DWG_SCREENED) {
        if (kahan) hipLaunchKernelGGL((gwn_direct_kernel<true, true>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w);
        else hipLaunchKernelGGL((gwn_direct_kernel<true, false>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w);
    } else {
        if (kahan) hipLaunchKernelGGL((gwn_direct_kernel<false, true>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w);
        else hipLaunchKernelGGL((gwn_direct_kernel<false, false>), grid, block, 0, ctx->stream, sa, sb, n, q, nq, sigma, out_w);
    }
    check_launch();
}

void gwn_direct(dwg_ctx* ctx, const float* pts, const float* nrm, const float* area, const float* radii, int64_t n,
                const float* q, int64_t nq, float sigma, uint32_t flags, float* out_w, float* out_g) {
    check_direct_args(n, nq, radii, flags, out_g);
    if (nq == 0) return;
    float4* sa = static_cast<float4*>(ctx->buf("direct_sa", sizeof(float4) * (n + 1)));
    float4* sb = static_cast<float4*>(ctx->buf("direct_sb", sizeof(float4) * (n + 1)));
    direct_sums(ctx, sa, sb, pts, nrm, area, radii, n, q, nq, sigma, flags, out_w, out_g, true);
}

// ---------------------------------------------------------------------------
// allocation-free direct sums: the packed sources live in a fixed device
// array instead of the context's growable buffers
// ---------------------------------------------------------------------------
// sa of source i at scratch[i], sb at scratch[kNoAllocSources + 1 + i]
// (8 MiB on every device the library runs on, held until the process ends)
static __device__ __align__(256) float4 scratch[2 * (kNoAllocSources + 1)];

namespace {
std::mutex g_scratch_mu;
std::vector<float4*> g_scratch_ptr;  // device address of `scratch` per device ordinal
}  // namespace

void noalloc_init(int device) {
    // the symbol's address differs per device: look it up with that device
    // current, once, and keep it (host code never uses the symbol itself)
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) throw Error(DWG_ERR_ARG, "bad device ordinal");
    if (size_t(device) < g_scratch_ptr.size() && g_scratch_ptr[size_t(device)]) return;
    int cur = -1;
    DWG_HIP_CHECK(hipGetDevice(&cur));
    DWG_REQUIRE(cur == device, "noalloc_init: the device must be current");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(DWG_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    if (g_scratch_ptr.size() <= size_t(device)) g_scratch_ptr.resize(size_t(device) + 1, nullptr);
    g_scratch_ptr[size_t(device)] = static_cast<float4*>(p);
}

void gwn_direct_noalloc(dwg_ctx* ctx, const float* pts, const float* nrm, const float* area, const float* radii,
                        int64_t n, const float* q, int64_t nq, float sigma, uint32_t flags, float* out_w,
                        float* out_g) {
    check_direct_args(n, nq, radii, flags, out_g);
    if (n > kNoAllocSources)
        throw Error(DWG_ERR_ARG, "dwg_gwn_direct_noalloc: " + std::to_string(n) + " sources exceed its limit of " +
                                     std::to_string(kNoAllocSources) + " (use dwg_gwn_direct)");
    if (nq == 0) return;
    float4* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (size_t(ctx->device) < g_scratch_ptr.size()) base = g_scratch_ptr[size_t(ctx->device)];
    }
    if (!base) throw Error(DWG_ERR_STATE, "allocation-free scratch of this device was not set up");
    direct_sums(ctx, base, base + (kNoAllocSources + 1), pts, nrm, area, radii, n, q, nq, sigma, flags, out_w, out_g,
                false);
}

}  // namespace dwg

// ---------------------------------------------------------------------------
// device octree
// ---------------------------------------------------------------------------
dwg_octree::~dwg_octree() {
    for (void* p : {(void*)info, (void*)kid, (void*)cm_r, (void*)cm_lo, (void*)ec, (void*)sa, (void*)sb, (void*)perm,
                    (void*)an64, (void*)cell, (void*)cm64, (void*)payload, (void*)radius})
        dwg::dev_free(p);
}

namespace dwg {

// EC/4π per node (FP64 prefix-sum difference, one rounding), into ec and
// A leaf with one point is walked like a far dipole: its traversal record
// holds the point (x, y, z), its payload the point's (A·n/4π, s²) and a zero
// residual, so the walk reads the same words for both and evaluates the
// same term, ρ² = max(r², s²) (an internal node's s² is the default
// screening radius's)
constexpr float kDipoleS2 = kScreenDefault * kScreenDefault;

__device__ __forceinline__ void write_point_leaf(int64_t i, const float4 a, const float4 b, float4* __restrict__ cm_r,
                                                 float4* __restrict__ payload) {
    payload[2 * i] = make_float4(b.x, b.y, b.z, a.w * a.w);
    payload[2 * i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    cm_r[i].x = a.x;
    cm_r[i].y = a.y;
    cm_r[i].z = a.z;
}

// the node payload: words 0-3 of an internal node (EC/4π, s²); a leaf with
// one point carries that point's term instead (write_point_leaf)
__global__ void node_ec_kernel(const int4* __restrict__ info, const Double3* __restrict__ scan, int64_t m,
                               const float4* __restrict__ sa, const float4* __restrict__ sb,
                               float4* __restrict__ ec, float4* __restrict__ payload, float4* __restrict__ cm_r) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int4 nf = info[i];
    const Double3 s = scan[nf.z] - scan[nf.y];
    const float x = float(s.x * double(kInv4Pi)), y = float(s.y * double(kInv4Pi)), z = float(s.z * double(kInv4Pi));
    ec[i] = make_float4(x, y, z, 0.f);
    if (nf.x >= 0) {
        payload[2 * i] = make_float4(x, y, z, kDipoleS2);
    } else if (nf.z - nf.y == 1) {
        write_point_leaf(i, sa[nf.y], sb[nf.y], cm_r, payload);
    }
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
    hipLaunchKernelGGL(node_ec_kernel, dim3(grid_for(m, 256)), dim3(256), 0, ctx->stream, T->info, T->an64, m, T->sa,
                       T->sb, T->ec, T->payload, T->cm_r);
    check_launch();
    T->payload_bf16 = false;
    T->updated = true;
}

// the point terms of one-point leaves live in their payload too
__global__ void leaf_payload_kernel(const int4* __restrict__ info, int64_t m, const float4* __restrict__ sa,
                                    const float4* __restrict__ sb, float4* __restrict__ payload,
                                    float4* __restrict__ cm_r) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int4 nf = info[i];
    if (nf.x < 0 && nf.z - nf.y == 1) write_point_leaf(i, sa[nf.y], sb[nf.y], cm_r, payload);
}

void octree_set_point_areas(dwg_ctx* ctx, dwg_octree* T, const float* nrm, const float* area) {
    DWG_REQUIRE(T->updated, "octree has no moments yet: call dwg_octree_update first");
    const int64_t n = T->n_points, m = T->n_nodes;
    ProfileScope ps(ctx, "octree_pack");
    hipLaunchKernelGGL(pack_sources_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, nullptr, nrm, area,
                       nullptr, T->perm, n, 0, nullptr, T->sb, nullptr);
    check_launch();
    hipLaunchKernelGGL(leaf_payload_kernel, dim3(grid_for(m, 256)), dim3(256), 0, ctx->stream, T->info, m, T->sa,
                       T->sb, T->payload, T->cm_r);
    check_launch();
}

// ---------------------------------------------------------------------------
// Barnes-Hut traversal, one wave per brick of grid points (kGridParts per
// lane) or per 64 explicit queries
// ---------------------------------------------------------------------------
constexpr int kTreeThreads = 64;
constexpr int kTreeWaves = kTreeThreads / 64;
constexpr int kMaxTreeLevel = 21;
// stack entries a walk of a tree of max_level levels needs at most: 7 left
// over per level and the 8 children of the deepest node opened
__host__ __device__ constexpr int tree_stack_entries(int max_level) { return 7 * max_level + 8; }
constexpr int kStack = tree_stack_entries(kMaxTreeLevel);  // the gradient walk's fixed stack
// grid walks: each lane evaluates kGridParts points, one in each 4×4×4 part
// of a brick of kPartsX × kPartsY × kPartsZ parts
constexpr int kPartsX = 2, kPartsY = 2, kPartsZ = 1;
constexpr int kGridParts = kPartsX * kPartsY * kPartsZ;
constexpr int kPointPts = 1;  // explicit queries per lane

struct TreeArgs {
    const int4* __restrict__ info;
    const int2* __restrict__ kid;
    const float4* __restrict__ cm_r;
    const float* __restrict__ payload;  // 8 words per node (see dwg_octree::payload)
    const float4* __restrict__ sa;
    const float4* __restrict__ sb;
};

// Wave-uniform reads of node records go through the scalar data cache: a
// pointer into the constant address space makes each such load an s_load
// (one copy into SGPRs).  Through a generic pointer the compiler emits a
// vector load, which returns 64 copies of the same bytes through the
// vector memory path.
typedef int NativeInt2 __attribute__((ext_vector_type(2)));
typedef float NativeFloat4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int2 scalar_load(const int2* p, int64_t i) {
    const NativeInt2 v = ((const __attribute__((address_space(4))) NativeInt2*)p)[i];
    return make_int2(v.x, v.y);
}

typedef int NativeInt4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 scalar_load(const int4* p, int64_t i) {
    const NativeInt4 v = ((const __attribute__((address_space(4))) NativeInt4*)p)[i];
    return make_int4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ float4 scalar_load(const float4* p, int64_t i) {
    const NativeFloat4 v = ((const __attribute__((address_space(4))) NativeFloat4*)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ float bf16_round(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}

struct QueryArgs {
    const float* q;          // explicit points
    const int* self_perm;    // the tree's own points in tree order, results to out[perm[slot]]
    bool grid;               // implicit grid (q and self_perm unused)
    int64_t nq;
    float ox, oy, oz, h;
    int rx, ry, rz;
    int bx, by, bz;  // bricks per axis (kPartsX·4 × kPartsY·4 × kPartsZ·4 points per brick)
    int64_t n_bricks;
    int stack;  // walk stack entries (tree_stack_entries of the tree's depth)
};


template <bool SCREENED, bool COUNT, bool BF16, bool GRID, bool NEAR64 = false>
__global__ __launch_bounds__(kTreeThreads) void gwn_tree_kernel(TreeArgs T, QueryArgs Q, float beta, float sigma,
                                                                 uint32_t flags, float* __restrict__ out,
                                                                 int64_t* __restrict__ terms_out) {
    // grid walks: each lane evaluates kGridParts points (one per 4×4×4 part
    // of the brick), so that a pop, its loads and its pushes serve them all
    constexpr int H = GRID ? kGridParts : kPointPts;
    static_assert(kTreeWaves == 1, "one wave per block: the stack lives in the block's dynamic LDS");
    // the stack: Q.stack entries of (record, H lane masks), in dynamic LDS
    // sized by the tree's depth (a shallow tree leaves room for more blocks)
    extern __shared__ uint64_t s_dyn[];
    int2* s_node_w = reinterpret_cast<int2*>(s_dyn);
    uint64_t* s_mask_w = s_dyn + Q.stack;
    __shared__ float4 s_pay[kTreeWaves][16];  // the 8 children's payloads of the node being opened
    __shared__ float4 s_cmr[kTreeWaves][16];  // and their (CM, w) records (twice)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t wid = blk * kTreeWaves + wave;
    int64_t qi[H];
    bool valid[H];
    float qx[H], qy[H], qz[H];
    if (!GRID) {
        // H consecutive runs of 64 queries
#pragma unroll
        for (int h = 0; h < H; ++h) {
            qi[h] = (wid * H + h) * 64 + lane;
            valid[h] = qi[h] < Q.nq;
            if (Q.self_perm) {  // consecutive tree slots: spatially coherent waves
                const float4 a = valid[h] ? T.sa[qi[h]] : make_float4(0.f, 0.f, 0.f, 0.f);
                qx[h] = a.x;
                qy[h] = a.y;
                qz[h] = a.z;
                qi[h] = valid[h] ? int64_t(Q.self_perm[qi[h]]) : 0;
            } else {
                qx[h] = valid[h] ? Q.q[3 * qi[h]] : 0.f;
                qy[h] = valid[h] ? Q.q[3 * qi[h] + 1] : 0.f;
                qz[h] = valid[h] ? Q.q[3 * qi[h] + 2] : 0.f;
            }
        }
    } else {
        // brick of kPartsX × kPartsY × kPartsZ parts of 4 × 4 × 4 points
        const int64_t br = wid;
        const int bxi = int(br % Q.bx), byi = int((br / Q.bx) % Q.by), bzi = int(br / (int64_t(Q.bx) * Q.by));
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int x = (bxi * kPartsX + h % kPartsX) * 4 + (lane & 3);
            const int y = (byi * kPartsY + (h / kPartsX) % kPartsY) * 4 + ((lane >> 2) & 3);
            const int z = (bzi * kPartsZ + h / (kPartsX * kPartsY)) * 4 + (lane >> 4);
            valid[h] = br < Q.n_bricks && x < Q.rx && y < Q.ry && z < Q.rz;
            qi[h] = int64_t(x) + int64_t(Q.rx) * (int64_t(y) + int64_t(Q.ry) * z);
            qx[h] = __builtin_fmaf(float(x), Q.h, Q.ox);
            qy[h] = __builtin_fmaf(float(y), Q.h, Q.oy);
            qz[h] = __builtin_fmaf(float(z), Q.h, Q.oz);
        }
    }
    uint64_t all[H];
    bool any = false;
#pragma unroll
    for (int h = 0; h < H; ++h) {
        all[h] = __builtin_amdgcn_ballot_w64(valid[h]);
        any = any || all[h] != 0;
    }
    if (!any) return;
    const float sl2e = sigma * kLog2e;
    float sum[H], comp[H];
    int64_t nterms[H];
    double near[H];  // DWG_NEAR_FP64: point terms inside the screening radius, in FP64
#pragma unroll
    for (int h = 0; h < H; ++h) {
        sum[h] = comp[h] = 0.f;
        nterms[h] = 0;
        near[h] = 0.0;
    }
    const bool kahan = flags & DWG_KAHAN;

    auto point_term = [&](const float4 a, const float4 b, bool on, float part, int h) {
        const float dx = a.x - qx[h], dy = a.y - qy[h], dz = a.z - qz[h];
        const float r2 = dist2(dx, dy, dz);
        if constexpr (NEAR64) {
            const bool nr = on && near_field(r2, a.w);
            if (__builtin_amdgcn_ballot_w64(nr)) {  // rare: lanes next to a source
                if (nr) near[h] += near_term64(a, b, qx[h], qy[h], qz[h], sigma);
            }
            on = on && !nr;
        }
        float inv;
        bool ok;
        const float x = tree_term_split<SCREENED>(dx, dy, dz, r2, a.w * a.w, b, sigma, sl2e, inv, ok);
        if (BF16) return __builtin_fmaf((on && ok) ? x : 0.f, inv, part);  // last product fused into the sum
        return part + ((on && ok) ? x * inv : 0.f);
    };
    auto add = [&](float part, int h) {
        if (kahan) kahan_add(sum[h], comp[h], part); else sum[h] += part;
    };
    // a leaf of several points, summed point by point for the parts whose
    // lanes opened it
    auto leaf_sum = [&](int b, int e, const uint64_t* m) {
        float part[H];
#pragma unroll
        for (int h = 0; h < H; ++h) part[h] = 0.f;
        for (int j = b; j < e; ++j) {
            const float4 pa = scalar_load(T.sa, j), pb = scalar_load(T.sb, j);
#pragma unroll
            for (int h = 0; h < H; ++h)
                if (H == 1 || m[h]) part[h] = point_term(pa, pb, __builtin_amdgcn_inverse_ballot_w64(m[h]), part[h], h);
        }
#pragma unroll
        for (int h = 0; h < H; ++h) {
            if (COUNT && __builtin_amdgcn_inverse_ballot_w64(m[h])) nterms[h] += e - b;
            if (H == 1 || m[h]) add(part[h], h);
        }
    };

    // the stack holds (traversal record of a node to open, lanes that open
    // it, per part): an internal node's first child, or a leaf's
    // (−count, begin)
    int sp = 0;
    const int4 root = scalar_load(T.info, 0);
    if (root.x < 0) {
        leaf_sum(root.y, root.z, all);
    } else {
        s_node_w[0] = make_int2(root.x, 0);
#pragma unroll
        for (int h = 0; h < H; ++h) s_mask_w[h] = all[h];
        sp = 1;
    }
    while (sp > 0) {
        --sp;
        const int2 top = s_node_w[sp];
        const int fc = uniform(top.x);
        uint64_t mask[H];
#pragma unroll
        for (int h = 0; h < H; ++h) mask[h] = uniform64(s_mask_w[sp * H + h]);
        if (fc < 0) {  // a leaf of several points: summed point by point
            leaf_sum(uniform(top.y), uniform(top.y) - fc, mask);
            continue;
        }
        float part[H];
#pragma unroll
        for (int h = 0; h < H; ++h) part[h] = 0.f;
        // the 8 children are contiguous: their records come in one batch of
        // scalar loads, their payloads in one coalesced vector load (word
        // 8c + w of child c in lane 8c + w), so a node costs one memory
        // round trip however many of its children are far or leaves
        int2 nks[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) nks[c] = scalar_load(T.kid, fc + c);
        // payload word 8c + w of child c: loaded by lane 8c + w, and (CM, w)
        // word 4c + w by lanes 4c + w and 32 + 4c + w (a second copy keeps the
        // store unconditional); both read back by the whole wave from LDS at
        // uniform addresses (broadcasts).  Records in LDS instead of SGPRs
        // leave the stack pointer room in the scalar registers
        const float pw = T.payload[int64_t(fc) * 8 + lane];
        const float cw = reinterpret_cast<const float*>(T.cm_r)[int64_t(fc) * 4 + (lane & 31)];
        reinterpret_cast<float*>(s_pay[wave])[lane] = pw;
        reinterpret_cast<float*>(s_cmr[wave])[lane] = cw;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int2 nk = nks[c];
            if (nk.x == 0) continue;  // empty
            // one test and one term for every kind of child: an internal
            // node is far when d² > (β·R)² (cm_r.w), a one-point leaf's
            // record makes the test its term's guard, a leaf of several
            // points is never far (∞) and is opened like an internal node
            const float4 cr = s_cmr[wave][c];
            uint64_t nm[H];
            bool push = false;
#pragma unroll
            for (int h = 0; h < H; ++h) {
                nm[h] = 0;
                if (H > 1 && !mask[h]) continue;  // no lane of this part opened the node
                const bool act = __builtin_amdgcn_inverse_ballot_w64(mask[h]);
                const float dx = cr.x - qx[h], dy = cr.y - qy[h], dz = cr.z - qz[h];
                const float r2 = dist2(dx, dy, dz);
                bool far = r2 > cr.w;
                if constexpr (NEAR64) {
                    // point terms inside the screening radius go to the FP64 sum
                    if (nk.x == -1) {
                        const bool nr = act && far && r2 < s_pay[wave][2 * c].w;
                        if (__builtin_amdgcn_ballot_w64(nr)) {  // rare: lanes next to a source
                            if (nr)
                                near[h] += near_term64(scalar_load(T.sa, nk.y), scalar_load(T.sb, nk.y), qx[h], qy[h],
                                                       qz[h], sigma);
                        }
                        far = far && !nr;
                    }
                }
                // lane masks straight from the compare: far for the lanes that
                // opened this node, near ones open the child
                const uint64_t farm = __builtin_amdgcn_ballot_w64(far);
                if (mask[h] & farm) {
                    const float4 e4 = s_pay[wave][2 * c];
                    if (BF16) {
                        // BF16 moments carry a relative error of 2^-9, far above
                        // the rounding of the centre: d = CM_f − q and its |d|²
                        // from the far test serve as they are, and the term's
                        // last product is fused into the sum.  r > 1e-10 holds
                        // for every far node here (far radius floor) and is the
                        // far test of a point.  The lanes that take the term are
                        // selected by the exec mask (mask & farm), not by a
                        // select on the result: one VALU instruction less per term
                        if (__builtin_amdgcn_inverse_ballot_w64(mask[h] & farm)) {
                            float inv;
                            bool ok;
                            const float x = tree_term_split<SCREENED>(dx, dy, dz, r2, e4.w, e4, sigma, sl2e, inv, ok);
                            part[h] = __builtin_fmaf(x, inv, part[h]);
                        }
                    } else {
                        // d = (CM_f − q) + δ: the first difference is exact for
                        // nearby values, so d carries a single rounding (δ = 0
                        // for a point)
                        const float4 lo = s_pay[wave][2 * c + 1];
                        const float ex = dx + lo.x, ey = dy + lo.y, ez = dz + lo.z;
                        float inv;
                        bool ok;
                        const float x =
                            tree_term_split<SCREENED>(ex, ey, ez, dist2(ex, ey, ez), e4.w, e4, sigma, sl2e, inv, ok);
                        // the plain dipole has no r > 1e-10 guard in the reference either
                        part[h] += ((act && far && ok) ? x * inv : 0.f);
                    }
                }
                if (COUNT && act && (far || nk.x == -1)) nterms[h] += 1;
                nm[h] = mask[h] & ~farm;
                push = push || nm[h] != 0;
            }
            if (push && nk.x != -1) {
                // every lane writes the same entry: no divergent branch, so
                // the stack pointer stays in a scalar register
                s_node_w[sp] = nk;
#pragma unroll
                for (int h = 0; h < H; ++h) s_mask_w[sp * H + h] = nm[h];
                ++sp;
            }
        }
#pragma unroll
        for (int h = 0; h < H; ++h)
            if (H == 1 || mask[h]) add(part[h], h);
    }
#pragma unroll
    for (int h = 0; h < H; ++h) {
        if (valid[h]) {
            if (out) out[qi[h]] = NEAR64 ? float(double(sum[h] - comp[h]) + near[h]) : sum[h] - comp[h];
            if (COUNT) terms_out[qi[h]] = nterms[h];
        }
    }
}

// Barnes-Hut value and gradient at query points (new; the reference has no
// gradient): the walk and far test of gwn_tree_kernel (FP32 path: far
// dipoles at (CM_f − q) + residual), each lane accumulating w and ∇w of its
// own term set.
template <bool SCREENED>
__global__ __launch_bounds__(kTreeThreads) void gwn_tree_grad_kernel(TreeArgs T, const float* __restrict__ q,
                                                                      int64_t nq, float sigma,
                                                                      float* __restrict__ out_w,
                                                                      float* __restrict__ out_g) {
    __shared__ int s_node[kTreeWaves][kStack];
    __shared__ uint64_t s_mask[kTreeWaves][kStack];
    __shared__ float4 s_pay[kTreeWaves][16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t qi = (int64_t(blockIdx.x) * kTreeWaves + wave) * 64 + lane;
    const bool valid = qi < nq;
    const float qx = valid ? q[3 * qi] : 0.f, qy = valid ? q[3 * qi + 1] : 0.f, qz = valid ? q[3 * qi + 2] : 0.f;
    const uint64_t all = __builtin_amdgcn_ballot_w64(valid);
    if (all == 0) return;
    const float sl2e = sigma * kLog2e;
    // the terms of one opened node sum into `part`, which folds into `acc`
    // with Kahan compensation: the gradient's near terms are large and
    // cancel, so plain FP32 accumulation over many nodes would lose digits
    float acc[4] = {0.f, 0.f, 0.f, 0.f}, comp[4] = {0.f, 0.f, 0.f, 0.f}, part[4] = {0.f, 0.f, 0.f, 0.f};
    auto add = [&](float dx, float dy, float dz, float sr, const float4 b, bool on) {
        float w, gx, gy, gz;
        gwn_term_grad_d<SCREENED>(dx, dy, dz, sr, b, sigma, sl2e, w, gx, gy, gz);
        part[0] += on ? w : 0.f;
        part[1] += on ? gx : 0.f;
        part[2] += on ? gy : 0.f;
        part[3] += on ? gz : 0.f;
    };
    auto fold = [&]() {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            kahan_add(acc[k], comp[k], part[k]);
            part[k] = 0.f;
        }
    };
    auto leaf = [&](int b, int e, bool on) {
        for (int j = b; j < e; ++j) {
            const float4 a = scalar_load(T.sa, j);
            add(a.x - qx, a.y - qy, a.z - qz, a.w, scalar_load(T.sb, j), on);
        }
    };
    int sp = 0;
    const int4 root = T.info[0];
    if (root.x < 0) {
        leaf(root.y, root.z, valid);
        fold();
    } else {
        if (lane == 0) {
            s_node[wave][0] = root.x;
            s_mask[wave][0] = all;
        }
        sp = 1;
    }
    while (sp > 0) {
        --sp;
        const int fc = uniform(s_node[wave][sp]);
        const uint64_t mask = uniform64(s_mask[wave][sp]);
        const bool act = (mask >> lane) & 1;
        reinterpret_cast<float*>(s_pay[wave])[lane] = T.payload[int64_t(fc) * 8 + lane];
        __builtin_amdgcn_wave_barrier();
        for (int c = 0; c < 8; ++c) {
            const int2 nk = scalar_load(T.kid, fc + c);
            if (nk.x == 0) continue;
            if (nk.x < 0) {
                leaf(nk.y, nk.y - nk.x, act);
                continue;
            }
            const float4 cr = scalar_load(T.cm_r, fc + c);
            const float dx = cr.x - qx, dy = cr.y - qy, dz = cr.z - qz;
            const bool far = dist2(dx, dy, dz) > cr.w;
            const uint64_t farm = __builtin_amdgcn_ballot_w64(far);
            if (mask & farm) {
                const float4 lo = s_pay[wave][2 * c + 1];
                add(dx + lo.x, dy + lo.y, dz + lo.z, kScreenDefault, s_pay[wave][2 * c], act && far);
            }
            const uint64_t nm = mask & ~farm;
            if (nm) {
                if (lane == 0) {
                    s_node[wave][sp] = nk.x;
                    s_mask[wave][sp] = nm;
                }
                ++sp;
            }
        }
        fold();
        __builtin_amdgcn_wave_barrier();
    }
    if (valid) {
        out_w[qi] = acc[0] - comp[0];
        if (out_g) {
            out_g[3 * qi] = acc[1] - comp[1];
            out_g[3 * qi + 1] = acc[2] - comp[2];
            out_g[3 * qi + 2] = acc[3] - comp[3];
        }
    }
}

static QueryArgs make_points(const float* q, int64_t nq) {
    QueryArgs Q{};
    Q.q = q;
    Q.nq = nq;
    return Q;
}

static QueryArgs make_self(const dwg_octree* T) {
    QueryArgs Q{};
    Q.self_perm = T->perm;
    Q.nq = T->n_points;
    return Q;
}

static QueryArgs make_grid(const dwg_grid* g) {
    QueryArgs Q{};
    Q.grid = true;
    Q.ox = g->origin[0];
    Q.oy = g->origin[1];
    Q.oz = g->origin[2];
    Q.h = g->h;
    Q.rx = g->res[0];
    Q.ry = g->res[1];
    Q.rz = g->res[2];
    Q.bx = (Q.rx + 4 * kPartsX - 1) / (4 * kPartsX);
    Q.by = (Q.ry + 4 * kPartsY - 1) / (4 * kPartsY);
    Q.bz = (Q.rz + 4 * kPartsZ - 1) / (4 * kPartsZ);
    Q.n_bricks = int64_t(Q.bx) * Q.by * Q.bz;
    Q.nq = int64_t(Q.rx) * Q.ry * Q.rz;
    return Q;
}

// cm_r.w = (β·R)² for the traversal's far test d² > (β·R)², with the
// rounding sequence float(β·R) then its square (the order the oracle uses),
// clamped below by `floor`.  A leaf with one point gets `point_guard`, so
// that the far test is the point term's own guard (r² > 1e-20 screened,
// none plain: −1); a leaf with several points ∞, so that it is always
// opened and its points summed one by one
__global__ void far_radius_kernel(const int2* __restrict__ kid, const float* __restrict__ r, int64_t m, float beta,
                                  float floor, float point_guard, float4* __restrict__ cm_r) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int k = kid[i].x;
    const float br = beta * r[i];
    cm_r[i].w = k == -1 ? point_guard : k < -1 ? INFINITY : fmaxf(br * br, floor);
}

// Far-radius floor of the mixed-precision walk: with d² > 4e-20 every far
// node has r = |d| > 2e-10, so the screened term's r > 1e-10 guard always
// holds there and the walk drops it.  Only a node whose points all lie within
// 1e-10 of each other (duplicates) has (β·R)² below the floor; it is opened
// instead, and its points' own guarded terms give the same sum.
constexpr float kFarFloorBf16 = 4e-20f;

// Far-field moments in the payload of every internal node: EC/4π in FP32,
// or rounded to BF16 (to nearest even, kept in FP32 words) for BF16
// evaluations
__global__ void payload_precision_kernel(const int2* __restrict__ kid, const float4* __restrict__ ec, int64_t m,
                                         int bf16, float4* __restrict__ payload) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m || kid[i].x <= 0) return;
    const float4 e = ec[i];
    payload[2 * i] = bf16 ? make_float4(bf16_round(e.x), bf16_round(e.y), bf16_round(e.z), kDipoleS2)
                          : make_float4(e.x, e.y, e.z, kDipoleS2);
}

static void launch_tree(dwg_ctx* ctx, const dwg_octree* Tc, const QueryArgs& Q, float beta, float sigma,
                        uint32_t flags, float* out, int64_t* terms, bool count) {
    DWG_REQUIRE(Tc && Tc->updated, "octree has no moments yet: call dwg_octree_update first");
    DWG_REQUIRE(Tc->max_level <= kMaxTreeLevel, "octree deeper than the traversal stack");
    auto* T = const_cast<dwg_octree*>(Tc);  // the far radii follow β; the tree itself is unchanged
    const bool want_bf16 = (flags & DWG_FAR_BF16) && !count;
    const float floor = want_bf16 ? kFarFloorBf16 : 0.f;
    const float guard = (flags & DWG_SCREENED) ? 1e-20f : -1.f;
    if (!(T->far_beta == beta) || T->far_floor != floor || T->far_guard != guard) {
        hipLaunchKernelGGL(far_radius_kernel, dim3(grid_for(T->n_nodes, 256)), dim3(256), 0, ctx->stream, T->kid,
                           T->radius, T->n_nodes, beta, floor, guard, T->cm_r);
        check_launch();
        T->far_beta = beta;
        T->far_floor = floor;
        T->far_guard = guard;
    }
    if (T->payload_bf16 != want_bf16) {
        hipLaunchKernelGGL(payload_precision_kernel, dim3(grid_for(T->n_nodes, 256)), dim3(256), 0, ctx->stream, T->kid,
                           T->ec, T->n_nodes, int(want_bf16), T->payload);
        check_launch();
        T->payload_bf16 = want_bf16;
    }
    const int64_t waves = Q.grid ? Q.n_bricks : (Q.nq + 64 * kPointPts - 1) / (64 * kPointPts);
    if (waves == 0) return;
    TreeArgs A{T->info, T->kid, T->cm_r, reinterpret_cast<const float*>(T->payload), T->sa, T->sb};
    dim3 grid(grid_for(waves, kTreeWaves)), block(kTreeThreads);
    // the walk's stack: one record and one lane mask per part for each entry
    QueryArgs Qs = Q;
    Qs.stack = tree_stack_entries(std::max(T->max_level, 1));
    const size_t lds = sizeof(uint64_t) * size_t(Qs.stack) * size_t(1 + (Q.grid ? kGridParts : kPointPts));
    const bool scr = flags & DWG_SCREENED, bf = flags & DWG_FAR_BF16;
    const char* name = count ? "gwn_tree_terms" : (!Q.grid ? (scr ? "gwn_tree_screened" : "gwn_tree_plain")
                                                           : (scr ? "gwn_tree_grid_screened" : "gwn_tree_grid_plain"));
    ProfileScope ps(ctx, name);
    const bool near64 = scr && (flags & DWG_NEAR_FP64);
#define DWG_TREE_LAUNCH(S, C, B, G, N) \
    hipLaunchKernelGGL((gwn_tree_kernel<S, C, B, G, N>), grid, block, lds, ctx->stream, A, Qs, beta, sigma, flags, out, terms)
#define DWG_TREE_LAUNCH_Q(S, C, B, N)       \
    do {                                    \
        if (!Q.grid)                        \
            DWG_TREE_LAUNCH(S, C, B, false, N); \
        else                                \
            DWG_TREE_LAUNCH(S, C, B, true, N);  \
    } while (0)
    if (count) DWG_TREE_LAUNCH_Q(false, true, false, false);
    else if (near64 && bf) DWG_TREE_LAUNCH_Q(true, false, true, true);
    else if (near64) DWG_TREE_LAUNCH_Q(true, false, false, true);
    else if (scr && bf) DWG_TREE_LAUNCH_Q(true, false, true, false);
    else if (scr) DWG_TREE_LAUNCH_Q(true, false, false, false);
    else if (bf) DWG_TREE_LAUNCH_Q(false, false, true, false);
    else DWG_TREE_LAUNCH_Q(false, false, false, false);
#undef DWG_TREE_LAUNCH_Q
#undef DWG_TREE_LAUNCH
    check_launch();
}

void gwn_tree(dwg_ctx* ctx, const dwg_octree* T, const float* q, int64_t nq, float beta, float sigma, uint32_t flags,
              float* out) {
    DWG_REQUIRE(nq >= 0 && (q || nq == 0), "bad query array");
    launch_tree(ctx, T, make_points(q, nq), beta, sigma, flags, out, nullptr, false);
}

void gwn_tree_self(dwg_ctx* ctx, const dwg_octree* T, float beta, float sigma, uint32_t flags, float* out) {
    DWG_REQUIRE(T, "null octree");
    launch_tree(ctx, T, make_self(T), beta, sigma, flags, out, nullptr, false);
}

void gwn_tree_grid(dwg_ctx* ctx, const dwg_octree* T, const dwg_grid* g, float beta, float sigma, uint32_t flags,
                   float* out) {
    DWG_REQUIRE(g && g->res[0] > 0 && g->res[1] > 0 && g->res[2] > 0 && g->h > 0, "bad grid");
    launch_tree(ctx, T, make_grid(g), beta, sigma, flags, out, nullptr, false);
}

void gwn_tree_grad(dwg_ctx* ctx, const dwg_octree* Tc, const float* q, int64_t nq, float beta, float sigma,
                   uint32_t flags, float* out_w, float* out_g) {
    DWG_REQUIRE(Tc && Tc->updated, "octree has no moments yet: call dwg_octree_update first");
    DWG_REQUIRE(Tc->max_level <= kMaxTreeLevel, "octree deeper than the traversal stack");
    DWG_REQUIRE(nq >= 0 && (q || nq == 0) && (out_w || nq == 0), "bad query or output array");
    auto* T = const_cast<dwg_octree*>(Tc);  // far radii follow β and the payload precision; the tree is unchanged
    const float guard = (flags & DWG_SCREENED) ? 1e-20f : -1.f;
    if (!(T->far_beta == beta) || T->far_floor != 0.f || T->far_guard != guard) {
        hipLaunchKernelGGL(far_radius_kernel, dim3(grid_for(T->n_nodes, 256)), dim3(256), 0, ctx->stream, T->kid,
                           T->radius, T->n_nodes, beta, 0.f, guard, T->cm_r);
        check_launch();
        T->far_beta = beta;
        T->far_floor = 0.f;
        T->far_guard = guard;
    }
    if (T->payload_bf16) {
        hipLaunchKernelGGL(payload_precision_kernel, dim3(grid_for(T->n_nodes, 256)), dim3(256), 0, ctx->stream, T->kid,
                           T->ec, T->n_nodes, 0, T->payload);
        check_launch();
        T->payload_bf16 = false;
    }
    if (nq == 0) return;
    TreeArgs A{T->info, T->kid, T->cm_r, reinterpret_cast<const float*>(T->payload), T->sa, T->sb};
    dim3 grid(grid_for((nq + 63) / 64, kTreeWaves)), block(kTreeThreads);
    ProfileScope ps(ctx, "gwn_tree_grad");
    if (flags & DWG_SCREENED)
        hipLaunchKernelGGL(gwn_tree_grad_kernel<true>, grid, block, 0, ctx->stream, A, q, nq, sigma, out_w, out_g);
    else
        hipLaunchKernelGGL(gwn_tree_grad_kernel<false>, grid, block, 0, ctx->stream, A, q, nq, sigma, out_w, out_g);
    check_launch();
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

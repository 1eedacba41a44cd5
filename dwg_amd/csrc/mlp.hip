// Learned normal initialisation (no reference counterpart): a 48→128→128→3
// MLP with GELU over k-NN patches, trained and run in BF16 with FP32
// accumulation on the matrix cores.
//
// Patches: the 16 nearest neighbours of a point among the patch points
// (exact kNN, (d², index) order; the first neighbour at distance 0 — the
// point itself when it is a patch point — is dropped), offsets p_j − p_i
// divided by the distance to the 16th neighbour, flattened to 48 values in
// distance order.  A cloud larger than the training clouds (50,000 points)
// takes its patch points from a hashed subsample of that size, so patches
// span the same fraction of the shape as in training.
//
// Every product is one launch of gemm_bf16: C = A·B with arbitrary element
// strides for A (M×K), B (K×N) and C (M×N), BF16 operands, FP32
// accumulation, one 32×32 output tile per wave on v_mfma_f32_32x32x16_bf16.
// GELU is the tanh form of Hendrycks & Gimpel (2016).  The loss is the mean
// over the batch of 1 − ŷ·t (oriented) or 1 − (ŷ·t)² (normal line, sign
// free), ŷ = y/|y|, t the unit target normal.  The initialisation orients
// the predicted line away from the bounding-box centre, as the Gauss map
// does (TurboNormalInit.cu:3-44).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>

#include "knn.h"
#include "mlp.h"
#include "scan.h"

dwg_mlp::~dwg_mlp() {
    for (void* p : {(void*)params, (void*)bf16, (void*)grad, (void*)adam_m, (void*)adam_v})
        dwg::dev_free(p);
}

namespace dwg {

typedef short bf16x8 __attribute__((ext_vector_type(8)));    // 8 BF16 bit patterns
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));  // the MFMA operand type

typedef float f32x16 __attribute__((ext_vector_type(16)));

// v_mfma_f32_32x32x16_bf16 (gfx950): 32×32 += 32×16 · 16×32 in one instruction
__device__ __forceinline__ f32x16 mfma_32x32x16(const bf16x8 a, const bf16x8 b, const f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8v, a), __builtin_bit_cast(bf16x8v, b), c,
                                                   0, 0, 0);
}

__device__ __forceinline__ uint16_t to_bf16(float f) {
    // round to nearest even; NaN stays NaN
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return uint16_t((u >> 16) | 0x40);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// ---------------------------------------------------------------------------
// patches
// ---------------------------------------------------------------------------
__global__ void patches_kernel(const int32_t* __restrict__ knn, const float* __restrict__ d2, int kk,
                               const float* __restrict__ ps, const float* __restrict__ q, int64_t n,
                               uint16_t* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int sel[kPatchK];
    int m = 0;
    float last_d2 = 0.f;
    bool self_skipped = false;
    for (int j = 0; j < kk && m < kPatchK; ++j) {
        const int32_t c = knn[i * kk + j];
        if (c < 0) break;
        if (!self_skipped && d2[i * kk + j] == 0.f) {
            self_skipped = true;
            continue;
        }
        sel[m++] = c;
        last_d2 = d2[i * kk + j];
    }
    const float r = last_d2 > 0.f ? sqrtf(last_d2) : 1.f;
    const float inv = 1.0f / r;
    for (int j = 0; j < kPatchK; ++j) {
        float v[3] = {0.f, 0.f, 0.f};
        if (j < m) {
            const int64_t c = sel[j];
            for (int d = 0; d < 3; ++d) v[d] = (ps[3 * c + d] - q[3 * i + d]) * inv;
        }
        for (int d = 0; d < 3; ++d) out[i * kIn + 3 * j + d] = to_bf16(v[d]);
    }
}

// keep point i of the patch subsample when its hashed index falls in the
// first 1/stride of the range
__global__ void subsample_flags_kernel(int64_t n, uint32_t stride, int64_t* __restrict__ flags) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        flags[i] = 0;
        return;
    }
    uint32_t h = uint32_t(i) * 0x9E3779B1u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    flags[i] = (h % stride) == 0 ? 1 : 0;
}

__global__ void subsample_gather_kernel(const int64_t* __restrict__ slot, const float* __restrict__ p, int64_t n,
                                        float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || slot[i + 1] == slot[i]) return;
    const int64_t j = slot[i];
    for (int d = 0; d < 3; ++d) out[3 * j + d] = p[3 * i + d];
}

// orient the predicted line away from c (n·(p − c) ≥ 0)
__global__ void orient_kernel(float* __restrict__ nrm, const float* __restrict__ p, int64_t n, float cx, float cy,
                              float cz) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = nrm[3 * i] * (p[3 * i] - cx) + nrm[3 * i + 1] * (p[3 * i + 1] - cy) +
                    nrm[3 * i + 2] * (p[3 * i + 2] - cz);
    if (d < 0.f)
        for (int k = 0; k < 3; ++k) nrm[3 * i + k] = -nrm[3 * i + k];
}

// ---------------------------------------------------------------------------
// strided BF16 GEMM on the matrix cores
// ---------------------------------------------------------------------------
struct GemmArgs {
    const uint16_t* A;
    int64_t sAi, sAk;
    const uint16_t* B;
    int64_t sBk, sBj;
    float* C;
    int64_t sCi, sCj;
    int64_t M, N, K;
    int accumulate;
};

// Lane layout of v_mfma_f32_32x32x16_bf16 (one wave, 32×32 += 32×16 · 16×32):
//   A: lane l holds A[l%32][8(l/32) .. 8(l/32)+7]
//   B: lane l holds B[8(l/32) .. 8(l/32)+7][l%32]
//   D: lane l, register r holds D[8(r/4) + 4(l/32) + r%4][l%32]
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g, int64_t tiles_n) {
    const int lane = threadIdx.x & 63;
    const int64_t tile = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int64_t tiles_m = (g.M + 31) / 32;
    if (tile >= tiles_m * tiles_n) return;
    const int64_t ti = tile / tiles_n, tj = tile % tiles_n;
    const int64_t row = ti * 32 + (lane & 31);
    const int64_t col = tj * 32 + (lane & 31);
    const int kq = 8 * (lane >> 5);
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    for (int64_t k0 = 0; k0 < g.K; k0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t k = k0 + kq + j;
            a[j] = (row < g.M && k < g.K) ? short(g.A[row * g.sAi + k * g.sAk]) : short(0);
            b[j] = (col < g.N && k < g.K) ? short(g.B[k * g.sBk + col * g.sBj]) : short(0);
        }
        acc = mfma_32x32x16(a, b, acc);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int64_t i = ti * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
        const int64_t j = tj * 32 + (lane & 31);
        if (i < g.M && j < g.N) {
            float* c = g.C + i * g.sCi + j * g.sCj;
            *c = g.accumulate ? *c + acc[r] : acc[r];
        }
    }
}

static void gemm(dwg_ctx* ctx, const GemmArgs& g) {
    const int64_t tn = (g.N + 31) / 32, tm = (g.M + 31) / 32;
    const int64_t tiles = tm * tn;
    if (tiles == 0) return;
    hipLaunchKernelGGL(gemm_bf16_kernel, dim3(grid_for(tiles, 4)), dim3(256), 0, ctx->stream, g, tn);
    check_launch();
}

// ---------------------------------------------------------------------------
// elementwise pieces
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gelu(float x, float* dgelu) {
    const float c = 0.7978845608028654f;  // sqrt(2/pi)
    const float u = c * (x + 0.044715f * x * x * x);
    const float t = tanhf(u);
    if (dgelu) *dgelu = 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * x * x);
    return 0.5f * x * (1.0f + t);
}

// z += bias; h = bf16(gelu(z)); z keeps the pre-activation for the backward pass
__global__ void bias_gelu_kernel(float* __restrict__ z, const float* __restrict__ bias, int64_t rows, int cols,
                                 uint16_t* __restrict__ h) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= rows * cols) return;
    const float v = z[t] + bias[t % cols];
    z[t] = v;
    h[t] = to_bf16(gelu(v, nullptr));
}

// grad of the pre-activation: dz = bf16(dh · gelu'(z))
__global__ void gelu_back_kernel(const float* __restrict__ dh, const float* __restrict__ z, int64_t count,
                                 uint16_t* __restrict__ dz, float* __restrict__ dz32) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count) return;
    float g;
    (void)gelu(z[t], &g);
    const float v = dh[t] * g;
    dz[t] = to_bf16(v);
    if (dz32) dz32[t] = v;
}

__global__ void add_bias_normalize_kernel(float* __restrict__ y, const float* __restrict__ b3, int64_t rows,
                                          int normalize) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    float v[3];
    for (int c = 0; c < 3; ++c) v[c] = y[3 * i + c] + b3[c];
    if (normalize) {
        const float s = 1.0f / (sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 1e-12f);
        for (int c = 0; c < 3; ++c) v[c] *= s;
    }
    for (int c = 0; c < 3; ++c) y[3 * i + c] = v[c];
}

// oriented:  loss = mean(1 − c),  dL/dy = −(t − c ŷ)/(|y|·B)
// line:      loss = mean(1 − c²), dL/dy = −2c (t − c ŷ)/(|y|·B),   c = ŷ·t
__global__ void loss_kernel(const float* __restrict__ y, const float* __restrict__ t, int64_t rows, int mode,
                            float* __restrict__ dy, uint16_t* __restrict__ dy16, double* __restrict__ loss_part) {
    __shared__ double lds[256];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    double l = 0;
    if (i < rows) {
        const float y0 = y[3 * i], y1 = y[3 * i + 1], y2 = y[3 * i + 2];
        const float ny = sqrtf(y0 * y0 + y1 * y1 + y2 * y2) + 1e-12f;
        const float h0 = y0 / ny, h1 = y1 / ny, h2 = y2 / ny;
        const float t0 = t[3 * i], t1 = t[3 * i + 1], t2 = t[3 * i + 2];
        const float c = h0 * t0 + h1 * t1 + h2 * t2;
        l = mode ? 1.0 - double(c) * double(c) : 1.0 - double(c);
        const float s = (mode ? -2.0f * c : -1.0f) / (ny * float(rows));
        const float g[3] = {s * (t0 - c * h0), s * (t1 - c * h1), s * (t2 - c * h2)};
        for (int k = 0; k < 3; ++k) {
            dy[3 * i + k] = g[k];
            dy16[3 * i + k] = to_bf16(g[k]);
        }
    }
    lds[threadIdx.x] = l;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = lds[0];
}

// column sums of a rows×cols FP32 matrix into out[cols]: one block per
// column, fixed-order tree reduction (deterministic)
__global__ __launch_bounds__(256) void colsum_kernel(const float* __restrict__ m, int64_t rows, int cols,
                                                     float* __restrict__ out) {
    __shared__ double lds[256];
    const int c = blockIdx.x;
    double s = 0;
    for (int64_t r = threadIdx.x; r < rows; r += 256) s += m[r * cols + c];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = float(lds[0]);
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, uint16_t* __restrict__ b16, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    b16[i] = to_bf16(pi);
}

__global__ void to_bf16_kernel(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = to_bf16(a[i]);
}

// batch gather with one rotation R (row-major 3×3) applied to every patch
// offset and to the target normal
__global__ void augment_kernel(const uint16_t* __restrict__ pin, const float* __restrict__ tin,
                               const int64_t* __restrict__ idx, int64_t batch, const float* __restrict__ R,
                               uint16_t* __restrict__ pout, float* __restrict__ tout) {
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const int64_t s = idx[b];
    float r[9];
    for (int k = 0; k < 9; ++k) r[k] = R[k];
    for (int j = 0; j < kPatchK + 1; ++j) {
        float v[3];
        if (j < kPatchK)
            for (int d = 0; d < 3; ++d) v[d] = from_bf16(pin[s * kIn + 3 * j + d]);
        else
            for (int d = 0; d < 3; ++d) v[d] = tin[3 * s + d];
        float w[3];
        for (int d = 0; d < 3; ++d) w[d] = r[3 * d] * v[0] + r[3 * d + 1] * v[1] + r[3 * d + 2] * v[2];
        if (j < kPatchK)
            for (int d = 0; d < 3; ++d) pout[b * kIn + 3 * j + d] = to_bf16(w[d]);
        else
            for (int d = 0; d < 3; ++d) tout[3 * b + d] = w[d];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static dwg_mlp* mlp_alloc() {
    auto* m = new dwg_mlp();
    try {
        m->params = static_cast<decltype(m->params)>(dev_malloc(sizeof(float) * kParams));
        m->bf16 = static_cast<decltype(m->bf16)>(dev_malloc(sizeof(uint16_t) * kParams));
        m->grad = static_cast<decltype(m->grad)>(dev_malloc(sizeof(float) * kParams));
        m->adam_m = static_cast<decltype(m->adam_m)>(dev_malloc(sizeof(float) * kParams));
        m->adam_v = static_cast<decltype(m->adam_v)>(dev_malloc(sizeof(float) * kParams));
    } catch (...) {
        delete m;
        throw;
    }
    return m;
}

static void mlp_set(dwg_ctx* ctx, dwg_mlp* m, const float* h) {
    DWG_HIP_CHECK(hipMemcpyAsync(m->params, h, sizeof(float) * kParams, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(kParams, 256)), dim3(256), 0, ctx->stream, m->params, m->bf16,
                       kParams);
    check_launch();
    fill_u32(ctx, m->adam_m, 0, kParams);
    fill_u32(ctx, m->adam_v, 0, kParams);
    fill_u32(ctx, m->grad, 0, kParams);
    m->step = 0;
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

dwg_mlp* mlp_create(dwg_ctx* ctx, uint64_t seed) {
    // uniform(±1/sqrt(fan_in)) weights and biases
    std::vector<float> h(kParams);
    std::mt19937_64 gen(seed);
    auto fill = [&](int64_t off, int64_t count, int fan_in) {
        std::uniform_real_distribution<float> u(-1.0f / std::sqrt(float(fan_in)), 1.0f / std::sqrt(float(fan_in)));
        for (int64_t i = 0; i < count; ++i) h[off + i] = u(gen);
    };
    fill(kOffW1, int64_t(kHid) * kIn, kIn);
    fill(kOffB1, kHid, kIn);
    fill(kOffW2, int64_t(kHid) * kHid, kHid);
    fill(kOffB2, kHid, kHid);
    fill(kOffW3, int64_t(kOut) * kHid, kHid);
    fill(kOffB3, kOut, kHid);
    dwg_mlp* m = mlp_alloc();
    try {
        mlp_set(ctx, m, h.data());
    } catch (...) {
        delete m;
        throw;
    }
    return m;
}

// Weights file: 8-byte magic "DWGMLP01", int32 dims {48,128,128,3}, then the
// FP32 parameter vector (W1, b1, W2, b2, W3, b3; row-major [out][in]).
static const char kMagic[8] = {'D', 'W', 'G', 'M', 'L', 'P', '0', '1'};

dwg_mlp* mlp_load(dwg_ctx* ctx, const char* path) {
    FILE* f = std::fopen(path, "rb");
    DWG_REQUIRE(f, std::string("cannot open ") + path);
    char magic[8];
    int32_t dims[4];
    std::vector<float> h(kParams);
    const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
                    std::fread(dims, 4, 4, f) == 4 && dims[0] == kIn && dims[1] == kHid && dims[2] == kHid &&
                    dims[3] == kOut && std::fread(h.data(), 4, size_t(kParams), f) == size_t(kParams);
    std::fclose(f);
    DWG_REQUIRE(ok, std::string("not a 48-128-128-3 weights file: ") + path);
    dwg_mlp* m = mlp_alloc();
    try {
        mlp_set(ctx, m, h.data());
    } catch (...) {
        delete m;
        throw;
    }
    return m;
}

void mlp_save(dwg_ctx* ctx, const dwg_mlp* m, const char* path) {
    std::vector<float> h(kParams);
    DWG_HIP_CHECK(hipMemcpyAsync(h.data(), m->params, sizeof(float) * kParams, hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    FILE* f = std::fopen(path, "wb");
    DWG_REQUIRE(f, std::string("cannot create ") + path);
    const int32_t dims[4] = {kIn, kHid, kHid, kOut};
    const bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(dims, 4, 4, f) == 4 &&
                    std::fwrite(h.data(), 4, size_t(kParams), f) == size_t(kParams);
    std::fclose(f);
    DWG_REQUIRE(ok, std::string("cannot write ") + path);
}

void mlp_patches_from(dwg_ctx* ctx, const dwg_knn_index* I, const float* src, int64_t m, const float* q, int64_t n,
                      uint16_t* out) {
    DWG_REQUIRE(I && I->n == m, "kNN index does not match the patch points");
    if (n == 0) return;
    const int kk = kPatchK + 1;
    int32_t* knn = static_cast<int32_t*>(ctx->buf("mlp_knn", sizeof(int32_t) * n * kk));
    float* d2 = static_cast<float*>(ctx->buf("mlp_d2", sizeof(float) * n * kk));
    knn_query(ctx, I, q, n, kk, nullptr, nullptr, knn, d2, "knn_query_patches");
    ProfileScope ps(ctx, "mlp_patches");
    hipLaunchKernelGGL(patches_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, knn, d2, kk, src, q, n, out);
    check_launch();
}

void mlp_patches(dwg_ctx* ctx, const dwg_knn_index* I, const float* p, int64_t n, uint16_t* out) {
    mlp_patches_from(ctx, I, p, n, p, n, out);
}

struct Acts {
    float *z1, *z2, *y;
    uint16_t *h1, *h2;
};

static Acts acts(dwg_ctx* ctx, int64_t rows) {
    Acts a;
    a.z1 = static_cast<float*>(ctx->buf("mlp_z1", sizeof(float) * rows * kHid));
    a.z2 = static_cast<float*>(ctx->buf("mlp_z2", sizeof(float) * rows * kHid));
    a.y = static_cast<float*>(ctx->buf("mlp_y", sizeof(float) * rows * kOut));
    a.h1 = static_cast<uint16_t*>(ctx->buf("mlp_h1", sizeof(uint16_t) * rows * kHid));
    a.h2 = static_cast<uint16_t*>(ctx->buf("mlp_h2", sizeof(uint16_t) * rows * kHid));
    return a;
}

// y = W3·gelu(W2·gelu(W1·x + b1) + b2) + b3 for `rows` patches (raw y in a.y)
static void forward(dwg_ctx* ctx, const dwg_mlp* m, const uint16_t* x, int64_t rows, const Acts& a) {
    const uint16_t* W = m->bf16;
    gemm(ctx, {x, kIn, 1, W + kOffW1, 1, kIn, a.z1, kHid, 1, rows, kHid, kIn, 0});
    hipLaunchKernelGGL(bias_gelu_kernel, dim3(grid_for(rows * kHid, 256)), dim3(256), 0, ctx->stream, a.z1,
                       m->params + kOffB1, rows, kHid, a.h1);
    check_launch();
    gemm(ctx, {a.h1, kHid, 1, W + kOffW2, 1, kHid, a.z2, kHid, 1, rows, kHid, kHid, 0});
    hipLaunchKernelGGL(bias_gelu_kernel, dim3(grid_for(rows * kHid, 256)), dim3(256), 0, ctx->stream, a.z2,
                       m->params + kOffB2, rows, kHid, a.h2);
    check_launch();
    gemm(ctx, {a.h2, kHid, 1, W + kOffW3, 1, kHid, a.y, kOut, 1, rows, kOut, kHid, 0});
}

// Inference of the whole network, 32 patches per wave: the three layers as
// the same sequence of v_mfma_f32_32x32x16_bf16 steps as the layer-by-layer
// GEMMs (so the same sums, bit for bit), bias and GELU applied to the
// accumulators, the BF16 activations staged in LDS between layers, and the
// output bias and normalisation at the end: one launch, no activations in
// HBM.
constexpr int kInferWaves = 4;
constexpr int kHPad = kHid + 8;  // LDS row of 136 BF16 values: 2-way bank conflicts at most

__device__ __forceinline__ void infer_layer(const f32x16 (&acc)[4], const float* __restrict__ bias, int lane,
                                            uint16_t (*h)[kHPad]) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int rr = 8 * (q / 4) + 4 * (lane >> 5) + (q % 4), cc = t * 32 + (lane & 31);
            h[rr][cc] = to_bf16(gelu(acc[t][q] + bias[cc], nullptr));
        }
}

__global__ __launch_bounds__(64 * kInferWaves) void mlp_infer_kernel(const uint16_t* __restrict__ x, int64_t n,
                                                                     const uint16_t* __restrict__ W,
                                                                     const float* __restrict__ P, int normalize,
                                                                     float* __restrict__ out) {
    __shared__ uint16_t hs[kInferWaves][32][kHPad];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint16_t(*h)[kHPad] = hs[wave];
    const int64_t row0 = (int64_t(blockIdx.x) * kInferWaves + wave) * 32;
    const int r = lane & 31, kq = 8 * (lane >> 5);
    const int64_t row = row0 + r;
    const bool ok = row < n;
    f32x16 acc[4];
    // layer 1: Z1 = X·W1ᵀ (K = 48)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[t][q] = 0.f;
    for (int k0 = 0; k0 < kIn; k0 += 16) {
        bf16x8 a;
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = ok ? short(x[row * kIn + k0 + kq + j]) : short(0);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf16x8 b;
#pragma unroll
            for (int j = 0; j < 8; ++j) b[j] = short(W[kOffW1 + int64_t(t * 32 + r) * kIn + k0 + kq + j]);
            acc[t] = mfma_32x32x16(a, b, acc[t]);
        }
    }
    infer_layer(acc, P + kOffB1, lane, h);
    __syncthreads();
    // layer 2: Z2 = H1·W2ᵀ (K = 128)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[t][q] = 0.f;
    for (int k0 = 0; k0 < kHid; k0 += 16) {
        bf16x8 a;
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = short(h[r][k0 + kq + j]);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf16x8 b;
#pragma unroll
            for (int j = 0; j < 8; ++j) b[j] = short(W[kOffW2 + int64_t(t * 32 + r) * kHid + k0 + kq + j]);
            acc[t] = mfma_32x32x16(a, b, acc[t]);
        }
    }
    __syncthreads();
    infer_layer(acc, P + kOffB2, lane, h);
    __syncthreads();
    // layer 3: Y = H2·W3ᵀ (3 columns of a 32-wide tile)
    f32x16 y;
#pragma unroll
    for (int q = 0; q < 16; ++q) y[q] = 0.f;
    for (int k0 = 0; k0 < kHid; k0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            a[j] = short(h[r][k0 + kq + j]);
            b[j] = r < kOut ? short(W[kOffW3 + int64_t(r) * kHid + k0 + kq + j]) : short(0);
        }
        y = mfma_32x32x16(a, b, y);
    }
    __syncthreads();
    // column c of row i sits in lane c (+32), register 8(i/8)... stage through LDS as floats
    float* ys = reinterpret_cast<float*>(&h[0][0]);  // 32 × 3 floats
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int rr = 8 * (q / 4) + 4 * (lane >> 5) + (q % 4);
        if (r < kOut) ys[rr * kOut + r] = y[q];
    }
    __syncthreads();
    if (lane < 32 && ok) {
        float v[3];
        for (int c = 0; c < 3; ++c) v[c] = ys[lane * kOut + c] + P[kOffB3 + c];
        if (normalize) {
            const float s = 1.0f / (sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 1e-12f);
            for (int c = 0; c < 3; ++c) v[c] *= s;
        }
        for (int c = 0; c < 3; ++c) out[row * 3 + c] = v[c];
    }
}

void mlp_forward(dwg_ctx* ctx, const dwg_mlp* m, const uint16_t* x, int64_t n, float* out, int normalize) {
    ProfileScope ps(ctx, "mlp_forward");
    if (n <= 0) return;
    const int64_t waves = (n + 31) / 32;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid_for(waves, kInferWaves)), dim3(64 * kInferWaves), 0, ctx->stream,
                       x, n, m->bf16, m->params, normalize, out);
    check_launch();
}

double mlp_loss_grad(dwg_ctx* ctx, dwg_mlp* m, const uint16_t* x, const float* target, int64_t B, int mode) {
    DWG_REQUIRE(B > 0, "empty batch");
    ProfileScope ps(ctx, "mlp_train");
    Acts a = acts(ctx, B);
    forward(ctx, m, x, B, a);
    hipLaunchKernelGGL(add_bias_normalize_kernel, dim3(grid_for(B, 256)), dim3(256), 0, ctx->stream, a.y,
                       m->params + kOffB3, B, 0);
    check_launch();
    float* dy = static_cast<float*>(ctx->buf("mlp_dy", sizeof(float) * B * kOut));
    uint16_t* dy16 = static_cast<uint16_t*>(ctx->buf("mlp_dy16", sizeof(uint16_t) * B * kOut));
    const unsigned lb = grid_for(B, 256);
    double* lpart = static_cast<double*>(ctx->buf("mlp_loss", sizeof(double) * (lb + 1)));
    hipLaunchKernelGGL(loss_kernel, dim3(lb), dim3(256), 0, ctx->stream, a.y, target, B, mode, dy, dy16, lpart);
    check_launch();
    float* G = m->grad;
    const uint16_t* W = m->bf16;
    // layer 3: dW3 = dYᵀ·H2, db3 = Σ dY, dH2 = dY·W3
    gemm(ctx, {dy16, 1, kOut, a.h2, kHid, 1, G + kOffW3, kHid, 1, kOut, kHid, B, 0});
    hipLaunchKernelGGL(colsum_kernel, dim3(kOut), dim3(256), 0, ctx->stream, dy, B, kOut, G + kOffB3);
    check_launch();
    float* dh = static_cast<float*>(ctx->buf("mlp_dh", sizeof(float) * B * kHid));
    float* dz32 = static_cast<float*>(ctx->buf("mlp_dz32", sizeof(float) * B * kHid));
    uint16_t* dz = static_cast<uint16_t*>(ctx->buf("mlp_dz", sizeof(uint16_t) * B * kHid));
    gemm(ctx, {dy16, kOut, 1, W + kOffW3, kHid, 1, dh, kHid, 1, B, kHid, kOut, 0});
    hipLaunchKernelGGL(gelu_back_kernel, dim3(grid_for(B * kHid, 256)), dim3(256), 0, ctx->stream, dh, a.z2, B * kHid,
                       dz, dz32);
    check_launch();
    // layer 2: dW2 = dZ2ᵀ·H1, db2 = Σ dZ2, dH1 = dZ2·W2
    gemm(ctx, {dz, 1, kHid, a.h1, kHid, 1, G + kOffW2, kHid, 1, kHid, kHid, B, 0});
    hipLaunchKernelGGL(colsum_kernel, dim3(kHid), dim3(256), 0, ctx->stream, dz32, B, kHid, G + kOffB2);
    check_launch();
    gemm(ctx, {dz, kHid, 1, W + kOffW2, kHid, 1, dh, kHid, 1, B, kHid, kHid, 0});
    hipLaunchKernelGGL(gelu_back_kernel, dim3(grid_for(B * kHid, 256)), dim3(256), 0, ctx->stream, dh, a.z1, B * kHid,
                       dz, dz32);
    check_launch();
    // layer 1: dW1 = dZ1ᵀ·X, db1 = Σ dZ1
    gemm(ctx, {dz, 1, kHid, x, kIn, 1, G + kOffW1, kIn, 1, kHid, kIn, B, 0});
    hipLaunchKernelGGL(colsum_kernel, dim3(kHid), dim3(256), 0, ctx->stream, dz32, B, kHid, G + kOffB1);
    check_launch();
    std::vector<double> h(lb);
    DWG_HIP_CHECK(hipMemcpyAsync(h.data(), lpart, sizeof(double) * lb, hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double l = 0;
    for (double v : h) l += v;
    return l / double(B);
}

void mlp_adam(dwg_ctx* ctx, dwg_mlp* m, float lr, float b1, float b2, float eps) {
    m->step += 1;
    const float c1 = 1.0f - std::pow(b1, float(m->step)), c2 = 1.0f - std::pow(b2, float(m->step));
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(kParams, 256)), dim3(256), 0, ctx->stream, m->params, m->grad,
                       m->adam_m, m->adam_v, m->bf16, kParams, lr, b1, b2, eps, c1, c2);
    check_launch();
}

void mlp_augment(dwg_ctx* ctx, const uint16_t* pin, const float* tin, const int64_t* idx, int64_t batch,
                 const float* h_rot, uint16_t* pout, float* tout) {
    float* R = static_cast<float*>(ctx->buf("mlp_rot", sizeof(float) * 16));
    DWG_HIP_CHECK(hipMemcpyAsync(R, h_rot, sizeof(float) * 9, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(augment_kernel, dim3(grid_for(batch, 256)), dim3(256), 0, ctx->stream, pin, tin, idx, batch, R,
                       pout, tout);
    check_launch();
}

void mlp_get(dwg_ctx* ctx, const dwg_mlp* m, float* h_params, float* h_grad) {
    if (h_params)
        DWG_HIP_CHECK(hipMemcpyAsync(h_params, m->params, sizeof(float) * kParams, hipMemcpyDeviceToHost, ctx->stream));
    if (h_grad) DWG_HIP_CHECK(hipMemcpyAsync(h_grad, m->grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void mlp_put(dwg_ctx* ctx, dwg_mlp* m, const float* h_params) { mlp_set(ctx, m, h_params); }

void mlp_init(dwg_ctx* ctx, const dwg_mlp* m, const float* p, int64_t n, const float c[3], float* out) {
    DWG_REQUIRE(n > 0, "no points");
    ProfileScope prof(ctx, "mlp_init");
    const float* src = p;
    int64_t ns = n;
    if (n > kTrainPoints) {
        const uint32_t stride = uint32_t((n + kTrainPoints - 1) / kTrainPoints);
        int64_t* slot = static_cast<int64_t*>(ctx->buf("mlp_sub_slot", sizeof(int64_t) * (n + 1)));
        hipLaunchKernelGGL(subsample_flags_kernel, dim3(grid_for(n + 1, 256)), dim3(256), 0, ctx->stream, n, stride,
                           slot);
        check_launch();
        exclusive_scan<int64_t>(ctx, slot, n);
        ns = read_scalar<int64_t>(ctx, slot + n);
        DWG_REQUIRE(ns > kPatchK, "patch subsample too small");
        float* sub = static_cast<float*>(ctx->buf("mlp_sub_pts", sizeof(float) * 3 * ns));
        hipLaunchKernelGGL(subsample_gather_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, slot, p, n, sub);
        check_launch();
        src = sub;
    }
    std::unique_ptr<dwg_knn_index> index(knn_build(ctx, src, ns, kPatchK + 1));
    uint16_t* patches = static_cast<uint16_t*>(ctx->buf("mlp_init_patches", sizeof(uint16_t) * kIn * n));
    mlp_patches_from(ctx, index.get(), src, ns, p, n, patches);
    mlp_forward(ctx, m, patches, n, out, 1);
    hipLaunchKernelGGL(orient_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, out, p, n, c[0], c[1], c[2]);
    check_launch();
}

void gemm_bf16_test(dwg_ctx* ctx, const uint16_t* A, int64_t sAi, int64_t sAk, const uint16_t* B, int64_t sBk,
                    int64_t sBj, float* C, int64_t M, int64_t N, int64_t K) {
    gemm(ctx, {A, sAi, sAk, B, sBk, sBj, C, N, 1, M, N, K, 0});
}

}  // namespace dwg

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {
template <class F>
int mlp_guard(dwg_ctx* ctx, F&& fn) {
    if (!ctx) return DWG_ERR_ARG;
    try {
        DWG_HIP_CHECK(hipSetDevice(ctx->device));
        fn();
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
}  // namespace

extern "C" {

int64_t dwg_mlp_num_params(void) { return dwg::kParams; }

int dwg_mlp_create(dwg_ctx* ctx, uint64_t seed, dwg_mlp** out) {
    if (!out) return DWG_ERR_ARG;
    *out = nullptr;
    return mlp_guard(ctx, [&] { *out = dwg::mlp_create(ctx, seed); });
}

int dwg_mlp_load(dwg_ctx* ctx, const char* path, dwg_mlp** out) {
    if (!out || !path) return DWG_ERR_ARG;
    *out = nullptr;
    return mlp_guard(ctx, [&] { *out = dwg::mlp_load(ctx, path); });
}

int dwg_mlp_save(dwg_ctx* ctx, const dwg_mlp* mlp, const char* path) {
    if (!mlp || !path) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] { dwg::mlp_save(ctx, mlp, path); });
}

void dwg_mlp_destroy(dwg_mlp* mlp) { delete mlp; }

int dwg_mlp_get(dwg_ctx* ctx, const dwg_mlp* mlp, float* h_params, float* h_grad) {
    if (!mlp) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] { dwg::mlp_get(ctx, mlp, h_params, h_grad); });
}

int dwg_mlp_set(dwg_ctx* ctx, dwg_mlp* mlp, const float* h_params) {
    if (!mlp || !h_params) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] { dwg::mlp_put(ctx, mlp, h_params); });
}

int dwg_mlp_patches(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_pts, int64_t n, uint16_t* d_patches) {
    if (!index || !d_pts || !d_patches) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] { dwg::mlp_patches(ctx, index, d_pts, n, d_patches); });
}

int dwg_mlp_patches_from(dwg_ctx* ctx, const dwg_knn_index* index, const float* d_src, int64_t n_src,
                         const float* d_q, int64_t nq, uint16_t* d_patches) {
    if (!index || !d_src || (nq > 0 && (!d_q || !d_patches))) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] { dwg::mlp_patches_from(ctx, index, d_src, n_src, d_q, nq, d_patches); });
}

int dwg_mlp_init(dwg_ctx* ctx, const dwg_mlp* mlp, const float* d_pts, int64_t n, const float* h_center,
                 float* d_out) {
    if (!mlp || !d_pts || !h_center || !d_out) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] { dwg::mlp_init(ctx, mlp, d_pts, n, h_center, d_out); });
}

int dwg_mlp_forward(dwg_ctx* ctx, const dwg_mlp* mlp, const uint16_t* d_patches, int64_t n, int normalize,
                    float* d_out) {
    if (!mlp || (n > 0 && (!d_patches || !d_out))) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] {
        if (n > 0) dwg::mlp_forward(ctx, mlp, d_patches, n, d_out, normalize);
    });
}

int dwg_mlp_loss_grad(dwg_ctx* ctx, dwg_mlp* mlp, const uint16_t* d_patches, const float* d_target, int64_t batch,
                      int mode, double* h_loss) {
    if (!mlp || !d_patches || !d_target || mode < 0 || mode > 1) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] {
        const double l = dwg::mlp_loss_grad(ctx, mlp, d_patches, d_target, batch, mode);
        if (h_loss) *h_loss = l;
    });
}

int dwg_mlp_adam(dwg_ctx* ctx, dwg_mlp* mlp, float lr, float beta1, float beta2, float eps) {
    if (!mlp) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] { dwg::mlp_adam(ctx, mlp, lr, beta1, beta2, eps); });
}

int dwg_mlp_augment(dwg_ctx* ctx, const uint16_t* d_patches, const float* d_target, const int64_t* d_index,
                    int64_t batch, const float* h_rotation, uint16_t* d_patches_out, float* d_target_out) {
    if (!d_patches || !d_target || !d_index || !h_rotation || !d_patches_out || !d_target_out) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] {
        dwg::mlp_augment(ctx, d_patches, d_target, d_index, batch, h_rotation, d_patches_out, d_target_out);
    });
}

int dwg_gemm_bf16(dwg_ctx* ctx, const uint16_t* d_a, int64_t s_ai, int64_t s_ak, const uint16_t* d_b, int64_t s_bk,
                  int64_t s_bj, float* d_c, int64_t m, int64_t n, int64_t k) {
    if (!d_a || !d_b || !d_c || m < 0 || n < 0 || k < 0) return DWG_ERR_ARG;
    return mlp_guard(ctx, [&] { dwg::gemm_bf16_test(ctx, d_a, s_ai, s_ak, d_b, s_bk, s_bj, d_c, m, n, k); });
}

}  // extern "C"

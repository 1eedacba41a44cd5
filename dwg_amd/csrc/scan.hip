// Device-wide exclusive scans and reductions (hand-written; replace the
// reference's thrust::reduce / CUB calls, TurboDWG.cu:541-542).
#include "scan.h"

namespace dwg {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <class T>
__device__ T block_exclusive_scan(T v, T* lds, T& total) {
    // wave-level inclusive scan via LDS (64 lanes), then across 4 waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
    lds[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < 64; off <<= 1) {
        T y = (lane >= off) ? lds[threadIdx.x - off] : T();
        __syncthreads();
        x = x + y;
        lds[threadIdx.x] = x;
        __syncthreads();
    }
    __shared__ T wave_tot[kScanThreads / 64];
    if (lane == 63) wave_tot[wave] = x;
    __syncthreads();
    T before = T();
    for (int w = 0; w < wave; ++w) before = before + wave_tot[w];
    total = T();
    for (int w = 0; w < kScanThreads / 64; ++w) total = total + wave_tot[w];
    __syncthreads();
    return before + x - v;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void tile_reduce(const T* __restrict__ d, int64_t n,
                                                             T* __restrict__ partial) {
    __shared__ T lds[kScanThreads];
    int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
    T s = T();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s = s + d[base + k];
    T total;
    (void)block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void tile_scan(T* __restrict__ d, int64_t n,
                                                           const T* __restrict__ partial) {
    __shared__ T lds[kScanThreads];
    int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
    T items[kScanItems];
    T s = T();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        items[k] = (base + k < n) ? d[base + k] : T();
        s = s + items[k];
    }
    T total;
    T run = block_exclusive_scan(s, lds, total) + partial[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) d[base + k] = run;
        run = run + items[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) d[n] = partial[blockIdx.x] + total;
}

template <class T>
__global__ void single_scan(T* __restrict__ d, int64_t n) {
    // tiny inputs: one thread (n ≤ kScanTile)
    T run = T();
    for (int64_t i = 0; i < n; ++i) {
        T v = d[i];
        d[i] = run;
        run = run + v;
    }
    d[n] = run;
}

__global__ void fill_u32_kernel(uint32_t* d, uint32_t v, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) d[i] = v;
}

__global__ __launch_bounds__(256) void reduce_f32_f64(const float* __restrict__ v, int64_t n,
                                                      double* __restrict__ partial) {
    __shared__ double lds[256];
    double s = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) s += v[i];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

__global__ void reduce_final_f64(const double* __restrict__ partial, int n, double* __restrict__ out) {
    __shared__ double lds[256];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = lds[0];
}

}  // namespace

template <class T>
void exclusive_scan(dwg_ctx* ctx, T* d, int64_t n, int level) {
    if (n <= kScanTile) {
        hipLaunchKernelGGL(single_scan<T>, dim3(1), dim3(1), 0, ctx->stream, d, n);
        check_launch();
        return;
    }
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    std::string name = "scan_partial_" + std::to_string(level) + "_" + std::to_string(sizeof(T));
    T* partial = static_cast<T*>(ctx->buf(name, sizeof(T) * (tiles + 1)));
    hipLaunchKernelGGL(tile_reduce<T>, dim3(unsigned(tiles)), dim3(kScanThreads), 0, ctx->stream, d, n, partial);
    check_launch();
    exclusive_scan<T>(ctx, partial, tiles, level + 1);
    hipLaunchKernelGGL(tile_scan<T>, dim3(unsigned(tiles)), dim3(kScanThreads), 0, ctx->stream, d, n, partial);
    check_launch();
}

template void exclusive_scan<int64_t>(dwg_ctx*, int64_t*, int64_t, int);
template void exclusive_scan<Double3>(dwg_ctx*, Double3*, int64_t, int);

void exclusive_scan_i64(dwg_ctx* ctx, int64_t* d, int64_t n) { exclusive_scan<int64_t>(ctx, d, n, 0); }

void fill_u32(dwg_ctx* ctx, void* d, uint32_t value, int64_t count) {
    if (count <= 0) return;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(count, 256)), dim3(256), 0, ctx->stream,
                       static_cast<uint32_t*>(d), value, count);
    check_launch();
}

double reduce_sum_f32(dwg_ctx* ctx, const float* d, int64_t n) {
    const int blocks = 1024;
    double* partial = static_cast<double*>(ctx->buf("reduce_partial", sizeof(double) * (blocks + 1)));
    hipLaunchKernelGGL(reduce_f32_f64, dim3(blocks), dim3(256), 0, ctx->stream, d, n, partial);
    check_launch();
    hipLaunchKernelGGL(reduce_final_f64, dim3(1), dim3(256), 0, ctx->stream, partial, blocks, partial + blocks);
    check_launch();
    return read_scalar<double>(ctx, partial + blocks);
}

}  // namespace dwg

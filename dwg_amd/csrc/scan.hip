// Device-wide exclusive scans and reductions (hand-written; replace the
// reference's thrust::reduce / CUB calls, TurboDWG.cu:541-542).
#include "scan.h"

#include <cstring>

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
__global__ __launch_bounds__(kScanThreads) void single_scan(T* __restrict__ d, int64_t n) {
    // inputs of at most kScanTile elements (the top level of a scan): one
    // block, kScanItems consecutive elements a thread
    __shared__ T lds[kScanThreads];
    const int64_t base = int64_t(threadIdx.x) * kScanItems;
    T items[kScanItems];
    T s = T();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        items[k] = (base + k < n) ? d[base + k] : T();
        s = s + items[k];
    }
    T total;
    T run = block_exclusive_scan(s, lds, total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) d[base + k] = run;
        run = run + items[k];
    }
    if (threadIdx.x == kScanThreads - 1) d[n] = total;
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

// Stable LSD radix sort of (uint64 key, int value) pairs, 8 bits a pass.
// A block owns a tile of kSortTile consecutive items, taken in rounds of
// 256 (item r·256 + t of the tile in round r, thread t).  Pass: every block
// counts its tile's digits; one exclusive scan over the counts laid out
// digit-major ([digit][block]) gives each (digit, block) its first output
// slot; every block then writes its items to slot + rank, the rank of an
// item among its tile's items with the same digit that precede it (earlier
// rounds, earlier waves of its round, lower lanes of its wave: the wave's
// lanes with the same digit come from 8 ballots), so equal digits keep their
// input order and the passes compose into a stable sort.
constexpr int kSortThreads = 256, kSortRounds = 16, kSortTile = kSortThreads * kSortRounds;
static_assert(kSortThreads == 256, "one thread per digit value");

__global__ __launch_bounds__(kSortThreads) void radix_count_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                                   int shift, int64_t nblocks,
                                                                   int64_t* __restrict__ counts) {
    __shared__ unsigned s_h[256];
    s_h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = int64_t(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        const int64_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[int64_t(threadIdx.x) * nblocks + blockIdx.x] = s_h[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(const uint64_t* __restrict__ kin,
                                                                     const int* __restrict__ vin, int64_t n, int shift,
                                                                     int64_t nblocks, const int64_t* __restrict__ first,
                                                                     uint64_t* __restrict__ kout,
                                                                     int* __restrict__ vout) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ int s_run[256];          // items of each digit in the tile's earlier rounds
    __shared__ int s_wave[kWaves][256];  // items of each digit in each wave of this round
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    s_run[threadIdx.x] = 0;
    for (int w = 0; w < kWaves; ++w) s_wave[w][threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = int64_t(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        const int64_t i = base + r * kSortThreads + threadIdx.x;
        const bool ok = i < n;
        const uint64_t k = ok ? kin[i] : 0;
        const int v = ok ? vin[i] : 0;
        const int dg = int((k >> shift) & 255u);
        uint64_t same = __builtin_amdgcn_ballot_w64(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __builtin_amdgcn_ballot_w64((dg >> b) & 1);
            same &= ((dg >> b) & 1) ? m : ~m;
        }
        const int below = __builtin_popcountll(same & ((1ull << lane) - 1ull));
        if (ok && below == 0) s_wave[wave][dg] = __builtin_popcountll(same);
        __syncthreads();
        if (ok) {
            int rank = s_run[dg] + below;
            for (int w = 0; w < wave; ++w) rank += s_wave[w][dg];
            const int64_t dst = first[int64_t(dg) * nblocks + blockIdx.x] + rank;
            kout[dst] = k;
            vout[dst] = v;
        }
        __syncthreads();
        int add = 0;
        for (int w = 0; w < kWaves; ++w) {
            add += s_wave[w][threadIdx.x];
            s_wave[w][threadIdx.x] = 0;
        }
        s_run[threadIdx.x] += add;
        __syncthreads();
    }
}

}  // namespace

void radix_sort_pairs(dwg_ctx* ctx, const uint64_t* kin, const int* vin, uint64_t* kout, int* vout, int64_t n,
                      int bits) {
    DWG_REQUIRE(n >= 0 && bits >= 1 && bits <= 64, "bad radix sort request");
    if (n == 0) return;
    const int64_t nblocks = (n + kSortTile - 1) / kSortTile;
    int64_t* first = static_cast<int64_t*>(ctx->buf("sort_first", sizeof(int64_t) * (256 * nblocks + 1)));
    uint64_t* kt[2] = {static_cast<uint64_t*>(ctx->buf("sort_k0", sizeof(uint64_t) * n)),
                       static_cast<uint64_t*>(ctx->buf("sort_k1", sizeof(uint64_t) * n))};
    int* vt[2] = {static_cast<int*>(ctx->buf("sort_v0", sizeof(int) * n)),
                  static_cast<int*>(ctx->buf("sort_v1", sizeof(int) * n))};
    const int passes = (bits + 7) / 8;
    const uint64_t* ks = kin;
    const int* vs = vin;
    for (int p = 0; p < passes; ++p) {
        uint64_t* kd = p == passes - 1 ? kout : kt[p & 1];
        int* vd = p == passes - 1 ? vout : vt[p & 1];
        hipLaunchKernelGGL(radix_count_kernel, dim3(unsigned(nblocks)), dim3(kSortThreads), 0, ctx->stream, ks, n, 8 * p,
                           nblocks, first);
        check_launch();
        exclusive_scan<int64_t>(ctx, first, 256 * nblocks);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(unsigned(nblocks)), dim3(kSortThreads), 0, ctx->stream, ks, vs, n,
                           8 * p, nblocks, first, kd, vd);
        check_launch();
        ks = kd;
        vs = vd;
    }
}

template <class T>
void exclusive_scan(dwg_ctx* ctx, T* d, int64_t n, int level) {
    if (n <= kScanTile) {
        hipLaunchKernelGGL(single_scan<T>, dim3(1), dim3(kScanThreads), 0, ctx->stream, d, n);
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

// the partial sums and their total (at partial[blocks]) queued on the stream
static const double* reduce_sum_f32_queue(dwg_ctx* ctx, const float* d, int64_t n) {
    const int blocks = 1024;
    double* partial = static_cast<double*>(ctx->buf("reduce_partial", sizeof(double) * (blocks + 1)));
    hipLaunchKernelGGL(reduce_f32_f64, dim3(blocks), dim3(256), 0, ctx->stream, d, n, partial);
    check_launch();
    hipLaunchKernelGGL(reduce_final_f64, dim3(1), dim3(256), 0, ctx->stream, partial, blocks, partial + blocks);
    check_launch();
    return partial + blocks;
}

double reduce_sum_f32(dwg_ctx* ctx, const float* d, int64_t n) {
    return read_scalar<double>(ctx, reduce_sum_f32_queue(ctx, d, n));
}

// the outstanding sum's place in the context's pinned page (read_scalar
// uses its start)
constexpr size_t kAsyncSumSlot = 2048;

void reduce_sum_f32_async(dwg_ctx* ctx, const float* d, int64_t n) {
    const double* total = reduce_sum_f32_queue(ctx, d, n);
    if (!ctx->pinned) DWG_HIP_CHECK(hipHostMalloc(&ctx->pinned, 4096, hipHostMallocDefault));
    if (!ctx->readback) DWG_HIP_CHECK(hipEventCreateWithFlags(&ctx->readback, hipEventDisableTiming));
    DWG_HIP_CHECK(hipMemcpyAsync(static_cast<char*>(ctx->pinned) + kAsyncSumSlot, total, sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipEventRecord(ctx->readback, ctx->stream));
}

double reduce_sum_f32_wait(dwg_ctx* ctx) {
    DWG_REQUIRE(ctx->readback, "no sum outstanding");
    DWG_HIP_CHECK(hipEventSynchronize(ctx->readback));
    double v;
    std::memcpy(&v, static_cast<const char*>(ctx->pinned) + kAsyncSumSlot, sizeof(v));
    return v;
}

}  // namespace dwg

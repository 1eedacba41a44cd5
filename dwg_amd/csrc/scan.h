// Device-wide scans and reductions.
#pragma once
#include "common.h"

namespace dwg {

struct Double3 {
    double x, y, z;  // trivial default ctor (LDS arrays); Double3() value-initialises to 0
    Double3() = default;
    __host__ __device__ Double3(double a, double b, double c) : x(a), y(b), z(c) {}
    __host__ __device__ Double3 operator+(const Double3& o) const { return Double3(x + o.x, y + o.y, z + o.z); }
    __host__ __device__ Double3 operator-(const Double3& o) const { return Double3(x - o.x, y - o.y, z - o.z); }
};

// In-place exclusive scan of d[0..n); d must hold n+1 elements, d[n] = total.
template <class T>
void exclusive_scan(dwg_ctx* ctx, T* d, int64_t n, int level = 0);

// Stable sort of n (key, value) pairs by the low `bits` bits of the key
// (higher bits must be 0): LSD radix, 8 bits a pass.  kout/vout must not
// alias kin/vin.
void radix_sort_pairs(dwg_ctx* ctx, const uint64_t* kin, const int* vin, uint64_t* kout, int* vout, int64_t n,
                      int bits);

// FP64 sum of a float array (synchronises, returns the host value).
double reduce_sum_f32(dwg_ctx* ctx, const float* d, int64_t n);
// The same sum without waiting: queued on the stream with its copy into
// pinned host memory; reduce_sum_f32_wait returns it once that copy is done
// (work queued in between keeps the GPU busy while the host waits).  One
// sum may be outstanding per context.
void reduce_sum_f32_async(dwg_ctx* ctx, const float* d, int64_t n);
double reduce_sum_f32_wait(dwg_ctx* ctx);

}  // namespace dwg

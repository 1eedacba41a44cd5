// Internal helpers shared by the libdwg_hip kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include "pool.h"

#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dwg_hip.h"

namespace dwg {

// Error carried from deep inside the library to the ABI boundary, where it
// becomes a status code plus a message (never crosses the C ABI).
struct Error : std::runtime_error {
    int code;
    int64_t value;
    Error(int c, const std::string& m, int64_t v = 0) : std::runtime_error(m), code(c), value(v) {}
};

#define DWG_HIP_CHECK(expr)                                                                   \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::dwg::Error(DWG_ERR_HIP, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

#define DWG_REQUIRE(cond, msg)                                                 \
    do {                                                                       \
        if (!(cond)) throw ::dwg::Error(DWG_ERR_ARG, std::string(msg));        \
    } while (0)

// Growable device buffer.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release() {
        dev_free(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void* ensure(size_t need) {
        if (need <= bytes && ptr) return ptr;
        release();
        size_t b = need < 256 ? 256 : need;
        ptr = dev_malloc(b);
        bytes = b;
        return ptr;
    }
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

struct KernelTimes {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0;
    int64_t launches = 0;
};

}  // namespace dwg

struct dwg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool profile = false;
    std::string profile_suffix;  // appended to recorded kernel names (phase tag)
    std::map<std::string, dwg::KernelTimes> times;
    std::map<std::string, int64_t> counters;  // event counts recorded while profiling
    // named scratch buffers owned by the context
    std::map<std::string, dwg::DevBuf> scratch;
    // pinned host staging for small read-backs
    void* pinned = nullptr;
    // marks the copy of an outstanding asynchronous read-back (reduce_sum_f32_async)
    hipEvent_t readback = nullptr;

    void* buf(const std::string& name, size_t bytes) { return scratch[name].ensure(bytes); }
    size_t buf_size(const std::string& name) const {
        const auto it = scratch.find(name);
        return it == scratch.end() ? 0 : it->second.bytes;
    }
    void count(const std::string& name, int64_t v) {
        if (profile) counters[name + profile_suffix] += v;
    }
};

namespace dwg {

// Brackets a kernel launch with events when profiling is on (a null
// context records nothing).
struct ProfileScope {
    dwg_ctx* ctx;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    ProfileScope(dwg_ctx* c, const char* n) : ctx(c), name(n) {
        if (ctx && ctx->profile) {
            DWG_HIP_CHECK(hipEventCreate(&a));
            DWG_HIP_CHECK(hipEventCreate(&b));
            DWG_HIP_CHECK(hipEventRecord(a, ctx->stream));
        }
    }
    ~ProfileScope() {
        if (ctx && ctx->profile && a) {
            (void)hipEventRecord(b, ctx->stream);
            ctx->times[std::string(name) + ctx->profile_suffix].pending.emplace_back(a, b);
        }
    }
};

inline void check_launch() { DWG_HIP_CHECK(hipGetLastError()); }

inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g > 0x7fffffff) throw Error(DWG_ERR_ARG, "launch grid too large");
    return unsigned(g < 1 ? 1 : g);
}

// Copies a small device value to the host through pinned staging and syncs.
template <class T>
T read_scalar(dwg_ctx* ctx, const T* d) {
    if (!ctx->pinned) DWG_HIP_CHECK(hipHostMalloc(&ctx->pinned, 4096, hipHostMallocDefault));
    DWG_HIP_CHECK(hipMemcpyAsync(ctx->pinned, d, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    DWG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return *static_cast<T*>(ctx->pinned);
}

// Fill kernel used instead of hipMemsetAsync so the library stays safe to
// record into HIP graphs.
void fill_u32(dwg_ctx* ctx, void* d, uint32_t value, int64_t count);

// Exclusive scan of int64 values in place; returns nothing, total at d[n].
// d must hold n+1 elements.
void exclusive_scan_i64(dwg_ctx* ctx, int64_t* d, int64_t n);

}  // namespace dwg

// ---- device-side helpers --------------------------------------------------
namespace dwg {

// The one float squared-distance formula used by every kernel and the CPU
// oracle: fma(dz, dz, fma(dy, dy, dx*dx)).
__device__ __forceinline__ float dist2(float dx, float dy, float dz) {
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// Block index remap so that each XCD (blocks are dealt round-robin over
// the 8 XCDs) gets a contiguous run of block indices: neighbouring bricks
// or query chunks then share one L2
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
    const int64_t per = nb / 8;
    if (b >= per * 8) return b;  // tail keeps its slot
    return (b % 8) * per + b / 8;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
    uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return (uint64_t(hi) << 32) | lo;
}

__device__ __forceinline__ void kahan_add(float& sum, float& comp, float x) {
    float y = x - comp;
    float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

constexpr float kInv4Pi = 0.0795774715459476679f;  // 1/(4π)
constexpr float kLog2e = 1.44269504088896340736f;
constexpr float kEpsPlain = 1e-8f;     // fast_winding_number.cpp:183-184
constexpr float kEpsScreened = 1e-7f;  // TurboLibiGWN.cu:555
constexpr float kScreenDefault = 0.0015f;  // TurboLibiGWN.cu:549

}  // namespace dwg

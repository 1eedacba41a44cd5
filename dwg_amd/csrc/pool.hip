// Per-device cache of released device blocks (see pool.h).
#include "pool.h"

#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.h"

namespace dwg {
namespace {

struct Pool {
    std::multimap<size_t, void*> ready;        // released and safe to reuse (by size)
    std::vector<std::pair<size_t, void*>> pending;  // released since the last device synchronisation
    std::unordered_map<void*, size_t> live;    // handed out: block → its size
};

std::mutex g_mu;
std::map<int, Pool> g_pools;  // by device ordinal

size_t round_size(size_t b) {
    constexpr size_t kBig = size_t(2) << 20;
    if (b < 256) return 256;
    if (b < (size_t(1) << 20)) return (b + 255) & ~size_t(255);
    return (b + kBig - 1) & ~(kBig - 1);
}

bool fits(size_t have, size_t want) { return have >= want && have <= want + want / 4; }

// a ready block for `b` bytes, or nullptr
void* take(Pool& P, size_t b) {
    auto it = P.ready.lower_bound(b);
    if (it == P.ready.end() || !fits(it->first, b)) return nullptr;
    void* p = it->second;
    P.live[p] = it->first;
    P.ready.erase(it);
    return p;
}

// every pending block of the current device becomes ready: the device is
// idle once the synchronisation returns
void settle(Pool& P) {
    if (P.pending.empty()) return;
    DWG_HIP_CHECK(hipDeviceSynchronize());
    for (auto& sp : P.pending) P.ready.emplace(sp.first, sp.second);
    P.pending.clear();
}

size_t drop_ready(Pool& P) {
    size_t freed = 0;
    for (auto& sp : P.ready) {
        (void)hipFree(sp.second);
        freed += sp.first;
    }
    P.ready.clear();
    return freed;
}

}  // namespace

void* dev_malloc(size_t bytes) {
    int dev = 0;
    DWG_HIP_CHECK(hipGetDevice(&dev));
    const size_t b = round_size(bytes);
    std::lock_guard<std::mutex> lock(g_mu);
    Pool& P = g_pools[dev];
    if (void* p = take(P, b)) return p;
    for (auto& sp : P.pending)
        if (fits(sp.first, b)) {
            settle(P);
            if (void* p = take(P, b)) return p;
            break;
        }
    void* p = nullptr;
    if (hipMalloc(&p, b) != hipSuccess) {
        (void)hipGetLastError();
        settle(P);  // out of memory: hand the cache back and try once more
        drop_ready(P);
        p = nullptr;
        const hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess)
            throw Error(DWG_ERR_HIP, "device allocation of " + std::to_string(b) + " bytes failed: " +
                                         hipGetErrorString(e));
    }
    P.live[p] = b;
    return p;
}

void dev_free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(g_mu);
    for (auto& dp : g_pools) {
        auto it = dp.second.live.find(p);
        if (it == dp.second.live.end()) continue;
        dp.second.pending.emplace_back(it->second, p);
        dp.second.live.erase(it);
        return;
    }
    (void)hipFree(p);  // not a block of the cache
}

size_t dev_cache_release() {
    std::lock_guard<std::mutex> lock(g_mu);
    int cur = 0;
    (void)hipGetDevice(&cur);
    size_t freed = 0;
    for (auto& dp : g_pools) {
        if (hipSetDevice(dp.first) != hipSuccess) continue;
        settle(dp.second);
        freed += drop_ready(dp.second);
    }
    (void)hipSetDevice(cur);
    return freed;
}

}  // namespace dwg

// Device allocations of the library, counted, and the fault hooks the error
// path tests use: the n-th next allocation (or kernel launch) fails as if
// the device were out of memory (or the launch were refused), so that every
// caller's clean-up and error report can be checked on a healthy device.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <unordered_map>

#include "instance.hpp"

namespace nsship {
namespace {

std::mutex g_mem_mu;
std::unordered_map<void*, size_t> g_live;
long long g_live_bytes = 0;
std::atomic<long long> g_alloc_countdown{-1};
std::atomic<long long> g_launch_countdown{-1};

// True on the armed call: the countdown is decremented by every call while
// it is >= 0, and fires when it passes 0 (it then rests at -1, disarmed).
bool fire(std::atomic<long long>& countdown) {
    long long v = countdown.load();
    while (v >= 0)
        if (countdown.compare_exchange_weak(v, v - 1)) return v == 0;
    return false;
}

}  // namespace

hipError_t dev_alloc(void** p, size_t bytes) {
    *p = nullptr;
    if (fire(g_alloc_countdown)) return hipErrorOutOfMemory;
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> lock(g_mem_mu);
        g_live[*p] = bytes;
        g_live_bytes += static_cast<long long>(bytes);
    }
    return e;
}

void dev_free(void* p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lock(g_mem_mu);
        const auto it = g_live.find(p);
        if (it != g_live.end()) {
            g_live_bytes -= static_cast<long long>(it->second);
            g_live.erase(it);
        }
    }
    (void)hipFree(p);
}

bool launch_fault() { return fire(g_launch_countdown); }

}  // namespace nsship

extern "C" int nss_hip_debug_fail_alloc(int n) {
    nsship::g_alloc_countdown.store(n < 0 ? -1 : n);
    return NSS_HIP_OK;
}

extern "C" int nss_hip_debug_fail_launch(int n) {
    nsship::g_launch_countdown.store(n < 0 ? -1 : n);
    return NSS_HIP_OK;
}

extern "C" int nss_hip_debug_live_allocations(long long* count, long long* bytes) {
    std::lock_guard<std::mutex> lock(nsship::g_mem_mu);
    if (count) *count = static_cast<long long>(nsship::g_live.size());
    if (bytes) *bytes = nsship::g_live_bytes;
    return NSS_HIP_OK;
}

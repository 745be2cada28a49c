// pool.hip — the process-wide cache of device blocks behind Arena and of
// the small pinned host areas behind Ctx.  Solves hand their blocks back
// here when they end (after synchronising their stream), and the next solve
// on the same device takes them instead of calling hipMalloc; hipFree (which
// waits for the whole device) stays out of the solve loop.
#include <cstdlib>
#include <map>
#include <mutex>

#include "common.hpp"

namespace rama {
namespace pool {

namespace {
std::mutex g_mu;
struct DevCache {
    std::multimap<size_t, void*> blocks;
    size_t bytes = 0;
};
std::map<int, DevCache> g_dev;
std::vector<void*> g_pinned;

size_t bound() {
    static const size_t b = [] {
        const char* e = std::getenv("RAMA_CACHE_BYTES");
        return e ? (size_t)std::strtoull(e, nullptr, 10) : ((size_t)32 << 30);
    }();
    return b;
}
}  // namespace

void* take(int device, size_t cls) {
    std::lock_guard<std::mutex> l(g_mu);
    auto d = g_dev.find(device);
    if (d == g_dev.end()) return nullptr;
    auto it = d->second.blocks.find(cls);
    if (it == d->second.blocks.end()) return nullptr;
    void* p = it->second;
    d->second.blocks.erase(it);
    d->second.bytes -= cls;
    return p;
}

void give(int device, size_t cls, void* p) {
    {
        std::lock_guard<std::mutex> l(g_mu);
        DevCache& c = g_dev[device];
        if (c.bytes + cls <= bound()) {
            c.blocks.emplace(cls, p);
            c.bytes += cls;
            return;
        }
    }
    (void)hipFree(p);
}

void release(int device) {
    std::multimap<size_t, void*> drop;
    {
        std::lock_guard<std::mutex> l(g_mu);
        for (auto& kv : g_dev) {
            if (device >= 0 && kv.first != device) continue;
            drop.insert(kv.second.blocks.begin(), kv.second.blocks.end());
            kv.second.blocks.clear();
            kv.second.bytes = 0;
        }
    }
    for (auto& kv : drop) (void)hipFree(kv.second);
}

void* take_pinned() {
    {
        std::lock_guard<std::mutex> l(g_mu);
        if (!g_pinned.empty()) {
            void* p = g_pinned.back();
            g_pinned.pop_back();
            return p;
        }
    }
    void* p = nullptr;
    RAMA_HIP(hipHostMalloc(&p, kPinnedBytes + kMirrorBytes, hipHostMallocPortable | hipHostMallocMapped));
    return p;
}

void give_pinned(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> l(g_mu);
    g_pinned.push_back(p);
}

}  // namespace pool
}  // namespace rama

extern "C" void rama_release_cached(int device) {
    if (device >= 0) {
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != device) {
            (void)hipSetDevice(device);
            (void)hipDeviceSynchronize();
            rama::pool::release(device);
            (void)hipSetDevice(cur);
            (void)hipGetLastError();
            return;
        }
    }
    (void)hipDeviceSynchronize();
    rama::pool::release(device);
    (void)hipGetLastError();
}

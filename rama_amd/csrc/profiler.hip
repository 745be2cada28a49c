// profiler.hip — event bookkeeping behind profiler.hpp and its C ABI:
//   rama_profile_enable(on), rama_profile_reset(),
//   rama_profile_count(), rama_profile_entry(k, name, launches, ms, bytes),
//   rama_profile_work(k, units, entries).
#include <atomic>
#include <cstdio>
#include <map>
#include <string>
#include <cstdlib>
#include <mutex>

#include "profiler.hpp"

namespace rama {

namespace {
const bool g_sync_sites = std::getenv("RAMA_SYNC_SITES") != nullptr;
struct SyncSites {
    std::mutex mu;
    std::map<std::pair<std::string, int>, long long> n;
    ~SyncSites() {
        for (auto& kv : n)
            std::fprintf(stderr, "sync %-40s line %5d  %lld\n", kv.first.first.c_str(), kv.first.second, kv.second);
    }
};
SyncSites& sync_sites() {
    static SyncSites s;
    return s;
}
}  // namespace

bool sync_sites_on() { return g_sync_sites; }
void note_sync_site(const char* fn, int line) {
    auto& s = sync_sites();
    std::lock_guard<std::mutex> l(s.mu);
    ++s.n[{fn, line}];
}

namespace prof {

namespace {
std::atomic<bool> g_on{false};
std::mutex g_mu;
thread_local const char* tl_site = nullptr;
const bool g_sites = std::getenv("RAMA_PROFILE_SITES") != nullptr;
struct Pending {
    std::string name;
    hipEvent_t a, b;
    double bytes, units, entries;
};
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_free;
struct Agg {
    std::string name;
    int64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
    double units = 0.0, entries = 0.0;
};
std::vector<Agg> g_agg;

void resolve_locked() {
    for (auto& p : g_pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            Agg* a = nullptr;
            for (auto& x : g_agg)
                if (x.name == p.name) a = &x;
            if (!a) {
                g_agg.push_back(Agg{p.name});
                a = &g_agg.back();
            }
            a->launches += 1;
            a->ms += ms;
            a->bytes += p.bytes;
            a->units += p.units;
            a->entries += p.entries;
        }
        (void)hipGetLastError();
        g_free.push_back(p.a);
        g_free.push_back(p.b);
    }
    g_pending.clear();
}
}  // namespace

bool enabled() { return g_on.load(std::memory_order_relaxed); }

std::string g_only;  // guarded by g_mu for writes; read while enabled
std::atomic<bool> g_filtered{false};

bool wants(const char* name) {
    if (!enabled()) return false;
    if (!g_filtered.load(std::memory_order_relaxed)) return true;
    std::lock_guard<std::mutex> l(g_mu);
    return g_only == name;
}

const char* current_site() { return tl_site; }
Site::Site(const char* s) : prev_(tl_site) { tl_site = s; }
Site::~Site() { tl_site = prev_; }

hipEvent_t take_event() {
    std::lock_guard<std::mutex> l(g_mu);
    if (!g_free.empty()) {
        hipEvent_t e = g_free.back();
        g_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    RAMA_HIP(hipEventCreate(&e));
    return e;
}

void push(const char* name, hipEvent_t a, hipEvent_t b, double bytes, double units, double entries) {
    std::lock_guard<std::mutex> l(g_mu);
    std::string key(name);
    if (g_sites && tl_site) key = key + "@" + tl_site;
    g_pending.push_back(Pending{std::move(key), a, b, bytes, units, entries});
    if (g_pending.size() > 4096) resolve_locked();
}

}  // namespace prof
}  // namespace rama

using namespace rama::prof;

extern "C" {

void rama_profile_enable(int on) {
    std::lock_guard<std::mutex> l(g_mu);
    resolve_locked();
    g_on.store(on != 0);
}

void rama_profile_only(const char* name) {
    std::lock_guard<std::mutex> l(g_mu);
    g_only = name ? name : "";
    g_filtered.store(!g_only.empty());
}

void rama_profile_reset(void) {
    std::lock_guard<std::mutex> l(g_mu);
    resolve_locked();
    g_agg.clear();
}

int64_t rama_profile_count(void) {
    std::lock_guard<std::mutex> l(g_mu);
    resolve_locked();
    return (int64_t)g_agg.size();
}

int rama_profile_entry(int64_t k, const char** name, int64_t* launches, double* ms, double* bytes) {
    std::lock_guard<std::mutex> l(g_mu);
    resolve_locked();
    if (k < 0 || k >= (int64_t)g_agg.size()) return RAMA_ERR_INVALID;
    *name = g_agg[k].name.c_str();
    *launches = g_agg[k].launches;
    *ms = g_agg[k].ms;
    *bytes = g_agg[k].bytes;
    return 0;
}

int rama_profile_work(int64_t k, double* units, double* entries) {
    std::lock_guard<std::mutex> l(g_mu);
    resolve_locked();
    if (k < 0 || k >= (int64_t)g_agg.size() || !units || !entries) return RAMA_ERR_INVALID;
    *units = g_agg[k].units;
    *entries = g_agg[k].entries;
    return 0;
}

}  // extern "C"

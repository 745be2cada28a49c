// common.hpp — runtime plumbing shared by every translation unit of
// librama_hip.so: HIP error handling, a caching device allocator, typed
// device buffers and the per-solve context (device, stream, pinned scalars).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/rama_hip.h"

namespace rama {

using i64 = int64_t;
using u64 = uint64_t;
using u32 = uint32_t;

// Error carried to the C ABI as (status, message).
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define RAMA_HIP(expr)                                                                          \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            throw ::rama::Error(_e == hipErrorOutOfMemory ? RAMA_ERR_OOM : RAMA_ERR_HIP,        \
                                std::string(#expr " failed: ") + hipGetErrorString(_e) + " at " \
                                + __FILE__ + ":" + std::to_string(__LINE__));                   \
    } while (0)

#define RAMA_REQUIRE(cond, msg)                                       \
    do {                                                              \
        if (!(cond)) throw ::rama::Error(RAMA_ERR_INVALID, (msg));    \
    } while (0)

// Host synchronisations counted per calling function and line when
// RAMA_SYNC_SITES is set, printed at exit (diagnostic).
bool sync_sites_on();
void note_sync_site(const char* fn, int line);

// Kernel launches made by this thread (diagnostic).
inline int64_t& launch_count() {
    static thread_local int64_t n = 0;
    return n;
}

// Launch-error check after a kernel launch.
#define RAMA_LAUNCHED()                  \
    do {                                 \
        ++::rama::launch_count();        \
        RAMA_HIP(hipGetLastError());     \
    } while (0)

inline unsigned grid_for(i64 n, int block = 256) {
    i64 g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// Process-wide cache of device blocks that finished solves handed back
// (their streams were synchronised first), so the next solve on the device
// reuses them instead of calling hipMalloc/hipFree.  Bounded by
// RAMA_CACHE_BYTES (default 32 GiB per device); rama_release_cached() and
// allocation failures empty it.
namespace pool {
constexpr size_t kPinnedBytes = 4096;  // read-back staging
constexpr size_t kMirrorBytes = 256;   // device-written result slots (Ctx::mirror)
void* take_pinned();                   // kPinnedBytes + kMirrorBytes, mapped into every device
void give_pinned(void* p);
void* take(int device, size_t cls);
void give(int device, size_t cls, void* p);  // frees when over the bound
void release(int device);                   // device < 0: every device
}  // namespace pool

// Caching device allocator.  Blocks are rounded to a size class (1/8 steps
// between powers of two) and recycled on the owning stream's order: a block
// released while kernels that use it are queued is only handed to later work
// on the same stream, so no synchronisation is needed.
class Arena {
   public:
    explicit Arena(int device) : device_(device) {}
    ~Arena() { trim(); }
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;

    // Fixed-region mode: every block is carved out of [base, base + bytes)
    // (first fit over the free extents in address order, 256-byte aligned,
    // neighbours merged on release); the driver and the process-wide cache
    // are never asked.  Exhausting the region is an RAMA_ERR_OOM error
    // raised on the host before any kernel sees the block.
    void use_region(void* base, size_t bytes) {
        rbase_ = static_cast<char*>(base);
        rbytes_ = bytes & ~(size_t)255;
        rfree_.clear();
        if (rbytes_) rfree_.emplace(0, rbytes_);
    }
    bool fixed() const { return rbase_ != nullptr; }

    void* alloc(size_t bytes) {
        if (bytes == 0) bytes = 1;
        if (rbase_) return region_alloc(bytes);
        const size_t cls = size_class(bytes);
        auto it = free_.find(cls);
        void* p = nullptr;
        if (it != free_.end()) {
            p = it->second;
            free_.erase(it);
        } else if ((p = pool::take(device_, cls)) != nullptr) {
            reserved_ += cls;
        } else {
            hipError_t e = hipMalloc(&p, cls);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                free_blocks();
                pool::release(device_);
                e = hipMalloc(&p, cls);
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error(RAMA_ERR_OOM, "device allocation of " + std::to_string(cls) + " bytes failed");
            }
            reserved_ += cls;
        }
        used_[p] = cls;
        live_ += cls;
        if (live_ > peak_) peak_ = live_;
        return p;
    }

    void release(void* p) {
        if (!p) return;
        auto it = used_.find(p);
        if (it == used_.end()) return;
        if (rbase_) {
            region_free(static_cast<char*>(p) - rbase_, it->second);
            live_ -= it->second;
            used_.erase(it);
            return;
        }
        free_.emplace(it->second, p);
        live_ -= it->second;
        used_.erase(it);
    }

    // Hand cached blocks to the process-wide cache.  Only call once the
    // stream that used them has been synchronised.
    void trim() {
        for (auto& kv : free_) {
            pool::give(device_, kv.first, kv.second);
            reserved_ -= kv.first;
        }
        free_.clear();
    }

    // Return cached blocks to the driver.
    void free_blocks() {
        for (auto& kv : free_) {
            (void)hipFree(kv.second);
            reserved_ -= kv.first;
        }
        free_.clear();
    }

    size_t peak() const { return peak_; }
    size_t live() const { return live_; }
    size_t reserved() const { return reserved_; }
    void reset_peak() { peak_ = live_; }

    static size_t size_class(size_t b) {
        if (b <= 4096) return 4096;
        size_t p = 1;
        while (p < b) p <<= 1;
        const size_t step = p >> 3;
        return ((b + step - 1) / step) * step;
    }

   private:
    void* region_alloc(size_t bytes) {
        const size_t need = (bytes + 255) & ~(size_t)255;
        for (auto it = rfree_.begin(); it != rfree_.end(); ++it) {
            if (it->second < need) continue;
            const size_t off = it->first, len = it->second;
            rfree_.erase(it);
            if (len > need) rfree_.emplace(off + need, len - need);
            void* p = rbase_ + off;
            used_[p] = need;
            live_ += need;
            if (live_ > peak_) peak_ = live_;
            return p;
        }
        throw Error(RAMA_ERR_OOM, "fixed device region of " + std::to_string(rbytes_) + " bytes exhausted: " +
                                      std::to_string(live_) + " bytes in use, " + std::to_string(need) +
                                      " more requested");
    }
    void region_free(size_t off, size_t len) {
        auto next = rfree_.lower_bound(off);
        if (next != rfree_.end() && off + len == next->first) {
            len += next->second;
            next = rfree_.erase(next);
        }
        if (next != rfree_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += len;
                return;
            }
        }
        rfree_.emplace(off, len);
    }

    int device_;
    std::multimap<size_t, void*> free_;
    std::unordered_map<void*, size_t> used_;
    size_t live_ = 0, peak_ = 0, reserved_ = 0;
    char* rbase_ = nullptr;               // fixed region (nullptr: caching mode)
    size_t rbytes_ = 0;
    std::map<size_t, size_t> rfree_;      // free extents of the region: offset -> length
};

// Typed device buffer owned by an Arena (move-only).
template <class T>
class DBuf {
   public:
    DBuf() = default;
    DBuf(Arena& a, i64 n) : a_(&a), n_(n) { p_ = n > 0 ? static_cast<T*>(a.alloc(sizeof(T) * (size_t)n)) : nullptr; }
    ~DBuf() { reset(); }
    DBuf(DBuf&& o) noexcept : a_(o.a_), p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            reset();
            a_ = o.a_; p_ = o.p_; n_ = o.n_;
            o.p_ = nullptr; o.n_ = 0;
        }
        return *this;
    }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;

    void reset() {
        if (p_ && a_) a_->release(p_);
        p_ = nullptr;
        n_ = 0;
    }
    // Logical shrink: keeps the allocation, changes the element count.
    void set_size(i64 n) { n_ = n; }
    T* get() const { return p_; }
    i64 size() const { return n_; }
    bool empty() const { return n_ == 0; }
    operator T*() const { return p_; }

   private:
    Arena* a_ = nullptr;
    T* p_ = nullptr;
    i64 n_ = 0;
};

// Per-solve execution context.
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    Arena arena;
    bool verbose = false;
    void* pinned = nullptr;  // small pinned host area for scalar read-back
    void* mirror_dev = nullptr;  // device address of the slots after it (null: allocation-free mode)
    // The last conflicted-cycle search on this context found no cycle (a
    // hint that the next dual pass leaves its graph unchanged).
    bool last_cycles_empty = false;

    Ctx(int dev, hipStream_t s, bool verb) : device(dev), arena(dev), verbose(verb) {
        RAMA_HIP(hipSetDevice(dev));
        if (s) {
            stream = s;
        } else {
            RAMA_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            own_stream = true;
        }
        pinned = pool::take_pinned();
        void* base = nullptr;
        RAMA_HIP(hipHostGetDevicePointer(&base, pinned, 0));
        mirror_dev = static_cast<unsigned char*>(base) + pool::kPinnedBytes;
    }
    // Allocation-free context: device memory comes only from the fixed
    // region [region, region + region_bytes), work runs on the caller's
    // stream as given (the null stream included), scalar read-backs land in
    // host memory inside the context and waits are stream synchronisations,
    // so nothing is allocated, created or pinned.
    Ctx(int dev, hipStream_t s, bool verb, void* region, size_t region_bytes)
        : device(dev), stream(s), arena(dev), verbose(verb), no_alloc(true) {
        RAMA_HIP(hipSetDevice(dev));
        arena.use_region(region, region_bytes);
        pinned = host_scalars;
    }
    ~Ctx() {
        if (stream || no_alloc) (void)hipStreamSynchronize(stream);
        drop_graphs();
        if (sync_ev) (void)hipEventDestroy(sync_ev);
        arena.trim();
        if (!no_alloc) pool::give_pinned(pinned);
        if (up_pinned) (void)hipHostFree(up_pinned);
        if (own_stream) (void)hipStreamDestroy(stream);
    }
    const bool no_alloc = false;
    alignas(64) unsigned char host_scalars[pool::kPinnedBytes];  // read-back area of an allocation-free context
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;

    template <class T>
    DBuf<T> buf(i64 n) { return DBuf<T>(arena, n); }

    int64_t syncs = 0;  // host round trips (diagnostic)

    // Onesweep look-back state: per-tile status words tagged with an epoch
    // so no launch needs the buffer cleared; a tile counter that the last
    // tile resets; a device error flag raised by a look-back that gave up.
    DBuf<uint64_t> scan_status;
    DBuf<int> scan_counter;  // [0] tile counter, [1] error flag
    uint32_t scan_epoch = 0;
    // Partials of the two-launch scan.
    DBuf<uint64_t> scan_partials;
    void* scan_scratch(size_t bytes) {
        if ((size_t)scan_partials.size() * 8 < bytes) scan_partials = DBuf<uint64_t>(arena, (int64_t)(bytes / 8 + 1));
        return scan_partials.get();
    }
    // Reduction partials and the ticket counter (reset by the last block).
    DBuf<uint64_t> red_partials;
    DBuf<unsigned> red_ticket;
    void* reduce_scratch(size_t bytes) {
        if ((size_t)red_partials.size() * 8 < bytes) red_partials = DBuf<uint64_t>(arena, (int64_t)(bytes / 8 + 1));
        return red_partials.get();
    }
    unsigned* reduce_ticket() {
        if (red_ticket.empty()) {
            red_ticket = DBuf<unsigned>(arena, 1);
            zero(red_ticket.get(), 1);
        }
        return red_ticket.get();
    }
    // Returns the status buffer (at least `tiles` words) and the epoch to use.
    uint64_t* scan_begin(int64_t tiles, uint32_t* epoch) {
        if (scan_counter.empty()) {
            scan_counter = DBuf<int>(arena, 2);
            zero(scan_counter.get(), 2);
        }
        if (scan_status.size() < tiles) {
            int64_t cap = 1024;
            while (cap < tiles) cap <<= 1;
            scan_status = DBuf<uint64_t>(arena, cap);
            zero(scan_status.get(), cap);
            scan_epoch = 0;
        }
        if (++scan_epoch >= 65535u) {
            zero(scan_status.get(), scan_status.size());
            scan_epoch = 1;
        }
        *epoch = scan_epoch;
        return scan_status.get();
    }
    // Host waits go through an event recorded on the stream and created
    // without hipEventBlockingSync, which the runtime waits on actively;
    // RAMA_SYNC_MODE=stream waits on the stream instead.
    hipEvent_t sync_ev = nullptr;
    void sync(const char* fn = __builtin_FUNCTION(), int line = __builtin_LINE()) {
        ++syncs;
        if (sync_sites_on()) note_sync_site(fn, line);
        static const bool by_stream = [] {
            const char* e = std::getenv("RAMA_SYNC_MODE");
            return e && std::strcmp(e, "stream") == 0;
        }();
        if (by_stream || no_alloc) {
            RAMA_HIP(hipStreamSynchronize(stream));
            drop_graphs();
            return;
        }
        if (!sync_ev) RAMA_HIP(hipEventCreateWithFlags(&sync_ev, hipEventDisableTiming));
        RAMA_HIP(hipEventRecord(sync_ev, stream));
        RAMA_HIP(hipEventSynchronize(sync_ev));
        drop_graphs();
    }

    // Runs `body` `times` times in a row.  body only launches kernels on
    // this stream (no host waits, no memsets, no allocation).  With
    // use_graph the first run is captured into a HIP graph and every run
    // replays it: one launch per run instead of one per kernel.  Capture is
    // skipped on the null stream, in the allocation-free mode (graph
    // instantiation allocates) and when the stream is being captured by the
    // caller.  The graphs live until the stream has been synchronised.
    template <class F>
    void repeat(int times, bool use_graph, F&& body) {
        if (times <= 0) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (use_graph && times > 1 && !no_alloc && stream != nullptr &&
            hipStreamIsCapturing(stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
            RAMA_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                body();
            } catch (...) {
                if (hipStreamEndCapture(stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
                (void)hipGetLastError();
                throw;
            }
            RAMA_HIP(hipStreamEndCapture(stream, &g));
            hipGraphExec_t ge = nullptr;
            const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                RAMA_HIP(e);
            }
            graphs.push_back(g);
            graph_execs.push_back(ge);
            for (int k = 0; k < times; ++k) RAMA_HIP(hipGraphLaunch(ge, stream));
            return;
        }
        (void)hipGetLastError();
        for (int k = 0; k < times; ++k) body();
    }
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> graph_execs;
    void drop_graphs() {
        for (hipGraphExec_t ge : graph_execs) (void)hipGraphExecDestroy(ge);
        for (hipGraph_t g : graphs) (void)hipGraphDestroy(g);
        graph_execs.clear();
        graphs.clear();
    }

    // Host-to-device upload through a pinned staging ring, asynchronous on
    // the stream.  A staging slot is reused only after a synchronisation
    // (every copy queued before it has run); a full ring synchronises.
    void* up_pinned = nullptr;
    size_t up_cap = 0, up_off = 0;
    int64_t up_sync = -1;
    template <class T>
    void upload(T* d, const T* h, i64 n, const char* fn = __builtin_FUNCTION(), int line = __builtin_LINE()) {
        const size_t bytes = sizeof(T) * (size_t)n;
        if (bytes == 0) return;
        if (sync_sites_on()) note_sync_site(fn, -line);  // uploads are counted at negative line numbers
        if (no_alloc) {  // no pinned staging: copy from the caller's memory and wait
            RAMA_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream));
            sync(fn, line);
            return;
        }
        const size_t need = (bytes + 255) & ~(size_t)255;
        if (up_sync != syncs) {
            up_off = 0;
            up_sync = syncs;
        }
        if (up_cap < need) {
            if (up_pinned) {
                sync();
                (void)hipHostFree(up_pinned);
                up_pinned = nullptr;
            }
            up_cap = std::max<size_t>((size_t)1 << 20, 2 * need);
            RAMA_HIP(hipHostMalloc(&up_pinned, up_cap, hipHostMallocDefault));
            up_off = 0;
            up_sync = syncs;
        } else if (up_off + need > up_cap) {
            sync();
            up_off = 0;
            up_sync = syncs;
        }
        char* slot = static_cast<char*>(up_pinned) + up_off;
        std::memcpy(slot, h, bytes);
        RAMA_HIP(hipMemcpyAsync(d, slot, bytes, hipMemcpyHostToDevice, stream));
        up_off += need;
    }

    // Result slots in pinned host memory mapped into the device's address
    // space.  A kernel whose result the host reads next stores it here as
    // well as in device memory, so the read-back is a wait with no copy
    // behind it (device stores to this memory are visible to the host once
    // the stream has been synchronised).  Null in the allocation-free mode,
    // whose host scalars are ordinary memory: callers then read the device
    // copy.  One set of slots per context: a result is read before the next
    // kernel that stores one is queued.
    template <class T>
    T* mirror() {
        static_assert(sizeof(T) <= pool::kMirrorBytes, "mirror slot too small");
        static const bool off = [] {  // RAMA_READ_MIRROR=0: read every result back by a copy
            const char* e = std::getenv("RAMA_READ_MIRROR");
            return e && std::strcmp(e, "0") == 0;
        }();
        return off ? nullptr : static_cast<T*>(mirror_dev);
    }
    // Host address of the mirror slots (for a value the host sets itself
    // when no kernel runs to store it).
    void* mirror_host() { return static_cast<unsigned char*>(pinned) + pool::kPinnedBytes; }
    // The n values a kernel stored at m = mirror<T>() (synchronises), or,
    // when m is null, the n device values at d.
    template <class T>
    void read_mirror(const T* d, const T* m, T* h, int n = 1, const char* fn = __builtin_FUNCTION(),
                     int line = __builtin_LINE()) {
        if (!m) {
            read(d, h, n, fn, line);
            return;
        }
        RAMA_REQUIRE(sizeof(T) * (size_t)n <= pool::kMirrorBytes, "mirror read too large");
        sync(fn, line);
        std::memcpy(h, mirror_host(), sizeof(T) * n);
    }

    // A flag that the next kernel raises (threads that find something store
    // 1) and the host then reads with flag_value: the mirror slot, cleared
    // by the host (no memset launch, no copy back), or in the
    // allocation-free mode the device word d, cleared on the stream.
    int* flag_slot(int* d) {
        int* m = mirror<int>();
        if (!m) {
            zero(d, 1);
            return d;
        }
        *static_cast<int*>(mirror_host()) = 0;
        return m;
    }
    int flag_value(const int* d, const int* f, const char* fn = __builtin_FUNCTION(), int line = __builtin_LINE()) {
        int h;
        read_mirror(d, f == d ? nullptr : f, &h, 1, fn, line);
        return h;
    }

    // Read `n` device scalars (contiguous) into host memory (synchronises).
    template <class T>
    void read(const T* d, T* h, int n = 1, const char* fn = __builtin_FUNCTION(), int line = __builtin_LINE()) {
        RAMA_REQUIRE(sizeof(T) * (size_t)n <= pool::kPinnedBytes, "scalar read-back too large");
        RAMA_HIP(hipMemcpyAsync(pinned, d, sizeof(T) * n, hipMemcpyDeviceToHost, stream));
        sync(fn, line);
        std::memcpy(h, pinned, sizeof(T) * n);
    }
    template <class T>
    T read1(const T* d, const char* fn = __builtin_FUNCTION(), int line = __builtin_LINE()) {
        T v;
        read(d, &v, 1, fn, line);
        return v;
    }
    template <class T>
    void write1(T* d, T v, const char* fn = __builtin_FUNCTION(), int line = __builtin_LINE()) {
        RAMA_HIP(hipMemcpyAsync(d, &v, sizeof(T), hipMemcpyHostToDevice, stream));
        sync(fn, line);
    }
    template <class T>
    void h2d(T* d, const T* h, i64 n) {
        if (n > 0) RAMA_HIP(hipMemcpyAsync(d, h, sizeof(T) * (size_t)n, hipMemcpyHostToDevice, stream));
    }
    template <class T>
    void d2h(T* h, const T* d, i64 n, const char* fn = __builtin_FUNCTION(), int line = __builtin_LINE()) {
        if (n > 0) RAMA_HIP(hipMemcpyAsync(h, d, sizeof(T) * (size_t)n, hipMemcpyDeviceToHost, stream));
        sync(fn, line);
    }
    template <class T>
    void d2d(T* dst, const T* src, i64 n) {
        if (n > 0) RAMA_HIP(hipMemcpyAsync(dst, src, sizeof(T) * (size_t)n, hipMemcpyDeviceToDevice, stream));
    }
    template <class T>
    void zero(T* d, i64 n) {
        if (n > 0) RAMA_HIP(hipMemsetAsync(d, 0, sizeof(T) * (size_t)n, stream));
    }
};

}  // namespace rama

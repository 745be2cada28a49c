// tuning.hip — the launch-shape table (tuning.hpp) and its C ABI, plus the
// sort timing entry the tuner measures sort paths with.
#include <dlfcn.h>

#include <chrono>
#include <fstream>
#include <sstream>

#include "rama_internal.hpp"
#include "tuning.hpp"

namespace rama {

void set_last_error(const std::string& s);

namespace tune {
namespace {

struct Row {
    int64_t max_size;  // -1: no upper limit
    int64_t value;
};

struct Spec {
    const char* name;
    int64_t lo, hi;  // accepted values
    std::vector<Row> defaults;
};

// Built-in defaults: the shapes measured best on the 12.47M-edge grid solve
// before any table existed.
const Spec& spec(int k) {
    static const Spec specs[KnobCount] = {
        {"sort_path_u64", 0, 2, {{(int64_t)64 * kRsTile, 2}, {-1, 0}}},
        {"sort_path_u32", 0, 2, {{(int64_t)64 * kRsTile, 2}, {-1, 0}}},
        {"sort_threads", 512, 1024, {{-1, 1024}}},
        {"scan_chunks", 64, 8192, {{-1, 2048}}},
        {"pick_blocks", 64, 65536, {{-1, 2048}}},
        {"match_blocks", 64, 65536, {{-1, 1024}}},
        {"mp_fused_max_blocks", 0, 65536, {{-1, 256}}},
        {"cycle_chunk_pct", 100, 1000, {{-1, 125}}},
    };
    return specs[k];
}

struct Table {
    std::vector<Row> rows[KnobCount];
    std::string source = "defaults";
};

std::mutex g_mu;
Table g_table;
bool g_loaded = false;

int knob_of(const std::string& s) {
    for (int k = 0; k < KnobCount; ++k)
        if (s == spec(k).name) return k;
    return -1;
}

void check_value(int k, int64_t v) {
    const Spec& sp = spec(k);
    RAMA_REQUIRE(v >= sp.lo && v <= sp.hi, std::string("tuning value out of range for ") + sp.name + ": " +
                                               std::to_string(v));
    if (k == SortThreads) RAMA_REQUIRE(v == 512 || v == 1024, "sort_threads must be 512 or 1024");
}

// Adds (or replaces) a row, keeping the rows ordered with -1 last.
void put_row(Table& t, int k, int64_t max_size, int64_t value) {
    RAMA_REQUIRE(max_size == -1 || max_size >= 0, "max_size must be -1 or >= 0");
    check_value(k, value);
    auto& rows = t.rows[k];
    for (auto& r : rows)
        if (r.max_size == max_size) {
            r.value = value;
            return;
        }
    rows.push_back(Row{max_size, value});
    std::sort(rows.begin(), rows.end(), [](const Row& a, const Row& b) {
        const uint64_t x = (uint64_t)a.max_size, y = (uint64_t)b.max_size;  // -1 sorts last
        return x < y;
    });
}

// Table file: '#' starts a comment; each other line is "knob max_size value".
Table parse(const std::string& path) {
    std::ifstream f(path);
    RAMA_REQUIRE((bool)f, "cannot open tuning table " + path);
    Table t;
    t.source = path;
    std::string line;
    int ln = 0;
    while (std::getline(f, line)) {
        ++ln;
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream is(line);
        std::string name;
        if (!(is >> name)) continue;
        int64_t ms = 0, v = 0;
        std::string extra;
        RAMA_REQUIRE((bool)(is >> ms >> v) && !(is >> extra),
                     path + ":" + std::to_string(ln) + ": expected 'knob max_size value'");
        const int k = knob_of(name);
        RAMA_REQUIRE(k >= 0, path + ":" + std::to_string(ln) + ": unknown knob " + name);
        put_row(t, k, ms, v);
    }
    return t;
}

std::string default_path() {
    const char* e = std::getenv("RAMA_TUNING_TABLE");
    if (e) return std::string(e);
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&default_path), &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        const size_t s = p.rfind('/');
        const std::string dir = s == std::string::npos ? std::string(".") : p.substr(0, s);
        return dir + "/tuning/gfx950.table";
    }
    return "";
}

// First use: the table file if there is one (a malformed file is reported
// on stderr once and the defaults apply).
void ensure_loaded_locked() {
    if (g_loaded) return;
    g_loaded = true;
    const std::string p = default_path();
    if (p.empty() || p == "none") return;
    std::ifstream probe(p);
    if (!probe) return;
    try {
        g_table = parse(p);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "rama: tuning table ignored: %s\n", ex.what());
        g_table = Table{};
        g_table.source = "defaults (" + p + " rejected)";
    }
}

int64_t lookup(const std::vector<Row>& rows, int64_t size) {
    for (const Row& r : rows)
        if (r.max_size < 0 || size <= r.max_size) return r.value;
    return -1;
}

}  // namespace

int64_t get(Knob k, int64_t size) {
    std::lock_guard<std::mutex> lk(g_mu);
    ensure_loaded_locked();
    const int64_t v = lookup(g_table.rows[k], size);
    return v >= 0 ? v : lookup(spec(k).defaults, size);
}

const char* name(Knob k) { return spec(k).name; }

}  // namespace tune

// The sort timing entry: random keys of `bits` bits (a hash of the index and
// the repetition), stable-sorted by the requested path and thread count.
namespace {

template <class K>
__global__ void k_tune_fill(K* keys, u32* vals, i64 n, int bits, u32 rep) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    const u64 mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u64 x = (u64)i * 0x9E3779B97F4A7C15ull + (u64)rep * 0xBF58476D1CE4E5B9ull + 0x94D049BB133111EBull;
        x ^= x >> 31;
        x *= 0xD6E8FEB86659FD93ull;
        x ^= x >> 32;
        keys[i] = (K)(x & mask);
        vals[i] = (u32)i;
    }
}

// Streaming copy: 16-byte non-temporal loads, U of them in flight per
// thread, grid-stride over whole vectors.  time_copy tries a few grid sizes
// and unroll depths and keeps the fastest: the practical HBM rate the
// roofline fractions are read against.
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
template <int U>
__global__ __launch_bounds__(kBlock) void k_stream_copy(const v4u* __restrict__ src, v4u* __restrict__ dst, i64 n16) {
    const i64 stride = (i64)gridDim.x * kBlock;
    i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    for (; i + (U - 1) * stride < n16; i += U * stride) {
        v4u v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(&src[i + u * stride]);
#pragma unroll
        for (int u = 0; u < U; ++u) dst[i + u * stride] = v[u];
    }
    for (; i < n16; i += stride) dst[i] = __builtin_nontemporal_load(&src[i]);
}

double time_copy(i64 bytes, int reps) {
    Ctx ctx(0, nullptr, false);
    const i64 n16 = bytes / 16;
    DBuf<v4u> a = ctx.buf<v4u>(n16), b = ctx.buf<v4u>(n16);
    int cus = 0;
    RAMA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx.device));
    hipEvent_t e0, e1;
    RAMA_HIP(hipEventCreate(&e0));
    RAMA_HIP(hipEventCreate(&e1));
    double best = 0.0;
    try {
        RAMA_HIP(hipMemsetAsync(a.get(), 1, (size_t)n16 * 16, ctx.stream));
        for (int shape = 0; shape < 8; ++shape) {
            // shapes 0-5: bounded grids (4, 8, 16 blocks per CU); 6-7: one
            // pass, every thread copying U vectors once
            const int per_cu = (shape % 3 == 0) ? 4 : (shape % 3 == 1 ? 8 : 16);
            const bool deep = shape >= 6 ? shape == 7 : shape >= 3;
            const i64 full = (n16 + (i64)kBlock * (deep ? 8 : 4) - 1) / ((i64)kBlock * (deep ? 8 : 4));
            const unsigned grid = shape >= 6 ? (unsigned)std::min<i64>(std::max<i64>(full, 1), (i64)1 << 30)
                                             : (unsigned)std::max(1, per_cu * cus);
            double ms_total = 0.0;
            for (int r = -1; r < reps; ++r) {  // one untimed run first
                RAMA_HIP(hipEventRecord(e0, ctx.stream));
                if (deep)
                    k_stream_copy<8><<<grid, kBlock, 0, ctx.stream>>>(a.get(), b.get(), n16);
                else
                    k_stream_copy<4><<<grid, kBlock, 0, ctx.stream>>>(a.get(), b.get(), n16);
                RAMA_LAUNCHED();
                RAMA_HIP(hipEventRecord(e1, ctx.stream));
                RAMA_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                RAMA_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (r >= 0) ms_total += ms;
            }
            best = std::max(best, 2.0 * (double)n16 * 16.0 * reps / (ms_total * 1e-3) / 1e9);
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

template <class K>
double time_sort(i64 n, int bits, int path, int threads, int reps) {
    Ctx ctx(0, nullptr, false);
    DBuf<K> keys = ctx.buf<K>(n), tk = ctx.buf<K>(n);
    DBuf<u32> vals = ctx.buf<u32>(n), tv = ctx.buf<u32>(n);
    hipEvent_t e0, e1;
    RAMA_HIP(hipEventCreate(&e0));
    RAMA_HIP(hipEventCreate(&e1));
    double total = 0.0;
    try {
        for (int r = -1; r < reps; ++r) {  // one untimed run first
            k_tune_fill<K><<<grid_for(std::min<i64>(n, (i64)1 << 22)), kBlock, 0, ctx.stream>>>(
                keys.get(), vals.get(), n, bits, (u32)(r + 1));
            RAMA_LAUNCHED();
            RAMA_HIP(hipEventRecord(e0, ctx.stream));
            radix_sort_pairs_shaped<K>(ctx, keys.get(), vals.get(), tk.get(), tv.get(), n, bits, path, threads);
            RAMA_HIP(hipEventRecord(e1, ctx.stream));
            RAMA_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            RAMA_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (r >= 0) total += ms;
        }
        scan_check(ctx);
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return total / std::max(reps, 1);
}

template <class F>
int guarded_call(F&& f) {
    try {
        set_last_error("");
        f();
        return RAMA_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return RAMA_ERR_INTERNAL;
    }
}

thread_local std::string g_source_copy;

}  // namespace
}  // namespace rama

extern "C" {

int rama_tuning_load(const char* path) {
    using namespace rama;
    return guarded_call([&] {
        tune::Table t;
        if (path && *path) t = tune::parse(path);
        std::lock_guard<std::mutex> lk(tune::g_mu);
        tune::g_table = std::move(t);
        tune::g_loaded = true;
    });
}

int rama_tuning_set(const char* knob, int64_t max_size, int64_t value) {
    using namespace rama;
    return guarded_call([&] {
        const int k = tune::knob_of(knob ? knob : "");
        RAMA_REQUIRE(k >= 0, std::string("unknown tuning knob ") + (knob ? knob : "(null)"));
        std::lock_guard<std::mutex> lk(tune::g_mu);
        tune::ensure_loaded_locked();
        tune::put_row(tune::g_table, k, max_size, value);
        if (tune::g_table.source.rfind("modified ", 0) != 0) tune::g_table.source = "modified " + tune::g_table.source;
    });
}

int64_t rama_tuning_get(const char* knob, int64_t size) {
    using namespace rama;
    const int k = tune::knob_of(knob ? knob : "");
    if (k < 0) return -1;
    return tune::get((tune::Knob)k, size);
}

const char* rama_tuning_source(void) {
    using namespace rama;
    std::lock_guard<std::mutex> lk(tune::g_mu);
    tune::ensure_loaded_locked();
    g_source_copy = tune::g_table.source;
    return g_source_copy.c_str();
}

int rama_tune_time_sort(int64_t n, int32_t key_bytes, int32_t bits, int32_t path, int32_t threads, int32_t reps,
                        double* ms_out) {
    using namespace rama;
    return guarded_call([&] {
        RAMA_REQUIRE(ms_out != nullptr, "NULL argument");
        RAMA_REQUIRE(n >= 2 && n < ((int64_t)1 << 31), "n out of range");
        RAMA_REQUIRE(key_bytes == 4 || key_bytes == 8, "key_bytes must be 4 or 8");
        RAMA_REQUIRE(bits > 0 && bits <= 8 * key_bytes, "bits out of range");
        RAMA_REQUIRE(path >= 0 && path <= 2, "path must be 0, 1 or 2");
        RAMA_REQUIRE(threads == 512 || threads == 1024, "threads must be 512 or 1024");
        RAMA_REQUIRE(reps >= 1 && reps <= 1000, "reps out of range");
        int nd = 0;
        if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) {
            (void)hipGetLastError();
            throw Error(RAMA_ERR_NO_DEVICE, "no HIP device available");
        }
        *ms_out = key_bytes == 8 ? time_sort<u64>(n, bits, path, threads, reps)
                                 : time_sort<u32>(n, bits, path, threads, reps);
    });
}

int rama_stream_copy_gbps(int64_t bytes, int32_t reps, double* gbps_out) {
    using namespace rama;
    return guarded_call([&] {
        RAMA_REQUIRE(gbps_out != nullptr, "NULL argument");
        RAMA_REQUIRE(bytes >= 16 && reps >= 1 && reps <= 1000, "bytes or reps out of range");
        int nd = 0;
        if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) {
            (void)hipGetLastError();
            throw Error(RAMA_ERR_NO_DEVICE, "no HIP device available");
        }
        *gbps_out = time_copy(bytes, reps);
    });
}

}  // extern "C"

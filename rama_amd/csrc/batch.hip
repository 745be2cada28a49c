// batch.hip — many instances on one GPU: a batched solve and a serving
// loop with continuous admission.
//
// Each instance is solved by the plain solver, unchanged, so its result is
// bit-identical alone and in any mix.  Throughput comes from overlap: a
// pool of worker threads, each with its own stream, caching arena and
// single-pass-scan state (reused across instances), keeps several solves in
// flight; while one worker waits on a read-back, the others' kernels run.
// Small instances are launch- and latency-bound (most of a round is
// dispatch gaps), so overlap is what fills the 256 CUs.
//
// Serving: requests enter a FIFO; a worker admits the head request when the
// device-memory estimate of everything in flight plus the request fits the
// HBM budget, uploads it (graph and dual state stay resident on the device
// for the whole solve) and publishes the labels when done.  Latency is
// measured from submission to labels back on the host.
#include <atomic>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "rama_internal.hpp"

namespace rama {
// the packed (lockstep) batch engine lives in lockstep.hip
void solve_packed(Ctx& ctx, PackedBatch& B, int workers);
// solve_common lives in capi.hip
void solve_common(Ctx& ctx, DBuf<int>& i, DBuf<int>& j, DBuf<float>& c, i64 m, const rama_options_c* copts,
                  DBuf<int>& labels, i64& n_labels, double* lb_out, int32_t* time_ms, rama_timeline** tl);
void set_last_error(const std::string& s);
}  // namespace rama

using namespace rama;

namespace {

// Device bytes a solve of m edges may hold at its peak (measured arena peaks
// stay below ~400 B/edge on grid and random graphs; rounded up).
constexpr double kBytesPerEdge = 512.0;
constexpr double kBytesFixed = 64.0 * 1024 * 1024;

double estimate_bytes(i64 m) { return kBytesFixed + kBytesPerEdge * (double)m; }

using Clock = std::chrono::steady_clock;

}  // namespace

struct rama_server {
    struct Request {
        int64_t ticket;
        std::vector<int32_t> i, j;
        std::vector<float> c;
        Clock::time_point submitted, admitted;
        double est = 0.0;
        int lock_id = -1;
        // result
        bool done = false;
        int status = 0;
        std::string error;
        std::vector<int32_t> labels;
        double lb = 0.0;
        double latency_ms = 0.0;
        double solve_ms = 0.0;
    };

    int device;
    rama_options_c opts;
    double budget;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<int64_t, std::shared_ptr<Request>> all;
    double in_flight = 0.0;
    int64_t next_ticket = 0;
    bool stopping = false;
    int64_t completed = 0;
    std::vector<std::thread> workers;

    // Continuous batching over lanes.  Each lane is a thread with its own
    // stream; lanes run side by side on the device.  An idle lane takes its
    // share of the queue: a lone request runs the plain solver (the lowest
    // latency for one instance), two or more become a lockstep pack that
    // shares every round's launches.  While every lane is busy, a lane with
    // a pack admits the queued requests between its rounds, where they run
    // their initial dual and join the next round.  Admission always keeps
    // the in-flight estimate within the HBM budget.  Options the lockstep
    // path does not carry, or RAMA_SERVE_MODE=workers, give one request per
    // lane.
    bool pack = false;
    int idle = 0;  // lanes waiting for work
    std::unordered_map<int, std::shared_ptr<Request>> by_id;  // lockstep ids of packed requests
    int next_id = 0;
    std::string dead;  // set once a lane has failed

    void finish(const std::shared_ptr<Request>& r) {
        {
            std::lock_guard<std::mutex> l(mu);
            in_flight -= r->est;
            r->done = true;
            ++completed;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }

    struct Sink : LockstepSink {
        rama_server* s;
        explicit Sink(rama_server* x) : s(x) {}
        void done(int id, int status, const std::string& error, const int32_t* labels, i64 n, double lb) override {
            std::shared_ptr<Request> r;
            {
                std::lock_guard<std::mutex> l(s->mu);
                auto it = s->by_id.find(id);
                if (it == s->by_id.end()) return;
                r = it->second;
                s->by_id.erase(it);
            }
            const auto t1 = Clock::now();
            r->status = status;
            r->error = error;
            r->lb = lb;
            if (status == 0 && n > 0) r->labels.assign(labels, labels + n);
            r->latency_ms = std::chrono::duration<double, std::milli>(t1 - r->submitted).count();
            r->solve_ms = std::chrono::duration<double, std::milli>(t1 - r->admitted).count();
            s->finish(r);
        }
    };
    Sink sink{this};

    rama_server(int dev, int lanes, double hbm_budget, const rama_options_c& o)
        : device(dev), opts(o), budget(hbm_budget) {
        const char* mode = std::getenv("RAMA_SERVE_MODE");
        pack = !(mode && std::string(mode) == "workers") && o.sanitize_graph == 0 && o.dump_timeline == 0 &&
               o.max_time_sec < 0 && o.verbose == 0;
        for (int w = 0; w < lanes; ++w) workers.emplace_back([this] { lane(); });
    }
    ~rama_server() {
        {
            std::lock_guard<std::mutex> l(mu);
            stopping = true;
        }
        cv_work.notify_all();
        for (auto& t : workers) t.join();
    }

    // Fails every packed and queued request and refuses later ones (a
    // lane's device state is lost).
    void fail_all(const std::string& msg) {
        std::vector<std::shared_ptr<Request>> rs;
        {
            std::lock_guard<std::mutex> l(mu);
            dead = msg.empty() ? "serving stopped after a device error" : msg;
            for (auto& kv : by_id) rs.push_back(kv.second);
            by_id.clear();
            for (auto& r : queue) rs.push_back(r);
            queue.clear();
            for (auto& r : rs) {
                r->status = RAMA_ERR_INTERNAL;
                r->error = dead;
                r->done = true;
            }
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }

    // Under mu: the head request may start now.
    bool fits() const {
        return !queue.empty() && (in_flight == 0.0 || in_flight + queue.front()->est <= budget);
    }
    std::shared_ptr<Request> pop_head() {
        auto r = queue.front();
        queue.pop_front();
        in_flight += r->est;
        r->admitted = Clock::now();
        return r;
    }

    void lane() {
        std::unique_ptr<Ctx> ctx;
        LockstepHandle* L = nullptr;
        try {
            ctx.reset(new Ctx(device, nullptr, false));
            if (pack) L = lockstep_create(*ctx, to_options(opts), 2, sink);
        } catch (const std::exception& e) {
            fail_all(e.what());
            return;
        }
        try {
            for (;;) {
                const int inpack = L ? lockstep_slots(L) : 0;
                std::vector<std::shared_ptr<Request>> group;
                {
                    std::unique_lock<std::mutex> l(mu);
                    if (inpack == 0) {
                        ++idle;
                        cv_work.wait(l, [&] { return fits() || (stopping && queue.empty()) || !dead.empty(); });
                        --idle;
                        if (!dead.empty() || !fits()) break;
                        // this lane's share of the queue (the other idle lanes take theirs)
                        size_t take = pack ? (queue.size() + (size_t)idle) / (size_t)(idle + 1) : 1;
                        take = std::min<size_t>(take, (size_t)kLockstepMaxSlots);
                        while (take-- > 0 && fits()) group.push_back(pop_head());
                    } else if (idle == 0) {
                        // every lane is busy: queued requests join this pack
                        while (fits() && inpack + (int)group.size() < kLockstepMaxSlots) group.push_back(pop_head());
                    }
                    if (group.size() > 1 || (!group.empty() && inpack > 0))
                        for (auto& r : group) {
                            r->lock_id = next_id++;
                            by_id[r->lock_id] = r;
                        }
                }
                if (group.size() == 1 && inpack == 0) {
                    process(*ctx, *group[0]);
                    finish(group[0]);
                    continue;
                }
                if (!group.empty()) admit(L, group);
                if (L && lockstep_slots(L) > 0) lockstep_round(L);
            }
        } catch (const std::exception& e) {
            fail_all(e.what());
        }
        if (L) lockstep_destroy(L);  // waits for the handed-off requests
    }

    // Packs the group's edges one after another and admits them.
    void admit(LockstepHandle* L, std::vector<std::shared_ptr<Request>>& group) {
        const int k = (int)group.size();
        std::vector<int64_t> off((size_t)k + 1, 0);
        for (int q = 0; q < k; ++q) off[q + 1] = off[q] + (int64_t)group[q]->c.size();
        std::vector<int32_t> gi((size_t)off[k]), gj((size_t)off[k]);
        std::vector<float> gc((size_t)off[k]);
        std::vector<int> ids((size_t)k);
        for (int q = 0; q < k; ++q) {
            Request& r = *group[q];
            std::copy(r.i.begin(), r.i.end(), gi.begin() + off[q]);
            std::copy(r.j.begin(), r.j.end(), gj.begin() + off[q]);
            std::copy(r.c.begin(), r.c.end(), gc.begin() + off[q]);
            ids[q] = r.lock_id;
            std::vector<int32_t>().swap(r.i);
            std::vector<int32_t>().swap(r.j);
            std::vector<float>().swap(r.c);
        }
        lockstep_admit(L, k, off.data(), gi.data(), gj.data(), gc.data(), ids.data());
    }

    void process(Ctx& ctx, Request& r) {
        const auto t0 = Clock::now();
        try {
            const i64 m = (i64)r.c.size();
            DBuf<int> di = ctx.buf<int>(m), dj = ctx.buf<int>(m);
            DBuf<float> dc = ctx.buf<float>(m);
            ctx.h2d(di.get(), r.i.data(), m);
            ctx.h2d(dj.get(), r.j.data(), m);
            ctx.h2d(dc.get(), r.c.data(), m);
            DBuf<int> labels;
            i64 n = 0;
            solve_common(ctx, di, dj, dc, m, &opts, labels, n, &r.lb, nullptr, nullptr);
            r.labels.resize((size_t)n);
            if (n > 0) ctx.d2h(r.labels.data(), labels.get(), n);
            r.status = 0;
        } catch (const Error& e) {
            r.status = e.status;
            r.error = e.what();
        } catch (const std::exception& e) {
            r.status = RAMA_ERR_INTERNAL;
            r.error = e.what();
        }
        const auto t1 = Clock::now();
        r.solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        r.latency_ms = std::chrono::duration<double, std::milli>(t1 - r.submitted).count();
        // inputs are no longer needed
        std::vector<int32_t>().swap(r.i);
        std::vector<int32_t>().swap(r.j);
        std::vector<float>().swap(r.c);
    }
};

extern "C" {

rama_server* rama_server_create(int device, int workers, int64_t hbm_budget_bytes, const rama_options_c* opts) {
    try {
        RAMA_REQUIRE(workers > 0 && workers <= 64, "workers must be in [1, 64]");
        RAMA_REQUIRE(opts != nullptr, "options must not be NULL");
        return new rama_server(device, workers, (double)hbm_budget_bytes, *opts);
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}

int64_t rama_server_submit(rama_server* s, const int32_t* i, const int32_t* j, const float* c, int64_t m) {
    if (!s || !i || !j || !c || m <= 0) {
        set_last_error("invalid request");
        return -1;
    }
    auto r = std::make_shared<rama_server::Request>();
    r->i.assign(i, i + m);
    r->j.assign(j, j + m);
    r->c.assign(c, c + m);
    r->est = estimate_bytes(m);
    r->submitted = Clock::now();
    {
        std::lock_guard<std::mutex> l(s->mu);
        if (!s->dead.empty()) {
            set_last_error(s->dead);
            return -1;
        }
        r->ticket = s->next_ticket++;
        s->queue.push_back(r);
        s->all[r->ticket] = r;
    }
    s->cv_work.notify_one();
    return r->ticket;
}

int rama_server_poll(rama_server* s, int64_t ticket) {
    if (!s) return -1;
    std::lock_guard<std::mutex> l(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return -1;
    return it->second->done ? 1 : 0;
}

int rama_server_wait(rama_server* s, int64_t ticket, int32_t* labels_out, int64_t cap, int64_t* n_labels,
                     double* lb, double* latency_ms, double* solve_ms) {
    if (!s) return RAMA_ERR_INVALID;
    std::shared_ptr<rama_server::Request> r;
    {
        std::unique_lock<std::mutex> l(s->mu);
        auto it = s->all.find(ticket);
        if (it == s->all.end()) {
            set_last_error("unknown ticket");
            return RAMA_ERR_INVALID;
        }
        r = it->second;
        // the wait releases the lock: submits may rehash the map and another
        // waiter on the same ticket may have removed it, so the iterator is
        // not reused; only the first waiter to wake takes the result
        s->cv_done.wait(l, [&] { return r->done; });
        if (s->all.erase(ticket) == 0) {
            set_last_error("ticket already collected by another wait");
            return RAMA_ERR_INVALID;
        }
    }
    if (r->status != 0) {
        set_last_error(r->error);
        return r->status;
    }
    if (n_labels) *n_labels = (int64_t)r->labels.size();
    if ((int64_t)r->labels.size() > cap) {
        set_last_error("label buffer too small");
        return RAMA_ERR_INVALID;
    }
    if (labels_out) std::copy(r->labels.begin(), r->labels.end(), labels_out);
    if (lb) *lb = r->lb;
    if (latency_ms) *latency_ms = r->latency_ms;
    if (solve_ms) *solve_ms = r->solve_ms;
    return RAMA_OK;
}

void rama_server_destroy(rama_server* s) { delete s; }

// Batched solve of k instances packed back to back: instance q owns edges
// [edge_off[q], edge_off[q+1]) of (i, j, c) with its own 0-based node ids,
// and receives its labels at labels_out + label_off[q] (capacity
// label_off[q+1] - label_off[q]).
int rama_solve_batch_host(int32_t k, const int64_t* edge_off, const int32_t* i, const int32_t* j, const float* c,
                          const rama_options_c* opts, int32_t workers, int32_t* labels_out, const int64_t* label_off,
                          int64_t* n_labels, double* lbs, int32_t* status) {
    try {
        RAMA_REQUIRE(k >= 0 && edge_off && i && j && c && opts && labels_out && label_off && n_labels && lbs,
                     "NULL argument");
        RAMA_REQUIRE(workers >= 1 && workers <= 64, "workers must be in [1, 64]");
        int dev_count = 0;
        if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) {
            (void)hipGetLastError();
            throw Error(RAMA_ERR_NO_DEVICE, "no HIP device available");
        }
        std::mutex emu;
        std::string first_error;
        // lockstep path: every instance in one packed graph (per-instance
        // results equal solving alone); options it does not carry (time
        // budget, timelines, sanitising, printing) use the per-instance path
        const char* mode = std::getenv("RAMA_BATCH_MODE");
        const bool packed = k >= 2 && !(mode && std::string(mode) == "workers") && opts->sanitize_graph == 0 &&
                            opts->dump_timeline == 0 && opts->max_time_sec < 0 && opts->verbose == 0;
        if (packed) {
            for (int32_t q = 0; q < k; ++q) {
                n_labels[q] = 0;
                lbs[q] = 0.0;
                if (status) status[q] = 0;
            }
            std::vector<int32_t> st((size_t)k, 0);
            PackedBatch B{k,        edge_off, i,     j,   c,    to_options(*opts), labels_out, label_off, n_labels,
                          lbs,      st.data(), &first_error, &emu};
            Ctx ctx(0, nullptr, false);
            solve_packed(ctx, B, workers);
            bool failed = false;
            for (int32_t q = 0; q < k; ++q) {
                if (status) status[q] = st[q];
                failed = failed || st[q] != 0;
            }
            if (failed) throw Error(RAMA_ERR_INTERNAL, "batch: " + first_error);
            return RAMA_OK;
        }
        std::atomic<int32_t> next{0};
        std::atomic<int> failed{0};
        std::vector<std::thread> pool;
        auto work = [&] {
            std::unique_ptr<Ctx> ctx;
            try {
                ctx.reset(new Ctx(0, nullptr, false));
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> l(emu);
                if (first_error.empty()) first_error = e.what();
                failed = 1;
                return;
            }
            for (;;) {
                const int32_t q = next.fetch_add(1);
                if (q >= k) return;
                const i64 e0 = edge_off[q], m = edge_off[q + 1] - edge_off[q];
                int st = 0;
                try {
                    DBuf<int> di = ctx->buf<int>(m), dj = ctx->buf<int>(m);
                    DBuf<float> dc = ctx->buf<float>(m);
                    ctx->h2d(di.get(), i + e0, m);
                    ctx->h2d(dj.get(), j + e0, m);
                    ctx->h2d(dc.get(), c + e0, m);
                    DBuf<int> lab;
                    i64 n = 0;
                    solve_common(*ctx, di, dj, dc, m, opts, lab, n, &lbs[q], nullptr, nullptr);
                    n_labels[q] = n;
                    const i64 cap = label_off[q + 1] - label_off[q];
                    if (n > cap) throw Error(RAMA_ERR_INVALID, "label buffer too small for instance " + std::to_string(q));
                    if (n > 0) ctx->d2h(labels_out + label_off[q], lab.get(), n);
                } catch (const Error& e) {
                    st = e.status;
                    std::lock_guard<std::mutex> l(emu);
                    if (first_error.empty()) first_error = e.what();
                } catch (const std::exception& e) {
                    st = RAMA_ERR_INTERNAL;
                    std::lock_guard<std::mutex> l(emu);
                    if (first_error.empty()) first_error = e.what();
                }
                if (status) status[q] = st;
                if (st) failed = 1;
            }
        };
        const int nw = std::min<int>(workers, std::max<int>(k, 1));
        for (int w = 0; w < nw; ++w) pool.emplace_back(work);
        for (auto& t : pool) t.join();
        if (failed) throw Error(RAMA_ERR_INTERNAL, "batch: " + first_error);
        return RAMA_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return RAMA_ERR_INTERNAL;
    }
}

}  // extern "C"

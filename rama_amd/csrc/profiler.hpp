// profiler.hpp — optional per-kernel timing with HIP events recorded on the
// launching stream.  Each instrumented launch carries the algorithmic bytes
// it must move (a roofline numerator), so achieved bandwidth = bytes / time.
// Disabled by default; enabling costs two event records per launch.
#pragma once

#include "common.hpp"

namespace rama {
namespace prof {

bool enabled();
bool wants(const char* name);  // enabled and not filtered out
// `units`: work items of a kernel without a byte roofline (cycle-search
// expansions); `entries`: neighbour-list entries its merge walks stream.
void push(const char* name, hipEvent_t a, hipEvent_t b, double bytes, double units = 0.0, double entries = 0.0);
hipEvent_t take_event();

// Call-site tag for shared primitives (sorts): with RAMA_PROFILE_SITES set,
// entries recorded inside a Site are named "kernel@site".
const char* current_site();
class Site {
   public:
    explicit Site(const char* s);
    ~Site();
    Site(const Site&) = delete;
    Site& operator=(const Site&) = delete;

   private:
    const char* prev_;
};

// Brackets the launches in its scope with events on ctx.stream.
class Scope {
   public:
    Scope(Ctx& ctx, const char* name, double bytes, double units = 0.0, double entries = 0.0)
        : ctx_(ctx), name_(name), bytes_(bytes), units_(units), entries_(entries) {
        if (wants(name)) {
            a_ = take_event();
            b_ = take_event();
            RAMA_HIP(hipEventRecord(a_, ctx_.stream));
        }
    }
    ~Scope() {
        if (a_) {
            (void)hipEventRecord(b_, ctx_.stream);
            push(name_, a_, b_, bytes_, units_, entries_);
        }
    }
    Scope(const Scope&) = delete;
    Scope& operator=(const Scope&) = delete;

   private:
    Ctx& ctx_;
    const char* name_;
    double bytes_, units_, entries_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

}  // namespace prof
}  // namespace rama

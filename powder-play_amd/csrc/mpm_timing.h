// Per-kernel timing with HIP events recorded on the launching stream.
// Enabled per context (mpm_set_timing); resolved after the stream drains.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace mpm {

struct KernelTimes {
    struct Pending {
        const char* name;
        hipEvent_t begin, end;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, unsigned long long>> totals;  // ms, launches

    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    // Call after the stream has completed every pending event.
    void resolve() {
        for (const Pending& p : pending) {
            float ms = 0.0f;
            if (p.begin && p.end && hipEventElapsedTime(&ms, p.begin, p.end) == hipSuccess) {
                auto& t = totals[p.name];
                t.first += ms;
                t.second += 1;
            }
            if (p.begin) pool.push_back(p.begin);
            if (p.end) pool.push_back(p.end);
        }
        pending.clear();
    }
    void reset() {
        resolve();
        totals.clear();
    }
    ~KernelTimes() {
        for (const Pending& p : pending) {
            if (p.begin) (void)hipEventDestroy(p.begin);
            if (p.end) (void)hipEventDestroy(p.end);
        }
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

struct TimedRegion {
    KernelTimes* t;
    const char* name;
    hipStream_t stream;
    hipEvent_t begin = nullptr;
    TimedRegion(KernelTimes* times, const char* n, hipStream_t s) : t(times), name(n), stream(s) {
        if (t) {
            begin = t->take();
            if (begin) (void)hipEventRecord(begin, stream);
        }
    }
    ~TimedRegion() {
        if (t) {
            hipEvent_t end = t->take();
            if (end) (void)hipEventRecord(end, stream);
            t->pending.push_back({name, begin, end});
        }
    }
};

}  // namespace mpm

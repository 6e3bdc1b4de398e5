// Per-kernel-family timing with HIP events recorded on the launching stream. A context binds its
// profiler for the duration of an API call; launch sites open a ProfileScope naming the kernel
// family and its algorithmic bytes. Disabled profilers cost one branch per scope.
#pragma once

#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ldh_internal.h"

namespace ldh {

struct Profiler {
    bool enabled = false;
    hipStream_t stream = nullptr;
    struct Pending {
        std::string name;
        hipEvent_t start, stop;
        double bytes;
    };
    struct Stat {
        int64_t launches = 0;
        double ms = 0.0;
        double bytes = 0.0;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, Stat> stats;

    hipEvent_t event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        LDH_HIP(hipEventCreate(&e), "profiler-event");
        return e;
    }
    void resolve() {
        for (auto& p : pending) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
                Stat& s = stats[p.name];
                s.launches += 1;
                s.ms += ms;
                s.bytes += p.bytes;
            } else {
                (void)hipGetLastError();
            }
            pool.push_back(p.start);
            pool.push_back(p.stop);
        }
        pending.clear();
    }
    int read(ldh_kernel_stat* out, int capacity) {
        resolve();
        int n = 0;
        for (const auto& kv : stats) {
            if (n >= capacity) break;
            ldh_kernel_stat& o = out[n++];
            std::memset(&o, 0, sizeof(o));
            std::strncpy(o.name, kv.first.c_str(), sizeof(o.name) - 1);
            o.launches = kv.second.launches;
            o.total_ms = kv.second.ms;
            o.algorithmic_bytes = kv.second.bytes;
        }
        return n;
    }
    void reset() {
        resolve();
        stats.clear();
    }
    ~Profiler() {
        resolve();
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

Profiler*& active_profiler();

struct ProfilerBinding {
    Profiler* prev;
    explicit ProfilerBinding(Profiler& p) : prev(active_profiler()) { active_profiler() = &p; }
    ~ProfilerBinding() { active_profiler() = prev; }
};

// Times the launches issued on the bound profiler's stream while in scope.
struct ProfileScope {
    Profiler* p = nullptr;
    hipEvent_t start = nullptr;
    std::string name;
    double bytes = 0.0;
    // An empty name records nothing (launch sequences that must not create events).
    ProfileScope(const std::string& n, double b) {
        if (n.empty()) return;
        Profiler* ap = active_profiler();
        if (!ap || !ap->enabled || !ap->stream) return;
        p = ap;
        name = n;
        bytes = b;
        start = p->event();
        (void)hipEventRecord(start, p->stream);
    }
    ~ProfileScope() {
        if (!p) return;
        hipEvent_t stop = p->event();
        (void)hipEventRecord(stop, p->stream);
        p->pending.push_back({name, start, stop, bytes});
    }
};

}  // namespace ldh

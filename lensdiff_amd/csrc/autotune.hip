// Variant autotuner (see autotune.h).
#include "autotune.h"

#include <cstdio>

namespace ldh {

Autotuner& Autotuner::instance() {
    static Autotuner t;
    return t;
}

int Autotuner::run(const std::string& key, int variants, hipStream_t s, const std::function<void(int)>& launch) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    const std::string k = key + "@dev" + std::to_string(dev);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(s, &cap);
    {
        std::lock_guard<std::mutex> g(mu_);
        if (forced_ >= 0) {
            const int v = forced_ < variants ? forced_ : variants - 1;
            launch(v);
            return v;
        }
        auto it = table_.find(k);
        // inside a graph capture nothing may synchronise: an untuned key takes variant 0
        if (it != table_.end() || variants <= 1 || cap != hipStreamCaptureStatusNone) {
            const int v = it != table_.end() ? it->second.choice : 0;
            launch(v);
            return v;
        }
    }
    Entry e;
    e.ms.assign(size_t(variants), 0.0f);
    hipEvent_t a, b;
    LDH_HIP(hipEventCreate(&a), "autotune-event");
    LDH_HIP(hipEventCreate(&b), "autotune-event");
    for (int v = 0; v < variants; ++v) {
        launch(v);  // warm: first launch of the variant
        LDH_HIP(hipEventRecord(a, s), "autotune-event");
        launch(v);
        LDH_HIP(hipEventRecord(b, s), "autotune-event");
        LDH_HIP(hipEventSynchronize(b), "autotune-event");
        float ms = 0.0f;
        LDH_HIP(hipEventElapsedTime(&ms, a, b), "autotune-event");
        e.ms[size_t(v)] = ms;
        if (ms < e.ms[size_t(e.choice)]) e.choice = v;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    std::lock_guard<std::mutex> g(mu_);
    table_.emplace(k, e);  // a concurrent tuner of the same key may have won: keep the first
    return e.choice;
}

std::string Autotuner::report() const {
    std::lock_guard<std::mutex> g(mu_);
    std::string out;
    char buf[64];
    for (const auto& kv : table_) {
        out += kv.first + ": v" + std::to_string(kv.second.choice) + " ";
        for (float ms : kv.second.ms) {
            std::snprintf(buf, sizeof(buf), " %.3f", ms);
            out += buf;
        }
        out += " ms\n";
    }
    return out;
}

void Autotuner::reset() {
    std::lock_guard<std::mutex> g(mu_);
    table_.clear();
}

void Autotuner::force(int variant) {
    std::lock_guard<std::mutex> g(mu_);
    table_.clear();
    forced_ = variant < 0 ? -1 : variant;
}

}  // namespace ldh

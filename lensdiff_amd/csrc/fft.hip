// FFT tables and the plain (unfused) transforms.
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>

#include "fft_io.h"
#include "profiler.h"

namespace ldh {

void FftTables::init(hipStream_t s) {
    std::vector<float2> h(kTwiddleEntries);
    for (const int t : {kTw2, kTw3, kTw9, kTw5}) {
        const int off = tw_offset(t);
        for (int j = 0; j < t; ++j) {
            // exp(-2 pi i j / t), computed in double and rounded once
            const double a = -2.0 * M_PI * double(j) / double(t);
            h[size_t(off + j)] = make_float2(float(std::cos(a)), float(std::sin(a)));
        }
    }
    twiddle.reset(sizeof(float2) * kTwiddleEntries);
    LDH_HIP(hipMemcpyAsync(twiddle.get(), h.data(), sizeof(float2) * kTwiddleEntries, hipMemcpyHostToDevice, s),
            "fft-table-upload");
    LDH_HIP(hipStreamSynchronize(s), "fft-table-upload");
}

static void check_dims(int width, int height) {
    if (!grid_length_ok(width) || width < 16 || !grid_length_ok(height)) fail(LDH_ERR_TOO_LARGE, "fft-size-unsupported");
}

void fft_rows(hipStream_t s, const FftTables& t, float2* data, int width, int rows, bool inverse) {
    rows_pass(s, t.twiddle.as<float2>(), width, rows, PlainRows{data, width}, inverse);
}

void fft_cols_forward(hipStream_t s, const FftTables& t, float2* data, int width, int height) {
    cols_forward(s, t.twiddle.as<float2>(), data, width, height, PlainLoad{});
}

void fft_cols_inverse(hipStream_t s, const FftTables& t, float2* data, int width, int height) {
    cols_inverse(s, t.twiddle.as<float2>(), data, data, width, height, PlainLoad{});
}

void fft2d_forward(hipStream_t s, const FftTables& t, float2* data, int width, int height) {
    check_dims(width, height);
    fft_rows(s, t, data, width, height, false);
    fft_cols_forward(s, t, data, width, height);
}

void fft2d_inverse(hipStream_t s, const FftTables& t, float2* data, int width, int height) {
    check_dims(width, height);
    fft_cols_inverse(s, t, data, width, height);
    fft_rows(s, t, data, width, height, true);
}

namespace {

struct SplitPlan {
    ColSplit chosen;
    std::vector<ColSplit> cand;
    std::vector<float> ms;  // per candidate: one forward + inverse column transform (empty: not timed)
};
std::mutex g_split_mu;    // the table
std::mutex g_tune_mu;     // one height tuned at a time; others wait for its entry
std::map<std::pair<int, int>, SplitPlan> g_splits;  // (device, height)

// Times every candidate on scratch memory (forward then inverse column transform, twice after one
// warm-up run); moves e.chosen (the balanced split on entry) to a clearly faster one, and leaves
// it where it was if anything fails.
void tune_splits(hipStream_t s, const float2* tw, int width, int height, SplitPlan& e) {
    Profiler* const bound = active_profiler();
    active_profiler() = nullptr;  // tuning launches are not the caller's work
    hipEvent_t a = nullptr, b = nullptr;
    try {
        DeviceBuffer scratch(sizeof(float2) * size_t(width) * size_t(height));
        float2* d = scratch.as<float2>();
        LDH_HIP(hipMemsetAsync(d, 0, sizeof(float2) * size_t(width) * size_t(height), s), "fft-split-tune");
        LDH_HIP(hipEventCreate(&a), "fft-split-tune");
        LDH_HIP(hipEventCreate(&b), "fft-split-tune");
        std::vector<float> ms;
        for (const ColSplit& c : e.cand) {
            auto run = [&] {
                cols_forward_split(s, tw, d, width, height, PlainLoad{}, ColRows{}, c);
                cols_inverse_split(s, tw, d, d, width, height, PlainLoad{}, 1 << 30, c);
            };
            run();
            LDH_HIP(hipEventRecord(a, s), "fft-split-tune");
            run();
            run();
            LDH_HIP(hipEventRecord(b, s), "fft-split-tune");
            LDH_HIP(hipEventSynchronize(b), "fft-split-tune");
            float t = 0.0f;
            LDH_HIP(hipEventElapsedTime(&t, a, b), "fft-split-tune");
            ms.push_back(t * 0.5f);
        }
        // the balanced split unless another is faster by more than 5% (timing noise would
        // otherwise pick among equals from run to run)
        size_t best = 0, balanced = 0;
        for (size_t i = 0; i < ms.size(); ++i) {
            if (ms[i] < ms[best]) best = i;
            if (e.cand[i].n1 == e.chosen.n1) balanced = i;
        }
        e.ms = ms;
        if (ms[best] < 0.95f * ms[balanced]) e.chosen = e.cand[best];
    } catch (const Error&) {
        // out of memory or a failed call: the balanced split, untimed
        e.ms.clear();
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    active_profiler() = bound;
}

}  // namespace

std::vector<ColSplit> col_splits(int height) {
    std::vector<ColSplit> out;
    if (col_tile_length_ok(height)) return out;
    for (const int a : {8, 12, 16, 24, 32, 36, 48, 64, 72, 96, 128, 144})
        if (height % a == 0 && col_tile_length_ok(height / a)) out.push_back(ColSplit{a, height / a});
    return out;
}

ColSplit col_plan(hipStream_t s, const float2* tw, int width, int height) {
    if (col_tile_length_ok(height)) return ColSplit{0, height};
    int dev = 0;
    LDH_HIP(hipGetDevice(&dev), "fft-split-plan");
    const auto key = std::make_pair(dev, height);
    {
        std::lock_guard<std::mutex> g(g_split_mu);
        const auto it = g_splits.find(key);
        if (it != g_splits.end()) return it->second.chosen;
    }
    std::lock_guard<std::mutex> tune_lock(g_tune_mu);
    {
        std::lock_guard<std::mutex> g(g_split_mu);
        const auto it = g_splits.find(key);
        if (it != g_splits.end()) return it->second.chosen;  // tuned while this thread waited
    }
    SplitPlan e;
    e.cand = col_splits(height);
    if (e.cand.empty()) fail(LDH_ERR_TOO_LARGE, "fft-height-unsupported");
    col_split(height, &e.chosen.n1, &e.chosen.n2);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    LDH_HIP(hipStreamIsCapturing(s, &cap), "fft-split-plan");
    if (cap == hipStreamCaptureStatusNone && e.cand.size() > 1) tune_splits(s, tw, width, height, e);
    std::lock_guard<std::mutex> g(g_split_mu);
    return g_splits.emplace(key, e).first->second.chosen;
}

std::string col_plan_report() {
    std::lock_guard<std::mutex> g(g_split_mu);
    std::string out;
    char buf[64];
    for (const auto& kv : g_splits) {
        const SplitPlan& e = kv.second;
        out += "fft-cols/" + std::to_string(kv.first.second) + "[";
        size_t chosen = 0;
        for (size_t i = 0; i < e.cand.size(); ++i) {
            if (i) out += ",";
            out += std::to_string(e.cand[i].n1) + "x" + std::to_string(e.cand[i].n2);
            if (e.cand[i].n1 == e.chosen.n1) chosen = i;
        }
        out += "]@dev" + std::to_string(kv.first.first) + ": v" + std::to_string(chosen) + " ";
        for (float t : e.ms) {
            std::snprintf(buf, sizeof(buf), " %.3f", t);
            out += buf;
        }
        out += " ms\n";
    }
    return out;
}

int fft_col_storage_row(hipStream_t s, const FftTables& t, int width, int height, int k) {
    return col_storage_row(col_plan(s, t.twiddle.as<float2>(), width, height), k);
}

Profiler*& active_profiler() {
    thread_local Profiler* p = nullptr;
    return p;
}

}  // namespace ldh

// FFT tables and the plain (unfused) transforms.
#include <cmath>

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

int fft_col_storage_row(int height, int k) { return col_storage_row(height, k); }

Profiler*& active_profiler() {
    thread_local Profiler* p = nullptr;
    return p;
}

}  // namespace ldh

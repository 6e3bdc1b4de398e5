// Sequential-FP32-sum reproduction and unit-sum normalisation on the device.
#pragma once

#include "ldh_internal.h"

namespace ldh {

constexpr int kSeqSumBlock = 4096;  // elements per block of the walk

// Where the segments of one sum launch live: segment g reads array g / per_array at offset
// (g % per_array) * stride, each element multiplied on load by 1.0f / prescale[g % per_array]
// of its array when that array has prescale totals (the values an in-place normalisation by
// those totals would have left), totals <= 0 meaning no scale.
struct SeqSegments {
    static constexpr int kMaxArrays = 4;
    const float* array[kMaxArrays] = {nullptr, nullptr, nullptr, nullptr};
    const float* prescale[kMaxArrays] = {nullptr, nullptr, nullptr, nullptr};
    int per_array = 1;
    int64_t stride = 0;
    __device__ const float* base(int g) const { return array[g / per_array] + int64_t(g % per_array) * stride; }
    __device__ float scale(int g) const {
        const float* t = prescale[g / per_array];
        if (t == nullptr) return 1.0f;
        const float v = t[g % per_array];
        return v > 0.0f ? 1.0f / v : 1.0f;
    }
};

class SeqSum {
public:
    // totals[seg] = the FP32 sequential sum of x[seg*stride .. seg*stride + n) (device buffer).
    // x must be non-negative.
    void run(hipStream_t s, const float* x, int64_t n, int64_t stride, int segments, float* totals);
    // The same over the segments described by g (segments = arrays used * per_array).
    void run(hipStream_t s, const SeqSegments& g, int64_t n, int segments, float* totals);
    // x *= 1.0f / totals[seg] where totals[seg] > 0 (LensDiffCpuReference.cpp:440-449).
    void scale(hipStream_t s, float* x, int64_t n, int64_t stride, int segments, const float* totals);
    // run(), then scale().
    void normalize(hipStream_t s, float* x, int64_t n, int64_t stride, int segments, float* totals);
    // Grows the block tables for sums of `segments` arrays of n ahead of a graph capture.
    void reserve(int64_t n, int segments);

private:
    DeviceBuffer bsum_, bins_, ebase_;
};

}  // namespace ldh

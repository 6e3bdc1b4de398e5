// Sequential-FP32-sum reproduction and unit-sum normalisation on the device.
#pragma once

#include "ldh_internal.h"

namespace ldh {

constexpr int kSeqSumBlock = 4096;  // elements per block of the walk

class SeqSum {
public:
    // totals[seg] = the FP32 sequential sum of x[seg*stride .. seg*stride + n) (device buffer).
    // x must be non-negative.
    void run(hipStream_t s, const float* x, int64_t n, int64_t stride, int segments, float* totals);
    // x *= 1.0f / totals[seg] where totals[seg] > 0 (LensDiffCpuReference.cpp:440-449).
    void scale(hipStream_t s, float* x, int64_t n, int64_t stride, int segments, const float* totals);
    // run(), then scale().
    void normalize(hipStream_t s, float* x, int64_t n, int64_t stride, int segments, float* totals);

private:
    DeviceBuffer bsum_, bins_, ebase_;
};

}  // namespace ldh

// Sequential-FP32-sum reproduction and unit-sum normalisation on the device.
#pragma once

#include "ldh_internal.h"

namespace ldh {

constexpr int kSeqSumBlock = 4096;  // elements per block of the walk

class SeqSum {
public:
    // totals[seg] = the FP32 sequential sum of x[seg*stride .. seg*stride + n) (device buffer).
    void run(hipStream_t s, const float* x, int64_t n, int64_t stride, int segments, float* totals);
    // run(), then x *= 1.0f / total per segment where total > 0 (LensDiffCpuReference.cpp:440-449).
    void normalize(hipStream_t s, float* x, int64_t n, int64_t stride, int segments, float* totals);

private:
    DeviceBuffer bsum_, bins_, elo_;
};

}  // namespace ldh

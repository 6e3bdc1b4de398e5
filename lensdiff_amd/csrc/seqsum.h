// Sequential-FP32-sum reproduction and unit-sum normalisation on the device.
#pragma once

#include "ldh_internal.h"

namespace ldh {

constexpr int kSeqSumBlock = 4096;     // elements per block of the walk
constexpr int kSeqSumMinBlock = 512;   // shortest block a caller may choose

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

// Kernels that produce a summed array can hand the sum its block sums: a workgroup of
// kSumThreads threads covers block blockIdx.x (kSeqSumBlock consecutive elements) of one
// segment, each thread adds the values it writes into FP64 shares, and every thread then calls
// store_block_sums, which writes the workgroup's totals to dst[0..N).
constexpr int kSumThreads = 256;
inline int seqsum_blocks(int64_t n) { return int((n + kSeqSumBlock - 1) / kSeqSumBlock); }

template <int N>
__device__ __forceinline__ void store_block_sums(double (&share)[N], double* const (&dst)[N]) {
    constexpr int kWaves = kSumThreads / 64;
    __shared__ double part[N][kWaves];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double v = share[j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((threadIdx.x & 63) == 0) part[j][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += part[threadIdx.x][w];
        *dst[threadIdx.x] = t;
    }
}

class SeqSum {
public:
    // totals[seg] = the FP32 sequential sum of x[seg*stride .. seg*stride + n) (device buffer).
    // x must be non-negative.
    void run(hipStream_t s, const float* x, int64_t n, int64_t stride, int segments, float* totals);
    // The same over the segments described by g (segments = arrays used * per_array). The sum
    // runs over blocks of `block` elements (kSeqSumMinBlock .. kSeqSumBlock); block_sums
    // (nullable, [segment][block]) are their FP64 sums when the producer of the data already
    // formed them (any summation order: they only place the blocks' binades and mark all-zero
    // blocks).
    void run(hipStream_t s, const SeqSegments& g, int64_t n, int segments, float* totals,
             const double* block_sums = nullptr, int block = kSeqSumBlock);
    // Segments of one array (as the first run) with their block sums from the producer.
    void run(hipStream_t s, const float* x, int64_t n, int64_t stride, int segments, float* totals,
             const double* block_sums);
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

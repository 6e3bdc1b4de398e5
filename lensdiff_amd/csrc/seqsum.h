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

// v moved across lanes by a DPP control (both 32-bit halves); lanes that the control or the row
// mask leaves without a source read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Inclusive scan over the 64 lanes with DPP: shifts by 1, 2, 4, 8 within each row of 16, then
// the last lane of row 0 / 2 into rows 1 / 3 (row_bcast:15) and of row 1 into rows 2 and 3
// (row_bcast:31). All 64 lanes must be active. The walk's additions are exact (multiples of one
// power of two), so their order does not matter; the other callers form FP64 block sums, which
// only place blocks and mark empty ones.
__device__ __forceinline__ double wave_inclusive_scan(double v, int) {
    v += dpp_move<0x111, 0xf>(v);  // row_shr:1
    v += dpp_move<0x112, 0xf>(v);  // row_shr:2
    v += dpp_move<0x114, 0xf>(v);  // row_shr:4
    v += dpp_move<0x118, 0xf>(v);  // row_shr:8
    v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// The sum over the 64 lanes, in every lane (the scan's last lane, read back as a scalar).
__device__ __forceinline__ double wave_total(double v) {
    v = wave_inclusive_scan(v, 0);
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

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
        const double v = wave_total(share[j]);
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

// Allocation-free sums (ldh_sequential_sums_noalloc): the block tables live in a static device
// array of the library, one per device. noalloc_scratch_init looks up its address on `device` (once
// per device, at context creation); sequential_sums_noalloc then sums `segments` device arrays of
// n floats at x + g * n into totals[g] on stream s without allocating, for at most 65536
// (segment, 4096-element block) pairs. Calls that overlap on one device share the tables and race.
void noalloc_scratch_init(int device);
void sequential_sums_noalloc(hipStream_t s, int device, const float* x, int64_t n, int segments, float* totals);

}  // namespace ldh

// Reproduces, in parallel, the value a sequential FP32 accumulation returns:
//     float s = 0; for (i = 0; i < n; ++i) s += x[i];
// which is how the reference normalises every PSF and kernel (std::accumulate with a float
// accumulator, LensDiffCpuReference.cpp:436-449). For the non-negative arrays summed here the
// running sum rises monotonically through binades [2^e, 2^(e+1)); while it stays inside one
// binade, adding x rounds x to a multiple of u = 2^(e-23), so
//     s_after_block = s_before + sum_i rint(x_i / u) * u
// whatever the order inside the block. The sum is therefore found by
//   1. per-block FP64 sums, to fix the 32-binade window the running sum passes through;
//   2. per-block, per-binade sums of the rounded contributions (exact in FP64);
//   3. a one-wavefront walk over the blocks in order: blocks whose whole contribution keeps the
//      running sum inside its binade are added at once; the few blocks where it crosses into
//      the next binade are replayed element by element, taking the crossing addition as a
//      true FP32 add.
// The result equals the sequential FP32 sum except where an addition is an exact rounding tie
// (ties are resolved on the addend alone), an error of one unit in the last place of the
// running sum. Contributions added while the running sum is below the window are taken as
// exact (their rounding is below 2^-54 of the total).
#include <climits>

#include "seqsum.h"

namespace ldh {

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = kSeqSumBlock / kThreads;
constexpr int kBinades = 32;
constexpr int kWindowAbove = 1;  // binades kept above the binade of the FP64 total

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double wave_inclusive_scan(double v, int lane) {
    for (int o = 1; o < 64; o <<= 1) {
        const double t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

__global__ void __launch_bounds__(kThreads)
block_sums_kernel(const float* __restrict__ x, int64_t n, int64_t stride, int nb, double* __restrict__ bsum) {
    const int seg = blockIdx.y, b = blockIdx.x;
    const float* base = x + seg * stride;
    const int64_t start = int64_t(b) * kSeqSumBlock;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
        const int64_t i = start + int64_t(k) * kThreads + threadIdx.x;
        if (i < n) acc += double(base[i]);
    }
    __shared__ double red[kThreads / 64];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) t += red[w];
        bsum[int64_t(seg) * nb + b] = t;
    }
}

__global__ void __launch_bounds__(kThreads)
window_kernel(const double* __restrict__ bsum, int nb, int* __restrict__ elo) {
    const int seg = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += kThreads) acc += bsum[int64_t(seg) * nb + b];
    __shared__ double red[kThreads / 64];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) t += red[w];
        elo[seg] = t > 0.0 ? ilogb(t) + kWindowAbove - (kBinades - 1) : INT_MIN;
    }
}

__global__ void __launch_bounds__(kThreads)
binade_sums_kernel(const float* __restrict__ x, int64_t n, int64_t stride, int nb, const int* __restrict__ elo,
                   double* __restrict__ bins) {
    const int seg = blockIdx.y, b = blockIdx.x;
    const int e_lo = elo[seg];
    if (e_lo == INT_MIN) return;
    const float* base = x + seg * stride;
    const int64_t start = int64_t(b) * kSeqSumBlock;
    float v[kPerThread];
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
        const int64_t i = start + int64_t(k) * kThreads + threadIdx.x;
        v[k] = i < n ? base[i] : 0.0f;
    }
    __shared__ double red[kBinades][kThreads / 64];
#pragma unroll 4
    for (int q = 0; q < kBinades; ++q) {
        // contributions counted in units of u = 2^(e-23): rint(x * 2^(23-e))
        const int shift = 23 - (e_lo + q);
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) acc += double(rintf(ldexpf(v[k], shift)));
        acc = wave_sum(acc);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (threadIdx.x < kBinades) {
        const int q = threadIdx.x;
        double t = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) t += red[q][w];
        bins[(int64_t(seg) * kBinades + q) * nb + b] = ldexp(t, (e_lo + q) - 23);
    }
}

// Contribution of x at binade slot q (q < 0: below the window, taken exactly).
__device__ __forceinline__ double contribution(float x, int e_lo, int q) {
    if (q < 0) return double(x);
    return ldexp(double(rintf(ldexpf(x, 23 - (e_lo + q)))), (e_lo + q) - 23);
}
__device__ __forceinline__ double slot_limit(int e_lo, int q) {
    return q < 0 ? ldexp(1.0, e_lo) : ldexp(1.0, e_lo + q + 1);
}
__device__ __forceinline__ int slot_of(double s, int e_lo) {
    if (s < ldexp(1.0, e_lo)) return -1;
    const int q = ilogb(s) - e_lo;
    return q < kBinades - 1 ? q : kBinades - 1;
}

// Replays one block element by element from running sum *s (slot *q). Lane-parallel within
// 64-element chunks; each crossing addition is a real FP32 add.
__device__ void replay_block(const float* __restrict__ base, int64_t n, int64_t block, int e_lo, double* s,
                             int* q, int lane) {
    const int64_t start = block * kSeqSumBlock;
    constexpr int kChunks = kSeqSumBlock / 64;
    float v[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
        const int64_t i = start + int64_t(c) * 64 + lane;
        v[c] = i < n ? base[i] : 0.0f;
    }
#pragma unroll 1
    for (int c = 0; c < kChunks; ++c) {
        int first = 0;
        while (true) {
            const double contrib = lane >= first ? contribution(v[c], e_lo, *q) : 0.0;
            const double incl = wave_inclusive_scan(contrib, lane);
            const double limit = slot_limit(e_lo, *q);
            const bool over = lane >= first && (*s + incl) >= limit;
            const unsigned long long mask = __ballot(over);
            if (mask == 0ull) {
                *s += __shfl(incl, 63);
                break;
            }
            const int l = __ffsll((long long)mask) - 1;
            const double before = __shfl(incl - contrib, l);
            const float xl = __shfl(v[c], l);
            const float sf = float(*s + before);   // the running sum is an FP32 value
            *s = double(sf + xl);                   // the crossing addition, in FP32
            *q = slot_of(*s, e_lo);
            first = l + 1;
            if (first >= 64) break;
        }
    }
}

__global__ void __launch_bounds__(64)
walk_kernel(const float* __restrict__ x, int64_t n, int64_t stride, int nb, const int* __restrict__ elo,
            const double* __restrict__ bsum, const double* __restrict__ bins, float* __restrict__ out) {
    const int seg = blockIdx.x;
    const int lane = threadIdx.x;
    const int e_lo = elo[seg];
    if (e_lo == INT_MIN) {
        if (lane == 0) out[seg] = 0.0f;
        return;
    }
    const float* base = x + seg * stride;
    const double* bs = bsum + int64_t(seg) * nb;
    const double* bn = bins + int64_t(seg) * kBinades * nb;
    double s = 0.0;
    int q = -1;
#pragma unroll 1
    for (int g0 = 0; g0 < nb; g0 += 64) {
        const int blk = g0 + lane;
        int first = 0;
        while (first < 64) {
            double col = 0.0;
            if (lane >= first && blk < nb) col = q < 0 ? bs[blk] : bn[int64_t(q) * nb + blk];
            const double incl = wave_inclusive_scan(col, lane);
            const bool over = lane >= first && blk < nb && (s + incl) >= slot_limit(e_lo, q);
            const unsigned long long mask = __ballot(over);
            if (mask == 0ull) {
                s += __shfl(incl, 63);
                break;
            }
            const int l = __ffsll((long long)mask) - 1;
            s += __shfl(incl - col, l);
            if (q >= 0) s = double(float(s));
            replay_block(base, n, g0 + l, e_lo, &s, &q, lane);
            first = l + 1;
        }
    }
    if (lane == 0) out[seg] = float(s);
}

__global__ void scale_by_inverse_kernel(float* __restrict__ x, int64_t n, int64_t stride,
                                        const float* __restrict__ totals, int only_positive) {
    const int seg = blockIdx.y;
    const float total = totals[seg];
    if (total <= 0.0f) return;  // the reference leaves a non-positive total untouched
    const float inv = 1.0f / total;
    float* base = x + seg * stride;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        base[i] *= inv;
    (void)only_positive;
}

}  // namespace

void SeqSum::run(hipStream_t s, const float* x, int64_t n, int64_t stride, int segments, float* totals) {
    if (n <= 0 || segments <= 0) return;
    const int nb = int(div_up(n, kSeqSumBlock));
    bsum_.reset(sizeof(double) * size_t(nb) * segments);
    bins_.reset(sizeof(double) * size_t(nb) * segments * kBinades);
    elo_.reset(sizeof(int) * size_t(segments));
    hipLaunchKernelGGL(block_sums_kernel, dim3(nb, segments), dim3(kThreads), 0, s, x, n, stride, nb,
                       bsum_.as<double>());
    hipLaunchKernelGGL(window_kernel, dim3(segments), dim3(kThreads), 0, s, bsum_.as<double>(), nb, elo_.as<int>());
    hipLaunchKernelGGL(binade_sums_kernel, dim3(nb, segments), dim3(kThreads), 0, s, x, n, stride, nb,
                       elo_.as<int>(), bins_.as<double>());
    hipLaunchKernelGGL(walk_kernel, dim3(segments), dim3(64), 0, s, x, n, stride, nb, elo_.as<int>(),
                       bsum_.as<double>(), bins_.as<double>(), totals);
    LDH_CHECK_LAUNCH("seqsum-launch");
}

void SeqSum::normalize(hipStream_t s, float* x, int64_t n, int64_t stride, int segments, float* totals) {
    run(s, x, n, stride, segments, totals);
    const int grid = int(std::min<int64_t>(div_up(n, 256 * 4), 2048));
    hipLaunchKernelGGL(scale_by_inverse_kernel, dim3(grid, segments), dim3(256), 0, s, x, n, stride, totals, 0);
    LDH_CHECK_LAUNCH("normalize-launch");
}

}  // namespace ldh

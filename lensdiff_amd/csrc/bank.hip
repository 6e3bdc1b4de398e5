// Bank construction on the device: resampling (makeResampledKernel / sampleSquareFiltered,
// LensDiffCpuReference.cpp:316-353, :1060-1086) and shaping (azimuthalMeanKernel,
// reshapeAnisotropy, positiveResidualKernel, estimateAdaptiveSupportRadius,
// paddedAdaptiveSupportRadius, cropKernelToRadius, applySupportBoundaryTaper, :957-1151),
// all batched over the wavelengths of one bank.
//
// Ring sums (azimuthal means, support-radius energies, the radial profile) are the reference's
// float sums in raster order, reproduced exactly by rings.hip. Every unit-sum normalisation goes
// through SeqSum so the scale matches the reference's sequential FP32 accumulation; the kernels
// that write a normalised array (resample, ring mean, reshape, residual, crop + taper) place its
// blocks for that sum as they write them.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "bank.h"
#include "host_params.h"
#include "profiler.h"

namespace ldh {

namespace {

// 1 / totals[seg] when totals is given and the total is positive, else 1 (a non-positive sum
// leaves an array unnormalised, LensDiffCpuReference.cpp:440-449).
__device__ __forceinline__ float inverse_total(const float* __restrict__ totals, int seg) {
    if (totals == nullptr) return 1.0f;
    const float t = totals[seg];
    return t > 0.0f ? 1.0f / t : 1.0f;
}

// Bilinear sample of img * norm at (x, y) of a size x size image (sampleSquareBilinear,
// LensDiffCpuReference.cpp:316-336): each texel is scaled before interpolation, as when the image
// was normalised in place first. The texels are held from (org, org) on, `side` per row (the whole
// image: org 0, side = size; a centred window of it otherwise).
__device__ __forceinline__ float bilinear_sq(const float* __restrict__ img, int size, int org, int side, float x,
                                             float y, float norm) {
    if (x < 0.0f || y < 0.0f || x > float(size - 1) || y > float(size - 1)) return 0.0f;
    const int x0 = int(floorf(x)), y0 = int(floorf(y));
    const int x1 = min(x0 + 1, size - 1), y1 = min(y0 + 1, size - 1);
    const float tx = x - float(x0), ty = y - float(y0);
    const size_t r0 = size_t(y0 - org) * side, r1 = size_t(y1 - org) * side;
    const float v00 = img[r0 + (x0 - org)] * norm, v10 = img[r0 + (x1 - org)] * norm;
    const float v01 = img[r1 + (x0 - org)] * norm, v11 = img[r1 + (x1 - org)] * norm;
    const float a = v00 + (v10 - v00) * tx;
    const float b = v01 + (v11 - v01) * tx;
    return a + (b - a) * ty;
}

// Elements per thread of a block; each thread's element j is block element j * kSumThreads + tid.
// Kernels that gather (ring means) or read several arrays per texel (crop + taper) issue all of
// a thread's loads before using them (K*K < 2^31).
constexpr int kPerSum = kSeqSumBlock / kSumThreads;

// Workgroups of kSumThreads threads over blocks of kSeqSumBlock texels (seqsum_blocks(K*K) per
// kernel); with an active placement (pl.bsum set) each workgroup places its block for the
// sequential sum of the output (segment lam). The raw PSFs are R x R images held whole (org 0,
// side R) or as their centred side x side windows from (org, org).
__global__ void __launch_bounds__(kSumThreads)
resample_kernel(const float* __restrict__ raw, int R, int org, int side, const float* __restrict__ scales,
                const float* __restrict__ raw_totals, int support, int count, float* __restrict__ out, SeqPlace pl) {
    const int K = 2 * support + 1;
    const int kk = K * K;
    SegBlock sb;
    if (!place_block_of(seqsum_blocks(kk), count, sb)) return;
    const int lam = sb.seg;
    // raw_totals: the raw PSFs are not yet normalised; texels are scaled by 1/total
    const float norm = inverse_total(raw_totals, lam);
    const float* img = raw + size_t(lam) * side * side;
    const float rc = (R & 1) == 0 ? float(R) * 0.5f : float(R - 1) * 0.5f;
    const float kc = float(K - 1) * 0.5f;
    const float inv = 1.0f / fmaxf(scales[lam], 0.05f);
    const int taps = min(max(int(ceilf(inv)), 2), 6);
    const float step = inv / float(taps);
    const int start = sb.block * kSeqSumBlock, end = min(kk, start + kSeqSumBlock);
    double share[1] = {0.0};
    float vals[1][kPerSum];
#pragma unroll
    for (int j = 0; j < kPerSum; ++j) {
        const int i = start + int(threadIdx.x) + j * kSumThreads;
        vals[0][j] = 0.0f;
        if (i >= end) continue;
        const int x = i % K, y = i / K;
        const float sx = rc + (float(x) - kc) * inv;
        const float sy = rc + (float(y) - kc) * inv;
        float v;
        if (inv <= 1.0f) {
            v = bilinear_sq(img, R, org, side, sx, sy, norm);
        } else {
            float sum = 0.0f;
            for (int ty = 0; ty < taps; ++ty) {
                const float oy = (float(ty) + 0.5f) * step - inv * 0.5f;
                for (int tx = 0; tx < taps; ++tx) {
                    const float ox = (float(tx) + 0.5f) * step - inv * 0.5f;
                    sum += bilinear_sq(img, R, org, side, sx + ox, sy + oy, norm);
                }
            }
            v = sum / float(taps * taps);
        }
        out[size_t(lam) * kk + i] = v;
        vals[0][j] = v;
        share[0] += double(v);
    }
    if (pl.bsum == nullptr) return;
    place_blocks<1, kPerSum>(pl, lam, 0, sb.block, vals, share);
}

// Ring index of pixel (x, y) about centre c: round(|d|) (azimuthal mean, radial profile) or
// ceil(|d|) (support radius), clamped to rmax. dx, dy are exact in float, sqrt is correctly
// rounded, so the ring of every pixel equals the reference's.
template <bool CEIL>
__device__ __forceinline__ int ring_of(int x, int y, float c, int rmax) {
    const float dx = float(x) - c, dy = float(y) - c;
    const float d = sqrtf(dx * dx + dy * dy);
    const int r = CEIL ? int(ceilf(d)) : int(roundf(d));
    return min(rmax, r);
}

// The elementwise shaping kernels below run workgroups of kSumThreads threads over blocks of
// kSeqSumBlock elements of one kernel (grid: place_grid(seqsum_blocks(K*K), kernels)) and place
// the blocks of each array they produce for its sequential sum.

// Advances (x, y) by kSumThreads texels along the rows of a K-wide image.
__device__ __forceinline__ void advance_texel(int& x, int& y, int K) {
    x += kSumThreads;
    while (x >= K) {
        x -= K;
        ++y;
    }
}

// Mean of each ring, sum / float(count) (count at least 1), per segment (:1106-1114).
__global__ void ring_means_kernel(const float* __restrict__ sums, const int* __restrict__ counts, int nr,
                                  float* __restrict__ means) {
    const int seg = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    means[int64_t(seg) * nr + r] = sums[int64_t(seg) * nr + r] / float(max(1, counts[r]));
}

// The azimuthal-mean image: each texel takes its ring's mean (the ring means precomputed once).
__global__ void __launch_bounds__(kSumThreads)
ring_mean_kernel(const float* __restrict__ means, int K, int rmax, int count, float* __restrict__ out, SeqPlace pl) {
    const int kk = K * K;
    SegBlock sb;
    if (!place_block_of(seqsum_blocks(kk), count, sb)) return;
    const int seg = sb.seg;
    const float c = float(K - 1) * 0.5f;
    const int nr = rmax + 1;
    const int start = sb.block * kSeqSumBlock, end = min(kk, start + kSeqSumBlock);
    const int i0 = start + int(threadIdx.x);
    int x = i0 % K, y = i0 / K;
    float mv[kPerSum];
#pragma unroll
    for (int j = 0; j < kPerSum; ++j) {
        mv[j] = means[int64_t(seg) * nr + ring_of<false>(x, y, c, rmax)];
        advance_texel(x, y, K);
    }
    double share[1] = {0.0};
    float vals[1][kPerSum];
#pragma unroll
    for (int j = 0; j < kPerSum; ++j) {
        const int i = i0 + j * kSumThreads;
        vals[0][j] = i < end ? mv[j] : 0.0f;
        if (i >= end) continue;
        out[int64_t(seg) * kk + i] = mv[j];
        share[0] += double(mv[j]);
    }
    place_blocks<1, kPerSum>(pl, seg, 0, sb.block, vals, share);
}

// reshapeAnisotropy on the normalised base and mean (each read through its totals).
__global__ void __launch_bounds__(kSumThreads)
reshape_kernel(const float* __restrict__ base, const float* __restrict__ base_tot, const float* __restrict__ mean,
               const float* __restrict__ mean_tot, float gain0, const float* __restrict__ gains, int64_t kk,
               int count, float* __restrict__ shaped, SeqPlace pl) {
    SegBlock sb;
    if (!place_block_of(seqsum_blocks(kk), count, sb)) return;
    const int seg = sb.seg;
    const float gain = gains != nullptr ? gains[seg] : gain0;
    const float inv_m = inverse_total(mean_tot, seg), inv_b = inverse_total(base_tot, seg);
    const int64_t start = int64_t(sb.block) * kSeqSumBlock, end = min(kk, start + kSeqSumBlock);
    double share[1] = {0.0};
    float vals[1][kPerSum];
#pragma unroll
    for (int j = 0; j < kPerSum; ++j) {
        const int64_t i = start + threadIdx.x + int64_t(j) * kSumThreads;
        vals[0][j] = 0.0f;
        if (i >= end) continue;
        const int64_t o = int64_t(seg) * kk + i;
        const float m = mean[o] * inv_m;
        const float b = base[o] * inv_b;
        const float v = fmaxf(0.0f, m + (b - m) * gain);
        shaped[o] = v;
        vals[0][j] = v;
        share[0] += double(v);
    }
    place_blocks<1, kPerSum>(pl, seg, 0, sb.block, vals, share);
}

// positiveResidualKernel of the normalised shaped and mean kernels (STRUCT), placed for its sum
// (segment seg); also places the normalised shaped kernel, the support estimate's total (segment
// count + seg with STRUCT, else seg).
template <bool STRUCT>
__global__ void __launch_bounds__(kSumThreads)
residual_kernel(const float* __restrict__ shaped, const float* __restrict__ shaped_tot, const float* __restrict__ mean,
                const float* __restrict__ mean_tot, int64_t kk, int count, float* __restrict__ out, SeqPlace pl) {
    constexpr int N = STRUCT ? 2 : 1;
    SegBlock sb;
    if (!place_block_of(seqsum_blocks(kk), count, sb)) return;
    const int seg = sb.seg;
    const float inv_s = inverse_total(shaped_tot, seg);
    const float inv_m = STRUCT ? inverse_total(mean_tot, seg) : 0.0f;
    const int64_t start = int64_t(sb.block) * kSeqSumBlock, end = min(kk, start + kSeqSumBlock);
    double share[N];
    float vals[N][kPerSum];
#pragma unroll
    for (int k = 0; k < N; ++k) share[k] = 0.0;
#pragma unroll
    for (int j = 0; j < kPerSum; ++j) {
        const int64_t i = start + threadIdx.x + int64_t(j) * kSumThreads;
#pragma unroll
        for (int k = 0; k < N; ++k) vals[k][j] = 0.0f;
        if (i >= end) continue;
        const int64_t o = int64_t(seg) * kk + i;
        const float sn = shaped[o] * inv_s;
        if constexpr (STRUCT) {
            const float v = fmaxf(0.0f, sn - mean[o] * inv_m);
            out[o] = v;
            vals[0][j] = v;
            share[0] += double(v);
        }
        vals[N - 1][j] = sn;
        share[N - 1] += double(sn);
    }
    place_blocks<N, kPerSum>(pl, seg, count, sb.block, vals, share);
}

// estimateAdaptiveSupportRadius + paddedAdaptiveSupportRadius (LensDiffCpuReference.cpp:984-1036),
// one workgroup per bin: the ring energies and peaks are staged in LDS by the workgroup, then one
// wave finishes. outsidePeak[r] = max(peak[r..rmax]) is a suffix-max scan in 64-ring chunks from
// the end. The reference's float running sum cum_r = cum_(r-1) + float(e_r) is kept exactly: a
// chunk's 64 sums are formed one after the other from the chunk's values read across lanes
// (readlane, no memory round trip per ring), then every lane tests its ring (captured =
// cum_r / total >= 0.9999 and outsidePeak[r + 1] <= 2e-4 peak) and the first ring that passes,
// the reference's break, is found by a ballot.
constexpr int kSupportThreads = 256;
__global__ void __launch_bounds__(kSupportThreads)
support_radius_kernel(const float* __restrict__ energy, const unsigned* __restrict__ peaks,
                      const float* __restrict__ totals, int rmax, int max_radius, int* __restrict__ radius) {
    extern __shared__ unsigned char smem[];
    const int seg = blockIdx.x;
    const int nr = rmax + 1;
    float* e = reinterpret_cast<float*>(smem);
    unsigned* pk = reinterpret_cast<unsigned*>(e + nr);
    for (int r = threadIdx.x; r < nr; r += kSupportThreads) {
        e[r] = energy[int64_t(seg) * nr + r];
        pk[r] = peaks[int64_t(seg) * nr + r];
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    const int lane = int(threadIdx.x);
    // suffix maxima, last chunk first; carry: the maximum beyond the chunk
    unsigned carry = 0u;
    for (int base = (rmax / 64) * 64; base >= 0; base -= 64) {
        const int r = base + lane;
        unsigned v = r <= rmax ? pk[r] : 0u;
#pragma unroll
        for (int d = 1; d < 64; d *= 2) {
            const unsigned t = __shfl_down(v, d);
            if (lane + d < 64) v = max(v, t);
        }
        v = max(v, carry);
        if (r <= rmax) pk[r] = v;
        carry = __shfl(v, 0);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float total = totals[seg];
    const float gpeak = __uint_as_float(pk[0]);
    int est = max(4, rmax);
    if (total > 1e-6f && gpeak > 1e-6f) {
        float cum = 0.0f;
        for (int base = 0; base <= rmax; base += 64) {
            const int r = base + lane;
            const float ev = r <= rmax ? e[r] : 0.0f;
            float mine = 0.0f;
#pragma unroll
            for (int j = 0; j < 64; ++j) {
                cum += __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ev), j));
                if (lane == j) mine = cum;
            }
            const float captured = mine / total;
            const float remaining = r + 1 <= rmax ? __uint_as_float(pk[r + 1]) : 0.0f;
            const unsigned long long hit = __ballot(r <= rmax && captured >= 0.9999f && remaining <= gpeak * 2e-4f);
            if (hit != 0ull) {
                est = max(4, base + __ffsll(static_cast<long long>(hit)) - 1 + 1);
                break;
            }
        }
    }
    if (lane != 0) return;
    const int pad = max(6, min(24, int(ceilf(float(max(1, est)) * 0.05f))));
    const int hi = max(4, max_radius);
    radius[seg] = min(max(est + pad, 4), hi);
}

// cropKernelToRadius + applySupportBoundaryTaper on the K x K grid, for the full (shaped), core
// (mean) and structure kernels of a bin at once, the first NS of them (they share the radius and
// so the taper weight): keeps the centred (2r+1)^2 square, multiplies by the cosine taper, zeroes
// everything else. Sources are read through their totals (normalised values); dst is
// [NS][count][K*K].
struct TaperSources {
    const float* src[3];
    const float* tot[3];
};
template <int NS>
__global__ void __launch_bounds__(kSumThreads)
crop_taper_kernel(TaperSources ts, const int* __restrict__ radius, int K, int count, float* __restrict__ dst,
                  SeqPlace pl) {
    const int64_t kk = int64_t(K) * K;
    SegBlock sb;
    if (!place_block_of(seqsum_blocks(kk), count, sb)) return;
    const int bin = sb.seg;
    const int r = min(max(radius[bin], 1), K / 2);
    const int c = K / 2;
    const int64_t start = int64_t(sb.block) * kSeqSumBlock, end = min(kk, start + kSeqSumBlock);
    float inv[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) inv[k] = inverse_total(ts.tot[k], bin);
    const float extent = fmaxf(1.0f, float(r));
    const float fade = fmaxf(6.0f, fminf(24.0f, extent * 0.04f));
    const float ramp = fmaxf(0.0f, extent - fade);
    // loads of the texels the taper keeps first (all of the thread's), then the taper: a texel of
    // the square whose distance is at least extent + 1 (so rr >= extent) gets weight 0 and is
    // not read
    const int i0 = int(start) + int(threadIdx.x);
    const int64_t o0 = int64_t(bin) * kk;
    const float far2 = (extent + 1.0f) * (extent + 1.0f);
    float sv[NS][kPerSum];
    {
        int x = i0 % K, y = i0 / K;
#pragma unroll
        for (int j = 0; j < kPerSum; ++j) {
            const float fx = float(x - c), fy = float(y - c);
            const bool inside = i0 + j * kSumThreads < end && abs(x - c) <= r && abs(y - c) <= r &&
                                fx * fx + fy * fy < far2;
#pragma unroll
            for (int k = 0; k < NS; ++k) sv[k][j] = inside ? ts.src[k][o0 + i0 + j * kSumThreads] : 0.0f;
            advance_texel(x, y, K);
        }
    }
    double share[NS];
    float vals[NS][kPerSum];
#pragma unroll
    for (int k = 0; k < NS; ++k) share[k] = 0.0;
    int x = i0 % K, y = i0 / K;
#pragma unroll
    for (int j = 0; j < kPerSum; ++j) {
        const int i = i0 + j * kSumThreads;
#pragma unroll
        for (int k = 0; k < NS; ++k) vals[k][j] = 0.0f;
        if (i >= end) continue;
        const bool inside = abs(x - c) <= r && abs(y - c) <= r;
        float w = 0.0f;
        if (inside) {
            const float dx = float(x - c), dy = float(y - c);
            const float rr = sqrtf(dx * dx + dy * dy);
            if (rr <= ramp) {
                w = 1.0f;  // t = 0 up to the ramp: cos(0) = 1 exactly
            } else if (rr < extent) {
                const float t = fminf(fmaxf((rr - ramp) / fmaxf(fade, 1e-6f), 0.0f), 1.0f);
                const float sm = t * t * (3.0f - 2.0f * t);
                w = fmaxf(0.0f, cosf(sm * 1.57079637f));
            }
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const float v = inside ? sv[k][j] * inv[k] * w : 0.0f;
            dst[(int64_t(k) * count + bin) * kk + i] = v;
            vals[k][j] = v;
            share[k] += double(v);
        }
        advance_texel(x, y, K);
    }
    place_blocks<NS, kPerSum>(pl, bin, count, sb.block, vals, share);
}

// x[seg] *= scales[seg] (segments [3][count] of K*K) over the crop square of the segment's bin
// only: crop_taper_kernel leaves every texel outside it at 0, which the scale keeps.
__global__ void scale_crop_kernel(float* __restrict__ x, int K, const int* __restrict__ radius,
                                  const float* __restrict__ scales) {
    const int seg = blockIdx.y, count = gridDim.y / 3, bin = seg % count;
    const float inv = scales[seg];
    if (inv == 1.0f) return;
    const int r = min(max(radius[bin], 1), K / 2), c = K / 2;
    const int lo = c - r, side = min(K - 1, c + r) - lo + 1;  // lo >= 0; K even: c + r may reach K
    float* base = x + int64_t(seg) * K * K + int64_t(lo) * K + lo;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < side * side; i += gridDim.x * blockDim.x) {
        const int y = i / side, xx = i - y * side;
        base[int64_t(y) * K + xx] *= inv;
    }
}

__global__ void inverse_totals_kernel(const float* __restrict__ totals, int n, float* __restrict__ scales) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) scales[i] = inverse_total(totals, i);
}

__global__ void fill_zero_kernel(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = 0u;
}

inline dim3 place_grid_of(int64_t n, int segs) { return dim3(place_grid(seqsum_blocks(n), segs)); }

// Segments [count] of stride kk in one array, scaled by 1 / prescale[seg] when prescale is given.
inline SeqSegments one_array(const float* x, int count, int64_t kk, const float* prescale = nullptr) {
    SeqSegments g;
    g.array[0] = x;
    g.prescale[0] = prescale;
    g.per_array = count;
    g.stride = kk;
    return g;
}

}  // namespace

void resample_kernels(hipStream_t s, const float* raw, int R, const float* d_scales, const float* d_raw_totals,
                      int support, int count, float* out, const SeqPlace* place, int window) {
    const int K = 2 * support + 1;
    const int side = window > 0 ? 2 * window + 1 : R, org = window > 0 ? R / 2 - window : 0;
    ProfileScope ps_("resample", 8.0 * double(K) * K * count);
    hipLaunchKernelGGL(resample_kernel, place_grid_of(int64_t(K) * K, count), dim3(kSumThreads), 0, s, raw, R, org,
                       side, d_scales, d_raw_totals, support, count, out, place != nullptr ? *place : SeqPlace{});
    LDH_CHECK_LAUNCH("resample");
}

void base_kernels(hipStream_t s, BankWorkspace& w, const float* raw, int R, const float* d_scales,
                  const float* d_raw_totals, int support, int count, float* out, float* totals, int window) {
    const int K = 2 * support + 1;
    const int64_t kk = int64_t(K) * K;
    w.reserve_shaping(s, K, count, support);
    const SeqPlace pl = w.seqsum.place(kk, count);
    resample_kernels(s, raw, R, d_scales, d_raw_totals, support, count, out, &pl, window);
    w.seqsum.walk(s, one_array(out, count, kk), kk, count, totals);
}

const RingPlan& BankWorkspace::ring_plan(hipStream_t s, int size, float center, int rmax, bool ceil) {
    // least recently used last out: a batch's set-up touches every plan its capture will use
    for (size_t i = 0; i < ring_plans.size(); ++i)
        if (ring_plans[i]->matches(size, center, rmax, ceil)) {
            std::unique_ptr<RingPlan> hit = std::move(ring_plans[i]);
            ring_plans.erase(ring_plans.begin() + std::ptrdiff_t(i));
            ring_plans.push_back(std::move(hit));
            return *ring_plans.back();
        }
    constexpr size_t kKeep = 32;  // two per kernel size
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    LDH_HIP(hipStreamIsCapturing(s, &cap), "ring-plan");
    if (cap != hipStreamCaptureStatusNone) fail(LDH_ERR_DEVICE, "ring-plan-built-during-capture");
    if (ring_plans.size() >= kKeep) ring_plans.erase(ring_plans.begin());
    ring_plans.push_back(std::make_unique<RingPlan>());
    prepare_ring_plan(s, *ring_plans.back(), size, center, rmax, ceil);
    return *ring_plans.back();
}

void BankWorkspace::reserve_shaping(hipStream_t s, int K, int count, int max_radius) {
    const int64_t kk = int64_t(K) * K;
    const int64_t n = kk * count;
    mean.reset(sizeof(float) * n);
    shaped.reset(sizeof(float) * n);
    structure.reset(sizeof(float) * n);
    // sums: mean, shaped, structure, normalised-shaped run (count each), final (3 * count)
    shape_totals.reset(sizeof(float) * 7 * count);
    const int nr = std::max(K / 2, max_radius) + 1;
    ring_sum.reset(sizeof(float) * nr * count);
    ring_peak.reset(sizeof(unsigned) * nr * count);
    ring_mean.reset(sizeof(float) * nr * count);
    seqsum.reserve(kk, 3 * count);
    const float c = float(K - 1) * 0.5f;
    ring_scratch.reserve(ring_plan(s, K, c, K / 2, false), count);
    ring_scratch.reserve(ring_plan(s, K, c, std::min(max_radius, K / 2), true), count);
}

void apply_kernel_scales(hipStream_t s, float* out3, int K, int count, const int* d_radius, const float* scales) {
    const int64_t kk = int64_t(K) * K;
    hipLaunchKernelGGL(scale_crop_kernel, dim3(unsigned(std::min<int64_t>(div_up(kk, 256), 1024)), unsigned(3 * count)),
                       dim3(256), 0, s, out3, K, d_radius, scales);
    LDH_CHECK_LAUNCH("kernel-scales");
}

void shape_kernels(hipStream_t s, BankWorkspace& w, const float* base, const float* base_totals, int K, int count,
                   double anisotropy, int max_radius, float* out3, int* d_radius, float* out_scales,
                   const float* d_gains, int sets) {
    if (sets < 1 || sets > 3) fail(LDH_ERR_INVALID_ARGUMENT, "kernel-sets");
    const int64_t kk = int64_t(K) * K;
    const int rmax_mean = K / 2;
    w.reserve_shaping(s, K, count, max_radius);
    const float kc = float(K - 1) * 0.5f;
    const RingPlan& mean_rings = w.ring_plan(s, K, kc, rmax_mean, false);
    float* t_mean = w.shape_totals.as<float>();
    float* t_shaped = t_mean + count;
    float* t_struct = t_shaped + count;
    float* t_run = t_struct + count;
    float* t_final = t_run + count;
    // Each normalisation below only sums its array; readers apply 1 / total as they load, which
    // gives the values an in-place normalisation would have left.
    // azimuthal mean of the normalised base
    ring_sums(s, mean_rings, w.ring_scratch, w.seqsum, base, kk, count, base_totals, w.ring_sum.as<float>(), nullptr);
    // Every producer below places the blocks of the array it writes for the sequential sum that
    // follows it (SeqSum::place / walk), so the sums read no data except in the walk's replays.
    const dim3 grid = place_grid_of(kk, count);
    {
        ProfileScope ps_("shape-ring-mean", 4.0 * double(kk) * count);
        float* means = w.ring_mean.as<float>();
        hipLaunchKernelGGL(ring_means_kernel, dim3(unsigned(div_up(rmax_mean + 1, 256)), unsigned(count)), dim3(256), 0,
                           s, w.ring_sum.as<float>(), mean_rings.count.as<int>(), rmax_mean + 1, means);
        hipLaunchKernelGGL(ring_mean_kernel, grid, dim3(kSumThreads), 0, s, means, K, rmax_mean, count,
                           w.mean.as<float>(), w.seqsum.place(kk, count));
        LDH_CHECK_LAUNCH("ring-mean");
    }
    w.seqsum.walk(s, one_array(w.mean.as<float>(), count, kk), kk, count, t_mean);
    // anisotropy reshape
    const float gain = 1.0f + std::max(0.0f, float(anisotropy)) * 4.0f;
    {
        ProfileScope ps_("shape-reshape", 12.0 * double(kk) * count);
        hipLaunchKernelGGL(reshape_kernel, grid, dim3(kSumThreads), 0, s, base, base_totals, w.mean.as<float>(),
                           t_mean, gain, d_gains, kk, count, w.shaped.as<float>(), w.seqsum.place(kk, count));
        LDH_CHECK_LAUNCH("reshape");
    }
    w.seqsum.walk(s, one_array(w.shaped.as<float>(), count, kk), kk, count, t_shaped);
    // positive residual (only when the structure kernels are formed); the pass also places the
    // normalised shaped kernels (the support estimate's total), so one walk sums both: t_struct,
    // then t_run
    const bool with_structure = sets == 3;
    {
        ProfileScope ps_("shape-residual", (with_structure ? 12.0 : 4.0) * double(kk) * count);
        if (with_structure)
            hipLaunchKernelGGL(residual_kernel<true>, grid, dim3(kSumThreads), 0, s, w.shaped.as<float>(), t_shaped,
                               w.mean.as<float>(), t_mean, kk, count, w.structure.as<float>(),
                               w.seqsum.place(kk, 2 * count));
        else
            hipLaunchKernelGGL(residual_kernel<false>, grid, dim3(kSumThreads), 0, s, w.shaped.as<float>(), t_shaped,
                               static_cast<const float*>(nullptr), static_cast<const float*>(nullptr), kk, count,
                               static_cast<float*>(nullptr), w.seqsum.place(kk, count));
        LDH_CHECK_LAUNCH("residual");
    }
    SeqSegments g;
    g.per_array = count;
    g.stride = kk;
    if (with_structure) {
        g.array[0] = w.structure.as<float>();
        g.array[1] = w.shaped.as<float>();
        g.prescale[1] = t_shaped;
        w.seqsum.walk(s, g, kk, 2 * count, t_struct);
    } else {
        g.array[0] = w.shaped.as<float>();
        g.prescale[0] = t_shaped;
        w.seqsum.walk(s, g, kk, count, t_run);
    }
    // adaptive support radius of the normalised shaped kernel
    const int rmax_sup = std::min(max_radius, K / 2);
    // The last ring's energy only enters the reference's running sum at r = rmax, whose padded
    // result is clamped to the maximum radius either way when rmax + 6 reaches it (always, for
    // K = 2 * max_radius + 1): its sequential sum is formed only when it can matter.
    const bool last_matters = rmax_sup + 6 < std::max(4, max_radius);
    ring_sums(s, w.ring_plan(s, K, kc, rmax_sup, true), w.ring_scratch, w.seqsum, w.shaped.as<float>(), kk, count,
              t_shaped, w.ring_sum.as<float>(), w.ring_peak.as<unsigned>(), last_matters);
    const size_t sup_lds = size_t(rmax_sup + 1) * (sizeof(float) + sizeof(unsigned));
    {
        ProfileScope ps_("shape-support-radius", 0.0);
        hipLaunchKernelGGL(support_radius_kernel, dim3(unsigned(count)), dim3(kSupportThreads), sup_lds, s,
                           w.ring_sum.as<float>(), w.ring_peak.as<unsigned>(), t_run, rmax_sup, max_radius, d_radius);
        LDH_CHECK_LAUNCH("support-radius");
    }
    // crop + taper of full (shaped), core (mean), structure (the first `sets`), then normalise them
    TaperSources ts{{w.shaped.as<float>(), w.mean.as<float>(), w.structure.as<float>()}, {t_shaped, t_mean, t_struct}};
    {
        ProfileScope ps_("shape-crop-taper", 8.0 * sets * double(kk) * count);
        auto crop = sets == 3 ? crop_taper_kernel<3> : sets == 2 ? crop_taper_kernel<2> : crop_taper_kernel<1>;
        hipLaunchKernelGGL(crop, grid, dim3(kSumThreads), 0, s, ts, d_radius, K, count, out3,
                           w.seqsum.place(kk, sets * count));
        LDH_CHECK_LAUNCH("crop-taper");
    }
    w.seqsum.walk(s, one_array(out3, sets * count, kk), kk, sets * count, t_final);
    // the unit-sum scales travel with the kernels; readers multiply as they load
    hipLaunchKernelGGL(inverse_totals_kernel, dim3(unsigned(div_up(sets * count, 256))), dim3(256), 0, s, t_final,
                       sets * count, out_scales);
    LDH_CHECK_LAUNCH("kernel-scales");
}

std::vector<float> radial_profile(hipStream_t s, BankWorkspace& w, const float* psf, int R) {
    const int rmax = R / 2;
    const int nr = rmax + 1;
    const float center = (R & 1) == 0 ? float(R) * 0.5f : float(R - 1) * 0.5f;
    // one profile per (P, R) and workspace: the plan is not kept
    RingPlan plan;
    prepare_ring_plan(s, plan, R, center, rmax, false);
    RingScratch scratch;
    DeviceBuffer sums_d(sizeof(float) * nr);
    ring_sums(s, plan, scratch, w.seqsum, psf, int64_t(R) * R, 1, nullptr, sums_d.as<float>(), nullptr);
    std::vector<float> sums(static_cast<size_t>(nr));
    LDH_HIP(hipMemcpyAsync(sums.data(), sums_d.get(), sizeof(float) * nr, hipMemcpyDeviceToHost, s), "profile");
    LDH_HIP(hipStreamSynchronize(s), "profile");
    std::vector<float> prof(static_cast<size_t>(nr));
    for (int r = 0; r < nr; ++r) {
        const int cnt = plan.host_off[size_t(r) + 1] - plan.host_off[size_t(r)];
        prof[size_t(r)] = sums[size_t(r)] / float(std::max(1, cnt));
    }
    return prof;
}

// estimateFirstMinimumRadius on a radial profile (LensDiffCpuReference.cpp:889-955).
float first_minimum_from_profile(const std::vector<float>& prof, int size) {
    const int n = int(prof.size());
    if (n < 4) return 1.0f;
    std::vector<float> sm(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        const float l = prof[size_t(i == 0 ? 0 : i - 1)];
        const float r = prof[size_t(i + 1 < n ? i + 1 : i)];
        sm[size_t(i)] = (l + 2.0f * prof[size_t(i)] + r) * 0.25f;
    }
    auto refine = [](float l, float c, float r) {
        const float den = l - 2.0f * c + r;
        if (std::fabs(den) < 1e-6f) return 0.0f;
        return std::max(-0.5f, std::min(0.5f * (l - r) / den, 0.5f));
    };
    int best = 1;
    float best_v = sm[1];
    int descents = sm[1] < sm[0] ? 1 : 0;
    for (int i = 2; i + 1 < n; ++i) {
        if (sm[size_t(i)] <= sm[size_t(i - 1)]) {
            descents += 1;
            if (sm[size_t(i)] < best_v) { best_v = sm[size_t(i)]; best = i; }
            continue;
        }
        if (descents >= 2 && best > 0 && best + 1 < n)
            return std::max(1.0f, float(best) + refine(sm[size_t(best - 1)], sm[size_t(best)], sm[size_t(best + 1)]));
        best = i;
        best_v = sm[size_t(i)];
        descents = 0;
    }
    const int limit = std::max(2, std::min(n - 1, std::max(4, size / 8)));
    best = 1;
    best_v = sm[1];
    for (int i = 2; i <= limit; ++i)
        if (sm[size_t(i)] < best_v) { best_v = sm[size_t(i)]; best = i; }
    if (best > 0 && best + 1 < n)
        return std::max(1.0f, float(best) + refine(sm[size_t(best - 1)], sm[size_t(best)], sm[size_t(best + 1)]));
    return float(best);
}

float clear_circle_first_zero(hipStream_t s, const FftTables& t, BankWorkspace& w, int P, int R) {
    for (const auto& e : w.first_zero)
        if (e.first.first == P && e.first.second == R) return e.second;
    ldh_params q;
    ldh_params_default(&q);
    q.aperture_mode = LDH_APERTURE_CIRCLE;
    q.apodization_mode = LDH_APOD_FLAT;
    q.roundness = 1.0;
    q.pupil_resolution = P;
    w.ref_amp.reset(sizeof(float) * P * P);
    w.ref_phase.reset(sizeof(float) * P * P);
    w.ref_raw.reset(sizeof(float) * size_t(R) * R);
    w.ref_total.reset(sizeof(float));
    pupil_amplitude(s, make_pupil_geom(q), w.ref_amp.as<float>(), true);
    LDH_HIP(hipMemsetAsync(w.ref_phase.get(), 0, sizeof(float) * P * P, s), "first-zero");
    w.plan.prepare(s, P, R);
    w.plan.raw_psfs(s, t, w.ref_amp.as<float>(), w.ref_phase.as<float>(), 1, w.ref_raw.as<float>(),
                    w.ref_total.as<float>(), w.seqsum, w.psf);
    const float fz = first_minimum_from_profile(radial_profile(s, w, w.ref_raw.as<float>(), R), R);
    w.first_zero.push_back({{P, R}, fz});
    return fz;
}

bool defocus_only_coefs(const std::vector<PhaseCoef>& coefs) {
    for (size_t i = 1; i < coefs.size(); ++i) {
        PhaseCoef a = coefs[0], b = coefs[i];
        a.c[0] = b.c[0] = 0.0f;
        a.enabled = b.enabled = 0;
        if (std::memcmp(&a, &b, sizeof(PhaseCoef)) != 0) return false;
    }
    return true;
}

void bank_wavelength_params(const ldh_params& p, const Sizes& sz, const std::vector<float>& wl, float first_zero,
                            std::vector<PhaseCoef>& coefs, std::vector<float>& scales) {
    const int L = int(wl.size());
    const float scale_base = float(std::max(1.0, sz.diffraction_scale_px)) / std::max(1.0f, first_zero);
    coefs.resize(size_t(L));
    scales.resize(size_t(L));
    for (int i = 0; i < L; ++i) {
        ldh_params q = p;
        const float off = (wl[size_t(i)] - 550.0f) / 110.0f;
        q.phase_defocus += p.chromatic_focus * off;
        if (std::abs(q.phase_defocus) > 1e-6) q.phase_mode = LDH_PHASE_ENABLED;
        coefs[size_t(i)] = make_phase_coef(q);
        const float chroma = std::max(0.25f, 1.0f + float(p.chromatic_spread) * off);
        scales[size_t(i)] = scale_base * (wl[size_t(i)] / 550.0f) * chroma;
    }
}

std::unique_ptr<DeviceBank> build_bank(hipStream_t s, const FftTables& t, BankWorkspace& w, const ldh_params& p,
                                       int sets) {
    if (sets < 1 || sets > 3) fail(LDH_ERR_INVALID_ARGUMENT, "kernel-sets");
    const Sizes sz = resolve_sizes(p);
    if (sz.max_radius > kMaxKernelRadius) fail(LDH_ERR_TOO_LARGE, "kernel-radius-unsupported");
    if (sz.raw > kMaxRawPsf) fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
    auto bank = std::make_unique<DeviceBank>();
    const std::vector<float> wl = wavelength_list(p);
    const int L = int(wl.size());
    const int P = sz.pupil, R = sz.raw, r = sz.max_radius, K = 2 * r + 1;
    bank->bins = L;
    bank->pupil = P;
    bank->raw = R;
    bank->max_radius = r;
    bank->kernel = K;
    bank->wavelengths = wl;

    const float first_zero = clear_circle_first_zero(s, t, w, P, R);

    // amplitude once, phase per wavelength (chromatic defocus offset, :1720-1727), both transposed
    w.amp.reset(sizeof(float) * P * P);
    pupil_amplitude(s, make_pupil_geom(p), w.amp.as<float>(), true);
    std::vector<PhaseCoef> coefs;
    std::vector<float> scales;
    bank_wavelength_params(p, sz, wl, first_zero, coefs, scales);
    w.coefs.reset(sizeof(PhaseCoef) * L);
    w.scales.reset(sizeof(float) * L);
    LDH_HIP(hipMemcpyAsync(w.coefs.get(), coefs.data(), sizeof(PhaseCoef) * L, hipMemcpyHostToDevice, s), "bank-upload");
    LDH_HIP(hipMemcpyAsync(w.scales.get(), scales.data(), sizeof(float) * L, hipMemcpyHostToDevice, s), "bank-upload");
    // the wavelengths' coefficients differ only in defocus: their phases are formed with the
    // fields (one evaluation of the other terms per texel), unless some other term differs
    const bool defocus_only = defocus_only_coefs(coefs);
    if (!defocus_only) {
        w.phase.reset(sizeof(float) * size_t(P) * P * L);
        pupil_phase(s, w.coefs.as<PhaseCoef>(), P, L, w.phase.as<float>(), true);
    }

    // raw PSFs, then per-wavelength resampled base kernels. The resample reads each raw PSF only
    // near its centre and renormalises the kernel, so where that window is small the raw PSFs are
    // synthesised on it alone and left unnormalised (the unit-sum scale cancels in the kernel's
    // own normalisation); otherwise whole, with their unit-sum scale applied inside the resample.
    w.plan.prepare(s, P, R);
    const int h = resample_window_half_width(scales, r);
    const bool windowed = p.raw_psf_form == LDH_RAW_PSF_AUTO && w.plan.prepare_window(s, h);
    const float* phase = defocus_only ? nullptr : w.phase.as<float>();
    const PhaseCoef* dcoefs = defocus_only ? w.coefs.as<PhaseCoef>() : nullptr;
    const int64_t kk = int64_t(K) * K;
    w.base.reset(sizeof(float) * kk * L);
    w.base_totals.reset(sizeof(float) * L);
    bank->raw_window = windowed ? h : 0;
    if (windowed) {
        const int n = w.plan.window_side();
        w.raw.reset(sizeof(float) * size_t(n) * n * L);
        w.plan.window_psfs(s, t, w.amp.as<float>(), phase, L, w.raw.as<float>(), w.psf, nullptr, dcoefs);
        base_kernels(s, w, w.raw.as<float>(), R, w.scales.as<float>(), nullptr, r, L, w.base.as<float>(),
                     w.base_totals.as<float>(), h);
    } else {
        w.raw.reset(sizeof(float) * size_t(R) * R * L);
        w.totals.reset(sizeof(float) * L);
        w.plan.raw_psfs(s, t, w.amp.as<float>(), phase, L, w.raw.as<float>(), w.totals.as<float>(), w.seqsum, w.psf,
                        false, nullptr, dcoefs);
        base_kernels(s, w, w.raw.as<float>(), R, w.scales.as<float>(), w.totals.as<float>(), r, L, w.base.as<float>(),
                     w.base_totals.as<float>());
    }

    // shaping into full / core / structure, the first `sets` (the base is read normalised through
    // its totals)
    bank->sets = sets;
    bank->kernels.reset(sizeof(float) * kk * L * sets);
    bank->store = bank->kernels.as<float>();
    bank->which_stride = kk * L;
    w.radius.reset(sizeof(int) * L);
    bank->own_scales.reset(sizeof(float) * sets * L);
    bank->scales = bank->own_scales.as<float>();
    bank->scales_which_stride = L;
    shape_kernels(s, w, w.base.as<float>(), w.base_totals.as<float>(), K, L, p.anisotropy_emphasis, r,
                  bank->kernels.as<float>(), w.radius.as<int>(), bank->scales, nullptr, sets);
    bank->radius.resize(size_t(L));
    bank->host_scales.resize(size_t(sets) * L);
    LDH_HIP(hipMemcpyAsync(bank->radius.data(), w.radius.get(), sizeof(int) * L, hipMemcpyDeviceToHost, s), "bank-radius");
    LDH_HIP(hipMemcpyAsync(bank->host_scales.data(), bank->scales, sizeof(float) * sets * L, hipMemcpyDeviceToHost, s),
            "bank-scales");
    LDH_HIP(hipStreamSynchronize(s), "bank-radius");
    bank->support = 4;
    for (int v : bank->radius) bank->support = std::max(bank->support, v);
    return bank;
}

}  // namespace ldh

// Bank construction on the device: resampling (makeResampledKernel / sampleSquareFiltered,
// LensDiffCpuReference.cpp:316-353, :1060-1086) and shaping (azimuthalMeanKernel,
// reshapeAnisotropy, positiveResidualKernel, estimateAdaptiveSupportRadius,
// paddedAdaptiveSupportRadius, cropKernelToRadius, applySupportBoundaryTaper, :957-1151),
// all batched over the wavelengths of one bank.
//
// Ring statistics (azimuthal sums/counts, support-radius energies/peaks) use per-workgroup LDS
// histograms flushed once per workgroup, with FP64 sums. Every unit-sum normalisation goes
// through SeqSum so the scale matches the reference's sequential FP32 accumulation.
#include <algorithm>
#include <cmath>

#include "bank.h"
#include "host_params.h"

namespace ldh {

namespace {

// Bilinear sample of img * norm: each texel is scaled before interpolation, as when the image
// was normalised in place first.
__device__ __forceinline__ float bilinear_sq(const float* __restrict__ img, int size, float x, float y, float norm) {
    if (x < 0.0f || y < 0.0f || x > float(size - 1) || y > float(size - 1)) return 0.0f;
    const int x0 = int(floorf(x)), y0 = int(floorf(y));
    const int x1 = min(x0 + 1, size - 1), y1 = min(y0 + 1, size - 1);
    const float tx = x - float(x0), ty = y - float(y0);
    const float v00 = img[size_t(y0) * size + x0] * norm, v10 = img[size_t(y0) * size + x1] * norm;
    const float v01 = img[size_t(y1) * size + x0] * norm, v11 = img[size_t(y1) * size + x1] * norm;
    const float a = v00 + (v10 - v00) * tx;
    const float b = v01 + (v11 - v01) * tx;
    return a + (b - a) * ty;
}

__global__ void resample_kernel(const float* __restrict__ raw, int R, const float* __restrict__ scales,
                                const float* __restrict__ raw_totals, int support, float* __restrict__ out) {
    const int lam = blockIdx.y;
    // raw_totals: the raw PSFs are not yet normalised; scale texels by 1/total (total > 0)
    float norm = 1.0f;
    if (raw_totals != nullptr && raw_totals[lam] > 0.0f) norm = 1.0f / raw_totals[lam];
    const int K = 2 * support + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * K) return;
    const int x = i % K, y = i / K;
    const float* img = raw + size_t(lam) * R * R;
    const float rc = (R & 1) == 0 ? float(R) * 0.5f : float(R - 1) * 0.5f;
    const float kc = float(K - 1) * 0.5f;
    const float inv = 1.0f / fmaxf(scales[lam], 0.05f);
    const float sx = rc + (float(x) - kc) * inv;
    const float sy = rc + (float(y) - kc) * inv;
    float v;
    if (inv <= 1.0f) {
        v = bilinear_sq(img, R, sx, sy, norm);
    } else {
        const int taps = min(max(int(ceilf(inv)), 2), 6);
        const float step = inv / float(taps);
        float sum = 0.0f;
        for (int ty = 0; ty < taps; ++ty) {
            const float oy = (float(ty) + 0.5f) * step - inv * 0.5f;
            for (int tx = 0; tx < taps; ++tx) {
                const float ox = (float(tx) + 0.5f) * step - inv * 0.5f;
                sum += bilinear_sq(img, R, sx + ox, sy + oy, norm);
            }
        }
        v = sum / float(taps * taps);
    }
    out[size_t(lam) * K * K + i] = v;
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

constexpr int kRingThreads = 256;
constexpr int kRingPixelsPerBlock = 16384;

// sums[r] += v, counts[r] += 1, peaks[r] = max(v) over the pixels of ring r, per segment.
template <bool CEIL>
__global__ void __launch_bounds__(kRingThreads)
ring_stats_kernel(const float* __restrict__ img, int size, float center, int rmax, double* __restrict__ sums,
                  int* __restrict__ counts, unsigned* __restrict__ peaks) {
    extern __shared__ unsigned char smem[];
    const int nr = rmax + 1;
    double* s_sum = reinterpret_cast<double*>(smem);
    int* s_cnt = reinterpret_cast<int*>(s_sum + nr);
    unsigned* s_pk = reinterpret_cast<unsigned*>(s_cnt + nr);
    for (int r = threadIdx.x; r < nr; r += kRingThreads) { s_sum[r] = 0.0; s_cnt[r] = 0; s_pk[r] = 0u; }
    __syncthreads();
    const int seg = blockIdx.y;
    const int64_t npx = int64_t(size) * size;
    const float* base = img + seg * npx;
    const int64_t p0 = int64_t(blockIdx.x) * kRingPixelsPerBlock;
    const int64_t p1 = min(npx, p0 + kRingPixelsPerBlock);
    for (int64_t p = p0 + threadIdx.x; p < p1; p += kRingThreads) {
        const int x = int(p % size), y = int(p / size);
        const int r = ring_of<CEIL>(x, y, center, rmax);
        const float v = base[p];
        atomicAdd(&s_sum[r], double(v));
        if (counts) atomicAdd(&s_cnt[r], 1);
        if (peaks) atomicMax(&s_pk[r], __float_as_uint(v));  // v >= 0: bit order is value order
    }
    __syncthreads();
    for (int r = threadIdx.x; r < nr; r += kRingThreads) {
        if (s_cnt[r] == 0 && s_sum[r] == 0.0 && s_pk[r] == 0u) continue;
        atomicAdd(&sums[int64_t(seg) * nr + r], s_sum[r]);
        if (counts) atomicAdd(&counts[int64_t(seg) * nr + r], s_cnt[r]);
        if (peaks) atomicMax(&peaks[int64_t(seg) * nr + r], s_pk[r]);
    }
}

__global__ void ring_mean_kernel(const double* __restrict__ sums, const int* __restrict__ counts, int K, int rmax,
                                 float* __restrict__ out) {
    const int seg = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * K) return;
    const float c = float(K - 1) * 0.5f;
    const int r = ring_of<false>(i % K, i / K, c, rmax);
    const int nr = rmax + 1;
    const float s = float(sums[int64_t(seg) * nr + r]);
    const int n = max(1, counts[int64_t(seg) * nr + r]);
    out[int64_t(seg) * K * K + i] = s / float(n);
}

__global__ void reshape_residual_kernel(const float* __restrict__ base, const float* __restrict__ mean, float gain,
                                        int64_t n, float* __restrict__ shaped) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float m = mean[i];
    shaped[i] = fmaxf(0.0f, m + (base[i] - m) * gain);
}

__global__ void residual_kernel(const float* __restrict__ shaped, const float* __restrict__ mean, int64_t n,
                                float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = fmaxf(0.0f, shaped[i] - mean[i]);
}

// estimateAdaptiveSupportRadius + paddedAdaptiveSupportRadius (LensDiffCpuReference.cpp:984-1036),
// one thread per bin. `peaks` is turned in place into the outside-peak suffix maxima.
__global__ void support_radius_kernel(const double* __restrict__ energy, unsigned* __restrict__ peaks,
                                      const float* __restrict__ totals, int rmax, int max_radius, int count,
                                      int* __restrict__ radius) {
    const int seg = blockIdx.x * blockDim.x + threadIdx.x;
    if (seg >= count) return;
    const int nr = rmax + 1;
    const double* e = energy + int64_t(seg) * nr;
    unsigned* pk = peaks + int64_t(seg) * nr;
    for (int r = rmax - 1; r >= 0; --r) pk[r] = max(pk[r], pk[r + 1]);  // outsidePeak[r]
    const float total = totals[seg];
    const float gpeak = __uint_as_float(pk[0]);
    int est = max(4, rmax);
    if (total > 1e-6f && gpeak > 1e-6f) {
        float cum = 0.0f;
        for (int r = 0; r <= rmax; ++r) {
            cum += float(e[r]);
            const float captured = cum / total;
            const float remaining = r + 1 <= rmax ? __uint_as_float(pk[r + 1]) : 0.0f;
            if (captured >= 0.9999f && remaining <= gpeak * 2e-4f) {
                est = max(4, r + 1);
                break;
            }
        }
    }
    const int pad = max(6, min(24, int(ceilf(float(max(1, est)) * 0.05f))));
    const int hi = max(4, max_radius);
    radius[seg] = min(max(est + pad, 4), hi);
}

// cropKernelToRadius + applySupportBoundaryTaper on the K x K grid: keeps the centred
// (2r+1)^2 square, multiplies by the cosine taper, zeroes everything else.
__global__ void crop_taper_kernel(const float* __restrict__ src, const int* __restrict__ radius, int K, int count,
                                  float* __restrict__ dst) {
    const int seg = blockIdx.y;  // which * count + bin
    const int bin = seg % count;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * K) return;
    const int r = min(max(radius[bin], 1), K / 2);
    const int c = K / 2;
    const int x = i % K, y = i / K;
    const int64_t off = int64_t(seg) * K * K + i;
    if (abs(x - c) > r || abs(y - c) > r) {
        dst[off] = 0.0f;
        return;
    }
    const float dx = float(x - c), dy = float(y - c);
    const float rr = sqrtf(dx * dx + dy * dy);
    const float extent = fmaxf(1.0f, float(r));
    const float fade = fmaxf(6.0f, fminf(24.0f, extent * 0.04f));
    const float start = fmaxf(0.0f, extent - fade);
    float w;
    if (rr >= extent) {
        w = 0.0f;
    } else {
        const float t = fminf(fmaxf((rr - start) / fmaxf(fade, 1e-6f), 0.0f), 1.0f);
        const float sm = t * t * (3.0f - 2.0f * t);
        w = fmaxf(0.0f, cosf(sm * 1.57079637f));
    }
    dst[off] = src[off] * w;
}

inline dim3 grid1(int64_t n, int segs = 1) { return dim3(unsigned(div_up(n, 256)), unsigned(segs)); }

template <bool CEIL>
void ring_stats(hipStream_t s, const float* img, int size, float center, int rmax, int segs, double* sums,
                int* counts, unsigned* peaks) {
    const int nr = rmax + 1;
    LDH_HIP(hipMemsetAsync(sums, 0, sizeof(double) * nr * segs, s), "ring-stats");
    if (counts) LDH_HIP(hipMemsetAsync(counts, 0, sizeof(int) * nr * segs, s), "ring-stats");
    if (peaks) LDH_HIP(hipMemsetAsync(peaks, 0, sizeof(unsigned) * nr * segs, s), "ring-stats");
    const size_t lds = size_t(nr) * (sizeof(double) + sizeof(int) + sizeof(unsigned));
    const int64_t npx = int64_t(size) * size;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&ring_stats_kernel<CEIL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr_set = true;
    }
    hipLaunchKernelGGL((ring_stats_kernel<CEIL>), dim3(unsigned(div_up(npx, kRingPixelsPerBlock)), unsigned(segs)),
                       dim3(kRingThreads), lds, s, img, size, center, rmax, sums, counts, peaks);
    LDH_CHECK_LAUNCH("ring-stats");
}

}  // namespace

void resample_kernels(hipStream_t s, const float* raw, int R, const float* d_scales, const float* d_raw_totals,
                      int support, int count, float* out) {
    const int K = 2 * support + 1;
    hipLaunchKernelGGL(resample_kernel, grid1(int64_t(K) * K, count), dim3(256), 0, s, raw, R, d_scales, d_raw_totals,
                       support, out);
    LDH_CHECK_LAUNCH("resample");
}

void shape_kernels(hipStream_t s, BankWorkspace& w, const float* base, int K, int count, double anisotropy,
                   int max_radius, float* out3, int* d_radius) {
    const int64_t kk = int64_t(K) * K;
    const int64_t n = kk * count;
    const int rmax_mean = K / 2;
    w.mean.reset(sizeof(float) * n);
    w.shaped.reset(sizeof(float) * n);
    w.structure.reset(sizeof(float) * n);
    w.totals.reset(sizeof(float) * 3 * count);
    const int nr = std::max(rmax_mean, max_radius) + 1;
    w.ring_sum.reset(sizeof(double) * nr * count);
    w.ring_cnt.reset(sizeof(int) * nr * count);
    w.ring_peak.reset(sizeof(unsigned) * nr * count);
    float* tot = w.totals.as<float>();
    // azimuthal mean, normalised
    ring_stats<false>(s, base, K, float(K - 1) * 0.5f, rmax_mean, count, w.ring_sum.as<double>(),
                      w.ring_cnt.as<int>(), nullptr);
    hipLaunchKernelGGL(ring_mean_kernel, grid1(kk, count), dim3(256), 0, s, w.ring_sum.as<double>(),
                       w.ring_cnt.as<int>(), K, rmax_mean, w.mean.as<float>());
    LDH_CHECK_LAUNCH("ring-mean");
    w.seqsum.normalize(s, w.mean.as<float>(), kk, kk, count, tot);
    // anisotropy reshape, normalised
    const float gain = 1.0f + std::max(0.0f, float(anisotropy)) * 4.0f;
    hipLaunchKernelGGL(reshape_residual_kernel, grid1(n), dim3(256), 0, s, base, w.mean.as<float>(), gain, n,
                       w.shaped.as<float>());
    LDH_CHECK_LAUNCH("reshape");
    w.seqsum.normalize(s, w.shaped.as<float>(), kk, kk, count, tot);
    // positive residual, normalised when its sum is positive
    hipLaunchKernelGGL(residual_kernel, grid1(n), dim3(256), 0, s, w.shaped.as<float>(), w.mean.as<float>(), n,
                       w.structure.as<float>());
    LDH_CHECK_LAUNCH("residual");
    w.seqsum.normalize(s, w.structure.as<float>(), kk, kk, count, tot);
    // adaptive support radius of the shaped kernel
    const int rmax_sup = std::min(max_radius, K / 2);
    ring_stats<true>(s, w.shaped.as<float>(), K, float(K - 1) * 0.5f, rmax_sup, count, w.ring_sum.as<double>(),
                     nullptr, w.ring_peak.as<unsigned>());
    w.seqsum.run(s, w.shaped.as<float>(), kk, kk, count, tot);
    hipLaunchKernelGGL(support_radius_kernel, dim3(unsigned(div_up(count, 64))), dim3(64), 0, s,
                       w.ring_sum.as<double>(), w.ring_peak.as<unsigned>(), tot, rmax_sup, max_radius, count, d_radius);
    LDH_CHECK_LAUNCH("support-radius");
    // crop + taper of full (shaped), core (mean), structure, then normalise all three
    float* full = out3;
    float* core = out3 + n;
    float* structure = out3 + 2 * n;
    hipLaunchKernelGGL(crop_taper_kernel, grid1(kk, count), dim3(256), 0, s, w.shaped.as<float>(), d_radius, K, count,
                       full);
    hipLaunchKernelGGL(crop_taper_kernel, grid1(kk, count), dim3(256), 0, s, w.mean.as<float>(), d_radius, K, count,
                       core);
    hipLaunchKernelGGL(crop_taper_kernel, grid1(kk, count), dim3(256), 0, s, w.structure.as<float>(), d_radius, K,
                       count, structure);
    LDH_CHECK_LAUNCH("crop-taper");
    w.totals.reset(sizeof(float) * 3 * count);
    w.seqsum.normalize(s, out3, kk, kk, 3 * count, w.totals.as<float>());
}

std::vector<float> radial_profile(hipStream_t s, BankWorkspace& w, const float* psf, int R) {
    const int rmax = R / 2;
    const int nr = rmax + 1;
    w.ring_sum.reset(sizeof(double) * nr);
    w.ring_cnt.reset(sizeof(int) * nr);
    const float center = (R & 1) == 0 ? float(R) * 0.5f : float(R - 1) * 0.5f;
    ring_stats<false>(s, psf, R, center, rmax, 1, w.ring_sum.as<double>(), w.ring_cnt.as<int>(), nullptr);
    std::vector<double> sums(static_cast<size_t>(nr));
    std::vector<int> counts(static_cast<size_t>(nr));
    LDH_HIP(hipMemcpyAsync(sums.data(), w.ring_sum.get(), sizeof(double) * nr, hipMemcpyDeviceToHost, s), "profile");
    LDH_HIP(hipMemcpyAsync(counts.data(), w.ring_cnt.get(), sizeof(int) * nr, hipMemcpyDeviceToHost, s), "profile");
    LDH_HIP(hipStreamSynchronize(s), "profile");
    std::vector<float> prof(static_cast<size_t>(nr));
    for (int r = 0; r < nr; ++r) prof[size_t(r)] = float(sums[size_t(r)]) / float(std::max(1, counts[size_t(r)]));
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

static float clear_circle_first_zero(hipStream_t s, const FftTables& t, BankWorkspace& w, int P, int R) {
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
    pupil_amplitude(s, make_pupil_geom(q), w.ref_amp.as<float>());
    LDH_HIP(hipMemsetAsync(w.ref_phase.get(), 0, sizeof(float) * P * P, s), "first-zero");
    w.plan.prepare(s, P, R);
    w.plan.raw_psfs(s, t, w.ref_amp.as<float>(), w.ref_phase.as<float>(), 1, w.ref_raw.as<float>(),
                    w.ref_total.as<float>(), w.seqsum);
    const float fz = first_minimum_from_profile(radial_profile(s, w, w.ref_raw.as<float>(), R), R);
    w.first_zero.push_back({{P, R}, fz});
    return fz;
}

std::unique_ptr<DeviceBank> build_bank(hipStream_t s, const FftTables& t, BankWorkspace& w, const ldh_params& p) {
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

    const float first_zero = std::max(1.0f, clear_circle_first_zero(s, t, w, P, R));
    const float scale_base = float(std::max(1.0, sz.diffraction_scale_px)) / first_zero;

    // amplitude once, phase per wavelength (chromatic defocus offset, :1720-1727)
    w.amp.reset(sizeof(float) * P * P);
    pupil_amplitude(s, make_pupil_geom(p), w.amp.as<float>());
    std::vector<PhaseCoef> coefs(static_cast<size_t>(L));
    std::vector<float> scales(static_cast<size_t>(L));
    for (int i = 0; i < L; ++i) {
        ldh_params q = p;
        const float off = (wl[size_t(i)] - 550.0f) / 110.0f;
        q.phase_defocus += p.chromatic_focus * off;
        if (std::abs(q.phase_defocus) > 1e-6) q.phase_mode = LDH_PHASE_ENABLED;
        coefs[size_t(i)] = make_phase_coef(q);
        const float chroma = std::max(0.25f, 1.0f + float(p.chromatic_spread) * off);
        scales[size_t(i)] = scale_base * (wl[size_t(i)] / 550.0f) * chroma;
    }
    w.coefs.reset(sizeof(PhaseCoef) * L);
    w.scales.reset(sizeof(float) * L);
    LDH_HIP(hipMemcpyAsync(w.coefs.get(), coefs.data(), sizeof(PhaseCoef) * L, hipMemcpyHostToDevice, s), "bank-upload");
    LDH_HIP(hipMemcpyAsync(w.scales.get(), scales.data(), sizeof(float) * L, hipMemcpyHostToDevice, s), "bank-upload");
    w.phase.reset(sizeof(float) * size_t(P) * P * L);
    pupil_phase(s, w.coefs.as<PhaseCoef>(), P, L, w.phase.as<float>());

    // raw PSFs, then per-wavelength resampled base kernels
    w.plan.prepare(s, P, R);
    w.raw.reset(sizeof(float) * size_t(R) * R * L);
    w.totals.reset(sizeof(float) * L);
    // the unit-sum scale of each raw PSF is applied inside the resample
    w.plan.raw_psfs(s, t, w.amp.as<float>(), w.phase.as<float>(), L, w.raw.as<float>(), w.totals.as<float>(), w.seqsum,
                    false);
    const int64_t kk = int64_t(K) * K;
    bank->base.reset(sizeof(float) * kk * L);
    resample_kernels(s, w.raw.as<float>(), R, w.scales.as<float>(), w.totals.as<float>(), r, L, bank->base.as<float>());
    w.base_totals.reset(sizeof(float) * L);
    w.seqsum.normalize(s, bank->base.as<float>(), kk, kk, L, w.base_totals.as<float>());

    // shaping into full / core / structure
    bank->kernels.reset(sizeof(float) * kk * L * 3);
    w.radius.reset(sizeof(int) * L);
    shape_kernels(s, w, bank->base.as<float>(), K, L, p.anisotropy_emphasis, r, bank->kernels.as<float>(),
                  w.radius.as<int>());
    bank->radius.resize(size_t(L));
    LDH_HIP(hipMemcpyAsync(bank->radius.data(), w.radius.get(), sizeof(int) * L, hipMemcpyDeviceToHost, s), "bank-radius");
    LDH_HIP(hipStreamSynchronize(s), "bank-radius");
    bank->support = 4;
    for (int v : bank->radius) bank->support = std::max(bank->support, v);
    return bank;
}

}  // namespace ldh

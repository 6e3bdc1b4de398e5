// Exact ring sums (see rings.h). A float running sum that stays inside one binade [2^e, 2^(e+1))
// adds each addend v as a multiple of u = 2^(e-23): fl(s + v) = s + rint(v / u) * u, unless v / u
// is exactly halfway between two integers (the tie then rounds to the even result, which depends
// on s) or the sum leaves the binade. A wavefront therefore adds 512 elements of a ring at once as
// an integer sum of rint(v / u) whenever none is a tie and the total stays below 2^(e+1), else
// each 64 of them at once under the same conditions, else those 64 one after the other in float,
// in order, on every lane. Either way the result is the reference's sequential float sum, bit for
// bit.
#include <algorithm>
#include <cmath>

#include "profiler.h"
#include "rings.h"

namespace ldh {

namespace {

constexpr int kRingWaves = 4;     // wavefronts (rings) per workgroup
constexpr int kRingGroups = 8;    // 64-element groups whose loads are issued together

__device__ __forceinline__ float inverse_total(const float* __restrict__ totals, int seg) {
    if (totals == nullptr) return 1.0f;
    const float t = totals[seg];
    return t > 0.0f ? 1.0f / t : 1.0f;
}

// Sum over the 64 lanes, returned in every lane: DPP row shifts within rows of 16, then the rows'
// totals broadcast forward (row_bcast:15 / :31), read back from lane 63.
__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
    int x = int(v);
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return unsigned(__builtin_amdgcn_readlane(x, 63));
}

__device__ __forceinline__ float lane_value(float v, int j) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), j));
}

// s (the same in every lane) followed by the 64 addends v of lanes 0..63, in lane order, with a
// float accumulator. Addends are >= 0 and finite.
__device__ __forceinline__ float add_group_sequential(float s, float v) {
    if (s >= 0x1p-100f) {
        int ex;
        (void)frexpf(s, &ex);              // s in [2^(ex-1), 2^ex): u = 2^(ex-24)
        const int sh = 24 - ex;
        const float q = ldexpf(v, sh);     // v / u, exact
        const bool bad = !(q < 0x1p24f) || (q - floorf(q) == 0.5f);
        const unsigned c = bad ? 0u : unsigned(rintf(q));
        const unsigned long long any_bad = __ballot(bad);
        const unsigned tot = wave_sum_u32(c);
        const unsigned ms = unsigned(ldexpf(s, sh));  // s / u in [2^23, 2^24)
        if (any_bad == 0ull && ms + tot < (1u << 24)) return ldexpf(float(ms + tot), -sh);
    }
#pragma unroll
    for (int j = 0; j < 64; ++j) s = s + lane_value(v, j);
    return s;
}

// The kRingGroups groups v (64 lanes each, in group then lane order) added to s at once: one
// integer sum of every rint(v / u) in s's binade when no element is a tie or too large and the
// total stays in the binade (the addends are >= 0, so every partial sum then does too, and each
// addition is exact as in add_group_sequential). Returns false, s unchanged, otherwise.
__device__ __forceinline__ bool add_batch(float& s, const float (&v)[kRingGroups]) {
    if (!(s >= 0x1p-100f)) return false;
    int ex;
    (void)frexpf(s, &ex);
    const int sh = 24 - ex;
    unsigned c = 0u;
    bool bad = false;
#pragma unroll
    for (int g = 0; g < kRingGroups; ++g) {
        const float q = ldexpf(v[g], sh);
        const bool b = !(q < 0x1p24f) || (q - floorf(q) == 0.5f);
        c += b ? 0u : unsigned(rintf(q));
        bad = bad || b;
    }
    // a lane's count below 2^24 keeps the wave's total below 2^30
    bad = bad || c >= (1u << 24);
    if (__ballot(bad) != 0ull) return false;
    const unsigned tot = wave_sum_u32(c);
    const unsigned ms = unsigned(ldexpf(s, sh));
    if (ms + tot >= (1u << 24)) return false;
    s = ldexpf(float(ms + tot), -sh);
    return true;
}

// One wavefront per (segment, ring r < rings), over the ring-ordered copy (ring r of segment g at
// ordered[g * pixels + off[r] ..]), reading one batch of groups ahead of the one it adds.
// Workgroups are spread over the eight XCDs in contiguous ranges of items (workgroup b runs on
// XCD b % 8), so the rings of one image share an L2.
__global__ void __launch_bounds__(64 * kRingWaves)
ring_wave_sums_kernel(const float* __restrict__ ordered, int64_t pixels, const int* __restrict__ off, int rings,
                      int items, int per_xcd, float* __restrict__ sums, unsigned* __restrict__ peaks, int nr) {
    const int b = int(blockIdx.x);
    const int logical = (b % 8) * per_xcd + b / 8;
    const int item = logical * kRingWaves + int(threadIdx.x >> 6);
    if (item >= items) return;
    const int lane = int(threadIdx.x) & 63;
    const int seg = item / rings, r = item % rings;
    const float* x = ordered + int64_t(seg) * pixels;
    const int i0 = off[r], i1 = off[r + 1];
    float s = 0.0f;
    unsigned pk = 0u;
    constexpr int kStep = 64 * kRingGroups;
    float cur[kRingGroups];
    auto load = [&](float (&v)[kRingGroups], int base) {
#pragma unroll
        for (int g = 0; g < kRingGroups; ++g) {
            const int i = base + 64 * g + lane;
            const float t = x[min(i, i1 - 1)];
            v[g] = i < i1 ? t : 0.0f;
        }
    };
    if (i0 < i1) load(cur, i0);
    for (int base = i0; base < i1; base += kStep) {
        float nxt[kRingGroups];
        if (base + kStep < i1) load(nxt, base + kStep);
        // groups past the ring's end read 0: they change neither the peak nor a batch's sum
#pragma unroll
        for (int g = 0; g < kRingGroups; ++g) pk = max(pk, __float_as_uint(cur[g]));
        if (!add_batch(s, cur)) {
#pragma unroll
            for (int g = 0; g < kRingGroups; ++g) {
                if (base + 64 * g >= i1) break;
                s = add_group_sequential(s, cur[g]);
            }
        }
#pragma unroll
        for (int g = 0; g < kRingGroups; ++g) cur[g] = nxt[g];
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) pk = max(pk, unsigned(__shfl_xor(int(pk), d)));
    if (lane == 0) {
        sums[int64_t(seg) * nr + r] = s;
        if (peaks) peaks[int64_t(seg) * nr + r] = pk;
    }
}

// The ring-ordered copy of every segment, dst[g * pixels + i] = img[g][order[i]] * (1 / total),
// and the last ring's peak (i >= last) by atomic maximum (bits of values >= 0 order as the
// values). A workgroup takes one band of kRingBand raster pixels and kBandSegs segments: each
// segment's band is read coalesced into LDS, and its pixels are stored to their ring-ordered
// positions in ascending order (the band's pixels of one ring are consecutive there), so a wave's
// stores fall in a few contiguous runs. The band's position lists are read once for all its
// segments; the next segment's band is loaded while the current one is stored.
constexpr int kBandThreads = 256;
constexpr int kBandPer = kRingBand / kBandThreads;
constexpr int kBandSegs = 8;
__global__ void __launch_bounds__(kBandThreads)
ring_band_gather_kernel(const float* __restrict__ img, int64_t seg_stride, const float* __restrict__ prescale,
                        const unsigned short* __restrict__ band_local, const int* __restrict__ band_out,
                        int64_t pixels, int64_t last, int bands, int segs, int per_xcd, float* __restrict__ dst,
                        unsigned* __restrict__ peak_slot, int nr) {
    const int b = int(blockIdx.x);
    const int logical = (b % 8) * per_xcd + b / 8;  // band-major (band, segment group)
    const int groups = (segs + kBandSegs - 1) / kBandSegs;
    if (logical >= bands * groups) return;
    const int band = logical / groups, g0 = (logical % groups) * kBandSegs;
    const int g1 = min(segs, g0 + kBandSegs);
    const int64_t p0 = int64_t(band) * kRingBand;
    const int len = int(min(int64_t(kRingBand), pixels - p0));
    const int t = int(threadIdx.x);
    __shared__ float tile[kRingBand];
    int loc[kBandPer], out[kBandPer];
#pragma unroll
    for (int k = 0; k < kBandPer; ++k) {
        const int e = t + k * kBandThreads;
        const int c = min(e, len - 1);
        loc[k] = int(band_local[p0 + c]);
        out[k] = e < len ? band_out[p0 + c] : -1;
    }
    float cur[kBandPer];
    auto load = [&](int g, float (&v)[kBandPer]) {
        const float* x = img + int64_t(g) * seg_stride + p0;
#pragma unroll
        for (int k = 0; k < kBandPer; ++k) v[k] = x[min(t + k * kBandThreads, len - 1)];
    };
    load(g0, cur);
    for (int g = g0; g < g1; ++g) {
        __syncthreads();  // the previous segment's band has been read
#pragma unroll
        for (int k = 0; k < kBandPer; ++k) tile[t + k * kBandThreads] = cur[k];
        __syncthreads();
        if (g + 1 < g1) load(g + 1, cur);
        const float inv = inverse_total(prescale, g);
        float* o = dst + int64_t(g) * pixels;
        unsigned pk = 0u;
#pragma unroll
        for (int k = 0; k < kBandPer; ++k) {
            if (out[k] >= 0) {
                const float v = tile[loc[k]] * inv;
                o[out[k]] = v;
                if (out[k] >= last) pk = max(pk, __float_as_uint(v));
            }
        }
        if (peak_slot != nullptr) {
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) pk = max(pk, unsigned(__shfl_xor(int(pk), d)));
            if ((t & 63) == 0 && pk != 0u) atomicMax(&peak_slot[int64_t(g) * nr], pk);
        }
    }
}

// Clears the last ring's peak slot of every segment (before the gather's atomic maxima).
__global__ void ring_corner_clear_kernel(unsigned* __restrict__ peak_slot, int nr, int segs) {
    const int g = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (g < segs) peak_slot[int64_t(g) * nr] = 0u;
}

// sums[g * nr + rmax] = the last ring's sequential sum (0 for an empty ring).
__global__ void ring_corner_store_kernel(const float* __restrict__ totals, float* __restrict__ sum_slot, int nr,
                                         int segs) {
    const int g = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (g < segs) sum_slot[int64_t(g) * nr] = totals != nullptr ? totals[g] : 0.0f;
}

// The ring of pixel (x, y) in float arithmetic, as the reference evaluates it: dx, dy exact,
// sqrt correctly rounded, then round (half away from zero) or ceil, clamped to rmax.
inline int host_ring(int x, int y, float c, int rmax, bool ceil) {
    const float dx = float(x) - c, dy = float(y) - c;
    const float d = std::sqrt(dx * dx + dy * dy);
    const int r = ceil ? int(std::ceil(d)) : int(std::round(d));
    return std::min(rmax, r);
}

}  // namespace

void prepare_ring_plan(hipStream_t s, RingPlan& p, int size, float center, int rmax, bool ceil) {
    if (p.matches(size, center, rmax, ceil)) return;
    if (size <= 0 || rmax < 0) fail(LDH_ERR_INVALID_ARGUMENT, "ring-plan-invalid");
    p.size = 0;
    p.rmax = -1;
    const int nr = rmax + 1;
    const int64_t n = int64_t(size) * size;
    std::vector<int> off(static_cast<size_t>(nr) + 1, 0);
    for (int y = 0; y < size; ++y)
        for (int x = 0; x < size; ++x) ++off[size_t(host_ring(x, y, center, rmax, ceil)) + 1];
    std::vector<int> count(static_cast<size_t>(nr));
    for (int r = 0; r < nr; ++r) {
        count[size_t(r)] = off[size_t(r) + 1];
        off[size_t(r) + 1] += off[size_t(r)];
    }
    std::vector<int> order(static_cast<size_t>(n));
    std::vector<int> next(off.begin(), off.end() - 1);
    for (int y = 0; y < size; ++y)
        for (int x = 0; x < size; ++x)
            order[size_t(next[size_t(host_ring(x, y, center, rmax, ceil))]++)] = y * size + x;
    // per band of kRingBand raster pixels, its pixels in ring order: band-local index and
    // ring-ordered position (walking the ring order appends to each band in ascending position)
    const int64_t bands = div_up(n, int64_t(kRingBand));
    std::vector<unsigned short> band_local(static_cast<size_t>(n));
    std::vector<int> band_out(static_cast<size_t>(n));
    std::vector<int> fill(static_cast<size_t>(bands), 0);
    for (int64_t i = 0; i < n; ++i) {
        const int px = order[size_t(i)];
        const int band = px / kRingBand;
        const size_t k = size_t(band) * kRingBand + size_t(fill[size_t(band)]++);
        band_local[k] = static_cast<unsigned short>(px - band * kRingBand);
        band_out[k] = int(i);
    }
    p.band_local.reset(sizeof(unsigned short) * size_t(n));
    p.band_out.reset(sizeof(int) * size_t(n));
    p.off.reset(sizeof(int) * (size_t(nr) + 1));
    p.count.reset(sizeof(int) * size_t(nr));
    LDH_HIP(hipMemcpyAsync(p.band_local.get(), band_local.data(), sizeof(unsigned short) * size_t(n),
                           hipMemcpyHostToDevice, s),
            "ring-plan");
    LDH_HIP(hipMemcpyAsync(p.band_out.get(), band_out.data(), sizeof(int) * size_t(n), hipMemcpyHostToDevice, s),
            "ring-plan");
    LDH_HIP(hipMemcpyAsync(p.off.get(), off.data(), sizeof(int) * (size_t(nr) + 1), hipMemcpyHostToDevice, s),
            "ring-plan");
    LDH_HIP(hipMemcpyAsync(p.count.get(), count.data(), sizeof(int) * size_t(nr), hipMemcpyHostToDevice, s),
            "ring-plan");
    LDH_HIP(hipStreamSynchronize(s), "ring-plan");  // the host vectors go out of scope
    p.host_off = std::move(off);
    p.size = size;
    p.center = center;
    p.rmax = rmax;
    p.ceil = ceil;
    p.pixels = n;
}

void RingScratch::reserve(const RingPlan& plan, int segs) {
    ordered.reset(sizeof(float) * size_t(plan.pixels) * segs);
    last_total.reset(sizeof(float) * size_t(segs));
}

void ring_sums(hipStream_t s, const RingPlan& plan, RingScratch& scratch, SeqSum& seqsum, const float* img,
               int64_t seg_stride, int segs, const float* prescale, float* sums, unsigned* peaks, bool last_sum) {
    const int rmax = plan.rmax, nr = rmax + 1;
    const int64_t n = plan.pixels, last = plan.corner_begin(), nlast = plan.corner_pixels();
    scratch.reserve(plan, segs);
    float* ordered = scratch.ordered.as<float>();
    const unsigned small = unsigned(div_up(segs, 256));
    {
        const int bands = int(div_up(n, int64_t(kRingBand)));
        const int groups = int(div_up(segs, kBandSegs));
        // each segment read and written once, the band lists (6 bytes a pixel) once per group
        ProfileScope ps("ring-gather", 8.0 * double(n) * segs + 6.0 * double(n) * groups);
        if (peaks) hipLaunchKernelGGL(ring_corner_clear_kernel, dim3(small), dim3(256), 0, s, peaks + rmax, nr, segs);
        const int per_xcd = int(div_up(int64_t(bands) * groups, 8));
        hipLaunchKernelGGL(ring_band_gather_kernel, dim3(unsigned(8 * per_xcd)), dim3(kBandThreads), 0, s, img,
                           seg_stride, prescale, plan.band_local.as<unsigned short>(), plan.band_out.as<int>(), n, last,
                           bands, segs, per_xcd, ordered, peaks ? peaks + rmax : nullptr, nr);
        LDH_CHECK_LAUNCH("ring-gather");
    }
    if (rmax > 0) {
        ProfileScope ps("ring-sums", 4.0 * double(last) * segs);
        const int items = rmax * segs;
        const int per_xcd = int(div_up(div_up(items, kRingWaves), 8));
        hipLaunchKernelGGL(ring_wave_sums_kernel, dim3(unsigned(8 * per_xcd)), dim3(64 * kRingWaves), 0, s, ordered, n,
                           plan.off.as<int>(), rmax, items, per_xcd, sums, peaks, nr);
        LDH_CHECK_LAUNCH("ring-sums");
    }
    const bool sum_last = last_sum && nlast > 0;
    if (sum_last) seqsum.run(s, ordered + last, nlast, n, segs, scratch.last_total.as<float>());
    hipLaunchKernelGGL(ring_corner_store_kernel, dim3(small), dim3(256), 0, s,
                       sum_last ? scratch.last_total.as<float>() : nullptr, sums + rmax, nr, segs);
    LDH_CHECK_LAUNCH("ring-last");
}

}  // namespace ldh

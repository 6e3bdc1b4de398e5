// Ring sums as the reference forms them: one float accumulator per ring, the pixels visited in
// raster order (azimuthalMeanKernel, LensDiffCpuReference.cpp:1094-1104; estimateAdaptiveSupport-
// Radius, :997-1005; radialMeanProfile, :869-878). The integer support radius compares the running
// float sum of these ring sums against thresholds, so the sums are reproduced bit for bit.
#pragma once

#include <vector>

#include "ldh_internal.h"
#include "seqsum.h"

namespace ldh {

// Raster pixels per band of the ring-order copy (64 KiB of LDS a workgroup). A band's pixels of
// one ring are stored as one run, so the run length grows with the band: at K = 2047 a band of
// 16384 covers 8 rows and a ring's run at its sides 16 pixels.
constexpr int kRingBand = 16384;

// Pixels of a size x size image grouped by ring about `center`: ring(x, y) = round(|d|)
// (ceil = false) or ceil(|d|) (ceil = true), clamped to rmax. The ring order lists the pixels of
// ring 0, then ring 1, ..., each ring's pixels in raster order, ring r at positions
// [off[r], off[r+1]); the last ring (rmax) holds every pixel the clamp sends there. On the device
// the order is kept per band of kRingBand raster pixels: band b's pixels, sorted by ring-order
// position, at [b * kRingBand, ...) of band_local (index within the band) and band_out (position).
struct RingPlan {
    int size = 0, rmax = -1;
    float center = 0.0f;
    bool ceil = false;
    int64_t pixels = 0;
    std::vector<int> host_off;  // rmax + 2 entries
    DeviceBuffer band_local;    // uint16[pixels]
    DeviceBuffer band_out;      // int32[pixels]
    DeviceBuffer off;           // int32[rmax + 2]: host_off on the device
    DeviceBuffer count;         // int32[rmax + 1]: ring sizes
    bool matches(int sz, float c, int rm, bool cl) const {
        return size == sz && center == c && rmax == rm && ceil == cl;
    }
    int64_t corner_begin() const { return host_off[size_t(rmax)]; }
    int64_t corner_pixels() const { return pixels - corner_begin(); }
};

// Builds (host counting sort, uploaded on stream s) unless the plan already matches.
void prepare_ring_plan(hipStream_t s, RingPlan& plan, int size, float center, int rmax, bool ceil);

// Scratch of ring_sums: the ring-ordered copy of the segments and the last ring's sums.
struct RingScratch {
    DeviceBuffer ordered, last_total;
    void reserve(const RingPlan& plan, int segs);
};

// For every segment g (image g at img + g * seg_stride, each pixel read as v * (1 / prescale[g])
// when prescale is given and positive, the values an in-place normalisation would leave):
//   sums[g * (rmax + 1) + r]  = the float sum of ring r's pixels added in raster order,
//   peaks[g * (rmax + 1) + r] = the bits of their maximum (nullable; values are >= 0).
// The segments are first copied in ring order (coalesced writes); rings below rmax then take one
// wavefront each over contiguous data, and the clamped last ring, which can hold a large part of
// the image, goes to the SeqSum engine. With last_sum false the last ring's sum is not formed
// (left 0) for callers whose result cannot depend on it; its peak always is. Pixel values must be
// >= 0 and finite.
void ring_sums(hipStream_t s, const RingPlan& plan, RingScratch& scratch, SeqSum& seqsum, const float* img,
               int64_t seg_stride, int segs, const float* prescale, float* sums, unsigned* peaks,
               bool last_sum = true);

}  // namespace ldh

// PSF bank: per-wavelength resampled kernels and their shaping (LensDiffCpuReference.cpp:
// :957-1151, :1616-1656, :1702-1734), batched over the bank's wavelengths.
#pragma once

#include <memory>
#include <vector>

#include "ldh_internal.h"
#include "psf.h"
#include "seqsum.h"

namespace ldh {

constexpr int kMaxKernelRadius = 2048;

// Device-resident bank. Kernels are stored K x K (K = 2*max_radius+1) with zeros outside each
// bin's cropped support, in one array [which][bin][K*K] (which: full, core, structure).
struct DeviceBank {
    int bins = 0;
    int pupil = 0, raw = 0;
    int max_radius = 0;   // r, K = 2r+1
    int kernel = 0;       // K
    int support = 4;      // LensDiffPsfBankCache::supportRadiusPx
    std::vector<float> wavelengths;
    std::vector<int> radius;       // cropped radius per bin
    DeviceBuffer kernels;          // float[3][bins][K*K]
    size_t bytes() const { return kernels.bytes(); }
    const float* kernel_ptr(int which, int bin) const {
        return kernels.as<float>() + (size_t(which) * bins + bin) * size_t(kernel) * kernel;
    }
};

// Scratch shared by bank builds on one stream.
struct BankWorkspace {
    PsfPlan plan;
    SeqSum seqsum;
    DeviceBuffer amp, phase, coefs, raw, totals, base, base_totals, shape_totals, scales, mean, shaped, structure, ring_sum, ring_cnt, ring_peak,
        radius, ref_amp, ref_phase, ref_raw, ref_total;
    // first-zero radius of the clear-circle PSF per (P, R) (GetLensDiffReferenceRawPsfCached)
    std::vector<std::pair<std::pair<int, int>, float>> first_zero;
};

float first_minimum_from_profile(const std::vector<float>& profile, int size);

// Steps of the build, exposed for the parity entry points.
// d_raw_totals (nullable): per-PSF sums of raw PSFs that are not yet normalised.
void resample_kernels(hipStream_t s, const float* raw, int R, const float* d_scales, const float* d_raw_totals,
                      int support, int count, float* out);
// Shapes `count` base kernels (count x K*K) into full/core/structure (3 x count x K*K) and
// radius[count] (device). `max_radius` as LensDiffCpuReference.cpp:1633. base_totals (nullable):
// per-kernel sums of a base that is not yet normalised (read as base / total).
void shape_kernels(hipStream_t s, BankWorkspace& w, const float* base, const float* base_totals, int K, int count,
                   double anisotropy, int max_radius, float* out3, int* d_radius);
// Radial mean profile of an R x R PSF about its centre (radialMeanProfile, :865-887), computed
// on the device and returned to the host.
std::vector<float> radial_profile(hipStream_t s, BankWorkspace& w, const float* psf, int R);

std::unique_ptr<DeviceBank> build_bank(hipStream_t s, const FftTables& t, BankWorkspace& w, const ldh_params& p);

}  // namespace ldh

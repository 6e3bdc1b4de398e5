// PSF bank: per-wavelength resampled kernels and their shaping (LensDiffCpuReference.cpp:
// :957-1151, :1616-1656, :1702-1734), batched over the bank's wavelengths.
#pragma once

#include <memory>
#include <vector>

#include "host_params.h"
#include "ldh_internal.h"
#include "psf.h"
#include "rings.h"
#include "seqsum.h"

namespace ldh {

constexpr int kMaxKernelRadius = 2048;

// Device-resident bank. Kernels are stored K x K (K = 2*max_radius+1) with zeros outside each
// bin's cropped support: kernel (which, bin) at store + which * which_stride + bin * K*K (which:
// full, core, structure). A bank built alone owns `kernels` ([which][bin][K*K]); the banks of a
// batched build share one store ([which][unit][K*K], each bank's bins consecutive units).
// The stored kernels are not yet unit-sum: kernel (which, bin) holds texels t whose normalised
// values are t * scale (scale = 1 / its sequential FP32 sum, or 1 for a non-positive sum, the
// values LensDiffCpuReference.cpp:440-449 leaves), applied by every reader.
struct DeviceBank {
    // kernel sets held: the first `sets` of (full, core, structure). A render that convolves only
    // the full kernels builds its bank with that set alone; the others are then not formed.
    int sets = 3;
    int bins = 0;
    int pupil = 0, raw = 0;
    int max_radius = 0;   // r, K = 2r+1
    int kernel = 0;       // K
    int support = 4;      // LensDiffPsfBankCache::supportRadiusPx
    int raw_window = 0;   // half-width of the synthesised raw-PSF block (0: whole R x R)
    std::vector<float> wavelengths;
    std::vector<int> radius;       // cropped radius per bin
    DeviceBuffer kernels;                   // own storage, float[3][bins][K*K]
    std::shared_ptr<DeviceBuffer> shared;   // or the batch's store
    float* store = nullptr;
    int64_t which_stride = 0;
    // unit-sum multipliers: scale (which, bin) at scales + which * scales_which_stride + bin
    DeviceBuffer own_scales;
    std::shared_ptr<DeviceBuffer> shared_scales;
    float* scales = nullptr;
    int64_t scales_which_stride = 0;
    std::vector<float> host_scales;  // [which][bin]
    size_t bytes() const { return size_t(sets) * bins * size_t(kernel) * kernel * sizeof(float); }
    void check_set(int which) const {
        if (which < 0 || which >= sets) fail(LDH_ERR_INVALID_ARGUMENT, "kernel-set-not-built");
    }
    const float* kernel_ptr(int which, int bin) const {
        check_set(which);
        return store + which * which_stride + int64_t(bin) * kernel * kernel;
    }
    const float* scale_ptr(int which, int bin) const {
        check_set(which);
        return scales + which * scales_which_stride + bin;
    }
    float host_scale(int which, int bin) const {
        check_set(which);
        return host_scales[size_t(which) * bins + size_t(bin)];
    }
};

// A kernel as its readers see it: texels `data`, each multiplied by *scale when scale is set.
struct KernelRef {
    const float* data = nullptr;
    const float* scale = nullptr;
};

// Scratch shared by bank builds on one stream.
struct BankWorkspace {
    PsfPlan plan;
    PsfScratch psf;
    SeqSum seqsum;
    DeviceBuffer amp, phase, coefs, raw, totals, base, base_totals, shape_totals, scales, mean, shaped, structure,
        ring_sum, ring_peak, ring_mean, radius, ref_amp, ref_phase, ref_raw, ref_total;
    // ring orders of the kernel sizes seen (azimuthal-mean and support rings), and their scratch
    std::vector<std::unique_ptr<RingPlan>> ring_plans;
    RingScratch ring_scratch;
    // The plan for (size, center, rmax, ceil), built on first use (a host step: call
    // reserve_shaping before a graph capture).
    const RingPlan& ring_plan(hipStream_t s, int size, float center, int rmax, bool ceil);
    // first-zero radius of the clear-circle PSF per (P, R) (GetLensDiffReferenceRawPsfCached)
    std::vector<std::pair<std::pair<int, int>, float>> first_zero;
    // Grows the shaping buffers for `count` kernels of side K, and builds their ring plans, ahead
    // of a graph capture.
    void reserve_shaping(hipStream_t s, int K, int count, int max_radius);
};

float first_minimum_from_profile(const std::vector<float>& profile, int size);

// Steps of the build, exposed for the parity entry points.
// d_raw_totals (nullable): per-PSF sums of raw PSFs that are not yet normalised.
// place (nullable): the output's blocks are placed for a sequential sum (SeqSum::place), whose
// walk the caller then runs.
// window > 0: the raw PSFs are held as their centred (2 window + 1)^2 blocks
// (PsfPlan::window_psfs), else whole (R x R).
void resample_kernels(hipStream_t s, const float* raw, int R, const float* d_scales, const float* d_raw_totals,
                      int support, int count, float* out, const SeqPlace* place = nullptr, int window = 0);
// resample_kernels, then the sequential sums of the base kernels into totals[count].
void base_kernels(hipStream_t s, BankWorkspace& w, const float* raw, int R, const float* d_scales,
                  const float* d_raw_totals, int support, int count, float* out, float* totals, int window = 0);
// Shapes `count` base kernels (count x K*K) into full/core/structure (sets x count x K*K, the
// first `sets` of the three) and radius[count] (device). `max_radius` as
// LensDiffCpuReference.cpp:1633. base_totals (nullable): per-kernel sums of a base that is not yet
// normalised (read as base / total). d_gains (nullable): per-kernel anisotropy gain
// 1 + 4 * max(0, emphasis), else from `anisotropy`. out3 is left unnormalised; out_scales (device
// [sets][count]) receives the unit-sum multipliers.
void shape_kernels(hipStream_t s, BankWorkspace& w, const float* base, const float* base_totals, int K, int count,
                   double anisotropy, int max_radius, float* out3, int* d_radius, float* out_scales,
                   const float* d_gains = nullptr, int sets = 3);
// out3 *= out_scales in place (normalised kernels for a caller that copies them out).
void apply_kernel_scales(hipStream_t s, float* out3, int K, int count, const int* d_radius, const float* scales);
// Radial mean profile of an R x R PSF about its centre (radialMeanProfile, :865-887), computed
// on the device and returned to the host.
std::vector<float> radial_profile(hipStream_t s, BankWorkspace& w, const float* psf, int R);

// The bank of p with its first `sets` kernel sets (1: full; 3: full, core, structure).
std::unique_ptr<DeviceBank> build_bank(hipStream_t s, const FftTables& t, BankWorkspace& w, const ldh_params& p,
                                       int sets = 3);

// First minimum of the clear-circle PSF for (P, R) (GetLensDiffReferenceRawPsfCached +
// estimateFirstMinimumRadius, LensDiffCpuReference.cpp:2018-2038): built once per workspace and
// (P, R), with a host synchronisation.
float clear_circle_first_zero(hipStream_t s, const FftTables& t, BankWorkspace& w, int P, int R);

// Per-wavelength phase coefficients and resample scales of a bank (the wavelength loop of
// buildPsfBankGlobalOnly, LensDiffCpuReference.cpp:1702-1734).
void bank_wavelength_params(const ldh_params& p, const Sizes& sz, const std::vector<float>& wl, float first_zero,
                            std::vector<PhaseCoef>& coefs, std::vector<float>& scales);
// Whether coefficient sets differ only in defocus (c[0]) and `enabled`, so that one texel
// evaluation of the other phase terms serves all of them.
bool defocus_only_coefs(const std::vector<PhaseCoef>& coefs);

// Batched build of many banks (batch.hip). mode: LDH_BATCH_* (include/lensdiff_hip.h).
struct BatchTiming {
    double exec_ms = 0.0, capture_ms = 0.0;
    int groups = 0, psfs = 0, graph_nodes = 0, mode = 0, branches = 0;
};
// Workspaces and capture streams of the extra branches of a graph batch, kept with a context so
// that their buffers and ring plans are reused from batch to batch. Branch 0 is the caller's
// workspace and stream; the streams here only take part in captures (the graph runs on the
// caller's stream).
struct BatchLanes {
    std::vector<std::unique_ptr<BankWorkspace>> work;
    std::vector<hipStream_t> streams;
    void ensure(int extra);
    BatchLanes() = default;
    BatchLanes(const BatchLanes&) = delete;
    BatchLanes& operator=(const BatchLanes&) = delete;
    ~BatchLanes();
};
constexpr int kGraphBranches = 3;
// lanes (nullable): with LDH_BATCH_GRAPH, the groups are spread over up to kGraphBranches
// parallel branches of the graph, each with its own workspace.
std::vector<std::shared_ptr<DeviceBank>> build_banks_batched(hipStream_t s, const FftTables& t, BankWorkspace& w,
                                                             const std::vector<ldh_params>& requests, int mode,
                                                             BatchTiming* timing, BatchLanes* lanes = nullptr);

}  // namespace ldh

// PSF synthesis plan and launchers.
#pragma once

#include "ldh_internal.h"
#include "pupil.h"
#include "seqsum.h"

namespace ldh {

constexpr int kMaxRawPsf = 8192;  // its Bluestein length (R + P - 1 <= 9215) fits one LDS-resident line

// Stage buffers of the two line passes (shareable by plans of different sizes).
struct PsfScratch {
    DeviceBuffer field, g, gt, row_sums;
    // raw: the output line length (R, or 2h + 1 for a window)
    void reserve(int pupil, int raw, int count) {
        field.reset(sizeof(float2) * size_t(pupil) * pupil * count);
        g.reset(sizeof(float2) * size_t(pupil) * raw * count);
        gt.reset(sizeof(float2) * size_t(pupil) * raw * count);
        row_sums.reset(sizeof(double) * size_t(raw) * count);
    }
};

// Per (P, R) state: chirp and Bluestein filter spectrum.
class PsfPlan {
public:
    void prepare(hipStream_t s, int pupil, int raw);
    // `count` raw PSFs from one amplitude (P*P) and `count` phase maps (count*P*P), both
    // transposed ([x][y]), into out (count*R*R, centred); totals: device[count], their
    // sequential-FP32 sums. With `normalize` each PSF is scaled to unit sum, else left for the
    // consumer to scale by 1/total.
    // amp_index (device, nullable): PSF l uses amplitude amp + amp_index[l] * P * P.
    // defocus_coefs (device, nullable): the `count` phase coefficient sets, for `pupils` pupils
    // whose sets differ only in defocus and `enabled` within a pupil; the phases are then formed
    // with the fields and neither `phase` nor amp_index is read: pupil k has amplitude
    // amp + k * P * P and the units unit0[k] .. unit0[k + 1] - 1 (device, pupils + 1 entries; null
    // for one pupil).
    void raw_psfs(hipStream_t s, const FftTables& t, const float* amp, const float* phase, int count,
                  float* out, float* totals, SeqSum& seqsum, PsfScratch& scratch, bool normalize = true,
                  const int* amp_index = nullptr, const PhaseCoef* defocus_coefs = nullptr,
                  const int* unit0 = nullptr, int pupils = 1);
    int pupil() const { return pupil_; }
    int raw() const { return raw_; }
    int bluestein_length() const { return m_; }
    bool bluestein() const { return bluestein_; }

    // Window form (after prepare): only the centred (2h+1) x (2h+1) block of each raw PSF, the
    // frequencies -h .. h about DC of both transforms, each line a chirp-z transform of the P
    // pupil values onto those 2h+1 frequencies (length >= P + 2h, with the same kernel
    // exp(-2 i pi_f n k / R) as the full form). A consumer that samples the raw PSF only near its
    // centre and renormalises (makeResampledKernel, LensDiffCpuReference.cpp:1060-1086) needs no
    // more. Returns false, preparing nothing, when the window would not be cheaper than the full
    // R x R form or would reach the PSF's edge.
    bool prepare_window(hipStream_t s, int h);
    // |F|^2 / M^2 of `count` windows into out[count][(2h+1)^2] (row-major, centred: element (i, j)
    // is raw pixel (R/2 - h + j, R/2 - h + i)), not normalised; arguments as raw_psfs.
    void window_psfs(hipStream_t s, const FftTables& t, const float* amp, const float* phase, int count, float* out,
                     PsfScratch& scratch, const int* amp_index = nullptr, const PhaseCoef* defocus_coefs = nullptr,
                     const int* unit0 = nullptr, int pupils = 1);
    int window() const { return h_; }
    int window_side() const { return 2 * h_ + 1; }
    int window_length() const { return mw_; }

private:
    int pupil_ = 0, raw_ = 0, m_ = 0, offset_ = 0;
    bool bluestein_ = false;
    DeviceBuffer chirp_, bhat_;
    int h_ = -1, mw_ = 0;
    DeviceBuffer wchirp_, wbhat_;
    // the pupil fields (shared by both forms)
    void fields(hipStream_t s, const float* amp, const float* phase, int count, const float2* chirp, int bluestein,
                PsfScratch& scratch, const int* amp_index, const PhaseCoef* defocus_coefs, const int* unit0,
                int pupils);
};

// Half-width of the raw-PSF window that resampling to a (2 support + 1)^2 kernel at each of
// `scales` reads (makeResampledKernel + sampleSquareFiltered, LensDiffCpuReference.cpp:316-353,
// :1060-1086): the bilinear taps of every sample position plus a margin of two texels.
int resample_window_half_width(const std::vector<float>& scales, int support);

// Pupil maps, row-major ([y][x]) or transposed ([x][y], the layout raw_psfs takes).
void pupil_amplitude(hipStream_t s, const PupilGeom& g, float* out, bool transposed);
// One amplitude per geometry (device array of `count`), all of size n: out[count][n*n].
void pupil_amplitudes(hipStream_t s, const PupilGeom* d_geoms, int n, int count, float* out, bool transposed);
void pupil_phase(hipStream_t s, const PhaseCoef* d_coefs, int n, int count, float* out, bool transposed);

}  // namespace ldh

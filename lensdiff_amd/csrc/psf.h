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

private:
    int pupil_ = 0, raw_ = 0, m_ = 0, offset_ = 0;
    bool bluestein_ = false;
    DeviceBuffer chirp_, bhat_;
};

// Pupil maps, row-major ([y][x]) or transposed ([x][y], the layout raw_psfs takes).
void pupil_amplitude(hipStream_t s, const PupilGeom& g, float* out, bool transposed);
// One amplitude per geometry (device array of `count`), all of size n: out[count][n*n].
void pupil_amplitudes(hipStream_t s, const PupilGeom* d_geoms, int n, int count, float* out, bool transposed);
void pupil_phase(hipStream_t s, const PhaseCoef* d_coefs, int n, int count, float* out, bool transposed);

}  // namespace ldh

// Per-frame stages: selection, FFT convolution of the highlights with the bank, spectral map
// and composite (RunLensDiffCpuReference, LensDiffCpuReference.cpp:2073-2261).
#pragma once

#include <vector>

#include "bank.h"
#include "host_params.h"
#include "psf.h"
#include "ldh_internal.h"
#include "seqsum.h"

namespace ldh {

// Spectral bin -> RGB matrices ([channel][bin]), natural and style (BuildLensDiffSpectrumConfig).
struct SpectrumConfig {
    int bins = 1;
    std::vector<float> natural, style;
};
SpectrumConfig build_spectrum_config(const ldh_params& p, const std::vector<float>& wavelengths);

// FFT grid of a frame convolved with kernels of support radius `support`: the reference's
// nextPow2(W + n - 1) x nextPow2(H + n - 1), n = 2*support+1 (LensDiffCpuReference.cpp:1168-1169).
// A half spectrum (the transform of a real plane) keeps columns 0 .. pw/2 of each row, the rest
// following by symmetry (X[-k] = conj X[k]), at a row pitch of hw columns (pw/2 + 1 rounded up
// to the column tile width; the extra columns hold zeros).
struct Grid {
    int w = 0, h = 0;     // frame
    int pw = 0, ph = 0;   // FFT grid
    int hw = 0;           // half-spectrum row pitch
    size_t cells() const { return size_t(pw) * ph; }
    size_t half_cells() const { return size_t(hw) * ph; }
};
Grid make_grid(int w, int h, int support);

// Kernel spectra of a set of real K x K kernels in the split-column layout of the forward
// transform: packed two per complex spectrum (k_a + i k_b), or with `half`, one half spectrum
// per kernel.
struct KernelSpectra {
    Grid grid;
    int kernels = 0;      // real kernels
    int pairs = 0;        // complex spectra (half: one per kernel)
    bool fp16 = false;
    bool half = false;
    DeviceBuffer data;    // float2 or __half2: [pairs][ph][pw], or with half [kernels][ph][hw]
    size_t stride() const { return half ? grid.half_cells() : grid.cells(); }
    size_t bytes() const { return data.bytes(); }
};
// kernels: K x K each (texels times their scale), packed in order (0,1), (2,3), ... (each pair
// transformed together and, with half, separated into two half spectra); `radius` bounds the
// non-zero support of every kernel.
void build_kernel_spectra(hipStream_t s, const FftTables& t, const std::vector<KernelRef>& kernels, int K,
                          int radius, const Grid& g, bool fp16, KernelSpectra& out, DeviceBuffer& scratch,
                          bool half = false);
// Combined RGB kernels of a spectral bank: out[c] = sum_b M[c][b] (k_b * scales[b]) for the
// natural (c=0..2) and, if with_style, style (c=3..5) matrices. out: device float[6][K*K],
// unit scale. bins <= LDH_MAX_WAVELENGTHS.
void combine_kernels(hipStream_t s, const float* kernels, const float* scales, int bins, int K,
                     const SpectrumConfig& cfg, bool with_style, float* out);

// Scratch and outputs of one frame on one stream.
struct FrameWork {
    DeviceBuffer src, dst, mask, image_spec, out_spec, planes, effect, lum, totals, factors, matrices, kspec_tmp,
        combined, max_buf, zone, post, gauss, work, redist, display, view_pupil, view_coef;
    KernelSpectra kspec;  // per-frame kernel spectra (plain path)
    float gauss_radius = -1.0f;  // the scatter kernel held in `gauss` (radius in px, side)
    int gauss_size = 0;
    SeqSum seqsum;
};

// Effect-producing convolution modes.
enum class EffectMode { PerBin, Precombined };

// The kernel sets (DeviceBank::sets) a render of p convolves with: 1 when only the full kernels
// are read (a physical look outside the core and structure views), 2 for the core view, else 3.
int kernel_sets_needed(const ldh_params& p);

// Renders one frame with the bank. `src`/`dst`: device RGBA float32, tightly packed, w x h.
// `cached`: kernel spectra matching the bank and grid (full kernels, bins or combined), or null
// to compute them in `fw.kspec` for this frame.
void render_frame(hipStream_t s, const FftTables& t, FrameWork& fw, const ldh_params& p, const DeviceBank& bank,
                  const float4* src, float4* dst, int w, int h, EffectMode mode, const KernelSpectra* cached);

// Renders one frame with field-dependent phase (the zone loop of RunLensDiffCpuReference,
// LensDiffCpuReference.cpp:2160-2170): banks[zy * 3 + zx] is the bank of zone (zx, zy); each
// zone's effect, core and structure planes are blended into the frame's with the zone's tent
// weights (accumulateWeightedImage, :1452-1475) before the Preserve rescale and composite.
// Every zone is convolved on zone_grid(w, h, banks), so the frame's spectra are formed once.
// cached (nullable, kFieldZones entries): each zone's kernel spectra on that grid, as for
// render_frame.
void render_frame_zones(hipStream_t s, const FftTables& t, FrameWork& fw, const ldh_params& p,
                        const std::vector<const DeviceBank*>& banks, const float4* src, float4* dst, int w, int h,
                        EffectMode mode, const std::vector<const KernelSpectra*>* cached = nullptr);
// The convolution grid of a field-phase frame: the largest support of its zone banks.
Grid zone_grid(int w, int h, const std::vector<const DeviceBank*>& banks);

// The static debug views (BuildLensDiffStaticDebugRgba, LensDiffCpuReference.cpp:1909-1994)
// at w x h into dst: pupil, phase and phase edge from the pupil of `p`, PSF and field PSF from the
// first bin of `bank`, the chromatic split from its last, middle and first bins. The caller
// passes the edge zone's parameters and bank for the phase-edge and field-PSF views of a
// field-dependent phase.
bool is_static_view(int view);
void render_static_view(hipStream_t s, FrameWork& fw, const ldh_params& p, int view, const DeviceBank& bank,
                        float4* dst, int w, int h);

// Which real kernels feed the effect, for building cached spectra.
std::vector<KernelRef> effect_kernels(const ldh_params& p, const DeviceBank& bank, EffectMode mode,
                                      const float* combined, int* which_out);

}  // namespace ldh

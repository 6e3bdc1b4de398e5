/*
 * lensdiff_oracle.h — C API of the CPU oracle: a restatement of the LensDiff reference CPU
 * path (source/src/core/LensDiffCpuReference.cpp, LensDiffPhase.cpp, LensDiffSpectrum.cpp,
 * LensDiffTransfer.cpp, common/color/ColorManagement.cpp).
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load this library, and only as the checker. The product path never calls it.
 *
 * Every function exists in two instantiations:
 *   _f32 — FP32, following the reference's float operation order (radix-2 FFT with the
 *          twiddle recurrence, sequential float sums), the "reference FP32" result;
 *   _f64 — the same algorithm in FP64, the "FP64 oracle" every result is also scored against.
 */
#ifndef LENSDIFF_ORACLE_H
#define LENSDIFF_ORACLE_H

#include "lensdiff_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

void orc_params_default(ldh_params* p);
int orc_resolve_sizes(const ldh_params* p, ldh_sizes* out);
/* Wavelength list the bank uses (fixed mode list or the explicit extension list). */
int orc_wavelengths(const ldh_params* p, float* out /* LDH_MAX_WAVELENGTHS */);
int orc_has_nonflat_phase(const ldh_params* p);

#define ORC_DECLARE(T, S)                                                                       \
    void orc_pupil_amplitude_##S(const ldh_params* p, T* out);                                  \
    void orc_pupil_phase_##S(const ldh_params* p, T* out);                                      \
    void orc_raw_psf_##S(const T* amp, const T* phase, int pupil, int raw, T* out);             \
    void orc_fft1d_##S(T* data, int n, int inverse);                                            \
    void orc_fft2d_##S(T* data, int w, int h, int inverse);                                     \
    T orc_first_minimum_radius_##S(const T* raw, int raw_size);                                 \
    void orc_resample_kernel_##S(const T* raw, int raw_size, T scale, int support, T* out);     \
    int orc_shape_kernels_##S(const ldh_params* p, const T* base, int k, T* full, T* core,       \
                              T* structure, int* out_size);                                     \
    void* orc_bank_build_##S(const ldh_params* p);                                              \
    void orc_spectrum_config_##S(const ldh_params* p, int n, const float* wl, T* natural,       \
                                 T* style);                                                     \
    void orc_map_bins_##S(const ldh_params* p, int n, const T* natural, const T* style,         \
                          const T* bins, long npx, T* rgb);                                     \
    void orc_selection_##S(const ldh_params* p, const float* rgba, int w, int h, T* mask,        \
                           T* redistributed_rgba, T* driver);                                   \
    void orc_convolve_##S(const T* plane, int w, int h, const T* kernel, int k, T* out);        \
    int orc_render_image_##S(const ldh_params* p, const float* src_rgba, int w, int h,          \
                             T* out_rgba);

ORC_DECLARE(float, f32)
ORC_DECLARE(double, f64)
#undef ORC_DECLARE

/* Bank handles (from orc_bank_build_f32 / _f64). `which` is LDH_KERNEL_*. */
int orc_bank_bin_count(const void* bank);
int orc_bank_support_radius(const void* bank);
int orc_bank_kernel_size(const void* bank, int bin);
float orc_bank_wavelength(const void* bank, int bin);
int orc_bank_raw_size(const void* bank);
void orc_bank_kernel_f32(const void* bank, int bin, int which, float* out);
void orc_bank_kernel_f64(const void* bank, int bin, int which, double* out);
/* The per-bin base (resampled, uncropped) kernels, K*K with K = 2*max_radius+1. */
int orc_bank_base_size(const void* bank);
void orc_bank_base_f32(const void* bank, int bin, float* out);
void orc_bank_base_f64(const void* bank, int bin, double* out);
void orc_bank_free(void* bank);

/* The plugin's view-based entry (RunLensDiffCpuReference semantics): strided host views,
 * render window, transfer encode. Returns 1 on success, 0 with `err` filled on failure. */
int orc_render_f32(const ldh_params* p, const ldh_render_request* req, char* err, int err_len);

/* Sequential float sum (std::accumulate with a float accumulator), for the sum emulation. */
float orc_sequential_sum_f32(const float* x, long n);

#ifdef __cplusplus
}
#endif
#endif

/*
 * lensdiff_hip.h — C ABI of the MI355X (gfx950) LensDiff diffraction backend.
 *
 * Plain C, POD-only: no C++ types, exceptions or std::string cross this boundary.
 * Every entry point returns an ldh_status (0 = LDH_OK). On failure the context keeps a
 * short kebab-case error tag, readable with ldh_last_error(), mirroring the reference's
 * `std::string* error` convention (source/src/core/LensDiffCpuReference.cpp:2077-2084).
 *
 * The plugin-facing C++ seam, RunLensDiffHip(request, params, cache, error), is a thin layer
 * above these calls (lensdiff_amd/plugin/LensDiffHip.h); it mirrors
 *   bool RunLensDiffCpuReference(const LensDiffRenderRequest&, const LensDiffParams&,
 *                                LensDiffPsfBankCache&, std::string*)
 *   (source/src/core/LensDiffCpuReference.h:64-67) and
 *   bool RunLensDiffCuda(...)  (source/src/cuda/LensDiffCuda.h:7-10).
 *
 * Enum values and field meanings follow source/src/core/LensDiffTypes.h:14-208.
 */
#ifndef LENSDIFF_HIP_H
#define LENSDIFF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDH_ABI_VERSION 1
/* Largest wavelength list a bank may carry (the reference caps fixed modes at 9 bins,
 * LensDiffTypes.h:12; lists longer than 9 are an extension of this backend). */
#define LDH_MAX_WAVELENGTHS 64

/* ---- status codes ------------------------------------------------------------------- */
typedef int32_t ldh_status;
enum {
    LDH_OK = 0,
    LDH_ERR_INVALID_ARGUMENT = 1,
    LDH_ERR_MISSING_BUFFER = 2,   /* "missing-source-or-destination-buffer" */
    LDH_ERR_INVALID_BOUNDS = 3,   /* "invalid-source-bounds" */
    LDH_ERR_UNSUPPORTED = 4,      /* a parameter this backend does not implement */
    LDH_ERR_OUT_OF_MEMORY = 5,    /* device allocation failed or budget exceeded */
    LDH_ERR_DEVICE = 6,           /* a HIP runtime call failed */
    LDH_ERR_NOT_READY = 7,        /* ldh_poll: request still in flight */
    LDH_ERR_TOO_LARGE = 8         /* a size beyond this backend's documented limits */
};

/* ---- enums (values identical to LensDiffTypes.h:14-91) ------------------------------ */
enum { LDH_EXTRACT_MAX_RGB = 0, LDH_EXTRACT_LUMA = 1 };
enum {
    LDH_APERTURE_CIRCLE = 0, LDH_APERTURE_POLYGON = 1, LDH_APERTURE_STAR = 2,
    LDH_APERTURE_SPIRAL = 3, LDH_APERTURE_CUSTOM = 4, LDH_APERTURE_HEXAGON = 5,
    LDH_APERTURE_SQUARE_GRID = 6, LDH_APERTURE_SNOWFLAKE = 7
};
enum { LDH_APOD_FLAT = 0, LDH_APOD_COSINE = 1, LDH_APOD_GAUSSIAN = 2 };
enum { LDH_SPECTRAL_MONO = 0, LDH_SPECTRAL_TRISTIMULUS = 1, LDH_SPECTRAL_5 = 2, LDH_SPECTRAL_9 = 3 };
enum { LDH_STYLE_NATURAL = 0, LDH_STYLE_CYAN_MAGENTA = 1, LDH_STYLE_WARM_COOL = 2 };
enum { LDH_LOOK_PHYSICAL = 0, LDH_LOOK_SPLIT = 1 };
enum { LDH_ENERGY_PRESERVE = 0, LDH_ENERGY_AUGMENT = 1 };
enum {
    LDH_VIEW_FINAL = 0, LDH_VIEW_SELECTION = 1, LDH_VIEW_PUPIL = 2, LDH_VIEW_PSF = 3,
    LDH_VIEW_OTF = 4, LDH_VIEW_CORE = 5, LDH_VIEW_STRUCTURE = 6, LDH_VIEW_EFFECT = 7,
    LDH_VIEW_PHASE = 8, LDH_VIEW_PHASE_EDGE = 9, LDH_VIEW_FIELD_PSF = 10,
    LDH_VIEW_CHROMATIC_SPLIT = 11, LDH_VIEW_CREATIVE_FRINGE = 12, LDH_VIEW_SCATTER = 13
};
enum { LDH_TRANSFER_LINEAR = 0, LDH_TRANSFER_DAVINCI_INTERMEDIATE = 1 };
enum { LDH_PHASE_OFF = 0, LDH_PHASE_ENABLED = 1 };
/* Which kernel of a bank bin (LensDiffPsfBin, LensDiffTypes.h:373-378). */
enum { LDH_KERNEL_FULL = 0, LDH_KERNEL_CORE = 1, LDH_KERNEL_STRUCTURE = 2 };
/* How the plain render convolves the spectral bins. */
enum {
    LDH_CONV_PER_BIN = 0,      /* one convolution per bin, then the per-pixel spectral map */
    LDH_CONV_PRECOMBINED = 1   /* bins pre-combined into natural/style RGB kernels */
};
/* Raw PSFs of a bank build. AUTO: where the resample reads only a centred block of each raw PSF
 * (makeResampledKernel, LensDiffCpuReference.cpp:1060-1086) and that block is cheaper, only the
 * block is synthesised and its unit-sum scale, which the kernel's own normalisation divides out,
 * is skipped; the kernels then follow the reference's to the size of its FP32 sum rounding
 * (up to about 5e-4 relative L2 at R = 4092). WHOLE: every raw PSF is synthesised R x R and
 * normalised by its sequential FP32 sum, as BuildLensDiffShiftedRawPsf (:831-859) does. */
enum { LDH_RAW_PSF_AUTO = 0, LDH_RAW_PSF_WHOLE = 1 };

/* ---- image views (LensDiffImageRect / LensDiffImageView, LensDiffTypes.h:93-111) ---- */
typedef struct ldh_rect {
    int32_t x1, y1, x2, y2;
} ldh_rect;

typedef struct ldh_image_view {
    void* data;                 /* host pointer (OFX has no HIP render suite) or device */
    int64_t row_bytes;          /* may be negative */
    ldh_rect bounds;
    int32_t origin_x, origin_y;
    int32_t components;         /* 4 (RGBA); 3 is accepted for reading */
    int32_t bytes_per_component;/* 4 (float32) */
} ldh_image_view;

/* LensDiffRenderRequest (LensDiffTypes.h:113-126), host-pointer form. */
typedef struct ldh_render_request {
    ldh_rect render_window;
    ldh_rect frame_bounds;
    ldh_image_view src;
    ldh_image_view dst;
    int32_t frame_short_side_px;
} ldh_render_request;

/* LensDiffParams (LensDiffTypes.h:128-208) plus this backend's extensions (last block). */
typedef struct ldh_params {
    int32_t extraction_mode, aperture_mode, apodization_mode, spectral_mode;
    int32_t spectrum_style, look_mode, energy_mode, debug_view;
    int32_t input_transfer, resolution_aware;

    double threshold, softness_stops, point_emphasis, core_preserve;

    int32_t blade_count, vane_count, pupil_resolution, frame_short_side_px;
    double roundness, rotation_deg, central_obstruction, vane_thickness;
    double diffraction_scale_px;   /* percent of the optics short side */
    double max_kernel_radius_px;   /* percent of the optics short side */

    int32_t phase_mode, chromatic_affects_luma;
    double phase_defocus, phase_astigmatism0, phase_astigmatism45, phase_coma_x, phase_coma_y;
    double phase_spherical, phase_trefoil_x, phase_trefoil_y;
    double phase_secondary_astigmatism0, phase_secondary_astigmatism45;
    double phase_quadrafoil0, phase_quadrafoil45, phase_secondary_coma_x, phase_secondary_coma_y;
    double pupil_decenter_x, pupil_decenter_y;
    double phase_field_strength, phase_field_edge_bias, phase_field_defocus;
    double phase_field_astig_radial, phase_field_astig_tangential;
    double phase_field_coma_radial, phase_field_coma_tangential, phase_field_spherical;
    double phase_field_trefoil_radial, phase_field_trefoil_tangential;
    double phase_field_secondary_astig_radial, phase_field_secondary_astig_tangential;
    double phase_field_quadrafoil_radial, phase_field_quadrafoil_tangential;
    double phase_field_secondary_coma_radial, phase_field_secondary_coma_tangential;

    double spectrum_force, spectrum_saturation;
    double chromatic_focus, chromatic_spread, creative_fringe, scatter_amount, scatter_radius;
    double effect_gain, core_compensation, anisotropy_emphasis;
    double core_gain, structure_gain, core_shoulder, structure_shoulder;

    /* ---- extensions (no counterpart in LensDiffParams) ---- */
    /* Explicit wavelength list; 0 = the spectral mode's fixed list
     * (LensDiffCpuReference.cpp:1602-1614). */
    int32_t wavelength_count;
    int32_t conv_mode;             /* LDH_CONV_* for spectral banks */
    float wavelengths_nm[LDH_MAX_WAVELENGTHS];
    /* Zernike terms the reference basis lacks (LensDiffPhase.cpp:182-197), in waves:
     * Noll 2 (tilt x), 3 (tilt y), 18, 19, 20, 21. */
    double zernike_extra[6];
    /* LDH_RAW_PSF_*: how bank builds form the raw PSFs they resample. */
    int32_t raw_psf_form;
    /* 1 = store cached kernel spectra as FP16 (FP32 accumulation). */
    int32_t spectra_fp16;
} ldh_params;

/* Fill *p with LensDiffParams' member defaults (LensDiffTypes.h:128-208). */
void ldh_params_default(ldh_params* p);

/* ---- context ------------------------------------------------------------------------ */
typedef struct ldh_context ldh_context;
typedef struct ldh_bank ldh_bank;

/* Create a context on HIP device `device`. `memory_budget_bytes` bounds the resident bank
 * cache (0 = 64 GiB). Contexts are internally synchronised. */
ldh_status ldh_context_create(int32_t device, uint64_t memory_budget_bytes, ldh_context** out);
ldh_status ldh_context_destroy(ldh_context* ctx);
const char* ldh_status_string(ldh_status status);
/* Last kebab-case error tag recorded on this context ("" if none). */
const char* ldh_last_error(const ldh_context* ctx);
ldh_status ldh_synchronize(ldh_context* ctx);

/* ---- sizes (LensDiffCpuReference.cpp:104-115, LensDiffTypes.h:292-320) ---------------- */
typedef struct ldh_sizes {
    int32_t pupil_size;            /* P = clamp(pupilResolution, 64, 1024) */
    int32_t raw_psf_size;          /* R = clamp(max(2P, 512, 4r), 128, 16384) */
    int32_t max_kernel_radius_px;  /* r = max(4, ceil(percent * 0.01 * S)) */
    int32_t optics_short_side_px;  /* S */
    double diffraction_scale_px;
} ldh_sizes;
ldh_status ldh_resolve_sizes(const ldh_params* p, ldh_sizes* out);

/* ---- PSF synthesis stages (host buffers in, host buffers out; for parity and debug) ---- */
/* BuildLensDiffPupilAmplitudeForPsf (LensDiffCpuReference.cpp:758-822): P*P floats. */
ldh_status ldh_pupil_amplitude(ldh_context* ctx, const ldh_params* p, float* out);
/* BuildLensDiffPupilPhaseForPsf (LensDiffPhase.cpp:215-243): P*P waves. */
ldh_status ldh_pupil_phase(ldh_context* ctx, const ldh_params* p, float* out);
/* BuildLensDiffShiftedRawPsf (LensDiffCpuReference.cpp:831-859): R*R floats, unit sum.
 * `phase_waves` may be NULL (flat phase). */
ldh_status ldh_raw_psf(ldh_context* ctx, const float* amplitude, const float* phase_waves,
                       int32_t pupil_size, int32_t raw_size, float* out);
/* The centred (2h+1) x (2h+1) block of the raw PSF alone, as bank builds synthesise it when the
 * resample reads no more (no counterpart in the reference, whose bank path computes the whole
 * R x R PSF, BuildLensDiffShiftedRawPsf, and samples it near its centre, :1060-1086): element
 * (i, j) is raw pixel (R/2 - h + j, R/2 - h + i), |F|^2 without the unit-sum normalisation (any
 * common factor). LDH_ERR_UNSUPPORTED ("raw-psf-window-not-applicable") where the block would
 * reach the PSF's edge or would not be cheaper than the whole PSF. */
ldh_status ldh_raw_psf_window(ldh_context* ctx, const float* amplitude, const float* phase_waves,
                              int32_t pupil_size, int32_t raw_size, int32_t half_width, float* out);
/* EstimateLensDiffFirstMinimumRadius (LensDiffCpuReference.cpp:897-955) of a raw PSF. */
ldh_status ldh_first_minimum_radius(ldh_context* ctx, const float* shifted_raw_psf,
                                    int32_t raw_size, float* out_radius);
/* makeResampledKernel (LensDiffCpuReference.cpp:1060-1086): K=2r+1, K*K floats. */
ldh_status ldh_resample_kernel(ldh_context* ctx, const float* shifted_raw_psf, int32_t raw_size,
                               float scale_factor, int32_t support_radius, float* out);
/* populatePsfBankCacheFromBaseKernels for one bin (LensDiffCpuReference.cpp:1616-1656):
 * base K*K kernel -> full/core/structure of side (2*radius+1). Each output needs K*K floats
 * of capacity; *out_size receives the cropped side. */
ldh_status ldh_shape_kernels(ldh_context* ctx, const ldh_params* p, const float* base_kernel,
                             int32_t kernel_size, float* full, float* core, float* structure,
                             int32_t* out_size);
/* sumKernel (LensDiffCpuReference.cpp:436-438): the value of a sequential FP32 accumulation of
 * x[0..n) (std::accumulate with a float accumulator), computed on the device. */
ldh_status ldh_sequential_sum(ldh_context* ctx, const float* x, int64_t n, float* out);
/* The same for `count` arrays of n floats stored back to back: out[i] = the sequential FP32
 * sum of x[i*n .. (i+1)*n). block: the device's block length (0 for the default, else 512 to
 * 4096); the value does not depend on it. The PSF synthesis sums each raw PSF over blocks of one
 * row (sumKernel per wavelength, LensDiffCpuReference.cpp:440-449). */
ldh_status ldh_sequential_sums(ldh_context* ctx, const float* x, int64_t n, int32_t count, int32_t block,
                               float* out);
/* In-place 2D complex FFT of interleaved complex64 (host buffer), rows then columns,
 * forward sign -1, inverse scaled by 1/(w*h) (LensDiffCpuReference.cpp:194-301). Sizes:
 * powers of two (width >= 16, height >= 8), 3*2^a from 48 and 9*2^a from 144, up to 16384;
 * the reference takes powers of two only. */
ldh_status ldh_fft2d(ldh_context* ctx, float* data, int32_t width, int32_t height, int32_t inverse);
/* ldh_sequential_sums without allocating (no counterpart in the reference; for callers that cannot
 * allocate device memory: code under HIP graph capture, serving loops, harnesses that forbid it).
 * x: `count` device arrays of n floats at x + i * n, each >= 0 and finite; totals: device float[count];
 * stream: a hipStream_t, or NULL for the context's stream. The work is only enqueued (synchronise
 * the stream before reading totals). The block tables live in a 1.75 MiB static device array of the
 * library, one per GPU, held until the process ends: count * ceil(n / 4096) must not exceed 65536
 * (LDH_ERR_TOO_LARGE, "noalloc-scratch-too-small"). Calls that overlap on one GPU share that array
 * and race; each call overwrites whatever an earlier call left in it. */
ldh_status ldh_sequential_sums_noalloc(ldh_context* ctx, const float* x, int64_t n, int32_t count, float* totals,
                                       void* stream);

/* ---- PSF banks (EnsureLensDiffPsfBank, LensDiffCpuReference.cpp:2068-2071) ------------- */
/* Batched bank synthesis (no counterpart in the reference, which builds one bank per render). */
enum {
    LDH_BATCH_PER_REQUEST = 0,  /* one bank build after another (the single-bank path) */
    LDH_BATCH_GROUPED = 1,      /* requests with equal pupil/raw/kernel sizes share every launch */
    LDH_BATCH_GRAPH = 2         /* the groups captured into one HIP graph as up to three parallel
                                   branches (each with its own workspace), launched once */
};
typedef struct ldh_batch_stats {
    double exec_ms;        /* device time of the build (HIP events around its launches) */
    double capture_ms;     /* host time to capture and instantiate the graph (0 without one) */
    int32_t groups;        /* launch groups (requests in LDH_BATCH_PER_REQUEST) */
    int32_t psfs;          /* (request, wavelength) PSFs built */
    int32_t graph_nodes;   /* nodes of the captured graph (0 without one) */
    int32_t mode;
    int32_t branches;      /* parallel branches of the graph (1 without one) */
} ldh_batch_stats;
typedef struct ldh_bank_info {
    int32_t bin_count;
    int32_t support_radius_px;     /* LensDiffPsfBankCache::supportRadiusPx */
    int32_t pupil_size;
    int32_t raw_psf_size;
    float wavelengths_nm[LDH_MAX_WAVELENGTHS];
    int32_t kernel_size[LDH_MAX_WAVELENGTHS];   /* cropped side per bin */
    int32_t raw_window_px;  /* half-width h of the centred (2h+1)^2 raw-PSF block the bins were
                               resampled from when only that block was synthesised; 0: whole R x R */
} ldh_bank_info;

/* Build or fetch the bank for `p` from the context's cache (refcounted). */
ldh_status ldh_bank_acquire(ldh_context* ctx, const ldh_params* p, ldh_bank** out);
ldh_status ldh_bank_release(ldh_context* ctx, ldh_bank* bank);
/* Builds `count` banks (params[i] -> banks[i], each released with ldh_bank_release) in the
 * given LDH_BATCH_* mode; all modes give the same banks. stats may be NULL. */
ldh_status ldh_bank_build_batch(ldh_context* ctx, const ldh_params* params, int32_t count, int32_t mode,
                                ldh_bank** banks, ldh_batch_stats* stats);
ldh_status ldh_bank_info_get(ldh_context* ctx, const ldh_bank* bank, ldh_bank_info* out);
/* Copy one kernel of one bin to host (capacity in floats). */
ldh_status ldh_bank_download(ldh_context* ctx, const ldh_bank* bank, int32_t bin, int32_t which,
                             float* out, int64_t capacity);

/* ---- render --------------------------------------------------------------------------- */
/* The plain per-frame path (RunLensDiffCpuReference semantics, LensDiffCpuReference.cpp:
 * 2073-2261): host views in and out, synchronous, single-entry bank cache per context. */
ldh_status ldh_render(ldh_context* ctx, const ldh_render_request* req, const ldh_params* p);

/* Same path with a device-resident, tightly packed RGBA float32 frame (w*h*4 floats) in
 * and out; runs on the context stream without host staging. */
ldh_status ldh_render_device(ldh_context* ctx, const ldh_params* p, const float* d_src,
                             float* d_dst, int32_t width, int32_t height);

/* ---- served form -------------------------------------------------------------------------
 * No counterpart in the reference (its plugin renders one frame per call with a single-entry
 * instance cache, LensDiff.cpp:2318-2319). A context can run a request loop: `slots` device
 * slots (one HIP stream, frame scratch and bank workspace each) take queued render requests in
 * arrival order as soon as a slot is free and the resident bank/spectra cache fits the memory
 * budget; banks and their kernel spectra stay resident in HBM keyed by the lens parameters
 * (LRU eviction of entries no request is using). Frames are device-resident RGBA float32. */
typedef struct ldh_request_stats {
    double latency_ms;   /* completion - submission */
    double queue_ms;     /* start - submission */
    double service_ms;   /* completion - start */
    int32_t bank_hit, spectra_hit;
    ldh_status status;
    char error[48];
} ldh_request_stats;

typedef struct ldh_server_stats {
    int64_t submitted, completed, bank_hits, bank_misses, spectra_hits, spectra_misses, evictions;
    uint64_t resident_bytes, budget_bytes;
    int32_t slots, cached_banks;
} ldh_server_stats;

ldh_status ldh_server_start(ldh_context* ctx, int32_t slots, uint64_t budget_bytes);
ldh_status ldh_server_stop(ldh_context* ctx);
ldh_status ldh_submit(ldh_context* ctx, const ldh_params* p, const float* d_src, float* d_dst, int32_t width,
                      int32_t height, int64_t* ticket);
/* LDH_ERR_NOT_READY while the request is queued or running. */
ldh_status ldh_poll(ldh_context* ctx, int64_t ticket, ldh_request_stats* out);
ldh_status ldh_wait(ldh_context* ctx, int64_t ticket, ldh_request_stats* out);
ldh_status ldh_server_stats_get(ldh_context* ctx, ldh_server_stats* out);

/* Device memory helpers for callers without a GPU framework. */
ldh_status ldh_device_alloc(ldh_context* ctx, uint64_t bytes, void** out);
ldh_status ldh_device_free(ldh_context* ctx, void* ptr);
ldh_status ldh_memcpy_h2d(ldh_context* ctx, void* dst, const void* src, uint64_t bytes);
ldh_status ldh_memcpy_d2h(ldh_context* ctx, void* dst, const void* src, uint64_t bytes);

/* ---- measurement ---------------------------------------------------------------------- */
/* Per-kernel-family timing recorded with HIP events on the launch stream. */
typedef struct ldh_kernel_stat {
    char name[48];
    int64_t launches;
    double total_ms;
    double algorithmic_bytes;      /* summed byte-model figure of those launches */
} ldh_kernel_stat;
ldh_status ldh_profiling_enable(ldh_context* ctx, int32_t enable);
/* Returns the number of families written (<= capacity). */
int32_t ldh_profiling_read(ldh_context* ctx, ldh_kernel_stat* out, int32_t capacity);
ldh_status ldh_profiling_reset(ldh_context* ctx);
/* Kernel-variant choices of the process-wide autotuner (no counterpart in the reference): one
 * line per tuned key, "key@devN: vK  t0 t1 ... ms". Writes at most capacity - 1 bytes plus a NUL
 * into buf (buf may be NULL when capacity is 0); returns the full length. */
int32_t ldh_tuning_report(char* buf, int32_t capacity);

/* Pins every autotuned kernel family to one variant (variant >= 0; a family with fewer variants
 * takes its last) or returns them to tuning on first use (variant = -1); clears the choices made
 * so far (no counterpart in the reference: lets a caller or a test run each variant). */
void ldh_tuning_force(int32_t variant);

/* Test hook (no counterpart in the reference): the next `count` times the library reaches the
 * named failure site it fails there as a device error would, with the tag "injected-<site>".
 * Sites: "served-bank" (a served request building its bank), "served-spectra" (a served request
 * building the kernel spectra of its FFT grid). count = 0 disarms the site. Process-wide. */
void ldh_debug_inject_failure(const char* site, int32_t count);

#ifdef __cplusplus
}
#endif

#endif /* LENSDIFF_HIP_H */

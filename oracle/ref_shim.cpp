// C-ABI shim over the reference CPU path, compiled with the reference's own sources where they
// lie (see oracle/Makefile, target `ref`). The output, oracle/_ref/liblensdiff_ref.so, is a
// checker for the oracle and the CPU baseline; it is never part of the product.
//
// The reference's hot-path helpers live in an anonymous namespace of
// source/src/core/LensDiffCpuReference.cpp, so that translation unit is included here to reach
// them; nothing of it is copied.

#include "LensDiffCpuReference.cpp"  // NOLINT: compiled from /root/reference, not vendored

#include "lensdiff_hip.h"

#include <chrono>
#include <cstring>

namespace {

LensDiffParams to_ref(const ldh_params& p) {
    LensDiffParams r;
    r.extractionMode = static_cast<LensDiffExtractionMode>(p.extraction_mode);
    r.apertureMode = static_cast<LensDiffApertureMode>(p.aperture_mode);
    r.apodizationMode = static_cast<LensDiffApodizationMode>(p.apodization_mode);
    r.spectralMode = static_cast<LensDiffSpectralMode>(p.spectral_mode);
    r.spectrumStyle = static_cast<LensDiffSpectrumStyle>(p.spectrum_style);
    r.lookMode = static_cast<LensDiffLookMode>(p.look_mode);
    r.energyMode = static_cast<LensDiffEnergyMode>(p.energy_mode);
    r.debugView = static_cast<LensDiffDebugView>(p.debug_view);
    r.inputTransfer = static_cast<LensDiffInputTransfer>(p.input_transfer);
    r.resolutionAware = p.resolution_aware != 0;
    r.threshold = p.threshold;
    r.softnessStops = p.softness_stops;
    r.pointEmphasis = p.point_emphasis;
    r.corePreserve = p.core_preserve;
    r.bladeCount = p.blade_count;
    r.roundness = p.roundness;
    r.rotationDeg = p.rotation_deg;
    r.centralObstruction = p.central_obstruction;
    r.vaneCount = p.vane_count;
    r.vaneThickness = p.vane_thickness;
    r.diffractionScalePx = p.diffraction_scale_px;
    r.pupilResolution = p.pupil_resolution;
    r.maxKernelRadiusPx = p.max_kernel_radius_px;
    r.frameShortSidePx = p.frame_short_side_px;
    r.phaseMode = static_cast<LensDiffPhaseMode>(p.phase_mode);
    r.phaseDefocus = p.phase_defocus;
    r.phaseAstigmatism0 = p.phase_astigmatism0;
    r.phaseAstigmatism45 = p.phase_astigmatism45;
    r.phaseComaX = p.phase_coma_x;
    r.phaseComaY = p.phase_coma_y;
    r.phaseSpherical = p.phase_spherical;
    r.phaseTrefoilX = p.phase_trefoil_x;
    r.phaseTrefoilY = p.phase_trefoil_y;
    r.phaseSecondaryAstigmatism0 = p.phase_secondary_astigmatism0;
    r.phaseSecondaryAstigmatism45 = p.phase_secondary_astigmatism45;
    r.phaseQuadrafoil0 = p.phase_quadrafoil0;
    r.phaseQuadrafoil45 = p.phase_quadrafoil45;
    r.phaseSecondaryComaX = p.phase_secondary_coma_x;
    r.phaseSecondaryComaY = p.phase_secondary_coma_y;
    r.pupilDecenterX = p.pupil_decenter_x;
    r.pupilDecenterY = p.pupil_decenter_y;
    r.phaseFieldStrength = p.phase_field_strength;
    r.phaseFieldEdgeBias = p.phase_field_edge_bias;
    r.phaseFieldDefocus = p.phase_field_defocus;
    r.phaseFieldAstigRadial = p.phase_field_astig_radial;
    r.phaseFieldAstigTangential = p.phase_field_astig_tangential;
    r.phaseFieldComaRadial = p.phase_field_coma_radial;
    r.phaseFieldComaTangential = p.phase_field_coma_tangential;
    r.phaseFieldSpherical = p.phase_field_spherical;
    r.phaseFieldTrefoilRadial = p.phase_field_trefoil_radial;
    r.phaseFieldTrefoilTangential = p.phase_field_trefoil_tangential;
    r.phaseFieldSecondaryAstigRadial = p.phase_field_secondary_astig_radial;
    r.phaseFieldSecondaryAstigTangential = p.phase_field_secondary_astig_tangential;
    r.phaseFieldQuadrafoilRadial = p.phase_field_quadrafoil_radial;
    r.phaseFieldQuadrafoilTangential = p.phase_field_quadrafoil_tangential;
    r.phaseFieldSecondaryComaRadial = p.phase_field_secondary_coma_radial;
    r.phaseFieldSecondaryComaTangential = p.phase_field_secondary_coma_tangential;
    r.spectrumForce = p.spectrum_force;
    r.spectrumSaturation = p.spectrum_saturation;
    r.chromaticAffectsLuma = p.chromatic_affects_luma != 0;
    r.chromaticFocus = p.chromatic_focus;
    r.chromaticSpread = p.chromatic_spread;
    r.creativeFringe = p.creative_fringe;
    r.scatterAmount = p.scatter_amount;
    r.scatterRadius = p.scatter_radius;
    r.effectGain = p.effect_gain;
    r.coreCompensation = p.core_compensation;
    r.anisotropyEmphasis = p.anisotropy_emphasis;
    r.coreGain = p.core_gain;
    r.structureGain = p.structure_gain;
    r.coreShoulder = p.core_shoulder;
    r.structureShoulder = p.structure_shoulder;
    return r;
}

LensDiffImageView to_ref(const ldh_image_view& v) {
    LensDiffImageView r;
    r.data = v.data;
    r.rowBytes = v.row_bytes;
    r.bounds = {v.bounds.x1, v.bounds.y1, v.bounds.x2, v.bounds.y2};
    r.originX = v.origin_x;
    r.originY = v.origin_y;
    r.components = v.components;
    r.bytesPerComponent = v.bytes_per_component;
    return r;
}

void copy_out(const std::vector<float>& v, float* out) { std::copy(v.begin(), v.end(), out); }

RgbaImage image_from(const float* rgba, int w, int h, LensDiffInputTransfer tf) {
    RgbaImage img(w, h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float* q = rgba + (size_t(y) * w + x) * 4;
            const WorkshopColor::Vec3f d = LensDiffTransfer::decodeToLinear({q[0], q[1], q[2]}, tf);
            float* o = img.pixel(x, y);
            o[0] = d.x; o[1] = d.y; o[2] = d.z; o[3] = q[3];
        }
    return img;
}

}  // namespace

extern "C" {

void ref_pupil_amplitude(const ldh_params* p, float* out) { copy_out(BuildLensDiffPupilAmplitudeForPsf(to_ref(*p)), out); }
void ref_pupil_phase(const ldh_params* p, float* out) { copy_out(BuildLensDiffPupilPhaseForPsf(to_ref(*p)), out); }

void ref_raw_psf(const float* amp, const float* phase, int pupil, int raw, float* out) {
    std::vector<float> a(amp, amp + size_t(pupil) * pupil);
    if (phase) {
        std::vector<float> ph(phase, phase + size_t(pupil) * pupil);
        copy_out(BuildLensDiffShiftedRawPsf(a, ph, pupil, raw), out);
    } else {
        copy_out(BuildLensDiffShiftedRawPsf(a, pupil, raw), out);
    }
}

void ref_fft1d(float* d, int n, int inverse) {
    std::vector<Complex> a(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) a[size_t(i)] = Complex(d[2 * i], d[2 * i + 1]);
    fft1d(a, inverse != 0);
    for (int i = 0; i < n; ++i) { d[2 * i] = a[size_t(i)].real(); d[2 * i + 1] = a[size_t(i)].imag(); }
}

void ref_fft2d(float* d, int w, int h, int inverse) {
    std::vector<Complex> a(size_t(w) * h);
    for (size_t i = 0; i < a.size(); ++i) a[i] = Complex(d[2 * i], d[2 * i + 1]);
    fft2d(a, w, h, inverse != 0);
    for (size_t i = 0; i < a.size(); ++i) { d[2 * i] = a[i].real(); d[2 * i + 1] = a[i].imag(); }
}

float ref_first_minimum_radius(const float* raw, int n) {
    std::vector<float> v(raw, raw + size_t(n) * n);
    return EstimateLensDiffFirstMinimumRadius(v, n);
}

void ref_resample_kernel(const float* raw, int n, float scale, int support, float* out) {
    std::vector<float> v(raw, raw + size_t(n) * n);
    copy_out(makeResampledKernel(v, n, scale, support).values, out);
}

int ref_shape_kernels(const ldh_params* p, const float* base, int k, float* full, float* core,
                      float* structure, int* out_size) {
    const LensDiffParams rp = to_ref(*p);
    LensDiffKernel b;
    b.size = k;
    b.values.assign(base, base + size_t(k) * k);
    std::vector<LensDiffKernel> bases{b};
    LensDiffPsfBankCache cache;
    FinalizeLensDiffPsfBankFromBaseKernels(rp, MakeLensDiffPsfBankKey(rp), {}, 0, {550.0f}, bases, cache);
    const LensDiffPsfBin& bin = cache.bins.front();
    copy_out(bin.full.values, full);
    copy_out(bin.core.values, core);
    copy_out(bin.structure.values, structure);
    *out_size = bin.full.size;
    return cache.supportRadiusPx;
}

void* ref_bank_build(const ldh_params* p) {
    auto* cache = new LensDiffPsfBankCache();
    EnsureLensDiffPsfBank(to_ref(*p), *cache);
    return cache;
}
int ref_bank_bin_count(const void* h) { return int(static_cast<const LensDiffPsfBankCache*>(h)->bins.size()); }
int ref_bank_support_radius(const void* h) { return static_cast<const LensDiffPsfBankCache*>(h)->supportRadiusPx; }
int ref_bank_kernel_size(const void* h, int bin) { return static_cast<const LensDiffPsfBankCache*>(h)->bins[size_t(bin)].full.size; }
float ref_bank_wavelength(const void* h, int bin) { return static_cast<const LensDiffPsfBankCache*>(h)->bins[size_t(bin)].wavelengthNm; }
void ref_bank_kernel(const void* h, int bin, int which, float* out) {
    const auto& b = static_cast<const LensDiffPsfBankCache*>(h)->bins[size_t(bin)];
    copy_out(which == LDH_KERNEL_CORE ? b.core.values : (which == LDH_KERNEL_STRUCTURE ? b.structure.values : b.full.values), out);
}
void ref_bank_free(void* h) { delete static_cast<LensDiffPsfBankCache*>(h); }

// n <= 9 bins; matrices out as [channel][bin] with stride n.
void ref_spectrum_config(const ldh_params* p, int n, const float* wl, float* nat, float* sty) {
    std::vector<LensDiffPsfBin> bins(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) bins[size_t(i)].wavelengthNm = wl[i];
    const LensDiffSpectrumConfig c = BuildLensDiffSpectrumConfig(to_ref(*p), bins);
    for (int ch = 0; ch < 3; ++ch)
        for (int i = 0; i < n; ++i) {
            nat[ch * n + i] = c.naturalMatrix[size_t(ch * kLensDiffMaxSpectralBins + i)];
            sty[ch * n + i] = c.styleMatrix[size_t(ch * kLensDiffMaxSpectralBins + i)];
        }
}

// bins: n planes of npx values; the config is rebuilt from the wavelengths.
void ref_map_bins(const ldh_params* p, int n, const float* wl, const float* bins, long npx, float* rgb) {
    std::vector<LensDiffPsfBin> b(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) b[size_t(i)].wavelengthNm = wl[i];
    const LensDiffParams rp = to_ref(*p);
    const LensDiffSpectrumConfig c = BuildLensDiffSpectrumConfig(rp, b);
    for (long i = 0; i < npx; ++i) {
        std::array<float, kLensDiffMaxSpectralBins> v{};
        for (int j = 0; j < n && j < kLensDiffMaxSpectralBins; ++j) v[size_t(j)] = bins[size_t(j) * size_t(npx) + size_t(i)];
        const auto o = MapLensDiffSpectralBins(v, rp, c);
        rgb[i * 3] = o[0]; rgb[i * 3 + 1] = o[1]; rgb[i * 3 + 2] = o[2];
    }
}

void ref_selection(const ldh_params* p, const float* rgba, int w, int h, float* mask, float* redist, float* driver) {
    const LensDiffParams rp = to_ref(*p);
    const RgbaImage img = image_from(rgba, w, h, rp.inputTransfer);
    const ScalarImage m = buildSelectionMask(img, rp);
    const float s = 1.0f - static_cast<float>(clampValue(rp.corePreserve, 0.0, 1.0));
    copy_out(m.pixels, mask);
    if (redist) copy_out(applyMask(img, m, s).pixels, redist);
    if (driver) copy_out(applyMaskScalarDriver(img, m, rp.extractionMode, s).pixels, driver);
}

void ref_convolve(const float* plane, int w, int h, const float* kernel, int k, float* out) {
    ScalarImage src(w, h);
    std::copy(plane, plane + size_t(w) * h, src.pixels.begin());
    LensDiffKernel ker;
    ker.size = k;
    ker.values.assign(kernel, kernel + size_t(k) * k);
    copy_out(convolveScalar(src, ker).pixels, out);
}

int ref_render(const ldh_params* p, const ldh_render_request* req, char* err, int err_len) {
    LensDiffRenderRequest r;
    r.renderWindow = {req->render_window.x1, req->render_window.y1, req->render_window.x2, req->render_window.y2};
    r.frameBounds = {req->frame_bounds.x1, req->frame_bounds.y1, req->frame_bounds.x2, req->frame_bounds.y2};
    r.src = to_ref(req->src);
    r.dst = to_ref(req->dst);
    r.frameShortSidePx = req->frame_short_side_px;
    LensDiffPsfBankCache cache;
    std::string e;
    const bool ok = RunLensDiffCpuReference(r, to_ref(*p), cache, &e);
    if (!ok && err && err_len > 0) { std::strncpy(err, e.c_str(), size_t(err_len - 1)); err[err_len - 1] = 0; }
    return ok ? 1 : 0;
}

// Whole-frame render of a packed RGBA image (src bounds = dst bounds = render window).
int ref_render_image(const ldh_params* p, const float* src, int w, int h, float* out) {
    ldh_render_request req;
    std::memset(&req, 0, sizeof(req));
    req.render_window = {0, 0, w, h};
    req.frame_bounds = {0, 0, w, h};
    req.src = {const_cast<float*>(src), int64_t(w) * 16, {0, 0, w, h}, 0, 0, 4, 4};
    req.dst = {out, int64_t(w) * 16, {0, 0, w, h}, 0, 0, 4, 4};
    req.frame_short_side_px = std::min(w, h);
    char err[64];
    return ref_render(p, &req, err, sizeof(err));
}

// Timed pieces of one frame, for the CPU baseline: one raw PSF + resample at (P, R, K) and one
// convolveScalar of a W x H plane with a K x K kernel. Wall seconds via steady_clock.
double ref_time_raw_psf(const ldh_params* p) {
    const LensDiffParams rp = to_ref(*p);
    const int pupil = GetLensDiffEffectivePupilResolution(rp.pupilResolution);
    const int maxr = std::max(4, ResolveLensDiffMaxKernelRadiusPx(rp));
    const int raw = ChooseLensDiffRawPsfSize(pupil, maxr);
    const std::vector<float> amp = BuildLensDiffPupilAmplitudeForPsf(rp);
    const std::vector<float> ph = BuildLensDiffPupilPhaseForPsf(rp);
    const auto t0 = std::chrono::steady_clock::now();
    const std::vector<float> psf = BuildLensDiffShiftedRawPsf(amp, ph, pupil, raw);
    const LensDiffKernel k = makeResampledKernel(psf, raw, 1.0f, maxr);
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count() + 0.0 * k.values[0];
}

double ref_time_convolve(int w, int h, int k) {
    ScalarImage src(w, h);
    for (size_t i = 0; i < src.pixels.size(); i += 997) src.pixels[i] = 1.0f;
    LensDiffKernel ker;
    ker.size = k;
    ker.values.assign(size_t(k) * k, 1.0f / float(size_t(k) * k));
    const auto t0 = std::chrono::steady_clock::now();
    const ScalarImage out = convolveScalar(src, ker);
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count() + 0.0 * out.pixels[0];
}

}  // extern "C"

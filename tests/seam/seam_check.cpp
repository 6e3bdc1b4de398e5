// Checks the plugin-side seam RunLensDiffHip (lensdiff_amd/plugin/LensDiffHip.cpp) as the plugin
// compiles it: against the plugin's LensDiffTypes.h / LensDiffCpuReference.h, linked to the
// plugin core and to liblensdiff_hip.so. Built by tests/seam/Makefile; run by
// tests/test_plugin_seam.py. Prints one "name result" line per case and exits non-zero when a
// case that does not depend on the device fails.
//
//   seam_check tags     the seam's early-exit tags, and what a well-formed request returns here
//   seam_check params   LensDiffParams -> ldh_params field mapping, sizes against the plugin's
//
// The seam's parameter translation lives in an anonymous namespace, so its translation unit is
// compiled into this program by inclusion.
#include "LensDiffHip.cpp"  // NOLINT: the seam under test

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

namespace {

int g_failures = 0;

void expect(const char* name, bool ok, const std::string& detail) {
    std::printf("%s %s%s%s\n", name, ok ? "ok" : "FAIL", detail.empty() ? "" : " ", detail.c_str());
    if (!ok) ++g_failures;
}

struct Frame {
    int width, height, components;
    std::vector<float> pixels;
    Frame(int w, int h, int c) : width(w), height(h), components(c), pixels(size_t(w) * h * c) {
        for (size_t i = 0; i < pixels.size(); ++i) pixels[i] = float((i * 37) % 101) / 25.0f;
    }
    LensDiffImageView view(int x0, int y0) {
        LensDiffImageView v;
        v.data = pixels.data();
        v.rowBytes = static_cast<std::ptrdiff_t>(width) * components * sizeof(float);
        v.bounds = {x0, y0, x0 + width, y0 + height};
        v.components = components;
        v.bytesPerComponent = sizeof(float);
        return v;
    }
};

LensDiffRenderRequest make_request(Frame& src, Frame& dst) {
    LensDiffRenderRequest r;
    r.src = src.view(0, 0);
    r.dst = dst.view(0, 0);
    r.renderWindow = r.dst.bounds;
    r.frameBounds = r.src.bounds;
    r.frameShortSidePx = src.height;
    return r;
}

LensDiffParams small_params() {
    LensDiffParams p;
    p.pupilResolution = 64;
    p.maxKernelRadiusPx = 1.0;
    p.frameShortSidePx = 24;
    p.threshold = 0.5;
    return p;
}

// RunLensDiffHip returns false with a tag and leaves the destination and the cache alone.
void tag_case(const char* name, const LensDiffRenderRequest& r, const char* want, const Frame& dst_before,
              const Frame& dst) {
    LensDiffPsfBankCache cache;
    std::string error;
    const bool ok = RunLensDiffHip(r, small_params(), cache, &error);
    const bool untouched = dst.pixels == dst_before.pixels && !cache.valid && cache.bins.empty();
    expect(name, !ok && error == want && untouched,
           "returned=" + std::to_string(ok) + " error=" + error + " untouched=" + std::to_string(untouched));
}

int run_tags() {
    Frame src(32, 24, 4), dst(32, 24, 4);
    const Frame before = dst;
    {
        LensDiffRenderRequest r = make_request(src, dst);
        r.src.data = nullptr;
        tag_case("null-source", r, "missing-source-or-destination-buffer", before, dst);
    }
    {
        LensDiffRenderRequest r = make_request(src, dst);
        r.dst.data = nullptr;
        tag_case("null-destination", r, "missing-source-or-destination-buffer", before, dst);
    }
    {
        LensDiffRenderRequest r = make_request(src, dst);
        r.src.bounds.x2 = r.src.bounds.x1;
        tag_case("empty-source-width", r, "invalid-source-bounds", before, dst);
    }
    {
        LensDiffRenderRequest r = make_request(src, dst);
        r.src.bounds.y2 = r.src.bounds.y1 - 3;
        tag_case("negative-source-height", r, "invalid-source-bounds", before, dst);
    }
    {
        // a well-formed request: without a HIP device the seam reports hip-device-unavailable and
        // writes nothing; with one it renders and fills the cache
        LensDiffRenderRequest r = make_request(src, dst);
        LensDiffPsfBankCache cache;
        std::string error;
        const bool ok = RunLensDiffHip(r, small_params(), cache, &error);
        const bool untouched = dst.pixels == before.pixels;
        std::printf("well-formed %s error=%s untouched=%d cache_valid=%d\n", ok ? "rendered" : "failed",
                    error.c_str(), int(untouched), int(cache.valid));
        if (!ok && !untouched) {
            expect("well-formed-failure-writes-nothing", false, "");
        }
        // a second call takes the same path (the shared context is not left half-made)
        std::string error2;
        LensDiffPsfBankCache cache2;
        const bool ok2 = RunLensDiffHip(r, small_params(), cache2, &error2);
        expect("well-formed-repeat", ok2 == ok && error2 == error, "error=" + error2);
    }
    return g_failures == 0 ? 0 : 1;
}

// Every LensDiffParams member the seam forwards gets a distinct value; each must arrive in the
// ldh_params field of the same name.
int run_params() {
    LensDiffParams p;
    double v = 0.0;
    auto next = [&v] { v += 0.3125; return v; };  // exact in binary, distinct per field
    p.extractionMode = LensDiffExtractionMode::Luma;
    p.apertureMode = LensDiffApertureMode::Snowflake;
    p.apodizationMode = LensDiffApodizationMode::Gaussian;
    p.spectralMode = LensDiffSpectralMode::Spectral9;
    p.spectrumStyle = LensDiffSpectrumStyle::WarmCool;
    p.lookMode = LensDiffLookMode::Physical;
    p.energyMode = LensDiffEnergyMode::Augment;
    p.debugView = LensDiffDebugView::ChromaticSplit;
    p.inputTransfer = LensDiffInputTransfer::DavinciIntermediate;
    p.resolutionAware = true;
    p.phaseMode = LensDiffPhaseMode::Enabled;
    p.chromaticAffectsLuma = true;
    p.bladeCount = 9;
    p.vaneCount = 5;
    p.pupilResolution = 384;
    p.frameShortSidePx = 1080;
    double* doubles[] = {
        &p.threshold, &p.softnessStops, &p.pointEmphasis, &p.corePreserve, &p.roundness, &p.rotationDeg,
        &p.centralObstruction, &p.vaneThickness, &p.diffractionScalePx, &p.maxKernelRadiusPx,
        &p.phaseDefocus, &p.phaseAstigmatism0, &p.phaseAstigmatism45, &p.phaseComaX, &p.phaseComaY,
        &p.phaseSpherical, &p.phaseTrefoilX, &p.phaseTrefoilY, &p.phaseSecondaryAstigmatism0,
        &p.phaseSecondaryAstigmatism45, &p.phaseQuadrafoil0, &p.phaseQuadrafoil45, &p.phaseSecondaryComaX,
        &p.phaseSecondaryComaY, &p.pupilDecenterX, &p.pupilDecenterY, &p.phaseFieldStrength,
        &p.phaseFieldEdgeBias, &p.phaseFieldDefocus, &p.phaseFieldAstigRadial, &p.phaseFieldAstigTangential,
        &p.phaseFieldComaRadial, &p.phaseFieldComaTangential, &p.phaseFieldSpherical,
        &p.phaseFieldTrefoilRadial, &p.phaseFieldTrefoilTangential, &p.phaseFieldSecondaryAstigRadial,
        &p.phaseFieldSecondaryAstigTangential, &p.phaseFieldQuadrafoilRadial,
        &p.phaseFieldQuadrafoilTangential, &p.phaseFieldSecondaryComaRadial,
        &p.phaseFieldSecondaryComaTangential, &p.spectrumForce, &p.spectrumSaturation, &p.chromaticFocus,
        &p.chromaticSpread, &p.creativeFringe, &p.scatterAmount, &p.scatterRadius, &p.effectGain,
        &p.coreCompensation, &p.anisotropyEmphasis, &p.coreGain, &p.structureGain, &p.coreShoulder,
        &p.structureShoulder,
    };
    for (double* d : doubles) *d = next();
    const ldh_params q = to_ldh(p);

#define SAME(ref, ldh) expect(#ref, double(p.ref) == double(q.ldh), std::to_string(double(q.ldh)))
#define SAME_ENUM(ref, ldh) expect(#ref, int(p.ref) == int(q.ldh), std::to_string(int(q.ldh)))
    SAME_ENUM(extractionMode, extraction_mode);
    SAME_ENUM(apertureMode, aperture_mode);
    SAME_ENUM(apodizationMode, apodization_mode);
    SAME_ENUM(spectralMode, spectral_mode);
    SAME_ENUM(spectrumStyle, spectrum_style);
    SAME_ENUM(lookMode, look_mode);
    SAME_ENUM(energyMode, energy_mode);
    SAME_ENUM(debugView, debug_view);
    SAME_ENUM(inputTransfer, input_transfer);
    SAME_ENUM(phaseMode, phase_mode);
    SAME(resolutionAware, resolution_aware);
    SAME(chromaticAffectsLuma, chromatic_affects_luma);
    SAME(bladeCount, blade_count);
    SAME(vaneCount, vane_count);
    SAME(pupilResolution, pupil_resolution);
    SAME(frameShortSidePx, frame_short_side_px);
    SAME(threshold, threshold);
    SAME(softnessStops, softness_stops);
    SAME(pointEmphasis, point_emphasis);
    SAME(corePreserve, core_preserve);
    SAME(roundness, roundness);
    SAME(rotationDeg, rotation_deg);
    SAME(centralObstruction, central_obstruction);
    SAME(vaneThickness, vane_thickness);
    SAME(diffractionScalePx, diffraction_scale_px);
    SAME(maxKernelRadiusPx, max_kernel_radius_px);
    SAME(phaseDefocus, phase_defocus);
    SAME(phaseAstigmatism0, phase_astigmatism0);
    SAME(phaseAstigmatism45, phase_astigmatism45);
    SAME(phaseComaX, phase_coma_x);
    SAME(phaseComaY, phase_coma_y);
    SAME(phaseSpherical, phase_spherical);
    SAME(phaseTrefoilX, phase_trefoil_x);
    SAME(phaseTrefoilY, phase_trefoil_y);
    SAME(phaseSecondaryAstigmatism0, phase_secondary_astigmatism0);
    SAME(phaseSecondaryAstigmatism45, phase_secondary_astigmatism45);
    SAME(phaseQuadrafoil0, phase_quadrafoil0);
    SAME(phaseQuadrafoil45, phase_quadrafoil45);
    SAME(phaseSecondaryComaX, phase_secondary_coma_x);
    SAME(phaseSecondaryComaY, phase_secondary_coma_y);
    SAME(pupilDecenterX, pupil_decenter_x);
    SAME(pupilDecenterY, pupil_decenter_y);
    SAME(phaseFieldStrength, phase_field_strength);
    SAME(phaseFieldEdgeBias, phase_field_edge_bias);
    SAME(phaseFieldDefocus, phase_field_defocus);
    SAME(phaseFieldAstigRadial, phase_field_astig_radial);
    SAME(phaseFieldAstigTangential, phase_field_astig_tangential);
    SAME(phaseFieldComaRadial, phase_field_coma_radial);
    SAME(phaseFieldComaTangential, phase_field_coma_tangential);
    SAME(phaseFieldSpherical, phase_field_spherical);
    SAME(phaseFieldTrefoilRadial, phase_field_trefoil_radial);
    SAME(phaseFieldTrefoilTangential, phase_field_trefoil_tangential);
    SAME(phaseFieldSecondaryAstigRadial, phase_field_secondary_astig_radial);
    SAME(phaseFieldSecondaryAstigTangential, phase_field_secondary_astig_tangential);
    SAME(phaseFieldQuadrafoilRadial, phase_field_quadrafoil_radial);
    SAME(phaseFieldQuadrafoilTangential, phase_field_quadrafoil_tangential);
    SAME(phaseFieldSecondaryComaRadial, phase_field_secondary_coma_radial);
    SAME(phaseFieldSecondaryComaTangential, phase_field_secondary_coma_tangential);
    SAME(spectrumForce, spectrum_force);
    SAME(spectrumSaturation, spectrum_saturation);
    SAME(chromaticFocus, chromatic_focus);
    SAME(chromaticSpread, chromatic_spread);
    SAME(creativeFringe, creative_fringe);
    SAME(scatterAmount, scatter_amount);
    SAME(scatterRadius, scatter_radius);
    SAME(effectGain, effect_gain);
    SAME(coreCompensation, core_compensation);
    SAME(anisotropyEmphasis, anisotropy_emphasis);
    SAME(coreGain, core_gain);
    SAME(structureGain, structure_gain);
    SAME(coreShoulder, core_shoulder);
    SAME(structureShoulder, structure_shoulder);
#undef SAME
#undef SAME_ENUM
    // the library's own extensions keep their defaults: no explicit wavelength list, no extra
    // Zernike terms
    ldh_params d;
    ldh_params_default(&d);
    bool ext_default = q.wavelength_count == d.wavelength_count && q.conv_mode == d.conv_mode &&
                       q.raw_psf_form == d.raw_psf_form && q.spectra_fp16 == d.spectra_fp16;
    for (int i = 0; i < 6; ++i) ext_default = ext_default && q.zernike_extra[i] == d.zernike_extra[i];
    expect("extensions-default", ext_default, "");

    // sizes: ldh_resolve_sizes on the translated parameters against the plugin's own resolvers
    // (LensDiffTypes.h:292-320, LensDiffCpuReference.cpp:104-115, :1710-1711)
    int mismatches = 0, cases = 0;
    const int pupils[] = {16, 63, 64, 65, 256, 384, 512, 1000, 1024, 1025, 4096};
    const double radii[] = {-2.0, 0.0, 0.05, 1.0, 2.963, 9.5, 47.38, 94.7, 250.0};
    const int shorts[] = {0, 1, 24, 720, 1080, 2160, 4320};
    for (int pupil : pupils)
        for (double rad : radii)
            for (int s : shorts)
                for (int aware = 0; aware < 2; ++aware) {
                    LensDiffParams r;
                    r.pupilResolution = pupil;
                    r.maxKernelRadiusPx = rad;
                    r.frameShortSidePx = s;
                    r.resolutionAware = aware != 0;
                    r.diffractionScalePx = rad * 0.37;
                    const ldh_params rq = to_ldh(r);
                    ldh_sizes got;
                    if (ldh_resolve_sizes(&rq, &got) != LDH_OK) {
                        ++mismatches;
                        continue;
                    }
                    const int pupil_size = GetLensDiffEffectivePupilResolution(r.pupilResolution);
                    const int max_radius = std::max(4, ResolveLensDiffMaxKernelRadiusPx(r));
                    const int raw = ChooseLensDiffRawPsfSize(pupil_size, max_radius);
                    const bool same = got.pupil_size == pupil_size && got.max_kernel_radius_px == max_radius &&
                                      got.raw_psf_size == raw &&
                                      got.optics_short_side_px == ResolveLensDiffOpticsShortSidePx(r) &&
                                      got.diffraction_scale_px == ResolveLensDiffDiffractionScalePx(r);
                    if (!same) {
                        ++mismatches;
                        std::printf("sizes-mismatch pupil=%d radius=%g short=%d aware=%d\n", pupil, rad, s, aware);
                    }
                    ++cases;
                }
    expect("sizes", mismatches == 0, std::to_string(cases) + " cases, " + std::to_string(mismatches) + " differ");

    // the bank key the seam caches under is the plugin's (MakeLensDiffPsfBankKey): equal
    // parameters give equal keys, a bank-relevant change gives a different key
    LensDiffParams a = small_params(), b = small_params();
    const bool equal_keys = MakeLensDiffPsfBankKey(a) == MakeLensDiffPsfBankKey(b);
    b.bladeCount = a.bladeCount + 1;
    expect("bank-key", equal_keys && !(MakeLensDiffPsfBankKey(a) == MakeLensDiffPsfBankKey(b)), "");
    return g_failures == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string what = argc > 1 ? argv[1] : "tags";
    if (what == "tags") return run_tags();
    if (what == "params") return run_params();
    std::fprintf(stderr, "usage: seam_check tags|params\n");
    return 2;
}

// RunLensDiffHip: the plugin-side adapter from the plugin's C++ render types to the C ABI of
// liblensdiff_hip.so (include/lensdiff_hip.h). Host views are passed through unchanged: OFX has
// no HIP render suite (ofxGPURender.h:541-766), so the library stages host <-> device itself.
//
// The instance cache (LensDiffPsfBankCache) keeps the reference's meaning: after a render it
// holds the bank for the current key (bins of full/core/structure kernels), downloaded from the
// device only when the key changed (MakeLensDiffPsfBankKey, LensDiffCpuReference.cpp:1836-1875).
#include "LensDiffHip.h"

#include <cstring>
#include <mutex>
#include <vector>

#include "LensDiffCpuReference.h"
#include "lensdiff_hip.h"

namespace {

ldh_params to_ldh(const LensDiffParams& p) {
    ldh_params q;
    ldh_params_default(&q);
    q.extraction_mode = static_cast<int32_t>(p.extractionMode);
    q.aperture_mode = static_cast<int32_t>(p.apertureMode);
    q.apodization_mode = static_cast<int32_t>(p.apodizationMode);
    q.spectral_mode = static_cast<int32_t>(p.spectralMode);
    q.spectrum_style = static_cast<int32_t>(p.spectrumStyle);
    q.look_mode = static_cast<int32_t>(p.lookMode);
    q.energy_mode = static_cast<int32_t>(p.energyMode);
    q.debug_view = static_cast<int32_t>(p.debugView);
    q.input_transfer = static_cast<int32_t>(p.inputTransfer);
    q.resolution_aware = p.resolutionAware ? 1 : 0;
    q.threshold = p.threshold;
    q.softness_stops = p.softnessStops;
    q.point_emphasis = p.pointEmphasis;
    q.core_preserve = p.corePreserve;
    q.blade_count = p.bladeCount;
    q.vane_count = p.vaneCount;
    q.pupil_resolution = p.pupilResolution;
    q.frame_short_side_px = p.frameShortSidePx;
    q.roundness = p.roundness;
    q.rotation_deg = p.rotationDeg;
    q.central_obstruction = p.centralObstruction;
    q.vane_thickness = p.vaneThickness;
    q.diffraction_scale_px = p.diffractionScalePx;
    q.max_kernel_radius_px = p.maxKernelRadiusPx;
    q.phase_mode = static_cast<int32_t>(p.phaseMode);
    q.chromatic_affects_luma = p.chromaticAffectsLuma ? 1 : 0;
    q.phase_defocus = p.phaseDefocus;
    q.phase_astigmatism0 = p.phaseAstigmatism0;
    q.phase_astigmatism45 = p.phaseAstigmatism45;
    q.phase_coma_x = p.phaseComaX;
    q.phase_coma_y = p.phaseComaY;
    q.phase_spherical = p.phaseSpherical;
    q.phase_trefoil_x = p.phaseTrefoilX;
    q.phase_trefoil_y = p.phaseTrefoilY;
    q.phase_secondary_astigmatism0 = p.phaseSecondaryAstigmatism0;
    q.phase_secondary_astigmatism45 = p.phaseSecondaryAstigmatism45;
    q.phase_quadrafoil0 = p.phaseQuadrafoil0;
    q.phase_quadrafoil45 = p.phaseQuadrafoil45;
    q.phase_secondary_coma_x = p.phaseSecondaryComaX;
    q.phase_secondary_coma_y = p.phaseSecondaryComaY;
    q.pupil_decenter_x = p.pupilDecenterX;
    q.pupil_decenter_y = p.pupilDecenterY;
    q.phase_field_strength = p.phaseFieldStrength;
    q.phase_field_edge_bias = p.phaseFieldEdgeBias;
    q.phase_field_defocus = p.phaseFieldDefocus;
    q.phase_field_astig_radial = p.phaseFieldAstigRadial;
    q.phase_field_astig_tangential = p.phaseFieldAstigTangential;
    q.phase_field_coma_radial = p.phaseFieldComaRadial;
    q.phase_field_coma_tangential = p.phaseFieldComaTangential;
    q.phase_field_spherical = p.phaseFieldSpherical;
    q.phase_field_trefoil_radial = p.phaseFieldTrefoilRadial;
    q.phase_field_trefoil_tangential = p.phaseFieldTrefoilTangential;
    q.phase_field_secondary_astig_radial = p.phaseFieldSecondaryAstigRadial;
    q.phase_field_secondary_astig_tangential = p.phaseFieldSecondaryAstigTangential;
    q.phase_field_quadrafoil_radial = p.phaseFieldQuadrafoilRadial;
    q.phase_field_quadrafoil_tangential = p.phaseFieldQuadrafoilTangential;
    q.phase_field_secondary_coma_radial = p.phaseFieldSecondaryComaRadial;
    q.phase_field_secondary_coma_tangential = p.phaseFieldSecondaryComaTangential;
    q.spectrum_force = p.spectrumForce;
    q.spectrum_saturation = p.spectrumSaturation;
    q.chromatic_focus = p.chromaticFocus;
    q.chromatic_spread = p.chromaticSpread;
    q.creative_fringe = p.creativeFringe;
    q.scatter_amount = p.scatterAmount;
    q.scatter_radius = p.scatterRadius;
    q.effect_gain = p.effectGain;
    q.core_compensation = p.coreCompensation;
    q.anisotropy_emphasis = p.anisotropyEmphasis;
    q.core_gain = p.coreGain;
    q.structure_gain = p.structureGain;
    q.core_shoulder = p.coreShoulder;
    q.structure_shoulder = p.structureShoulder;
    return q;
}

ldh_image_view to_ldh(const LensDiffImageView& v) {
    ldh_image_view o;
    o.data = v.data;
    o.row_bytes = static_cast<int64_t>(v.rowBytes);
    o.bounds = {v.bounds.x1, v.bounds.y1, v.bounds.x2, v.bounds.y2};
    o.origin_x = v.originX;
    o.origin_y = v.originY;
    o.components = v.components;
    o.bytes_per_component = v.bytesPerComponent;
    return o;
}

// One context per process on device 0, created on first use (the plugin renders one instance
// at a time per its cache mutex, LensDiff.cpp:3821; the context serialises concurrent calls).
ldh_context* shared_context(std::string* error) {
    static std::mutex mu;
    static ldh_context* ctx = nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ctx && ldh_context_create(0, 0, &ctx) != LDH_OK) {
        if (error) *error = "hip-device-unavailable";
        ctx = nullptr;
    }
    return ctx;
}

bool fill_cache(ldh_context* ctx, const LensDiffParams& params, LensDiffPsfBankCache& cache, std::string* error) {
    const LensDiffPsfBankKey key = MakeLensDiffPsfBankKey(params);
    if (cache.valid && cache.key == key) return true;
    ldh_params q = to_ldh(params);
    // the cache's bins are the global bank (buildPsfBankGlobalOnly, LensDiffCpuReference.cpp:1769);
    // the zone banks of field-dependent phase stay inside the context
    q.phase_field_strength = 0.0;
    ldh_bank* bank = nullptr;
    if (ldh_bank_acquire(ctx, &q, &bank) != LDH_OK) {
        if (error) *error = ldh_last_error(ctx);
        return false;
    }
    ldh_bank_info info;
    bool ok = ldh_bank_info_get(ctx, bank, &info) == LDH_OK;
    LensDiffPsfBankCache fresh;
    fresh.valid = ok;
    fresh.key = key;
    fresh.supportRadiusPx = info.support_radius_px;
    for (int i = 0; ok && i < info.bin_count; ++i) {
        LensDiffPsfBin bin;
        bin.wavelengthNm = info.wavelengths_nm[i];
        LensDiffKernel* kernels[3] = {&bin.full, &bin.core, &bin.structure};
        for (int which = 0; ok && which < 3; ++which) {
            const int n = info.kernel_size[i];
            kernels[which]->size = n;
            kernels[which]->values.assign(static_cast<std::size_t>(n) * n, 0.0f);
            ok = ldh_bank_download(ctx, bank, i, which, kernels[which]->values.data(),
                                   static_cast<int64_t>(n) * n) == LDH_OK;
        }
        fresh.bins.push_back(std::move(bin));
    }
    ldh_bank_release(ctx, bank);
    if (!ok) {
        if (error) *error = ldh_last_error(ctx);
        return false;
    }
    cache = std::move(fresh);
    return true;
}

}  // namespace

bool RunLensDiffHip(const LensDiffRenderRequest& request,
                    const LensDiffParams& params,
                    LensDiffPsfBankCache& cache,
                    std::string* error) {
    if (request.src.data == nullptr || request.dst.data == nullptr) {
        if (error) *error = "missing-source-or-destination-buffer";
        return false;
    }
    if (request.src.bounds.width() <= 0 || request.src.bounds.height() <= 0) {
        if (error) *error = "invalid-source-bounds";
        return false;
    }
    ldh_context* ctx = shared_context(error);
    if (!ctx) return false;
    ldh_render_request r;
    std::memset(&r, 0, sizeof(r));
    r.render_window = {request.renderWindow.x1, request.renderWindow.y1, request.renderWindow.x2,
                       request.renderWindow.y2};
    r.frame_bounds = {request.frameBounds.x1, request.frameBounds.y1, request.frameBounds.x2,
                      request.frameBounds.y2};
    r.src = to_ldh(request.src);
    r.dst = to_ldh(request.dst);
    r.frame_short_side_px = request.frameShortSidePx;
    const ldh_params q = to_ldh(params);
    if (ldh_render(ctx, &r, &q) != LDH_OK) {
        if (error) *error = ldh_last_error(ctx);
        return false;
    }
    return fill_cache(ctx, params, cache, error);
}

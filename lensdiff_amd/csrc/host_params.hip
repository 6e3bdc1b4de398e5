// Host-side parameter resolution for the HIP backend.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <iterator>

#include "host_params.h"

namespace ldh {

static int optics_short_side(const ldh_params& p) {
    return p.resolution_aware ? 2160 : std::max(1, p.frame_short_side_px);
}
static double percent_to_px(double percent, int short_side) {
    return std::max(0.0, percent) * 0.01 * double(std::max(1, short_side));
}

Sizes resolve_sizes(const ldh_params& p) {
    Sizes s;
    s.pupil = std::max(64, std::min(p.pupil_resolution, 1024));
    s.optics_short_side = optics_short_side(p);
    const int r = std::max(0, int(std::ceil(percent_to_px(p.max_kernel_radius_px, s.optics_short_side))));
    s.max_radius = std::max(4, r);
    s.raw = std::max(128, std::min(std::max(s.pupil * 2, std::max(512, s.max_radius * 4)), 16384));
    s.diffraction_scale_px = percent_to_px(p.diffraction_scale_px, s.optics_short_side);
    return s;
}

std::vector<float> wavelength_list(const ldh_params& p) {
    if (p.wavelength_count > 0) {
        const int n = std::min(p.wavelength_count, LDH_MAX_WAVELENGTHS);
        return std::vector<float>(p.wavelengths_nm, p.wavelengths_nm + n);
    }
    switch (p.spectral_mode) {
        case LDH_SPECTRAL_9: return {420.f, 450.f, 480.f, 510.f, 540.f, 570.f, 600.f, 630.f, 660.f};
        case LDH_SPECTRAL_5: return {440.f, 490.f, 540.f, 590.f, 640.f};
        case LDH_SPECTRAL_TRISTIMULUS: return {460.f, 550.f, 610.f};
        default: return {550.f};
    }
}

bool is_spectral(const ldh_params& p) {
    return p.wavelength_count > 0 || p.spectral_mode != LDH_SPECTRAL_MONO;
}

static bool has_extra_terms(const ldh_params& p) {
    for (double c : p.zernike_extra)
        if (std::abs(c) > 1e-6) return true;
    return false;
}

// HasLensDiffNonFlatPhase (LensDiffPhase.cpp:20-38), extended by the extra Zernike terms.
bool params_nonflat_phase(const ldh_params& p) {
    return p.phase_mode == LDH_PHASE_ENABLED &&
           (std::abs(p.phase_defocus) > 1e-6 || std::abs(p.phase_astigmatism0) > 1e-6 ||
            std::abs(p.phase_astigmatism45) > 1e-6 || std::abs(p.phase_coma_x) > 1e-6 ||
            std::abs(p.phase_coma_y) > 1e-6 || std::abs(p.phase_spherical) > 1e-6 ||
            std::abs(p.phase_trefoil_x) > 1e-6 || std::abs(p.phase_trefoil_y) > 1e-6 ||
            std::abs(p.phase_secondary_astigmatism0) > 1e-6 || std::abs(p.phase_secondary_astigmatism45) > 1e-6 ||
            std::abs(p.phase_quadrafoil0) > 1e-6 || std::abs(p.phase_quadrafoil45) > 1e-6 ||
            std::abs(p.phase_secondary_coma_x) > 1e-6 || std::abs(p.phase_secondary_coma_y) > 1e-6 ||
            std::abs(p.pupil_decenter_x) > 1e-6 || std::abs(p.pupil_decenter_y) > 1e-6 || has_extra_terms(p));
}

PupilGeom make_pupil_geom(const ldh_params& p) {
    PupilGeom g{};
    g.size = std::max(64, std::min(p.pupil_resolution, 1024));
    g.mode = p.aperture_mode;
    g.blades = std::max(3, p.blade_count);
    g.vanes = p.vane_count;
    g.apodization = p.apodization_mode;
    g.obstruction = float(std::max(0.0, std::min(0.95, p.central_obstruction))) * 0.86f;
    g.rotation = float(p.rotation_deg * double(kPiF) / 180.0);
    g.roundness = float(std::max(0.0, std::min(p.roundness, 1.0)));
    g.vane_thickness = float(std::max(0.0, p.vane_thickness));
    g.star_inner = 0.18f + g.roundness * 0.62f;
    g.decenter_x = float(p.pupil_decenter_x);
    g.decenter_y = float(p.pupil_decenter_y);
    return g;
}

PhaseCoef make_phase_coef(const ldh_params& p) {
    PhaseCoef c{};
    c.enabled = params_nonflat_phase(p) ? 1 : 0;
    c.has_extra = has_extra_terms(p) ? 1 : 0;
    const double v[14] = {p.phase_defocus, p.phase_astigmatism0, p.phase_astigmatism45, p.phase_coma_x,
                          p.phase_coma_y, p.phase_spherical, p.phase_trefoil_x, p.phase_trefoil_y,
                          p.phase_secondary_astigmatism0, p.phase_secondary_astigmatism45, p.phase_quadrafoil0,
                          p.phase_quadrafoil45, p.phase_secondary_coma_x, p.phase_secondary_coma_y};
    for (int i = 0; i < 14; ++i) c.c[i] = float(v[i]);
    for (int i = 0; i < 6; ++i) c.extra[i] = float(p.zernike_extra[i]);
    c.decenter_x = float(p.pupil_decenter_x);
    c.decenter_y = float(p.pupil_decenter_y);
    const float rot = float(p.rotation_deg * double(kPiF) / 180.0);
    c.cos_rot = std::cos(-rot);
    c.sin_rot = std::sin(-rot);
    return c;
}

BankKey make_bank_key(const ldh_params& p) {
    const bool custom = p.aperture_mode == LDH_APERTURE_CUSTOM;
    const bool phase = params_nonflat_phase(p);
    const Sizes sz = resolve_sizes(p);
    BankKey k;
    auto add = [&](double v) { k.values.push_back(v); };
    add(p.aperture_mode); add(p.apodization_mode); add(p.spectral_mode);
    add(custom ? 0 : p.blade_count); add(custom ? 0 : p.vane_count); add(sz.pupil);
    add(sz.optics_short_side); add(p.max_kernel_radius_px);
    add(custom ? 0.0 : p.roundness); add(p.rotation_deg); add(custom ? 0.0 : p.central_obstruction);
    add(custom ? 0.0 : p.vane_thickness); add(p.diffraction_scale_px); add(p.anisotropy_emphasis);
    const double ph[16] = {p.phase_defocus, p.phase_astigmatism0, p.phase_astigmatism45, p.phase_coma_x,
                           p.phase_coma_y, p.phase_spherical, p.phase_trefoil_x, p.phase_trefoil_y,
                           p.phase_secondary_astigmatism0, p.phase_secondary_astigmatism45, p.phase_quadrafoil0,
                           p.phase_quadrafoil45, p.phase_secondary_coma_x, p.phase_secondary_coma_y,
                           p.pupil_decenter_x, p.pupil_decenter_y};
    for (double v : ph) add(phase ? v : 0.0);
    add(p.chromatic_focus); add(p.chromatic_spread);
    for (double v : p.zernike_extra) add(phase ? v : 0.0);
    const std::vector<float> wl = wavelength_list(p);
    add(double(wl.size()));
    for (float v : wl) add(double(v));
    add(p.raw_psf_form);  // the two forms round differently
    return k;
}

double working_scale(const ldh_params& p) {
    return double(optics_short_side(p)) / double(std::max(1, p.frame_short_side_px));
}
bool resolution_aware_active(const ldh_params& p) {
    return p.resolution_aware && std::abs(working_scale(p) - 1.0) > 1e-6;
}

double scatter_radius_px(const ldh_params& p) { return percent_to_px(p.scatter_radius, optics_short_side(p)); }
double creative_fringe_px(const ldh_params& p) { return percent_to_px(p.creative_fringe, optics_short_side(p)); }

bool has_field_phase(const ldh_params& p) {
    if (p.phase_mode != LDH_PHASE_ENABLED || !(std::abs(p.phase_field_strength) > 1e-6)) return false;
    const double terms[14] = {p.phase_field_defocus, p.phase_field_astig_radial, p.phase_field_astig_tangential,
                              p.phase_field_coma_radial, p.phase_field_coma_tangential, p.phase_field_spherical,
                              p.phase_field_trefoil_radial, p.phase_field_trefoil_tangential,
                              p.phase_field_secondary_astig_radial, p.phase_field_secondary_astig_tangential,
                              p.phase_field_quadrafoil_radial, p.phase_field_quadrafoil_tangential,
                              p.phase_field_secondary_coma_radial, p.phase_field_secondary_coma_tangential};
    return std::any_of(std::begin(terms), std::end(terms), [](double v) { return std::abs(v) > 1e-6; });
}

ldh_params field_zone_params(const ldh_params& p, int zx, int zy) {
    ldh_params z = p;
    if (!has_field_phase(p)) return z;
    const float nx = float(zx - 1), ny = float(zy - 1);
    const double radial = std::min(1.0, std::sqrt(double(nx) * nx + double(ny) * ny) / std::sqrt(2.0));
    const double bias = std::min(std::max(p.phase_field_edge_bias, 0.0), 1.0);
    const double scale = p.phase_field_strength * std::pow(radial, 1.0 + bias * 2.0);
    if (std::abs(scale) <= 1e-9) return z;
    const double angle = std::atan2(double(ny), double(nx));
    z.phase_defocus += p.phase_field_defocus * scale;
    z.phase_spherical += p.phase_field_spherical * scale;
    // (radial, tangential) field pairs rotated by harmonic * angle into the (x, y) pupil terms
    struct Pair {
        double radial, tangential;
        int harmonic;
        double *x, *y;
    };
    const Pair pairs[6] = {
        {p.phase_field_astig_radial, p.phase_field_astig_tangential, 2, &z.phase_astigmatism0, &z.phase_astigmatism45},
        {p.phase_field_coma_radial, p.phase_field_coma_tangential, 1, &z.phase_coma_x, &z.phase_coma_y},
        {p.phase_field_trefoil_radial, p.phase_field_trefoil_tangential, 3, &z.phase_trefoil_x, &z.phase_trefoil_y},
        {p.phase_field_secondary_astig_radial, p.phase_field_secondary_astig_tangential, 2,
         &z.phase_secondary_astigmatism0, &z.phase_secondary_astigmatism45},
        {p.phase_field_quadrafoil_radial, p.phase_field_quadrafoil_tangential, 4, &z.phase_quadrafoil0,
         &z.phase_quadrafoil45},
        {p.phase_field_secondary_coma_radial, p.phase_field_secondary_coma_tangential, 1, &z.phase_secondary_coma_x,
         &z.phase_secondary_coma_y},
    };
    for (const Pair& q : pairs) {
        const double r = q.radial * scale, t = q.tangential * scale;
        const double a = double(q.harmonic) * angle;
        const double c = std::cos(a), s = std::sin(a);
        *q.x += r * c - t * s;
        *q.y += r * s + t * c;
    }
    return z;
}

std::string unsupported_reason(const ldh_params& p, bool allow_field_zones) {
    if (p.aperture_mode == LDH_APERTURE_CUSTOM) return "custom-aperture-not-supported";
    if (p.aperture_mode < 0 || p.aperture_mode > LDH_APERTURE_SNOWFLAKE) return "invalid-aperture-mode";
    if (!allow_field_zones && has_field_phase(p)) return "field-zones-not-supported";
    if (p.wavelength_count < 0 || p.wavelength_count > LDH_MAX_WAVELENGTHS) return "invalid-wavelength-count";
    if (p.raw_psf_form != LDH_RAW_PSF_AUTO && p.raw_psf_form != LDH_RAW_PSF_WHOLE) return "invalid-raw-psf-form";
    switch (p.debug_view) {
        case LDH_VIEW_FINAL: case LDH_VIEW_SELECTION: case LDH_VIEW_EFFECT: case LDH_VIEW_CORE:
        case LDH_VIEW_STRUCTURE: case LDH_VIEW_CREATIVE_FRINGE: case LDH_VIEW_SCATTER: case LDH_VIEW_PUPIL:
        case LDH_VIEW_PSF: case LDH_VIEW_OTF: case LDH_VIEW_PHASE: case LDH_VIEW_PHASE_EDGE: case LDH_VIEW_FIELD_PSF:
        case LDH_VIEW_CHROMATIC_SPLIT: break;
        default: return "debug-view-not-supported";
    }
    return "";
}

}  // namespace ldh

extern "C" void ldh_params_default(ldh_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->extraction_mode = LDH_EXTRACT_MAX_RGB;
    p->aperture_mode = LDH_APERTURE_CIRCLE;
    p->apodization_mode = LDH_APOD_FLAT;
    p->spectral_mode = LDH_SPECTRAL_MONO;
    p->spectrum_style = LDH_STYLE_NATURAL;
    p->look_mode = LDH_LOOK_SPLIT;
    p->energy_mode = LDH_ENERGY_PRESERVE;
    p->debug_view = LDH_VIEW_FINAL;
    p->input_transfer = LDH_TRANSFER_LINEAR;
    p->threshold = 4.0;
    p->softness_stops = 1.0;
    p->core_preserve = 0.1;
    p->blade_count = 6;
    p->roundness = 0.5;
    p->vane_thickness = 0.01;
    p->diffraction_scale_px = 24.0;
    p->pupil_resolution = 256;
    p->max_kernel_radius_px = 64.0;
    p->frame_short_side_px = 2160;
    p->phase_mode = LDH_PHASE_OFF;
    p->spectrum_saturation = 1.0;
    p->effect_gain = 1.0;
    p->core_compensation = 1.0;
    p->core_gain = 1.0;
    p->structure_gain = 1.0;
}

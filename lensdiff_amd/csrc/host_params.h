// Host-side parameter resolution (LensDiffTypes.h:292-328, LensDiffCpuReference.cpp:104-115,
// :1602-1614, :1836-1875) and the bank cache key.
#pragma once

#include <string>
#include <vector>

#include "ldh_internal.h"
#include "pupil.h"

namespace ldh {

struct Sizes {
    int pupil = 0;        // P
    int raw = 0;          // R
    int max_radius = 0;   // max(4, ResolveLensDiffMaxKernelRadiusPx)
    int optics_short_side = 0;
    double diffraction_scale_px = 0.0;
};

Sizes resolve_sizes(const ldh_params& p);
std::vector<float> wavelength_list(const ldh_params& p);
bool is_spectral(const ldh_params& p);

// Identity of a bank: every field MakeLensDiffPsfBankKey keeps, with the same zeroing rules,
// plus this backend's extensions (wavelength list, extra Zernike terms). Compared bytewise.
struct BankKey {
    std::vector<double> values;
    bool operator==(const BankKey& o) const { return values == o.values; }
};
BankKey make_bank_key(const ldh_params& p);

// Rejects parameters outside this path with a kebab-case tag, or returns "". Field-dependent
// phase is accepted only where the caller renders field zones (allow_field_zones).
std::string unsupported_reason(const ldh_params& p, bool allow_field_zones = false);

// Effect working scale optics / frame short side (ResolveLensDiffEffectWorkingScale,
// LensDiffTypes.h:303-306); a resolution-aware render resamples the frame by it when it is away
// from 1.
double working_scale(const ldh_params& p);
bool resolution_aware_active(const ldh_params& p);

// Scatter radius and creative-fringe shift in pixels (ResolveLensDiffScatterRadiusPx,
// ResolveLensDiffCreativeFringePx, LensDiffTypes.h:322-328).
double scatter_radius_px(const ldh_params& p);
double creative_fringe_px(const ldh_params& p);

// Field-dependent phase (HasLensDiffFieldPhase, LensDiffPhase.cpp:40-57): the frame is rendered
// with one bank per zone of a 3x3 grid over the frame.
constexpr int kFieldZones = 9;
bool has_field_phase(const ldh_params& p);
// The parameters of zone (zx, zy), zx, zy in 0..2 (ResolveLensDiffFieldZoneParams at normalised
// position (zx - 1, zy - 1), LensDiffPhase.cpp:98-168).
ldh_params field_zone_params(const ldh_params& p, int zx, int zy);

}  // namespace ldh

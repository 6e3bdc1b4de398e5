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

// Rejects parameters outside this path with a kebab-case tag, or returns "".
std::string unsupported_reason(const ldh_params& p);

}  // namespace ldh

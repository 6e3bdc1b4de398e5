// CPU oracle for the LensDiff diffraction path — TEST INFRASTRUCTURE ONLY.
//
// A restatement of the reference CPU path, written from its behaviour:
//   source/src/core/LensDiffCpuReference.cpp  (pupil, FFT, raw PSF, bank, convolution, render)
//   source/src/core/LensDiffPhase.cpp         (Zernike phase)
//   source/src/core/LensDiffSpectrum.cpp      (spectral bin -> RGB)
//   source/src/core/LensDiffTransfer.cpp      (DaVinci Intermediate encode)
//   common/color/ColorManagement.cpp          (CIE 1931 CMFs, Rec.709 matrix, decode)
// Each function names the reference lines it follows. The FP32 instantiation keeps the
// reference's float operation order (that order is what the golden vectors pin); the FP64
// instantiation runs the same algorithm in double and is the FP64 oracle.
//
// Extensions beyond the reference (explicit wavelength lists longer than 9, Noll 2/3/18-21)
// follow the closed forms stated where they are implemented.
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use this library.

#include "lensdiff_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

namespace orc {

template <typename T> constexpr T kPi = T(3.14159265358979323846);
template <typename T> constexpr T kGray18 = T(0.18);

template <typename T> T clampv(T v, T lo, T hi) { return std::max(lo, std::min(v, hi)); }
template <typename T> T sat01(T v) { return clampv(v, T(0), T(1)); }
template <typename T> T luma709(T r, T g, T b) { return T(0.2126) * r + T(0.7152) * g + T(0.0722) * b; }
// LensDiffCpuReference.cpp:75-78
template <typename T> T smooth01(T t) {
    const T x = sat01(t);
    return x * x * (T(3) - T(2) * x);
}

// ---- sizes (LensDiffCpuReference.cpp:92-115, LensDiffTypes.h:292-320) --------------------
inline int pow2_at_least(int v) {
    int n = 1;
    while (n < v) n <<= 1;
    return n;
}
inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
inline int pupil_size_of(int requested) { return clampv(requested, 64, 1024); }
template <typename T> T spectrum_center(int size) {
    return (size & 1) == 0 ? T(size) * T(0.5) : T(size - 1) * T(0.5);
}
inline int raw_size_of(int pupil, int max_radius) {
    return clampv(std::max(pupil * 2, std::max(512, max_radius * 4)), 128, 16384);
}
inline int optics_short_side(const ldh_params& p) {
    return p.resolution_aware ? 2160 : std::max(1, p.frame_short_side_px);
}
inline double percent_to_px(double percent, int short_side) {
    return std::max(0.0, percent) * 0.01 * double(std::max(1, short_side));
}
inline int max_kernel_radius_px(const ldh_params& p) {
    return std::max(0, int(std::ceil(percent_to_px(p.max_kernel_radius_px, optics_short_side(p)))));
}
inline double diffraction_scale_px(const ldh_params& p) {
    return percent_to_px(p.diffraction_scale_px, optics_short_side(p));
}

// ---- sequential accumulation (std::accumulate with a T accumulator, :436-449) -------------
template <typename T> T seq_sum(const std::vector<T>& v) {
    T s = T(0);
    for (T x : v) s += x;
    return s;
}
template <typename T> void normalize_unit(std::vector<T>& v) {
    const T total = seq_sum(v);
    if (total <= T(0)) return;
    const T inv = T(1) / total;
    for (T& x : v) x *= inv;
}

// ---- aperture geometry (LensDiffCpuReference.cpp:451-680) --------------------------------
template <typename T>
T polygon_metric(T nx, T ny, int blades, T roundness, T rot, T outer) {
    const T radius = std::sqrt(nx * nx + ny * ny);
    const T circle = radius / std::max(outer, T(1e-5));
    if (blades < 3) return circle;
    const T angle = std::atan2(ny, nx) - rot;
    const T sector = T(2) * kPi<T> / T(blades);
    T local = std::fmod(angle, sector);
    if (local < T(0)) local += sector;
    local -= sector * T(0.5);
    const T poly_radius = outer * std::cos(kPi<T> / T(blades)) / std::max(std::cos(local), T(1e-4));
    const T poly = radius / std::max(poly_radius, T(1e-5));
    return circle * roundness + poly * (T(1) - roundness);
}

template <typename T>
bool vane_blocks(T nx, T ny, int vanes, T thickness, T rot, T outer) {
    if (vanes <= 0 || thickness <= T(0)) return false;
    const T half = thickness * outer;
    const int lines = std::max(1, vanes);
    for (int i = 0; i < lines; ++i) {
        const T a = rot + T(i) * kPi<T> / T(lines);
        const T cs = std::cos(a), sn = std::sin(a);
        const T along = nx * cs + ny * sn;
        const T across = -nx * sn + ny * cs;
        if (std::abs(across) <= half && std::abs(along) <= outer) return true;
    }
    return false;
}

template <typename T>
T star_metric(T nx, T ny, int points, T inner_ratio, T rot, T outer) {
    const T radius = std::sqrt(nx * nx + ny * ny);
    if (radius <= T(1e-6)) return T(0);
    const int n = std::max(3, points);
    const T inner = outer * clampv(inner_ratio, T(0.1), T(0.95));
    const T angle = std::atan2(ny, nx) - rot + kPi<T> * T(0.5);
    const T sector = T(2) * kPi<T> / T(n);
    T local = std::fmod(angle, sector);
    if (local < T(0)) local += sector;
    const T half = sector * T(0.5);
    const T t = local <= half ? local / std::max(half, T(1e-6)) : (sector - local) / std::max(half, T(1e-6));
    const T boundary = inner + (outer - inner) * clampv(t, T(0), T(1));
    return radius / std::max(boundary, T(1e-5));
}

template <typename T>
T segment_distance(T px, T py, T ax, T ay, T bx, T by) {
    const T vx = bx - ax, vy = by - ay;
    const T len2 = vx * vx + vy * vy;
    if (len2 <= T(1e-10)) {
        const T dx = px - ax, dy = py - ay;
        return std::sqrt(dx * dx + dy * dy);
    }
    const T t = clampv(((px - ax) * vx + (py - ay) * vy) / len2, T(0), T(1));
    const T cx = ax + t * vx, cy = ay + t * vy;
    const T dx = px - cx, dy = py - cy;
    return std::sqrt(dx * dx + dy * dy);
}

template <typename T>
bool square_grid_open(T nx, T ny, int blades, T roundness, T rot, T outer) {
    const T cs = std::cos(-rot), sn = std::sin(-rot);
    const T rx = nx * cs - ny * sn;
    const T ry = nx * sn + ny * cs;
    const T half = outer * T(0.82);
    if (std::abs(rx) > half || std::abs(ry) > half) return false;
    const int div = std::max(3, blades);
    const T pitch = (T(2) * half) / T(div);
    const T bar = pitch * (T(0.06) + (T(1) - clampv(roundness, T(0), T(1))) * T(0.12));
    auto line_dist = [&](T v) {
        const T w = std::fmod(v + half, pitch);
        const T pos = w < T(0) ? w + pitch : w;
        return std::min(pos, pitch - pos);
    };
    const T dx = line_dist(rx), dy = line_dist(ry);
    const T ex = std::abs(std::abs(rx) - half), ey = std::abs(std::abs(ry) - half);
    return dx <= bar || dy <= bar || ex <= bar || ey <= bar;
}

template <typename T>
bool snowflake_open(T nx, T ny, int blades, T roundness, T rot, T outer) {
    const T radius = std::sqrt(nx * nx + ny * ny);
    if (radius > outer) return false;
    const int levels = clampv(std::max(2, blades / 2), 2, 5);
    const T main_len = outer * T(0.92);
    const T branch_angle = kPi<T> / T(5.5);
    const T thick = outer * (T(0.045) + (T(1) - clampv(roundness, T(0), T(1))) * T(0.05));
    T best = std::numeric_limits<T>::max();
    for (int arm = 0; arm < 6; ++arm) {
        const T a = rot - kPi<T> * T(0.5) + T(arm) * (kPi<T> / T(3));
        const T cs = std::cos(a), sn = std::sin(a);
        best = std::min(best, segment_distance(nx, ny, T(0), T(0), cs * main_len, sn * main_len));
        for (int lv = 0; lv < levels; ++lv) {
            const T t = T(0.34) + T(lv) * (T(0.44) / T(std::max(1, levels - 1)));
            const T mx = cs * main_len * t, my = sn * main_len * t;
            const T blen = outer * (T(0.16) + T(0.03) * T(lv));
            for (int side = -1; side <= 1; side += 2) {
                const T th = a + T(side) * branch_angle;
                best = std::min(best, segment_distance(nx, ny, mx, my, mx + std::cos(th) * blen,
                                                       my + std::sin(th) * blen));
            }
        }
    }
    return best <= thick;
}

template <typename T>
bool spiral_open(T nx, T ny, int blades, T roundness, T rot, T outer) {
    const T radius = std::sqrt(nx * nx + ny * ny);
    if (radius > outer) return false;
    const T rn = radius / std::max(outer, T(1e-6));
    const T angle = std::atan2(ny, nx) - rot;
    const T twist = T(4) + (T(1) - clampv(roundness, T(0), T(1))) * T(8);
    const T opening = T(0.18) + clampv(roundness, T(0), T(1)) * T(0.26);
    const T ph = T(std::max(3, blades)) * angle + twist * rn * (T(2) * kPi<T>);
    const T band = T(0.5) + T(0.5) * std::cos(ph);
    return band >= (T(1) - opening);
}

template <typename T> T apodization(int mode, T rn) {
    if (mode == LDH_APOD_COSINE) return std::cos(std::min(rn, T(1)) * (kPi<T> * T(0.5)));
    if (mode == LDH_APOD_GAUSSIAN) return std::exp(T(-4) * rn * rn);
    return T(1);
}

// buildPupilAmplitude (LensDiffCpuReference.cpp:758-822). Custom apertures are not part of
// this path (the image loader is unsupported on Linux); they yield a dark pupil.
template <typename T> std::vector<T> pupil_amplitude(const ldh_params& p) {
    const int n = pupil_size_of(p.pupil_resolution);
    std::vector<T> out(size_t(n) * n, T(0));
    if (p.aperture_mode == LDH_APERTURE_CUSTOM) return out;
    const T outer = T(0.86);
    const T obstruction = T(std::max(0.0, std::min(0.95, p.central_obstruction))) * outer;
    const T rot = T(p.rotation_deg * double(kPi<T>) / 180.0);
    const T roundness = T(clampv(p.roundness, 0.0, 1.0));
    const T vane_thickness = T(std::max(0.0, p.vane_thickness));
    const T star_inner = T(0.18) + roundness * T(0.62);
    const int mode = p.aperture_mode;
    const int blades = std::max(3, p.blade_count);
    for (int y = 0; y < n; ++y) {
        for (int x = 0; x < n; ++x) {
            const T nx = ((T(x) + T(0.5)) / T(n) - T(0.5)) * T(2);
            const T ny = ((T(y) + T(0.5)) / T(n) - T(0.5)) * T(2);
            const T dx = nx - T(p.pupil_decenter_x);
            const T dy = ny - T(p.pupil_decenter_y);
            const T radius = std::sqrt(dx * dx + dy * dy);
            bool inside;
            if (mode == LDH_APERTURE_POLYGON || mode == LDH_APERTURE_HEXAGON) {
                const int sides = mode == LDH_APERTURE_HEXAGON ? 6 : blades;
                inside = polygon_metric(dx, dy, sides, roundness, rot, outer) <= T(1);
            } else if (mode == LDH_APERTURE_SQUARE_GRID) {
                inside = square_grid_open(dx, dy, blades, roundness, rot, outer);
            } else if (mode == LDH_APERTURE_SNOWFLAKE) {
                inside = snowflake_open(dx, dy, blades, roundness, rot, outer);
            } else if (mode == LDH_APERTURE_STAR) {
                inside = star_metric(dx, dy, blades, star_inner, rot, outer) <= T(1);
            } else if (mode == LDH_APERTURE_SPIRAL) {
                inside = spiral_open(dx, dy, blades, roundness, rot, outer);
            } else {
                // Circle; the polygon metric is evaluated as the reference does (:807-808).
                const T metric = polygon_metric(dx, dy, blades, roundness, rot, outer);
                inside = mode == LDH_APERTURE_CIRCLE ? radius <= outer : metric <= T(1);
            }
            if (!inside || radius < obstruction) continue;
            if (vane_blocks(dx, dy, p.vane_count, vane_thickness, rot, outer)) continue;
            out[size_t(y) * n + x] = apodization(p.apodization_mode, radius / outer);
        }
    }
    return out;
}

// ---- Zernike phase (LensDiffPhase.cpp:9-16, :20-38, :170-243) -----------------------------
inline bool has_extra_terms(const ldh_params& p) {
    for (double c : p.zernike_extra) if (std::abs(c) > 1e-6) return true;
    return false;
}
inline bool nonflat_phase(const ldh_params& p) {
    return p.phase_mode == LDH_PHASE_ENABLED &&
           (std::abs(p.phase_defocus) > 1e-6 || std::abs(p.phase_astigmatism0) > 1e-6 ||
            std::abs(p.phase_astigmatism45) > 1e-6 || std::abs(p.phase_coma_x) > 1e-6 ||
            std::abs(p.phase_coma_y) > 1e-6 || std::abs(p.phase_spherical) > 1e-6 ||
            std::abs(p.phase_trefoil_x) > 1e-6 || std::abs(p.phase_trefoil_y) > 1e-6 ||
            std::abs(p.phase_secondary_astigmatism0) > 1e-6 ||
            std::abs(p.phase_secondary_astigmatism45) > 1e-6 ||
            std::abs(p.phase_quadrafoil0) > 1e-6 || std::abs(p.phase_quadrafoil45) > 1e-6 ||
            std::abs(p.phase_secondary_coma_x) > 1e-6 || std::abs(p.phase_secondary_coma_y) > 1e-6 ||
            std::abs(p.pupil_decenter_x) > 1e-6 || std::abs(p.pupil_decenter_y) > 1e-6 ||
            has_extra_terms(p));
}

// ---- Field zones (LensDiffPhase.cpp:40-57, :98-168) ---------------------------------------
// Field-dependent phase: the frame is rendered with nine banks, one per zone of a 3x3 grid.
inline bool field_phase(const ldh_params& p) {
    if (p.phase_mode != LDH_PHASE_ENABLED || std::abs(p.phase_field_strength) <= 1e-6) return false;
    const double f[14] = {p.phase_field_defocus, p.phase_field_astig_radial, p.phase_field_astig_tangential,
                          p.phase_field_coma_radial, p.phase_field_coma_tangential, p.phase_field_spherical,
                          p.phase_field_trefoil_radial, p.phase_field_trefoil_tangential,
                          p.phase_field_secondary_astig_radial, p.phase_field_secondary_astig_tangential,
                          p.phase_field_quadrafoil_radial, p.phase_field_quadrafoil_tangential,
                          p.phase_field_secondary_coma_radial, p.phase_field_secondary_coma_tangential};
    for (double v : f)
        if (std::abs(v) > 1e-6) return true;
    return false;
}

// A (radial, tangential) pair of field terms rotated by harmonic * angle onto the pupil's
// (x, y) pair of the same Zernike order.
inline void add_rotated(double radial, double tangential, double angle, int harmonic, double& x, double& y) {
    const double a = double(harmonic) * angle;
    const double c = std::cos(a), s = std::sin(a);
    x += radial * c - tangential * s;
    y += radial * s + tangential * c;
}

// Parameters of the zone at normalised field position (nx, ny) in {-1, 0, 1}^2: the global
// terms plus the field terms scaled by strength * r^(1 + 2 edge_bias), r = |(nx, ny)| / sqrt 2.
inline ldh_params field_zone_params(const ldh_params& p, float nx, float ny) {
    ldh_params z = p;
    if (!field_phase(p)) return z;
    const double r = std::min(1.0, std::sqrt(double(nx) * nx + double(ny) * ny) / std::sqrt(2.0));
    const double bias = clampv(p.phase_field_edge_bias, 0.0, 1.0);
    const double k = p.phase_field_strength * std::pow(r, 1.0 + bias * 2.0);
    if (std::abs(k) <= 1e-9) return z;
    const double angle = std::atan2(double(ny), double(nx));
    z.phase_defocus += p.phase_field_defocus * k;
    z.phase_spherical += p.phase_field_spherical * k;
    add_rotated(p.phase_field_astig_radial * k, p.phase_field_astig_tangential * k, angle, 2, z.phase_astigmatism0,
                z.phase_astigmatism45);
    add_rotated(p.phase_field_coma_radial * k, p.phase_field_coma_tangential * k, angle, 1, z.phase_coma_x,
                z.phase_coma_y);
    add_rotated(p.phase_field_trefoil_radial * k, p.phase_field_trefoil_tangential * k, angle, 3, z.phase_trefoil_x,
                z.phase_trefoil_y);
    add_rotated(p.phase_field_secondary_astig_radial * k, p.phase_field_secondary_astig_tangential * k, angle, 2,
                z.phase_secondary_astigmatism0, z.phase_secondary_astigmatism45);
    add_rotated(p.phase_field_quadrafoil_radial * k, p.phase_field_quadrafoil_tangential * k, angle, 4,
                z.phase_quadrafoil0, z.phase_quadrafoil45);
    add_rotated(p.phase_field_secondary_coma_radial * k, p.phase_field_secondary_coma_tangential * k, angle, 1,
                z.phase_secondary_coma_x, z.phase_secondary_coma_y);
    return z;
}

// Waves at normalised pupil point (px, py). The 14 reference terms are summed in the
// reference's order (LensDiffPhase.cpp:199-212); the extension terms (Noll 1976 normalisation,
// Z2 = 2x, Z3 = 2y, Z18/19 = sqrt(12)(5r^2-4) r^3 cos/sin 3t, Z20/21 = sqrt(12) r^5 cos/sin 5t)
// are added after them.
template <typename T> T phase_waves_at(const ldh_params& p, T px, T py) {
    if (!nonflat_phase(p)) return T(0);
    const T sx = px - T(p.pupil_decenter_x);
    const T sy = py - T(p.pupil_decenter_y);
    const T r2 = sx * sx + sy * sy;
    if (r2 > T(1)) return T(0);
    const T n3 = T(1.7320508075688772), n6 = T(2.4494897427831781), n8 = T(2.8284271247461901);
    const T n5 = T(2.2360679774997898), n10 = T(3.1622776601683795), n12 = T(3.4641016151377544);
    const T z4 = n3 * (T(2) * r2 - T(1));
    const T z6 = n6 * (sx * sx - sy * sy);
    const T z5 = n6 * (T(2) * sx * sy);
    const T z8 = n8 * ((T(3) * r2 - T(2)) * sx);
    const T z7 = n8 * ((T(3) * r2 - T(2)) * sy);
    const T z11 = n5 * (T(6) * r2 * r2 - T(6) * r2 + T(1));
    const T z10 = n8 * (sx * sx * sx - T(3) * sx * sy * sy);
    const T z9 = n8 * (T(3) * sx * sx * sy - sy * sy * sy);
    const T sa = (T(4) * r2 - T(3));
    const T z12 = n10 * sa * (sx * sx - sy * sy);
    const T z13 = n10 * sa * (T(2) * sx * sy);
    const T z14 = n10 * (sx * sx * sx * sx - T(6) * sx * sx * sy * sy + sy * sy * sy * sy);
    const T z15 = n10 * (T(4) * sx * sy * (sx * sx - sy * sy));
    const T sc = (T(10) * r2 * r2 - T(12) * r2 + T(3));
    const T z16 = n12 * sc * sx;
    const T z17 = n12 * sc * sy;
    T w = T(p.phase_defocus) * z4 + T(p.phase_astigmatism0) * z6 + T(p.phase_astigmatism45) * z5 +
          T(p.phase_coma_x) * z8 + T(p.phase_coma_y) * z7 + T(p.phase_spherical) * z11 +
          T(p.phase_trefoil_x) * z10 + T(p.phase_trefoil_y) * z9 +
          T(p.phase_secondary_astigmatism0) * z12 + T(p.phase_secondary_astigmatism45) * z13 +
          T(p.phase_quadrafoil0) * z14 + T(p.phase_quadrafoil45) * z15 +
          T(p.phase_secondary_coma_x) * z16 + T(p.phase_secondary_coma_y) * z17;
    if (has_extra_terms(p)) {
        const T c3 = sx * sx * sx - T(3) * sx * sy * sy;       // r^3 cos 3t
        const T s3 = T(3) * sx * sx * sy - sy * sy * sy;       // r^3 sin 3t
        const T x2 = sx * sx, y2 = sy * sy;
        const T c5 = sx * (x2 * x2 - T(10) * x2 * y2 + T(5) * y2 * y2);   // r^5 cos 5t
        const T s5 = sy * (T(5) * x2 * x2 - T(10) * x2 * y2 + y2 * y2);   // r^5 sin 5t
        const T r53 = (T(5) * r2 - T(4));
        w += T(p.zernike_extra[0]) * (T(2) * sx) + T(p.zernike_extra[1]) * (T(2) * sy) +
             T(p.zernike_extra[2]) * (n12 * r53 * c3) + T(p.zernike_extra[3]) * (n12 * r53 * s3) +
             T(p.zernike_extra[4]) * (n12 * c5) + T(p.zernike_extra[5]) * (n12 * s5);
    }
    return w;
}

// BuildLensDiffPupilPhaseWaves (LensDiffPhase.cpp:215-243) with outerRadius 0.86 and the
// bank's rotation (LensDiffCpuReference.cpp:824-829).
template <typename T> std::vector<T> pupil_phase(const ldh_params& p) {
    const int n = pupil_size_of(p.pupil_resolution);
    std::vector<T> out(size_t(n) * n, T(0));
    const T outer = T(0.86);
    if (!nonflat_phase(p)) return out;
    const T rot = T(p.rotation_deg * double(kPi<T>) / 180.0);
    const T cs = std::cos(-rot), sn = std::sin(-rot);
    for (int y = 0; y < n; ++y) {
        for (int x = 0; x < n; ++x) {
            const T nx = ((T(x) + T(0.5)) / T(n) - T(0.5)) * T(2);
            const T ny = ((T(y) + T(0.5)) / T(n) - T(0.5)) * T(2);
            const T radius = std::sqrt(nx * nx + ny * ny);
            if (radius > outer) continue;
            const T rx = nx * cs - ny * sn;
            const T ry = nx * sn + ny * cs;
            out[size_t(y) * n + x] = phase_waves_at(p, rx / outer, ry / outer);
        }
    }
    return out;
}

// ---- FFT (LensDiffCpuReference.cpp:194-301) ------------------------------------------------
// Iterative radix-2 DIT after a bit-reversal permutation; the twiddle of each butterfly
// group is advanced by repeated multiplication (w *= w_len), as the reference does.
template <typename T> void fft_radix2(std::vector<std::complex<T>>& a, bool inverse) {
    using C = std::complex<T>;
    const int n = int(a.size());
    if (!is_pow2(n)) return;
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const T ang = T(2) * kPi<T> / T(len) * (inverse ? T(1) : T(-1));
        const C step(std::cos(ang), std::sin(ang));
        const int half = len / 2;
        for (int base = 0; base < n; base += len) {
            C w(T(1), T(0));
            for (int k = 0; k < half; ++k) {
                const C u = a[base + k];
                const C v = a[base + k + half] * w;
                a[base + k] = u + v;
                a[base + k + half] = u - v;
                w *= step;
            }
        }
    }
    if (inverse) {
        const T s = T(1) / T(n);
        for (auto& z : a) z *= s;
    }
}

// Bluestein for non-power-of-two lengths: chirp from a double-precision phase using the
// float value of pi in the FP32 path (LensDiffCpuReference.cpp:244-274).
template <typename T> void fft_any(std::vector<std::complex<T>>& x, bool inverse) {
    using C = std::complex<T>;
    const int n = int(x.size());
    if (n <= 1) return;
    if (is_pow2(n)) { fft_radix2(x, inverse); return; }
    const int m = pow2_at_least(n * 2 - 1);
    std::vector<C> a(static_cast<size_t>(m), C(0, 0)), b(size_t(m), C(0, 0));
    const double sgn = inverse ? 1.0 : -1.0;
    const double pi = double(kPi<T>);
    for (int i = 0; i < n; ++i) {
        const double ph = sgn * pi * double(i) * double(i) / double(n);
        const C chirp(T(std::cos(ph)), T(std::sin(ph)));
        const C conj(T(std::cos(-ph)), T(std::sin(-ph)));
        a[size_t(i)] = x[size_t(i)] * chirp;
        b[size_t(i)] = conj;
        if (i != 0) b[size_t(m - i)] = conj;
    }
    fft_radix2(a, false);
    fft_radix2(b, false);
    for (int i = 0; i < m; ++i) a[size_t(i)] *= b[size_t(i)];
    fft_radix2(a, true);
    const T scale = inverse ? (T(1) / T(n)) : T(1);
    for (int i = 0; i < n; ++i) {
        const double ph = sgn * pi * double(i) * double(i) / double(n);
        const C chirp(T(std::cos(ph)), T(std::sin(ph)));
        x[size_t(i)] = a[size_t(i)] * chirp * scale;
    }
}

// Rows, then columns (LensDiffCpuReference.cpp:277-301). Every line is transformed on its own
// with the same arithmetic whichever thread takes it, so spreading the lines over OpenMP threads
// leaves the result bit for bit the reference's sequential one.
template <typename T> void fft_2d(std::vector<std::complex<T>>& d, int w, int h, bool inverse) {
#pragma omp parallel
    {
        std::vector<std::complex<T>> line(static_cast<size_t>(w));
#pragma omp for schedule(static)
        for (int y = 0; y < h; ++y) {
            for (int x = 0; x < w; ++x) line[size_t(x)] = d[size_t(y) * w + x];
            fft_any(line, inverse);
            for (int x = 0; x < w; ++x) d[size_t(y) * w + x] = line[size_t(x)];
        }
        line.resize(size_t(h));
#pragma omp for schedule(static)
        for (int x = 0; x < w; ++x) {
            for (int y = 0; y < h; ++y) line[size_t(y)] = d[size_t(y) * w + x];
            fft_any(line, inverse);
            for (int y = 0; y < h; ++y) d[size_t(y) * w + x] = line[size_t(y)];
        }
    }
}

// ---- raw PSF (LensDiffCpuReference.cpp:831-863, fftShiftSquare :303-314) -------------------
template <typename T>
std::vector<T> raw_psf(const std::vector<T>& amp, const T* phase, int pupil, int raw) {
    using C = std::complex<T>;
    std::vector<C> f(size_t(raw) * raw, C(0, 0));
    const int off = std::max(0, (raw - pupil) / 2);
    for (int y = 0; y < pupil; ++y) {
        for (int x = 0; x < pupil; ++x) {
            const size_t i = size_t(y) * pupil + x;
            const T a = amp[i];
            T rad = T(0);
            if (phase != nullptr && a > T(0)) rad = phase[i] * T(2) * kPi<T>;
            f[size_t(y + off) * raw + (x + off)] = C(a * std::cos(rad), a * std::sin(rad));
        }
    }
    fft_2d(f, raw, raw, false);
    std::vector<T> inten(size_t(raw) * raw);
    for (size_t i = 0; i < inten.size(); ++i) inten[i] = std::norm(f[i]);
    std::vector<T> shifted(inten.size());
    const int half = raw / 2;
    for (int y = 0; y < raw; ++y)
        for (int x = 0; x < raw; ++x)
            shifted[size_t(y) * raw + x] = inten[size_t((y + half) % raw) * raw + size_t((x + half) % raw)];
    normalize_unit(shifted);
    return shifted;
}

// ---- first minimum (LensDiffCpuReference.cpp:865-955) ------------------------------------
template <typename T> std::vector<T> radial_profile(const T* img, int size) {
    const int rmax = size / 2;
    std::vector<T> sums(static_cast<size_t>(rmax + 1), T(0));
    std::vector<int> counts(static_cast<size_t>(rmax + 1), 0);
    const T c = spectrum_center<T>(size);
    for (int y = 0; y < size; ++y)
        for (int x = 0; x < size; ++x) {
            const T dx = T(x) - c, dy = T(y) - c;
            const int r = std::min(rmax, int(std::round(std::sqrt(dx * dx + dy * dy))));
            sums[size_t(r)] += img[size_t(y) * size + x];
            counts[size_t(r)] += 1;
        }
    std::vector<T> prof(static_cast<size_t>(rmax + 1));
    for (int r = 0; r <= rmax; ++r) prof[size_t(r)] = sums[size_t(r)] / T(std::max(1, counts[size_t(r)]));
    return prof;
}

template <typename T> T parabola_offset(T l, T c, T r) {
    const T den = l - T(2) * c + r;
    if (std::abs(den) < T(1e-6)) return T(0);
    return clampv(T(0.5) * (l - r) / den, T(-0.5), T(0.5));
}

template <typename T> T first_minimum(const T* raw, int size) {
    const std::vector<T> prof = radial_profile(raw, size);
    const int n = int(prof.size());
    if (n < 4) return T(1);
    std::vector<T> s(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        const T l = prof[size_t(i == 0 ? 0 : i - 1)];
        const T r = prof[size_t(i + 1 < n ? i + 1 : i)];
        s[size_t(i)] = (l + T(2) * prof[size_t(i)] + r) * T(0.25);
    }
    int best = 1;
    T best_v = s[1];
    int descents = s[1] < s[0] ? 1 : 0;
    for (int i = 2; i + 1 < n; ++i) {
        if (s[size_t(i)] <= s[size_t(i - 1)]) {
            descents += 1;
            if (s[size_t(i)] < best_v) { best_v = s[size_t(i)]; best = i; }
            continue;
        }
        if (descents >= 2 && best > 0 && best + 1 < n) {
            const T o = parabola_offset(s[size_t(best - 1)], s[size_t(best)], s[size_t(best + 1)]);
            return std::max(T(1), T(best) + o);
        }
        best = i;
        best_v = s[size_t(i)];
        descents = 0;
    }
    const int limit = std::max(2, std::min(n - 1, std::max(4, size / 8)));
    best = 1;
    best_v = s[1];
    for (int i = 2; i <= limit; ++i)
        if (s[size_t(i)] < best_v) { best_v = s[size_t(i)]; best = i; }
    if (best > 0 && best + 1 < n) {
        const T o = parabola_offset(s[size_t(best - 1)], s[size_t(best)], s[size_t(best + 1)]);
        return std::max(T(1), T(best) + o);
    }
    return T(best);
}

// ---- resampling (LensDiffCpuReference.cpp:316-353, :1060-1086) ---------------------------
template <typename T> T bilinear(const T* img, int size, T x, T y) {
    if (x < T(0) || y < T(0) || x > T(size - 1) || y > T(size - 1)) return T(0);
    const int x0 = int(std::floor(x)), y0 = int(std::floor(y));
    const int x1 = std::min(x0 + 1, size - 1), y1 = std::min(y0 + 1, size - 1);
    const T tx = x - T(x0), ty = y - T(y0);
    const T v00 = img[size_t(y0) * size + x0], v10 = img[size_t(y0) * size + x1];
    const T v01 = img[size_t(y1) * size + x0], v11 = img[size_t(y1) * size + x1];
    const T a = v00 + (v10 - v00) * tx;
    const T b = v01 + (v11 - v01) * tx;
    return a + (b - a) * ty;
}

template <typename T> T box_bilinear(const T* img, int size, T x, T y, T footprint) {
    if (footprint <= T(1)) return bilinear(img, size, x, y);
    const int taps = clampv(int(std::ceil(footprint)), 2, 6);
    const T step = footprint / T(taps);
    T sum = T(0);
    for (int j = 0; j < taps; ++j) {
        const T oy = (T(j) + T(0.5)) * step - footprint * T(0.5);
        for (int i = 0; i < taps; ++i) {
            const T ox = (T(i) + T(0.5)) * step - footprint * T(0.5);
            sum += bilinear(img, size, x + ox, y + oy);
        }
    }
    return sum / T(taps * taps);
}

template <typename T> std::vector<T> resample_kernel(const T* raw, int raw_size, T scale, int support) {
    const int k = support * 2 + 1;
    std::vector<T> out(size_t(k) * k, T(0));
    const T rc = spectrum_center<T>(raw_size);
    const T kc = T(k - 1) * T(0.5);
    const T inv = T(1) / std::max(scale, T(0.05));
    for (int y = 0; y < k; ++y)
        for (int x = 0; x < k; ++x) {
            const T sx = rc + (T(x) - kc) * inv;
            const T sy = rc + (T(y) - kc) * inv;
            out[size_t(y) * k + x] = box_bilinear(raw, raw_size, sx, sy, inv);
        }
    normalize_unit(out);
    return out;
}

// ---- bank shaping (LensDiffCpuReference.cpp:957-1151) ------------------------------------
template <typename T> struct Kernel {
    int size = 0;
    std::vector<T> v;
};

template <typename T> Kernel<T> ring_mean(const Kernel<T>& k) {
    Kernel<T> m{k.size, std::vector<T>(k.v.size(), T(0))};
    const int rmax = k.size / 2;
    std::vector<T> sums(static_cast<size_t>(rmax + 1), T(0));
    std::vector<int> counts(static_cast<size_t>(rmax + 1), 0);
    const T c = T(k.size - 1) * T(0.5);
    std::vector<int> ring(k.v.size());
    for (int y = 0; y < k.size; ++y)
        for (int x = 0; x < k.size; ++x) {
            const T dx = T(x) - c, dy = T(y) - c;
            const int r = std::min(rmax, int(std::round(std::sqrt(dx * dx + dy * dy))));
            ring[size_t(y) * k.size + x] = r;
            sums[size_t(r)] += k.v[size_t(y) * k.size + x];
            counts[size_t(r)] += 1;
        }
    for (size_t i = 0; i < k.v.size(); ++i) {
        const int r = ring[i];
        m.v[i] = sums[size_t(r)] / T(std::max(1, counts[size_t(r)]));
    }
    normalize_unit(m.v);
    return m;
}

template <typename T> Kernel<T> emphasise_anisotropy(const Kernel<T>& k, const Kernel<T>& mean, T emphasis) {
    Kernel<T> s{k.size, std::vector<T>(k.v.size(), T(0))};
    const T gain = T(1) + std::max(T(0), emphasis) * T(4);
    for (size_t i = 0; i < k.v.size(); ++i) {
        const T mv = mean.v[i];
        s.v[i] = std::max(T(0), mv + (k.v[i] - mv) * gain);
    }
    normalize_unit(s.v);
    return s;
}

template <typename T> Kernel<T> positive_residual(const Kernel<T>& full, const Kernel<T>& mean) {
    Kernel<T> r{full.size, std::vector<T>(full.v.size(), T(0))};
    for (size_t i = 0; i < full.v.size(); ++i) r.v[i] = std::max(T(0), full.v[i] - mean.v[i]);
    if (seq_sum(r.v) > T(0)) normalize_unit(r.v);
    return r;
}

template <typename T> int support_radius(const Kernel<T>& k, int max_radius) {
    if (k.size <= 0 || k.v.empty()) return std::max(4, max_radius);
    const int rmax = std::min(max_radius, k.size / 2);
    const T c = T(k.size - 1) * T(0.5);
    std::vector<T> energy(static_cast<size_t>(rmax + 1), T(0)), peak(size_t(rmax + 1), T(0));
    T total = T(0), gpeak = T(0);
    for (int y = 0; y < k.size; ++y)
        for (int x = 0; x < k.size; ++x) {
            const T v = k.v[size_t(y) * k.size + x];
            const T dx = T(x) - c, dy = T(y) - c;
            const int r = std::min(rmax, int(std::ceil(std::sqrt(dx * dx + dy * dy))));
            energy[size_t(r)] += v;
            peak[size_t(r)] = std::max(peak[size_t(r)], v);
            total += v;
            gpeak = std::max(gpeak, v);
        }
    if (total <= T(1e-6) || gpeak <= T(1e-6)) return std::max(4, rmax);
    std::vector<T> outside(static_cast<size_t>(rmax + 2), T(0));
    for (int r = rmax; r >= 0; --r) outside[size_t(r)] = std::max(peak[size_t(r)], outside[size_t(r + 1)]);
    T cum = T(0);
    for (int r = 0; r <= rmax; ++r) {
        cum += energy[size_t(r)];
        const T captured = cum / total;
        const T remaining = outside[size_t(std::min(rmax + 1, r + 1))];
        if (captured >= T(0.9999) && remaining <= gpeak * T(2e-4)) return std::max(4, r + 1);
    }
    return std::max(4, rmax);
}

inline int padded_radius(int est, int max_radius) {
    const int pad = std::max(6, std::min(24, int(std::ceil(float(std::max(1, est)) * 0.05f))));
    return clampv(est + pad, 4, std::max(4, max_radius));
}

template <typename T> void boundary_taper(Kernel<T>& k, int radius) {
    if (k.size <= 0 || k.v.empty() || radius <= 0) return;
    const T c = T(k.size - 1) * T(0.5);
    const T extent = std::max(T(1), T(radius));
    const T fade = std::max(T(6), std::min(T(24), extent * T(0.04)));
    const T start = std::max(T(0), extent - fade);
    for (int y = 0; y < k.size; ++y)
        for (int x = 0; x < k.size; ++x) {
            const T dx = T(x) - c, dy = T(y) - c;
            const T rr = std::sqrt(dx * dx + dy * dy);
            T w;
            if (rr >= extent) {
                w = T(0);
            } else {
                const T t = clampv((rr - start) / std::max(fade, T(1e-6)), T(0), T(1));
                w = std::max(T(0), std::cos(smooth01(t) * T(kPi<double> * 0.5)));
            }
            k.v[size_t(y) * k.size + x] *= w;
        }
}

template <typename T> Kernel<T> crop_to_radius(const Kernel<T>& k, int radius) {
    const int r = clampv(radius, 1, k.size / 2);
    const int n = r * 2 + 1;
    const int c = k.size / 2;
    Kernel<T> out{n, std::vector<T>(size_t(n) * n, T(0))};
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x)
            out.v[size_t(y) * n + x] = k.v[size_t(c - r + y) * k.size + size_t(c - r + x)];
    boundary_taper(out, r);
    normalize_unit(out.v);
    return out;
}

template <typename T> struct Bin {
    float wavelength = 550.0f;
    Kernel<T> full, core, structure;
};

// One bin of populatePsfBankCacheFromBaseKernels (LensDiffCpuReference.cpp:1630-1654).
template <typename T> Bin<T> shape_bin(const ldh_params& p, const Kernel<T>& base, float wl, int* radius_out) {
    Kernel<T> mean = ring_mean(base);
    Kernel<T> shaped = emphasise_anisotropy(base, mean, T(p.anisotropy_emphasis));
    Kernel<T> structure = positive_residual(shaped, mean);
    const int max_r = std::max(4, max_kernel_radius_px(p));
    const int r = padded_radius(support_radius(shaped, max_r), max_r);
    Bin<T> b;
    b.wavelength = wl;
    b.core = crop_to_radius(mean, r);
    b.full = crop_to_radius(shaped, r);
    b.structure = crop_to_radius(structure, r);
    if (radius_out) *radius_out = r;
    return b;
}

// wavelengthsForMode (LensDiffCpuReference.cpp:1602-1614) or the explicit extension list.
inline std::vector<float> wavelength_list(const ldh_params& p) {
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

inline ldh_params clear_circle_params(int pupil) {
    ldh_params q;
    orc_params_default(&q);
    q.aperture_mode = LDH_APERTURE_CIRCLE;
    q.apodization_mode = LDH_APOD_FLAT;
    q.roundness = 1.0;
    q.pupil_resolution = pupil;
    return q;
}

// GetLensDiffReferenceRawPsfCached (LensDiffCpuReference.cpp:2018-2038): process-wide,
// keyed by (P, R), guarded by a mutex.
template <typename T> std::shared_ptr<const std::vector<T>> clear_circle_psf(int pupil, int raw) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::shared_ptr<const std::vector<T>>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({pupil, raw});
    if (it != cache.end()) return it->second;
    const ldh_params q = clear_circle_params(pupil);
    const std::vector<T> amp = pupil_amplitude<T>(q);
    auto psf = std::make_shared<const std::vector<T>>(raw_psf<T>(amp, nullptr, pupil, raw));
    cache.emplace(std::make_pair(pupil, raw), psf);
    return psf;
}

template <typename T> struct Bank {
    std::vector<Bin<T>> bins;
    std::vector<Kernel<T>> base;
    int support = 4;
    int raw_size = 0;
};

// buildPsfBankGlobalOnly (LensDiffCpuReference.cpp:1702-1734).
template <typename T> Bank<T>* build_bank(const ldh_params& p) {
    auto* bank = new Bank<T>();
    const int pupil = pupil_size_of(p.pupil_resolution);
    const int max_r = std::max(4, max_kernel_radius_px(p));
    const int raw = raw_size_of(pupil, max_r);
    bank->raw_size = raw;
    const std::vector<T> amp = pupil_amplitude<T>(p);
    const auto ref = clear_circle_psf<T>(pupil, raw);
    const T zero = std::max(T(1), first_minimum(ref->data(), raw));
    const T scale_base = T(std::max(1.0, diffraction_scale_px(p))) / zero;
    for (float wl : wavelength_list(p)) {
        ldh_params q = p;
        const float off = (wl - 550.0f) / 110.0f;
        q.phase_defocus += p.chromatic_focus * off;
        if (std::abs(q.phase_defocus) > 1e-6) q.phase_mode = LDH_PHASE_ENABLED;
        const std::vector<T> ph = pupil_phase<T>(q);
        const std::vector<T> rp = raw_psf<T>(amp, ph.data(), pupil, raw);
        const T chroma = std::max(T(0.25), T(1) + T(p.chromatic_spread) * T(off));
        const T scale = scale_base * (T(wl) / T(550)) * chroma;
        Kernel<T> k{max_r * 2 + 1, resample_kernel<T>(rp.data(), raw, scale, max_r)};
        bank->base.push_back(k);
    }
    // the bins are shaped independently of one another (populatePsfBankCacheFromBaseKernels,
    // :1616-1656, one bin per loop iteration); threads take whole bins
    const std::vector<float> wls = wavelength_list(p);
    const int nb = int(wls.size());
    bank->bins.resize(size_t(nb));
    std::vector<int> radii(size_t(nb), 0);
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nb; ++i) bank->bins[size_t(i)] = shape_bin<T>(p, bank->base[size_t(i)], wls[size_t(i)], &radii[size_t(i)]);
    for (int r : radii) bank->support = std::max(bank->support, r);
    return bank;
}

// ---- spectrum mapping (LensDiffSpectrum.cpp, ColorManagement.cpp) --------------------------
struct V3 { float x, y, z; };

// CIE 1931 2-degree colour matching functions at 5 nm from 380 nm (public CIE table; the
// reference's 82-entry table repeats the 380 nm row in its last slot, ColorManagement.cpp:48-90).
const V3 kCmf[82] = {
    {0.001368f, 0.000039f, 0.006450f}, {0.002236f, 0.000064f, 0.010550f}, {0.004243f, 0.000120f, 0.020050f},
    {0.007650f, 0.000217f, 0.036210f}, {0.014310f, 0.000396f, 0.067850f}, {0.023190f, 0.000640f, 0.110200f},
    {0.043510f, 0.001210f, 0.207400f}, {0.077630f, 0.002180f, 0.371300f}, {0.134380f, 0.004000f, 0.645600f},
    {0.214770f, 0.007300f, 1.039050f}, {0.283900f, 0.011600f, 1.385600f}, {0.328500f, 0.016840f, 1.622960f},
    {0.348280f, 0.023000f, 1.747060f}, {0.348060f, 0.029800f, 1.782600f}, {0.336200f, 0.038000f, 1.772110f},
    {0.318700f, 0.048000f, 1.744100f}, {0.290800f, 0.060000f, 1.669200f}, {0.251100f, 0.073900f, 1.528100f},
    {0.195360f, 0.090980f, 1.287640f}, {0.142100f, 0.112600f, 1.041900f}, {0.095640f, 0.139020f, 0.812950f},
    {0.057950f, 0.169300f, 0.616200f}, {0.032010f, 0.208020f, 0.465180f}, {0.014700f, 0.258600f, 0.353300f},
    {0.004900f, 0.323000f, 0.272000f}, {0.002400f, 0.407300f, 0.212300f}, {0.009300f, 0.503000f, 0.158200f},
    {0.029100f, 0.608200f, 0.111700f}, {0.063270f, 0.710000f, 0.078250f}, {0.109600f, 0.793200f, 0.057250f},
    {0.165500f, 0.862000f, 0.042160f}, {0.225750f, 0.914850f, 0.029840f}, {0.290400f, 0.954000f, 0.020300f},
    {0.359700f, 0.980300f, 0.013400f}, {0.433450f, 0.994950f, 0.008750f}, {0.512050f, 1.000000f, 0.005750f},
    {0.594500f, 0.995000f, 0.003900f}, {0.678400f, 0.978600f, 0.002750f}, {0.762100f, 0.952000f, 0.002100f},
    {0.842500f, 0.915400f, 0.001800f}, {0.916300f, 0.870000f, 0.001650f}, {0.978600f, 0.816300f, 0.001400f},
    {1.026300f, 0.757000f, 0.001100f}, {1.056700f, 0.694900f, 0.001000f}, {1.062200f, 0.631000f, 0.000800f},
    {1.045600f, 0.566800f, 0.000600f}, {1.002600f, 0.503000f, 0.000340f}, {0.938400f, 0.441200f, 0.000240f},
    {0.854450f, 0.381000f, 0.000190f}, {0.751400f, 0.321000f, 0.000100f}, {0.642400f, 0.265000f, 0.000050f},
    {0.541900f, 0.217000f, 0.000030f}, {0.447900f, 0.175000f, 0.000020f}, {0.360800f, 0.138200f, 0.000010f},
    {0.283500f, 0.107000f, 0.000000f}, {0.218700f, 0.081600f, 0.000000f}, {0.164900f, 0.061000f, 0.000000f},
    {0.121200f, 0.044580f, 0.000000f}, {0.087400f, 0.032000f, 0.000000f}, {0.063600f, 0.023200f, 0.000000f},
    {0.046770f, 0.017000f, 0.000000f}, {0.032900f, 0.011920f, 0.000000f}, {0.022700f, 0.008210f, 0.000000f},
    {0.015840f, 0.005723f, 0.000000f}, {0.011359f, 0.004102f, 0.000000f}, {0.008111f, 0.002929f, 0.000000f},
    {0.005790f, 0.002091f, 0.000000f}, {0.004109f, 0.001484f, 0.000000f}, {0.002899f, 0.001047f, 0.000000f},
    {0.002049f, 0.000740f, 0.000000f}, {0.001440f, 0.000520f, 0.000000f}, {0.001000f, 0.000361f, 0.000000f},
    {0.000690f, 0.000249f, 0.000000f}, {0.000476f, 0.000172f, 0.000000f}, {0.000332f, 0.000120f, 0.000000f},
    {0.000235f, 0.000085f, 0.000000f}, {0.000166f, 0.000060f, 0.000000f}, {0.000117f, 0.000042f, 0.000000f},
    {0.000083f, 0.000030f, 0.000000f}, {0.000059f, 0.000021f, 0.000000f}, {0.000042f, 0.000015f, 0.000000f},
    {0.001368f, 0.000039f, 0.006450f},
};

struct M3 { float m[3][3]; };

// 3x3 inverse by cofactors in float (ColorManagement.cpp:118-150).
inline M3 inverse3(const M3& a, float det_eps) {
    const float c00 = a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1];
    const float c01 = a.m[0][2] * a.m[2][1] - a.m[0][1] * a.m[2][2];
    const float c02 = a.m[0][1] * a.m[1][2] - a.m[0][2] * a.m[1][1];
    const float c10 = a.m[1][2] * a.m[2][0] - a.m[1][0] * a.m[2][2];
    const float c11 = a.m[0][0] * a.m[2][2] - a.m[0][2] * a.m[2][0];
    const float c12 = a.m[0][2] * a.m[1][0] - a.m[0][0] * a.m[1][2];
    const float c20 = a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0];
    const float c21 = a.m[0][1] * a.m[2][0] - a.m[0][0] * a.m[2][1];
    const float c22 = a.m[0][0] * a.m[1][1] - a.m[0][1] * a.m[1][0];
    const float det = a.m[0][0] * c00 + a.m[0][1] * c10 + a.m[0][2] * c20;
    M3 o{};
    if (std::fabs(det) <= det_eps) return o;
    const float id = 1.0f / det;
    const float c[3][3] = {{c00, c01, c02}, {c10, c11, c12}, {c20, c21, c22}};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) o.m[i][j] = c[i][j] * id;
    return o;
}
inline V3 mul3(const M3& a, V3 v) {
    return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
            a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
            a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
inline V3 xy_to_xyz(float x, float y) {
    if (std::fabs(y) <= 1e-8f) return {x, 1.0f, 1.0f - x};
    return {x * 1.0f / y, 1.0f, (1.0f - x - y) * 1.0f / y};
}
// Rec.709 XYZ->RGB: primaries matrix scaled to the D65 white, then inverted
// (ColorManagement.cpp:266-299).
inline M3 rec709_xyz_to_rgb() {
    const V3 r = xy_to_xyz(0.64f, 0.33f), g = xy_to_xyz(0.30f, 0.60f), b = xy_to_xyz(0.15f, 0.06f);
    const V3 w = xy_to_xyz(0.3127f, 0.3290f);
    M3 prim{{{r.x, g.x, b.x}, {r.y, g.y, b.y}, {r.z, g.z, b.z}}};
    const V3 s = mul3(inverse3(prim, 1e-12f), w);
    M3 out = prim;
    for (int row = 0; row < 3; ++row) { out.m[row][0] *= s.x; out.m[row][1] *= s.y; out.m[row][2] *= s.z; }
    return inverse3(out, 1e-12f);
}

inline float luma3(const V3& v) { return 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z; }
inline V3 scale3(const V3& v, float s) { return {v.x * s, v.y * s, v.z * s}; }
inline V3 to_luma(const V3& v, float target) {
    const float cur = luma3(v);
    if (cur <= 1e-6f || target <= 0.0f) return v;
    return scale3(v, target / cur);
}
// LensDiffSpectrum.cpp:49-72
inline V3 natural_column(float wl) {
    static const M3 m = rec709_xyz_to_rgb();
    const int idx = clampv(int(std::lround((wl - 380.0f) / 5.0f)), 0, 81);
    V3 c = mul3(m, kCmf[idx]);
    c = {std::max(0.0f, c.x), std::max(0.0f, c.y), std::max(0.0f, c.z)};
    if (c.x + c.y + c.z <= 1e-6f) {
        if (wl <= 500.0f) c = {0.08f, 0.55f, 1.0f};
        else if (wl <= 580.0f) c = {0.42f, 1.0f, 0.24f};
        else c = {1.0f, 0.22f, 0.08f};
    }
    return c;
}
// LensDiffSpectrum.cpp:74-128
inline std::vector<V3> balance_to_white3(const std::vector<V3>& c) {
    const float b[3][3] = {{c[0].x, c[1].x, c[2].x}, {c[0].y, c[1].y, c[2].y}, {c[0].z, c[1].z, c[2].z}};
    M3 bm{};
    std::memcpy(bm.m, b, sizeof(b));
    const float c00 = b[1][1] * b[2][2] - b[1][2] * b[2][1];
    const float c10 = b[1][2] * b[2][0] - b[1][0] * b[2][2];
    const float c20 = b[1][0] * b[2][1] - b[1][1] * b[2][0];
    const float det = b[0][0] * c00 + b[0][1] * c10 + b[0][2] * c20;
    if (std::fabs(det) <= 1e-8f) return c;
    const M3 inv = inverse3(bm, 1e-8f);
    float wgt[3];
    for (int i = 0; i < 3; ++i) {
        wgt[i] = inv.m[i][0] * 1.0f + inv.m[i][1] * 1.0f + inv.m[i][2] * 1.0f;
        if (!std::isfinite(wgt[i]) || wgt[i] <= 0.0f) return c;
    }
    return {scale3(c[0], wgt[0]), scale3(c[1], wgt[1]), scale3(c[2], wgt[2])};
}
inline V3 sum_cols(const std::vector<V3>& c) {
    V3 s{0, 0, 0};
    for (const V3& v : c) s = {s.x + v.x, s.y + v.y, s.z + v.z};
    return s;
}
// LensDiffSpectrum.cpp:130-167
inline std::vector<V3> balance_neutral(const std::vector<V3>& c) {
    if (c.empty()) return c;
    const V3 s = sum_cols(c);
    if (std::max(s.x, std::max(s.y, s.z)) <= 1e-6f) return c;
    const float t = (s.x + s.y + s.z) / 3.0f;
    if (t <= 1e-6f) return c;
    const V3 k{t / std::max(s.x, 1e-6f), t / std::max(s.y, 1e-6f), t / std::max(s.z, 1e-6f)};
    std::vector<V3> out(c.size());
    for (size_t i = 0; i < c.size(); ++i)
        out[i] = to_luma(V3{c[i].x * k.x, c[i].y * k.y, c[i].z * k.z}, luma3(c[i]));
    return out;
}
// LensDiffSpectrum.cpp:198-233
inline std::vector<V3> balance_to_sum(const std::vector<V3>& c, const V3& target) {
    if (c.empty()) return c;
    const V3 s = sum_cols(c);
    const float cur = (s.x + s.y + s.z) / 3.0f;
    const float tav = (target.x + target.y + target.z) / 3.0f;
    if (cur <= 1e-6f || tav <= 1e-6f) return balance_neutral(c);
    const V3 nt{target.x * (cur / tav), target.y * (cur / tav), target.z * (cur / tav)};
    const V3 k{nt.x / std::max(s.x, 1e-6f), nt.y / std::max(s.y, 1e-6f), nt.z / std::max(s.z, 1e-6f)};
    std::vector<V3> out(c.size());
    for (size_t i = 0; i < c.size(); ++i)
        out[i] = to_luma(V3{c[i].x * k.x, c[i].y * k.y, c[i].z * k.z}, luma3(c[i]));
    return out;
}
// LensDiffSpectrum.cpp:235-269
inline V3 style_column(float wl, int style, const V3& nat) {
    V3 target = nat;
    float strength = 0.0f;
    if (style == LDH_STYLE_CYAN_MAGENTA) {
        target = wl <= 500.0f ? V3{0.18f, 0.84f, 1.0f} : (wl <= 580.0f ? V3{0.90f, 0.89f, 0.96f} : V3{0.98f, 0.68f, 0.82f});
        strength = 0.56f;
    } else if (style == LDH_STYLE_WARM_COOL) {
        target = wl <= 500.0f ? V3{0.24f, 0.74f, 1.0f} : (wl <= 580.0f ? V3{0.92f, 0.86f, 0.76f} : V3{1.0f, 0.76f, 0.48f});
        strength = 0.52f;
    }
    const float t = clampv(strength, 0.0f, 1.0f);
    const V3 mixed{nat.x * (1.0f - t) + target.x * t, nat.y * (1.0f - t) + target.y * t,
                   nat.z * (1.0f - t) + target.z * t};
    return to_luma(mixed, luma3(nat));
}
const float kCanon5[5] = {440.f, 490.f, 540.f, 590.f, 640.f};
const float kCanon9[9] = {420.f, 450.f, 480.f, 510.f, 540.f, 570.f, 600.f, 630.f, 660.f};
// LensDiffSpectrum.cpp:271-303
inline std::vector<V3> canon5_columns(int style) {
    std::vector<V3> c(5);
    for (int i = 0; i < 5; ++i) c[size_t(i)] = natural_column(kCanon5[i]);
    c = balance_neutral(c);
    if (style != LDH_STYLE_NATURAL) {
        std::vector<V3> s(5);
        for (int i = 0; i < 5; ++i) s[size_t(i)] = style_column(kCanon5[i], style, c[size_t(i)]);
        c = balance_neutral(s);
    }
    return c;
}
// LensDiffSpectrum.cpp:305-332
inline V3 interp_column(float wl, const std::vector<V3>& ref) {
    if (wl <= kCanon5[0]) return ref.front();
    if (wl >= kCanon5[4]) return ref.back();
    for (int i = 1; i < 5; ++i) {
        if (wl <= kCanon5[i]) {
            const float span = std::max(kCanon5[i] - kCanon5[i - 1], 1e-6f);
            const float t = clampv((wl - kCanon5[i - 1]) / span, 0.0f, 1.0f);
            const V3& a = ref[size_t(i - 1)];
            const V3& b = ref[size_t(i)];
            return {a.x * (1.0f - t) + b.x * t, a.y * (1.0f - t) + b.y * t, a.z * (1.0f - t) + b.z * t};
        }
    }
    return ref.back();
}

// BuildLensDiffSpectrumConfig (LensDiffSpectrum.cpp:380-460) for n bins. For n <= 9 this is the
// reference's behaviour; for n > 9 (an extension: the reference clamps to 9 bins) every bin gets
// a column and the ">= 9 bins" balancing branch applies to all of them. Matrices are stored
// row-major [channel][bin].
inline void spectrum_config(const ldh_params& p, int n, const float* wl, std::vector<float>& nat_m,
                            std::vector<float>& sty_m) {
    nat_m.assign(size_t(3 * n), 0.0f);
    sty_m.assign(size_t(3 * n), 0.0f);
    if (n <= 1) {
        for (int c = 0; c < 3; ++c) { nat_m[size_t(c * n)] = 1.0f; sty_m[size_t(c * n)] = 1.0f; }
        return;
    }
    bool canon9 = n == 9;
    for (int i = 0; canon9 && i < 9; ++i) canon9 = std::fabs(wl[i] - kCanon9[i]) <= 1e-3f;
    std::vector<V3> nat(static_cast<size_t>(n));
    if (canon9) {
        const std::vector<V3> ref = canon5_columns(LDH_STYLE_NATURAL);
        for (int i = 0; i < n; ++i) nat[size_t(i)] = interp_column(wl[i], ref);
    } else {
        for (int i = 0; i < n; ++i) nat[size_t(i)] = natural_column(wl[i]);
    }
    if (n == 3) nat = balance_to_white3(nat);
    else if (n >= 9) nat = balance_to_sum(nat, sum_cols(canon5_columns(LDH_STYLE_NATURAL)));
    else nat = balance_neutral(nat);
    std::vector<V3> sty = nat;
    if (p.spectrum_style != LDH_STYLE_NATURAL) {
        if (canon9) {
            const std::vector<V3> ref = canon5_columns(p.spectrum_style);
            for (int i = 0; i < n; ++i) sty[size_t(i)] = interp_column(wl[i], ref);
        } else {
            for (int i = 0; i < n; ++i) sty[size_t(i)] = style_column(wl[i], p.spectrum_style, nat[size_t(i)]);
        }
        if (n == 3) sty = balance_to_white3(sty);
        else if (n >= 9) sty = balance_to_sum(sty, sum_cols(canon5_columns(p.spectrum_style)));
        else sty = balance_neutral(sty);
    }
    for (int i = 0; i < n; ++i) {
        nat_m[size_t(0 * n + i)] = nat[size_t(i)].x; nat_m[size_t(1 * n + i)] = nat[size_t(i)].y;
        nat_m[size_t(2 * n + i)] = nat[size_t(i)].z;
        sty_m[size_t(0 * n + i)] = sty[size_t(i)].x; sty_m[size_t(1 * n + i)] = sty[size_t(i)].y;
        sty_m[size_t(2 * n + i)] = sty[size_t(i)].z;
    }
}

// MapLensDiffSpectralBins (LensDiffSpectrum.cpp:362-376, :462-497) for one pixel.
template <typename T>
void map_pixel(const ldh_params& p, int n, const T* nat_m, const T* sty_m, const T* v, T* rgb) {
    T nat[3], sty[3];
    if (n <= 1) {
        nat[0] = nat[1] = nat[2] = v[0];
        sty[0] = sty[1] = sty[2] = v[0];
    } else {
        for (int c = 0; c < 3; ++c) {
            T a = T(0), b = T(0);
            for (int i = 0; i < n; ++i) {
                a += nat_m[c * n + i] * v[i];
                b += sty_m[c * n + i] * v[i];
            }
            nat[c] = a;
            sty[c] = b;
        }
    }
    const T force = clampv(T(p.spectrum_force), T(0), T(1));
    T out[3];
    for (int c = 0; c < 3; ++c) out[c] = std::max(T(0), nat[c] * (T(1) - force) + sty[c] * force);
    const T s = T(std::max(0.0, p.spectrum_saturation));
    const T gray = T(0.2126) * out[0] + T(0.7152) * out[1] + T(0.0722) * out[2];
    for (int c = 0; c < 3; ++c) out[c] = gray + (out[c] - gray) * s;
    if (!p.chromatic_affects_luma) {
        const T target = T(0.2126) * nat[0] + T(0.7152) * nat[1] + T(0.0722) * nat[2];
        const T cur = T(0.2126) * out[0] + T(0.7152) * out[1] + T(0.0722) * out[2];
        if (cur > T(1e-6)) {
            const T k = target / cur;
            for (int c = 0; c < 3; ++c) out[c] *= k;
        }
    }
    for (int c = 0; c < 3; ++c) rgb[c] = out[c];
}

// ---- frame stages (LensDiffCpuReference.cpp:1153-1338) -------------------------------------
template <typename T> struct Plane {
    int w = 0, h = 0;
    std::vector<T> v;
    Plane() = default;
    Plane(int w_, int h_) : w(w_), h(h_), v(size_t(w_) * h_, T(0)) {}
};
template <typename T> struct Rgba {
    int w = 0, h = 0;
    std::vector<T> v;
    Rgba() = default;
    Rgba(int w_, int h_) : w(w_), h(h_), v(size_t(w_) * h_ * 4, T(0)) {}
    T* px(int x, int y) { return v.data() + (size_t(y) * w + x) * 4; }
    const T* px(int x, int y) const { return v.data() + (size_t(y) * w + x) * 4; }
};

// convolveScalar (LensDiffCpuReference.cpp:1153-1193): zero-padded linear convolution on a
// nextPow2(W+k-1) x nextPow2(H+k-1) grid, kernel centre wrapped to the origin, clamp >= 0.
template <typename T> Plane<T> convolve(const Plane<T>& src, const Kernel<T>& k) {
    using C = std::complex<T>;
    const int pw = pow2_at_least(src.w + k.size - 1);
    const int ph = pow2_at_least(src.h + k.size - 1);
    std::vector<C> img(size_t(pw) * ph, C(0, 0));
    for (int y = 0; y < src.h; ++y)
        for (int x = 0; x < src.w; ++x) img[size_t(y) * pw + x] = C(src.v[size_t(y) * src.w + x], T(0));
    fft_2d(img, pw, ph, false);
    std::vector<C> ker(size_t(pw) * ph, C(0, 0));
    const int c = k.size / 2;
    for (int y = 0; y < k.size; ++y)
        for (int x = 0; x < k.size; ++x) {
            const int px = (x - c + pw) % pw;
            const int py = (y - c + ph) % ph;
            ker[size_t(py) * pw + px] = C(k.v[size_t(y) * k.size + x], T(0));
        }
    fft_2d(ker, pw, ph, false);
    for (size_t i = 0; i < img.size(); ++i) img[i] *= ker[i];
    fft_2d(img, pw, ph, true);
    Plane<T> out(src.w, src.h);
    for (int y = 0; y < src.h; ++y)
        for (int x = 0; x < src.w; ++x) out.v[size_t(y) * src.w + x] = std::max(T(0), img[size_t(y) * pw + x].real());
    return out;
}

template <typename T> Rgba<T> convolve_rgb(const Rgba<T>& src, const Kernel<T>& k) {
    Rgba<T> out(src.w, src.h);
    for (int c = 0; c < 3; ++c) {
        Plane<T> ch(src.w, src.h);
        for (size_t i = 0; i < ch.v.size(); ++i) ch.v[i] = src.v[i * 4 + size_t(c)];
        const Plane<T> r = convolve(ch, k);
        for (size_t i = 0; i < ch.v.size(); ++i) out.v[i * 4 + size_t(c)] = r.v[i];
    }
    for (size_t i = 0; i < size_t(src.w) * src.h; ++i) out.v[i * 4 + 3] = src.v[i * 4 + 3];
    return out;
}

// buildSelectionMask / applyMask / applyMaskScalarDriver (LensDiffCpuReference.cpp:1219-1279).
template <typename T> Plane<T> selection_mask(const Rgba<T>& src, const ldh_params& p) {
    Plane<T> m(src.w, src.h);
    const T soft = T(std::max(0.01, p.softness_stops));
    const T thr = T(p.threshold);
    const T thr_lin = kGray18<T> * std::pow(T(2), thr);
    for (int y = 0; y < src.h; ++y)
        for (int x = 0; x < src.w; ++x) {
            const T* q = src.px(x, y);
            const T mx = std::max({q[0], q[1], q[2]});
            const T sig = p.extraction_mode == LDH_EXTRACT_LUMA ? luma709(q[0], q[1], q[2]) : mx;
            T gate = T(0);
            if (sig > T(0)) {
                const T stops = std::log2(std::max(sig, T(1e-6)) / kGray18<T>);
                const T e0 = thr - soft * T(0.5);
                const T e1 = thr + soft * T(0.5);
                const T t = sat01((stops - e0) / std::max(e1 - e0, T(1e-4)));
                gate = t * t * (T(3) - T(2) * t);
            }
            const T boost = T(1) + T(p.point_emphasis) * std::max(T(0), mx / std::max(thr_lin, T(1e-4)) - T(1));
            m.v[size_t(y) * src.w + x] = sat01(gate * boost);
        }
    return m;
}

template <typename T> Rgba<T> masked(const Rgba<T>& src, const Plane<T>& m, T scale) {
    Rgba<T> out(src.w, src.h);
    for (size_t i = 0; i < m.v.size(); ++i) {
        const T k = m.v[i] * scale;
        out.v[i * 4 + 0] = src.v[i * 4 + 0] * k;
        out.v[i * 4 + 1] = src.v[i * 4 + 1] * k;
        out.v[i * 4 + 2] = src.v[i * 4 + 2] * k;
        out.v[i * 4 + 3] = src.v[i * 4 + 3];
    }
    return out;
}

template <typename T> Plane<T> masked_driver(const Rgba<T>& src, const Plane<T>& m, int mode, T scale) {
    Plane<T> out(src.w, src.h);
    for (size_t i = 0; i < m.v.size(); ++i) {
        const T* q = src.v.data() + i * 4;
        const T sig = mode == LDH_EXTRACT_LUMA ? luma709(q[0], q[1], q[2]) : std::max({q[0], q[1], q[2]});
        out.v[i] = sig * m.v[i] * scale;
    }
    return out;
}

template <typename T> T luma_total(const Rgba<T>& img) {
    T s = T(0);
    for (size_t i = 0; i < size_t(img.w) * img.h; ++i) s += luma709(img.v[i * 4], img.v[i * 4 + 1], img.v[i * 4 + 2]);
    return s;
}

template <typename T> T shoulder_of(T v, T sh) {
    if (sh <= T(0)) return std::max(T(0), v);
    const T x = std::max(T(0), v);
    return sh * (T(1) - std::exp(-x / sh));
}
template <typename T> void apply_shoulder(Rgba<T>& img, T sh) {
    if (sh <= T(0)) return;
    for (size_t i = 0; i < size_t(img.w) * img.h; ++i)
        for (int c = 0; c < 3; ++c) img.v[i * 4 + size_t(c)] = shoulder_of(img.v[i * 4 + size_t(c)], sh);
}
template <typename T> Rgba<T> add_rgb(const Rgba<T>& a, const Rgba<T>& b, T ga, T gb) {
    Rgba<T> o(a.w, a.h);
    for (size_t i = 0; i < size_t(a.w) * a.h; ++i) {
        for (int c = 0; c < 3; ++c) o.v[i * 4 + size_t(c)] = a.v[i * 4 + size_t(c)] * ga + b.v[i * 4 + size_t(c)] * gb;
        o.v[i * 4 + 3] = a.v[i * 4 + 3];
    }
    return o;
}

enum class Which { Full, Core, Structure };

// buildSpectralEffect (LensDiffCpuReference.cpp:1298-1338), generalised to n bins.
template <typename T>
Rgba<T> spectral_effect(const Plane<T>& driver, const std::vector<Bin<T>>& bins, const ldh_params& p,
                        const std::vector<T>& nat_m, const std::vector<T>& sty_m, Which which) {
    const int n = int(bins.size());
    std::vector<Plane<T>> conv;
    for (const auto& b : bins) {
        const Kernel<T>& k = which == Which::Core ? b.core : (which == Which::Structure ? b.structure : b.full);
        conv.push_back(convolve(driver, k));
    }
    Rgba<T> out(driver.w, driver.h);
    std::vector<T> v(size_t(std::max(1, n)));
    for (size_t i = 0; i < driver.v.size(); ++i) {
        for (int j = 0; j < n; ++j) v[size_t(j)] = conv[size_t(j)].v[i];
        T rgb[3];
        map_pixel<T>(p, n, nat_m.data(), sty_m.data(), v.data(), rgb);
        out.v[i * 4 + 0] = rgb[0]; out.v[i * 4 + 1] = rgb[1]; out.v[i * 4 + 2] = rgb[2];
        out.v[i * 4 + 3] = T(1);
    }
    return out;
}

// buildGaussianKernel (LensDiffCpuReference.cpp:1680-1700).
template <typename T> Kernel<T> gaussian_kernel(T radius_px) {
    const int r = std::max(1, int(std::ceil(std::max(T(0.5), radius_px))));
    Kernel<T> k{r * 2 + 1, std::vector<T>(size_t(r * 2 + 1) * (r * 2 + 1), T(0))};
    const T sigma = std::max(T(0.5), radius_px * T(0.5));
    T sum = T(0);
    for (int y = -r; y <= r; ++y)
        for (int x = -r; x <= r; ++x) {
            const T v = std::exp(-(T(x * x + y * y) / (T(2) * sigma * sigma)));
            k.v[size_t(y + r) * k.size + size_t(x + r)] = v;
            sum += v;
        }
    if (sum > T(0)) for (T& v : k.v) v /= sum;
    return k;
}

template <typename T> std::array<T, 4> sample_clamped(const Rgba<T>& img, T fx, T fy) {
    const T x = clampv(fx, T(0), T(std::max(0, img.w - 1)));
    const T y = clampv(fy, T(0), T(std::max(0, img.h - 1)));
    const int x0 = int(std::floor(x)), y0 = int(std::floor(y));
    const int x1 = std::min(x0 + 1, img.w - 1), y1 = std::min(y0 + 1, img.h - 1);
    const T tx = x - T(x0), ty = y - T(y0);
    std::array<T, 4> o{};
    const T *a = img.px(x0, y0), *b = img.px(x1, y0), *c = img.px(x0, y1), *d = img.px(x1, y1);
    for (int i = 0; i < 4; ++i) {
        const T u = a[i] * (T(1) - tx) + b[i] * tx;
        const T w = c[i] * (T(1) - tx) + d[i] * tx;
        o[size_t(i)] = u * (T(1) - ty) + w * ty;
    }
    return o;
}

// applyCreativeFringe (LensDiffCpuReference.cpp:1499-1534).
template <typename T> Rgba<T> creative_fringe(const Rgba<T>& img, T amount, Rgba<T>* preview) {
    if (amount <= T(1e-6)) { if (preview) *preview = Rgba<T>(img.w, img.h); return img; }
    Rgba<T> out(img.w, img.h), pv(img.w, img.h);
    const T cx = T(std::max(0, img.w - 1)) * T(0.5), cy = T(std::max(0, img.h - 1)) * T(0.5);
    for (int y = 0; y < img.h; ++y)
        for (int x = 0; x < img.w; ++x) {
            const T dx = T(x) - cx, dy = T(y) - cy;
            const T len = std::sqrt(dx * dx + dy * dy);
            const T il = len > T(1e-6) ? T(1) / len : T(0);
            const T sx = dx * il * amount, sy = dy * il * amount;
            const auto r = sample_clamped(img, T(x) + sx, T(y) + sy);
            const auto g = sample_clamped(img, T(x), T(y));
            const auto b = sample_clamped(img, T(x) - sx, T(y) - sy);
            const T* s = img.px(x, y);
            T* d = out.px(x, y);
            T* q = pv.px(x, y);
            d[0] = r[0]; d[1] = g[1]; d[2] = b[2]; d[3] = s[3];
            q[0] = std::abs(d[0] - s[0]); q[1] = std::abs(d[1] - s[1]); q[2] = std::abs(d[2] - s[2]); q[3] = T(1);
        }
    if (preview) *preview = std::move(pv);
    return out;
}

// resampleRgbaImage / resampleScalarImage (LensDiffCpuReference.cpp:355-434).
template <typename T> Rgba<T> resample_rgba(const Rgba<T>& src, int ow, int oh) {
    if (src.w == ow && src.h == oh) return src;
    Rgba<T> out(ow, oh);
    const T sx = T(src.w) / T(std::max(1, ow)), sy = T(src.h) / T(std::max(1, oh));
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            const T fx = (T(x) + T(0.5)) * sx - T(0.5), fy = (T(y) + T(0.5)) * sy - T(0.5);
            T* o = out.px(x, y);
            if (fx < T(0) || fy < T(0) || fx > T(src.w - 1) || fy > T(src.h - 1)) { o[0] = o[1] = o[2] = o[3] = T(0); continue; }
            const int x0 = int(std::floor(fx)), y0 = int(std::floor(fy));
            const int x1 = std::min(x0 + 1, src.w - 1), y1 = std::min(y0 + 1, src.h - 1);
            const T tx = fx - T(x0), ty = fy - T(y0);
            const T *a = src.px(x0, y0), *b = src.px(x1, y0), *c = src.px(x0, y1), *d = src.px(x1, y1);
            for (int i = 0; i < 4; ++i) {
                const T u = a[i] + (b[i] - a[i]) * tx;
                const T w = c[i] + (d[i] - c[i]) * tx;
                o[i] = u + (w - u) * ty;
            }
        }
    return out;
}
template <typename T> Plane<T> resample_plane(const Plane<T>& src, int ow, int oh) {
    if (src.w == ow && src.h == oh) return src;
    Plane<T> out(ow, oh);
    const T sx = T(src.w) / T(std::max(1, ow)), sy = T(src.h) / T(std::max(1, oh));
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            const T fx = (T(x) + T(0.5)) * sx - T(0.5), fy = (T(y) + T(0.5)) * sy - T(0.5);
            out.v[size_t(y) * ow + x] = bilinear_rect(src, fx, fy);
        }
    return out;
}
template <typename T> T bilinear_rect(const Plane<T>& s, T x, T y) {
    if (x < T(0) || y < T(0) || x > T(s.w - 1) || y > T(s.h - 1)) return T(0);
    const int x0 = int(std::floor(x)), y0 = int(std::floor(y));
    const int x1 = std::min(x0 + 1, s.w - 1), y1 = std::min(y0 + 1, s.h - 1);
    const T tx = x - T(x0), ty = y - T(y0);
    const T v00 = s.v[size_t(y0) * s.w + x0], v10 = s.v[size_t(y0) * s.w + x1];
    const T v01 = s.v[size_t(y1) * s.w + x0], v11 = s.v[size_t(y1) * s.w + x1];
    const T a = v00 + (v10 - v00) * tx, b = v01 + (v11 - v01) * tx;
    return a + (b - a) * ty;
}

// Transfer decode/encode (ColorManagement.cpp:156-169, LensDiffTransfer.cpp:9-21).
inline float decode_channel(float x, int tf) {
    if (tf == LDH_TRANSFER_DAVINCI_INTERMEDIATE)
        return x <= 0.02740668f ? x / 10.44426855f : std::exp2(x / 0.07329248f - 7.0f) - 0.0075f;
    return x;
}
template <typename T> T encode_channel(T x, int tf) {
    if (tf == LDH_TRANSFER_DAVINCI_INTERMEDIATE)
        return x <= T(0.00262409f) ? (x * T(10.44426855f))
                                   : ((std::log2(std::max(x, T(0)) + T(0.0075f)) + T(7)) * T(0.07329248f));
    return x;
}

// ---- Static debug views (BuildLensDiffStaticDebugRgba, LensDiffCpuReference.cpp:1909-1994) ----
// sampleSquareBilinear (:316-335): zero outside the square.
template <typename T> T sample_square(const std::vector<T>& img, int n, T x, T y) {
    if (x < T(0) || y < T(0) || x > T(n - 1) || y > T(n - 1)) return T(0);
    const int x0 = int(std::floor(x)), y0 = int(std::floor(y));
    const int x1 = std::min(x0 + 1, n - 1), y1 = std::min(y0 + 1, n - 1);
    const T tx = x - T(x0), ty = y - T(y0);
    const T v00 = img[size_t(y0) * n + x0], v10 = img[size_t(y0) * n + x1];
    const T v01 = img[size_t(y1) * n + x0], v11 = img[size_t(y1) * n + x1];
    const T a = v00 + (v10 - v00) * tx, b = v01 + (v11 - v01) * tx;
    return a + (b - a) * ty;
}

// Where output texel (x, y) of the centred square falls in an n x n image, or false outside the
// square (the mapping of visualizeSquareCentered / visualizeSignedPhaseCentered).
template <typename T> bool square_coords(int x, int y, int ow, int oh, int n, T& sx, T& sy) {
    const int side = std::min(ow, oh), xo = (ow - side) / 2, yo = (oh - side) / 2;
    if (x < xo || x >= xo + side || y < yo || y >= yo + side) return false;
    const T nx = T(x - xo) / T(std::max(1, side - 1)), ny = T(y - yo) / T(std::max(1, side - 1));
    sx = nx * T(n - 1);
    sy = ny * T(n - 1);
    return true;
}

// visualizeSquareCentered (:1416-1450): grey, v / max or log1p(1024 v) / log1p(1024 max).
template <typename T> Rgba<T> view_square(const std::vector<T>& sq, int n, int ow, int oh, bool log_view) {
    Rgba<T> out(ow, oh);
    T mx = T(0);
    for (T v : sq) mx = std::max(mx, v);
    const T lmax = std::log1p(std::max(T(0), mx) * T(1024));
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            T d = T(0), sx, sy;
            if (square_coords(x, y, ow, oh, n, sx, sy)) {
                const T v = sample_square(sq, n, sx, sy);
                if (log_view) d = lmax > T(0) ? std::log1p(std::max(T(0), v) * T(1024)) / lmax : T(0);
                else d = mx > T(0) ? v / mx : T(0);
            }
            T* q = out.px(x, y);
            q[0] = q[1] = q[2] = d;
            q[3] = T(1);
        }
    return out;
}

// visualizeSignedPhaseCentered (:1536-1593): inside the pupil, neutral grey pushed towards warm
// for positive and cool for negative waves, by |phase| / max |phase|.
template <typename T>
Rgba<T> view_phase(const std::vector<T>& ph, const std::vector<T>& pupil, int n, int ow, int oh) {
    Rgba<T> out(ow, oh);
    T mabs = T(0);
    for (T v : ph) mabs = std::max(mabs, std::abs(v));
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            T* q = out.px(x, y);
            q[0] = q[1] = q[2] = T(0);
            q[3] = T(1);
            T sx, sy;
            if (!square_coords(x, y, ow, oh, n, sx, sy)) continue;
            if (sample_square(pupil, n, sx, sy) <= T(1e-4f)) continue;
            const T v = sample_square(ph, n, sx, sy);
            const T u = mabs > T(0) ? clampv(v / mabs, T(-1), T(1)) : T(0);
            const T m = std::abs(u), g = T(0.18f);
            if (m <= T(1e-6f)) {
                q[0] = q[1] = q[2] = g;
            } else if (u > T(0)) {
                q[0] = g + T(0.82f) * m;
                q[1] = g + T(0.28f) * m;
                q[2] = std::max(T(0), g - T(0.12f) * m);
            } else {
                q[0] = std::max(T(0), g - T(0.10f) * m);
                q[1] = g + T(0.34f) * m;
                q[2] = g + T(0.80f) * m;
            }
        }
    return out;
}

// computeOtfMagnitude (:1403-1414): |DFT| of the kernel wrapped to the origin on its own n x n
// grid (paddedSpectrumFromKernel, :1153-1165), centred by fftShiftSquare (:303-314).
template <typename T> std::vector<T> otf_magnitude(const Kernel<T>& k) {
    const int n = std::max(1, k.size), c = k.size / 2;
    std::vector<std::complex<T>> d(size_t(n) * n);
    for (int y = 0; y < k.size; ++y)
        for (int x = 0; x < k.size; ++x)
            d[size_t((y - c + n) % n) * n + size_t((x - c + n) % n)] = std::complex<T>(k.v[size_t(y) * k.size + x], T(0));
    fft_2d(d, n, n, false);
    std::vector<T> out(size_t(n) * n);
    const int half = n / 2;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x)
            out[size_t(y) * n + x] = std::abs(d[size_t((y + half) % n) * n + size_t((x + half) % n)]);
    return out;
}

inline bool static_view_kind(int view) {
    return view == LDH_VIEW_PUPIL || view == LDH_VIEW_PSF || view == LDH_VIEW_OTF || view == LDH_VIEW_PHASE ||
           view == LDH_VIEW_PHASE_EDGE || view == LDH_VIEW_FIELD_PSF || view == LDH_VIEW_CHROMATIC_SPLIT;
}

// The static view of the bank (global parameters) or, for the phase-edge and field-PSF views with
// field-dependent phase, of the last zone (2, 2), at ow x oh.
template <typename T> Rgba<T> static_view(const ldh_params& p, const Bank<T>& bank, int ow, int oh) {
    const int n = pupil_size_of(p.pupil_resolution);
    const bool edge = field_phase(p) && (p.debug_view == LDH_VIEW_PHASE_EDGE || p.debug_view == LDH_VIEW_FIELD_PSF);
    const ldh_params q = edge ? field_zone_params(p, 1.0f, 1.0f) : p;
    switch (p.debug_view) {
        case LDH_VIEW_PUPIL: return view_square(pupil_amplitude<T>(p), n, ow, oh, false);
        case LDH_VIEW_PHASE:
        case LDH_VIEW_PHASE_EDGE: return view_phase(pupil_phase<T>(q), pupil_amplitude<T>(q), n, ow, oh);
        case LDH_VIEW_PSF: return view_square(bank.bins.front().full.v, bank.bins.front().full.size, ow, oh, true);
        case LDH_VIEW_OTF: {
            const Kernel<T>& k = bank.bins.front().full;
            return view_square(otf_magnitude(k), std::max(1, k.size), ow, oh, true);
        }
        case LDH_VIEW_FIELD_PSF: {
            if (!edge) return view_square(bank.bins.front().full.v, bank.bins.front().full.size, ow, oh, true);
            std::unique_ptr<Bank<T>> zb(build_bank<T>(q));
            return view_square(zb->bins.front().full.v, zb->bins.front().full.size, ow, oh, true);
        }
        default: {  // chromatic split: last bin in red, middle in green, first in blue
            const auto& bins = bank.bins;
            const Kernel<T>* k[3] = {&bins.back().full, &bins[bins.size() / 2].full, &bins.front().full};
            Rgba<T> out(ow, oh);
            for (int c = 0; c < 3; ++c) {
                const Rgba<T> im = view_square(k[c]->v, k[c]->size, ow, oh, true);
                for (size_t i = 0; i < size_t(ow) * oh; ++i) out.v[i * 4 + size_t(c)] = im.v[i * 4 + size_t(c)];
            }
            for (size_t i = 0; i < size_t(ow) * oh; ++i) out.v[i * 4 + 3] = T(1);
            return out;
        }
    }
}

// accumulateWeightedImage (LensDiffCpuReference.cpp:1452-1475): dst += src * wx * wy for the
// zone's tent weights over the frame (zone centres at x = 0, (w-1)/2, w-1 and likewise in y);
// pixels with a zero weight are left alone, the others take src's alpha.
template <typename T> void accumulate_weighted(const Rgba<T>& src, Rgba<T>& dst, int zx, int zy) {
    const T dx = T(std::max(1, src.w - 1)), dy = T(std::max(1, src.h - 1));
    for (int y = 0; y < src.h; ++y) {
        const T py = (T(y) / dy) * T(2);
        const T wy = std::max(T(0), T(1) - std::abs(py - T(zy)));
        if (wy <= T(0)) continue;
        for (int x = 0; x < src.w; ++x) {
            const T px = (T(x) / dx) * T(2);
            const T wx = std::max(T(0), T(1) - std::abs(px - T(zx)));
            const T wt = wx * wy;
            if (wt <= T(0)) continue;
            const T* s = src.px(x, y);
            T* d = dst.px(x, y);
            for (int c = 0; c < 3; ++c) d[c] += s[c] * wt;
            d[3] = s[3];
        }
    }
}

// The effect and composite of RunLensDiffCpuReference (LensDiffCpuReference.cpp:2086-2220) on
// a decoded, tightly packed RGBA image. Returns the output image (before encode) or the debug
// view; `ok` false for parameters outside this path.
template <typename T>
Rgba<T> render_core(const ldh_params& p, const Rgba<T>& src, bool* ok) {
    *ok = true;
    std::unique_ptr<Bank<T>> bank(build_bank<T>(p));
    // static views replace the output at the frame's size (renderDebugView, :1800-1807)
    if (static_view_kind(p.debug_view)) return static_view(p, *bank, src.w, src.h);
    std::vector<float> wl;
    for (const auto& b : bank->bins) wl.push_back(b.wavelength);
    const int n = int(wl.size());
    std::vector<float> nm, sm;
    spectrum_config(p, n, wl.data(), nm, sm);
    const std::vector<T> nat_m(nm.begin(), nm.end()), sty_m(sm.begin(), sm.end());

    const double ws = double(optics_short_side(p)) / double(std::max(1, p.frame_short_side_px));
    const bool ra = p.resolution_aware && std::abs(ws - 1.0) > 1e-6;
    const int ww = ra ? std::max(1, int(std::lround(src.w * ws))) : src.w;
    const int wh = ra ? std::max(1, int(std::lround(src.h * ws))) : src.h;
    const Rgba<T> work = ra ? resample_rgba(src, ww, wh) : src;
    const Plane<T> mask = selection_mask(work, p);
    const T redist_scale = T(1) - T(clampv(p.core_preserve, 0.0, 1.0));
    const T protect = std::max(T(0.2), T(clampv(p.core_preserve, 0.0, 1.0)));
    const T max_sub = redist_scale > T(1e-6) ? (T(1) - protect) / redist_scale : T(0);
    const Rgba<T> redist = masked(work, mask, redist_scale);
    const Plane<T> driver = masked_driver(work, mask, p.extraction_mode, redist_scale);
    const T gain = p.energy_mode == LDH_ENERGY_PRESERVE ? T(clampv(p.effect_gain, 0.0, 1.0))
                                                        : T(std::max(0.0, p.effect_gain));
    const T comp = p.energy_mode == LDH_ENERGY_PRESERVE ? gain : T(std::max(0.0, p.core_compensation));

    // renderEffectFromBins (LensDiffCpuReference.cpp:2113-2154) for one bank's bins
    auto effect_from = [&](const std::vector<Bin<T>>& bins, Rgba<T>& core, Rgba<T>& structure, Rgba<T>& effect) {
        const Bin<T>& b0 = bins.front();
        const bool mono = p.wavelength_count <= 0 && p.spectral_mode == LDH_SPECTRAL_MONO;
        if (mono) {
            if (p.look_mode == LDH_LOOK_SPLIT) {
                core = convolve_rgb(redist, b0.core);
                structure = convolve_rgb(redist, b0.structure);
                apply_shoulder(core, T(p.core_shoulder));
                apply_shoulder(structure, T(p.structure_shoulder));
                effect = add_rgb(core, structure, T(std::max(0.0, p.core_gain)), T(std::max(0.0, p.structure_gain)));
            } else {
                effect = convolve_rgb(redist, b0.full);
                if (p.debug_view == LDH_VIEW_CORE) core = convolve_rgb(redist, b0.core);
                if (p.debug_view == LDH_VIEW_STRUCTURE) structure = convolve_rgb(redist, b0.structure);
            }
        } else {
            if (p.look_mode == LDH_LOOK_SPLIT) {
                core = spectral_effect(driver, bins, p, nat_m, sty_m, Which::Core);
                structure = spectral_effect(driver, bins, p, nat_m, sty_m, Which::Structure);
                apply_shoulder(core, T(p.core_shoulder));
                apply_shoulder(structure, T(p.structure_shoulder));
                effect = add_rgb(core, structure, T(std::max(0.0, p.core_gain)), T(std::max(0.0, p.structure_gain)));
            } else {
                effect = spectral_effect(driver, bins, p, nat_m, sty_m, Which::Full);
                if (p.debug_view == LDH_VIEW_CORE) core = spectral_effect(driver, bins, p, nat_m, sty_m, Which::Core);
                if (p.debug_view == LDH_VIEW_STRUCTURE) structure = spectral_effect(driver, bins, p, nat_m, sty_m, Which::Structure);
            }
        }
    };
    Rgba<T> core(work.w, work.h), structure(work.w, work.h), effect(work.w, work.h);
    if (!field_phase(p)) {
        effect_from(bank->bins, core, structure, effect);
    } else {
        // nine zone banks, each zone's images blended in with its tent weights
        // (EnsureLensDiffPsfBank :1756-1790, the zone loop :2160-2170)
        for (int zy = 0; zy < 3; ++zy)
            for (int zx = 0; zx < 3; ++zx) {
                std::unique_ptr<Bank<T>> zb(build_bank<T>(field_zone_params(p, float(zx - 1), float(zy - 1))));
                Rgba<T> zc(work.w, work.h), zs(work.w, work.h), ze(work.w, work.h);
                effect_from(zb->bins, zc, zs, ze);
                accumulate_weighted(zc, core, zx, zy);
                accumulate_weighted(zs, structure, zx, zy);
                accumulate_weighted(ze, effect, zx, zy);
            }
    }
    if (p.energy_mode == LDH_ENERGY_PRESERVE) {
        const T in_e = luma_total(redist), eff_e = luma_total(effect);
        if (eff_e > T(1e-6)) for (T& v : effect.v) v *= in_e / eff_e;
    }
    Rgba<T> scatter(work.w, work.h);
    const double scatter_px = percent_to_px(p.scatter_radius, optics_short_side(p));
    if (p.scatter_amount > 1e-6 && scatter_px > 0.25) {
        scatter = convolve_rgb(effect, gaussian_kernel(T(scatter_px)));
        effect = add_rgb(effect, scatter, T(1), T(std::max(0.0, p.scatter_amount)));
    }
    Rgba<T> fringe_pv(work.w, work.h);
    effect = creative_fringe(effect, T(std::max(0.0, percent_to_px(p.creative_fringe, optics_short_side(p)))), &fringe_pv);

    auto disp = [&](const Rgba<T>& im) { return ra ? resample_rgba(im, src.w, src.h) : im; };
    const Rgba<T> redist_d = disp(redist);
    const Rgba<T> effect_d = disp(effect);
    Rgba<T> out(src.w, src.h);
    for (size_t i = 0; i < size_t(src.w) * src.h; ++i) {
        for (int c = 0; c < 3; ++c) {
            const T s = src.v[i * 4 + size_t(c)], r = redist_d.v[i * 4 + size_t(c)], e = effect_d.v[i * 4 + size_t(c)];
            const T raw = std::max(T(0), s - comp * r + gain * e);
            const T flo = std::max(T(0), s - r * max_sub);
            out.v[i * 4 + size_t(c)] = std::max(raw, flo);
        }
        out.v[i * 4 + 3] = src.v[i * 4 + 3];
    }
    switch (p.debug_view) {
        case LDH_VIEW_FINAL: return out;
        case LDH_VIEW_SELECTION: {
            const Plane<T> md = ra ? resample_plane(mask, src.w, src.h) : mask;
            T mx = T(0);
            for (T v : md.v) mx = std::max(mx, v);
            const T im = mx > T(0) ? T(1) / mx : T(0);
            Rgba<T> d(src.w, src.h);
            for (size_t i = 0; i < md.v.size(); ++i) {
                const T v = sat01(md.v[i] * im);
                d.v[i * 4] = d.v[i * 4 + 1] = d.v[i * 4 + 2] = v;
                d.v[i * 4 + 3] = T(1);
            }
            return d;
        }
        case LDH_VIEW_CORE: return disp(core);
        case LDH_VIEW_STRUCTURE: return disp(structure);
        case LDH_VIEW_EFFECT: return effect_d;
        case LDH_VIEW_CREATIVE_FRINGE: return disp(fringe_pv);
        case LDH_VIEW_SCATTER: return disp(scatter);
        default: *ok = false; return out;
    }
}

}  // namespace orc

using namespace orc;

// ---- ldh_params defaults (LensDiffTypes.h:128-208) ----------------------------------------
// The oracle carries its own copy so it does not link against the product library.
extern "C" void orc_params_default(ldh_params* p) {
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

extern "C" {

int orc_resolve_sizes(const ldh_params* p, ldh_sizes* o) {
    o->pupil_size = pupil_size_of(p->pupil_resolution);
    o->max_kernel_radius_px = std::max(4, max_kernel_radius_px(*p));
    o->raw_psf_size = raw_size_of(o->pupil_size, o->max_kernel_radius_px);
    o->optics_short_side_px = optics_short_side(*p);
    o->diffraction_scale_px = diffraction_scale_px(*p);
    return 0;
}

int orc_wavelengths(const ldh_params* p, float* out) {
    const auto w = wavelength_list(*p);
    for (size_t i = 0; i < w.size(); ++i) out[i] = w[i];
    return int(w.size());
}

int orc_has_nonflat_phase(const ldh_params* p) { return nonflat_phase(*p) ? 1 : 0; }

float orc_sequential_sum_f32(const float* x, long n) {
    float s = 0.0f;
    for (long i = 0; i < n; ++i) s += x[i];
    return s;
}

}  // extern "C"

namespace {
template <typename T> void copy_out(const std::vector<T>& v, T* out) { std::copy(v.begin(), v.end(), out); }

template <typename T> void fft1d_api(T* data, int n, int inverse) {
    std::vector<std::complex<T>> a(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) a[size_t(i)] = {data[2 * i], data[2 * i + 1]};
    fft_any(a, inverse != 0);
    for (int i = 0; i < n; ++i) { data[2 * i] = a[size_t(i)].real(); data[2 * i + 1] = a[size_t(i)].imag(); }
}
template <typename T> void fft2d_api(T* data, int w, int h, int inverse) {
    std::vector<std::complex<T>> a(size_t(w) * h);
    for (size_t i = 0; i < a.size(); ++i) a[i] = {data[2 * i], data[2 * i + 1]};
    fft_2d(a, w, h, inverse != 0);
    for (size_t i = 0; i < a.size(); ++i) { data[2 * i] = a[i].real(); data[2 * i + 1] = a[i].imag(); }
}
template <typename T> int shape_api(const ldh_params* p, const T* base, int k, T* full, T* core, T* structure, int* out_size) {
    Kernel<T> b{k, std::vector<T>(base, base + size_t(k) * k)};
    int r = 0;
    const Bin<T> bin = shape_bin<T>(*p, b, 550.0f, &r);
    copy_out(bin.full.v, full);
    copy_out(bin.core.v, core);
    copy_out(bin.structure.v, structure);
    *out_size = bin.full.size;
    return r;
}
template <typename T> void spectrum_api(const ldh_params* p, int n, const float* wl, T* nat, T* sty) {
    std::vector<float> a, b;
    spectrum_config(*p, n, wl, a, b);
    for (size_t i = 0; i < a.size(); ++i) { nat[i] = T(a[i]); sty[i] = T(b[i]); }
}
template <typename T> void map_api(const ldh_params* p, int n, const T* nat, const T* sty, const T* bins, long npx, T* rgb) {
    std::vector<T> v(size_t(std::max(1, n)));
    for (long i = 0; i < npx; ++i) {
        for (int j = 0; j < n; ++j) v[size_t(j)] = bins[size_t(j) * size_t(npx) + size_t(i)];
        map_pixel<T>(*p, n, nat, sty, v.data(), rgb + size_t(i) * 3);
    }
}
template <typename T> Rgba<T> image_from(const float* rgba, int w, int h, int tf) {
    Rgba<T> img(w, h);
    for (size_t i = 0; i < size_t(w) * h; ++i) {
        for (int c = 0; c < 3; ++c) img.v[i * 4 + size_t(c)] = T(decode_channel(rgba[i * 4 + size_t(c)], tf));
        img.v[i * 4 + 3] = T(rgba[i * 4 + 3]);
    }
    return img;
}
template <typename T> void selection_api(const ldh_params* p, const float* rgba, int w, int h, T* mask, T* redist, T* driver) {
    const Rgba<T> img = image_from<T>(rgba, w, h, p->input_transfer);
    const Plane<T> m = selection_mask(img, *p);
    const T s = T(1) - T(clampv(p->core_preserve, 0.0, 1.0));
    copy_out(m.v, mask);
    if (redist) copy_out(masked(img, m, s).v, redist);
    if (driver) copy_out(masked_driver(img, m, p->extraction_mode, s).v, driver);
}
template <typename T> void convolve_api(const T* plane, int w, int h, const T* kernel, int k, T* out) {
    Plane<T> src(w, h);
    std::copy(plane, plane + size_t(w) * h, src.v.begin());
    Kernel<T> ker{k, std::vector<T>(kernel, kernel + size_t(k) * k)};
    copy_out(convolve(src, ker).v, out);
}
template <typename T> int render_image_api(const ldh_params* p, const float* src, int w, int h, T* out) {
    if (w <= 0 || h <= 0 || src == nullptr) return 0;
    bool ok = true;
    Rgba<T> r = render_core<T>(*p, image_from<T>(src, w, h, p->input_transfer), &ok);
    if (!ok) return 0;
    // the frame write of RunLensDiffCpuReference (:2235-2258): encode the Final view, alpha 1
    // for debug views
    for (size_t i = 0; i < size_t(w) * h; ++i) {
        if (p->debug_view == LDH_VIEW_FINAL) {
            for (int c = 0; c < 3; ++c) r.v[i * 4 + size_t(c)] = encode_channel(r.v[i * 4 + size_t(c)], p->input_transfer);
            r.v[i * 4 + 3] = T(src[i * 4 + 3]);
        } else {
            r.v[i * 4 + 3] = T(1);
        }
    }
    copy_out(r.v, out);
    return 1;
}

struct BankHandle {
    int precision;  // 32 or 64
    void* bank;
};
template <typename T> Bank<T>* bank_of(const void* h) { return static_cast<Bank<T>*>(static_cast<const BankHandle*>(h)->bank); }
}  // namespace

#define ORC_DEFINE(T, S)                                                                                       \
    extern "C" void orc_pupil_amplitude_##S(const ldh_params* p, T* out) { copy_out(pupil_amplitude<T>(*p), out); } \
    extern "C" void orc_pupil_phase_##S(const ldh_params* p, T* out) { copy_out(pupil_phase<T>(*p), out); }        \
    extern "C" void orc_raw_psf_##S(const T* amp, const T* phase, int pupil, int raw, T* out) {                   \
        std::vector<T> a(amp, amp + size_t(pupil) * pupil);                                                     \
        copy_out(raw_psf<T>(a, phase, pupil, raw), out);                                                        \
    }                                                                                                          \
    extern "C" void orc_fft1d_##S(T* d, int n, int inv) { fft1d_api<T>(d, n, inv); }                           \
    extern "C" void orc_fft2d_##S(T* d, int w, int h, int inv) { fft2d_api<T>(d, w, h, inv); }                 \
    extern "C" T orc_first_minimum_radius_##S(const T* raw, int n) { return first_minimum<T>(raw, n); }        \
    extern "C" void orc_resample_kernel_##S(const T* raw, int n, T scale, int support, T* out) {               \
        copy_out(resample_kernel<T>(raw, n, scale, support), out);                                             \
    }                                                                                                          \
    extern "C" int orc_shape_kernels_##S(const ldh_params* p, const T* base, int k, T* f, T* c, T* s, int* sz) { \
        return shape_api<T>(p, base, k, f, c, s, sz);                                                           \
    }                                                                                                          \
    extern "C" void* orc_bank_build_##S(const ldh_params* p) {                                                 \
        return new BankHandle{int(sizeof(T) * 8), build_bank<T>(*p)};                                           \
    }                                                                                                          \
    extern "C" void orc_spectrum_config_##S(const ldh_params* p, int n, const float* wl, T* a, T* b) {         \
        spectrum_api<T>(p, n, wl, a, b);                                                                       \
    }                                                                                                          \
    extern "C" void orc_map_bins_##S(const ldh_params* p, int n, const T* a, const T* b, const T* bins,        \
                                     long npx, T* rgb) { map_api<T>(p, n, a, b, bins, npx, rgb); }             \
    extern "C" void orc_selection_##S(const ldh_params* p, const float* rgba, int w, int h, T* m, T* r, T* d) { \
        selection_api<T>(p, rgba, w, h, m, r, d);                                                              \
    }                                                                                                          \
    extern "C" void orc_convolve_##S(const T* pl, int w, int h, const T* k, int ks, T* out) {                  \
        convolve_api<T>(pl, w, h, k, ks, out);                                                                 \
    }                                                                                                          \
    extern "C" int orc_render_image_##S(const ldh_params* p, const float* src, int w, int h, T* out) {         \
        return render_image_api<T>(p, src, w, h, out);                                                         \
    }

ORC_DEFINE(float, f32)
ORC_DEFINE(double, f64)
#undef ORC_DEFINE

extern "C" {

int orc_bank_bin_count(const void* h) {
    const auto* b = static_cast<const BankHandle*>(h);
    return b->precision == 32 ? int(bank_of<float>(h)->bins.size()) : int(bank_of<double>(h)->bins.size());
}
int orc_bank_support_radius(const void* h) {
    const auto* b = static_cast<const BankHandle*>(h);
    return b->precision == 32 ? bank_of<float>(h)->support : bank_of<double>(h)->support;
}
int orc_bank_raw_size(const void* h) {
    const auto* b = static_cast<const BankHandle*>(h);
    return b->precision == 32 ? bank_of<float>(h)->raw_size : bank_of<double>(h)->raw_size;
}
int orc_bank_kernel_size(const void* h, int bin) {
    const auto* b = static_cast<const BankHandle*>(h);
    return b->precision == 32 ? bank_of<float>(h)->bins[size_t(bin)].full.size
                              : bank_of<double>(h)->bins[size_t(bin)].full.size;
}
float orc_bank_wavelength(const void* h, int bin) {
    const auto* b = static_cast<const BankHandle*>(h);
    return b->precision == 32 ? bank_of<float>(h)->bins[size_t(bin)].wavelength
                              : bank_of<double>(h)->bins[size_t(bin)].wavelength;
}
int orc_bank_base_size(const void* h) {
    const auto* b = static_cast<const BankHandle*>(h);
    return b->precision == 32 ? bank_of<float>(h)->base[0].size : bank_of<double>(h)->base[0].size;
}
void orc_bank_kernel_f32(const void* h, int bin, int which, float* out) {
    const auto& b = bank_of<float>(h)->bins[size_t(bin)];
    copy_out(which == LDH_KERNEL_CORE ? b.core.v : (which == LDH_KERNEL_STRUCTURE ? b.structure.v : b.full.v), out);
}
void orc_bank_kernel_f64(const void* h, int bin, int which, double* out) {
    const auto& b = bank_of<double>(h)->bins[size_t(bin)];
    copy_out(which == LDH_KERNEL_CORE ? b.core.v : (which == LDH_KERNEL_STRUCTURE ? b.structure.v : b.full.v), out);
}
void orc_bank_base_f32(const void* h, int bin, float* out) { copy_out(bank_of<float>(h)->base[size_t(bin)].v, out); }
void orc_bank_base_f64(const void* h, int bin, double* out) { copy_out(bank_of<double>(h)->base[size_t(bin)].v, out); }
void orc_bank_free(void* h) {
    auto* b = static_cast<BankHandle*>(h);
    if (!b) return;
    if (b->precision == 32) delete static_cast<Bank<float>*>(b->bank);
    else delete static_cast<Bank<double>*>(b->bank);
    delete b;
}

// RunLensDiffCpuReference's I/O (LensDiffCpuReference.cpp:128-192, :2073-2085, :2235-2258).
int orc_render_f32(const ldh_params* p, const ldh_render_request* req, char* err, int err_len) {
    auto fail = [&](const char* tag) {
        if (err && err_len > 0) { std::strncpy(err, tag, size_t(err_len - 1)); err[err_len - 1] = 0; }
        return 0;
    };
    if (req->src.data == nullptr || req->dst.data == nullptr) return fail("missing-source-or-destination-buffer");
    const ldh_rect sb = req->src.bounds;
    const int w = sb.x2 - sb.x1, h = sb.y2 - sb.y1;
    if (w <= 0 || h <= 0) return fail("invalid-source-bounds");
    std::vector<float> src(size_t(w) * h * 4, 0.0f);
    const auto* base = static_cast<const uint8_t*>(req->src.data);
    const int64_t pix = int64_t(req->src.components) * req->src.bytes_per_component;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float* q = reinterpret_cast<const float*>(base + int64_t(y) * req->src.row_bytes + int64_t(x) * pix);
            float* o = src.data() + (size_t(y) * w + x) * 4;
            o[0] = q[0]; o[1] = q[1]; o[2] = q[2];
            o[3] = req->src.components > 3 ? q[3] : 1.0f;
        }
    bool ok = true;
    const Rgba<float> out = render_core<float>(*p, image_from<float>(src.data(), w, h, p->input_transfer), &ok);
    if (!ok) return fail("unsupported-debug-view");
    const ldh_rect db = req->dst.bounds;
    const ldh_rect rw = req->render_window;
    const int x1 = std::max(rw.x1, db.x1), y1 = std::max(rw.y1, db.y1);
    const int x2 = std::max(x1, std::min(rw.x2, db.x2)), y2 = std::max(y1, std::min(rw.y2, db.y2));
    auto* dbase = static_cast<uint8_t*>(req->dst.data);
    const int64_t dpix = int64_t(req->dst.components) * req->dst.bytes_per_component;
    for (int y = y1; y < y2; ++y)
        for (int x = x1; x < x2; ++x) {
            const int sx = x - sb.x1, sy = y - sb.y1;
            float rgba[4] = {0.0f, 0.0f, 0.0f, 1.0f};
            if (sx >= 0 && sx < w && sy >= 0 && sy < h) {
                const float* q = out.px(sx, sy);
                const bool fin = p->debug_view == LDH_VIEW_FINAL;
                for (int c = 0; c < 3; ++c) rgba[c] = fin ? encode_channel(q[c], p->input_transfer) : q[c];
                rgba[3] = fin ? src[(size_t(sy) * w + sx) * 4 + 3] : 1.0f;
            }
            float* d = reinterpret_cast<float*>(dbase + int64_t(y - db.y1) * req->dst.row_bytes + int64_t(x - db.x1) * dpix);
            d[0] = rgba[0]; d[1] = rgba[1]; d[2] = rgba[2];
            if (req->dst.components > 3) d[3] = rgba[3];
        }
    return 1;
}

}  // extern "C"

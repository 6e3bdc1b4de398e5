// Spectral bin -> RGB matrices (BuildLensDiffSpectrumConfig, LensDiffSpectrum.cpp:380-460) and
// the colour data they need (CIE 1931 2-degree CMFs at 5 nm and the Rec.709 XYZ->RGB matrix,
// ColorManagement.cpp:48-90, :266-299). Host code; the result is uploaded once per bank.
//
// For up to 9 bins this is the reference's behaviour. Longer wavelength lists (this backend's
// extension; the reference clamps to 9 bins, LensDiffSpectrum.cpp:383) give every bin a column
// and balance all of them with the reference's ">= 9 bins" branch.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "frame.h"

namespace ldh {

namespace {

struct V3 { float x, y, z; };
struct M3 { float m[3][3]; };

// CIE 1931 2-degree colour matching functions, 380..780 nm at 5 nm (public CIE data); slot 81
// repeats the 380 nm entry as the reference's table does.
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

bool invert(const M3& a, float eps, M3* out) {
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
    if (std::fabs(det) <= eps) { std::memset(out, 0, sizeof(*out)); return false; }
    const float id = 1.0f / det;
    const float c[3][3] = {{c00, c01, c02}, {c10, c11, c12}, {c20, c21, c22}};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out->m[i][j] = c[i][j] * id;
    return true;
}
V3 mul(const M3& a, V3 v) {
    return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z, a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
            a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
V3 xyz_of(float x, float y) {
    if (std::fabs(y) <= 1e-8f) return {x, 1.0f, 1.0f - x};
    return {x * 1.0f / y, 1.0f, (1.0f - x - y) * 1.0f / y};
}
M3 xyz_to_rec709() {
    const V3 r = xyz_of(0.64f, 0.33f), g = xyz_of(0.30f, 0.60f), b = xyz_of(0.15f, 0.06f), w = xyz_of(0.3127f, 0.3290f);
    M3 prim{{{r.x, g.x, b.x}, {r.y, g.y, b.y}, {r.z, g.z, b.z}}};
    M3 inv;
    invert(prim, 1e-12f, &inv);
    const V3 sc = mul(inv, w);
    for (int row = 0; row < 3; ++row) { prim.m[row][0] *= sc.x; prim.m[row][1] *= sc.y; prim.m[row][2] *= sc.z; }
    M3 out;
    invert(prim, 1e-12f, &out);
    return out;
}
float luma(const V3& v) { return 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z; }
V3 scaled(const V3& v, float s) { return {v.x * s, v.y * s, v.z * s}; }
V3 to_luma(const V3& v, float target) {
    const float cur = luma(v);
    if (cur <= 1e-6f || target <= 0.0f) return v;
    return scaled(v, target / cur);
}
V3 natural_column(float wl) {
    static const M3 m = xyz_to_rec709();
    const int idx = std::max(0, std::min(int(std::lround((wl - 380.0f) / 5.0f)), 81));
    V3 c = mul(m, kCmf[idx]);
    c = {std::max(0.0f, c.x), std::max(0.0f, c.y), std::max(0.0f, c.z)};
    if (c.x + c.y + c.z <= 1e-6f) {
        if (wl <= 500.0f) c = {0.08f, 0.55f, 1.0f};
        else if (wl <= 580.0f) c = {0.42f, 1.0f, 0.24f};
        else c = {1.0f, 0.22f, 0.08f};
    }
    return c;
}
V3 sum_of(const std::vector<V3>& c) {
    V3 s{0, 0, 0};
    for (const V3& v : c) s = {s.x + v.x, s.y + v.y, s.z + v.z};
    return s;
}
std::vector<V3> to_white(const std::vector<V3>& c) {
    M3 b{{{c[0].x, c[1].x, c[2].x}, {c[0].y, c[1].y, c[2].y}, {c[0].z, c[1].z, c[2].z}}};
    M3 inv;
    if (!invert(b, 1e-8f, &inv)) return c;
    float w[3];
    for (int i = 0; i < 3; ++i) {
        w[i] = inv.m[i][0] * 1.0f + inv.m[i][1] * 1.0f + inv.m[i][2] * 1.0f;
        if (!std::isfinite(w[i]) || w[i] <= 0.0f) return c;
    }
    return {scaled(c[0], w[0]), scaled(c[1], w[1]), scaled(c[2], w[2])};
}
std::vector<V3> rescale_channels(const std::vector<V3>& c, V3 k) {
    std::vector<V3> out(c.size());
    for (size_t i = 0; i < c.size(); ++i) out[i] = to_luma(V3{c[i].x * k.x, c[i].y * k.y, c[i].z * k.z}, luma(c[i]));
    return out;
}
std::vector<V3> to_neutral(const std::vector<V3>& c) {
    if (c.empty()) return c;
    const V3 s = sum_of(c);
    if (std::max(s.x, std::max(s.y, s.z)) <= 1e-6f) return c;
    const float t = (s.x + s.y + s.z) / 3.0f;
    if (t <= 1e-6f) return c;
    return rescale_channels(c, {t / std::max(s.x, 1e-6f), t / std::max(s.y, 1e-6f), t / std::max(s.z, 1e-6f)});
}
std::vector<V3> to_sum(const std::vector<V3>& c, V3 target) {
    if (c.empty()) return c;
    const V3 s = sum_of(c);
    const float cur = (s.x + s.y + s.z) / 3.0f, tav = (target.x + target.y + target.z) / 3.0f;
    if (cur <= 1e-6f || tav <= 1e-6f) return to_neutral(c);
    const V3 nt{target.x * (cur / tav), target.y * (cur / tav), target.z * (cur / tav)};
    return rescale_channels(c, {nt.x / std::max(s.x, 1e-6f), nt.y / std::max(s.y, 1e-6f), nt.z / std::max(s.z, 1e-6f)});
}
V3 style_column(float wl, int style, const V3& nat) {
    V3 target = nat;
    float strength = 0.0f;
    if (style == LDH_STYLE_CYAN_MAGENTA) {
        target = wl <= 500.0f ? V3{0.18f, 0.84f, 1.0f} : (wl <= 580.0f ? V3{0.90f, 0.89f, 0.96f} : V3{0.98f, 0.68f, 0.82f});
        strength = 0.56f;
    } else if (style == LDH_STYLE_WARM_COOL) {
        target = wl <= 500.0f ? V3{0.24f, 0.74f, 1.0f} : (wl <= 580.0f ? V3{0.92f, 0.86f, 0.76f} : V3{1.0f, 0.76f, 0.48f});
        strength = 0.52f;
    }
    const float t = std::max(0.0f, std::min(strength, 1.0f));
    return to_luma({nat.x * (1.0f - t) + target.x * t, nat.y * (1.0f - t) + target.y * t, nat.z * (1.0f - t) + target.z * t},
                   luma(nat));
}
const float kCanon5[5] = {440.f, 490.f, 540.f, 590.f, 640.f};
const float kCanon9[9] = {420.f, 450.f, 480.f, 510.f, 540.f, 570.f, 600.f, 630.f, 660.f};
std::vector<V3> canonical5(int style) {
    std::vector<V3> c(5);
    for (int i = 0; i < 5; ++i) c[size_t(i)] = natural_column(kCanon5[i]);
    c = to_neutral(c);
    if (style != LDH_STYLE_NATURAL) {
        std::vector<V3> st(5);
        for (int i = 0; i < 5; ++i) st[size_t(i)] = style_column(kCanon5[i], style, c[size_t(i)]);
        c = to_neutral(st);
    }
    return c;
}
V3 interpolate(float wl, const std::vector<V3>& ref) {
    if (wl <= kCanon5[0]) return ref.front();
    if (wl >= kCanon5[4]) return ref.back();
    for (int i = 1; i < 5; ++i)
        if (wl <= kCanon5[i]) {
            const float span = std::max(kCanon5[i] - kCanon5[i - 1], 1e-6f);
            const float t = std::max(0.0f, std::min((wl - kCanon5[i - 1]) / span, 1.0f));
            const V3& a = ref[size_t(i - 1)];
            const V3& b = ref[size_t(i)];
            return {a.x * (1.0f - t) + b.x * t, a.y * (1.0f - t) + b.y * t, a.z * (1.0f - t) + b.z * t};
        }
    return ref.back();
}
std::vector<V3> balance(const std::vector<V3>& c, int n, int style) {
    if (n == 3) return to_white(c);
    if (n >= 9) return to_sum(c, sum_of(canonical5(style)));
    return to_neutral(c);
}

}  // namespace

SpectrumConfig build_spectrum_config(const ldh_params& p, const std::vector<float>& wl) {
    SpectrumConfig cfg;
    const int n = std::max(1, int(wl.size()));
    cfg.bins = n;
    cfg.natural.assign(size_t(3 * n), 0.0f);
    cfg.style.assign(size_t(3 * n), 0.0f);
    if (n <= 1) {
        for (int c = 0; c < 3; ++c) cfg.natural[size_t(c * n)] = cfg.style[size_t(c * n)] = 1.0f;
        return cfg;
    }
    bool canon9 = n == 9;
    for (int i = 0; canon9 && i < 9; ++i) canon9 = std::fabs(wl[size_t(i)] - kCanon9[i]) <= 1e-3f;
    std::vector<V3> nat(static_cast<size_t>(n));
    if (canon9) {
        const std::vector<V3> ref = canonical5(LDH_STYLE_NATURAL);
        for (int i = 0; i < n; ++i) nat[size_t(i)] = interpolate(wl[size_t(i)], ref);
    } else {
        for (int i = 0; i < n; ++i) nat[size_t(i)] = natural_column(wl[size_t(i)]);
    }
    nat = balance(nat, n, LDH_STYLE_NATURAL);
    std::vector<V3> sty = nat;
    if (p.spectrum_style != LDH_STYLE_NATURAL) {
        if (canon9) {
            const std::vector<V3> ref = canonical5(p.spectrum_style);
            for (int i = 0; i < n; ++i) sty[size_t(i)] = interpolate(wl[size_t(i)], ref);
        } else {
            for (int i = 0; i < n; ++i) sty[size_t(i)] = style_column(wl[size_t(i)], p.spectrum_style, nat[size_t(i)]);
        }
        sty = balance(sty, n, p.spectrum_style);
    }
    for (int i = 0; i < n; ++i) {
        cfg.natural[size_t(i)] = nat[size_t(i)].x;
        cfg.natural[size_t(n + i)] = nat[size_t(i)].y;
        cfg.natural[size_t(2 * n + i)] = nat[size_t(i)].z;
        cfg.style[size_t(i)] = sty[size_t(i)].x;
        cfg.style[size_t(n + i)] = sty[size_t(i)].y;
        cfg.style[size_t(2 * n + i)] = sty[size_t(i)].z;
    }
    return cfg;
}

}  // namespace ldh

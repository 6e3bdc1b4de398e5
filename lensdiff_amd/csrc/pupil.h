// Pupil amplitude and Zernike phase, evaluated per pupil pixel on the device.
// Geometry follows buildPupilAmplitude and its helpers (LensDiffCpuReference.cpp:451-680,
// :758-822); the phase follows EvaluateLensDiffPhaseWaves / BuildLensDiffPupilPhaseWaves
// (LensDiffPhase.cpp:170-243). Parameters arrive pre-cast on the host exactly as the reference
// casts them, and the float expressions keep the reference's operation order (the library is
// built without FMA contraction).
#pragma once

#include "ldh_internal.h"

namespace ldh {

constexpr float kPiF = 3.14159265358979323846f;

struct PupilGeom {
    int size;        // P
    int mode;        // LDH_APERTURE_*
    int blades;      // max(3, bladeCount)
    int vanes;
    int apodization;
    float obstruction;     // clamp(centralObstruction, 0, 0.95) * 0.86
    float rotation;        // radians, float(rotationDeg * pi_f / 180)
    float roundness;       // clamp(roundness, 0, 1)
    float vane_thickness;  // max(0, vaneThickness)
    float star_inner;      // 0.18 + roundness * 0.62
    float decenter_x, decenter_y;
};

// Coefficients in the reference's summation order (LensDiffPhase.cpp:199-212):
// defocus, astig0, astig45, comaX, comaY, spherical, trefoilX, trefoilY, secAstig0, secAstig45,
// quad0, quad45, secComaX, secComaY; then the extension terms Noll 2, 3, 18, 19, 20, 21.
struct PhaseCoef {
    int enabled;     // HasLensDiffNonFlatPhase
    int has_extra;
    float c[14];
    float extra[6];
    float decenter_x, decenter_y;
    float cos_rot, sin_rot;  // cos(-rotation), sin(-rotation)
};

__device__ __forceinline__ float clampf_(float v, float lo, float hi) { return fmaxf(lo, fminf(v, hi)); }

__device__ inline float polygon_metric_d(float nx, float ny, int blades, float roundness, float rot, float outer) {
    const float radius = sqrtf(nx * nx + ny * ny);
    const float circle = radius / fmaxf(outer, 1e-5f);
    if (blades < 3) return circle;
    const float angle = atan2f(ny, nx) - rot;
    const float sector = 2.0f * kPiF / float(blades);
    float local = fmodf(angle, sector);
    if (local < 0.0f) local += sector;
    local -= sector * 0.5f;
    const float poly_radius = outer * cosf(kPiF / float(blades)) / fmaxf(cosf(local), 1e-4f);
    const float poly = radius / fmaxf(poly_radius, 1e-5f);
    return circle * roundness + poly * (1.0f - roundness);
}

__device__ inline bool vane_blocks_d(float nx, float ny, int vanes, float thickness, float rot, float outer) {
    if (vanes <= 0 || thickness <= 0.0f) return false;
    const float half = thickness * outer;
    const int lines = vanes > 1 ? vanes : 1;
    for (int i = 0; i < lines; ++i) {
        const float a = rot + float(i) * kPiF / float(lines);
        const float cs = cosf(a), sn = sinf(a);
        const float along = nx * cs + ny * sn;
        const float across = -nx * sn + ny * cs;
        if (fabsf(across) <= half && fabsf(along) <= outer) return true;
    }
    return false;
}

__device__ inline float star_metric_d(float nx, float ny, int points, float inner_ratio, float rot, float outer) {
    const float radius = sqrtf(nx * nx + ny * ny);
    if (radius <= 1e-6f) return 0.0f;
    const int n = points > 3 ? points : 3;
    const float inner = outer * clampf_(inner_ratio, 0.1f, 0.95f);
    const float angle = atan2f(ny, nx) - rot + kPiF * 0.5f;
    const float sector = 2.0f * kPiF / float(n);
    float local = fmodf(angle, sector);
    if (local < 0.0f) local += sector;
    const float half = sector * 0.5f;
    const float t = local <= half ? local / fmaxf(half, 1e-6f) : (sector - local) / fmaxf(half, 1e-6f);
    const float boundary = inner + (outer - inner) * clampf_(t, 0.0f, 1.0f);
    return radius / fmaxf(boundary, 1e-5f);
}

__device__ inline float segment_distance_d(float px, float py, float ax, float ay, float bx, float by) {
    const float vx = bx - ax, vy = by - ay;
    const float len2 = vx * vx + vy * vy;
    if (len2 <= 1e-10f) {
        const float dx = px - ax, dy = py - ay;
        return sqrtf(dx * dx + dy * dy);
    }
    const float t = clampf_(((px - ax) * vx + (py - ay) * vy) / len2, 0.0f, 1.0f);
    const float cx = ax + t * vx, cy = ay + t * vy;
    const float dx = px - cx, dy = py - cy;
    return sqrtf(dx * dx + dy * dy);
}

__device__ inline bool square_grid_open_d(float nx, float ny, int blades, float roundness, float rot, float outer) {
    const float cs = cosf(-rot), sn = sinf(-rot);
    const float rx = nx * cs - ny * sn;
    const float ry = nx * sn + ny * cs;
    const float half = outer * 0.82f;
    if (fabsf(rx) > half || fabsf(ry) > half) return false;
    const int div = blades > 3 ? blades : 3;
    const float pitch = (2.0f * half) / float(div);
    const float bar = pitch * (0.06f + (1.0f - clampf_(roundness, 0.0f, 1.0f)) * 0.12f);
    float w = fmodf(rx + half, pitch);
    float pos = w < 0.0f ? w + pitch : w;
    const float dx = fminf(pos, pitch - pos);
    w = fmodf(ry + half, pitch);
    pos = w < 0.0f ? w + pitch : w;
    const float dy = fminf(pos, pitch - pos);
    const float ex = fabsf(fabsf(rx) - half), ey = fabsf(fabsf(ry) - half);
    return dx <= bar || dy <= bar || ex <= bar || ey <= bar;
}

__device__ inline bool snowflake_open_d(float nx, float ny, int blades, float roundness, float rot, float outer) {
    const float radius = sqrtf(nx * nx + ny * ny);
    if (radius > outer) return false;
    int levels = blades / 2 > 2 ? blades / 2 : 2;
    levels = levels < 5 ? levels : 5;
    const float main_len = outer * 0.92f;
    const float branch_angle = kPiF / 5.5f;
    const float thick = outer * (0.045f + (1.0f - clampf_(roundness, 0.0f, 1.0f)) * 0.05f);
    float best = 3.402823466e38f;
    for (int arm = 0; arm < 6; ++arm) {
        const float a = rot - kPiF * 0.5f + float(arm) * (kPiF / 3.0f);
        const float cs = cosf(a), sn = sinf(a);
        best = fminf(best, segment_distance_d(nx, ny, 0.0f, 0.0f, cs * main_len, sn * main_len));
        for (int lv = 0; lv < levels; ++lv) {
            const float t = 0.34f + float(lv) * (0.44f / float(levels - 1 > 1 ? levels - 1 : 1));
            const float mx = cs * main_len * t, my = sn * main_len * t;
            const float blen = outer * (0.16f + 0.03f * float(lv));
            for (int side = -1; side <= 1; side += 2) {
                const float th = a + float(side) * branch_angle;
                best = fminf(best, segment_distance_d(nx, ny, mx, my, mx + cosf(th) * blen, my + sinf(th) * blen));
            }
        }
    }
    return best <= thick;
}

__device__ inline bool spiral_open_d(float nx, float ny, int blades, float roundness, float rot, float outer) {
    const float radius = sqrtf(nx * nx + ny * ny);
    if (radius > outer) return false;
    const float rn = radius / fmaxf(outer, 1e-6f);
    const float angle = atan2f(ny, nx) - rot;
    const float twist = 4.0f + (1.0f - clampf_(roundness, 0.0f, 1.0f)) * 8.0f;
    const float opening = 0.18f + clampf_(roundness, 0.0f, 1.0f) * 0.26f;
    const float ph = float(blades > 3 ? blades : 3) * angle + twist * rn * (2.0f * kPiF);
    const float band = 0.5f + 0.5f * cosf(ph);
    return band >= (1.0f - opening);
}

__device__ inline float apodization_d(int mode, float rn) {
    if (mode == LDH_APOD_COSINE) return cosf(fminf(rn, 1.0f) * (kPiF * 0.5f));
    if (mode == LDH_APOD_GAUSSIAN) return expf(-4.0f * rn * rn);
    return 1.0f;
}

__device__ inline float pupil_amplitude_at(const PupilGeom& g, int x, int y) {
    const float outer = 0.86f;
    const float nx = ((float(x) + 0.5f) / float(g.size) - 0.5f) * 2.0f;
    const float ny = ((float(y) + 0.5f) / float(g.size) - 0.5f) * 2.0f;
    const float dx = nx - g.decenter_x;
    const float dy = ny - g.decenter_y;
    const float radius = sqrtf(dx * dx + dy * dy);
    bool inside;
    switch (g.mode) {
        case LDH_APERTURE_POLYGON:
            inside = polygon_metric_d(dx, dy, g.blades, g.roundness, g.rotation, outer) <= 1.0f;
            break;
        case LDH_APERTURE_HEXAGON:
            inside = polygon_metric_d(dx, dy, 6, g.roundness, g.rotation, outer) <= 1.0f;
            break;
        case LDH_APERTURE_SQUARE_GRID:
            inside = square_grid_open_d(dx, dy, g.blades, g.roundness, g.rotation, outer);
            break;
        case LDH_APERTURE_SNOWFLAKE:
            inside = snowflake_open_d(dx, dy, g.blades, g.roundness, g.rotation, outer);
            break;
        case LDH_APERTURE_STAR:
            inside = star_metric_d(dx, dy, g.blades, g.star_inner, g.rotation, outer) <= 1.0f;
            break;
        case LDH_APERTURE_SPIRAL:
            inside = spiral_open_d(dx, dy, g.blades, g.roundness, g.rotation, outer);
            break;
        case LDH_APERTURE_CIRCLE:
            inside = radius <= outer;
            break;
        default:
            inside = false;
            break;
    }
    if (!inside || radius < g.obstruction) return 0.0f;
    if (vane_blocks_d(dx, dy, g.vanes, g.vane_thickness, g.rotation, outer)) return 0.0f;
    return apodization_d(g.apodization, radius / outer);
}

// The terms of phase_waves_at at one texel, for coefficient sets that differ only in defocus
// (c[0]) and `enabled`: the other products c_j z_j and the extra terms' sum are formed once, and
// phase(c0) adds them in the same order as phase_waves_at, so the result is the same float.
struct PhaseTexel {
    bool inside;
    float z4;
    float prod[13];  // c[1..13] times their terms, in phase_waves_at's order
    float extra;     // the extra terms' sum (0 without them)
    bool has_extra;
    __device__ float phase(const PhaseCoef& p) const {
        if (!p.enabled || !inside) return 0.0f;
        float w = p.c[0] * z4;
#pragma unroll
        for (int j = 0; j < 13; ++j) w = w + prod[j];
        if (has_extra) w += extra;
        return w;
    }
};
__device__ inline PhaseTexel phase_texel(const PhaseCoef& p, int x, int y, int size) {
    PhaseTexel t{};
    t.has_extra = p.has_extra != 0;
    const float outer = 0.86f;
    const float nx = ((float(x) + 0.5f) / float(size) - 0.5f) * 2.0f;
    const float ny = ((float(y) + 0.5f) / float(size) - 0.5f) * 2.0f;
    const float radius = sqrtf(nx * nx + ny * ny);
    const float rx = nx * p.cos_rot - ny * p.sin_rot;
    const float ry = nx * p.sin_rot + ny * p.cos_rot;
    const float sx = rx / outer - p.decenter_x;
    const float sy = ry / outer - p.decenter_y;
    const float r2 = sx * sx + sy * sy;
    t.inside = !(radius > outer) && !(r2 > 1.0f);
    const float n3 = 1.7320508075688772f, n6 = 2.4494897427831781f, n8 = 2.8284271247461901f;
    const float n5 = 2.2360679774997898f, n10 = 3.1622776601683795f, n12 = 3.4641016151377544f;
    t.z4 = n3 * (2.0f * r2 - 1.0f);
    const float z6 = n6 * (sx * sx - sy * sy);
    const float z5 = n6 * (2.0f * sx * sy);
    const float z8 = n8 * ((3.0f * r2 - 2.0f) * sx);
    const float z7 = n8 * ((3.0f * r2 - 2.0f) * sy);
    const float z11 = n5 * (6.0f * r2 * r2 - 6.0f * r2 + 1.0f);
    const float z10 = n8 * (sx * sx * sx - 3.0f * sx * sy * sy);
    const float z9 = n8 * (3.0f * sx * sx * sy - sy * sy * sy);
    const float sa = (4.0f * r2 - 3.0f);
    const float z12 = n10 * sa * (sx * sx - sy * sy);
    const float z13 = n10 * sa * (2.0f * sx * sy);
    const float z14 = n10 * (sx * sx * sx * sx - 6.0f * sx * sx * sy * sy + sy * sy * sy * sy);
    const float z15 = n10 * (4.0f * sx * sy * (sx * sx - sy * sy));
    const float sc = (10.0f * r2 * r2 - 12.0f * r2 + 3.0f);
    const float z16 = n12 * sc * sx;
    const float z17 = n12 * sc * sy;
    const float z[13] = {z6, z5, z8, z7, z11, z10, z9, z12, z13, z14, z15, z16, z17};
#pragma unroll
    for (int j = 0; j < 13; ++j) t.prod[j] = p.c[j + 1] * z[j];
    if (t.has_extra) {
        const float c3 = sx * sx * sx - 3.0f * sx * sy * sy;
        const float s3 = 3.0f * sx * sx * sy - sy * sy * sy;
        const float x2 = sx * sx, y2 = sy * sy;
        const float c5 = sx * (x2 * x2 - 10.0f * x2 * y2 + 5.0f * y2 * y2);
        const float s5 = sy * (5.0f * x2 * x2 - 10.0f * x2 * y2 + y2 * y2);
        const float r53 = (5.0f * r2 - 4.0f);
        t.extra = p.extra[0] * (2.0f * sx) + p.extra[1] * (2.0f * sy) + p.extra[2] * (n12 * r53 * c3) +
                  p.extra[3] * (n12 * r53 * s3) + p.extra[4] * (n12 * c5) + p.extra[5] * (n12 * s5);
    }
    return t;
}

__device__ inline float phase_waves_at(const PhaseCoef& p, int x, int y, int size) {
    if (!p.enabled) return 0.0f;
    const float outer = 0.86f;
    const float nx = ((float(x) + 0.5f) / float(size) - 0.5f) * 2.0f;
    const float ny = ((float(y) + 0.5f) / float(size) - 0.5f) * 2.0f;
    const float radius = sqrtf(nx * nx + ny * ny);
    if (radius > outer) return 0.0f;
    const float rx = nx * p.cos_rot - ny * p.sin_rot;
    const float ry = nx * p.sin_rot + ny * p.cos_rot;
    const float sx = rx / outer - p.decenter_x;
    const float sy = ry / outer - p.decenter_y;
    const float r2 = sx * sx + sy * sy;
    if (r2 > 1.0f) return 0.0f;
    const float n3 = 1.7320508075688772f, n6 = 2.4494897427831781f, n8 = 2.8284271247461901f;
    const float n5 = 2.2360679774997898f, n10 = 3.1622776601683795f, n12 = 3.4641016151377544f;
    const float z4 = n3 * (2.0f * r2 - 1.0f);
    const float z6 = n6 * (sx * sx - sy * sy);
    const float z5 = n6 * (2.0f * sx * sy);
    const float z8 = n8 * ((3.0f * r2 - 2.0f) * sx);
    const float z7 = n8 * ((3.0f * r2 - 2.0f) * sy);
    const float z11 = n5 * (6.0f * r2 * r2 - 6.0f * r2 + 1.0f);
    const float z10 = n8 * (sx * sx * sx - 3.0f * sx * sy * sy);
    const float z9 = n8 * (3.0f * sx * sx * sy - sy * sy * sy);
    const float sa = (4.0f * r2 - 3.0f);
    const float z12 = n10 * sa * (sx * sx - sy * sy);
    const float z13 = n10 * sa * (2.0f * sx * sy);
    const float z14 = n10 * (sx * sx * sx * sx - 6.0f * sx * sx * sy * sy + sy * sy * sy * sy);
    const float z15 = n10 * (4.0f * sx * sy * (sx * sx - sy * sy));
    const float sc = (10.0f * r2 * r2 - 12.0f * r2 + 3.0f);
    const float z16 = n12 * sc * sx;
    const float z17 = n12 * sc * sy;
    float w = p.c[0] * z4 + p.c[1] * z6 + p.c[2] * z5 + p.c[3] * z8 + p.c[4] * z7 + p.c[5] * z11 +
              p.c[6] * z10 + p.c[7] * z9 + p.c[8] * z12 + p.c[9] * z13 + p.c[10] * z14 + p.c[11] * z15 +
              p.c[12] * z16 + p.c[13] * z17;
    if (p.has_extra) {
        const float c3 = sx * sx * sx - 3.0f * sx * sy * sy;
        const float s3 = 3.0f * sx * sx * sy - sy * sy * sy;
        const float x2 = sx * sx, y2 = sy * sy;
        const float c5 = sx * (x2 * x2 - 10.0f * x2 * y2 + 5.0f * y2 * y2);
        const float s5 = sy * (5.0f * x2 * x2 - 10.0f * x2 * y2 + y2 * y2);
        const float r53 = (5.0f * r2 - 4.0f);
        w += p.extra[0] * (2.0f * sx) + p.extra[1] * (2.0f * sy) + p.extra[2] * (n12 * r53 * c3) +
             p.extra[3] * (n12 * r53 * s3) + p.extra[4] * (n12 * c5) + p.extra[5] * (n12 * s5);
    }
    return w;
}

// Host-side construction with the reference's casts.
PupilGeom make_pupil_geom(const ldh_params& p);
PhaseCoef make_phase_coef(const ldh_params& p);
bool params_nonflat_phase(const ldh_params& p);

}  // namespace ldh

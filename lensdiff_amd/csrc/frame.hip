// Per-frame path on the device:
//   selection mask (buildSelectionMask, LensDiffCpuReference.cpp:1219-1249)
//   -> image spectra: the row pass reads the frame and forms the redistributed channels or the
//      scalar driver on the fly (applyMask / applyMaskScalarDriver, :1251-1279), zero-padded to
//      the reference's nextPow2 grid (:1168-1169)
//   -> kernel spectra: the row pass wraps each kernel's centre to the origin
//      (paddedSpectrumFromKernel, :1153-1165); two real kernels share one complex transform
//   -> per (image spectrum, kernel spectrum): inverse transform whose first pass multiplies the
//      spectra and whose last pass crops to the frame, clamps at zero and writes the planes
//      (convolveScalar, :1167-1193)
//   -> spectral map (MapLensDiffSpectralBins, LensDiffSpectrum.cpp:462-497), split-look
//      shoulder/gains (:1340-1368), Preserve rescale (:2172-2178) and the composite (:2200-2220).
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

#include "fft_io.h"
#include "frame.h"
#include "host_params.h"

namespace ldh {

namespace {

__device__ __forceinline__ float decode_channel(float x, int tf) {
    if (tf == LDH_TRANSFER_DAVINCI_INTERMEDIATE)
        return x <= 0.02740668f ? x / 10.44426855f : exp2f(x / 0.07329248f - 7.0f) - 0.0075f;
    return x;
}
__device__ __forceinline__ float encode_channel(float x, int tf) {
    if (tf == LDH_TRANSFER_DAVINCI_INTERMEDIATE)
        return x <= 0.00262409f ? (x * 10.44426855f) : ((log2f(fmaxf(x, 0.0f) + 0.0075f) + 7.0f) * 0.07329248f);
    return x;
}
__device__ __forceinline__ float4 decoded(const float4* src, size_t i, int tf) {
    const float4 q = src[i];
    return make_float4(decode_channel(q.x, tf), decode_channel(q.y, tf), decode_channel(q.z, tf), q.w);
}
__device__ __forceinline__ float luma709(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }
__device__ __forceinline__ float max3(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }

struct SelectParams {
    int tf, extraction;
    float threshold_stops, softness, threshold_linear, point_emphasis;
};

// The selection mask of one decoded pixel (buildSelectionMask, LensDiffCpuReference.cpp): a
// smoothstep gate on the signal's stops above the threshold, times the point-emphasis boost,
// clamped to [0, 1]. A function of the pixel alone, so passes that read the pixel evaluate it.
__device__ __forceinline__ float select_decoded(float4 q, const SelectParams& sp) {
    const float mx = max3(q.x, q.y, q.z);
    const float sig = sp.extraction == LDH_EXTRACT_LUMA ? luma709(q.x, q.y, q.z) : mx;
    float gate = 0.0f;
    if (sig > 0.0f) {
        const float stops = log2f(fmaxf(sig, 1e-6f) / 0.18f);
        const float e0 = sp.threshold_stops - sp.softness * 0.5f;
        const float e1 = sp.threshold_stops + sp.softness * 0.5f;
        const float t = fminf(fmaxf((stops - e0) / fmaxf(e1 - e0, 1e-4f), 0.0f), 1.0f);
        gate = t * t * (3.0f - 2.0f * t);
    }
    const float boost = 1.0f + sp.point_emphasis * fmaxf(0.0f, mx / fmaxf(sp.threshold_linear, 1e-4f) - 1.0f);
    return fminf(fmaxf(gate * boost, 0.0f), 1.0f);
}

__global__ void selection_kernel(const float4* __restrict__ src, int64_t npx, SelectParams sp, float* __restrict__ mask) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    mask[i] = select_decoded(decoded(src, size_t(i), sp.tf), sp);
}

// Row pass of an image spectrum: reads the frame, writes the grid.
// The scalar driver of a pixel (applyMaskScalarDriver, LensDiffCpuReference.cpp:1266-1279):
// signal (max RGB or Rec.709 luma of the decoded pixel) times mask times scale.
__device__ __forceinline__ float driver_decoded(float4 q, float m, int extraction, float scale) {
    const float sig = extraction == LDH_EXTRACT_LUMA ? luma709(q.x, q.y, q.z) : max3(q.x, q.y, q.z);
    return sig * m * scale;
}
__device__ __forceinline__ float driver_value(float4 raw, float m, int tf, int extraction, float scale) {
    const float4 q = make_float4(decode_channel(raw.x, tf), decode_channel(raw.y, tf), decode_channel(raw.z, tf), raw.w);
    return driver_decoded(q, m, extraction, scale);
}

// Separation of the transform Z of z = a + i b (a, b real) into half spectra:
// A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = (Z[k] - conj Z[N-k]) / 2i.
__device__ __forceinline__ void split_real_pair(float2 zk, float2 zm, float2& a, float2& b) {
    a = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
    b = make_float2(0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x));
}

struct FrameRows {
    const float4* src;
    const float* mask;
    float2* out;
    int w, h, pw, tf, extraction;
    int driver;     // 1: scalar driver; 0: channels ca (+ i cb)
    int ca, cb;     // cb < 0: zero imaginary part
    float scale;
    // fetch: the pixel and its mask value, branch-free (a point outside the frame reads pixel 0);
    // finish: decode and weigh (0 outside the frame).
    struct Raw {
        float4 q;
        float m;
        bool in;
    };
    __device__ Raw fetch(int row, int x) const {
        const bool in = row < h && x < w;
        const size_t i = in ? size_t(row) * w + x : 0;
        return Raw{src[i], mask[i], in};
    }
    __device__ float2 finish(const Raw& t) const {
        if (!t.in) return make_float2(0.0f, 0.0f);
        const float4 q = make_float4(decode_channel(t.q.x, tf), decode_channel(t.q.y, tf), decode_channel(t.q.z, tf), t.q.w);
        const float m = t.m;
        if (driver) return make_float2(driver_value(t.q, m, tf, extraction, scale), 0.0f);
        const float k = m * scale;
        const float c[3] = {q.x, q.y, q.z};
        return make_float2(c[ca] * k, cb >= 0 ? c[cb] * k : 0.0f);
    }
    __device__ void store(int row, int x, float2 v) const { out[size_t(row) * pw + x] = v; }
};

// Row pass of a kernel spectrum: kernel centre wrapped to the origin, two real kernels packed.
struct KernelRows {
    const float* ka;
    const float* kb;
    const float* sa;  // texel multipliers (nullable: 1)
    const float* sb;
    float2* out;
    int K, c, r, pw, ph;
    // fetch: the texels of both kernels, branch-free (a point outside the support reads the centre
    // texel; a missing second kernel reads the first); finish: scales, and 0 outside the support.
    struct Raw {
        float a, b;
        bool in;
    };
    __device__ Raw fetch(int row, int x) const {
        const int dy = row < ph / 2 ? row : row - ph;
        const int dx = x < pw / 2 ? x : x - pw;
        const bool in = dx >= -r && dx <= r && dy >= -r && dy <= r;
        const size_t i = in ? size_t(c + dy) * K + size_t(c + dx) : size_t(c) * K + size_t(c);
        const float* kbp = kb ? kb : ka;
        return Raw{ka[i], kbp[i], in};
    }
    __device__ float2 finish(const Raw& t) const {
        if (!t.in) return make_float2(0.0f, 0.0f);
        const float a = sa ? t.a * *sa : t.a;
        const float b = kb ? (sb ? t.b * *sb : t.b) : 0.0f;
        return make_float2(a, b);
    }
    __device__ void store(int row, int x, float2 v) const { out[size_t(row) * pw + x] = v; }
};

// Row pass of the scalar driver into a half spectrum: driver rows 2p and 2p + 1 transformed
// together as a + i b and separated into rows 2p, 2p + 1 (pitch hw; columns past pw/2 zero).
// Rows from h on are zero and not stored. The mask is evaluated from the pixel (select_decoded).
struct DriverRowsHalf {
    const float4* src;
    SelectParams sp;
    float2* out;
    int w, h, hw, extraction;
    float scale;
    static constexpr bool kStoreLds = true;
    static constexpr int kFetchBatch = 6;  // two pixels a load: fewer in flight, more waves
    struct Raw {
        float4 qa, qb;
        bool ina, inb;
    };
    __device__ Raw fetch(int p, int x) const {
        const int ra = 2 * p, rb = 2 * p + 1;
        const bool ina = ra < h && x < w, inb = rb < h && x < w;
        const size_t ia = ina ? size_t(ra) * w + x : 0, ib = inb ? size_t(rb) * w + x : 0;
        return Raw{src[ia], src[ib], ina, inb};
    }
    __device__ float value(float4 raw) const {
        const float4 q = make_float4(decode_channel(raw.x, sp.tf), decode_channel(raw.y, sp.tf),
                                     decode_channel(raw.z, sp.tf), raw.w);
        return driver_decoded(q, select_decoded(q, sp), extraction, scale);
    }
    __device__ float2 finish(const Raw& t) const {
        return make_float2(t.ina ? value(t.qa) : 0.0f, t.inb ? value(t.qb) : 0.0f);
    }
    __device__ void store(int, int, float2) const {}
    __device__ void store_sym(int p, int k, float2 zk, float2 zm, int n) const {
        float2 a = make_float2(0.0f, 0.0f), b = a;
        if (k <= n / 2) split_real_pair(zk, zm, a, b);
        out[size_t(2 * p) * hw + k] = a;
        if (2 * p + 1 < h) out[size_t(2 * p + 1) * hw + k] = b;
    }
};

// Kernel rows of two real kernels transformed together (KernelRows' loads), separated into the
// half spectra of each (outa, outb: pitch hw).
struct KernelRowsHalf : KernelRows {
    float2* outb;
    int hw;
    static constexpr bool kStoreLds = true;
    __device__ void store_sym(int row, int k, float2 zk, float2 zm, int n) const {
        float2 a = make_float2(0.0f, 0.0f), b = a;
        if (k <= n / 2) split_real_pair(zk, zm, a, b);
        out[size_t(row) * hw + k] = a;
        if (outb) outb[size_t(row) * hw + k] = b;
    }
};

// Last inverse pass from half spectra: the rows of two real planes' half spectra a, b (pitch hw,
// b nullable) extended by symmetry to Z = A + i B, so one inverse row transform gives plane a in
// its real and plane b in its imaginary part; cropped, scaled and clamped as ExtractRows.
struct ExtractRowsHalf {
    const float2* a;
    const float2* b;
    float* oa;
    float* ob;
    int w, h, hw, n;
    float scale;
    struct Raw {
        float2 xa, xb;
        bool mirror;
    };
    __device__ Raw fetch(int row, int x) const {
        const bool mirror = x > n / 2;
        const size_t i = size_t(row) * hw + (mirror ? n - x : x);
        return Raw{a[i], b ? b[i] : make_float2(0.0f, 0.0f), mirror};
    }
    __device__ float2 finish(const Raw& t) const {
        const float s = t.mirror ? -1.0f : 1.0f;  // conjugates past pw/2
        return make_float2(t.xa.x - s * t.xb.y, s * t.xa.y + t.xb.x);
    }
    __device__ void store(int row, int x, float2 v) const {
        if (row >= h || x >= w) return;
        const size_t i = size_t(row) * w + x;
        oa[i] = fmaxf(0.0f, v.x * scale);
        if (ob) ob[i] = fmaxf(0.0f, v.y * scale);
    }
};

// Last inverse pass: crop to the frame, scale by 1/(pw*ph) (exact), clamp at zero.
struct ExtractRows {
    float2* data;
    float* oa;
    float* ob;
    int w, h, pw;
    float scale;
    using Raw = float2;
    __device__ float2 fetch(int row, int x) const { return data[size_t(row) * pw + x]; }
    __device__ float2 finish(float2 v) const { return v; }
    __device__ void store(int row, int x, float2 v) const {
        if (row >= h || x >= w) return;
        const size_t i = size_t(row) * w + x;
        oa[i] = fmaxf(0.0f, v.x * scale);
        if (ob) ob[i] = fmaxf(0.0f, v.y * scale);
    }
};

// Row pass of the forward transform of two real planes packed as a + i b (b nullable: 0), zero
// outside the frame; fetch is branch-free (a point outside reads pixel 0).
struct PlaneRows {
    const float* a;
    const float* b;
    float2* out;
    int w, h, pw;
    struct Raw {
        float a, b;
        bool in;
    };
    __device__ Raw fetch(int row, int x) const {
        const bool in = row < h && x < w;
        const size_t i = in ? size_t(row) * w + x : 0;
        return Raw{a[i], b ? b[i] : 0.0f, in};
    }
    __device__ float2 finish(const Raw& t) const { return t.in ? make_float2(t.a, t.b) : make_float2(0.0f, 0.0f); }
    __device__ void store(int row, int x, float2 v) const { out[size_t(row) * pw + x] = v; }
};

__global__ void to_half_kernel(const float2* __restrict__ in, __half2* __restrict__ out, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = __float22half2_rn(in[i]);
}

// MapLensDiffSpectralBins per pixel over `n` bin planes -> 3 effect planes.
__global__ void spectral_map_kernel(const float* __restrict__ planes, int n, int64_t npx,
                                    const float* __restrict__ nat_m, const float* __restrict__ sty_m, float force,
                                    float saturation, int affects_luma, float* __restrict__ eff) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    float nat[3], sty[3];
    if (n <= 1) {
        const float v = planes[i];
        nat[0] = nat[1] = nat[2] = sty[0] = sty[1] = sty[2] = v;
    } else {
        for (int c = 0; c < 3; ++c) { nat[c] = 0.0f; sty[c] = 0.0f; }
        for (int b = 0; b < n; ++b) {
            const float v = planes[int64_t(b) * npx + i];
            for (int c = 0; c < 3; ++c) {
                nat[c] += nat_m[c * n + b] * v;
                sty[c] += sty_m[c * n + b] * v;
            }
        }
    }
    float o[3];
    for (int c = 0; c < 3; ++c) o[c] = fmaxf(0.0f, nat[c] * (1.0f - force) + sty[c] * force);
    const float gray = 0.2126f * o[0] + 0.7152f * o[1] + 0.0722f * o[2];
    for (int c = 0; c < 3; ++c) o[c] = gray + (o[c] - gray) * saturation;
    if (!affects_luma) {
        const float target = 0.2126f * nat[0] + 0.7152f * nat[1] + 0.0722f * nat[2];
        const float cur = 0.2126f * o[0] + 0.7152f * o[1] + 0.0722f * o[2];
        if (cur > 1e-6f) {
            const float k = target / cur;
            for (int c = 0; c < 3; ++c) o[c] *= k;
        }
    }
    for (int c = 0; c < 3; ++c) eff[int64_t(c) * npx + i] = o[c];
}

// The map of one pixel when the bins were pre-combined: natural R,G,B and style R,G,B in, the
// effect R,G,B out (the blend, saturation and luma steps of spectral_map_kernel).
struct PrecombinedMap {
    float force, saturation;
    int affects_luma;
    __device__ void apply(const float nat[3], const float sty[3], float o[3]) const {
        for (int c = 0; c < 3; ++c) o[c] = fmaxf(0.0f, nat[c] * (1.0f - force) + sty[c] * force);
        const float gray = 0.2126f * o[0] + 0.7152f * o[1] + 0.0722f * o[2];
        for (int c = 0; c < 3; ++c) o[c] = gray + (o[c] - gray) * saturation;
        if (!affects_luma) {
            const float target = 0.2126f * nat[0] + 0.7152f * nat[1] + 0.0722f * nat[2];
            const float cur = 0.2126f * o[0] + 0.7152f * o[1] + 0.0722f * o[2];
            if (cur > 1e-6f) {
                const float k = target / cur;
                for (int c = 0; c < 3; ++c) o[c] *= k;
            }
        }
    }
};

// The same map over planes holding natural R,G,B (and style R,G,B).
__global__ void precombined_map_kernel(const float* __restrict__ planes, int with_style, int64_t npx,
                                       PrecombinedMap map, float* __restrict__ eff) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    float nat[3], sty[3];
    for (int c = 0; c < 3; ++c) {
        nat[c] = planes[int64_t(c) * npx + i];
        sty[c] = with_style ? planes[int64_t(3 + c) * npx + i] : nat[c];
    }
    float o[3];
    map.apply(nat, sty, o);
    for (int c = 0; c < 3; ++c) eff[int64_t(c) * npx + i] = o[c];
}

__device__ __forceinline__ float soft_shoulder(float v, float sh) {
    const float x = fmaxf(0.0f, v);
    return sh * (1.0f - expf(-x / sh));
}

// Split look: effect = shoulder(core) * core_gain + shoulder(structure) * structure_gain.
// With write_back, the shouldered core and structure replace their planes, as applyShoulder
// leaves the reference's images (LensDiffCpuReference.cpp:2126-2132) for the core and structure
// views.
__global__ void split_combine_kernel(float* __restrict__ core, float* __restrict__ structure, int64_t npx,
                                     float core_sh, float struct_sh, float core_gain, float struct_gain,
                                     int write_back, float* __restrict__ eff) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    for (int c = 0; c < 3; ++c) {
        float a = core[int64_t(c) * npx + i];
        float b = structure[int64_t(c) * npx + i];
        if (core_sh > 0.0f) {
            a = soft_shoulder(a, core_sh);
            if (write_back) core[int64_t(c) * npx + i] = a;
        }
        if (struct_sh > 0.0f) {
            b = soft_shoulder(b, struct_sh);
            if (write_back) structure[int64_t(c) * npx + i] = b;
        }
        eff[int64_t(c) * npx + i] = a * core_gain + b * struct_gain;
    }
}

__global__ void zero_kernel(float* __restrict__ x, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        x[i] = 0.0f;
}

// accumulateWeightedImage (LensDiffCpuReference.cpp:1452-1475) for zone (zx, zy): acc += zone *
// wx * wy over `planes` planes, the tent weights centred on x = zx (w-1)/2 and y = zy (h-1)/2;
// pixels where the weight is zero are not touched.
__global__ void zone_accumulate_kernel(const float* __restrict__ zone, float* __restrict__ acc, int planes, int w,
                                       int h, int zx, int zy) {
    const int64_t npx = int64_t(w) * h;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    const int x = int(i % w), y = int(i / w);
    const float py = (float(y) / float(max(1, h - 1))) * 2.0f;
    const float wy = fmaxf(0.0f, 1.0f - fabsf(py - float(zy)));
    const float px = (float(x) / float(max(1, w - 1))) * 2.0f;
    const float wx = fmaxf(0.0f, 1.0f - fabsf(px - float(zx)));
    const float wt = wx * wy;
    if (!(wy > 0.0f) || !(wt > 0.0f)) return;
    for (int c = 0; c < planes; ++c) acc[int64_t(c) * npx + i] += zone[int64_t(c) * npx + i] * wt;
}

// Luma of the redistributed image (segment 0) and of the effect (segment 1), for Preserve.
__global__ void preserve_luma_kernel(const float4* __restrict__ src, const float* __restrict__ mask, int tf,
                                     float scale, const float* __restrict__ eff, int64_t npx, float* __restrict__ lum) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    const float4 q = decoded(src, size_t(i), tf);
    const float m = mask[i] * scale;
    lum[i] = luma709(q.x * m, q.y * m, q.z * m);
    lum[npx + i] = luma709(eff[i], eff[npx + i], eff[2 * npx + i]);
}

__global__ void preserve_factor_kernel(const float* __restrict__ totals, float* __restrict__ factor) {
    const float in_e = totals[0], eff_e = totals[1];
    factor[0] = eff_e > 1e-6f ? in_e / eff_e : 1.0f;
    factor[1] = eff_e > 1e-6f ? 1.0f : 0.0f;
}

__global__ void max_kernel(const float* __restrict__ x, int64_t n, unsigned* __restrict__ out) {
    unsigned m = 0u;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        m = max(m, __float_as_uint(fmaxf(x[i], 0.0f)));
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// max |x| as float bits (non-negative floats order like their bit patterns)
__global__ void max_abs_kernel(const float* __restrict__ x, int64_t n, unsigned* __restrict__ out) {
    unsigned m = 0u;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        m = max(m, __float_as_uint(fabsf(x[i])));
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// An n x n square for the static views: texel (x, y) at data[y * pitch + x], multiplied by
// *scale when scale is set; maxbits holds the bits of the largest stored texel.
struct SquareSrc {
    const float* data = nullptr;
    int n = 1, pitch = 1;
    const float* scale = nullptr;
    const unsigned* maxbits = nullptr;
};

// sampleSquareBilinear (LensDiffCpuReference.cpp:316-335) of a square, zero outside it.
__device__ float sample_square(const SquareSrc& q, float sc, float x, float y) {
    const int n = q.n;
    if (x < 0.0f || y < 0.0f || x > float(n - 1) || y > float(n - 1)) return 0.0f;
    const int x0 = int(floorf(x)), y0 = int(floorf(y));
    const int x1 = min(x0 + 1, n - 1), y1 = min(y0 + 1, n - 1);
    const float tx = x - float(x0), ty = y - float(y0);
    const float v00 = q.data[int64_t(y0) * q.pitch + x0] * sc, v10 = q.data[int64_t(y0) * q.pitch + x1] * sc;
    const float v01 = q.data[int64_t(y1) * q.pitch + x0] * sc, v11 = q.data[int64_t(y1) * q.pitch + x1] * sc;
    const float a = v00 + (v10 - v00) * tx, b = v01 + (v11 - v01) * tx;
    return a + (b - a) * ty;
}

// Where output texel (x, y) falls in the centred square of side min(ow, oh) mapped onto an n x n
// image (visualizeSquareCentered, :1416-1450), or false outside the square.
__device__ bool square_coords(int x, int y, int ow, int oh, int n, float& sx, float& sy) {
    const int side = min(ow, oh), xo = (ow - side) / 2, yo = (oh - side) / 2;
    if (x < xo || x >= xo + side || y < yo || y >= yo + side) return false;
    sx = (float(x - xo) / float(max(1, side - 1))) * float(n - 1);
    sy = (float(y - yo) / float(max(1, side - 1))) * float(n - 1);
    return true;
}

// visualizeSquareCentered: channel c of dst from square src[c] (the same square three times for
// the grey views; last, middle and first bin for the chromatic split, :1951-1972), v / max or
// log1p(1024 v) / log1p(1024 max); alpha 1.
__global__ void square_view_kernel(SquareSrc r, SquareSrc g, SquareSrc b, int log_view, float4* __restrict__ dst,
                                   int ow, int oh) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(ow) * oh) return;
    const int x = int(i % ow), y = int(i / ow);
    const SquareSrc* src[3] = {&r, &g, &b};
    float d[3];
    for (int c = 0; c < 3; ++c) {
        const SquareSrc& q = *src[c];
        const float sc = q.scale ? *q.scale : 1.0f;
        const float mx = __uint_as_float(*q.maxbits) * sc;
        float sx, sy;
        d[c] = 0.0f;
        if (!square_coords(x, y, ow, oh, q.n, sx, sy)) continue;
        const float v = sample_square(q, sc, sx, sy);
        if (log_view) {
            const float lmax = log1pf(fmaxf(0.0f, mx) * 1024.0f);
            d[c] = lmax > 0.0f ? log1pf(fmaxf(0.0f, v) * 1024.0f) / lmax : 0.0f;
        } else {
            d[c] = mx > 0.0f ? v / mx : 0.0f;
        }
    }
    dst[i] = make_float4(d[0], d[1], d[2], 1.0f);
}

// visualizeSignedPhaseCentered (:1536-1593): inside the pupil (amplitude > 1e-4), neutral grey
// moved towards warm for positive and cool for negative waves by |phase| / max |phase|.
__global__ void phase_view_kernel(SquareSrc phase, SquareSrc pupil, float4* __restrict__ dst, int ow, int oh) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(ow) * oh) return;
    const int x = int(i % ow), y = int(i / ow);
    float4 o = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    float sx, sy;
    if (square_coords(x, y, ow, oh, phase.n, sx, sy) && sample_square(pupil, 1.0f, sx, sy) > 1e-4f) {
        const float mabs = __uint_as_float(*phase.maxbits);
        const float v = sample_square(phase, 1.0f, sx, sy);
        const float u = mabs > 0.0f ? fminf(fmaxf(v / mabs, -1.0f), 1.0f) : 0.0f;
        const float m = fabsf(u), g = 0.18f;
        if (m <= 1e-6f) o = make_float4(g, g, g, 1.0f);
        else if (u > 0.0f) o = make_float4(g + 0.82f * m, g + 0.28f * m, fmaxf(0.0f, g - 0.12f * m), 1.0f);
        else o = make_float4(fmaxf(0.0f, g - 0.10f * m), g + 0.34f * m, g + 0.80f * m, 1.0f);
    }
    dst[i] = o;
}

// OTF view (computeOtfMagnitude, LensDiffCpuReference.cpp:1403-1414): the n x n DFT of the kernel
// wrapped to the origin, as the reference's FFT computes it for odd n (Bluestein with its chirp
// angle pi i^2 / n taken with pi rounded to float, which is the DFT with kernel
// exp(-2 i pi_f x k / n)). Evaluated directly in FP32, one output per thread. That kernel is not
// periodic in x k mod n (pi_f != pi), so each term's twiddle is the table value
// exp(-2 i pi (x k mod n) / n) (computed in double on the host, rounded once) times
// exp(-2 i (pi_f - pi) x k / n), whose angle stays below 8e-4 for n <= 4097 and is formed from its
// series (1 - a^2 / 2, -a; the next terms are below 1e-11). Rows first: thread kx of workgroup
// (kx block, wrapped row py) sums row (py + c) % n of the normalised kernel, staged in LDS in
// wrapped order.
__device__ __forceinline__ float2 otf_twiddle(const float2* __restrict__ tw, int idx, int m, float eps) {
    const float2 t = tw[idx];
    const float a = eps * float(m);
    const float cr = 1.0f - 0.5f * a * a;
    return make_float2(t.x * cr + t.y * a, t.y * cr - t.x * a);
}

__global__ void otf_rows_kernel(const float* __restrict__ k, int pitch, int n, const float* __restrict__ scale,
                                const float2* __restrict__ tw, float eps, float2* __restrict__ out) {
    extern __shared__ float otf_row[];
    const int py = blockIdx.y, c = n / 2;
    const float sc = scale ? *scale : 1.0f;
    const float* row = k + int64_t((py + c) % n) * pitch;
    for (int px = threadIdx.x; px < n; px += blockDim.x) otf_row[px] = row[(px + c) % n] * sc;
    __syncthreads();
    const int kx = blockIdx.x * blockDim.x + threadIdx.x;
    if (kx >= n) return;
    float ar = 0.0f, ai = 0.0f;
    int idx = 0, m = 0;  // m = kx px, idx = m mod n
    for (int px = 0; px < n; ++px) {
        const float2 w = otf_twiddle(tw, idx, m, eps);
        const float v = otf_row[px];
        ar += v * w.x;
        ai += v * w.y;
        m += kx;
        idx += kx;
        if (idx >= n) idx -= n;
    }
    out[int64_t(py) * n + kx] = make_float2(ar, ai);
}

// Columns: thread ky of workgroup (ky block, column kx) sums column kx of the row transforms
// (staged in LDS) against the twiddles of ky py, and writes |X| at its fftShiftSquare position
// (:303-314).
__global__ void otf_cols_kernel(const float2* __restrict__ rows, int n, const float2* __restrict__ tw, float eps,
                                float* __restrict__ mag) {
    extern __shared__ float2 otf_col[];
    const int kx = blockIdx.y;
    for (int py = threadIdx.x; py < n; py += blockDim.x) otf_col[py] = rows[int64_t(py) * n + kx];
    __syncthreads();
    const int ky = blockIdx.x * blockDim.x + threadIdx.x;
    if (ky >= n) return;
    float ar = 0.0f, ai = 0.0f;
    int idx = 0, m = 0;
    for (int py = 0; py < n; ++py) {
        const float2 w = otf_twiddle(tw, idx, m, eps);
        const float2 v = otf_col[py];
        ar += v.x * w.x - v.y * w.y;
        ai += v.x * w.y + v.y * w.x;
        m += ky;
        idx += ky;
        if (idx >= n) idx -= n;
    }
    const int half = n / 2;
    const int ys = (ky - half + n) % n, xs = (kx - half + n) % n;
    mag[int64_t(ys) * n + xs] = sqrtf(ar * ar + ai * ai);
}

// Bilinear sample at (x, y) in the form of sampleScalarBilinear / sampleRgbaBilinearImage
// (LensDiffCpuReference.cpp:355-394): zero outside [0, w-1] x [0, h-1].
struct Bilinear {
    bool in = false;
    int64_t i00 = 0, i10 = 0, i01 = 0, i11 = 0;
    float tx = 0.0f, ty = 0.0f;
    __device__ Bilinear(float x, float y, int w, int h) {
        if (x < 0.0f || y < 0.0f || x > float(w - 1) || y > float(h - 1)) return;
        in = true;
        const int x0 = int(floorf(x)), y0 = int(floorf(y));
        const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
        tx = x - float(x0);
        ty = y - float(y0);
        i00 = int64_t(y0) * w + x0;
        i10 = int64_t(y0) * w + x1;
        i01 = int64_t(y1) * w + x0;
        i11 = int64_t(y1) * w + x1;
    }
    __device__ float mix(float v00, float v10, float v01, float v11) const {
        const float a = v00 + (v10 - v00) * tx;
        const float b = v01 + (v11 - v01) * tx;
        return a + (b - a) * ty;
    }
};

// Source position of output texel x when w texels are resampled to ow (resampleRgbaImage,
// resampleScalarImage, :396-434).
__device__ __forceinline__ float resample_at(int x, int w, int ow) {
    return (float(x) + 0.5f) * (float(w) / float(max(1, ow))) - 0.5f;
}

// The decoded frame (loadSourceRgba) resampled to the working size ow x oh
// (LensDiffCpuReference.cpp:2090-2093).
__global__ void resample_frame_kernel(const float4* __restrict__ src, int w, int h, int tf, float4* __restrict__ out,
                                      int ow, int oh) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(ow) * oh) return;
    const int x = int(i % ow), y = int(i / ow);
    if (ow == w && oh == h) {  // resampleRgbaImage returns the image unchanged at its own size
        out[i] = decoded(src, size_t(i), tf);
        return;
    }
    const Bilinear b(resample_at(x, w, ow), resample_at(y, h, oh), w, h);
    if (!b.in) {
        out[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float4 p00 = decoded(src, size_t(b.i00), tf), p10 = decoded(src, size_t(b.i10), tf);
    const float4 p01 = decoded(src, size_t(b.i01), tf), p11 = decoded(src, size_t(b.i11), tf);
    out[i] = make_float4(b.mix(p00.x, p10.x, p01.x, p11.x), b.mix(p00.y, p10.y, p01.y, p11.y),
                         b.mix(p00.z, p10.z, p01.z, p11.z), b.mix(p00.w, p10.w, p01.w, p11.w));
}

// `planes` planes of w x h resampled to ow x oh (the display images of a resolution-aware
// render, :2192-2198).
__global__ void resample_planes_kernel(const float* __restrict__ in, int w, int h, int planes,
                                       float* __restrict__ out, int ow, int oh) {
    const int64_t n = int64_t(ow) * oh;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = int(i % ow), y = int(i / ow);
    const int64_t nin = int64_t(w) * h;
    if (ow == w && oh == h) {
        for (int c = 0; c < planes; ++c) out[int64_t(c) * n + i] = in[int64_t(c) * nin + i];
        return;
    }
    const Bilinear b(resample_at(x, w, ow), resample_at(y, h, oh), w, h);
    for (int c = 0; c < planes; ++c) {
        const float* pl = in + int64_t(c) * nin;
        out[int64_t(c) * n + i] = b.in ? b.mix(pl[b.i00], pl[b.i10], pl[b.i01], pl[b.i11]) : 0.0f;
    }
}

// The redistributed image (applyMask, :1251-1265) as 3 planes: frame * (mask * scale).
__global__ void redist_planes_kernel(const float4* __restrict__ work, const float* __restrict__ mask, float scale,
                                     int64_t n, float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q = work[i];
    const float k = mask[i] * scale;
    out[i] = q.x * k;
    out[n + i] = q.y * k;
    out[2 * n + i] = q.z * k;
}

struct CompositeParams {
    int tf, view;
    float scale;        // redistribution scale 1 - corePreserve
    float comp, gain;   // core compensation, effect gain
    float max_sub;      // (1 - protect) / scale
};

// The redistributed image of one pixel: the decoded frame s times mask * (1 - corePreserve).
__device__ __forceinline__ void redistributed(float4 s, float mask, const CompositeParams& cp, float r[3]) {
    const float m = mask * cp.scale;
    r[0] = s.x * m;
    r[1] = s.y * m;
    r[2] = s.z * m;
}

// The final image of one pixel from the decoded frame s, its redistributed image r and the effect
// e: max(s - comp * r + gain * e, s - r * max_sub), clamped at zero and encoded; alpha from s.
__device__ __forceinline__ float4 composite_final(float4 s, const float r[3], const float e[3],
                                                  const CompositeParams& cp) {
    const float sv[3] = {s.x, s.y, s.z};
    float o[3];
    for (int c = 0; c < 3; ++c) {
        const float raw = fmaxf(0.0f, sv[c] - cp.comp * r[c] + cp.gain * e[c]);
        const float flo = fmaxf(0.0f, sv[c] - r[c] * cp.max_sub);
        o[c] = encode_channel(fmaxf(raw, flo), cp.tf);
    }
    return make_float4(o[0], o[1], o[2], s.w);
}

// eff *= factor[0] where the Preserve rescale applies (factor[1] > 0): scaleImage on the effect
// before the post passes read it (LensDiffCpuReference.cpp:2172-2178).
__global__ void scale_planes_kernel(float* __restrict__ eff, int64_t n, const float* __restrict__ factor) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !(factor[1] > 0.0f)) return;
    eff[i] *= factor[0];
}

// eff = eff * 1 + add * gain over 3 planes (addImages, :2185).
__global__ void add_scaled_kernel(float* __restrict__ eff, const float* __restrict__ add, int64_t n, float gain) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    eff[i] = eff[i] * 1.0f + add[i] * gain;
}

// Bilinear sample of one plane at (fx, fy) clamped to the frame (sampleRgbaBilinear, :1477-1497).
__device__ __forceinline__ float sample_plane(const float* __restrict__ pl, int w, int h, float fx, float fy) {
    const float x = fminf(fmaxf(fx, 0.0f), float(max(0, w - 1)));
    const float y = fminf(fmaxf(fy, 0.0f), float(max(0, h - 1)));
    const int x0 = int(floorf(x)), y0 = int(floorf(y));
    const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
    const float tx = x - float(x0), ty = y - float(y0);
    const float a = pl[int64_t(y0) * w + x0] * (1.0f - tx) + pl[int64_t(y0) * w + x1] * tx;
    const float b = pl[int64_t(y1) * w + x0] * (1.0f - tx) + pl[int64_t(y1) * w + x1] * tx;
    return a * (1.0f - ty) + b * ty;
}

// applyCreativeFringe (:1499-1534): red sampled `amount` pixels outward along the ray from the
// frame centre, blue as far inward, green in place; preview = |fringed - effect| per channel.
__global__ void fringe_kernel(const float* __restrict__ eff, float* __restrict__ out, float* __restrict__ preview,
                              int w, int h, float amount) {
    const int64_t npx = int64_t(w) * h;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    const int x = int(i % w), y = int(i / w);
    const float cx = float(max(0, w - 1)) * 0.5f, cy = float(max(0, h - 1)) * 0.5f;
    const float dx = float(x) - cx, dy = float(y) - cy;
    const float len = sqrtf(dx * dx + dy * dy);
    const float il = len > 1e-6f ? 1.0f / len : 0.0f;
    const float sx = dx * il * amount, sy = dy * il * amount;
    const float v[3] = {sample_plane(eff, w, h, float(x) + sx, float(y) + sy),
                        sample_plane(eff + npx, w, h, float(x), float(y)),
                        sample_plane(eff + 2 * npx, w, h, float(x) - sx, float(y) - sy)};
    for (int c = 0; c < 3; ++c) {
        out[int64_t(c) * npx + i] = v[c];
        preview[int64_t(c) * npx + i] = fabsf(v[c] - eff[int64_t(c) * npx + i]);
    }
}

// redist (nullable): the redistributed image as 3 planes (a resolution-aware render resamples it
// from the working frame); otherwise it is formed here from the frame and the mask.
__global__ void composite_kernel(const float4* __restrict__ src, const float* __restrict__ mask,
                                 const float* __restrict__ eff, const float* __restrict__ factor,
                                 const unsigned* __restrict__ mask_max, int64_t npx, CompositeParams cp,
                                 float4* __restrict__ dst, const float* __restrict__ redist) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    const float4 s = decoded(src, size_t(i), cp.tf);
    if (cp.view == LDH_VIEW_SELECTION) {
        const float mx = __uint_as_float(*mask_max);
        const float im = mx > 0.0f ? 1.0f / mx : 0.0f;
        const float d = fminf(fmaxf(mask[i] * im, 0.0f), 1.0f);
        dst[i] = make_float4(d, d, d, 1.0f);
        return;
    }
    const bool scaled = factor != nullptr && factor[1] > 0.0f;
    float e[3];
    for (int c = 0; c < 3; ++c) {
        e[c] = eff[int64_t(c) * npx + i];
        if (scaled) e[c] *= factor[0];
    }
    if (cp.view != LDH_VIEW_FINAL) {
        dst[i] = make_float4(e[0], e[1], e[2], 1.0f);
        return;
    }
    float r[3];
    if (redist) {
        for (int c = 0; c < 3; ++c) r[c] = redist[int64_t(c) * npx + i];
    } else {
        redistributed(s, mask[i], cp, r);
    }
    dst[i] = composite_final(s, r, e, cp);
}

// The rows of a combine: m[r * bins + b] for r < ROWS, passed by value.
struct CombineMatrix {
    float m[6 * LDH_MAX_WAVELENGTHS];
};

// out[r] = sum_b m[r][b] * (kernels[b] * scales[b]) (ROWS = 3 or 6), accumulated over b in order
// for every r; each kernel texel is read once, eight bins' texels loaded before they are used.
template <int ROWS>
__global__ void __launch_bounds__(256)
combine_kernel(const float* __restrict__ kernels, const float* __restrict__ scales, int bins, int64_t kk,
               CombineMatrix cm, float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= kk) return;
    float acc[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) acc[r] = 0.0f;
    int b = 0;
    for (; b + 8 <= bins; b += 8) {
        float k[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) k[j] = kernels[int64_t(b + j) * kk + i];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float v = k[j] * scales[b + j];
#pragma unroll
            for (int r = 0; r < ROWS; ++r) acc[r] += cm.m[r * bins + b + j] * v;
        }
    }
    for (; b < bins; ++b) {
        const float v = kernels[int64_t(b) * kk + i] * scales[b];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) acc[r] += cm.m[r * bins + b] * v;
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) out[int64_t(r) * kk + i] = acc[r];
}

inline dim3 grid1(int64_t n) { return dim3(unsigned(div_up(n, 256))); }

}  // namespace

Grid make_grid(int w, int h, int support) {
    Grid g;
    g.w = w;
    g.h = h;
    const int n = 2 * support + 1;
    // Any grid of at least (w + n - 1) x (h + n - 1) holds the linear convolution without wrap,
    // so the cropped result is the reference's (which pads to powers of two, :1168-1169); the
    // smallest 2^a * 3^b length keeps the transforms short.
    g.pw = grid_length_at_least(std::max(16, w + n - 1));
    g.ph = grid_length_at_least(std::max(16, h + n - 1));
    if (g.pw == 0 || g.ph == 0) fail(LDH_ERR_TOO_LARGE, "frame-too-large-for-fft-grid");
    g.hw = (g.pw / 2 + 1 + kColWidth - 1) / kColWidth * kColWidth;
    return g;
}

void build_kernel_spectra(hipStream_t s, const FftTables& t, const std::vector<KernelRef>& kernels, int K,
                          int radius, const Grid& g, bool fp16, KernelSpectra& out, DeviceBuffer& scratch, bool half) {
    const int n = int(kernels.size());
    const int pairs = (n + 1) / 2;
    out.grid = g;
    out.kernels = n;
    out.pairs = half ? n : pairs;
    out.fp16 = fp16;
    out.half = half;
    const size_t cells = g.cells();
    const float2* tw = t.twiddle.as<float2>();
    if (half) {
        // each pair transformed together along the rows, separated into two half spectra
        const size_t hc = g.half_cells();
        out.data.reset((fp16 ? sizeof(__half2) : sizeof(float2)) * hc * n);
        if (fp16) scratch.reset(sizeof(float2) * hc * 2);
        // only the rows [0, r] and [ph - r, ph) of the wrapped kernels are non-zero
        ColRows zero;
        zero.zero_lo = radius + 1;
        zero.zero_hi = g.ph - radius;
        for (int p = 0; p < pairs; ++p) {
            const bool two = 2 * p + 1 < n;
            float2* da = fp16 ? scratch.as<float2>() : out.data.as<float2>() + hc * (2 * p);
            float2* db = !two ? nullptr : fp16 ? scratch.as<float2>() + hc : out.data.as<float2>() + hc * (2 * p + 1);
            const KernelRef ka = kernels[size_t(2 * p)];
            const KernelRef kb = two ? kernels[size_t(2 * p + 1)] : KernelRef{};
            KernelRowsHalf op;
            static_cast<KernelRows&>(op) = KernelRows{ka.data, kb.data, ka.scale, kb.scale, da, K, K / 2, radius, g.pw,
                                                      g.ph};
            op.outb = db;
            op.hw = g.hw;
            const int band = 2 * radius + 1;
            rows_pass(s, tw, g.pw, band, op, false, radius + 1, g.ph - band);
            if (!fp16) continue;  // FP32 spectra: every kernel's columns below, in the same launches
            for (int j = 0; j < (two ? 2 : 1); ++j) {
                float2* d = j == 0 ? da : db;
                cols_forward(s, tw, d, g.hw, g.ph, PlainLoad{}, zero);
                hipLaunchKernelGGL(to_half_kernel, grid1(int64_t(hc)), dim3(256), 0, s, d,
                                   out.data.as<__half2>() + hc * (2 * p + j), int64_t(hc));
                LDH_CHECK_LAUNCH("spectra-to-half");
            }
        }
        if (!fp16) cols_forward_channels(s, tw, out.data.as<float2>(), hc, n, g.hw, g.ph, zero);
        return;
    }
    out.data.reset((fp16 ? sizeof(__half2) : sizeof(float2)) * cells * pairs);
    if (fp16) scratch.reset(sizeof(float2) * cells);
    for (int p = 0; p < pairs; ++p) {
        float2* dst = fp16 ? scratch.as<float2>() : out.data.as<float2>() + cells * p;
        const KernelRef ka = kernels[size_t(2 * p)];
        const KernelRef kb = 2 * p + 1 < n ? kernels[size_t(2 * p + 1)] : KernelRef{};
        KernelRows op{ka.data, kb.data, ka.scale, kb.scale, dst, K, K / 2, radius, g.pw, g.ph};
        // only the rows [0, r] and [ph - r, ph) of the wrapped kernel are non-zero
        const int band = 2 * radius + 1;
        rows_pass(s, tw, g.pw, band, op, false, radius + 1, g.ph - band);
        ColRows zero;
        zero.zero_lo = radius + 1;
        zero.zero_hi = g.ph - radius;
        cols_forward(s, tw, dst, g.pw, g.ph, PlainLoad{}, zero);
        if (fp16) {
            hipLaunchKernelGGL(to_half_kernel, grid1(int64_t(cells)), dim3(256), 0, s, dst,
                               reinterpret_cast<__half2*>(out.data.as<__half2>()) + cells * p, int64_t(cells));
            LDH_CHECK_LAUNCH("spectra-to-half");
        }
    }
}

void combine_kernels(hipStream_t s, const float* kernels, const float* scales, int bins, int K,
                     const SpectrumConfig& cfg, bool with_style, float* out) {
    if (bins < 1 || bins > LDH_MAX_WAVELENGTHS) fail(LDH_ERR_INVALID_ARGUMENT, "combine-bins");
    const int rows = with_style ? 6 : 3;
    CombineMatrix cm{};
    for (int r = 0; r < 3; ++r)
        for (int b = 0; b < bins; ++b) {
            cm.m[r * bins + b] = cfg.natural[size_t(r * bins + b)];
            if (with_style) cm.m[(3 + r) * bins + b] = cfg.style[size_t(r * bins + b)];
        }
    const int64_t kk = int64_t(K) * K;
    ProfileScope ps_("spectra-combine", 4.0 * double(kk) * (bins + rows));
    if (with_style)
        hipLaunchKernelGGL(combine_kernel<6>, grid1(kk), dim3(256), 0, s, kernels, scales, bins, kk, cm, out);
    else
        hipLaunchKernelGGL(combine_kernel<3>, grid1(kk), dim3(256), 0, s, kernels, scales, bins, kk, cm, out);
    LDH_CHECK_LAUNCH("combine-kernels");
}

namespace {

// Which kernel set an effect needs.
enum Set { kFull = LDH_KERNEL_FULL, kCore = LDH_KERNEL_CORE, kStructure = LDH_KERNEL_STRUCTURE };

struct Plan {
    bool spectral = false;
    bool split = false;
    std::vector<int> sets;  // kernel sets convolved, in order
};

bool core_or_structure_view(const ldh_params& p) {
    return p.debug_view == LDH_VIEW_CORE || p.debug_view == LDH_VIEW_STRUCTURE;
}

Plan plan_of(const ldh_params& p) {
    Plan pl;
    pl.spectral = is_spectral(p);
    pl.split = p.look_mode == LDH_LOOK_SPLIT;
    if (pl.split) {
        pl.sets = {kCore, kStructure};
    } else if (p.debug_view == LDH_VIEW_CORE) {
        pl.sets = {kCore};
    } else if (p.debug_view == LDH_VIEW_STRUCTURE) {
        pl.sets = {kStructure};
    } else {
        pl.sets = {kFull};
    }
    if (pl.split && p.debug_view == LDH_VIEW_CORE) pl.sets = {kCore, kStructure};
    return pl;
}

}  // namespace

int kernel_sets_needed(const ldh_params& p) {
    int sets = 1;
    for (int s : plan_of(p).sets) sets = std::max(sets, s + 1);
    return sets;
}

std::vector<KernelRef> effect_kernels(const ldh_params& p, const DeviceBank& bank, EffectMode mode,
                                      const float* combined, int* which_out) {
    const Plan pl = plan_of(p);
    std::vector<KernelRef> ks;
    const int which = pl.sets.front();
    if (which_out) *which_out = which;
    if (!pl.spectral) {
        ks.push_back({bank.kernel_ptr(which, 0), bank.scale_ptr(which, 0)});
    } else if (mode == EffectMode::PerBin) {
        for (int b = 0; b < bank.bins; ++b) ks.push_back({bank.kernel_ptr(which, b), bank.scale_ptr(which, b)});
    } else {
        const int64_t kk = int64_t(bank.kernel) * bank.kernel;
        for (int r = 0; r < 6; ++r) ks.push_back({combined + r * kk, nullptr});
    }
    return ks;
}

namespace {

// Step 1: the selection mask of the frame into fw.mask.
// Whether effect_stage takes the pre-combined half-spectrum path (effect_precombined): the
// spectral full-kernel effect in EffectMode::Precombined.
bool precombined_path(const ldh_params& p, EffectMode mode) {
    const Plan pl = plan_of(p);
    return mode == EffectMode::Precombined && pl.spectral && !pl.split && pl.sets.front() == kFull;
}

SelectParams select_params(const ldh_params& p) {
    SelectParams sp;
    sp.tf = p.input_transfer;
    sp.extraction = p.extraction_mode;
    sp.softness = float(std::max(0.01, p.softness_stops));
    sp.threshold_stops = float(p.threshold);
    sp.threshold_linear = 0.18f * std::pow(2.0f, sp.threshold_stops);
    sp.point_emphasis = float(p.point_emphasis);
    return sp;
}

void select_stage(hipStream_t s, FrameWork& fw, const ldh_params& p, const float4* src, int64_t npx) {
    // 1. selection mask
    fw.mask.reset(sizeof(float) * npx);
    const SelectParams sp = select_params(p);
    ProfileScope ps_("frame-select", 20.0 * double(npx));
    hipLaunchKernelGGL(selection_kernel, grid1(npx), dim3(256), 0, s, src, npx, sp, fw.mask.as<float>());
    LDH_CHECK_LAUNCH("selection");
}

// The final composite folded into the effect's last pass: the output frame and the composite's
// parameters (for a frame whose effect goes to the output unchanged: no Preserve rescale, no
// scatter or fringe pass, no display resampling, the final view).
struct FusedFinal {
    float4* dst;
    CompositeParams cp;
};

// The last row pass of the pre-combined effect with the map and the composite folded in: the
// transform yields channels NK - 2 and NK - 1 (ExtractRowsHalf's pair), the channels before them
// are read from `planes` (channel c at planes + c * npx); each pixel then goes through
// PrecombinedMap and composite_final exactly as precombined_map_kernel and composite_kernel take
// it, with its selection mask evaluated from the pixel, and only the output pixel is stored.
// NK = 3: natural R, G, B; NK = 6: then style R, G, B.
template <int NK>
struct FinalRowsHalf : ExtractRowsHalf {
    const float* planes;
    const float4* src;
    SelectParams sp;  // the mask is evaluated from the pixel (select_decoded)
    float4* dst;
    int64_t npx;
    PrecombinedMap map;
    CompositeParams cp;
    __device__ void store(int row, int x, float2 v) const {
        if (row >= h || x >= w) return;
        const size_t i = size_t(row) * w + x;
        const float4 q = decoded(src, i, cp.tf);
        const float mk = select_decoded(q, sp);
        float ch[NK];
#pragma unroll
        for (int c = 0; c < NK - 2; ++c) ch[c] = planes[int64_t(c) * npx + i];
        ch[NK - 2] = fmaxf(0.0f, v.x * scale);
        ch[NK - 1] = fmaxf(0.0f, v.y * scale);
        float e[3], r[3];
        map.apply(ch, NK == 6 ? ch + 3 : ch, e);
        redistributed(q, mk, cp, r);
        dst[i] = composite_final(q, r, e, cp);
    }
};

// Steps 2-4 for the spectral full-kernel effect with pre-combined RGB kernels (natural, and style
// when styled), on half spectra: the driver is real, so its transform is stored for columns
// 0 .. pw/2 only (DriverRowsHalf, two rows per row transform); each combined kernel likewise;
// every channel's product is inverted over the half-width columns (one set of launches for all
// channels; unless keep_image asks for the driver's spectrum to stay in fw.image_spec for later
// zones, its last forward column pass runs fused with the first inverse one), and channels are
// grouped in pairs (an odd count starts with one alone) for the last row pass (ExtractRowsHalf),
// which yields two real planes per row transform. The planes then go through the pre-combined
// spectral map into effect9[0..3) or, with `fin`, the last group's pass maps and composites into
// fin->dst.
void effect_precombined(hipStream_t s, const FftTables& t, FrameWork& fw, const ldh_params& p, const DeviceBank& bank,
                        const float4* src, int w, int h, const Grid& g, const SpectrumConfig& cfg, bool with_style,
                        float force, float redist_scale, const KernelSpectra* cached, float* effect9,
                        bool image_ready, bool keep_image, const FusedFinal* fin) {
    const int64_t npx = int64_t(w) * h;
    const size_t hc = g.half_cells();
    const float2* tw = t.twiddle.as<float2>();
    const int nk = with_style ? 6 : 3;
    // 2. the driver's half spectrum
    fw.image_spec.reset(sizeof(float2) * hc);
    float2* ispec = fw.image_spec.as<float2>();
    // rows from h on are zero; the column transform runs here when the spectrum is kept for later
    // zones, else with the inverse passes below (cols_convolve_channels)
    ColRows zero;
    zero.zero_lo = h;
    zero.zero_hi = g.ph;
    const bool fuse_columns = !image_ready && !keep_image;
    if (!image_ready) {
        DriverRowsHalf op{src, select_params(p), ispec, w, h, g.hw, p.extraction_mode, redist_scale};
        rows_pass(s, tw, g.pw, (h + 1) / 2, op, false);
        if (!fuse_columns) cols_forward(s, tw, ispec, g.hw, g.ph, PlainLoad{}, zero);
    }
    // 3. half spectra of the combined kernels (cached with the lens, or built for this frame)
    const KernelSpectra* ks = cached;
    if (ks == nullptr) {
        const int64_t kk = int64_t(bank.kernel) * bank.kernel;
        fw.combined.reset(sizeof(float) * kk * 6);
        combine_kernels(s, bank.kernel_ptr(LDH_KERNEL_FULL, 0), bank.scale_ptr(LDH_KERNEL_FULL, 0), bank.bins,
                        bank.kernel, cfg, with_style, fw.combined.as<float>());
        std::vector<KernelRef> refs;
        for (int r = 0; r < nk; ++r) refs.push_back({fw.combined.as<float>() + r * kk, nullptr});
        build_kernel_spectra(s, t, refs, bank.kernel, bank.support, g, false, fw.kspec, fw.kspec_tmp, true);
        ks = &fw.kspec;
    }
    if (!ks->half || ks->grid.pw != g.pw || ks->grid.ph != g.ph || ks->kernels != nk)
        fail(LDH_ERR_INVALID_ARGUMENT, "cached-spectra-mismatch");
    // 4. every channel's product and inverse columns in the same launches (the driver's spectrum
    // read about once for all); per group: one row pass to one or two planes
    fw.out_spec.reset(sizeof(float2) * hc * nk);
    float2* outs = fw.out_spec.as<float2>();
    // only rows < h reach the crop
    if (fuse_columns) {
        if (ks->fp16)
            cols_convolve_channels(s, tw, ispec, zero, outs, hc, nk, g.hw, g.ph,
                                   MulSpectrumHalfLoad{ks->data.as<__half2>(), hc}, h);
        else
            cols_convolve_channels(s, tw, ispec, zero, outs, hc, nk, g.hw, g.ph,
                                   MulSpectrumLoad{ks->data.as<float2>(), hc}, h);
    } else if (ks->fp16) {
        cols_inverse_channels(s, tw, ispec, outs, hc, nk, g.hw, g.ph, MulSpectrumHalfLoad{ks->data.as<__half2>(), hc}, h);
    } else {
        cols_inverse_channels(s, tw, ispec, outs, hc, nk, g.hw, g.ph, MulSpectrumLoad{ks->data.as<float2>(), hc}, h);
    }
    fw.planes.reset(sizeof(float) * npx * 6);
    float* planes = fw.planes.as<float>();
    const float inv_cells = 1.0f / float(g.cells());
    const PrecombinedMap map{force, float(std::max(0.0, p.spectrum_saturation)), p.chromatic_affects_luma};
    for (int c = 0; c < nk;) {
        const int m = c == 0 && nk % 2 == 1 ? 1 : 2;
        const ExtractRowsHalf ex{outs + hc * c, m == 2 ? outs + hc * (c + 1) : nullptr, planes + npx * c,
                                 m == 2 ? planes + npx * (c + 1) : nullptr, w, h, g.hw, g.pw, inv_cells};
        c += m;
        if (c < nk || !fin) {
            rows_pass(s, tw, g.pw, h, ex, true);
            continue;
        }
        // spectra in, then per pixel the earlier planes, frame and mask in and the output pixel out
        const double bytes = 16.0 * double(g.hw) * h + (4.0 * (nk - 2) + 32.0) * double(npx);
        if (nk == 6)
            rows_pass(s, tw, g.pw, h, FinalRowsHalf<6>{ex, planes, src, select_params(p), fin->dst, npx, map, fin->cp},
                      true, 1 << 30, 0, "fft-row-pass-final", bytes);
        else
            rows_pass(s, tw, g.pw, h, FinalRowsHalf<3>{ex, planes, src, select_params(p), fin->dst, npx, map, fin->cp},
                      true, 1 << 30, 0, "fft-row-pass-final", bytes);
        return;
    }
    ProfileScope ps_map("frame-spectral-map", (4.0 * nk + 12.0) * double(npx));
    hipLaunchKernelGGL(precombined_map_kernel, grid1(npx), dim3(256), 0, s, planes, with_style ? 1 : 0, npx, map,
                       effect9);
    LDH_CHECK_LAUNCH("effect-map");
}

// Steps 2-4 for one bank: the effect planes [effect, core, structure] (3 each; only the effect
// planes unless the look is split) into effect9 (npx * 9 floats). shared_grid (nullable): the
// grid to convolve on instead of the bank's own (the zones of a field-dependent phase share one);
// image_ready: fw.image_spec already holds the frame's spectra on that grid. fin (nullable): the
// final composite to fold into the last pass where the effect takes the pre-combined path; returns
// whether it was (fin->dst then holds the output frame and effect9 is untouched).
bool effect_stage(hipStream_t s, const FftTables& t, FrameWork& fw, const ldh_params& p, const DeviceBank& bank,
                  const float4* src, int w, int h, EffectMode mode, const KernelSpectra* cached, float* effect9,
                  const Grid* shared_grid = nullptr, bool image_ready = false, const FusedFinal* fin = nullptr) {
    const Plan pl = plan_of(p);
    const int64_t npx = int64_t(w) * h;
    const Grid g = shared_grid != nullptr ? *shared_grid : make_grid(w, h, bank.support);
    const size_t cells = g.cells();
    const float2* tw = t.twiddle.as<float2>();
    const int n = bank.bins;
    const SpectrumConfig cfg = build_spectrum_config(p, bank.wavelengths);
    const float force = std::max(0.0f, std::min(float(p.spectrum_force), 1.0f));
    const bool with_style = p.spectrum_style != LDH_STYLE_NATURAL && force > 0.0f;
    if (!precombined_path(p, mode)) mode = EffectMode::PerBin;  // pre-combination: the spectral full-kernel effect
    const float redist_scale = 1.0f - float(std::max(0.0, std::min(p.core_preserve, 1.0)));

    if (mode == EffectMode::Precombined) {
        effect_precombined(s, t, fw, p, bank, src, w, h, g, cfg, with_style, force, redist_scale, cached, effect9,
                           image_ready, shared_grid != nullptr, fin);
        return fin != nullptr;
    }

    // 2. image spectra: mono -> (R + iG), (B); spectral -> driver
    const int nimg = pl.spectral ? 1 : 2;
    fw.image_spec.reset(sizeof(float2) * cells * nimg);
    for (int j = 0; j < nimg && !image_ready; ++j) {
        float2* spec = fw.image_spec.as<float2>() + cells * j;
        FrameRows op{src, fw.mask.as<float>(), spec, w, h, g.pw, p.input_transfer, p.extraction_mode,
                     pl.spectral ? 1 : 0, j == 0 ? 0 : 2, j == 0 ? 1 : -1, redist_scale};
        // rows >= h of the padded image are zero: not transformed, read as zero by the columns
        rows_pass(s, tw, g.pw, h, op, false);
        ColRows zero;
        zero.zero_lo = h;
        zero.zero_hi = g.ph;
        cols_forward(s, tw, spec, g.pw, g.ph, PlainLoad{}, zero);
    }

    // 3-4. per kernel set: spectra x image spectra -> planes -> 3 effect planes
    fw.out_spec.reset(sizeof(float2) * cells * nimg);
    const int nplanes = pl.spectral ? n : 3;
    fw.planes.reset(sizeof(float) * npx * std::max(nplanes, 3));
    fw.matrices.reset(sizeof(float) * 6 * n);
    LDH_HIP(hipMemcpyAsync(fw.matrices.as<float>(), cfg.natural.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice, s),
            "matrices");
    LDH_HIP(hipMemcpyAsync(fw.matrices.as<float>() + 3 * n, cfg.style.data(), sizeof(float) * 3 * n,
                           hipMemcpyHostToDevice, s), "matrices");
    const float inv_cells = 1.0f / float(cells);
    float* planes = fw.planes.as<float>();

    for (size_t si = 0; si < pl.sets.size(); ++si) {
        const int set = pl.sets[si];
        // kernels of this set
        std::vector<KernelRef> ks;
        const bool use_cache = cached != nullptr && si == 0 && pl.sets.size() == 1;
        if (use_cache) {
            // the cached spectra stand for the kernels; only their count matters here
            ks.assign(size_t(cached->kernels), KernelRef{});
        } else if (!pl.spectral) {
            ks.push_back({bank.kernel_ptr(set, 0), bank.scale_ptr(set, 0)});
        } else {
            for (int b = 0; b < n; ++b) ks.push_back({bank.kernel_ptr(set, b), bank.scale_ptr(set, b)});
        }
        const int pairs = pl.spectral ? int((ks.size() + 1) / 2) : 1;
        if (use_cache && (cached->grid.pw != g.pw || cached->grid.ph != g.ph || cached->pairs != pairs))
            fail(LDH_ERR_INVALID_ARGUMENT, "cached-spectra-mismatch");
        for (int pi = 0; pi < pairs; ++pi) {
            // kernel spectrum of this pair: cached, or its row pass here and its columns fused
            // with the products below (cols_convolve_channels: the kernel's transform never
            // stored); every product with an image spectrum inverted (only rows < h reach the crop)
            if (use_cache) {
                const bool half = cached->fp16;
                for (int j = 0; j < nimg; ++j) {
                    const float2* ispec = fw.image_spec.as<float2>() + cells * j;
                    float2* o = fw.out_spec.as<float2>() + cells * j;
                    if (half)
                        cols_inverse(s, tw, ispec, o, g.pw, g.ph, MulSpectrumHalfLoad{cached->data.as<__half2>() + cells * pi},
                                     h);
                    else
                        cols_inverse(s, tw, ispec, o, g.pw, g.ph, MulSpectrumLoad{cached->data.as<float2>() + cells * pi}, h);
                }
            } else {
                const KernelRef ka = ks[size_t(pl.spectral ? 2 * pi : 0)];
                const KernelRef kb = pl.spectral && size_t(2 * pi + 1) < ks.size() ? ks[size_t(2 * pi + 1)] : KernelRef{};
                fw.kspec_tmp.reset(sizeof(float2) * cells);
                float2* krow = fw.kspec_tmp.as<float2>();
                const int r = bank.support;
                KernelRows op{ka.data, kb.data, ka.scale, kb.scale, krow, bank.kernel, bank.kernel / 2, r, g.pw, g.ph};
                rows_pass(s, tw, g.pw, 2 * r + 1, op, false, r + 1, g.ph - (2 * r + 1));
                ColRows kz;  // only the rows [0, r] and [ph - r, ph) of the wrapped kernels are non-zero
                kz.zero_lo = r + 1;
                kz.zero_hi = g.ph - r;
                cols_convolve_channels(s, tw, krow, kz, fw.out_spec.as<float2>(), cells, nimg, g.pw, g.ph,
                                       MulSpectrumLoad{fw.image_spec.as<float2>(), cells}, h);
            }
            for (int j = 0; j < nimg; ++j) {
                float2* o = fw.out_spec.as<float2>() + cells * j;
                float* oa;
                float* ob;
                if (!pl.spectral) {
                    oa = planes + npx * (j == 0 ? 0 : 2);
                    ob = j == 0 ? planes + npx : nullptr;
                } else {
                    oa = planes + npx * (2 * pi);
                    ob = size_t(2 * pi + 1) < ks.size() ? planes + npx * (2 * pi + 1) : nullptr;
                }
                ExtractRows ex{o, oa, ob, w, h, g.pw, inv_cells};
                rows_pass(s, tw, g.pw, h, ex, true);
            }
        }
        // effect planes of this set
        float* eff = effect9 + npx * 3 * (pl.split ? 1 + int(si) : 0);
        if (!pl.spectral) {
            LDH_HIP(hipMemcpyAsync(eff, planes, sizeof(float) * npx * 3, hipMemcpyDeviceToDevice, s), "effect-copy");
        } else {
            ProfileScope ps_map("frame-spectral-map", (4.0 * n + 12.0) * double(npx));
            hipLaunchKernelGGL(spectral_map_kernel, grid1(npx), dim3(256), 0, s, planes, n, npx, fw.matrices.as<float>(),
                               fw.matrices.as<float>() + 3 * n, force, float(std::max(0.0, p.spectrum_saturation)),
                               p.chromatic_affects_luma, eff);
        }
        LDH_CHECK_LAUNCH("effect-map");
    }
    if (pl.split) {
        float* core = effect9 + npx * 3;
        float* structure = effect9 + npx * 6;
        hipLaunchKernelGGL(split_combine_kernel, grid1(npx), dim3(256), 0, s, core, structure, npx,
                           float(p.core_shoulder), float(p.structure_shoulder), float(std::max(0.0, p.core_gain)),
                           float(std::max(0.0, p.structure_gain)), core_or_structure_view(p) ? 1 : 0, effect9);
        LDH_CHECK_LAUNCH("split-combine");
    }
    return false;
}

// buildGaussianKernel (LensDiffCpuReference.cpp:1680-1700) on the host: unnormalised weights
// exp(-(x^2+y^2) / (2 sigma^2)) summed in raster order in float, then each divided by the sum.
std::vector<float> gaussian_kernel(float radius_px, int& size) {
    const int r = std::max(1, int(std::ceil(std::max(0.5f, radius_px))));
    size = 2 * r + 1;
    std::vector<float> k(size_t(size) * size);
    const float sigma = std::max(0.5f, radius_px * 0.5f);
    const float two_s2 = 2.0f * sigma * sigma;
    float sum = 0.0f;
    for (int y = -r; y <= r; ++y)
        for (int x = -r; x <= r; ++x) {
            const float v = std::exp(-(float(x * x + y * y) / two_s2));
            k[size_t(y + r) * size + size_t(x + r)] = v;
            sum += v;
        }
    if (sum > 0.0f)
        for (float& v : k) v /= sum;
    return k;
}

// The scatter image (convolveRgba of the effect with the Gaussian, :2183-2184): effect planes
// (R + iG) and (B) transformed on the smallest grid for the kernel's support, multiplied by the
// kernel's spectrum in the first inverse pass, cropped and clamped at zero into out3.
void scatter_pass(hipStream_t s, const FftTables& t, FrameWork& fw, float radius_px, const float* eff, int w, int h,
                  float* out3) {
    // the kernel is built on the host once per radius and kept on the device
    if (fw.gauss_radius != radius_px || fw.gauss_size == 0) {
        int K = 0;
        const std::vector<float> k = gaussian_kernel(radius_px, K);
        fw.gauss.reset(sizeof(float) * k.size());
        LDH_HIP(hipMemcpyAsync(fw.gauss.get(), k.data(), sizeof(float) * k.size(), hipMemcpyHostToDevice, s),
                "scatter-kernel");
        LDH_HIP(hipStreamSynchronize(s), "scatter-kernel");
        fw.gauss_radius = radius_px;
        fw.gauss_size = K;
    }
    const int K = fw.gauss_size;
    const int r = K / 2;
    const int64_t npx = int64_t(w) * h;
    const Grid g = make_grid(w, h, r);
    const size_t cells = g.cells();
    const float2* tw = t.twiddle.as<float2>();
    ProfileScope ps_("frame-scatter", 0.0);
    build_kernel_spectra(s, t, {KernelRef{fw.gauss.as<float>(), nullptr}}, K, r, g, false, fw.kspec, fw.kspec_tmp);
    fw.image_spec.reset(sizeof(float2) * cells * 2);
    fw.out_spec.reset(sizeof(float2) * cells);
    for (int j = 0; j < 2; ++j) {
        float2* spec = fw.image_spec.as<float2>() + cells * j;
        PlaneRows op{eff + npx * (j == 0 ? 0 : 2), j == 0 ? eff + npx : nullptr, spec, w, h, g.pw};
        rows_pass(s, tw, g.pw, h, op, false);
        ColRows zero;
        zero.zero_lo = h;
        zero.zero_hi = g.ph;
        cols_forward(s, tw, spec, g.pw, g.ph, PlainLoad{}, zero);
        float2* o = fw.out_spec.as<float2>();
        cols_inverse(s, tw, spec, o, g.pw, g.ph, MulSpectrumLoad{fw.kspec.data.as<float2>()}, h);
        ExtractRows ex{o, out3 + npx * (j == 0 ? 0 : 2), j == 0 ? out3 + npx : nullptr, w, h, g.pw,
                       1.0f / float(cells)};
        rows_pass(s, tw, g.pw, h, ex, true);
    }
}

// The frame a resolution-aware render was resampled from: its images are resampled back to it
// and composited there.
struct Display {
    const float4* src;
    int w, h, tf;
};

bool scatter_on(const ldh_params& p) { return p.scatter_amount > 1e-6 && scatter_radius_px(p) > 0.25; }
float fringe_px(const ldh_params& p) { return float(std::max(0.0, creative_fringe_px(p))); }

// The composite's parameters (LensDiffCpuReference.cpp:2096-2109, used by the final mix at
// :2203-2217): Preserve caps the gain at 1 and compensates by it; protect = max(0.2, corePreserve).
CompositeParams composite_params(const ldh_params& p) {
    const bool preserve = p.energy_mode == LDH_ENERGY_PRESERVE;
    const float redist_scale = 1.0f - float(std::max(0.0, std::min(p.core_preserve, 1.0)));
    const float gain = preserve ? float(std::max(0.0, std::min(p.effect_gain, 1.0))) : float(std::max(0.0, p.effect_gain));
    const float protect = std::max(0.2f, float(std::max(0.0, std::min(p.core_preserve, 1.0))));
    CompositeParams cp;
    cp.tf = p.input_transfer;
    cp.view = p.debug_view;
    cp.scale = redist_scale;
    cp.comp = preserve ? gain : float(std::max(0.0, p.core_compensation));
    cp.gain = gain;
    cp.max_sub = redist_scale > 1e-6f ? (1.0f - protect) / redist_scale : 0.0f;
    return cp;
}

// Steps 5-6 on the effect planes in fw.effect of the (working) frame src, w x h. With `disp`, the
// effect (or view) planes, the redistributed image and the mask are resampled to the display
// frame and the composite runs there (LensDiffCpuReference.cpp:2192-2220).
void finish_stage(hipStream_t s, const FftTables& t, FrameWork& fw, const ldh_params& p, const float4* src,
                  float4* dst, int w, int h, const Display* disp) {
    const int64_t npx = int64_t(w) * h;
    const Plan pl = plan_of(p);
    const float redist_scale = 1.0f - float(std::max(0.0, std::min(p.core_preserve, 1.0)));
    const float* effect = fw.effect.as<float>();
    if (pl.split) {
        if (p.debug_view == LDH_VIEW_CORE) effect = effect + npx * 3;
        else if (p.debug_view == LDH_VIEW_STRUCTURE) effect = effect + npx * 6;
    }

    // 5. Preserve rescale
    const bool preserve = p.energy_mode == LDH_ENERGY_PRESERVE;
    const float* factor = nullptr;
    if (preserve) {
        fw.lum.reset(sizeof(float) * npx * 2);
        fw.totals.reset(sizeof(float) * 2);
        fw.factors.reset(sizeof(float) * 2);
        hipLaunchKernelGGL(preserve_luma_kernel, grid1(npx), dim3(256), 0, s, src, fw.mask.as<float>(), p.input_transfer,
                           redist_scale, fw.effect.as<float>(), npx, fw.lum.as<float>());
        fw.seqsum.run(s, fw.lum.as<float>(), npx, npx, 2, fw.totals.as<float>());
        hipLaunchKernelGGL(preserve_factor_kernel, dim3(1), dim3(1), 0, s, fw.totals.as<float>(), fw.factors.as<float>());
        LDH_CHECK_LAUNCH("preserve");
        factor = fw.factors.as<float>();
    }

    // 5b. scatter and creative fringe on the rescaled effect (:2179-2190); they change only the
    // final image and the effect, scatter and fringe views
    const bool scatter = scatter_on(p);
    const float fringe = fringe_px(p);
    const bool post_view = p.debug_view == LDH_VIEW_SCATTER || p.debug_view == LDH_VIEW_CREATIVE_FRINGE;
    const bool post_out = p.debug_view == LDH_VIEW_FINAL || p.debug_view == LDH_VIEW_EFFECT || post_view;
    const float* post_effect = fw.effect.as<float>();  // the effect after the post passes
    const float* view_planes = nullptr;                // the scatter or fringe view
    if (post_out && (scatter || fringe > 1e-6f || post_view || disp)) {
        fw.post.reset(sizeof(float) * npx * 9);  // [scatter image][fringe preview][fringed effect]
        float* eff = fw.effect.as<float>();
        float* scat = fw.post.as<float>();
        float* preview = scat + npx * 3;
        float* fringed = scat + npx * 6;
        if (factor) {
            hipLaunchKernelGGL(scale_planes_kernel, grid1(npx * 3), dim3(256), 0, s, eff, npx * 3, factor);
            LDH_CHECK_LAUNCH("preserve-scale");
            factor = nullptr;
        }
        if (scatter) {
            scatter_pass(s, t, fw, float(scatter_radius_px(p)), eff, w, h, scat);
            hipLaunchKernelGGL(add_scaled_kernel, grid1(npx * 3), dim3(256), 0, s, eff, scat, npx * 3,
                               float(std::max(0.0, p.scatter_amount)));
            LDH_CHECK_LAUNCH("scatter-add");
        } else if (p.debug_view == LDH_VIEW_SCATTER) {
            hipLaunchKernelGGL(zero_kernel, dim3(unsigned(std::min<int64_t>(div_up(npx * 3, 256), 4096))), dim3(256), 0,
                               s, scat, npx * 3);
            LDH_CHECK_LAUNCH("scatter-clear");
        }
        if (fringe > 1e-6f) {
            ProfileScope ps_("frame-fringe", 40.0 * double(npx));
            hipLaunchKernelGGL(fringe_kernel, grid1(npx), dim3(256), 0, s, eff, fringed, preview, w, h, fringe);
            LDH_CHECK_LAUNCH("fringe");
            post_effect = fringed;
        } else if (p.debug_view == LDH_VIEW_CREATIVE_FRINGE) {
            hipLaunchKernelGGL(zero_kernel, dim3(unsigned(std::min<int64_t>(div_up(npx * 3, 256), 4096))), dim3(256), 0,
                               s, preview, npx * 3);
            LDH_CHECK_LAUNCH("fringe-clear");
        }
        if (p.debug_view == LDH_VIEW_SCATTER) view_planes = scat;
        if (p.debug_view == LDH_VIEW_CREATIVE_FRINGE) view_planes = preview;
    }

    // 6. composite
    const unsigned* mask_max = nullptr;
    if (p.debug_view == LDH_VIEW_SELECTION && !disp) {
        fw.max_buf.reset(sizeof(unsigned));
        LDH_HIP(hipMemsetAsync(fw.max_buf.get(), 0, sizeof(unsigned), s), "mask-max");
        hipLaunchKernelGGL(max_kernel, dim3(256), dim3(256), 0, s, fw.mask.as<float>(), npx, fw.max_buf.as<unsigned>());
        mask_max = fw.max_buf.as<unsigned>();
    }
    CompositeParams cp = composite_params(p);
    const float* eff_out = (p.debug_view == LDH_VIEW_FINAL || p.debug_view == LDH_VIEW_EFFECT || !pl.split)
                               ? post_effect : effect;
    if (p.debug_view == LDH_VIEW_CORE || p.debug_view == LDH_VIEW_STRUCTURE) {
        eff_out = effect;
        factor = nullptr;  // the reference rescales only the effect image
    }
    if (view_planes) {
        eff_out = view_planes;
        factor = nullptr;
    }
    if (disp) {
        // the display images at the frame's size; the Preserve factor is already in the effect
        const int64_t n = int64_t(disp->w) * disp->h;
        fw.display.reset(sizeof(float) * n * 7);  // [effect or view][redistributed][mask]
        float* d_eff = fw.display.as<float>();
        float* d_red = d_eff + n * 3;
        float* d_mask = d_eff + n * 6;
        {
            ProfileScope ps_("frame-display-resample", 4.0 * 7.0 * double(n));
            if (p.debug_view == LDH_VIEW_SELECTION) {
                hipLaunchKernelGGL(resample_planes_kernel, grid1(n), dim3(256), 0, s, fw.mask.as<float>(), w, h, 1,
                                   d_mask, disp->w, disp->h);
                fw.max_buf.reset(sizeof(unsigned));
                LDH_HIP(hipMemsetAsync(fw.max_buf.get(), 0, sizeof(unsigned), s), "mask-max");
                hipLaunchKernelGGL(max_kernel, dim3(256), dim3(256), 0, s, d_mask, n, fw.max_buf.as<unsigned>());
                mask_max = fw.max_buf.as<unsigned>();
            } else {
                hipLaunchKernelGGL(resample_planes_kernel, grid1(n), dim3(256), 0, s, eff_out, w, h, 3, d_eff, disp->w,
                                   disp->h);
            }
            if (p.debug_view == LDH_VIEW_FINAL) {
                fw.redist.reset(sizeof(float) * npx * 3);
                hipLaunchKernelGGL(redist_planes_kernel, grid1(npx), dim3(256), 0, s, src, fw.mask.as<float>(),
                                   redist_scale, npx, fw.redist.as<float>());
                hipLaunchKernelGGL(resample_planes_kernel, grid1(n), dim3(256), 0, s, fw.redist.as<float>(), w, h, 3,
                                   d_red, disp->w, disp->h);
            }
            LDH_CHECK_LAUNCH("display-resample");
        }
        cp.tf = disp->tf;
        ProfileScope ps_("frame-composite", 56.0 * double(n));
        hipLaunchKernelGGL(composite_kernel, grid1(n), dim3(256), 0, s, disp->src, d_mask, d_eff,
                           static_cast<const float*>(nullptr), mask_max, n, cp, dst, d_red);
        LDH_CHECK_LAUNCH("composite");
        return;
    }
    ProfileScope ps_("frame-composite", 56.0 * double(npx));
    hipLaunchKernelGGL(composite_kernel, grid1(npx), dim3(256), 0, s, src, fw.mask.as<float>(), eff_out, factor,
                       mask_max, npx, cp, dst, static_cast<const float*>(nullptr));
    LDH_CHECK_LAUNCH("composite");
}

}  // namespace

namespace {

// Steps 1-4 on the frame src (w x h): the selection mask, then the effect planes of the bank or,
// with zone banks, of every zone blended into fw.effect. With one bank, fin (nullable) is passed
// on to effect_stage; returns whether the final composite was folded in there.
bool effect_planes(hipStream_t s, const FftTables& t, FrameWork& fw, const ldh_params& p, const DeviceBank* bank,
                   const std::vector<const DeviceBank*>* zones, const float4* src, int w, int h, EffectMode mode,
                   const KernelSpectra* cached, const std::vector<const KernelSpectra*>* zone_cached,
                   const FusedFinal* fin = nullptr) {
    const int64_t npx = int64_t(w) * h;
    // the folded pre-combined path evaluates the mask wherever it reads a pixel
    if (!(fin && !zones && precombined_path(p, mode))) select_stage(s, fw, p, src, npx);
    fw.effect.reset(sizeof(float) * npx * 9);
    if (!zones) return effect_stage(s, t, fw, p, *bank, src, w, h, mode, cached, fw.effect.as<float>(), nullptr, false, fin);
    // the core and structure planes of a split look reach the output only in their views
    const int planes = plan_of(p).split && core_or_structure_view(p) ? 9 : 3;
    fw.zone.reset(sizeof(float) * npx * 9);
    hipLaunchKernelGGL(zero_kernel, dim3(unsigned(std::min<int64_t>(div_up(npx * planes, 256), 4096))), dim3(256), 0, s,
                       fw.effect.as<float>(), npx * planes);
    LDH_CHECK_LAUNCH("zone-clear");
    // one grid for every zone (the largest zone support): the frame's spectra are formed once
    const Grid gz = zone_grid(w, h, *zones);
    for (int z = 0; z < kFieldZones; ++z) {
        const KernelSpectra* zc = zone_cached != nullptr ? (*zone_cached)[size_t(z)] : nullptr;
        effect_stage(s, t, fw, p, *(*zones)[size_t(z)], src, w, h, mode, zc, fw.zone.as<float>(), &gz, z > 0);
        ProfileScope ps_("frame-zone-blend", 12.0 * planes * double(npx));
        hipLaunchKernelGGL(zone_accumulate_kernel, grid1(npx), dim3(256), 0, s, fw.zone.as<float>(),
                           fw.effect.as<float>(), planes, w, h, z % 3, z / 3);
        LDH_CHECK_LAUNCH("zone-accumulate");
    }
    return false;
}

// The whole frame; a resolution-aware render (working scale optics / frame short side away from
// 1, LensDiffCpuReference.cpp:2088-2093) runs steps 1-5 on the decoded frame resampled to the
// working size and composites its resampled images at the frame's size.
void render_impl(hipStream_t s, const FftTables& t, FrameWork& fw, const ldh_params& p, const DeviceBank* bank,
                 const std::vector<const DeviceBank*>* zones, const float4* src, float4* dst, int w, int h,
                 EffectMode mode, const KernelSpectra* cached,
                 const std::vector<const KernelSpectra*>* zone_cached = nullptr) {
    const double ws = working_scale(p);
    if (!resolution_aware_active(p)) {
        // a final image with nothing between the effect and the composite: the composite runs in
        // the effect's last pass where the effect takes the pre-combined path
        const bool direct = !zones && p.debug_view == LDH_VIEW_FINAL && p.energy_mode != LDH_ENERGY_PRESERVE &&
                            !scatter_on(p) && !(fringe_px(p) > 1e-6f);
        const FusedFinal fin{dst, composite_params(p)};
        if (effect_planes(s, t, fw, p, bank, zones, src, w, h, mode, cached, zone_cached, direct ? &fin : nullptr))
            return;
        finish_stage(s, t, fw, p, src, dst, w, h, nullptr);
        return;
    }
    const int ww = std::max(1, int(std::lround(w * ws))), wh = std::max(1, int(std::lround(h * ws)));
    const int64_t nw = int64_t(ww) * wh;
    fw.work.reset(sizeof(float4) * nw);
    hipLaunchKernelGGL(resample_frame_kernel, grid1(nw), dim3(256), 0, s, src, w, h, p.input_transfer,
                       fw.work.as<float4>(), ww, wh);
    LDH_CHECK_LAUNCH("working-resample");
    ldh_params lin = p;  // the working frame is already decoded
    lin.input_transfer = LDH_TRANSFER_LINEAR;
    effect_planes(s, t, fw, lin, bank, zones, fw.work.as<float4>(), ww, wh, mode, cached, zone_cached);
    const Display d{src, w, h, p.input_transfer};
    finish_stage(s, t, fw, lin, fw.work.as<float4>(), dst, ww, wh, &d);
}

}  // namespace

void render_frame(hipStream_t s, const FftTables& t, FrameWork& fw, const ldh_params& p, const DeviceBank& bank,
                  const float4* src, float4* dst, int w, int h, EffectMode mode, const KernelSpectra* cached) {
    render_impl(s, t, fw, p, &bank, nullptr, src, dst, w, h, mode, cached);
}

void render_frame_zones(hipStream_t s, const FftTables& t, FrameWork& fw, const ldh_params& p,
                        const std::vector<const DeviceBank*>& banks, const float4* src, float4* dst, int w, int h,
                        EffectMode mode, const std::vector<const KernelSpectra*>* cached) {
    if (banks.size() != size_t(kFieldZones)) fail(LDH_ERR_INVALID_ARGUMENT, "field-zone-bank-count");
    if (cached != nullptr && cached->size() != size_t(kFieldZones)) fail(LDH_ERR_INVALID_ARGUMENT, "field-zone-spectra-count");
    render_impl(s, t, fw, p, nullptr, &banks, src, dst, w, h, mode, nullptr, cached);
}

Grid zone_grid(int w, int h, const std::vector<const DeviceBank*>& banks) {
    int support = 0;
    for (const DeviceBank* b : banks) support = std::max(support, b->support);
    return make_grid(w, h, support);
}

bool is_static_view(int view) {
    return view == LDH_VIEW_PUPIL || view == LDH_VIEW_PSF || view == LDH_VIEW_OTF || view == LDH_VIEW_PHASE ||
           view == LDH_VIEW_PHASE_EDGE || view == LDH_VIEW_FIELD_PSF || view == LDH_VIEW_CHROMATIC_SPLIT;
}

void render_static_view(hipStream_t s, FrameWork& fw, const ldh_params& p, int view, const DeviceBank& bank,
                        float4* dst, int w, int h) {
    const int64_t npx = int64_t(w) * h;
    fw.max_buf.reset(sizeof(unsigned) * 3);
    unsigned* mx = fw.max_buf.as<unsigned>();
    LDH_HIP(hipMemsetAsync(mx, 0, sizeof(unsigned) * 3, s), "view-max");
    ProfileScope ps_("frame-static-view", 16.0 * double(npx));
    if (view == LDH_VIEW_PUPIL || view == LDH_VIEW_PHASE || view == LDH_VIEW_PHASE_EDGE) {
        // the pupil amplitude and phase waves at the bank's pupil size (cache.pupilDisplay /
        // phaseDisplay, :1627-1630)
        const PupilGeom geom = make_pupil_geom(p);
        const int n = geom.size;
        fw.view_pupil.reset(sizeof(float) * size_t(n) * n * 2);
        float* amp = fw.view_pupil.as<float>();
        float* ph = amp + size_t(n) * n;
        pupil_amplitude(s, geom, amp, false);
        SquareSrc a{amp, n, n, nullptr, mx};
        hipLaunchKernelGGL(max_kernel, dim3(64), dim3(256), 0, s, amp, int64_t(n) * n, mx);
        if (view == LDH_VIEW_PUPIL) {
            hipLaunchKernelGGL(square_view_kernel, grid1(npx), dim3(256), 0, s, a, a, a, 0, dst, w, h);
        } else {
            const PhaseCoef c = make_phase_coef(p);
            fw.view_coef.reset(sizeof(c));
            LDH_HIP(hipMemcpyAsync(fw.view_coef.get(), &c, sizeof(c), hipMemcpyHostToDevice, s), "view-phase");
            pupil_phase(s, fw.view_coef.as<PhaseCoef>(), n, 1, ph, false);
            hipLaunchKernelGGL(max_abs_kernel, dim3(64), dim3(256), 0, s, ph, int64_t(n) * n, mx + 1);
            SquareSrc q{ph, n, n, nullptr, mx + 1};
            hipLaunchKernelGGL(phase_view_kernel, grid1(npx), dim3(256), 0, s, q, a, dst, w, h);
        }
        LDH_CHECK_LAUNCH("static-view");
        return;
    }
    // kernels of the bank: the first bin's full kernel, or last / middle / first for the split
    auto square = [&](int bin, unsigned* bits) {
        const int K = bank.kernel, c = K / 2;
        const int rc = std::min(std::max(bank.radius[size_t(bin)], 1), c);
        const float* k = bank.kernel_ptr(kFull, bin);
        hipLaunchKernelGGL(max_kernel, dim3(256), dim3(256), 0, s, k, int64_t(K) * K, bits);
        return SquareSrc{k + int64_t(c - rc) * K + (c - rc), 2 * rc + 1, K, bank.scale_ptr(kFull, bin), bits};
    };
    if (view == LDH_VIEW_OTF) {
        const SquareSrc k = square(0, mx);
        const int n = k.n;
        fw.view_pupil.reset(sizeof(float2) * size_t(n) * n + sizeof(float) * size_t(n) * n);
        float2* rows = fw.view_pupil.as<float2>();
        float* mag = reinterpret_cast<float*>(rows + size_t(n) * n);
        const unsigned nb = unsigned(div_up(n, 256));
        if (n > 4097) fail(LDH_ERR_TOO_LARGE, "otf-view-kernel-too-large");
        // exp(-2 i pi j / n) in double, rounded once; eps = 2 (pi_f - pi) / n
        std::vector<float2> tw(static_cast<size_t>(n));
        for (int j = 0; j < n; ++j) {
            const double a = -2.0 * M_PI * double(j) / double(n);
            tw[size_t(j)] = make_float2(float(std::cos(a)), float(std::sin(a)));
        }
        const float eps = float(2.0 * (double(float(M_PI)) - M_PI) / double(n));
        fw.view_coef.reset(sizeof(float2) * size_t(n));
        LDH_HIP(hipMemcpyAsync(fw.view_coef.get(), tw.data(), sizeof(float2) * size_t(n), hipMemcpyHostToDevice, s),
                "view-otf");
        hipLaunchKernelGGL(otf_rows_kernel, dim3(nb, unsigned(n)), dim3(256), sizeof(float) * n, s, k.data, k.pitch, n,
                           k.scale, fw.view_coef.as<float2>(), eps, rows);
        hipLaunchKernelGGL(otf_cols_kernel, dim3(nb, unsigned(n)), dim3(256), sizeof(float2) * n, s, rows, n,
                           fw.view_coef.as<float2>(), eps, mag);
        LDH_HIP(hipStreamSynchronize(s), "view-otf");  // the host table goes out of scope
        hipLaunchKernelGGL(max_kernel, dim3(256), dim3(256), 0, s, mag, int64_t(n) * n, mx + 1);
        const SquareSrc m{mag, n, n, nullptr, mx + 1};
        hipLaunchKernelGGL(square_view_kernel, grid1(npx), dim3(256), 0, s, m, m, m, 1, dst, w, h);
        LDH_CHECK_LAUNCH("static-view");
        return;
    }
    if (view == LDH_VIEW_CHROMATIC_SPLIT) {
        const SquareSrc r = square(bank.bins - 1, mx), g = square(bank.bins / 2, mx + 1), b = square(0, mx + 2);
        hipLaunchKernelGGL(square_view_kernel, grid1(npx), dim3(256), 0, s, r, g, b, 1, dst, w, h);
    } else {
        const SquareSrc k = square(0, mx);
        hipLaunchKernelGGL(square_view_kernel, grid1(npx), dim3(256), 0, s, k, k, k, 1, dst, w, h);
    }
    LDH_CHECK_LAUNCH("static-view");
}

}  // namespace ldh

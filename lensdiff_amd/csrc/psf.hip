// PSF synthesis for a batch of wavelengths (buildShiftedRawPsf, LensDiffCpuReference.cpp:831-859,
// driven per wavelength by buildPsfBankGlobalOnly :1702-1734).
//
// The R x R transform only has a P x P non-zero block (the pupil, embedded at offset
// floor((R-P)/2)), so it is computed as two separable line stages, columns first:
//   pass 1: for each pupil column x, the R-point DFT of its P embedded values -> H[x][ky]
//   pass 2: for each ky, the R-point DFT over the P columns of H             -> F[ky][kx]
// The pupil amplitude and phase are generated transposed ([x][y]) and turned into the complex
// field (times the chirp) by one elementwise kernel, so pass 1 reads contiguous lines; pass 2's
// line is one row of F, whose |F|^2 lands straight in its row of the centred raw PSF.
// Each R-point DFT is done in LDS. For power-of-two R it is a plain Stockham FFT of length R.
// Otherwise it is Bluestein's algorithm with the reference's chirp exp(-i pi j^2 / R) evaluated
// in double with pi rounded to float (LensDiffCpuReference.cpp:244-274), which makes the
// transform sum_n x[n] exp(-2 i pi_f k n / R): the reference's sampled spectrum. Moving the
// pupil from offset floor((R-P)/2) to 0 multiplies each row transform by exp(-2 i pi_f kx off / R)
// and each column transform by exp(-2 i pi_f ky off / R): unit-modulus factors that |F|^2
// drops. With the input at 0 the circular convolution only has to hold lags -(P-1) .. R-1, so
// the Bluestein length is the smallest line length M >= R + P - 1 (5120 for R = 4092, P = 1024,
// where the reference uses 8192).
// The sequential-sum normalisation then scales each raw PSF to unit energy.
#include <cmath>

#include "autotune.h"
#include "fft_line.h"
#include "profiler.h"
#include "psf.h"
#include "pupil.h"

namespace ldh {

namespace {

// Stage-0 input of the row transform of pupil row y: A * (cos 2 pi phi, sin 2 pi phi) where
// A > 0, at line position x + off, times the chirp for Bluestein.
// The complex pupil field of row-line y at position x: A * (cos 2 pi phi, sin 2 pi phi) where
// A > 0, times the chirp for Bluestein (one elementwise kernel for all wavelengths).
__global__ void pupil_field_kernel(const float* __restrict__ amp, const int* __restrict__ amp_index,
                                   const float* __restrict__ phase, int P, const float2* __restrict__ chirp,
                                   int bluestein, float2* __restrict__ field) {
    const int lam = blockIdx.y;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t pp = int64_t(P) * P;
    if (i >= pp) return;
    const int64_t a0 = amp_index != nullptr ? int64_t(amp_index[lam]) * pp : 0;
    const int x = int(i % P);
    const float a = amp[a0 + i];
    float rad = 0.0f;
    if (a > 0.0f) rad = phase[int64_t(lam) * pp + i] * 2.0f * kPiF;
    float2 v = make_float2(a * cosf(rad), a * sinf(rad));
    if (bluestein) v = cmul_fixed(v, chirp[x]);
    field[int64_t(lam) * pp + i] = v;
}

// The complex fields of the wavelengths of one or more pupils (workgroup row k: pupil k, its
// amplitude amp + k * P * P and units unit0[k] .. unit0[k + 1] - 1; unit0 null: one pupil of
// `units` units), where a pupil's phase coefficient sets differ only in defocus (the chromatic
// focus offset): each texel's other phase terms are formed once (phase_texel), and each unit's
// phase, amplitude * (cos, sin)(2 pi phase) and chirp follow as in pupil_phase_kernel and
// pupil_field_kernel. Transposed layout ([x][y]) as there.
__global__ void pupil_field_defocus_kernel(const float* __restrict__ amp, const PhaseCoef* __restrict__ coefs,
                                           const int* __restrict__ unit0, int units, int P,
                                           const float2* __restrict__ chirp, int bluestein, float2* __restrict__ field) {
    const int k = blockIdx.y;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t pp = int64_t(P) * P;
    if (i >= pp) return;
    const int u0 = unit0 ? unit0[k] : 0, u1 = unit0 ? unit0[k + 1] : units;
    const int x = int(i / P), y = int(i % P);
    const float a = amp[int64_t(k) * pp + i];
    const float2 c = bluestein ? chirp[y] : make_float2(1.0f, 0.0f);
    const PhaseTexel t = phase_texel(coefs[u0], x, y, P);
    for (int u = u0; u < u1; ++u) {
        float rad = 0.0f;
        if (a > 0.0f) rad = t.phase(coefs[u]) * 2.0f * kPiF;
        float2 v = make_float2(a * cosf(rad), a * sinf(rad));
        if (bluestein) v = cmul_fixed(v, c);  // as in pupil_field_kernel, bit for bit
        field[int64_t(u) * pp + i] = v;
    }
}

// Stage-0 input of the row transform of row-line y: its P field values at positions j - off.
// Branch-free: every lane loads (out-of-range positions read element 0) so the stage's loads
// issue back to back.
struct RowInput {
    const float2* __restrict__ row;
    int P, off;
    __device__ float2 operator()(int j) const {
        const int x = j - off;
        const bool in = x >= 0 && x < P;
        const float2 v = row[in ? x : 0];
        return in ? v : make_float2(0.0f, 0.0f);
    }
};

// Stage-0 input of the column transform of frequency row ky: the P pass-1 values, which already
// carry the chirp of their position (the row pass applies it). Branch-free: every lane loads
// (out-of-range positions read element 0) so the stage's loads issue back to back.
struct ColInput {
    const float2* __restrict__ col;
    int P, off;
    __device__ float2 operator()(int j) const {
        const int y = j - off;
        const bool in = y >= 0 && y < P;
        const float2 v = col[in ? y : 0];
        return in ? v : make_float2(0.0f, 0.0f);
    }
};

// Output k < R of a row line times `post`: the column transform's chirp at this line's position
// times 1/M (Bluestein), or 1. Bluestein's output chirp c_k is left out: every value the column
// transform of frequency row k reads carries the same c_k, so it only multiplies that row of F by
// a unit-modulus factor, which |F|^2 drops.
struct RowOutput {
    static constexpr bool kPrefetch = false;
    float2* __restrict__ out;
    int R;
    float2 post;
    int bluestein;
    __device__ void operator()(int k, float2 v) const {
        if (k < R) out[k] = bluestein ? cmul(v, post) : v;
    }
};
// |F|^2 of output k into the centred row: column (k - R/2) mod R (fftShiftSquare, :303-314),
// also added in FP64 to the thread's share of the row sum. Bluestein's output chirp has unit
// modulus, so |F|^2 is |v|^2 / M^2 without it (scale2 = 1 / M^2, or 1 for a plain FFT).
struct ColOutput {
    static constexpr bool kPrefetch = false;
    float* __restrict__ row;
    float scale2;
    int R, shift;  // R - R/2
    double* acc;
    __device__ void operator()(int k, float2 v) const {
        if (k < R) {
            int x = k + shift;
            if (x >= R) x -= R;
            const float p = fmaf(v.x, v.x, v.y * v.y) * scale2;
            row[x] = p;
            *acc += double(p);
        }
    }
};

// The R-point transform of one line: a length-M FFT for power-of-two R, else Bluestein's
// forward FFT, filter multiply (a pass over the line in LDS) and inverse FFT.
template <int M, int B, class In, class Out>
__device__ __forceinline__ void line_transform(float2* lds, const float2* __restrict__ tw,
                                               const float2* __restrict__ bhat, bool bluestein, int tid,
                                               const In& in, const Out& out, int extent) {
    const GlobalIn<In> gin{in};
    const GlobalOut<Out> gout{out};
    const StageTwiddles<M, B> st = stage_twiddles<M, B>(tw, tid);  // shared by both transforms
    if (bluestein) {
        line_fft<M, B, false>(lds, tw, tid, gin, LdsOut{lds}, st, extent);
        line_multiply<M, LinePlan<M, B>::NT>(lds, bhat, tid);
        line_fft<M, B, true>(lds, tw, tid, LdsIn{lds}, gout, st);
    } else {
        line_fft<M, B, false>(lds, tw, tid, gin, gout, st, extent);
    }
}

// B: radix of the stages after the first (16 or 8); W: minimum waves per SIMD the register
// allocation must allow (1: no constraint). The autotuner picks (B, W) for long lines.
template <int M, int B, int W>
__global__ void __launch_bounds__((LinePlan<M, B>::NT), W)
psf_rows_kernel(const float2* __restrict__ field, int P, int nout, int off, const float2* __restrict__ chirp,
                const float2* __restrict__ bhat, int bluestein, float2* __restrict__ G, const float2* __restrict__ tw) {
    extern __shared__ float2 lds[];
    const int y = blockIdx.x, lam = blockIdx.y, tid = threadIdx.x;
    const RowInput in{field + (size_t(lam) * P + y) * P, P, off};
    const float inv_m = 1.0f / float(M);
    const float2 post = bluestein ? make_float2(chirp[y].x * inv_m, chirp[y].y * inv_m) : make_float2(1.0f, 0.0f);
    const RowOutput out{G + (size_t(lam) * P + y) * nout, nout, post, bluestein};
    line_transform<M, B>(lds, tw, bhat, bluestein != 0, tid, in, out, off + P);
}

// nout output rows of nout values each, line ky landing in row (ky + shift) mod nout with its
// values shifted alike (the full form's fftShiftSquare: shift = R - R/2; a window: 0).
// Also writes the FP64 sum of each output row to row_sums[lam * nout + y] (nullable): the block
// sums of a sequential sum whose blocks are the rows.
template <int M, int B, int W>
__global__ void __launch_bounds__((LinePlan<M, B>::NT), W)
psf_cols_kernel(const float2* __restrict__ GT, int P, int nout, int shift, int off, const float2* __restrict__ chirp,
                const float2* __restrict__ bhat, int bluestein, float* __restrict__ IT,
                const float2* __restrict__ tw, double* __restrict__ row_sums) {
    extern __shared__ float2 lds[];
    constexpr int kWaves = LinePlan<M, B>::NT / 64;
    __shared__ double part[kWaves];
    const int ky = blockIdx.x, lam = blockIdx.y, tid = threadIdx.x;
    int y = ky + shift;
    if (y >= nout) y -= nout;
    double acc = 0.0;
    const ColInput in{GT + (size_t(lam) * nout + ky) * P, P, off};
    const float inv_m = 1.0f / float(M);
    const ColOutput out{IT + (size_t(lam) * nout + y) * nout, bluestein ? inv_m * inv_m : 1.0f, nout, shift, &acc};
    line_transform<M, B>(lds, tw, bhat, bluestein != 0, tid, in, out, off + P);
    if (row_sums == nullptr) return;
    acc = wave_total(acc);
    if ((tid & 63) == 0) part[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < kWaves; ++i) t += part[i];
        row_sums[size_t(lam) * nout + y] = t;
    }
}

// Batched transpose of complex [rows][cols] -> [cols][rows] through a 32x33 LDS tile.
__global__ void __launch_bounds__(256)
transpose_c64_kernel(const float2* __restrict__ in, float2* __restrict__ out, int rows, int cols) {
    __shared__ float2 tile[32][33];
    const int lam = blockIdx.z;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const float2* src = in + size_t(lam) * rows * cols;
    float2* dst = out + size_t(lam) * rows * cols;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = src[size_t(r) * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) dst[size_t(c) * rows + r] = tile[tx][k];
    }
}

// Pixel (x, y) of output element i: row-major out[y][x], or transposed out[x][y].
__device__ __forceinline__ void pixel_of(int i, int n, int transposed, int* x, int* y) {
    if (transposed) { *x = i / n; *y = i % n; }
    else { *x = i % n; *y = i / n; }
}

__global__ void pupil_amplitude_kernel(PupilGeom g, int transposed, float* __restrict__ out) {
    const int n = g.size;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    int x, y;
    pixel_of(i, n, transposed, &x, &y);
    out[i] = pupil_amplitude_at(g, x, y);
}

__global__ void pupil_amplitude_batch_kernel(const PupilGeom* __restrict__ geoms, int n, int transposed,
                                             float* __restrict__ out) {
    const int k = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    const PupilGeom g = geoms[k];
    int x, y;
    pixel_of(i, n, transposed, &x, &y);
    out[size_t(k) * n * n + i] = pupil_amplitude_at(g, x, y);
}

__global__ void pupil_phase_kernel(const PhaseCoef* __restrict__ coefs, int n, int transposed, float* __restrict__ out) {
    const int lam = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    const PhaseCoef c = coefs[lam];
    int x, y;
    pixel_of(i, n, transposed, &x, &y);
    out[size_t(lam) * n * n + i] = phase_waves_at(c, x, y, n);
}

// Double-precision transforms on the host (Bluestein filter tables only): radix-2 for the
// power-of-two part, then direct DFTs over the factor 3 or 9 (M = 2^a * 3^b).
void host_fft_pow2(std::vector<double>& re, std::vector<double>& im) {
    const int n = int(re.size());
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { std::swap(re[i], re[j]); std::swap(im[i], im[j]); }
    }
    for (int len = 2; len <= n; len <<= 1) {
        for (int k = 0; k < len / 2; ++k) {
            const double a = -2.0 * M_PI * double(k) / double(len);
            const double wr = std::cos(a), wi = std::sin(a);
            for (int b = 0; b < n; b += len) {
                const int i0 = b + k, i1 = b + k + len / 2;
                const double xr = re[i1] * wr - im[i1] * wi, xi = re[i1] * wi + im[i1] * wr;
                re[i1] = re[i0] - xr; im[i1] = im[i0] - xi;
                re[i0] += xr; im[i0] += xi;
            }
        }
    }
}

// Forward DFT of length n = 2^a * q (q = 1, 3, 5 or 9) in double: q interleaved radix-2 FFTs of
// length 2^a, then twiddles and length-q DFTs.
void host_dft(std::vector<double>& re, std::vector<double>& im) {
    const int n = int(re.size());
    int q = n;
    while ((q & 1) == 0) q >>= 1;
    const int m = n / q;
    const std::vector<double> line(static_cast<size_t>(m), 0.0);
    std::vector<std::vector<double>> sr(static_cast<size_t>(q), line), si(static_cast<size_t>(q), line);
    for (int r = 0; r < q; ++r)
        for (int j = 0; j < m; ++j) { sr[size_t(r)][size_t(j)] = re[size_t(j * q + r)]; si[size_t(r)][size_t(j)] = im[size_t(j * q + r)]; }
    for (int r = 0; r < q; ++r) host_fft_pow2(sr[size_t(r)], si[size_t(r)]);
    // X[k1 + m k2] = sum_r W_n^(r (k1 + m k2)) S_r[k1]
    for (int k = 0; k < n; ++k) {
        double xr = 0.0, xi = 0.0;
        const int k1 = k % m;
        for (int r = 0; r < q; ++r) {
            const double a = -2.0 * M_PI * double((int64_t(r) * k) % n) / double(n);
            const double c = std::cos(a), sn = std::sin(a);
            const double vr = sr[size_t(r)][size_t(k1)], vi = si[size_t(r)][size_t(k1)];
            xr += vr * c - vi * sn;
            xi += vr * sn + vi * c;
        }
        re[size_t(k)] = xr;
        im[size_t(k)] = xi;
    }
}

// Variants of the line kernels: (B, W) = (16, 1), (16, 3), (16, 4), (8, 8), (8, 3).
// Tuning keys carry the line length and the batch (wavelength count rounded up to a power of
// two, 32 and more alike): a one-wavelength launch does not fill the device the way a
// 31-wavelength bank does.
constexpr int kPsfVariants = 5;
// Long lines are register-bound: they get every variant.
template <int M> constexpr int psf_variants() { return M >= 2048 ? kPsfVariants : 1; }

template <int M, int B, int W>
void launch_rows_v(hipStream_t s, const float2* field, int P, int nout, int off, const float2* chirp,
                   const float2* bhat, int bluestein, float2* G, const float2* tw, int L) {
    allow_large_lds(reinterpret_cast<const void*>(&psf_rows_kernel<M, B, W>), (M + M / 32 + 1) * 8);
    hipLaunchKernelGGL((psf_rows_kernel<M, B, W>), dim3(P, L), dim3(LinePlan<M, B>::NT), (M + M / 32 + 1) * 8, s, field,
                       P, nout, off, chirp, bhat, bluestein, G, tw);
}
template <int M, int B, int W>
void launch_cols_v(hipStream_t s, const float2* GT, int P, int nout, int shift, int off, const float2* chirp,
                   const float2* bhat, int bluestein, float* IT, const float2* tw, double* row_sums, int L) {
    allow_large_lds(reinterpret_cast<const void*>(&psf_cols_kernel<M, B, W>), (M + M / 32 + 1) * 8);
    hipLaunchKernelGGL((psf_cols_kernel<M, B, W>), dim3(nout, L), dim3(LinePlan<M, B>::NT), (M + M / 32 + 1) * 8, s,
                       GT, P, nout, shift, off, chirp, bhat, bluestein, IT, tw, row_sums);
}

std::string tune_key(const char* kind, int M, int L) {
    return std::string(kind) + "/" + std::to_string(M) + "/L" + std::to_string(std::min(32, pow2_at_least(L)));
}

template <int M>
bool launch_rows(int m, hipStream_t s, const float2* field, int P, int R, int off, const float2* chirp,
                 const float2* bhat, int bluestein, float2* G, const float2* tw, int L) {
    if (m != M) return false;
    Autotuner::instance().run(tune_key("psf-rows", M, L), psf_variants<M>(), s, [&](int v) {
        if constexpr (psf_variants<M>() > 1) {
            if (v == 1) return launch_rows_v<M, 16, 3>(s, field, P, R, off, chirp, bhat, bluestein, G, tw, L);
            if (v == 2) return launch_rows_v<M, 16, 4>(s, field, P, R, off, chirp, bhat, bluestein, G, tw, L);
            if (v == 3) return launch_rows_v<M, 8, 8>(s, field, P, R, off, chirp, bhat, bluestein, G, tw, L);
            if (v == 4) return launch_rows_v<M, 8, 3>(s, field, P, R, off, chirp, bhat, bluestein, G, tw, L);
        }
        launch_rows_v<M, 16, 1>(s, field, P, R, off, chirp, bhat, bluestein, G, tw, L);
    });
    return true;
}
template <int M>
bool launch_cols(int m, hipStream_t s, const float2* GT, int P, int nout, int shift, int off, const float2* chirp,
                 const float2* bhat, int bluestein, float* IT, const float2* tw, double* row_sums, int L) {
    if (m != M) return false;
    Autotuner::instance().run(tune_key("psf-cols", M, L), psf_variants<M>(), s, [&](int v) {
        if constexpr (psf_variants<M>() > 1) {
            if (v == 1)
                return launch_cols_v<M, 16, 3>(s, GT, P, nout, shift, off, chirp, bhat, bluestein, IT, tw, row_sums, L);
            if (v == 2)
                return launch_cols_v<M, 16, 4>(s, GT, P, nout, shift, off, chirp, bhat, bluestein, IT, tw, row_sums, L);
            if (v == 3)
                return launch_cols_v<M, 8, 8>(s, GT, P, nout, shift, off, chirp, bhat, bluestein, IT, tw, row_sums, L);
            if (v == 4)
                return launch_cols_v<M, 8, 3>(s, GT, P, nout, shift, off, chirp, bhat, bluestein, IT, tw, row_sums, L);
        }
        launch_cols_v<M, 16, 1>(s, GT, P, nout, shift, off, chirp, bhat, bluestein, IT, tw, row_sums, L);
    });
    return true;
}

// Line lengths of the PSF kernels: powers of two (plain transforms of R, and Bluestein) and
// 3 * 2^a, 5 * 2^a, 9 * 2^a Bluestein lengths with a composite-first plan.
#define LDH_PSF_LENGTHS(X)                                                                                  \
    X(128) || X(144) || X(160) || X(192) || X(256) || X(288) || X(320) || X(384) || X(512) || X(576) ||     \
        X(640) || X(768) || X(1024) || X(1280) || X(1536) || X(2048) || X(2304) || X(2560) || X(3072) ||    \
        X(4096) || X(4608) || X(5120) || X(6144) || X(8192) || X(9216) || X(12288) || X(16384)
constexpr int kPsfLengths[] = {128,  144,  160,  192,  256,  288,  320,  384,  512,  576,  640,   768,  1024, 1280,
                               1536, 2048, 2304, 2560, 3072, 4096, 4608, 5120, 6144, 8192, 9216, 12288, 16384};

int psf_length_at_least(int v) {
    for (int n : kPsfLengths)
        if (n >= v) return n;
    return 0;
}

}  // namespace

void PsfPlan::prepare(hipStream_t s, int pupil, int raw) {
    if (pupil == pupil_ && raw == raw_) return;
    if (raw > kMaxRawPsf) fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
    pupil_ = pupil;
    raw_ = raw;
    h_ = -1;  // a window belongs to its (P, R)
    mw_ = 0;
    bluestein_ = !is_pow2(raw);
    // Bluestein: input at 0 (see above), lags -(P-1) .. R-1 must not wrap
    m_ = bluestein_ ? psf_length_at_least(raw + std::min(pupil, raw) - 1) : raw;
    if (m_ == 0) fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
    offset_ = bluestein_ ? 0 : std::max(0, (raw - pupil) / 2);
    if (!bluestein_) return;
    // chirp[j] = exp(-i * pi_f * j^2 / R), the reference's forward chirp (sign -1), with the
    // phase formed in double from pi rounded to float.
    const double pi_f = double(kPiF);
    const int lags_neg = std::min(pupil, raw);  // b[M - j] for j in [1, P)
    std::vector<float2> chirp(static_cast<size_t>(raw));
    std::vector<double> br(static_cast<size_t>(m_), 0.0), bi(static_cast<size_t>(m_), 0.0);
    for (int j = 0; j < raw; ++j) {
        const double ph = -1.0 * pi_f * double(j) * double(j) / double(raw);
        chirp[size_t(j)] = make_float2(float(std::cos(ph)), float(std::sin(ph)));
        const float cr = float(std::cos(-ph)), ci = float(std::sin(-ph));
        br[size_t(j)] = cr; bi[size_t(j)] = ci;
        if (j != 0 && j < lags_neg) { br[size_t(m_ - j)] = cr; bi[size_t(m_ - j)] = ci; }
    }
    host_dft(br, bi);
    std::vector<float2> bhat(static_cast<size_t>(m_));
    for (int k = 0; k < m_; ++k) bhat[size_t(k)] = make_float2(float(br[size_t(k)]), float(bi[size_t(k)]));
    chirp_.reset(sizeof(float2) * raw);
    bhat_.reset(sizeof(float2) * m_);
    LDH_HIP(hipMemcpyAsync(chirp_.get(), chirp.data(), sizeof(float2) * raw, hipMemcpyHostToDevice, s), "psf-chirp");
    LDH_HIP(hipMemcpyAsync(bhat_.get(), bhat.data(), sizeof(float2) * m_, hipMemcpyHostToDevice, s), "psf-chirp");
    LDH_HIP(hipStreamSynchronize(s), "psf-chirp");
}

void pupil_amplitude(hipStream_t s, const PupilGeom& g, float* out, bool transposed) {
    const int n = g.size;
    hipLaunchKernelGGL(pupil_amplitude_kernel, dim3(div_up(int64_t(n) * n, 256)), dim3(256), 0, s, g,
                       transposed ? 1 : 0, out);
    LDH_CHECK_LAUNCH("pupil-amplitude");
}

void pupil_amplitudes(hipStream_t s, const PupilGeom* d_geoms, int n, int count, float* out, bool transposed) {
    hipLaunchKernelGGL(pupil_amplitude_batch_kernel, dim3(div_up(int64_t(n) * n, 256), count), dim3(256), 0, s, d_geoms,
                       n, transposed ? 1 : 0, out);
    LDH_CHECK_LAUNCH("pupil-amplitude");
}

void pupil_phase(hipStream_t s, const PhaseCoef* d_coefs, int n, int count, float* out, bool transposed) {
    hipLaunchKernelGGL(pupil_phase_kernel, dim3(div_up(int64_t(n) * n, 256), count), dim3(256), 0, s, d_coefs, n,
                       transposed ? 1 : 0, out);
    LDH_CHECK_LAUNCH("pupil-phase");
}

void PsfPlan::fields(hipStream_t s, const float* amp, const float* phase, int count, const float2* chirp,
                     int bluestein, PsfScratch& scratch, const int* amp_index, const PhaseCoef* defocus_coefs,
                     const int* unit0, int pupils) {
    const int P = pupil_, L = count;
    // amplitude (shared by the wavelengths of a pupil) and phase in, complex field out
    if (defocus_coefs != nullptr) {
        ProfileScope ps("psf-field", (4.0 * pupils + 8.0 * L) * P * P);
        hipLaunchKernelGGL(pupil_field_defocus_kernel, dim3(unsigned(div_up(int64_t(P) * P, 256)), unsigned(pupils)),
                           dim3(256), 0, s, amp, defocus_coefs, unit0, L, P, chirp, bluestein,
                           scratch.field.as<float2>());
    } else {
        ProfileScope ps("psf-field", 12.0 * P * P * L);
        hipLaunchKernelGGL(pupil_field_kernel, dim3(unsigned(div_up(int64_t(P) * P, 256)), unsigned(L)), dim3(256), 0,
                           s, amp, amp_index, phase, P, chirp, bluestein, scratch.field.as<float2>());
    }
    LDH_CHECK_LAUNCH("psf-field");
}

void PsfPlan::raw_psfs(hipStream_t s, const FftTables& t, const float* amp, const float* phase, int count,
                       float* out, float* totals, SeqSum& seqsum, PsfScratch& scratch, bool normalize,
                       const int* amp_index, const PhaseCoef* defocus_coefs, const int* unit0, int pupils) {
    const int P = pupil_, R = raw_, L = count;
    scratch.reserve(P, R, L);
    DeviceBuffer& g_ = scratch.g;
    DeviceBuffer& gt_ = scratch.gt;
    const float2* tw = t.twiddle.as<float2>();
    const float2* chirp = chirp_.as<float2>();
    const float2* bhat = bhat_.as<float2>();
    const int blu = bluestein_ ? 1 : 0;
    // rows of at least kSeqSumMinBlock elements serve as the blocks of the sequential sum, with
    // their sums formed by the column pass
    const bool row_blocks = R >= kSeqSumMinBlock && R <= kSeqSumBlock;
    double* row_sums = row_blocks ? scratch.row_sums.as<double>() : nullptr;
    fields(s, amp, phase, L, chirp, blu, scratch, amp_index, defocus_coefs, unit0, pupils);
    {
        // algorithmic bytes: the field in, row transforms out
        ProfileScope ps("psf-rows", (8.0 * P * P + 8.0 * P * R) * L);
#define LDH_PSF_ROWS(MM) launch_rows<MM>(m_, s, scratch.field.as<float2>(), P, R, offset_, chirp, bhat, blu, g_.as<float2>(), tw, L)
        if (!(LDH_PSF_LENGTHS(LDH_PSF_ROWS))) fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
#undef LDH_PSF_ROWS
        LDH_CHECK_LAUNCH("psf-rows");
    }
    {
        ProfileScope ps_t("psf-transpose", 16.0 * P * R * L);
        hipLaunchKernelGGL(transpose_c64_kernel, dim3(div_up(R, 32), div_up(P, 32), L), dim3(256), 0, s,
                           g_.as<float2>(), gt_.as<float2>(), P, R);
        LDH_CHECK_LAUNCH("psf-transpose");
    }
    {
        // algorithmic bytes: the P pass-1 values of each output row in, |F|^2 out
        ProfileScope ps("psf-cols", (8.0 * P * R + 4.0 * R * R) * L);
#define LDH_PSF_COLS(MM) \
    launch_cols<MM>(m_, s, gt_.as<float2>(), P, R, R - R / 2, offset_, chirp, bhat, blu, out, tw, row_sums, L)
        if (!(LDH_PSF_LENGTHS(LDH_PSF_COLS))) fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
#undef LDH_PSF_COLS
        LDH_CHECK_LAUNCH("psf-cols");
    }
    SeqSegments g;
    g.array[0] = out;
    g.per_array = L;
    g.stride = int64_t(R) * R;
    seqsum.run(s, g, int64_t(R) * R, L, totals, row_sums, row_blocks ? R : kSeqSumBlock);
    if (normalize) seqsum.scale(s, out, int64_t(R) * R, int64_t(R) * R, L, totals);
}

int resample_window_half_width(const std::vector<float>& scales, int support) {
    const float kc = float(support);  // (K - 1) / 2
    int h = 0;
    for (float sc : scales) {
        // the resample's own float arithmetic (resample_kernel): sample positions reach
        // centre +- kc * inv, the box filter half a footprint further, bilinear one texel more
        const float inv = 1.0f / std::max(sc, 0.05f);
        const float extent = kc * inv + (inv > 1.0f ? inv * 0.5f : 0.0f);
        h = std::max(h, int(std::ceil(extent)) + 2);
    }
    return h;
}

bool PsfPlan::prepare_window(hipStream_t s, int h) {
    const int P = pupil_, R = raw_;
    const int n = 2 * h + 1;
    // the window must stay inside the PSF (R is even: centre R/2) and its lines must be cheaper
    // than the full form's P + R lines of length m_
    if (h < 1 || R / 2 - h < 1 || R / 2 + h > R - 2) return false;
    const int mw = psf_length_at_least(P + n - 1);
    if (mw == 0 || double(P + n) * mw >= 0.8 * double(P + R) * m_) return false;
    if (h == h_) return true;
    // chirp-z onto k = k0 + m, m in [0, n), k0 = -h, with W = exp(-2 i pi_f / R):
    //   X[k] = W^(k^2/2) sum_j (x_j W^(j^2/2)) W^(-(k-j)^2/2)
    // pre-chirp W^(j^2/2) = exp(-i pi_f j^2 / R) for j < P; filter g[d] = exp(+i pi_f (k0+d)^2 / R)
    // for lags d in [-(P-1), n-1], laid out circularly over mw >= P + n - 1 points and transformed
    // in double; the post-chirp W^(k^2/2) has unit modulus and |F|^2 drops it.
    const double pi_f = double(kPiF);
    std::vector<float2> chirp(static_cast<size_t>(P));
    for (int j = 0; j < P; ++j) {
        const double ph = -pi_f * double(j) * double(j) / double(R);
        chirp[size_t(j)] = make_float2(float(std::cos(ph)), float(std::sin(ph)));
    }
    std::vector<double> br(static_cast<size_t>(mw), 0.0), bi(static_cast<size_t>(mw), 0.0);
    for (int d = -(P - 1); d <= n - 1; ++d) {
        const double k = double(d - h);
        const double ph = pi_f * k * k / double(R);
        const size_t at = size_t(d >= 0 ? d : mw + d);
        // float, as the full form's table entries, before the double transform
        br[at] = float(std::cos(ph));
        bi[at] = float(std::sin(ph));
    }
    host_dft(br, bi);
    std::vector<float2> bhat(static_cast<size_t>(mw));
    for (int k = 0; k < mw; ++k) bhat[size_t(k)] = make_float2(float(br[size_t(k)]), float(bi[size_t(k)]));
    wchirp_.reset(sizeof(float2) * P);
    wbhat_.reset(sizeof(float2) * mw);
    LDH_HIP(hipMemcpyAsync(wchirp_.get(), chirp.data(), sizeof(float2) * P, hipMemcpyHostToDevice, s), "psf-window");
    LDH_HIP(hipMemcpyAsync(wbhat_.get(), bhat.data(), sizeof(float2) * mw, hipMemcpyHostToDevice, s), "psf-window");
    LDH_HIP(hipStreamSynchronize(s), "psf-window");
    h_ = h;
    mw_ = mw;
    return true;
}

void PsfPlan::window_psfs(hipStream_t s, const FftTables& t, const float* amp, const float* phase, int count,
                          float* out, PsfScratch& scratch, const int* amp_index, const PhaseCoef* defocus_coefs,
                          const int* unit0, int pupils) {
    if (h_ < 1) fail(LDH_ERR_DEVICE, "psf-window-not-prepared");
    const int P = pupil_, L = count, n = 2 * h_ + 1;
    scratch.reserve(P, n, L);
    const float2* tw = t.twiddle.as<float2>();
    const float2* chirp = wchirp_.as<float2>();
    const float2* bhat = wbhat_.as<float2>();
    fields(s, amp, phase, L, chirp, 1, scratch, amp_index, defocus_coefs, unit0, pupils);
    {
        ProfileScope ps("psf-rows", (8.0 * P * P + 8.0 * P * n) * L);
#define LDH_PSF_ROWS(MM) launch_rows<MM>(mw_, s, scratch.field.as<float2>(), P, n, 0, chirp, bhat, 1, scratch.g.as<float2>(), tw, L)
        if (!(LDH_PSF_LENGTHS(LDH_PSF_ROWS))) fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
#undef LDH_PSF_ROWS
        LDH_CHECK_LAUNCH("psf-rows");
    }
    {
        ProfileScope ps_t("psf-transpose", 16.0 * P * n * L);
        hipLaunchKernelGGL(transpose_c64_kernel, dim3(div_up(n, 32), div_up(P, 32), L), dim3(256), 0, s,
                           scratch.g.as<float2>(), scratch.gt.as<float2>(), P, n);
        LDH_CHECK_LAUNCH("psf-transpose");
    }
    {
        ProfileScope ps("psf-cols", (8.0 * P * n + 4.0 * n * n) * L);
#define LDH_PSF_COLS(MM) \
    launch_cols<MM>(mw_, s, scratch.gt.as<float2>(), P, n, 0, 0, chirp, bhat, 1, out, tw, nullptr, L)
        if (!(LDH_PSF_LENGTHS(LDH_PSF_COLS))) fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
#undef LDH_PSF_COLS
        LDH_CHECK_LAUNCH("psf-cols");
    }
}

}  // namespace ldh

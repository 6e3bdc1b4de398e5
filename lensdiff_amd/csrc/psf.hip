// PSF synthesis for a batch of wavelengths (buildShiftedRawPsf, LensDiffCpuReference.cpp:831-859,
// driven per wavelength by buildPsfBankGlobalOnly :1702-1734).
//
// The R x R transform only has a P x P non-zero block (the pupil, embedded at offset
// floor((R-P)/2)), so it is computed as two separable line stages:
//   rows:    for each pupil row y, the R-point DFT of its P embedded values  -> G[y][kx]
//   columns: for each kx, the R-point DFT over the P pupil rows of G        -> F[ky][kx]
// Each R-point DFT is done in LDS. For power-of-two R it is a plain Stockham FFT of length R.
// Otherwise it is Bluestein's algorithm with a power-of-two M = nextPow2(2R-1), using the
// reference's chirp exp(-i pi j^2 / R) evaluated in double with pi rounded to float
// (LensDiffCpuReference.cpp:244-274) so the sampled spectrum matches the reference's.
// The column stage writes |F|^2 transposed; a shift-transpose pass then produces the centred
// raw PSF, which the sequential-sum normalisation scales to unit energy.
#include <cmath>

#include "fft_core.h"
#include "psf.h"
#include "pupil.h"

namespace ldh {

namespace {

template <int M>
__device__ __forceinline__ void lds_zero(float2* lds, int tid) {
    constexpr int NT = TileShape<M, true>::NT;
    for (int j = tid; j < M; j += NT) lds[lds_pad(j)] = make_float2(0.0f, 0.0f);
}

// After the line holds its (chirped) input: forward DFT of length M, and for Bluestein the
// multiply by B^ and the inverse DFT. Leaves X[k] * (1/M) (Bluestein) in LDS, unchirped.
template <int M>
__device__ __forceinline__ void line_transform(float2* lds, const float2* __restrict__ tw,
                                               const float2* __restrict__ bhat, bool bluestein, int tid) {
    constexpr int NT = TileShape<M, true>::NT;
    fft_stages<M, 1, 1, NT, false, M, 1>(lds, tw, tid);
    if (bluestein) {
        for (int k = tid; k < M; k += NT) lds[lds_pad(k)] = cmul(lds[lds_pad(k)], bhat[k]);
        __syncthreads();
        fft_stages<M, 1, 1, NT, true, M, 1>(lds, tw, tid);
    }
}

template <int M>
__global__ void __launch_bounds__((TileShape<M, true>::NT))
psf_rows_kernel(const float* __restrict__ amp, const float* __restrict__ phase, int P, int R, int off,
                const float2* __restrict__ chirp, const float2* __restrict__ bhat, int bluestein,
                float2* __restrict__ G, const float2* __restrict__ tw) {
    constexpr int NT = TileShape<M, true>::NT;
    extern __shared__ float2 lds[];
    const int y = blockIdx.x, lam = blockIdx.y, tid = threadIdx.x;
    lds_zero<M>(lds, tid);
    __syncthreads();
    const float* arow = amp + size_t(y) * P;
    const float* prow = phase + (size_t(lam) * P + y) * P;
    for (int x = tid; x < P; x += NT) {
        const float a = arow[x];
        float rad = 0.0f;
        if (a > 0.0f) rad = prow[x] * 2.0f * kPiF;
        float2 v = make_float2(a * cosf(rad), a * sinf(rad));
        const int j = x + off;
        if (bluestein) v = cmul(v, chirp[j]);
        lds[lds_pad(j)] = v;
    }
    __syncthreads();
    line_transform<M>(lds, tw, bhat, bluestein != 0, tid);
    float2* out = G + (size_t(lam) * P + y) * R;
    const float inv_m = 1.0f / float(M);
    for (int k = tid; k < R; k += NT) {
        float2 v = lds[lds_pad(k)];
        if (bluestein) v = cmul(make_float2(v.x * inv_m, v.y * inv_m), chirp[k]);
        out[k] = v;
    }
}

template <int M>
__global__ void __launch_bounds__((TileShape<M, true>::NT))
psf_cols_kernel(const float2* __restrict__ GT, int P, int R, int off, const float2* __restrict__ chirp,
                const float2* __restrict__ bhat, int bluestein, float* __restrict__ IT,
                const float2* __restrict__ tw) {
    constexpr int NT = TileShape<M, true>::NT;
    extern __shared__ float2 lds[];
    const int kx = blockIdx.x, lam = blockIdx.y, tid = threadIdx.x;
    lds_zero<M>(lds, tid);
    __syncthreads();
    const float2* in = GT + (size_t(lam) * R + kx) * P;
    for (int yy = tid; yy < P; yy += NT) {
        float2 v = in[yy];
        const int j = yy + off;
        if (bluestein) v = cmul(v, chirp[j]);
        lds[lds_pad(j)] = v;
    }
    __syncthreads();
    line_transform<M>(lds, tw, bhat, bluestein != 0, tid);
    float* out = IT + (size_t(lam) * R + kx) * R;
    const float inv_m = 1.0f / float(M);
    for (int k = tid; k < R; k += NT) {
        float2 v = lds[lds_pad(k)];
        if (bluestein) v = cmul(make_float2(v.x * inv_m, v.y * inv_m), chirp[k]);
        out[k] = v.x * v.x + v.y * v.y;
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

// S[y][x] = IT[(x + h) % R][(y + h) % R]: undoes the column stage's transpose and applies the
// centring shift by h = R/2 (fftShiftSquare, LensDiffCpuReference.cpp:303-314).
__global__ void __launch_bounds__(256)
shift_transpose_kernel(const float* __restrict__ IT, float* __restrict__ S, int R) {
    __shared__ float tile[32][33];
    const int lam = blockIdx.z;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const int h = R / 2;
    const float* src = IT + size_t(lam) * R * R;
    float* dst = S + size_t(lam) * R * R;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // load: tile[k][tx] = IT[a = (x0+k+h)%R][b = (y0+tx+h)%R]
    for (int k = ty; k < 32; k += 8) {
        const int x = x0 + k, y = y0 + tx;
        if (x < R && y < R) tile[k][tx] = src[size_t((x + h) % R) * R + (y + h) % R];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int y = y0 + k, x = x0 + tx;
        if (x < R && y < R) dst[size_t(y) * R + x] = tile[tx][k];
    }
}

__global__ void pupil_amplitude_kernel(PupilGeom g, float* __restrict__ out) {
    const int n = g.size;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    out[i] = pupil_amplitude_at(g, i % n, i / n);
}

__global__ void pupil_phase_kernel(const PhaseCoef* __restrict__ coefs, int n, float* __restrict__ out) {
    const int lam = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    const PhaseCoef c = coefs[lam];
    out[size_t(lam) * n * n + i] = phase_waves_at(c, i % n, i / n, n);
}

// Double-precision iterative radix-2 FFT on the host (tables only).
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

template <int M>
void launch_rows(hipStream_t s, const float* amp, const float* phase, int P, int R, int off,
                 const float2* chirp, const float2* bhat, int bluestein, float2* G, const float2* tw, int L) {
    using S = TileShape<M, true>;
    allow_large_lds(reinterpret_cast<const void*>(&psf_rows_kernel<M>), (M + M / 32 + 1) * 8);
    hipLaunchKernelGGL((psf_rows_kernel<M>), dim3(P, L), dim3(S::NT), (M + M / 32 + 1) * 8, s, amp, phase, P, R,
                       off, chirp, bhat, bluestein, G, tw);
}
template <int M>
void launch_cols(hipStream_t s, const float2* GT, int P, int R, int off, const float2* chirp, const float2* bhat,
                 int bluestein, float* IT, const float2* tw, int L) {
    using S = TileShape<M, true>;
    allow_large_lds(reinterpret_cast<const void*>(&psf_cols_kernel<M>), (M + M / 32 + 1) * 8);
    hipLaunchKernelGGL((psf_cols_kernel<M>), dim3(R, L), dim3(S::NT), (M + M / 32 + 1) * 8, s, GT, P, R, off,
                       chirp, bhat, bluestein, IT, tw);
}

}  // namespace

void PsfPlan::prepare(hipStream_t s, int pupil, int raw) {
    if (pupil == pupil_ && raw == raw_) return;
    if (raw > kMaxRawPsf) fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
    pupil_ = pupil;
    raw_ = raw;
    bluestein_ = !is_pow2(raw);
    m_ = bluestein_ ? pow2_at_least(2 * raw - 1) : raw;
    offset_ = std::max(0, (raw - pupil) / 2);
    if (!bluestein_) return;
    // chirp[j] = exp(-i * pi_f * j^2 / R), the reference's forward chirp (sign -1), with the
    // phase formed in double from pi rounded to float.
    const double pi_f = double(kPiF);
    std::vector<float2> chirp(static_cast<size_t>(raw));
    std::vector<double> br(static_cast<size_t>(m_), 0.0), bi(static_cast<size_t>(m_), 0.0);
    for (int j = 0; j < raw; ++j) {
        const double ph = -1.0 * pi_f * double(j) * double(j) / double(raw);
        chirp[size_t(j)] = make_float2(float(std::cos(ph)), float(std::sin(ph)));
        const float cr = float(std::cos(-ph)), ci = float(std::sin(-ph));
        br[size_t(j)] = cr; bi[size_t(j)] = ci;
        if (j != 0) { br[size_t(m_ - j)] = cr; bi[size_t(m_ - j)] = ci; }
    }
    host_fft_pow2(br, bi);
    std::vector<float2> bhat(static_cast<size_t>(m_));
    for (int k = 0; k < m_; ++k) bhat[size_t(k)] = make_float2(float(br[size_t(k)]), float(bi[size_t(k)]));
    chirp_.reset(sizeof(float2) * raw);
    bhat_.reset(sizeof(float2) * m_);
    LDH_HIP(hipMemcpyAsync(chirp_.get(), chirp.data(), sizeof(float2) * raw, hipMemcpyHostToDevice, s), "psf-chirp");
    LDH_HIP(hipMemcpyAsync(bhat_.get(), bhat.data(), sizeof(float2) * m_, hipMemcpyHostToDevice, s), "psf-chirp");
    LDH_HIP(hipStreamSynchronize(s), "psf-chirp");
}

void pupil_amplitude(hipStream_t s, const PupilGeom& g, float* out) {
    const int n = g.size;
    hipLaunchKernelGGL(pupil_amplitude_kernel, dim3(div_up(int64_t(n) * n, 256)), dim3(256), 0, s, g, out);
    LDH_CHECK_LAUNCH("pupil-amplitude");
}

void pupil_phase(hipStream_t s, const PhaseCoef* d_coefs, int n, int count, float* out) {
    hipLaunchKernelGGL(pupil_phase_kernel, dim3(div_up(int64_t(n) * n, 256), count), dim3(256), 0, s, d_coefs, n, out);
    LDH_CHECK_LAUNCH("pupil-phase");
}

void PsfPlan::raw_psfs(hipStream_t s, const FftTables& t, const float* amp, const float* phase, int count,
                       float* out, float* totals, SeqSum& seqsum, bool normalize) {
    const int P = pupil_, R = raw_, L = count;
    g_.reset(sizeof(float2) * size_t(P) * R * L);
    gt_.reset(sizeof(float2) * size_t(P) * R * L);
    it_.reset(sizeof(float) * size_t(R) * R * L);
    const float2* tw = t.twiddle.as<float2>();
    const float2* chirp = chirp_.as<float2>();
    const float2* bhat = bhat_.as<float2>();
    const int blu = bluestein_ ? 1 : 0;
    switch (m_) {
#define LDH_PSF_M(MM)                                                                                \
    case MM:                                                                                         \
        launch_rows<MM>(s, amp, phase, P, R, offset_, chirp, bhat, blu, g_.as<float2>(), tw, L);   \
        break;
        LDH_PSF_M(128) LDH_PSF_M(256) LDH_PSF_M(512) LDH_PSF_M(1024) LDH_PSF_M(2048) LDH_PSF_M(4096)
        LDH_PSF_M(8192) LDH_PSF_M(16384)
#undef LDH_PSF_M
        default: fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
    }
    LDH_CHECK_LAUNCH("psf-rows");
    hipLaunchKernelGGL(transpose_c64_kernel, dim3(div_up(R, 32), div_up(P, 32), L), dim3(256), 0, s,
                       g_.as<float2>(), gt_.as<float2>(), P, R);
    LDH_CHECK_LAUNCH("psf-transpose");
    switch (m_) {
#define LDH_PSF_M(MM)                                                                                  \
    case MM:                                                                                           \
        launch_cols<MM>(s, gt_.as<float2>(), P, R, offset_, chirp, bhat, blu, it_.as<float>(), tw, L); \
        break;
        LDH_PSF_M(128) LDH_PSF_M(256) LDH_PSF_M(512) LDH_PSF_M(1024) LDH_PSF_M(2048) LDH_PSF_M(4096)
        LDH_PSF_M(8192) LDH_PSF_M(16384)
#undef LDH_PSF_M
        default: fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
    }
    LDH_CHECK_LAUNCH("psf-cols");
    hipLaunchKernelGGL(shift_transpose_kernel, dim3(div_up(R, 32), div_up(R, 32), L), dim3(256), 0, s,
                       it_.as<float>(), out, R);
    LDH_CHECK_LAUNCH("psf-shift");
    if (normalize)
        seqsum.normalize(s, out, int64_t(R) * R, int64_t(R) * R, L, totals);
    else
        seqsum.run(s, out, int64_t(R) * R, int64_t(R) * R, L, totals);
}

}  // namespace ldh

// IO policies for the FFT tile kernel and the row / column pass launchers built on them.
//
// A 2D transform of a H x W grid (row pitch W) is one row pass plus a column transform. Column
// lengths that are column tile lengths take one pass; longer ones are split N = N1 * N2
// (col_split) into two passes over tiles of kColWidth adjacent columns:
//   forward  F1: length-N1 DFTs over rows n2 + N2*n1, then twiddle W_N^(n2*k1), in place;
//            F2: length-N2 DFTs over the contiguous rows N2*k1 + n2, in place.
//   The spectrum row of frequency k = k1 + N1*k2 is then stored at row N2*k1 + k2.
//   inverse  I2: length-N2 inverse DFTs over rows N2*k1 + k2, then twiddle W_N^(-k1*n2);
//            I1: length-N1 inverse DFTs over rows n2 + N2*k1, giving natural order.
// Pointwise spectral work only needs matching layouts, so convolution never reorders.
#pragma once

#include <hip/hip_fp16.h>

#include <algorithm>
#include <string>
#include <vector>

#include "fft_core.h"
#include "profiler.h"

namespace ldh {

// ---- element ops ---------------------------------------------------------------------------
// extra_bytes: bytes an op reads per element besides the element itself (byte model).
// pair(d, idx): elements idx and idx + 1 (idx even: one 16-byte load per array).
// at(c): the op for channel c of a multi-channel pass (its own arrays at c times a stride).
struct PlainLoad {
    static constexpr double extra_bytes = 0.0;
    __device__ PlainLoad at(int) const { return *this; }
    __device__ float2 operator()(const float2* d, size_t idx) const { return d[idx]; }
    __device__ float4 pair(const float2* d, size_t idx) const { return *reinterpret_cast<const float4*>(d + idx); }
};
// Spectrum multiply fused into the first inverse pass: v = d[idx] * k[idx]; channel c multiplies
// by the spectrum at k + c * chan.
struct MulSpectrumLoad {
    static constexpr double extra_bytes = 8.0;
    const float2* k;
    size_t chan = 0;
    __device__ MulSpectrumLoad at(int c) const { return MulSpectrumLoad{k + size_t(c) * chan, chan}; }
    __device__ float2 operator()(const float2* d, size_t idx) const { return cmul(d[idx], k[idx]); }
    // the product of elements a (two, as pair() loads them) with the spectrum at idx, idx + 1
    __device__ float4 mul_pair(float4 a, size_t idx) const {
        const float4 b = *reinterpret_cast<const float4*>(k + idx);
        const float2 lo = cmul(make_float2(a.x, a.y), make_float2(b.x, b.y));
        const float2 hi = cmul(make_float2(a.z, a.w), make_float2(b.z, b.w));
        return make_float4(lo.x, lo.y, hi.x, hi.y);
    }
    __device__ float4 pair(const float2* d, size_t idx) const {
        const float4 a = *reinterpret_cast<const float4*>(d + idx);
        const float4 b = *reinterpret_cast<const float4*>(k + idx);
        const float2 lo = cmul(make_float2(a.x, a.y), make_float2(b.x, b.y));
        const float2 hi = cmul(make_float2(a.z, a.w), make_float2(b.z, b.w));
        return make_float4(lo.x, lo.y, hi.x, hi.y);
    }
};
// Same with the kernel spectrum stored as FP16 pairs (FP32 arithmetic).
struct MulSpectrumHalfLoad {
    static constexpr double extra_bytes = 4.0;
    const __half2* k;
    size_t chan = 0;
    __device__ MulSpectrumHalfLoad at(int c) const { return MulSpectrumHalfLoad{k + size_t(c) * chan, chan}; }
    __device__ float4 mul_pair(float4 a, size_t idx) const {
        const __half2 k0 = k[idx], k1 = k[idx + 1];
        const float2 lo = cmul(make_float2(a.x, a.y), __half22float2(k0));
        const float2 hi = cmul(make_float2(a.z, a.w), __half22float2(k1));
        return make_float4(lo.x, lo.y, hi.x, hi.y);
    }
    __device__ float2 operator()(const float2* d, size_t idx) const {
        const float2 kk = __half22float2(k[idx]);
        return cmul(d[idx], kk);
    }
    __device__ float4 pair(const float2* d, size_t idx) const {
        const float4 a = *reinterpret_cast<const float4*>(d + idx);
        const __half2 k0 = k[idx], k1 = k[idx + 1];
        const float2 lo = cmul(make_float2(a.x, a.y), __half22float2(k0));
        const float2 hi = cmul(make_float2(a.z, a.w), __half22float2(k1));
        return make_float4(lo.x, lo.y, hi.x, hi.y);
    }
};

// ---- column passes ----------------------------------------------------------------------------
// Rows a column pass may skip: loads of rows in [zero_lo, zero_hi) read as zero without touching
// memory (rows the caller knows are zero), and rows >= store_rows are not stored (rows nobody
// reads afterwards).
struct ColRows {
    int zero_lo = 0, zero_hi = 0;
    int store_rows = 1 << 30;
};

// row(g, i) = g * g_row_mul + i * i_row_mul; the twiddle exponent of an output point is g * i.
// A pass over nchan channels (arrays of equal shape) treats them as one grid nchan times as wide:
// column group cg of channel c is tile column nchan * cg + c, so a column group's channels are
// neighbouring tiles and share, through L2, what they read in common (src_chan = 0).
template <class Load, int TW>
struct ColIO {
    const float2* src;   // loads (may equal data); channel c at src + c * src_chan
    float2* data;        // stores; channel c at data + c * dst_chan
    const float2* tw;
    int pitch;
    int ncolgroups;   // nchan * width / kColWidth
    int groups;       // line groups (g values) per tile
    int gcount;       // line groups in the pass: tiles past it (a partial last tile) do nothing
    int g_row_mul, i_row_mul;
    int tw_stride;    // table length / N for the twiddles W_N^(g i) of the split
    ColRows rows;
    Load op;
    int nchan = 1;
    size_t src_chan = 0, dst_chan = 0;
    __device__ int row_of(int tile, int line, int i, int* g_out, int* col, int* ch) const {
        const int t = tile % ncolgroups;
        const int cg = t / nchan;
        *ch = t - cg * nchan;
        const int g = (tile / ncolgroups) * groups + line / kColWidth;
        *g_out = g;
        *col = cg * kColWidth + line % kColWidth;
        return g * g_row_mul + i * i_row_mul;
    }
    // Every lane loads: a point that reads as zero (past the last line group, or in the zero band)
    // loads element 0 of the source instead and is replaced by 0, so the loads of a batch issue
    // back to back instead of each behind its own branch.
    using Raw = float2;
    __device__ float2 finish(float2 v) const { return v; }
    __device__ float2 fetch(int tile, int line, int i) const {
        int g, col, ch;
        const int row = row_of(tile, line, i, &g, &col, &ch);
        const bool zero = g >= gcount || (row >= rows.zero_lo && row < rows.zero_hi);
        const float2 v = op.at(ch)(src + size_t(ch) * src_chan, zero ? size_t(0) : size_t(row) * size_t(pitch) + size_t(col));
        return zero ? make_float2(0.0f, 0.0f) : v;
    }
    // The same for the two adjacent columns line, line + 1 (line even): one 16-byte access.
    static constexpr bool kPairs = true;
    __device__ float4 fetch_pair(int tile, int line, int i) const {
        int g, col, ch;
        const int row = row_of(tile, line, i, &g, &col, &ch);
        const bool zero = g >= gcount || (row >= rows.zero_lo && row < rows.zero_hi);
        const float4 v = op.at(ch).pair(src + size_t(ch) * src_chan,
                                        zero ? size_t(0) : size_t(row) * size_t(pitch) + size_t(col));
        return zero ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : v;
    }
    __device__ void store_pair(int tile, int line, int i, float2 a, float2 b, float2 w) const {
        int g, col, ch;
        const int row = row_of(tile, line, i, &g, &col, &ch);
        if (g >= gcount || row >= rows.store_rows) return;
        if constexpr (TW != 0) {
            a = cmul(a, w);
            b = cmul(b, w);
        }
        *reinterpret_cast<float4*>(data + size_t(ch) * dst_chan + size_t(row) * size_t(pitch) + size_t(col)) =
            make_float4(a.x, a.y, b.x, b.y);
    }
    // The split's twiddle W_N^(g i) of a stored point (conjugated for TW < 0), loaded by the tile
    // kernel before its stages; anything for a point that is not stored.
    static constexpr bool kStoreTwiddle = TW != 0;
    __device__ float2 twiddle(int tile, int line, int i) const {
        int g, col, ch;
        (void)row_of(tile, line, i, &g, &col, &ch);
        float2 w = tw[(g < gcount ? g : 0) * i * tw_stride];
        if (TW < 0) w.y = -w.y;
        return w;
    }
    __device__ void store(int tile, int line, int i, float2 v, float2 w) const {
        int g, col, ch;
        const int row = row_of(tile, line, i, &g, &col, &ch);
        if (g >= gcount || row >= rows.store_rows) return;
        if constexpr (TW != 0) v = cmul(v, w);
        data[size_t(ch) * dst_chan + size_t(row) * size_t(pitch) + size_t(col)] = v;
    }
};

// One column pass over nchan channels (ColIO): channel c reads src + c * src_chan through
// op.at(c) and writes data + c * dst_chan. Strides keep 16-byte alignment.
template <int TW, class Load>
void col_pass(hipStream_t s, const float2* tw, const float2* src, float2* data, int width, int npts, int gcount,
              int g_row_mul, int i_row_mul, int height, const Load& op, ColRows rows, bool inverse, int nchan = 1,
              size_t src_chan = 0, size_t dst_chan = 0) {
    const int nl = col_tile_lines(npts);
    const int groups = nl / kColWidth;
    // column tiles move 16-byte pairs of adjacent elements
    if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(data)) & 15) != 0 || (src_chan | dst_chan) % 2)
        fail(LDH_ERR_INVALID_ARGUMENT, "fft-column-data-not-16-byte-aligned");
    if (nchan < 1) fail(LDH_ERR_INVALID_ARGUMENT, "fft-column-channels");
    ColIO<Load, TW> io{src, data, tw + tw_offset(height), width, nchan * (width / kColWidth), groups, gcount,
                       g_row_mul, i_row_mul, tw_table(height) / height, rows, op, nchan, src_chan, dst_chan};
    const int ntiles = nchan * (width / kColWidth) * int(div_up(gcount, groups));
    // algorithmic bytes: elements read (outside the zero band; a source the channels share once)
    // and written (below store_rows)
    const double all = double(width) * double(height);
    const double zero = double(std::max(0, std::min(rows.zero_hi, height) - std::max(rows.zero_lo, 0)));
    const double stored = double(std::min(rows.store_rows, height));
    const double loaded = all - zero * double(width);
    const double src_reads = src_chan != 0 ? double(nchan) : 1.0;
    ProfileScope ps(std::string("fft-col-pass") + (inverse ? "-inv" : "-fwd"),
                    8.0 * loaded * src_reads + Load::extra_bytes * loaded * nchan + 8.0 * stored * double(width) * nchan);
    if (inverse) launch_tiles<false, true>(s, npts, io, tw, ntiles);
    else launch_tiles<false, false>(s, npts, io, tw, ntiles);
}

inline void check_cols(int width, int height) {
    if (width < kColWidth || width % kColWidth) fail(LDH_ERR_INVALID_ARGUMENT, "fft-width-not-multiple-of-16");
    if (!grid_length_ok(height)) fail(LDH_ERR_INVALID_ARGUMENT, "fft-height-unsupported");
}

// ---- column split plans ---------------------------------------------------------------------------
// A height that is not a column tile length is transformed in two passes N = N1 x N2; every pair
// of tile lengths whose product is the height (in either order) computes the same transform, with
// different tile shapes, occupancy and twiddle traffic. The split also fixes the row order of the
// spectrum (col_storage_row), so it is chosen once per height and device for the life of the
// process: the first column transform of a height times every split (forward and inverse passes
// over scratch memory of the caller's width, HIP events on its stream) and keeps the fastest. A
// height first met while the stream is capturing a graph keeps the balanced split of col_split.
struct ColSplit {
    int n1 = 0, n2 = 0;  // n1 == 0: one pass of length n2
};
std::vector<ColSplit> col_splits(int height);
ColSplit col_plan(hipStream_t s, const float2* tw, int width, int height);
// One line per planned height: "fft-cols/<height>[n1xn2,...]@dev<d>: v<chosen> <ms per split> ms".
std::string col_plan_report();
__host__ __device__ inline int col_storage_row(ColSplit sp, int k) {
    return sp.n1 == 0 ? k : sp.n2 * (k % sp.n1) + k / sp.n1;
}

// Forward column transform (in place) with split sp. `op` applies to the loads of the first
// pass, which also reads the rows [zero_lo, zero_hi) of `first` as zero.
template <class Load>
void cols_forward_split(hipStream_t s, const float2* tw, float2* data, int width, int height, const Load& op,
                        ColRows first, ColSplit sp) {
    first.store_rows = 1 << 30;
    const int n1 = sp.n1, n2 = sp.n2;
    if (n1 == 0) {
        col_pass<0>(s, tw, data, data, width, height, 1, 0, 1, height, op, first, false);
        return;
    }
    col_pass<1>(s, tw, data, data, width, n1, n2, 1, n2, height, op, first, false);
    col_pass<0>(s, tw, data, data, width, n2, n1, n2, 1, height, PlainLoad{}, ColRows{}, false);
}

// Inverse column transform (unscaled) with split sp. The first pass loads from `src` through
// `op` and every pass stores into `data` (src == data for in place); the last pass stores only
// rows < store_rows.
template <class Load>
void cols_inverse_split(hipStream_t s, const float2* tw, const float2* src, float2* data, int width, int height,
                        const Load& op, int store_rows, ColSplit sp) {
    ColRows last;
    last.store_rows = store_rows;
    const int n1 = sp.n1, n2 = sp.n2;
    if (n1 == 0) {
        col_pass<0>(s, tw, src, data, width, height, 1, 0, 1, height, op, last, true);
        return;
    }
    col_pass<-1>(s, tw, src, data, width, n2, n1, n2, 1, height, op, ColRows{}, true);
    col_pass<0>(s, tw, data, data, width, n1, n2, 1, n2, height, PlainLoad{}, last, true);
}

// The same with the height's planned split (col_plan).
template <class Load>
void cols_forward(hipStream_t s, const float2* tw, float2* data, int width, int height, const Load& op,
                  ColRows first = ColRows{}) {
    check_cols(width, height);
    cols_forward_split(s, tw, data, width, height, op, first, col_plan(s, tw, width, height));
}
template <class Load>
void cols_inverse(hipStream_t s, const float2* tw, const float2* src, float2* data, int width, int height,
                  const Load& op, int store_rows = 1 << 30) {
    check_cols(width, height);
    cols_inverse_split(s, tw, src, data, width, height, op, store_rows, col_plan(s, tw, width, height));
}

// The forward column transforms of nchan channels of one shape in place (channel c at
// data + c * chan) in the same launches, their rows [first.zero_lo, first.zero_hi) read as zero.
inline void cols_forward_channels(hipStream_t s, const float2* tw, float2* data, size_t chan, int nchan, int width,
                                  int height, ColRows first) {
    check_cols(width, height);
    const ColSplit sp = col_plan(s, tw, width, height);
    first.store_rows = 1 << 30;
    const int n1 = sp.n1, n2 = sp.n2;
    if (n1 == 0) {
        col_pass<0>(s, tw, data, data, width, height, 1, 0, 1, height, PlainLoad{}, first, false, nchan, chan, chan);
        return;
    }
    col_pass<1>(s, tw, data, data, width, n1, n2, 1, n2, height, PlainLoad{}, first, false, nchan, chan, chan);
    col_pass<0>(s, tw, data, data, width, n2, n1, n2, 1, height, PlainLoad{}, ColRows{}, false, nchan, chan, chan);
}

// The inverse column transforms of nchan channels in the same launches, all from one source:
// channel c is the transform of src through op.at(c), stored at data + c * dst_chan (only rows
// < store_rows of the last pass). Neighbouring tiles carry a column group's channels, so the
// source is read from memory about once.
template <class Load>
void cols_inverse_channels(hipStream_t s, const float2* tw, const float2* src, float2* data, size_t dst_chan,
                           int nchan, int width, int height, const Load& op, int store_rows) {
    check_cols(width, height);
    const ColSplit sp = col_plan(s, tw, width, height);
    ColRows last;
    last.store_rows = store_rows;
    const int n1 = sp.n1, n2 = sp.n2;
    if (n1 == 0) {
        col_pass<0>(s, tw, src, data, width, height, 1, 0, 1, height, op, last, true, nchan, 0, dst_chan);
        return;
    }
    col_pass<-1>(s, tw, src, data, width, n2, n1, n2, 1, height, op, ColRows{}, true, nchan, 0, dst_chan);
    col_pass<0>(s, tw, data, data, width, n1, n2, 1, n2, height, PlainLoad{}, last, true, nchan, dst_chan, dst_chan);
}

// ---- forward and inverse column passes fused for a convolution -----------------------------------
// With a split N1 x N2, the forward pass F2 and the inverse pass I2 transform the same row groups
// (rows N2 * k1 + j, j < N2), so one tile can run F2, then for every channel c of a product the
// multiply by c's spectrum, I2 and I2's conjugate twiddle: the forward result stays in registers
// and LDS. The operations are those of the two passes run apart (the spectrum product as
// MulSpectrumLoad forms it, the twiddles as the tile kernel forms them); complex products may
// contract to FMA differently in another kernel (cmul), so results agree to rounding (5.5e-9
// rel L2 on a UHD frame).
struct FwdMulInvGeom {
    const float2* src;   // the forward pass's input (F1's output)
    float2* data;        // channel c's I2 output at data + c * dst_chan
    size_t dst_chan;
    int nchan;
    int pitch, ncolgroups;   // ncolgroups = width / kColWidth
    int groups, gcount;      // line groups per tile; line groups in the pass (N1)
    int g_row_mul;           // N2 (i_row_mul 1)
    int tw_stride;           // table length / height
    const float2* twn;       // the height's twiddle table (W_height^m at m * tw_stride)
};

template <int N>
constexpr bool fwd_mul_inv_ok() {
    using S = TileShape<N, false>;
    constexpr int NP = S::NL / 2;
    return (S::ELEMS / 2) % S::NT == 0 && S::NT % NP == 0;
}

// ONE: a single channel (gm.nchan == 1), so the forward result is dead after its product.
template <int N, bool ONE, class Load>
__global__ void __launch_bounds__((TileShape<N, false>::NT))
fft_fwd_mul_inv_kernel(FwdMulInvGeom gm, Load op, const float2* __restrict__ tw, int ntiles) {
    using S = TileShape<N, false>;
    constexpr int NL = S::NL, NT = S::NT, E = S::ELEMS;
    constexpr int NP = NL / 2;
    constexpr int IP = E / 2 / NT;
    static_assert(fwd_mul_inv_ok<N>(), "column pairs");
    extern __shared__ float2 lds[];
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int tid = threadIdx.x;
    float2 tw_mine = make_float2(1.0f, 0.0f);
    if constexpr (S::TWL) {
        if (tid < N) tw_mine = tw[tw_offset(N) + tid * (tw_table(N) / N)];
    }
    const int cg = tile % gm.ncolgroups;
    const int lp = (tid % NP) * 2, i0 = tid / NP;
    const int g = (tile / gm.ncolgroups) * gm.groups + lp / kColWidth;
    const int col = cg * kColWidth + lp % kColWidth;
    const bool live = g < gm.gcount;
    size_t idx[IP];
#pragma unroll
    for (int c = 0; c < IP; ++c) idx[c] = size_t(g * gm.g_row_mul + i0 + c * (NT / NP)) * size_t(gm.pitch) + size_t(col);
    // F2's loads (a tile past the last line group reads element 0 and is dropped)
    float4 x[IP];
#pragma unroll
    for (int c = 0; c < IP; ++c) x[c] = *reinterpret_cast<const float4*>(gm.src + (live ? idx[c] : 0));
#pragma unroll
    for (int c = 0; c < IP; ++c) {
        const int at = lp + (i0 + c * (NT / NP)) * NL;
        lds[lds_pad(at)] = live ? make_float2(x[c].x, x[c].y) : make_float2(0.0f, 0.0f);
        lds[lds_pad(at + 1)] = live ? make_float2(x[c].z, x[c].w) : make_float2(0.0f, 0.0f);
    }
    float2* twl = lds + S::DATA_WORDS;
    if constexpr (S::TWL) {
        if (tid < N) twl[tid] = tw_mine;
    }
    // I2's store twiddles conj(W_height^(g i)), the same for every channel, formed as the tile
    // kernel forms them for column pairs: conj W^(g i0) times powers of conj W^(g D), D = NT / NP
    auto twd = [&](int i) {
        float2 t = gm.twn[(live ? g : 0) * i * gm.tw_stride];
        t.y = -t.y;
        return t;
    };
    float2 w[IP];
    {
        constexpr int D = NT / NP;
        const float2 base = twd(i0);
        float2 pw[IP];
        pw[0] = make_float2(1.0f, 0.0f);
        if constexpr (IP > 1) pw[1] = twd(D);
#pragma unroll
        for (int c = 2; c < IP; ++c) {
            const int h = c >= 8 ? 8 : c >= 4 ? 4 : c >= 2 ? 2 : 1;
            pw[c] = c == h ? cmul(pw[h / 2], pw[h / 2]) : cmul(pw[c - h], pw[h]);
        }
#pragma unroll
        for (int c = 0; c < IP; ++c) w[c] = c == 0 ? base : cmul(base, pw[c]);
    }
    __syncthreads();
    fft_stages<N, 1, NL, NT, false, 1, NL, S::TWL>(lds, tw, twl, tid);
#pragma unroll
    for (int c = 0; c < IP; ++c) {
        const int at = lp + (i0 + c * (NT / NP)) * NL;
        const float2 a = lds[lds_pad(at)], b = lds[lds_pad(at + 1)];
        x[c] = make_float4(a.x, a.y, b.x, b.y);
    }
    const int nchan = ONE ? 1 : gm.nchan;
    for (int ch = 0; ch < nchan; ++ch) {
        const Load opc = op.at(ch);
        float4 y[IP];
#pragma unroll
        for (int c = 0; c < IP; ++c) y[c] = opc.mul_pair(x[c], live ? idx[c] : 0);
        __syncthreads();  // the previous readers of the tile are done
#pragma unroll
        for (int c = 0; c < IP; ++c) {
            const int at = lp + (i0 + c * (NT / NP)) * NL;
            lds[lds_pad(at)] = make_float2(y[c].x, y[c].y);
            lds[lds_pad(at + 1)] = make_float2(y[c].z, y[c].w);
        }
        __syncthreads();
        fft_stages<N, 1, NL, NT, true, 1, NL, S::TWL>(lds, tw, twl, tid);
        if (live) {
            float2* out = gm.data + size_t(ch) * gm.dst_chan;
#pragma unroll
            for (int c = 0; c < IP; ++c) {
                const int at = lp + (i0 + c * (NT / NP)) * NL;
                const float2 a = cmul(lds[lds_pad(at)], w[c]), b = cmul(lds[lds_pad(at + 1)], w[c]);
                *reinterpret_cast<float4*>(out + idx[c]) = make_float4(a.x, a.y, b.x, b.y);
            }
        }
    }
}

template <int N, class Load>
bool launch_fwd_mul_inv_if(hipStream_t s, int n, const FwdMulInvGeom& gm, const Load& op, const float2* tw,
                           int ntiles) {
    if constexpr (!fwd_mul_inv_ok<N>()) {
        return false;
    } else {
        if (n != N) return false;
        using S = TileShape<N, false>;
        if (gm.nchan == 1) {
            allow_large_lds(reinterpret_cast<const void*>(&fft_fwd_mul_inv_kernel<N, true, Load>), S::LDS_BYTES);
            hipLaunchKernelGGL((fft_fwd_mul_inv_kernel<N, true, Load>), dim3(ntiles), dim3(S::NT), S::LDS_BYTES, s, gm,
                               op, tw, ntiles);
        } else {
            allow_large_lds(reinterpret_cast<const void*>(&fft_fwd_mul_inv_kernel<N, false, Load>), S::LDS_BYTES);
            hipLaunchKernelGGL((fft_fwd_mul_inv_kernel<N, false, Load>), dim3(ntiles), dim3(S::NT), S::LDS_BYTES, s, gm,
                               op, tw, ntiles);
        }
        return true;
    }
}

// The column work of a convolution of one transformed source with nchan spectra: the forward
// transform of `drv` (whose rows [first.zero_lo, first.zero_hi) read as zero), the product with
// op.at(c)'s spectrum and the inverse transform for every channel c, stored at data + c * dst_chan
// (rows < store_rows of the last pass). With a split whose N2 tile moves column pairs, F2 and I2
// run fused (fft_fwd_mul_inv_kernel) and the forward result is never stored; otherwise the passes
// run apart and `drv` holds the source's spectrum afterwards.
template <class Load>
void cols_convolve_channels(hipStream_t s, const float2* tw, float2* drv, ColRows first, float2* data,
                            size_t dst_chan, int nchan, int width, int height, const Load& op, int store_rows) {
    check_cols(width, height);
    const ColSplit sp = col_plan(s, tw, width, height);
    const int n1 = sp.n1, n2 = sp.n2;
    if (n1 != 0 && nchan >= 1) {
        const int nl = col_tile_lines(n2);
        FwdMulInvGeom gm{drv, data, dst_chan, nchan, width, width / kColWidth, nl / kColWidth, n1, n2,
                         tw_table(height) / height, tw + tw_offset(height)};
        const int ntiles = (width / kColWidth) * int(div_up(n1, nl / kColWidth));
        first.store_rows = 1 << 30;
        col_pass<1>(s, tw, drv, drv, width, n1, n2, 1, n2, height, PlainLoad{}, first, false);
        const double cells = double(width) * double(height);
        bool fused;
        {
            ProfileScope ps("fft-col-pass-fwd-inv", 8.0 * cells + (Load::extra_bytes + 8.0) * cells * nchan);
            fused = launch_fwd_mul_inv_if<8>(s, n2, gm, op, tw, ntiles) || launch_fwd_mul_inv_if<12>(s, n2, gm, op, tw, ntiles) ||
                    launch_fwd_mul_inv_if<16>(s, n2, gm, op, tw, ntiles) || launch_fwd_mul_inv_if<24>(s, n2, gm, op, tw, ntiles) ||
                    launch_fwd_mul_inv_if<32>(s, n2, gm, op, tw, ntiles) || launch_fwd_mul_inv_if<36>(s, n2, gm, op, tw, ntiles) ||
                    launch_fwd_mul_inv_if<48>(s, n2, gm, op, tw, ntiles) || launch_fwd_mul_inv_if<64>(s, n2, gm, op, tw, ntiles) ||
                    launch_fwd_mul_inv_if<72>(s, n2, gm, op, tw, ntiles) || launch_fwd_mul_inv_if<96>(s, n2, gm, op, tw, ntiles) ||
                    launch_fwd_mul_inv_if<128>(s, n2, gm, op, tw, ntiles) || launch_fwd_mul_inv_if<144>(s, n2, gm, op, tw, ntiles);
        }
        if (fused) {
            LDH_CHECK_LAUNCH("fft-col-fwd-inv");
            ColRows last;
            last.store_rows = store_rows;
            col_pass<0>(s, tw, data, data, width, n1, n2, 1, n2, height, PlainLoad{}, last, true, nchan, dst_chan,
                        dst_chan);
            return;
        }
        // no fused tile for this N2: F2 apart (F1 is done)
        col_pass<0>(s, tw, drv, drv, width, n2, n1, n2, 1, height, PlainLoad{}, ColRows{}, false);
        cols_inverse_channels(s, tw, drv, data, dst_chan, nchan, width, height, op, store_rows);
        return;
    }
    cols_forward_split(s, tw, drv, width, height, PlainLoad{}, first, sp);
    cols_inverse_channels(s, tw, drv, data, dst_chan, nchan, width, height, op, store_rows);
}

// ---- row passes ---------------------------------------------------------------------------------
// RowOp supplies load(row, x) and store(row, x, v). The pass covers `nrows` logical rows; logical
// row l is grid row l below `split` and l + gap from there (so a band of zero rows in the middle
// of the grid is skipped).
template <class RowOp>
struct RowIO {
    RowOp op;
    int lines_per_tile;
    int nrows;
    int split, gap;
    __device__ int row_of(int tile, int line) const {
        const int l = tile * lines_per_tile + line;
        return l < split ? l : l + gap;
    }
    // Row ops split a point's load into fetch (memory only, no branches) and finish (the
    // arithmetic). A line past the last (a partial last tile) fetches row 0 and reads as zero.
    struct Raw {
        typename RowOp::Raw r;
        bool in;
    };
    __device__ Raw fetch(int tile, int line, int i) const {
        const int l = tile * lines_per_tile + line;
        const bool in = l < nrows;
        return Raw{op.fetch(in ? row_of(tile, line) : 0, i), in};
    }
    __device__ float2 finish(const Raw& t) const {
        const float2 v = op.finish(t.r);
        return t.in ? v : make_float2(0.0f, 0.0f);
    }
    static constexpr bool kStoreTwiddle = false;
    static constexpr bool kPairs = false;
    static constexpr int kFetchBatch = io_fetch_batch<RowOp>::value;
    __device__ float4 fetch_pair(int, int, int) const { return make_float4(0.0f, 0.0f, 0.0f, 0.0f); }
    __device__ void store_pair(int, int, int, float2, float2, float2) const {}
    __device__ float2 twiddle(int, int, int) const { return make_float2(1.0f, 0.0f); }
    __device__ void store(int tile, int line, int i, float2 v, float2) const {
        const int l = tile * lines_per_tile + line;
        if (l < nrows) op.store(row_of(tile, line), i, v);
    }
    // Row ops with kStoreLds store by symmetry: for k = 0 .. op.hw - 1 of each line,
    // op.store_sym(row, k, Z[k], Z[(N - k) % N], N) (k < N; hw <= N).
    static constexpr bool kStoreLds = io_stores_lds<RowOp>::value;
    template <int N, int NL, int NT>
    __device__ void store_lds(int tile, int tid, const float2* lds) const {
        const int hw = op.hw;
        for (int e = tid; e < NL * hw; e += NT) {
            const int line = e / hw, k = e - line * hw;
            if (tile * lines_per_tile + line >= nrows) continue;
            const float2 zk = lds[lds_pad(line * N + k)];
            const float2 zm = lds[lds_pad(line * N + (N - k) % N)];
            op.store_sym(row_of(tile, line), k, zk, zm, N);
        }
    }
};

// The pass is timed as "fft-row-pass-inv" / "-fwd" with 16 bytes per point, or, given `family`,
// under that name with `bytes` (a pass whose loads and stores carry other data).
template <class RowOp>
void rows_pass(hipStream_t s, const float2* tw, int width, int nrows, const RowOp& op, bool inverse,
               int split = 1 << 30, int gap = 0, const char* family = nullptr, double bytes = 0.0) {
    const int nl = row_tile_lines(width);
    RowIO<RowOp> io{op, nl, nrows, split, gap};
    const int ntiles = int(div_up(nrows, nl));
    ProfileScope ps(family ? std::string(family) : std::string("fft-row-pass") + (inverse ? "-inv" : "-fwd"),
                    family ? bytes : 16.0 * double(width) * double(nrows));
    if (inverse) launch_tiles<true, true>(s, width, io, tw, ntiles);
    else launch_tiles<true, false>(s, width, io, tw, ntiles);
}

struct PlainRows {
    float2* data;
    int pitch;
    using Raw = float2;
    __device__ float2 fetch(int row, int x) const { return data[size_t(row) * pitch + x]; }
    __device__ float2 finish(float2 v) const { return v; }
    __device__ void store(int row, int x, float2 v) const { data[size_t(row) * pitch + x] = v; }
};

}  // namespace ldh

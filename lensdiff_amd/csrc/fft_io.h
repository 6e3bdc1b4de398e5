// IO policies for the FFT tile kernel and the row / column pass launchers built on them.
//
// A 2D transform of a H x W grid (row pitch W) is one row pass plus a column transform. Column
// lengths up to 64 take one pass; longer ones are split N = N1 * N2 into two passes over tiles
// of 16 adjacent columns (128-byte row segments):
//   forward  F1: length-N1 DFTs over rows n2 + N2*n1, then twiddle W_N^(n2*k1), in place;
//            F2: length-N2 DFTs over the contiguous rows N2*k1 + n2, in place.
//   The spectrum row of frequency k = k1 + N1*k2 is then stored at row N2*k1 + k2.
//   inverse  I2: length-N2 inverse DFTs over rows N2*k1 + k2, then twiddle W_N^(-k1*n2);
//            I1: length-N1 inverse DFTs over rows n2 + N2*k1, giving natural order.
// Pointwise spectral work only needs matching layouts, so convolution never reorders.
#pragma once

#include <hip/hip_fp16.h>

#include "fft_core.h"
#include "profiler.h"

namespace ldh {

// ---- element ops ---------------------------------------------------------------------------
struct PlainLoad {
    __device__ float2 operator()(const float2* d, size_t idx) const { return d[idx]; }
};
// Spectrum multiply fused into the first inverse pass: v = d[idx] * k[idx].
struct MulSpectrumLoad {
    const float2* k;
    __device__ float2 operator()(const float2* d, size_t idx) const { return cmul(d[idx], k[idx]); }
};
// Same with the kernel spectrum stored as FP16 pairs (FP32 arithmetic).
struct MulSpectrumHalfLoad {
    const __half2* k;
    __device__ float2 operator()(const float2* d, size_t idx) const {
        const float2 kk = __half22float2(k[idx]);
        return cmul(d[idx], kk);
    }
};

// ---- column passes ----------------------------------------------------------------------------
// row(g, i) = g * g_row_mul + i * i_row_mul; the twiddle exponent of an output point is g * i.
template <class Load, int TW>
struct ColIO {
    const float2* src;   // loads (may equal data)
    float2* data;        // stores
    const float2* tw;
    int pitch;
    int ncolgroups;   // width / 16
    int groups;       // line groups (g values) per tile
    int g_row_mul, i_row_mul;
    int tw_shift;
    Load op;
    __device__ size_t index(int tile, int line, int i, int* g_out) const {
        const int cg = tile % ncolgroups;
        const int g = (tile / ncolgroups) * groups + (line >> 4);
        *g_out = g;
        const int row = g * g_row_mul + i * i_row_mul;
        return size_t(row) * size_t(pitch) + size_t(cg * 16 + (line & 15));
    }
    __device__ float2 load(int tile, int line, int i) const {
        int g;
        const size_t idx = index(tile, line, i, &g);
        return op(src, idx);
    }
    __device__ void store(int tile, int line, int i, float2 v) const {
        int g;
        const size_t idx = index(tile, line, i, &g);
        if constexpr (TW != 0) {
            float2 w = tw[(g * i) << tw_shift];
            if (TW < 0) w.y = -w.y;
            v = cmul(v, w);
        }
        data[idx] = v;
    }
};

template <int TW, class Load>
void col_pass(hipStream_t s, const float2* tw, const float2* src, float2* data, int width, int npts, int gcount,
              int g_row_mul, int i_row_mul, int tw_log2n, const Load& op, bool inverse) {
    const int nl = col_tile_lines(npts);
    const int groups = nl / 16;
    ColIO<Load, TW> io{src, data, tw, width, width / 16, groups, g_row_mul, i_row_mul,
                       kMaxFftLog2 - tw_log2n, op};
    const int ntiles = (width / 16) * int(div_up(gcount, groups));
    // algorithmic bytes of one column pass: every element read and written once
    ProfileScope ps(std::string("fft-col-pass") + (inverse ? "-inv" : "-fwd"),
                    16.0 * double(width) * double(npts) * double(gcount));
    // Column tiles see lengths 8..128: single passes up to 64, split factors 8..128.
    if (inverse) launch_tiles<false, true, 3, 7>(s, npts, io, tw, ntiles);
    else launch_tiles<false, false, 3, 7>(s, npts, io, tw, ntiles);
}

// Forward column transform (in place). `op` applies to the loads of the first pass.
template <class Load>
void cols_forward(hipStream_t s, const float2* tw, float2* data, int width, int height, const Load& op) {
    if (width < 16 || (width & 15)) fail(LDH_ERR_INVALID_ARGUMENT, "fft-width-not-multiple-of-16");
    if (height < 8) fail(LDH_ERR_INVALID_ARGUMENT, "fft-height-below-8");
    if (height <= 64) {
        col_pass<0>(s, tw, data, data, width, height, 1, 0, 1, ilog2(height), op, false);
        return;
    }
    int n1, n2;
    col_split(height, &n1, &n2);
    col_pass<1>(s, tw, data, data, width, n1, n2, 1, n2, ilog2(height), op, false);
    col_pass<0>(s, tw, data, data, width, n2, n1, n2, 1, ilog2(height), PlainLoad{}, false);
}

// Inverse column transform (unscaled). The first pass loads from `src` through `op` and every
// pass stores into `data` (src == data for in place).
template <class Load>
void cols_inverse(hipStream_t s, const float2* tw, const float2* src, float2* data, int width, int height,
                  const Load& op) {
    if (width < 16 || (width & 15)) fail(LDH_ERR_INVALID_ARGUMENT, "fft-width-not-multiple-of-16");
    if (height < 8) fail(LDH_ERR_INVALID_ARGUMENT, "fft-height-below-8");
    if (height <= 64) {
        col_pass<0>(s, tw, src, data, width, height, 1, 0, 1, ilog2(height), op, true);
        return;
    }
    int n1, n2;
    col_split(height, &n1, &n2);
    col_pass<-1>(s, tw, src, data, width, n2, n1, n2, 1, ilog2(height), op, true);
    col_pass<0>(s, tw, data, data, width, n1, n2, 1, n2, ilog2(height), PlainLoad{}, true);
}

// ---- row passes ---------------------------------------------------------------------------------
// RowOp supplies load(row, x) and store(row, x, v); rows beyond `nrows` are skipped.
template <class RowOp>
struct RowIO {
    RowOp op;
    int lines_per_tile;
    int nrows;
    __device__ float2 load(int tile, int line, int i) const {
        const int row = tile * lines_per_tile + line;
        return row < nrows ? op.load(row, i) : make_float2(0.0f, 0.0f);
    }
    __device__ void store(int tile, int line, int i, float2 v) const {
        const int row = tile * lines_per_tile + line;
        if (row < nrows) op.store(row, i, v);
    }
};

template <class RowOp>
void rows_pass(hipStream_t s, const float2* tw, int width, int nrows, const RowOp& op, bool inverse) {
    const int nl = row_tile_lines(width);
    RowIO<RowOp> io{op, nl, nrows};
    const int ntiles = int(div_up(nrows, nl));
    ProfileScope ps(std::string("fft-row-pass") + (inverse ? "-inv" : "-fwd"), 16.0 * double(width) * double(nrows));
    if (inverse) launch_tiles<true, true, 4, kMaxFftLog2>(s, width, io, tw, ntiles);
    else launch_tiles<true, false, 4, kMaxFftLog2>(s, width, io, tw, ntiles);
}

struct PlainRows {
    float2* data;
    int pitch;
    __device__ float2 load(int row, int x) const { return data[size_t(row) * pitch + x]; }
    __device__ void store(int row, int x, float2 v) const { data[size_t(row) * pitch + x] = v; }
};

}  // namespace ldh

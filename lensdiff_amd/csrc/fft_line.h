// Single-line FFTs held in LDS for the whole transform (the PSF line kernels), planned as one
// composite first stage and S stages of radix B = 16 or 8: M = R0 * B^S with R0 <= 48
// (5120 = 20*16*16 = 10*8*8*8, 6144 = 24*16*16 = 12*8*8*8, 8192 = 32*16*16 = 16*8*8*8). Radix 16 means fewer passes over
// LDS; radix 8 means fewer registers per thread and more threads per line; which is faster is
// left to the autotuner. Stage inputs and outputs go through policies, so the first
// stage can read its (mostly zero) input straight from global memory and the last stage can
// write straight to global memory, and pointwise work is fused into whichever stage touches
// the data.
#pragma once

#include "fft_core.h"

namespace ldh {

// exp(-2 pi i m / R) for the composite radices: cos and sin of 2 pi m / R rounded to float.
template <int R> struct Roots;
template <> struct Roots<5> {
    static constexpr float c[5] = {1.000000000e+00f, 3.090169944e-01f, -8.090169944e-01f, -8.090169944e-01f, 3.090169944e-01f};
    static constexpr float s[5] = {0.000000000e+00f, 9.510565163e-01f, 5.877852523e-01f, -5.877852523e-01f, -9.510565163e-01f};
};
template <> struct Roots<10> {
    static constexpr float c[10] = {1.000000000e+00f, 8.090169944e-01f, 3.090169944e-01f, -3.090169944e-01f, -8.090169944e-01f, -1.000000000e+00f, -8.090169944e-01f, -3.090169944e-01f, 3.090169944e-01f, 8.090169944e-01f};
    static constexpr float s[10] = {0.000000000e+00f, 5.877852523e-01f, 9.510565163e-01f, 9.510565163e-01f, 5.877852523e-01f, 1.224646799e-16f, -5.877852523e-01f, -9.510565163e-01f, -9.510565163e-01f, -5.877852523e-01f};
};
template <> struct Roots<20> {
    static constexpr float c[20] = {1.000000000e+00f, 9.510565163e-01f, 8.090169944e-01f, 5.877852523e-01f, 3.090169944e-01f, 6.123233996e-17f, -3.090169944e-01f, -5.877852523e-01f, -8.090169944e-01f, -9.510565163e-01f, -1.000000000e+00f, -9.510565163e-01f, -8.090169944e-01f, -5.877852523e-01f, -3.090169944e-01f, -1.836970199e-16f, 3.090169944e-01f, 5.877852523e-01f, 8.090169944e-01f, 9.510565163e-01f};
    static constexpr float s[20] = {0.000000000e+00f, 3.090169944e-01f, 5.877852523e-01f, 8.090169944e-01f, 9.510565163e-01f, 1.000000000e+00f, 9.510565163e-01f, 8.090169944e-01f, 5.877852523e-01f, 3.090169944e-01f, 1.224646799e-16f, -3.090169944e-01f, -5.877852523e-01f, -8.090169944e-01f, -9.510565163e-01f, -1.000000000e+00f, -9.510565163e-01f, -8.090169944e-01f, -5.877852523e-01f, -3.090169944e-01f};
};
template <> struct Roots<40> {
    static constexpr float c[40] = {1.000000000e+00f, 9.876883406e-01f, 9.510565163e-01f, 8.910065242e-01f, 8.090169944e-01f, 7.071067812e-01f, 5.877852523e-01f, 4.539904997e-01f, 3.090169944e-01f, 1.564344650e-01f, 6.123233996e-17f, -1.564344650e-01f, -3.090169944e-01f, -4.539904997e-01f, -5.877852523e-01f, -7.071067812e-01f, -8.090169944e-01f, -8.910065242e-01f, -9.510565163e-01f, -9.876883406e-01f, -1.000000000e+00f, -9.876883406e-01f, -9.510565163e-01f, -8.910065242e-01f, -8.090169944e-01f, -7.071067812e-01f, -5.877852523e-01f, -4.539904997e-01f, -3.090169944e-01f, -1.564344650e-01f, -1.836970199e-16f, 1.564344650e-01f, 3.090169944e-01f, 4.539904997e-01f, 5.877852523e-01f, 7.071067812e-01f, 8.090169944e-01f, 8.910065242e-01f, 9.510565163e-01f, 9.876883406e-01f};
    static constexpr float s[40] = {0.000000000e+00f, 1.564344650e-01f, 3.090169944e-01f, 4.539904997e-01f, 5.877852523e-01f, 7.071067812e-01f, 8.090169944e-01f, 8.910065242e-01f, 9.510565163e-01f, 9.876883406e-01f, 1.000000000e+00f, 9.876883406e-01f, 9.510565163e-01f, 8.910065242e-01f, 8.090169944e-01f, 7.071067812e-01f, 5.877852523e-01f, 4.539904997e-01f, 3.090169944e-01f, 1.564344650e-01f, 1.224646799e-16f, -1.564344650e-01f, -3.090169944e-01f, -4.539904997e-01f, -5.877852523e-01f, -7.071067812e-01f, -8.090169944e-01f, -8.910065242e-01f, -9.510565163e-01f, -9.876883406e-01f, -1.000000000e+00f, -9.876883406e-01f, -9.510565163e-01f, -8.910065242e-01f, -8.090169944e-01f, -7.071067812e-01f, -5.877852523e-01f, -4.539904997e-01f, -3.090169944e-01f, -1.564344650e-01f};
};
template <> struct Roots<6> {
    static constexpr float c[6] = {1.000000000e+00f, 5.000000000e-01f, -5.000000000e-01f, -1.000000000e+00f, -5.000000000e-01f, 5.000000000e-01f};
    static constexpr float s[6] = {0.000000000e+00f, 8.660254038e-01f, 8.660254038e-01f, 1.224646799e-16f, -8.660254038e-01f, -8.660254038e-01f};
};
template <> struct Roots<9> {
    static constexpr float c[9] = {1.000000000e+00f, 7.660444431e-01f, 1.736481777e-01f, -5.000000000e-01f, -9.396926208e-01f, -9.396926208e-01f, -5.000000000e-01f, 1.736481777e-01f, 7.660444431e-01f};
    static constexpr float s[9] = {0.000000000e+00f, 6.427876097e-01f, 9.848077530e-01f, 8.660254038e-01f, 3.420201433e-01f, -3.420201433e-01f, -8.660254038e-01f, -9.848077530e-01f, -6.427876097e-01f};
};
template <> struct Roots<12> {
    static constexpr float c[12] = {1.000000000e+00f, 8.660254038e-01f, 5.000000000e-01f, 6.123233996e-17f, -5.000000000e-01f, -8.660254038e-01f, -1.000000000e+00f, -8.660254038e-01f, -5.000000000e-01f, -1.836970199e-16f, 5.000000000e-01f, 8.660254038e-01f};
    static constexpr float s[12] = {0.000000000e+00f, 5.000000000e-01f, 8.660254038e-01f, 1.000000000e+00f, 8.660254038e-01f, 5.000000000e-01f, 1.224646799e-16f, -5.000000000e-01f, -8.660254038e-01f, -1.000000000e+00f, -8.660254038e-01f, -5.000000000e-01f};
};
template <> struct Roots<18> {
    static constexpr float c[18] = {1.000000000e+00f, 9.396926208e-01f, 7.660444431e-01f, 5.000000000e-01f, 1.736481777e-01f, -1.736481777e-01f, -5.000000000e-01f, -7.660444431e-01f, -9.396926208e-01f, -1.000000000e+00f, -9.396926208e-01f, -7.660444431e-01f, -5.000000000e-01f, -1.736481777e-01f, 1.736481777e-01f, 5.000000000e-01f, 7.660444431e-01f, 9.396926208e-01f};
    static constexpr float s[18] = {0.000000000e+00f, 3.420201433e-01f, 6.427876097e-01f, 8.660254038e-01f, 9.848077530e-01f, 9.848077530e-01f, 8.660254038e-01f, 6.427876097e-01f, 3.420201433e-01f, 1.224646799e-16f, -3.420201433e-01f, -6.427876097e-01f, -8.660254038e-01f, -9.848077530e-01f, -9.848077530e-01f, -8.660254038e-01f, -6.427876097e-01f, -3.420201433e-01f};
};
template <> struct Roots<24> {
    static constexpr float c[24] = {1.000000000e+00f, 9.659258263e-01f, 8.660254038e-01f, 7.071067812e-01f, 5.000000000e-01f, 2.588190451e-01f, 6.123233996e-17f, -2.588190451e-01f, -5.000000000e-01f, -7.071067812e-01f, -8.660254038e-01f, -9.659258263e-01f, -1.000000000e+00f, -9.659258263e-01f, -8.660254038e-01f, -7.071067812e-01f, -5.000000000e-01f, -2.588190451e-01f, -1.836970199e-16f, 2.588190451e-01f, 5.000000000e-01f, 7.071067812e-01f, 8.660254038e-01f, 9.659258263e-01f};
    static constexpr float s[24] = {0.000000000e+00f, 2.588190451e-01f, 5.000000000e-01f, 7.071067812e-01f, 8.660254038e-01f, 9.659258263e-01f, 1.000000000e+00f, 9.659258263e-01f, 8.660254038e-01f, 7.071067812e-01f, 5.000000000e-01f, 2.588190451e-01f, 1.224646799e-16f, -2.588190451e-01f, -5.000000000e-01f, -7.071067812e-01f, -8.660254038e-01f, -9.659258263e-01f, -1.000000000e+00f, -9.659258263e-01f, -8.660254038e-01f, -7.071067812e-01f, -5.000000000e-01f, -2.588190451e-01f};
};
template <> struct Roots<32> {
    static constexpr float c[32] = {1.000000000e+00f, 9.807852804e-01f, 9.238795325e-01f, 8.314696123e-01f, 7.071067812e-01f, 5.555702330e-01f, 3.826834324e-01f, 1.950903220e-01f, 6.123233996e-17f, -1.950903220e-01f, -3.826834324e-01f, -5.555702330e-01f, -7.071067812e-01f, -8.314696123e-01f, -9.238795325e-01f, -9.807852804e-01f, -1.000000000e+00f, -9.807852804e-01f, -9.238795325e-01f, -8.314696123e-01f, -7.071067812e-01f, -5.555702330e-01f, -3.826834324e-01f, -1.950903220e-01f, -1.836970199e-16f, 1.950903220e-01f, 3.826834324e-01f, 5.555702330e-01f, 7.071067812e-01f, 8.314696123e-01f, 9.238795325e-01f, 9.807852804e-01f};
    static constexpr float s[32] = {0.000000000e+00f, 1.950903220e-01f, 3.826834324e-01f, 5.555702330e-01f, 7.071067812e-01f, 8.314696123e-01f, 9.238795325e-01f, 9.807852804e-01f, 1.000000000e+00f, 9.807852804e-01f, 9.238795325e-01f, 8.314696123e-01f, 7.071067812e-01f, 5.555702330e-01f, 3.826834324e-01f, 1.950903220e-01f, 1.224646799e-16f, -1.950903220e-01f, -3.826834324e-01f, -5.555702330e-01f, -7.071067812e-01f, -8.314696123e-01f, -9.238795325e-01f, -9.807852804e-01f, -1.000000000e+00f, -9.807852804e-01f, -9.238795325e-01f, -8.314696123e-01f, -7.071067812e-01f, -5.555702330e-01f, -3.826834324e-01f, -1.950903220e-01f};
};
template <> struct Roots<36> {
    static constexpr float c[36] = {1.000000000e+00f, 9.848077530e-01f, 9.396926208e-01f, 8.660254038e-01f, 7.660444431e-01f, 6.427876097e-01f, 5.000000000e-01f, 3.420201433e-01f, 1.736481777e-01f, 6.123233996e-17f, -1.736481777e-01f, -3.420201433e-01f, -5.000000000e-01f, -6.427876097e-01f, -7.660444431e-01f, -8.660254038e-01f, -9.396926208e-01f, -9.848077530e-01f, -1.000000000e+00f, -9.848077530e-01f, -9.396926208e-01f, -8.660254038e-01f, -7.660444431e-01f, -6.427876097e-01f, -5.000000000e-01f, -3.420201433e-01f, -1.736481777e-01f, -1.836970199e-16f, 1.736481777e-01f, 3.420201433e-01f, 5.000000000e-01f, 6.427876097e-01f, 7.660444431e-01f, 8.660254038e-01f, 9.396926208e-01f, 9.848077530e-01f};
    static constexpr float s[36] = {0.000000000e+00f, 1.736481777e-01f, 3.420201433e-01f, 5.000000000e-01f, 6.427876097e-01f, 7.660444431e-01f, 8.660254038e-01f, 9.396926208e-01f, 9.848077530e-01f, 1.000000000e+00f, 9.848077530e-01f, 9.396926208e-01f, 8.660254038e-01f, 7.660444431e-01f, 6.427876097e-01f, 5.000000000e-01f, 3.420201433e-01f, 1.736481777e-01f, 1.224646799e-16f, -1.736481777e-01f, -3.420201433e-01f, -5.000000000e-01f, -6.427876097e-01f, -7.660444431e-01f, -8.660254038e-01f, -9.396926208e-01f, -9.848077530e-01f, -1.000000000e+00f, -9.848077530e-01f, -9.396926208e-01f, -8.660254038e-01f, -7.660444431e-01f, -6.427876097e-01f, -5.000000000e-01f, -3.420201433e-01f, -1.736481777e-01f};
};
template <> struct Roots<48> {
    static constexpr float c[48] = {1.000000000e+00f, 9.914448614e-01f, 9.659258263e-01f, 9.238795325e-01f, 8.660254038e-01f, 7.933533403e-01f, 7.071067812e-01f, 6.087614290e-01f, 5.000000000e-01f, 3.826834324e-01f, 2.588190451e-01f, 1.305261922e-01f, 6.123233996e-17f, -1.305261922e-01f, -2.588190451e-01f, -3.826834324e-01f, -5.000000000e-01f, -6.087614290e-01f, -7.071067812e-01f, -7.933533403e-01f, -8.660254038e-01f, -9.238795325e-01f, -9.659258263e-01f, -9.914448614e-01f, -1.000000000e+00f, -9.914448614e-01f, -9.659258263e-01f, -9.238795325e-01f, -8.660254038e-01f, -7.933533403e-01f, -7.071067812e-01f, -6.087614290e-01f, -5.000000000e-01f, -3.826834324e-01f, -2.588190451e-01f, -1.305261922e-01f, -1.836970199e-16f, 1.305261922e-01f, 2.588190451e-01f, 3.826834324e-01f, 5.000000000e-01f, 6.087614290e-01f, 7.071067812e-01f, 7.933533403e-01f, 8.660254038e-01f, 9.238795325e-01f, 9.659258263e-01f, 9.914448614e-01f};
    static constexpr float s[48] = {0.000000000e+00f, 1.305261922e-01f, 2.588190451e-01f, 3.826834324e-01f, 5.000000000e-01f, 6.087614290e-01f, 7.071067812e-01f, 7.933533403e-01f, 8.660254038e-01f, 9.238795325e-01f, 9.659258263e-01f, 9.914448614e-01f, 1.000000000e+00f, 9.914448614e-01f, 9.659258263e-01f, 9.238795325e-01f, 8.660254038e-01f, 7.933533403e-01f, 7.071067812e-01f, 6.087614290e-01f, 5.000000000e-01f, 3.826834324e-01f, 2.588190451e-01f, 1.305261922e-01f, 1.224646799e-16f, -1.305261922e-01f, -2.588190451e-01f, -3.826834324e-01f, -5.000000000e-01f, -6.087614290e-01f, -7.071067812e-01f, -7.933533403e-01f, -8.660254038e-01f, -9.238795325e-01f, -9.659258263e-01f, -9.914448614e-01f, -1.000000000e+00f, -9.914448614e-01f, -9.659258263e-01f, -9.238795325e-01f, -8.660254038e-01f, -7.933533403e-01f, -7.071067812e-01f, -6.087614290e-01f, -5.000000000e-01f, -3.826834324e-01f, -2.588190451e-01f, -1.305261922e-01f};
};

template <int R, bool INV> __device__ __forceinline__ float2 root(int m) {
    const int i = m % R;
    return make_float2(Roots<R>::c[i], INV ? Roots<R>::s[i] : -Roots<R>::s[i]);
}

template <int R, bool INV> __device__ __forceinline__ void dft_any(float2* v);

// R = A * B in registers: n = B n1 + n2, k = k1 + A k2:
//   X[k1 + A k2] = sum_n2 W_B^(n2 k2) W_R^(n2 k1) sum_n1 v[B n1 + n2] W_A^(n1 k1)
template <int A, int B, bool INV> __device__ __forceinline__ void dft_ab(float2* v) {
#pragma clang fp contract(fast)
    constexpr int R = A * B;
    float2 y[R];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
        float2 a[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) a[n1] = v[B * n1 + n2];
        dft_any<A, INV>(a);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) y[n2 * A + k1] = (n2 == 0 || k1 == 0) ? a[k1] : cmul(a[k1], root<R, INV>(n2 * k1));
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        float2 b[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) b[n2] = y[n2 * A + k1];
        dft_any<B, INV>(b);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = b[k2];
    }
}

template <int R, bool INV> __device__ __forceinline__ void dft_any(float2* v) {
    if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 16) dft<R, INV>(v);
    else if constexpr (R == 6) dft_ab<2, 3, INV>(v);
    else if constexpr (R == 10) dft_ab<2, 5, INV>(v);
    else if constexpr (R == 20) dft_ab<4, 5, INV>(v);
    else if constexpr (R == 40) dft_ab<8, 5, INV>(v);
    else if constexpr (R == 9) dft_ab<3, 3, INV>(v);
    else if constexpr (R == 12) dft_ab<4, 3, INV>(v);
    else if constexpr (R == 18) dft_ab<2, 9, INV>(v);
    else if constexpr (R == 24) dft_ab<8, 3, INV>(v);
    else if constexpr (R == 32) dft_ab<2, 16, INV>(v);
    else if constexpr (R == 36) dft_ab<4, 9, INV>(v);
    else dft_ab<16, 3, INV>(v);  // 48
}

// The composite first stage of a line whose input is zero past `extent`: butterfly u reads
// positions u + r * (M / R0), so when extent <= NZ * (M / R0) only its first NZ inputs can be
// non-zero, and with Q = R0 / NZ
//     X[Q k1 + k2] = sum_{n < NZ} (x[n] W_R0^(n k2)) W_NZ^(n k1)
// (W_R0^Q = W_NZ): Q twiddled NZ-point DFTs replace the R0-point DFT. NZ is the smallest divisor
// of R0 with 5 NZ >= R0 (for R >= 4P a Bluestein line of M >= R + P - 1 points holds its P inputs
// in about the first fifth), for the radices with a root table.
constexpr bool has_roots(int r) {
    return r == 5 || r == 6 || r == 9 || r == 10 || r == 12 || r == 18 || r == 20 || r == 24 || r == 32 || r == 36 ||
           r == 40 || r == 48;
}
constexpr int prune_nz(int r0) {
    if (!has_roots(r0)) return r0;
    for (int d = 1; d < r0; ++d)
        if (r0 % d == 0 && 5 * d >= r0) return d;
    return r0;
}
template <int R0, int NZ, bool INV> __device__ __forceinline__ void dft_pruned(float2* v) {
#pragma clang fp contract(fast)
    constexpr int Q = R0 / NZ;
    static_assert(Q * NZ == R0, "NZ divides R0");
    float2 x[NZ];
#pragma unroll
    for (int n = 0; n < NZ; ++n) x[n] = v[n];
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) {
        float2 a[NZ];
#pragma unroll
        for (int n = 0; n < NZ; ++n) a[n] = (n == 0 || k2 == 0) ? x[n] : cmul(x[n], root<R0, INV>(n * k2));
        if constexpr (NZ > 1) dft_any<NZ, INV>(a);
#pragma unroll
        for (int k1 = 0; k1 < NZ; ++k1) v[Q * k1 + k2] = a[k1];
    }
}

// M = R0 * B^S with the fewest stages S >= 1 and 2 <= R0 <= 48 (S = 0 when M <= 48).
constexpr int line_stages(int M, int B) {
    int bs = 1;
    for (int s = 0; s <= 6; ++s) {
        if (M % bs == 0 && M / bs <= 48 && M / bs >= 2 && (s > 0 || M <= 48)) return s;
        bs *= B;
    }
    return -1;
}
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }
constexpr bool radix_ok(int r) {
    return r == 2 || r == 3 || r == 4 || r == 5 || r == 6 || r == 8 || r == 9 || r == 10 || r == 12 || r == 16 ||
           r == 18 || r == 20 || r == 24 || r == 32 || r == 36 || r == 40 || r == 48;
}
template <int M, int B = 16> struct LinePlan {
    static constexpr int S = line_stages(M, B);
    static constexpr int R0 = S >= 0 ? M / ipow(B, S) : 0;
    static constexpr bool ok = S >= 0 && radix_ok(R0);
    // threads: one butterfly each in the widest stage, at most 512 (radix 16) or 1024 (radix 8)
    static constexpr int bf = (S > 0 && M / B > M / R0) ? M / B : M / R0;
    static constexpr int cap = B == 16 ? 512 : 1024;
    static constexpr int NT = bf < 64 ? 64 : (bf > cap ? cap : (bf + 63) / 64 * 64);
};

// Stage policies. A stage reads elements base + r * STEP through in.get<STEP>(base, pbase, r)
// and writes them through out.put<STEP>(base, pbase, r, v), with pbase = lds_pad(base); `lds`
// tells whether a policy touches the line in LDS. Global policies only need operator().
// An output policy with `prefetch` multiplies each element by a per-element constant from global
// memory: the stage loads the constants (pre<STEP>(base, r)) before its butterflies, so their
// latency overlaps the arithmetic, and writes with put_pre<STEP>(base, pbase, r, v, c).
template <class F> struct GlobalIn {
    static constexpr bool lds = false;
    F f;
    template <int STEP> __device__ float2 get(int base, int, int r) const { return f(base + r * STEP); }
};
template <class F> struct GlobalOut {
    static constexpr bool lds = false;
    static constexpr bool prefetch = F::kPrefetch;
    F f;
    template <int STEP> __device__ void put(int base, int, int r, float2 v) const { f(base + r * STEP, v); }
    template <int STEP> __device__ float2 pre(int base, int r) const { return f.pre(base + r * STEP); }
    template <int STEP> __device__ void put_pre(int base, int, int r, float2 v, float2 c) const {
        f.put_pre(base + r * STEP, v, c);
    }
};
struct LdsIn {
    static constexpr bool lds = true;
    float2* line;
    template <int STEP> __device__ float2 get(int base, int pbase, int r) const {
        return line[lds_pad_step<STEP>(base, pbase, r)];
    }
};
struct LdsOut {
    static constexpr bool lds = true;
    static constexpr bool prefetch = false;
    float2* line;
    template <int STEP> __device__ void put(int base, int pbase, int r, float2 v) const {
        line[lds_pad_step<STEP>(base, pbase, r)] = v;
    }
};
__host__ __device__ constexpr int top_bit(int r) { return r >= 16 ? 16 : r >= 8 ? 8 : r >= 4 ? 4 : r >= 2 ? 2 : 1; }

// W^k (forward sign) of each radix-B stage's butterfly for thread tid, k = tid % Ns: loaded
// together before the first transform (and reused by a second one), so no stage waits for its
// twiddle. Used where every thread does one butterfly per stage.
template <int M, int B> struct StageTwiddles {
    static constexpr int S = LinePlan<M, B>::S;
    float2 w[S > 0 ? S : 1];
};
template <int M, int B>
__device__ __forceinline__ StageTwiddles<M, B> stage_twiddles(const float2* __restrict__ tw, int tid) {
    StageTwiddles<M, B> st;
    const float2* __restrict__ t = tw + tw_offset(M);
    int ns = LinePlan<M, B>::R0;
#pragma unroll
    for (int s = 0; s < StageTwiddles<M, B>::S; ++s) {
        st.w[s] = t[(tid % ns) * (tw_table(M) / (ns * B))];
        ns *= B;
    }
    return st;
}

// One Stockham stage over a single line of N points: sub-length Ns, radix R. wk: the stage's
// W^k when PRE (one butterfly per thread), else loaded here.
// NZ < R (first stage only): inputs r >= NZ are zero and not read.
template <int N, int Ns, int R, int NT, bool INV, bool PRE = false, int NZ = R, class In, class Out>
__device__ __forceinline__ void line_stage(const float2* __restrict__ tw, int tid, const In& in, const Out& out,
                                           float2 wk = make_float2(1.0f, 0.0f)) {
    static_assert(NZ == R || Ns == 1, "only the first stage is pruned");
    constexpr int BF = N / R;
    constexpr int PER = (BF + NT - 1) / NT;
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int u = tid + p * NT;
        if (BF % NT == 0 || u < BF) {
            const int pu = lds_pad(u);
#pragma unroll
            for (int r = 0; r < NZ; ++r) v[p][r] = in.template get<BF>(u, pu, r);
        }
    }
    if constexpr (In::lds && Out::lds) __syncthreads();  // every read before the in-place writes
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int u = tid + p * NT;
        if (BF % NT == 0 || u < BF) {
            const int k = u % Ns;
            const int base = (u - k) * R + k;
            float2 c[Out::prefetch ? R : 1];
            if constexpr (Out::prefetch) {
#pragma unroll
                for (int r = 0; r < R; ++r) c[r] = out.template pre<Ns>(base, r);
            }
            if constexpr (Ns > 1) {
                constexpr int stride = tw_table(N) / (Ns * R);
                const float2* __restrict__ t = tw + tw_offset(N);
                static_assert((R & (R - 1)) == 0, "tail stages are radix 2^b");
                // W^k from the table (one load), W^(2^b k) by squaring, the other powers as
                // products W^(k r) = W^(k (r - h)) W^(k h), h the top bit of r: at most six
                // rounded products per twiddle (loads, not arithmetic, limit these kernels).
                float2 w[R];
                if constexpr (PRE && PER == 1) w[1] = wk;
                else w[1] = t[k * stride];
                if (INV) w[1].y = -w[1].y;
#pragma unroll
                for (int h = 2; h < R; h *= 2) w[h] = cmul(w[h / 2], w[h / 2]);
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const int h = top_bit(r);
                    if (r != h) w[r] = cmul(w[r - h], w[h]);
                    v[p][r] = cmul(v[p][r], w[r]);
                }
            }
            if constexpr (NZ < R) dft_pruned<R, NZ, INV>(v[p]);
            else dft_any<R, INV>(v[p]);
            const int pbase = lds_pad(base);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (Out::prefetch) out.template put_pre<Ns>(base, pbase, r, v[p][r], c[r]);
                else out.template put<Ns>(base, pbase, r, v[p][r]);
            }
        }
    }
    if constexpr (Out::lds) __syncthreads();
}

// Radix-B stages from sub-length Ns (stage SI of the tail) to M; the last writes through `out`.
template <int M, int B, int Ns, int SI, int NT, bool INV, class Out>
__device__ __forceinline__ void line_tail(const float2* __restrict__ tw, int tid, const LdsIn& li, const LdsOut& lo,
                                          const Out& out, const StageTwiddles<M, B>& st) {
    if constexpr (Ns * B == M) {
        line_stage<M, Ns, B, NT, INV, true>(tw, tid, li, out, st.w[SI]);
    } else {
        line_stage<M, Ns, B, NT, INV, true>(tw, tid, li, lo, st.w[SI]);
        line_tail<M, B, Ns * B, SI + 1, NT, INV>(tw, tid, li, lo, out, st);
    }
}

// line[i] *= f[i] over the M points of a line in LDS (padded), NT threads: every thread's
// loads are issued before its multiplies (a pass of its own, when the transforms' registers are
// free). Ends with a barrier.
template <int M, int NT>
__device__ __forceinline__ void line_multiply(float2* line, const float2* __restrict__ f, int tid) {
    constexpr int IT = (M + NT - 1) / NT;
    float2 c[IT], v[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int i = tid + it * NT;
        if (M % NT == 0 || i < M) {
            c[it] = f[i];
            v[it] = line[lds_pad(i)];
        }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int i = tid + it * NT;
        if (M % NT == 0 || i < M) line[lds_pad(i)] = cmul(v[it], c[it]);
    }
    __syncthreads();
}

// The whole line transform: stage 0 reads through `in`, the last stage writes through `out`,
// the stages between keep the line in LDS (`line`, padded). st: stage_twiddles<M, B>(tw, tid).
// extent: the input is zero at positions >= extent (uniform over the workgroup).
template <int M, int B, bool INV, class In, class Out>
__device__ __forceinline__ void line_fft(float2* line, const float2* __restrict__ tw, int tid, const In& in,
                                         const Out& out, const StageTwiddles<M, B>& st, int extent = M) {
    using P = LinePlan<M, B>;
    static_assert(P::ok, "line length has no composite-first plan");
    constexpr int NT = P::NT, R0 = P::R0, S = P::S;
    const LdsIn li{line};
    const LdsOut lo{line};
    if constexpr (S == 0) {
        line_stage<M, 1, R0, NT, INV>(tw, tid, in, out);
    } else {
        constexpr int NZ = prune_nz(R0);
        if constexpr (NZ < R0) {
            if (extent <= NZ * (M / R0)) line_stage<M, 1, R0, NT, INV, false, NZ>(tw, tid, in, lo);
            else line_stage<M, 1, R0, NT, INV>(tw, tid, in, lo);
        } else {
            line_stage<M, 1, R0, NT, INV>(tw, tid, in, lo);
        }
        line_tail<M, B, R0, 0, NT, INV>(tw, tid, li, lo, out, st);
    }
}

}  // namespace ldh

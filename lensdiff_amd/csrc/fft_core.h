// Device-side FFT building blocks for gfx950: register DFTs of radix 2/4/8/16, Stockham stages
// over a tile held in LDS, and a tile kernel whose global loads and stores are supplied by an
// IO policy (so pointwise work is fused into the first and last pass).
//
// A tile holds NL lines of N points. Element (line, i) lives at LDS index line*LS + i*IS,
// padded by one float2 every 32 entries to keep the first stage's strided writes spread
// across LDS banks. Row tiles use (LS, IS) = (N, 1); column tiles (lines adjacent in global
// memory) use (1, NL).
#pragma once

#include <type_traits>

#include "ldh_internal.h"

namespace ldh {

__device__ __forceinline__ int lds_pad(int x) { return x + (x >> 5); }

// Padded position of element base + r * STEP, given pbase = lds_pad(base). When STEP is a
// multiple of 32 the padding is linear in r, so the offsets fold into the LDS instructions'
// immediate field instead of costing address arithmetic per access.
template <int STEP> __device__ __forceinline__ int lds_pad_step(int base, int pbase, int r) {
    if constexpr (STEP % 32 == 0) return pbase + r * (STEP + STEP / 32);
    else return lds_pad(base + r * STEP);
}

// multiply by -i (forward) or +i (inverse)
template <bool INV> __device__ __forceinline__ float2 rot90(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV> __device__ __forceinline__ void dft2(float2& a, float2& b) {
    const float2 t = a;
    a = make_float2(t.x + b.x, t.y + b.y);
    b = make_float2(t.x - b.x, t.y - b.y);
}

template <bool INV> __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    const float2 t0 = make_float2(a0.x + a2.x, a0.y + a2.y);
    const float2 t1 = make_float2(a0.x - a2.x, a0.y - a2.y);
    const float2 t2 = make_float2(a1.x + a3.x, a1.y + a3.y);
    const float2 t3 = rot90<INV>(make_float2(a1.x - a3.x, a1.y - a3.y));
    a0 = make_float2(t0.x + t2.x, t0.y + t2.y);
    a2 = make_float2(t0.x - t2.x, t0.y - t2.y);
    a1 = make_float2(t1.x + t3.x, t1.y + t3.y);
    a3 = make_float2(t1.x - t3.x, t1.y - t3.y);
}

// exp(-+2 pi i j / 16) for j = 0..15 (cos, sin of 2 pi j / 16)
__device__ __forceinline__ float2 w16(int j, bool inv) {
    constexpr float c[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                             0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                             -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
                             0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};
    const float cs = c[j & 15];
    const float sn = c[(j + 12) & 15];  // sin(2 pi j/16) = cos(2 pi (j-4)/16)
    return make_float2(cs, inv ? sn : -sn);
}

template <bool INV> __device__ __forceinline__ void dft8(float2* v) {
    // even/odd split: E = DFT4(v0,v2,v4,v6), O = DFT4(v1,v3,v5,v7)
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    const float h = 0.70710678118654752f;
    // W8^1, W8^2, W8^3 applied to o1, o2, o3
    const float2 t1 = INV ? make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y))
                          : make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
    const float2 t2 = rot90<INV>(o2);
    const float2 t3 = INV ? make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y))
                          : make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
    v[0] = make_float2(e0.x + o0.x, e0.y + o0.y);
    v[4] = make_float2(e0.x - o0.x, e0.y - o0.y);
    v[1] = make_float2(e1.x + t1.x, e1.y + t1.y);
    v[5] = make_float2(e1.x - t1.x, e1.y - t1.y);
    v[2] = make_float2(e2.x + t2.x, e2.y + t2.y);
    v[6] = make_float2(e2.x - t2.x, e2.y - t2.y);
    v[3] = make_float2(e3.x + t3.x, e3.y + t3.y);
    v[7] = make_float2(e3.x - t3.x, e3.y - t3.y);
}

template <bool INV> __device__ __forceinline__ void dft16(float2* v) {
    // X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 v[4 n1 + n2] W4^(n1 k1)
    float2 y[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        float2 a0 = v[n2], a1 = v[4 + n2], a2 = v[8 + n2], a3 = v[12 + n2];
        dft4<INV>(a0, a1, a2, a3);
        y[n2 * 4 + 0] = a0;
        y[n2 * 4 + 1] = n2 ? cmul(a1, w16(n2 * 1, INV)) : a1;
        y[n2 * 4 + 2] = n2 ? cmul(a2, w16(n2 * 2, INV)) : a2;
        y[n2 * 4 + 3] = n2 ? cmul(a3, w16(n2 * 3, INV)) : a3;
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 b0 = y[0 * 4 + k1], b1 = y[1 * 4 + k1], b2 = y[2 * 4 + k1], b3 = y[3 * 4 + k1];
        dft4<INV>(b0, b1, b2, b3);
        v[k1 + 0] = b0;
        v[k1 + 4] = b1;
        v[k1 + 8] = b2;
        v[k1 + 12] = b3;
    }
}

template <bool INV> __device__ __forceinline__ void dft3(float2& a0, float2& a1, float2& a2) {
#pragma clang fp contract(fast)
    const float s = 0.86602540378443865f;  // sin(2 pi / 3)
    const float2 t = make_float2(a1.x + a2.x, a1.y + a2.y);
    const float2 m = make_float2(a0.x - 0.5f * t.x, a0.y - 0.5f * t.y);
    const float2 d = rot90<INV>(make_float2(a1.x - a2.x, a1.y - a2.y));
    a0 = make_float2(a0.x + t.x, a0.y + t.y);
    a1 = make_float2(m.x + s * d.x, m.y + s * d.y);
    a2 = make_float2(m.x - s * d.x, m.y - s * d.y);
}

// 5 points: X0 = a0 + t1 + t2, X1,4 = b1 -+ i d1, X2,3 = b2 -+ i d2 (forward) with
// t1 = a1 + a4, t2 = a2 + a3, b1 = a0 + c1 t1 + c2 t2, b2 = a0 + c2 t1 + c1 t2,
// d1 = s1 (a1 - a4) + s2 (a2 - a3), d2 = s2 (a1 - a4) - s1 (a2 - a3), ck = cos(2 pi k / 5),
// sk = sin(2 pi k / 5).
template <bool INV>
__device__ __forceinline__ void dft5(float2& a0, float2& a1, float2& a2, float2& a3, float2& a4) {
#pragma clang fp contract(fast)
    const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
    const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
    const float2 t1 = make_float2(a1.x + a4.x, a1.y + a4.y), t2 = make_float2(a2.x + a3.x, a2.y + a3.y);
    const float2 t3 = make_float2(a1.x - a4.x, a1.y - a4.y), t4 = make_float2(a2.x - a3.x, a2.y - a3.y);
    const float2 b1 = make_float2(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
    const float2 b2 = make_float2(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
    const float2 d1 = rot90<INV>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    const float2 d2 = rot90<INV>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    a0 = make_float2(a0.x + t1.x + t2.x, a0.y + t1.y + t2.y);
    a1 = make_float2(b1.x + d1.x, b1.y + d1.y);
    a4 = make_float2(b1.x - d1.x, b1.y - d1.y);
    a2 = make_float2(b2.x + d2.x, b2.y + d2.y);
    a3 = make_float2(b2.x - d2.x, b2.y - d2.y);
}

template <int R, bool INV> __device__ __forceinline__ void dft(float2* v) {
    if constexpr (R == 2) dft2<INV>(v[0], v[1]);
    else if constexpr (R == 3) dft3<INV>(v[0], v[1], v[2]);
    else if constexpr (R == 5) dft5<INV>(v[0], v[1], v[2], v[3], v[4]);
    else if constexpr (R == 4) dft4<INV>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 8) dft8<INV>(v);
    else dft16<INV>(v);
}

// Radix of the stage that starts at sub-length Ns: the largest power of two up to 16 that
// divides what remains, then 3 for the odd factors (lengths 2^a * 3^b).
template <int N, int Ns> struct StageRadix {
    static constexpr int rem = N / Ns;
    static constexpr int p2 = rem & -rem;
    static constexpr int value = p2 >= 16 ? 16 : (p2 >= 2 ? p2 : 3);
};

// One Stockham stage over all NL lines of the tile, executed by NT threads.
// twl (TWL): the N twiddles W_N^j, j < N, in LDS; else the stage reads the global table.
template <int N, int Ns, int NL, int NT, bool INV, int LS, int IS, bool TWL>
__device__ __forceinline__ void fft_stage(float2* lds, const float2* __restrict__ tw, const float2* twl, int tid) {
    constexpr int R = StageRadix<N, Ns>::value;
    constexpr int BF = N / R;
    constexpr int TOTAL = NL * BF;
    constexpr int PER = (TOTAL + NT - 1) / NT;
    constexpr bool kPow2 = (R & (R - 1)) == 0;
    constexpr int stride = tw_table(N) / (Ns * R);
    constexpr int q = N / (Ns * R);  // W_(Ns R)^k = W_N^(k q)
    // one table twiddle per butterfly and its powers by squaring and products, except for an
    // odd radix with the table in LDS (every power read from the table)
    constexpr bool kOne = kPow2 || !TWL;
    const float2* __restrict__ t = tw + tw_offset(N);
    float2 v[PER][R];
    int line_of[PER], j_of[PER];
    // the stage's twiddles are read first, so they arrive while the LDS reads and the barrier
    // complete
    float2 wt[PER][Ns > 1 ? (kOne ? 1 : R - 1) : 1];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int u = tid + p * NT;
        int line, j;
        if constexpr (IS == 1) { line = u / BF; j = u % BF; }
        else { line = u % NL; j = u / NL; }
        line_of[p] = line;
        j_of[p] = j;
        if constexpr (Ns > 1) {
            // every lane loads (a lane past TOTAL reads entry 0), so the loads issue together
            const int k = (TOTAL % NT == 0 || u < TOTAL) ? j % Ns : 0;
            if constexpr (kOne) wt[p][0] = TWL ? twl[k * q] : t[k * stride];
            else {
#pragma unroll
                for (int r = 1; r < R; ++r) wt[p][r - 1] = twl[k * r * q];
            }
        }
    }
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int u = tid + p * NT;
        const int line = line_of[p], j = j_of[p];
        if (TOTAL % NT == 0 || u < TOTAL) {
            const int b = line * LS + j * IS;
            const int pb = lds_pad(b);
#pragma unroll
            for (int r = 0; r < R; ++r) v[p][r] = lds[lds_pad_step<BF * IS>(b, pb, r)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int u = tid + p * NT;
        if (TOTAL % NT == 0 || u < TOTAL) {
            const int j = j_of[p];
            const int k = j % Ns;
            if constexpr (Ns > 1) {
                if constexpr (kOne) {
                    // W^k from the table, W^(2^b k) by squaring, the other powers as products
                    // W^(k (r - h)) W^(k h), h the top bit of r
                    float2 w[R];
                    w[1] = wt[p][0];
                    if (INV) w[1].y = -w[1].y;
#pragma unroll
                    for (int h = 2; h < R; h *= 2) w[h] = cmul(w[h / 2], w[h / 2]);
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        const int h = r >= 8 ? 8 : r >= 4 ? 4 : r >= 2 ? 2 : 1;
                        if (r != h) w[r] = cmul(w[r - h], w[h]);
                        v[p][r] = cmul(v[p][r], w[r]);
                    }
                } else {
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        float2 w = wt[p][r - 1];
                        if (INV) w.y = -w.y;
                        v[p][r] = cmul(v[p][r], w);
                    }
                }
            }
            dft<R, INV>(v[p]);
            const int b = line_of[p] * LS + ((j - k) * R + k) * IS;
            const int pb = lds_pad(b);
#pragma unroll
            for (int r = 0; r < R; ++r) lds[lds_pad_step<Ns * IS>(b, pb, r)] = v[p][r];
        }
    }
    __syncthreads();
}

template <int N, int Ns, int NL, int NT, bool INV, int LS, int IS, bool TWL>
__device__ __forceinline__ void fft_stages(float2* lds, const float2* __restrict__ tw, const float2* twl, int tid) {
    if constexpr (Ns < N) {
        fft_stage<N, Ns, NL, NT, INV, LS, IS, TWL>(lds, tw, twl, tid);
        fft_stages<N, Ns * StageRadix<N, Ns>::value, NL, NT, INV, LS, IS, TWL>(lds, tw, twl, tid);
    }
}

// Adjacent columns per column tile (kColWidth * 8 bytes contiguous per row).
constexpr int kColWidth = 16;

// Tile geometry for a line length and layout.
template <int N, bool ROWS> struct TileShape {
    // rows: at least 1024 elements per tile; columns: kColWidth adjacent columns times enough
    // line groups for 1024 elements.
    static constexpr int NL = ROWS ? (N >= 1024 ? 1 : 1024 / N) : (kColWidth * (N >= 64 ? 1 : 64 / N));
    static constexpr int ELEMS = NL * N;
    static constexpr int NT0 = ELEMS / 8;  // 8 elements per thread (more threads hide the loads)
    static constexpr int NT1 = NT0 < 64 ? 64 : (NT0 > 512 ? 512 : NT0);
    static constexpr int NT = (NT1 + 63) / 64 * 64;
    static constexpr int DATA_WORDS = ELEMS + ELEMS / 32 + 1;
    // column tiles whose N has a factor 3 (radix-3 stages read two twiddles per butterfly) keep
    // the N twiddles W_N^j in LDS after the data; power-of-two tiles read one per butterfly from
    // the global table
    static constexpr bool TWL = !ROWS && (N & (N - 1)) != 0;
    static constexpr int LDS_BYTES = (DATA_WORDS + (TWL ? N : 0)) * 8;
};

// Maps the hardware workgroup id to a logical tile so that consecutive logical tiles run on
// the same XCD (workgroups are dealt round-robin over the 8 XCDs).
__device__ __forceinline__ int xcd_tile(int bid, int nblocks) {
    if ((nblocks & 7) != 0) return bid;
    return (bid & 7) * (nblocks >> 3) + (bid >> 3);
}

// An IO may store a tile's finished lines itself, reading any positions of the transform from
// LDS (IO::kStoreLds and IO::store_lds<N, NL, NT>(tile, tid, lds); element i of line l is at
// lds[lds_pad(l * N + i)] for row tiles).
// Loads a thread issues before its first finish (IO::kFetchBatch, default 16).
template <class T, class = void>
struct io_fetch_batch : std::integral_constant<int, 16> {};
template <class T>
struct io_fetch_batch<T, std::void_t<decltype(T::kFetchBatch)>> : std::integral_constant<int, T::kFetchBatch> {};

template <class T, class = void>
struct io_stores_lds : std::false_type {};
template <class T>
struct io_stores_lds<T, std::void_t<decltype(T::kStoreLds)>> : std::integral_constant<bool, T::kStoreLds> {};

// The tile kernel. IO supplies:
//   __device__ float2 load(int tile, int line, int i) const;
//   __device__ void store(int tile, int line, int i, float2 v) const;
template <int N, bool ROWS, bool INV, class IO>
__global__ void __launch_bounds__((TileShape<N, ROWS>::NT))
fft_tile_kernel(IO io, const float2* __restrict__ tw, int ntiles) {
    using S = TileShape<N, ROWS>;
    constexpr int NL = S::NL, NT = S::NT, E = S::ELEMS;
    constexpr int LS = ROWS ? N : 1;
    constexpr int IS = ROWS ? 1 : NL;
    extern __shared__ float2 lds[];
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int tid = threadIdx.x;
    // the column tile's N twiddles: read first (one per thread), stored to LDS with the data
    static_assert(!S::TWL || N <= NT, "one twiddle per thread");
    float2 tw_mine = make_float2(1.0f, 0.0f);
    if constexpr (S::TWL) {
        if (tid < N) tw_mine = tw[tw_offset(N) + tid * (tw_table(N) / N)];
    }
    // Column tiles move adjacent column pairs (16 bytes per lane and access): pair p of the tile
    // is columns 2 (p % NP), 2 (p % NP) + 1 of point p / NP. Needs every thread to own the same
    // number of pairs and keep its column pair (NT a multiple of NP).
    constexpr int NP = NL / 2;
    constexpr bool kPairs = !ROWS && IO::kPairs && (E / 2) % NT == 0 && NT % NP == 0;
    constexpr int IT = (E + NT - 1) / NT;
    constexpr int IP = kPairs ? E / 2 / NT : 1;  // pairs per thread
    if constexpr (kPairs) {
        const int lp = (tid % NP) * 2, i0 = tid / NP;
        float4 t[IP];
#pragma unroll
        for (int c = 0; c < IP; ++c) t[c] = io.fetch_pair(tile, lp, i0 + c * (NT / NP));
#pragma unroll
        for (int c = 0; c < IP; ++c) {
            const int at = lp + (i0 + c * (NT / NP)) * NL;
            lds[lds_pad(at)] = make_float2(t[c].x, t[c].y);
            lds[lds_pad(at + 1)] = make_float2(t[c].z, t[c].w);
        }
    } else {
        // loads in batches of up to kBatch per thread: every fetch of a batch (memory only) is issued
        // before the first finish (the arithmetic, which may branch on launch-uniform settings) and
        // the LDS writes
        constexpr int kBatch = IT < io_fetch_batch<IO>::value ? IT : io_fetch_batch<IO>::value;
#pragma unroll
        for (int c0 = 0; c0 < IT; c0 += kBatch) {
            typename IO::Raw t[kBatch];
#pragma unroll
            for (int c = 0; c < kBatch; ++c) {
                const int e = tid + (c0 + c) * NT;
                if ((E % NT == 0 && IT % kBatch == 0) || (c0 + c < IT && e < E)) {
                    int line, i;
                    if constexpr (ROWS) { line = e / N; i = e % N; }
                    else { line = e % NL; i = e / NL; }
                    t[c] = io.fetch(tile, line, i);
                }
            }
#pragma unroll
            for (int c = 0; c < kBatch; ++c) {
                const int e = tid + (c0 + c) * NT;
                if ((E % NT == 0 && IT % kBatch == 0) || (c0 + c < IT && e < E)) {
                    int line, i;
                    if constexpr (ROWS) { line = e / N; i = e % N; }
                    else { line = e % NL; i = e / NL; }
                    lds[lds_pad(line * LS + i * IS)] = io.finish(t[c]);
                }
            }
        }
    }
    float2* twl = lds + S::DATA_WORDS;
    if constexpr (S::TWL) {
        if (tid < N) twl[tid] = tw_mine;
    }
    __syncthreads();
    // store-side twiddles loaded before the stages, so their latency overlaps the transform
    float2 w[IO::kStoreTwiddle ? IT : 1];
    if constexpr (IO::kStoreTwiddle && kPairs) {
        // a thread's pairs keep their columns (so g) and step i by D = NT / NP
        constexpr int D = NT / NP;
        const int lp = (tid % NP) * 2, i0 = tid / NP;
        const float2 base = io.twiddle(tile, lp, i0);
        float2 pw[IP];
        pw[0] = make_float2(1.0f, 0.0f);
        if constexpr (IP > 1) pw[1] = io.twiddle(tile, lp, D);
#pragma unroll
        for (int c = 2; c < IP; ++c) {
            const int h = c >= 8 ? 8 : c >= 4 ? 4 : c >= 2 ? 2 : 1;
            pw[c] = c == h ? cmul(pw[h / 2], pw[h / 2]) : cmul(pw[c - h], pw[h]);
        }
#pragma unroll
        for (int c = 0; c < IP; ++c) w[c] = c == 0 ? base : cmul(base, pw[c]);
    } else if constexpr (IO::kStoreTwiddle) {
        if constexpr (!ROWS && NT % NL == 0) {
            // a thread's points share their line (so g) and step i by D = NT / NL: W^(g i) for
            // i = i0 + c D is W^(g i0) times (W^(g D))^c, two table reads and products
            constexpr int D = NT / NL;
            const int line = tid % NL, i0 = tid / NL;
            const float2 base = io.twiddle(tile, line, i0);
            float2 pw[IT];
            pw[0] = make_float2(1.0f, 0.0f);
            if constexpr (IT > 1) pw[1] = io.twiddle(tile, line, D);
#pragma unroll
            for (int c = 2; c < IT; ++c) {
                const int h = c >= 8 ? 8 : c >= 4 ? 4 : c >= 2 ? 2 : 1;
                pw[c] = c == h ? cmul(pw[h / 2], pw[h / 2]) : cmul(pw[c - h], pw[h]);
            }
#pragma unroll
            for (int c = 0; c < IT; ++c) w[c] = c == 0 ? base : cmul(base, pw[c]);
        } else {
#pragma unroll
            for (int c = 0; c < IT; ++c) {
                const int e = tid + c * NT;
                if (E % NT == 0 || e < E) {
                    int line, i;
                    if constexpr (ROWS) { line = e / N; i = e % N; }
                    else { line = e % NL; i = e / NL; }
                    w[c] = io.twiddle(tile, line, i);
                }
            }
        }
    }
    fft_stages<N, 1, NL, NT, INV, LS, IS, S::TWL>(lds, tw, twl, tid);
    if constexpr (io_stores_lds<IO>::value) {
        io.template store_lds<N, NL, NT>(tile, tid, lds);
    } else if constexpr (kPairs) {
        const int lp = (tid % NP) * 2, i0 = tid / NP;
#pragma unroll
        for (int c = 0; c < IP; ++c) {
            const int i = i0 + c * (NT / NP);
            const int at = lp + i * NL;
            io.store_pair(tile, lp, i, lds[lds_pad(at)], lds[lds_pad(at + 1)], w[IO::kStoreTwiddle ? c : 0]);
        }
    } else {
#pragma unroll
        for (int c = 0; c < IT; ++c) {
            const int e = tid + c * NT;
            if (E % NT == 0 || e < E) {
                int line, i;
                if constexpr (ROWS) { line = e / N; i = e % N; }
                else { line = e % NL; i = e / NL; }
                io.store(tile, line, i, lds[lds_pad(line * LS + i * IS)], w[IO::kStoreTwiddle ? c : 0]);
            }
        }
    }
}

// Kernels whose dynamic LDS exceeds 64 KiB must opt in before launch (gfx950 has 160 KiB/CU).
inline void allow_large_lds(const void* fn, int bytes) {
    if (bytes <= 65536) return;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
        fail(LDH_ERR_DEVICE, "lds-attribute-failed");
}

template <bool ROWS, bool INV, int N, class IO>
bool launch_tile_if(hipStream_t s, int n, const IO& io, const float2* tw, int ntiles) {
    if (n != N) return false;
    using S = TileShape<N, ROWS>;
    allow_large_lds(reinterpret_cast<const void*>(&fft_tile_kernel<N, ROWS, INV, IO>), S::LDS_BYTES);
    hipLaunchKernelGGL((fft_tile_kernel<N, ROWS, INV, IO>), dim3(ntiles), dim3(S::NT), S::LDS_BYTES, s, io, tw,
                       ntiles);
    return true;
}

// Line lengths the tile kernels are built for. Rows: every grid length from 16. Columns: the
// single-pass heights and split factors (col_split).
template <bool ROWS, bool INV, class IO>
void launch_tiles(hipStream_t s, int n, const IO& io, const float2* tw, int ntiles) {
    bool ok;
    if constexpr (ROWS) {
        ok = launch_tile_if<ROWS, INV, 16>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 32>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 48>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 64>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 96>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 128>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 144>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 192>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 256>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 288>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 384>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 512>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 576>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 768>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 1024>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 1152>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 1536>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 2048>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 2304>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 3072>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 4096>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 4608>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 6144>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 8192>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 9216>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 12288>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 16384>(s, n, io, tw, ntiles);
    } else {
        ok = launch_tile_if<ROWS, INV, 8>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 12>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 16>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 24>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 32>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 36>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 48>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 64>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 72>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 96>(s, n, io, tw, ntiles) ||
             launch_tile_if<ROWS, INV, 128>(s, n, io, tw, ntiles) || launch_tile_if<ROWS, INV, 144>(s, n, io, tw, ntiles);
    }
    if (!ok) fail(LDH_ERR_TOO_LARGE, "fft-size-unsupported");
    LDH_CHECK_LAUNCH("fft-tile-launch");
}

inline int row_tile_lines(int n) { return n >= 1024 ? 1 : 1024 / n; }
inline int col_tile_lines(int n) { return kColWidth * (n >= 64 ? 1 : 64 / n); }

// Column tile lengths: one pass for heights up to 144 in this set, else two factors from it.
__host__ __device__ constexpr bool col_tile_length_ok(int n) {
    return n == 8 || n == 12 || n == 16 || n == 24 || n == 32 || n == 36 || n == 48 || n == 64 || n == 72 ||
           n == 96 || n == 128 || n == 144;
}

// The balanced split of a column length into N1 x N2 (both column tile lengths, N1 <= N2) for
// the two-pass column transform, n1 = 0 when the height takes a single pass: the split a height
// keeps when it is not tuned (col_plan).
__host__ __device__ inline void col_split(int n, int* n1, int* n2) {
    *n1 = 0;
    *n2 = n;
    if (col_tile_length_ok(n)) return;
    int best = 0;
    const int cand[12] = {8, 12, 16, 24, 32, 36, 48, 64, 72, 96, 128, 144};
    for (int i = 0; i < 12; ++i) {
        const int a = cand[i];
        if (n % a != 0 || !col_tile_length_ok(n / a) || a > n / a) continue;
        best = a;  // the largest a <= n / a: the most balanced split
    }
    *n1 = best;
    *n2 = best ? n / best : n;
}

}  // namespace ldh

// Sequential-FP32-sum reproduction and unit-sum normalisation on the device.
#pragma once

#include <climits>

#include "ldh_internal.h"

namespace ldh {

constexpr int kSeqSumBlock = 4096;     // elements per block of the walk
constexpr int kSeqSumMinBlock = 512;   // shortest block a caller may choose

// Where the segments of one sum launch live: segment g reads array g / per_array at offset
// (g % per_array) * stride, each element multiplied on load by 1.0f / prescale[g % per_array]
// of its array when that array has prescale totals (the values an in-place normalisation by
// those totals would have left), totals <= 0 meaning no scale.
struct SeqSegments {
    static constexpr int kMaxArrays = 4;
    const float* array[kMaxArrays] = {nullptr, nullptr, nullptr, nullptr};
    const float* prescale[kMaxArrays] = {nullptr, nullptr, nullptr, nullptr};
    int per_array = 1;
    int64_t stride = 0;
    __device__ const float* base(int g) const { return array[g / per_array] + int64_t(g % per_array) * stride; }
    __device__ float scale(int g) const {
        const float* t = prescale[g / per_array];
        if (t == nullptr) return 1.0f;
        const float v = t[g % per_array];
        return v > 0.0f ? 1.0f / v : 1.0f;
    }
};

// v moved across lanes by a DPP control (both 32-bit halves); lanes that the control or the row
// mask leaves without a source read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Inclusive scan over the 64 lanes with DPP: shifts by 1, 2, 4, 8 within each row of 16, then
// the last lane of row 0 / 2 into rows 1 / 3 (row_bcast:15) and of row 1 into rows 2 and 3
// (row_bcast:31). All 64 lanes must be active. The walk's additions are exact (multiples of one
// power of two), so their order does not matter; the other callers form FP64 block sums, which
// only place blocks and mark empty ones.
__device__ __forceinline__ double wave_inclusive_scan(double v, int) {
    v += dpp_move<0x111, 0xf>(v);  // row_shr:1
    v += dpp_move<0x112, 0xf>(v);  // row_shr:2
    v += dpp_move<0x114, 0xf>(v);  // row_shr:4
    v += dpp_move<0x118, 0xf>(v);  // row_shr:8
    v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// The sum over the 64 lanes, in every lane (the scan's last lane, read back as a scalar).
__device__ __forceinline__ double wave_total(double v) {
    v = wave_inclusive_scan(v, 0);
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Producers of summed arrays run workgroups of kSumThreads threads over blocks of kSeqSumBlock
// elements and place them for the sum in the pass that writes them (place_blocks below).
constexpr int kSumThreads = 256;
__host__ __device__ inline int seqsum_blocks(int64_t n) { return int((n + kSeqSumBlock - 1) / kSeqSumBlock); }

// ---- single-pass block placement ---------------------------------------------------------------
// A producer whose output is summed next can place its blocks itself, in the pass that writes
// them: each workgroup (one kSeqSumBlock block of up to four segments) publishes its block total,
// finds the exclusive prefix of the segment by a decoupled look-back over the totals and prefixes
// the earlier workgroups publish, and forms the block's two binade sums from the values still in
// its registers. SeqSum::walk then finishes the sum with no prefix pass and no second read of the
// data. The prefix only places blocks (an FP32 estimate, as the FP64 one was): a look-back that
// has waited kPlaceMaxPolls polls without an answer leaves its block unplaced, which the walk
// replays, and words of other runs carry other tags, so the estimate can cost time but never
// change the value. Workgroups publish their totals before they wait and are dispatched in
// order, so the earliest waiting workgroup always has its predecessor running or done.
constexpr int kSeqNoBase = INT_MIN;   // running sum placed at zero (or unplaced): the walk replays
constexpr double kSeqTied = -1.0;     // a binade sum holding a rounding tie
constexpr int kSeqBins = 2;           // binades per block (the running sum starts in eb or eb+1)
constexpr int kPlaceMaxPolls = 4096;

struct SeqPlace {
    double* bsum = nullptr;                // [segment][nb] FP64 block sums
    double* bins = nullptr;                // [segment][nb][kSeqBins] binade sums
    int* ebase = nullptr;                  // [segment][nb] base binades
    unsigned long long* status = nullptr;  // [segment][nb] look-back words
    unsigned tag = 0;                      // this run's tag (< 2^21)
    int nb = 0;                            // blocks per segment
};

// Look-back word: high half 0x7FC00000 | tag << 1 | (1: inclusive prefix, 0: block total only),
// a float NaN pattern no summed data holds; low half the float value.
__device__ __forceinline__ unsigned place_word_hi(unsigned tag, bool prefix) {
    return 0x7FC00000u | (tag << 1) | (prefix ? 1u : 0u);
}
__device__ __forceinline__ void place_publish(unsigned long long* p, unsigned hi, float v) {
    const unsigned long long w = (static_cast<unsigned long long>(hi) << 32) | __float_as_uint(v);
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of block b of one segment (st: the segment's words), run by one whole wave:
// lane k reads the word of block j - k for a window of 64 predecessors; the closest one with a
// prefix ends the look-back, with every block between it and b showing at least its total.
// Returns -1 when the wait exceeds kPlaceMaxPolls polls.
__device__ __forceinline__ float place_lookback(unsigned long long* st, unsigned tag, int b, float total) {
    const int lane = threadIdx.x & 63;
    const unsigned hi_total = place_word_hi(tag, false), hi_prefix = place_word_hi(tag, true);
    if (b == 0) {
        if (lane == 0) place_publish(st, hi_prefix, total);
        return 0.0f;
    }
    if (lane == 0) place_publish(st + b, hi_total, total);
    float excl = 0.0f;
    int j = b - 1;
    int polls = 0;
    for (;;) {
        const int idx = j - lane;
        const unsigned long long w =
            idx >= 0 ? __hip_atomic_load(st + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
        const unsigned hi = unsigned(w >> 32);
        const bool has_prefix = idx < 0 || hi == hi_prefix;
        const bool ready = has_prefix || hi == hi_total;
        const unsigned long long pmask = __ballot(has_prefix);
        const int first = pmask != 0ull ? __ffsll(static_cast<long long>(pmask)) - 1 : 64;
        const unsigned long long upto = first >= 63 ? ~0ull : (2ull << first) - 1ull;  // lanes 0..first
        if ((__ballot(!ready) & upto) != 0ull) {
            if (++polls > kPlaceMaxPolls) return -1.0f;
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        float v = (lane <= first && idx >= 0) ? __uint_as_float(unsigned(w)) : 0.0f;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
        excl += v;
        if (first < 64) return excl;
        j -= 64;
    }
}

// Producers that place their blocks run a 1-D grid of place_grid(nb, segments) workgroups of
// one block each, block-major: workgroup p takes block p / segments of segment p % segments, so
// a block's predecessor in its segment was dispatched `segments` workgroups before it and has
// usually published its total, often its prefix, by the time the block looks back.
inline unsigned place_grid(int nb, int segments) { return unsigned(nb) * unsigned(segments); }
struct SegBlock {
    int seg, block;
};
__device__ __forceinline__ bool place_block_of(int nb, int segments, SegBlock& sb) {
    const int p = int(blockIdx.x);
    sb.seg = p % segments;
    sb.block = p / segments;
    return sb.block < nb;
}

// Places block b of N segments (seg0 + k * seg_stride, k < N) from the producer's
// values: v[k][j] is the j-th value the thread wrote into segment k's block (0 for positions
// past the end), share[k] the thread's FP64 sum of them. Writes the block's FP64 sum, base binade
// and binade sums into pl. Every thread of the kSumThreads workgroup calls it.
template <int N, int PER>
__device__ void place_blocks(const SeqPlace& pl, int seg0, int seg_stride, int b, const float (&v)[N][PER],
                             const double (&share)[N]) {
    constexpr int kWaves = kSumThreads / 64;
    static_assert(N <= kWaves, "one look-back wave per segment");
    __shared__ double part[N][kWaves];
    __shared__ double total_sh[N];
    __shared__ int eb_sh[N];
    __shared__ double red[N][kSeqBins][kWaves];
    __shared__ int tied[N][kSeqBins][kWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double t = wave_total(share[k]);
        if (lane == 0) part[k][wave] = t;
    }
    __syncthreads();
    if (wave < N) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += part[wave][w];
        const size_t row = size_t(seg0 + wave * seg_stride) * size_t(pl.nb);
        unsigned long long* st = pl.status + row;
        const float excl = place_lookback(st, pl.tag, b, float(t));
        int eb = kSeqNoBase;
        if (excl >= 0.0f) {
            if (lane == 0) place_publish(st + b, place_word_hi(pl.tag, true), excl + float(t));
            const double lo = double(excl) * (1.0 - 1.0 / 64.0);
            eb = lo > 0.0 ? ilogb(lo) : kSeqNoBase;
        }
        if (lane == 0) {
            pl.bsum[row + b] = t;
            pl.ebase[row + b] = eb;
            total_sh[wave] = t;
            eb_sh[wave] = eb;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int eb = eb_sh[k];
        if (eb == kSeqNoBase || total_sh[k] == 0.0) continue;  // the walk does not look these up
#pragma unroll
        for (int q = 0; q < kSeqBins; ++q) {
            // contributions counted in units of u = 2^(e-23): rint(x * 2^(23-e))
            const int shift = 23 - (eb + q);
            double acc = 0.0;
            bool tie = false;
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const float t = ldexpf(v[k][j], shift);
                const float r = rintf(t);
                tie |= fabsf(t - r) == 0.5f;
                acc += double(r);
            }
            acc = wave_total(acc);
            const bool any_tie = __ballot(tie) != 0ull;
            if (lane == 0) {
                red[k][q][wave] = acc;
                tied[k][q][wave] = any_tie ? 1 : 0;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < N * kSeqBins) {
        const int k = threadIdx.x / kSeqBins, q = threadIdx.x % kSeqBins;
        const int eb = eb_sh[k];
        if (eb != kSeqNoBase && total_sh[k] != 0.0) {
            double t = 0.0;
            int tie = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                t += red[k][q][w];
                tie |= tied[k][q][w];
            }
            const size_t row = size_t(seg0 + k * seg_stride) * size_t(pl.nb);
            pl.bins[(row + b) * kSeqBins + q] = tie ? kSeqTied : ldexp(t, (eb + q) - 23);
        }
    }
}

class SeqSum {
public:
    // totals[seg] = the FP32 sequential sum of x[seg*stride .. seg*stride + n) (device buffer).
    // x must be non-negative.
    void run(hipStream_t s, const float* x, int64_t n, int64_t stride, int segments, float* totals);
    // The same over the segments described by g (segments = arrays used * per_array). The sum
    // runs over blocks of `block` elements (kSeqSumMinBlock .. kSeqSumBlock); block_sums
    // (nullable, [segment][block]) are their FP64 sums when the producer of the data already
    // formed them (any summation order: they only place the blocks' binades and mark all-zero
    // blocks).
    void run(hipStream_t s, const SeqSegments& g, int64_t n, int segments, float* totals,
             const double* block_sums = nullptr, int block = kSeqSumBlock);
    // Segments of one array (as the first run) with their block sums from the producer.
    void run(hipStream_t s, const float* x, int64_t n, int64_t stride, int segments, float* totals,
             const double* block_sums);
    // x *= 1.0f / totals[seg] where totals[seg] > 0 (LensDiffCpuReference.cpp:440-449).
    void scale(hipStream_t s, float* x, int64_t n, int64_t stride, int segments, const float* totals);
    // run(), then scale().
    void normalize(hipStream_t s, float* x, int64_t n, int64_t stride, int segments, float* totals);
    // Grows the block tables for sums of `segments` arrays of n ahead of a graph capture.
    void reserve(int64_t n, int segments);
    // Single-pass placement: place() hands out the tables of a run of `segments` sums over n
    // elements in kSeqSumBlock blocks, with a fresh tag, for a producer that calls place_blocks
    // in every block of every segment; walk() then finishes that run on the same stream
    // (totals[seg] as run() would give). g describes where the producer wrote the data.
    SeqPlace place(int64_t n, int segments);
    void walk(hipStream_t s, const SeqSegments& g, int64_t n, int segments, float* totals);

private:
    DeviceBuffer bsum_, bins_, ebase_, status_;
    int64_t placed_n_ = -1;
    int placed_segments_ = 0;
};

// Allocation-free sums (ldh_sequential_sums_noalloc): the block tables live in a static device
// array of the library, one per device. noalloc_scratch_init looks up its address on `device` (once
// per device, at context creation); sequential_sums_noalloc then sums `segments` device arrays of
// n floats at x + g * n into totals[g] on stream s without allocating, for at most 65536
// (segment, 4096-element block) pairs. Calls that overlap on one device share the tables and race.
void noalloc_scratch_init(int device);
void sequential_sums_noalloc(hipStream_t s, int device, const float* x, int64_t n, int segments, float* totals);

}  // namespace ldh

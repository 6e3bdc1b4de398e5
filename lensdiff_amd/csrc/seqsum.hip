// Reproduces, in parallel, the value a sequential FP32 accumulation returns:
//     float s = 0; for (i = 0; i < n; ++i) s += x[i];
// which is how the reference normalises every PSF and kernel (std::accumulate with a float
// accumulator, LensDiffCpuReference.cpp:436-449). For the non-negative arrays summed here the
// running sum rises monotonically through binades [2^e, 2^(e+1)); while it stays inside one
// binade, adding x rounds x to a multiple of u = 2^(e-23), so
//     s_after_block = s_before + sum_i rint(x_i / u) * u
// whatever the order inside the block. The sum is found in four steps:
//   1. per-block FP64 sums;
//   2. their exclusive prefix per segment, which places the running sum of each block to well
//      within a binade; each block gets a base binade eb just below that estimate;
//   3. per block, the sums of rounded contributions for the two binades eb, eb+1 (exact in
//      FP64);
//   4. a one-wavefront walk over the blocks in order: a block whose binade sum keeps the
//      running sum inside its binade is added at once; a block where the running sum crosses
//      into the next binade, or whose binade lies outside the two computed, is replayed by the
//      whole workgroup: its 16-element chunks are added through a scan of their rounded
//      contributions, and the chunk holding the crossing is added element by element in true
//      FP32.
// An addend that lands exactly halfway between two multiples of u rounds by the parity of the
// running sum, which the block sums cannot know: binade sums and chunks holding such a tie are
// marked and always go the element-by-element way. The estimate only decides which blocks are
// replayed, never the value, so the result is the sequential FP32 sum, bit for bit, for any
// non-negative finite input.
// Steps 1-3 run here for sums whose data no kernel of ours is producing (SeqSum::run). A kernel
// that writes a summed array does them itself in the same pass (place_blocks in seqsum.h: the
// block's total, its prefix by a decoupled look-back, its two binade sums from registers), and
// SeqSum::walk runs step 4 alone.
#include <atomic>
#include <climits>
#include <cstdint>
#include <map>
#include <mutex>

#include "profiler.h"
#include "seqsum.h"

namespace ldh {

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = kSeqSumBlock / kThreads;
constexpr int kBins = kSeqBins;            // binades per block (the running sum starts in eb or eb+1)
constexpr int kNoBase = kSeqNoBase;        // running sum estimated at zero: replay the block
constexpr double kTied = kSeqTied;         // a binade sum holding a rounding tie


// Blocks are `bl` consecutive elements (bl <= kSeqSumBlock); block b covers
// [b * bl, min(n, (b + 1) * bl)).
__global__ void __launch_bounds__(kThreads)
block_sums_kernel(SeqSegments g, int64_t n, int bl, int nb, double* __restrict__ bsum) {
    const int seg = blockIdx.y, b = blockIdx.x;
    const float* base = g.base(seg);
    const float sc = g.scale(seg);
    const int64_t start = int64_t(b) * bl;
    const int64_t end = min(n, start + bl);
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
        const int64_t i = start + int64_t(k) * kThreads + threadIdx.x;
        if (i < end) acc += double(base[i] * sc);
    }
    __shared__ double red[kThreads / 64];
    acc = wave_total(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) t += red[w];
        bsum[int64_t(seg) * nb + b] = t;
    }
}

// Exclusive prefix of the block sums of one segment; base binade of each block placed 1/64
// below the prefix so that the running sum (within a fraction of a percent of the exact prefix)
// starts inside the two binades computed for the block (a block whose running sum starts higher is
// replayed, which only costs time).
__global__ void __launch_bounds__(kThreads)
prefix_kernel(const double* __restrict__ bsum, int nb, int* __restrict__ ebase) {
    const int seg = blockIdx.x, t = threadIdx.x;
    const double* bs = bsum + int64_t(seg) * nb;
    int* eb = ebase + int64_t(seg) * nb;
    const int per = (nb + kThreads - 1) / kThreads;
    const int b0 = min(nb, t * per), b1 = min(nb, b0 + per);
    double local = 0.0;
    for (int b = b0; b < b1; ++b) local += bs[b];
    __shared__ double part[kThreads];
    part[t] = local;
    __syncthreads();
    if (t == 0) {
        double run = 0.0;
        for (int i = 0; i < kThreads; ++i) {
            const double v = part[i];
            part[i] = run;
            run += v;
        }
    }
    __syncthreads();
    double run = part[t];
    for (int b = b0; b < b1; ++b) {
        const double lo = run * (1.0 - 1.0 / 64.0);
        eb[b] = lo > 0.0 ? ilogb(lo) : kNoBase;
        run += bs[b];
    }
}

__global__ void __launch_bounds__(kThreads)
binade_sums_kernel(SeqSegments g, int64_t n, int bl, int nb, const double* __restrict__ bsum,
                   const int* __restrict__ ebase, double* __restrict__ bins) {
    const int seg = blockIdx.y, b = blockIdx.x;
    const int eb = ebase[int64_t(seg) * nb + b];
    // the walk adds nothing for an all-zero block (block sum 0) and replays blocks without a base
    if (eb == kNoBase || bsum[int64_t(seg) * nb + b] == 0.0) return;
    const float* base = g.base(seg);
    const float sc = g.scale(seg);
    const int64_t start = int64_t(b) * bl;
    const int64_t end = min(n, start + bl);
    float v[kPerThread];
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
        const int64_t i = start + int64_t(k) * kThreads + threadIdx.x;
        v[k] = i < end ? base[i] : 0.0f;
    }
    // scaled after every load is issued (scaling at each load serialises the loads)
    if (sc != 1.0f) {
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) v[k] *= sc;
    }
    __shared__ double red[kBins][kThreads / 64];
    __shared__ int tied[kBins][kThreads / 64];
#pragma unroll
    for (int q = 0; q < kBins; ++q) {
        // contributions counted in units of u = 2^(e-23): rint(x * 2^(23-e))
        const int shift = 23 - (eb + q);
        double acc = 0.0;
        bool tie = false;
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const float t = ldexpf(v[k], shift);
            const float r = rintf(t);
            tie |= fabsf(t - r) == 0.5f;
            acc += double(r);
        }
        acc = wave_total(acc);
        const bool any_tie = __ballot(tie) != 0ull;
        if ((threadIdx.x & 63) == 0) {
            red[q][threadIdx.x >> 6] = acc;
            tied[q][threadIdx.x >> 6] = any_tie ? 1 : 0;
        }
    }
    __syncthreads();
    if (threadIdx.x < kBins) {
        const int q = threadIdx.x;
        double t = 0.0;
        int tie = 0;
        for (int w = 0; w < kThreads / 64; ++w) {
            t += red[q][w];
            tie |= tied[q][w];
        }
        // a rounding tie depends on the running sum's last bit: marked for replay
        bins[(int64_t(seg) * nb + b) * kBins + q] = tie ? kTied : ldexp(t, (eb + q) - 23);
    }
}

// Running-sum state of the walk: s is always an FP32 value; e = ilogb(s), or kNoBase at s = 0.
struct Run {
    double s;
    int e;
};

__device__ __forceinline__ double binade_limit(int e) { return ldexp(1.0, e + 1); }

// Sum of the rounded contributions of a chunk of C elements at binade e (exact values at s = 0),
// or kTied when one of them is a rounding tie.
template <int C>
__device__ __forceinline__ double chunk_contribution(const float* v, int e) {
    double acc = 0.0;
    if (e == kNoBase) {
#pragma unroll
        for (int j = 0; j < C; ++j) acc += double(v[j]);
        return acc;
    }
    const int shift = 23 - e;
    bool tie = false;
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const float t = ldexpf(v[j], shift);
        const float r = rintf(t);
        tie |= fabsf(t - r) == 0.5f;
        acc += double(r);
    }
    return tie ? kTied : ldexp(acc, e - 23);
}

// The walk runs one workgroup of kWalkThreads threads per segment. Every wave executes the
// block-level walk (identical data, identical decisions); the replay of a block spreads its
// kSeqSumBlock elements over all threads, kWalkChunk consecutive elements each.
constexpr int kWalkThreads = 256;
constexpr int kWalkWaves = kWalkThreads / 64;
constexpr int kWalkChunk = kSeqSumBlock / kWalkThreads;
constexpr int kNone = 1 << 30;

struct WalkShared {
    double wave_total[kWalkWaves];
    double incl;  // scan value of the chunk just resolved element by element
    int first_over[kWalkWaves];
    float sum;
};

// Adds one block element by element in effect: thread t holds chunk t (kWalkChunk consecutive
// elements). Chunks that keep the running sum inside its binade are added through a scan of
// their rounded contributions; the chunk where it leaves the binade (or where s leaves 0) is
// added by its own thread in true FP32, in order; so is a chunk holding a rounding tie. Every
// thread of the workgroup calls it with the same r and leaves with the same r.
__device__ void replay_block(const float* __restrict__ base, float sc, int64_t n, int bl, int64_t block, Run& r,
                             WalkShared& sh) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t start = block * bl + int64_t(tid) * kWalkChunk;
    const int64_t end = min(n, block * bl + bl);
    float v[kWalkChunk];
    if (start + kWalkChunk <= end && (reinterpret_cast<uintptr_t>(base + start) & 15) == 0) {
        const float4* p4 = reinterpret_cast<const float4*>(base + start);
#pragma unroll
        for (int j = 0; j < kWalkChunk / 4; ++j) {
            const float4 q = p4[j];
            v[4 * j] = q.x; v[4 * j + 1] = q.y; v[4 * j + 2] = q.z; v[4 * j + 3] = q.w;
        }
    } else {
        // a partial chunk: clamped loads (always in the block) issued together, then masked
#pragma unroll
        for (int j = 0; j < kWalkChunk; ++j) v[j] = base[min(start + j, end - 1)];
#pragma unroll
        for (int j = 0; j < kWalkChunk; ++j)
            if (start + j >= end) v[j] = 0.0f;
    }
    if (sc != 1.0f) {
#pragma unroll
        for (int j = 0; j < kWalkChunk; ++j) v[j] *= sc;
    }
    // Chunk contributions and their scan hold for one binade (ce); a chunk resolved element by
    // element that leaves the running sum in that binade (a rounding tie) only moves the origin
    // of the scan (base), so the next chunks need no new contributions or scan.
    int first = 0;
    int ce = kNoBase + 1;  // no contributions yet (never a binade or kNoBase)
    double c = 0.0, incl = 0.0, total = 0.0, base_incl = 0.0;
    bool tie = false;
    while (first < kWalkThreads) {
        if (r.e != ce) {
            c = tid >= first ? chunk_contribution<kWalkChunk>(v, r.e) : 0.0;
            tie = c == kTied;
            if (tie) c = 0.0;
            // workgroup inclusive scan: within the wave, then the totals of the waves before
            incl = wave_inclusive_scan(c, lane);
            if (lane == 63) sh.wave_total[wave] = incl;
            __syncthreads();
            total = 0.0;
#pragma unroll
            for (int w = 0; w < kWalkWaves; ++w) {
                if (w < wave) incl += sh.wave_total[w];
                total += sh.wave_total[w];
            }
            base_incl = 0.0;
            ce = r.e;
        }
        // the running sum after chunk t (t >= first, nothing over before it) is r.s + rel
        const double rel = incl - base_incl;
        const bool over =
            tid >= first && (tie || (r.e == kNoBase ? rel > 0.0 : r.s + rel >= binade_limit(r.e)));
        const unsigned long long mask = __ballot(over);
        if (lane == 0) sh.first_over[wave] = mask ? wave * 64 + __ffsll((long long)mask) - 1 : kNone;
        __syncthreads();
        int l = kNone;
#pragma unroll
        for (int w = 0; w < kWalkWaves; ++w) l = min(l, sh.first_over[w]);
        if (l == kNone) {
            r.s += total - base_incl;
            return;
        }
        if (tid == l) {
            float sf = float(r.s + (rel - c));  // exact: multiples of u below the binade limit
#pragma unroll
            for (int j = 0; j < kWalkChunk; ++j) sf = sf + v[j];
            sh.sum = sf;
            sh.incl = incl;
        }
        __syncthreads();
        const float sf = sh.sum;
        r.s = double(sf);
        r.e = sf > 0.0f ? ilogb(sf) : kNoBase;
        base_incl = sh.incl;
        first = l + 1;
    }
}

__global__ void __launch_bounds__(kWalkThreads)
walk_kernel(SeqSegments g, int64_t n, int bl, int nb, const double* __restrict__ bsum, const int* __restrict__ ebase,
            const double* __restrict__ bins, float* __restrict__ out) {
    __shared__ WalkShared sh;
    const int seg = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const float* base = g.base(seg);
    const float sc = g.scale(seg);
    const double* bs = bsum + int64_t(seg) * nb;
    const int* eb = ebase + int64_t(seg) * nb;
    const double* bn = bins + int64_t(seg) * nb * kBins;
    Run r{0.0, kNoBase};
    // per-lane block data of the current group, loaded one group ahead without branches (lanes
    // past the last block read it and are masked off), so the loads stay in flight while the
    // current group is walked
    auto load = [&](int blk, double& sum, int& e0, double (&b)[kBins]) {
        const int c = min(blk, nb - 1);
        sum = bs[c];
        e0 = eb[c];
#pragma unroll
        for (int q = 0; q < kBins; ++q) b[q] = bn[int64_t(c) * kBins + q];
        if (blk >= nb) {
            sum = 0.0;
            e0 = kNoBase;
        }
    };
    double cur_sum, nxt_sum;
    int cur_e0, nxt_e0;
    double cur_b[kBins] = {}, nxt_b[kBins] = {};
    load(lane, cur_sum, cur_e0, cur_b);
#pragma unroll 1
    for (int g0 = 0; g0 < nb; g0 += 64) {
        load(g0 + 64 + lane, nxt_sum, nxt_e0, nxt_b);
        const int blk = g0 + lane;
        int first = 0;
        while (first < 64) {
            // contribution of this lane's block at the current binade, or a forced replay
            double col = 0.0;
            bool force = false;
            if (lane >= first && blk < nb && cur_sum != 0.0) {
                const int q = r.e - cur_e0;
                if (r.e != kNoBase && cur_e0 != kNoBase && q >= 0 && q < kBins) {
                    col = cur_b[0];
#pragma unroll
                    for (int k = 1; k < kBins; ++k)
                        if (q == k) col = cur_b[k];
                    if (col == kTied) {
                        col = 0.0;
                        force = true;
                    }
                } else {
                    force = true;
                }
            }
            const double incl = wave_inclusive_scan(col, lane);
            const bool over = force || (lane >= first && blk < nb && r.e != kNoBase && r.s + incl >= binade_limit(r.e));
            const unsigned long long mask = __ballot(over);
            if (mask == 0ull) {
                r.s += __shfl(incl, 63);
                break;
            }
            const int l = __ffsll((long long)mask) - 1;
            r.s += __shfl(incl - col, l);
            replay_block(base, sc, n, bl, g0 + l, r, sh);
            first = l + 1;
        }
        cur_sum = nxt_sum;
        cur_e0 = nxt_e0;
#pragma unroll
        for (int q = 0; q < kBins; ++q) cur_b[q] = nxt_b[q];
    }
    if (threadIdx.x == 0) out[seg] = float(r.s);
}

__global__ void scale_by_inverse_kernel(float* __restrict__ x, int64_t n, int64_t stride,
                                        const float* __restrict__ totals) {
    const int seg = blockIdx.y;
    const float total = totals[seg];
    if (total <= 0.0f) return;  // the reference leaves a non-positive total untouched
    const float inv = 1.0f / total;
    float* base = x + seg * stride;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        base[i] *= inv;
}

}  // namespace

void SeqSum::run(hipStream_t s, const float* x, int64_t n, int64_t stride, int segments, float* totals) {
    SeqSegments g;
    g.array[0] = x;
    g.per_array = segments;
    g.stride = stride;
    run(s, g, n, segments, totals);
}

void SeqSum::run(hipStream_t s, const float* x, int64_t n, int64_t stride, int segments, float* totals,
                 const double* block_sums) {
    SeqSegments g;
    g.array[0] = x;
    g.per_array = segments;
    g.stride = stride;
    run(s, g, n, segments, totals, block_sums);
}

void SeqSum::reserve(int64_t n, int segments) {
    const int nb = int(div_up(n, kSeqSumMinBlock));  // room for the shortest blocks
    bsum_.reset(sizeof(double) * size_t(nb) * segments);
    bins_.reset(sizeof(double) * size_t(nb) * segments * kBins);
    ebase_.reset(sizeof(int) * size_t(nb) * segments);
    status_.reset(sizeof(unsigned long long) * size_t(seqsum_blocks(n)) * segments);
}


namespace {

// The block tables of one run: FP64 block sums, the per-block base binade, and each block's sums
// for its two binades. Every entry a kernel reads was written earlier in the same run (blocks
// whose sum is 0 or whose binade is unknown are never looked up), so the tables need no clearing.
struct Tables {
    double* bsum;
    double* bins;
    int* ebase;
};

void launch_run(hipStream_t s, const SeqSegments& g, int64_t n, int segments, float* totals,
                const double* block_sums, int block, const Tables& t, bool profiled) {
    const int nb = int(div_up(n, block));
    const double data_bytes = 4.0 * double(n) * segments;  // one read of the summed data
    const double* bsum = block_sums != nullptr ? block_sums : t.bsum;
    {
        ProfileScope ps(profiled ? "seqsum-block-sums" : "", block_sums != nullptr ? 0.0 : data_bytes);
        if (block_sums == nullptr)
            hipLaunchKernelGGL(block_sums_kernel, dim3(nb, segments), dim3(kThreads), 0, s, g, n, block, nb, t.bsum);
        hipLaunchKernelGGL(prefix_kernel, dim3(segments), dim3(kThreads), 0, s, bsum, nb, t.ebase);
    }
    {
        ProfileScope ps(profiled ? "seqsum-binade-sums" : "", data_bytes);
        hipLaunchKernelGGL(binade_sums_kernel, dim3(nb, segments), dim3(kThreads), 0, s, g, n, block, nb, bsum,
                           t.ebase, t.bins);
    }
    {
        // reads the block tables and replays a few blocks: latency-bound, no byte model
        ProfileScope ps(profiled ? "seqsum-walk" : "", 0.0);
        hipLaunchKernelGGL(walk_kernel, dim3(segments), dim3(kWalkThreads), 0, s, g, n, block, nb, bsum, t.ebase,
                           t.bins, totals);
    }
    LDH_CHECK_LAUNCH("seqsum-launch");
}

void check_run(const SeqSegments& g, int segments, int block) {
    if (segments > g.per_array * SeqSegments::kMaxArrays) fail(LDH_ERR_INVALID_ARGUMENT, "seqsum-segments");
    if (block < kSeqSumMinBlock || block > kSeqSumBlock) fail(LDH_ERR_INVALID_ARGUMENT, "seqsum-block");
}

// Scratch of the allocation-free sums (sequential_sums_noalloc): the block tables of at most
// kNoAllocBlocks (segment, block) pairs, 1.75 MiB per device, held until the process ends. Host
// code never uses this name as an address: it takes the device address that hipGetSymbolAddress
// returns for the current device (noalloc_scratch).
constexpr int64_t kNoAllocBlocks = int64_t(1) << 16;
constexpr size_t kNoAllocBytes = size_t(kNoAllocBlocks) * (2 * sizeof(double) + sizeof(double) + sizeof(int));
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;  // device -> address of its copy of `scratch`

}  // namespace

void SeqSum::run(hipStream_t s, const SeqSegments& g, int64_t n, int segments, float* totals, const double* block_sums,
                 int block) {
    if (n <= 0 || segments <= 0) return;
    check_run(g, segments, block);
    reserve(n, segments);
    launch_run(s, g, n, segments, totals, block_sums, block,
               Tables{bsum_.as<double>(), bins_.as<double>(), ebase_.as<int>()}, true);
}

void noalloc_scratch_init(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(device)) return;
    int prev = -1;
    LDH_HIP(hipGetDevice(&prev), "noalloc-scratch");
    if (prev != device) LDH_HIP(hipSetDevice(device), "noalloc-scratch");
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev != device) (void)hipSetDevice(prev);
    LDH_HIP(e, "noalloc-scratch-address");
    if (ptr == nullptr) fail(LDH_ERR_DEVICE, "noalloc-scratch-address");
    g_scratch_ptr[device] = ptr;
}

void sequential_sums_noalloc(hipStream_t s, int device, const float* x, int64_t n, int segments, float* totals) {
    if (n <= 0 || segments <= 0) return;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        auto it = g_scratch_ptr.find(device);
        if (it == g_scratch_ptr.end()) fail(LDH_ERR_DEVICE, "noalloc-scratch-not-initialised");
        base = it->second;
    }
    const int64_t pairs = div_up(n, kSeqSumBlock) * segments;
    if (pairs > kNoAllocBlocks) fail(LDH_ERR_TOO_LARGE, "noalloc-scratch-too-small");
    SeqSegments g;
    g.array[0] = x;
    g.per_array = segments;
    g.stride = n;
    check_run(g, segments, kSeqSumBlock);
    // [bsum: pairs doubles][bins: 2 pairs doubles][ebase: pairs ints], each table 256-byte aligned
    auto align = [](size_t b) { return (b + 255) / 256 * 256; };
    unsigned char* p = static_cast<unsigned char*>(base);
    Tables t;
    t.bsum = reinterpret_cast<double*>(p);
    t.bins = reinterpret_cast<double*>(p + align(sizeof(double) * size_t(pairs)));
    t.ebase = reinterpret_cast<int*>(p + align(sizeof(double) * size_t(pairs)) + align(2 * sizeof(double) * size_t(pairs)));
    launch_run(s, g, n, segments, totals, nullptr, kSeqSumBlock, t, false);
}

namespace {
// Tags of placement runs, process-wide so that a status buffer handed from one workspace to
// another by the pool cannot hold words of the same tag.
std::atomic<unsigned> g_place_tag{0};
}  // namespace

SeqPlace SeqSum::place(int64_t n, int segments) {
    if (n <= 0 || segments <= 0) fail(LDH_ERR_INVALID_ARGUMENT, "seqsum-place-size");
    reserve(n, segments);
    SeqPlace pl;
    pl.bsum = bsum_.as<double>();
    pl.bins = bins_.as<double>();
    pl.ebase = ebase_.as<int>();
    pl.status = status_.as<unsigned long long>();
    pl.tag = (g_place_tag.fetch_add(1, std::memory_order_relaxed) & ((1u << 21) - 1u));
    pl.nb = seqsum_blocks(n);
    placed_n_ = n;
    placed_segments_ = segments;
    return pl;
}

void SeqSum::walk(hipStream_t s, const SeqSegments& g, int64_t n, int segments, float* totals) {
    if (n != placed_n_ || segments != placed_segments_) fail(LDH_ERR_INVALID_ARGUMENT, "seqsum-walk-without-placement");
    check_run(g, segments, kSeqSumBlock);
    placed_n_ = -1;
    const int nb = seqsum_blocks(n);
    ProfileScope ps("seqsum-walk", 0.0);
    hipLaunchKernelGGL(walk_kernel, dim3(segments), dim3(kWalkThreads), 0, s, g, n, kSeqSumBlock, nb,
                       bsum_.as<double>(), ebase_.as<int>(), bins_.as<double>(), totals);
    LDH_CHECK_LAUNCH("seqsum-walk");
}

void SeqSum::scale(hipStream_t s, float* x, int64_t n, int64_t stride, int segments, const float* totals) {
    const int grid = int(std::min<int64_t>(div_up(n, 256 * 4), 2048));
    ProfileScope ps("seqsum-scale", 8.0 * double(n) * segments);
    hipLaunchKernelGGL(scale_by_inverse_kernel, dim3(grid, segments), dim3(256), 0, s, x, n, stride, totals);
    LDH_CHECK_LAUNCH("normalize-launch");
}

void SeqSum::normalize(hipStream_t s, float* x, int64_t n, int64_t stride, int segments, float* totals) {
    run(s, x, n, stride, segments, totals);
    scale(s, x, n, stride, segments, totals);
}

}  // namespace ldh

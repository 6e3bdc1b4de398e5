// Internal declarations shared by the HIP translation units of liblensdiff_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "lensdiff_hip.h"

namespace ldh {

// ---- error handling ----------------------------------------------------------------------
struct Error {
    ldh_status status;
    std::string tag;  // kebab-case, as the reference's backends report (LensDiffCuda.cu:711-729)
};

#define LDH_HIP(call, tag)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            throw ::ldh::Error{LDH_ERR_DEVICE, std::string(tag) + ":" + hipGetErrorName(e_)}; \
        }                                                                                  \
    } while (0)

#define LDH_CHECK_LAUNCH(tag) LDH_HIP(hipGetLastError(), tag)

inline void fail(ldh_status s, const std::string& tag) { throw Error{s, tag}; }

// Test hook (ldh_debug_inject_failure): throws Error{LDH_ERR_DEVICE, "injected-<site>"} while
// failures are armed for `site`, consuming one each time.
void debug_failure_point(const char* site);

// ---- small device helpers ------------------------------------------------------------------
// Complex products may contract to FMA (the library is built with -ffp-contract=off so that the
// pupil, selection and ring arithmetic keep their float semantics; transforms gain from FMA).
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
#pragma clang fp contract(fast)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * b with the fusion written out (real: fma(a.x, b.x, -a.y b.y), imaginary:
// fma(a.x, b.y, a.y b.x)): the same bits wherever it is inlined, where cmul leaves the choice of
// which product to fuse to the compiler.
__host__ __device__ inline float2 cmul_fixed(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__host__ __device__ inline float2 cmulc(float2 a, float2 b) {  // a * conj(b)
#pragma clang fp contract(fast)
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

inline int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}
inline bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
inline int pow2_at_least(int v) {
    int n = 1;
    while (n < v) n <<= 1;
    return n;
}
inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- device memory ---------------------------------------------------------------------------
// Pooled device allocations (pool.hip). pool_alloc returns a block of at least `bytes` on the
// current device (its size and device in granted / device), or null when the device is full;
// pool_release hands it back, reusable once the work queued before the release on every
// registered stream of that device has completed. Streams that run library work are registered
// while they live (unregister after synchronising them, before destroying them).
void* pool_alloc(size_t bytes, size_t* granted, int* device);
void pool_release(void* p, size_t granted, int device);
void pool_register_stream(int device, hipStream_t s);
void pool_unregister_stream(int device, hipStream_t s);
// Returns the device's cached blocks to the driver (after a device synchronisation).
void pool_trim(int device);
size_t pool_cached_bytes(int device);

// A device allocation owned by RAII, from the pool. Allocation failures become
// LDH_ERR_OUT_OF_MEMORY. bytes() is the granted size (at least the requested one).
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { reset(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_), dev_(o.dev_) {
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            ptr_ = o.ptr_;
            bytes_ = o.bytes_;
            dev_ = o.dev_;
            o.ptr_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    void reset(size_t bytes) {
        if (bytes <= bytes_ && ptr_) return;
        release();
        if (bytes == 0) return;
        size_t granted = 0;
        int dev = 0;
        ptr_ = pool_alloc(bytes, &granted, &dev);
        if (ptr_ == nullptr) fail(LDH_ERR_OUT_OF_MEMORY, "device-allocation-failed");
        bytes_ = granted;
        dev_ = dev;
    }
    void release() {
        if (ptr_) pool_release(ptr_, bytes_, dev_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    void* get() const { return ptr_; }
    size_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
    int dev_ = 0;
};

// ---- timing of kernel families with HIP events (on the launch stream) -----------------------
struct Profiler;

// ---- FFT (fft.hip) ---------------------------------------------------------------------------
// Transform lengths are 2^a * 3^b with b <= 2, up to kMaxFft (the PSF line transforms also
// take 5 * 2^a). Twiddles come from one buffer holding four tables, one per length family:
// exp(-2 pi i j / T) for T = 16384 (powers of two), 12288 (3 * 2^a), 9216 (9 * 2^a) and 10240
// (5 * 2^a). A length N reads its family's table at stride T / N, so the last stage of every
// transform reads consecutive entries.
constexpr int kMaxFftLog2 = 14;
constexpr int kMaxFft = 1 << kMaxFftLog2;   // 16384 points
constexpr int kTw2 = 16384, kTw3 = 12288, kTw9 = 9216, kTw5 = 10240;
constexpr int kTwOff2 = 0, kTwOff3 = kTw2, kTwOff9 = kTw2 + kTw3, kTwOff5 = kTw2 + kTw3 + kTw9;
constexpr int kTwiddleEntries = kTw2 + kTw3 + kTw9 + kTw5;

// odd part of n (1, 3, 5 or 9 for supported lengths)
__host__ __device__ constexpr int odd_part(int n) { return (n & 1) ? n : odd_part(n >> 1); }
// Table length and offset of the family of length n.
__host__ __device__ constexpr int tw_table(int n) {
    return odd_part(n) == 1 ? kTw2 : odd_part(n) == 3 ? kTw3 : odd_part(n) == 5 ? kTw5 : kTw9;
}
__host__ __device__ constexpr int tw_offset(int n) {
    return odd_part(n) == 1 ? kTwOff2 : odd_part(n) == 3 ? kTwOff3 : odd_part(n) == 5 ? kTwOff5 : kTwOff9;
}

// 2^a * 3^b, b <= 2, in [2, kMaxFft]
__host__ __device__ constexpr bool fft_length_ok(int n) {
    if (n < 2 || n > kMaxFft) return false;
    while ((n & 1) == 0) n >>= 1;
    return n == 1 || n == 3 || n == 9;
}
// Grid lengths: powers of two from 8, 3 * 2^a from 48, 9 * 2^a from 144 (the lengths the
// row and column kernels are built for; rows also need at least 16).
__host__ __device__ constexpr bool grid_length_ok(int n) {
    if (!fft_length_ok(n)) return false;
    int odd = n;
    while ((odd & 1) == 0) odd >>= 1;
    return odd == 1 ? n >= 8 : (odd == 3 ? n >= 48 : n >= 144);
}
// Smallest transform length >= v (at least 128; 0 when none fits kMaxFft).
inline int fft_length_at_least(int v) {
    int best = 0;
    for (int odd : {1, 3, 9})
        for (int n = odd; n <= kMaxFft; n *= 2)
            if (n >= v && n >= 128 && fft_length_ok(n) && (best == 0 || n < best)) best = n;
    return best;
}
// Smallest grid length >= v (0 when none fits kMaxFft).
inline int grid_length_at_least(int v) {
    int best = 0;
    for (int odd : {1, 3, 9})
        for (int n = odd; n <= kMaxFft; n *= 2)
            if (n >= v && grid_length_ok(n) && (best == 0 || n < best)) best = n;
    return best;
}

struct FftTables {
    DeviceBuffer twiddle;  // float2[kTwiddleEntries]: the four family tables
    void init(hipStream_t s);
};

// 2D complex FFT of a height x width array (row pitch = width), both grid lengths. Forward
// leaves the spectrum rows in the column transform's digit-split order (fft_col_storage_row);
// the matching inverse takes that order back. Inverse is unscaled.
void fft2d_forward(hipStream_t s, const FftTables& t, float2* data, int width, int height);
void fft2d_inverse(hipStream_t s, const FftTables& t, float2* data, int width, int height);
// Row or column passes only (used by fused pipelines).
void fft_rows(hipStream_t s, const FftTables& t, float2* data, int width, int rows, bool inverse);
void fft_cols_forward(hipStream_t s, const FftTables& t, float2* data, int width, int height);
void fft_cols_inverse(hipStream_t s, const FftTables& t, float2* data, int width, int height);
// Map a natural frequency row index to the storage row used by the split column transform of
// a width x height grid (the height's planned split, tuned on first use on stream s).
int fft_col_storage_row(hipStream_t s, const FftTables& t, int width, int height, int k);

}  // namespace ldh

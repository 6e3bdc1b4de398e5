// Device memory pool behind DeviceBuffer. Large hipMalloc / hipFree calls take milliseconds to
// seconds on a busy device (page-table work in the driver, and hipFree waits for the device), and
// the served loop creates and evicts banks and kernel spectra of gigabytes all the time, so freed
// blocks are kept and handed out again instead.
//
// A released block can still be read or written by work queued before its release on any of the
// library's streams. It therefore waits in a pending list, with one event recorded on every
// stream registered for its device at release time, and becomes reusable once all those events
// have completed. Streams are registered by the contexts and server slots that own them. A stream
// that is being captured into a graph gets no event (recording would enter the graph); a block
// released then waits for a device synchronisation, which the pool performs before it falls back
// to hipMalloc.
#include <map>
#include <mutex>
#include <vector>

#include "ldh_internal.h"

namespace ldh {

namespace {

constexpr size_t kSmall = size_t(1) << 20;   // below: 4 KiB granules, above: 2 MiB
constexpr size_t kSmallGranule = size_t(4) << 10;
constexpr size_t kLargeGranule = size_t(2) << 20;

size_t granted_size(size_t bytes) {
    const size_t g = bytes < kSmall ? kSmallGranule : kLargeGranule;
    return (bytes + g - 1) / g * g;
}

struct Pending {
    void* ptr;
    size_t size;
    std::vector<hipEvent_t> events;
    bool needs_sync;
};

struct DevicePool {
    std::multimap<size_t, void*> free_blocks;
    std::vector<Pending> pending;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> spare_events;
    size_t cached = 0;  // bytes in free + pending blocks

    hipEvent_t event() {
        if (!spare_events.empty()) {
            hipEvent_t e = spare_events.back();
            spare_events.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return e;
    }
    // Pending blocks whose events have all completed become free.
    void collect(bool synced) {
        for (size_t i = 0; i < pending.size();) {
            Pending& p = pending[i];
            bool done = synced || !p.needs_sync;
            if (done && !synced)
                for (hipEvent_t e : p.events) {
                    const hipError_t q = hipEventQuery(e);
                    if (q == hipErrorNotReady) {
                        done = false;
                        break;
                    }
                    if (q != hipSuccess) (void)hipGetLastError();
                }
            if (!done) {
                ++i;
                continue;
            }
            for (hipEvent_t e : p.events) spare_events.push_back(e);
            free_blocks.emplace(p.size, p.ptr);
            pending[i] = pending.back();
            pending.pop_back();
        }
    }
    void* take(size_t size) {
        // best fit within a quarter (or, for small blocks, a factor of two) of the request
        auto it = free_blocks.lower_bound(size);
        if (it == free_blocks.end()) return nullptr;
        const size_t limit = size < kSmall ? 2 * size : size + size / 4;
        if (it->first > limit) return nullptr;
        void* p = it->second;
        cached -= it->first;
        free_blocks.erase(it);
        return p;
    }
};

std::mutex g_pool_mu;
// never destroyed: blocks stay mapped until the process ends
std::map<int, DevicePool>& pools() {
    static auto* p = new std::map<int, DevicePool>();
    return *p;
}

int current_device() {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) (void)hipGetLastError();
    return d;
}

}  // namespace

void* pool_alloc(size_t bytes, size_t* granted, int* device) {
    const int dev = current_device();
    const size_t size = std::max(granted_size(bytes), kSmallGranule);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    DevicePool& pool = pools()[dev];
    pool.collect(false);
    void* p = pool.take(size);
    if (p == nullptr) {
        // blocks that only wait on the device, then a fresh allocation; when the device is full,
        // the cached blocks go back first
        if (!pool.pending.empty()) {
            if (hipDeviceSynchronize() != hipSuccess) (void)hipGetLastError();
            pool.collect(true);
            p = pool.take(size);
        }
        if (p == nullptr && hipMalloc(&p, size) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            for (auto& kv : pool.free_blocks) (void)hipFree(kv.second);
            pool.free_blocks.clear();
            pool.cached = 0;
            for (const Pending& q : pool.pending) pool.cached += q.size;
            if (hipMalloc(&p, size) != hipSuccess) {
                (void)hipGetLastError();
                p = nullptr;
            }
        }
    }
    if (p == nullptr) return nullptr;
    *granted = size;
    *device = dev;
    return p;
}

void pool_release(void* p, size_t granted, int device) {
    if (p == nullptr) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    DevicePool& pool = pools()[device];
    Pending pend{p, granted, {}, false};
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) (void)hipGetLastError();
    if (prev != device) (void)hipSetDevice(device);
    for (hipStream_t s : pool.streams) {
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cap) != hipSuccess) {
            (void)hipGetLastError();
            pend.needs_sync = true;
            continue;
        }
        if (cap != hipStreamCaptureStatusNone) {
            pend.needs_sync = true;
            continue;
        }
        hipEvent_t e = pool.event();
        if (e == nullptr || hipEventRecord(e, s) != hipSuccess) {
            (void)hipGetLastError();
            if (e) pool.spare_events.push_back(e);
            pend.needs_sync = true;
            continue;
        }
        pend.events.push_back(e);
    }
    if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
    pool.cached += granted;
    pool.pending.push_back(std::move(pend));
}

void pool_register_stream(int device, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    pools()[device].streams.push_back(s);
}

void pool_unregister_stream(int device, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto& v = pools()[device].streams;
    for (size_t i = 0; i < v.size(); ++i)
        if (v[i] == s) {
            v[i] = v.back();
            v.pop_back();
            break;
        }
}

void pool_trim(int device) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    DevicePool& pool = pools()[device];
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) (void)hipGetLastError();
    if (prev != device) (void)hipSetDevice(device);
    if (!pool.pending.empty()) {
        if (hipDeviceSynchronize() != hipSuccess) (void)hipGetLastError();
        pool.collect(true);
    }
    for (auto& kv : pool.free_blocks) (void)hipFree(kv.second);
    pool.free_blocks.clear();
    pool.cached = 0;
    if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
}

size_t pool_cached_bytes(int device) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    return pools()[device].cached;
}

}  // namespace ldh

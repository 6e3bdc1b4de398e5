// Served form: a request loop over several device slots (one HIP stream, frame scratch and bank
// workspace each) that keeps PSF banks and their kernel spectra resident in HBM, keyed by the
// lens parameters, admits requests under a memory budget and lets requests that arrive while
// others run proceed on free slots, so their work overlaps on the device.
#pragma once

#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "bank.h"
#include "frame.h"
#include "host_params.h"
#include "profiler.h"

namespace ldh {

struct ServedRequest {
    int64_t ticket = 0;
    ldh_params params{};
    const float4* src = nullptr;
    float4* dst = nullptr;
    int w = 0, h = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
    bool bank_hit = false, spectra_hit = false;
    int passed = 0;  // times a later request was admitted ahead of this one at the queue's head
    ldh_status status = LDH_OK;
    std::string error;
};

struct SpectraKey {
    int pw, ph, kernels, fp16;
    bool operator==(const SpectraKey& o) const {
        return pw == o.pw && ph == o.ph && kernels == o.kernels && fp16 == o.fp16;
    }
};

struct CacheEntry {
    BankKey key;
    std::shared_ptr<DeviceBank> bank;
    // the nine zone banks of a field-dependent phase, zone (zx, zy) at zy * 3 + zx (bank unset)
    std::vector<std::shared_ptr<DeviceBank>> zones;
    // precombined natural (+ style) kernels, float[n][K*K] with n = 3 (6 with style), per zone
    // (zone z at z * n * K * K)
    DeviceBuffer combined;
    bool has_combined = false;
    bool combined_style = false;
    // Kernel spectra for one FFT grid. The record is shared with the requests waiting for it, so
    // a failed build can hand its error to them after it has left the entry.
    struct Spectra {
        SpectraKey key;
        KernelSpectra spec;
        std::vector<KernelSpectra> zones;  // per zone, for a field-dependent phase (spec unused)
        bool ready = false;
        bool failed = false;
        std::string error;
    };
    std::vector<std::shared_ptr<Spectra>> spectra;
    bool ready = false;     // bank (and combined kernels) built
    bool failed = false;
    std::string error;
    int pins = 0;
    int uses = 0;  // requests admitted with this entry
    uint64_t last_use = 0;
    size_t bytes() const;
};

class Server {
public:
    Server(const FftTables& fft, int device, int slots, uint64_t budget);
    ~Server();
    int64_t submit(const ldh_params& p, const float4* src, float4* dst, int w, int h);
    // true when done (fills *out)
    bool poll(int64_t ticket, ServedRequest* out, bool wait);
    ldh_server_stats stats();

private:
    struct Slot;
    void worker(Slot* s);
    ServedRequest* admit(Slot* s, std::shared_ptr<CacheEntry>* entry, bool* builder);
    size_t resident_bytes_locked() const;
    bool evict_locked(size_t need);

    const FftTables& fft_;
    int device_;
    uint64_t budget_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stopping_ = false;
    int64_t next_ticket_ = 1;
    uint64_t clock_ = 0;
    std::deque<ServedRequest*> queue_;
    std::map<int64_t, std::unique_ptr<ServedRequest>> requests_;
    std::list<std::shared_ptr<CacheEntry>> cache_;
    std::vector<std::unique_ptr<Slot>> slots_;
    ldh_server_stats stats_{};

public:
    // Per-slot kernel timing (HIP events on each slot's stream), merged on read.
    void set_profiling(bool on);
    void merge_profiles(std::map<std::string, Profiler::Stat>& into);
};

}  // namespace ldh

// The served request loop (see server.h). Each slot is a worker thread with its own HIP stream,
// frame scratch and bank workspace; the FFT tables are shared read-only. A request's bank and
// the kernel spectra for its FFT grid come from the resident cache; the first request needing
// an entry builds it on its own stream while requests for other entries keep running on the
// other slots, and later requests for the same entry wait for it to be ready.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <functional>

#include "server.h"

namespace ldh {

namespace {
double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

bool spectral_fullkernel(const ldh_params& p) {
    return is_spectral(p) && p.look_mode == LDH_LOOK_PHYSICAL && p.debug_view != LDH_VIEW_CORE &&
           p.debug_view != LDH_VIEW_STRUCTURE;
}
bool with_style(const ldh_params& p) {
    return p.spectrum_style != LDH_STYLE_NATURAL && std::max(0.0, std::min(p.spectrum_force, 1.0)) > 0.0;
}
// Real kernels whose spectra serve this request (per zone with a field-dependent phase):
// precombined RGB (3, or 6 with style) for a spectral bank, the single full kernel for a mono
// bank; 0 = not cacheable (split look).
int cached_kernel_count(const ldh_params& p) {
    if (spectral_fullkernel(p)) return with_style(p) ? 6 : 3;
    if (!is_spectral(p) && p.look_mode == LDH_LOOK_PHYSICAL && p.debug_view != LDH_VIEW_CORE &&
        p.debug_view != LDH_VIEW_STRUCTURE)
        return 1;
    return 0;
}

// The kernel sets of a request's bank: the zone banks of a field-dependent phase hold all three;
// a single bank the sets its renders convolve with (a static view reads the full kernels only).
int served_sets(const ldh_params& p) {
    if (has_field_phase(p)) return 3;
    return is_static_view(p.debug_view) ? 1 : kernel_sets_needed(p);
}

size_t bank_bytes_estimate(const ldh_params& p) {
    const Sizes s = resolve_sizes(p);
    const size_t kk = size_t(2 * s.max_radius + 1) * size_t(2 * s.max_radius + 1);
    const size_t combined = spectral_fullkernel(p) ? (with_style(p) ? 6 : 3) : 0;
    return kk * 4 * (size_t(served_sets(p)) * wavelength_list(p).size() + combined) *
           (has_field_phase(p) ? kFieldZones : 1);
}

// The parameters of the nine zone banks of a field-dependent phase, each built from its resolved
// global terms (buildPsfBankGlobalOnly).
std::vector<ldh_params> zone_requests(const ldh_params& p) {
    std::vector<ldh_params> zones;
    for (int z = 0; z < kFieldZones; ++z) {
        zones.push_back(field_zone_params(p, z % 3, z / 3));
        zones.back().phase_field_strength = 0.0;
    }
    return zones;
}

// Cache identity of a request: its bank key and kernel sets, or its zones' keys in order.
BankKey served_key(const ldh_params& p) {
    if (!has_field_phase(p)) {
        BankKey k = make_bank_key(p);
        k.values.push_back(double(served_sets(p)));
        return k;
    }
    BankKey k;
    k.values.push_back(double(kFieldZones));
    for (const ldh_params& z : zone_requests(p)) {
        const BankKey zk = make_bank_key(z);
        k.values.insert(k.values.end(), zk.values.begin(), zk.values.end());
    }
    return k;
}

// The grid size a request convolves at: its working size when resolution-aware.
void served_frame_size(const ldh_params& p, int w, int h, int& gw, int& gh) {
    gw = w;
    gh = h;
    if (resolution_aware_active(p)) {
        const double ws = working_scale(p);
        gw = std::max(1, int(std::lround(w * ws)));
        gh = std::max(1, int(std::lround(h * ws)));
    }
}
size_t spectra_bytes(const Grid& g, int kernels, bool fp16) {
    return g.cells() * size_t((kernels + 1) / 2) * (fp16 ? 4 : 8);
}
// Kernel sets whose spectra a request keeps: one per zone with a field-dependent phase.
int spectra_sets(const ldh_params& p) { return has_field_phase(p) ? kFieldZones : 1; }
// The support a request's grid is sized for: its bank's, its zones' largest, or before they are
// built the kernel radius bound.
int entry_support(const CacheEntry* e, const ldh_params& p) {
    if (e == nullptr || !e->ready) return resolve_sizes(p).max_radius;
    if (e->bank) return e->bank->support;
    int support = 0;
    for (const auto& z : e->zones) support = std::max(support, z->support);
    return support;
}
}  // namespace

size_t CacheEntry::bytes() const {
    size_t b = (bank ? bank->bytes() : 0) + combined.bytes();
    for (const auto& z : zones)
        if (z) b += z->bytes();
    for (const auto& s : spectra)
        if (s->ready) {  // a record in the making is written without the lock
            b += s->spec.bytes();
            for (const auto& z : s->zones) b += z.bytes();
        }
    return b;
}

struct Server::Slot {
    int index = 0;
    hipStream_t stream = nullptr;
    FrameWork fw;
    BankWorkspace bw;
    DeviceBuffer tmp;
    Profiler prof;
    std::mutex prof_mu;
    std::thread thread;
};

void Server::set_profiling(bool on) {
    for (auto& s : slots_) {
        std::lock_guard<std::mutex> lk(s->prof_mu);
        s->prof.enabled = on;
    }
}

void Server::merge_profiles(std::map<std::string, Profiler::Stat>& into) {
    for (auto& s : slots_) {
        std::lock_guard<std::mutex> lk(s->prof_mu);
        (void)hipStreamSynchronize(s->stream);
        s->prof.resolve();
        for (const auto& kv : s->prof.stats) {
            Profiler::Stat& d = into[kv.first];
            d.launches += kv.second.launches;
            d.ms += kv.second.ms;
            d.bytes += kv.second.bytes;
        }
        s->prof.stats.clear();
    }
}

Server::Server(const FftTables& fft, int device, int slots, uint64_t budget)
    : fft_(fft), device_(device), budget_(budget) {
    stats_.slots = slots;
    stats_.budget_bytes = budget;
    for (int i = 0; i < slots; ++i) {
        auto s = std::make_unique<Slot>();
        s->index = i;
        LDH_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "server-stream");
        pool_register_stream(device, s->stream);
        s->prof.stream = s->stream;
        slots_.push_back(std::move(s));
    }
    for (auto& s : slots_) {
        Slot* sp = s.get();
        sp->thread = std::thread([this, sp] { worker(sp); });
    }
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stopping_ = true;
    }
    cv_.notify_all();
    for (auto& s : slots_) {
        if (s->thread.joinable()) s->thread.join();
        (void)hipStreamSynchronize(s->stream);
    }
    cache_.clear();
    for (auto& s : slots_) {
        s->fw = FrameWork();
        s->bw = BankWorkspace();
        s->tmp.release();
        pool_unregister_stream(device_, s->stream);
        (void)hipStreamDestroy(s->stream);
    }
}

int64_t Server::submit(const ldh_params& p, const float4* src, float4* dst, int w, int h) {
    auto r = std::make_unique<ServedRequest>();
    r->params = p;
    r->src = src;
    r->dst = dst;
    r->w = w;
    r->h = h;
    r->t_submit = now_ms();
    std::lock_guard<std::mutex> lk(mu_);
    r->ticket = next_ticket_++;
    ServedRequest* raw = r.get();
    requests_[raw->ticket] = std::move(r);
    queue_.push_back(raw);
    stats_.submitted += 1;
    cv_.notify_all();
    return raw->ticket;
}

bool Server::poll(int64_t ticket, ServedRequest* out, bool wait) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = requests_.find(ticket);
    if (it == requests_.end()) throw Error{LDH_ERR_INVALID_ARGUMENT, "unknown-ticket"};
    if (wait) cv_.wait(lk, [&] { return it->second->done; });
    if (!it->second->done) return false;
    *out = *it->second;
    requests_.erase(it);
    return true;
}

ldh_server_stats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    ldh_server_stats s = stats_;
    s.resident_bytes = resident_bytes_locked();
    s.cached_banks = int32_t(cache_.size());
    return s;
}

size_t Server::resident_bytes_locked() const {
    size_t b = 0;
    for (const auto& e : cache_) b += e->bytes();
    return b;
}

// Segmented LRU: unpinned entries that one request has used so far go first, least recently
// used first, then the others in the same order. A stream of lenses seen once then cycles through
// the first segment and leaves the lenses that recur resident.
bool Server::evict_locked(size_t need) {
    size_t resident = resident_bytes_locked();
    for (int pass = 0; pass < 2; ++pass)
        for (auto it = cache_.end(); resident + need > budget_ && it != cache_.begin();) {
            --it;
            if ((*it)->pins == 0 && (pass == 1 || (*it)->uses <= 1)) {
                resident -= (*it)->bytes();
                it = cache_.erase(it);
                stats_.evictions += 1;
            }
        }
    return resident + need <= budget_;
}

constexpr int kAdmitWindow = 8;  // queued requests a free slot considers
constexpr int kMaxPassed = 16;   // admissions ahead of the oldest request before it goes first

// Picks the next request for slot s among the oldest queued ones (kAdmitWindow of them; only the
// oldest once later ones have gone ahead of it kMaxPassed times). A request whose bank or spectra
// another slot is still building is left queued: a slot taking it would only wait. Of the others,
// the oldest one whose cache needs fit the budget as it stands goes first; failing that, the
// oldest one that fits after evicting unpinned entries (or any, when nothing is in flight). Pins
// the request's cache entry, creating it, with the caller as builder, when absent.
ServedRequest* Server::admit(Slot* s, std::shared_ptr<CacheEntry>* entry, bool* builder) {
    std::unique_lock<std::mutex> lk(mu_);
    while (true) {
        if (stopping_) return nullptr;
        bool in_flight = false;
        for (const auto& c : cache_) in_flight = in_flight || c->pins > 0;
        const size_t resident = resident_bytes_locked();
        const int window = (!queue_.empty() && queue_.front()->passed >= kMaxPassed) ? 1 : kAdmitWindow;
        struct Candidate {
            std::deque<ServedRequest*>::iterator it;
            std::shared_ptr<CacheEntry> e;
            size_t need;
        };
        std::vector<Candidate> cands;
        int scanned = 0;
        for (auto q = queue_.begin(); q != queue_.end() && scanned < window; ++q, ++scanned) {
            ServedRequest* r = *q;
            const BankKey key = served_key(r->params);
            std::shared_ptr<CacheEntry> e;
            for (const auto& c : cache_)
                if (c->key == key) {
                    e = c;
                    break;
                }
            if (e && !e->ready) continue;  // its bank is being built
            size_t need = e ? 0 : bank_bytes_estimate(r->params);
            const int nk = cached_kernel_count(r->params);
            if (nk > 0 && (!e || nk == 1 || (e->has_combined && e->combined_style == with_style(r->params)))) {
                const int support = entry_support(e.get(), r->params);
                int gw = 0, gh = 0;
                served_frame_size(r->params, r->w, r->h, gw, gh);
                const Grid g = make_grid(gw, gh, support);
                const SpectraKey sk{g.pw, g.ph, nk, r->params.spectra_fp16 != 0 ? 1 : 0};
                std::shared_ptr<CacheEntry::Spectra> spec;
                if (e)
                    for (const auto& sp : e->spectra)
                        if (sp->key == sk) spec = sp;
                if (spec && !spec->ready) continue;  // its spectra are being built
                if (!spec) need += spectra_bytes(g, nk, r->params.spectra_fp16 != 0) * size_t(spectra_sets(r->params));
            }
            cands.push_back({q, e, need});
        }
        const Candidate* take = nullptr;
        for (const auto& c : cands)
            if (c.need == 0 || resident + c.need <= budget_) {
                take = &c;
                break;
            }
        if (!take)
            for (const auto& c : cands)
                if (evict_locked(c.need) || !in_flight) {
                    take = &c;
                    break;
                }
        if (take) {
            ServedRequest* r = *take->it;
            std::shared_ptr<CacheEntry> e = take->e;
            if (take->it != queue_.begin()) queue_.front()->passed += 1;
            queue_.erase(take->it);
            *builder = !e;
            if (!e) {
                e = std::make_shared<CacheEntry>();
                e->key = served_key(r->params);
                cache_.push_front(e);
                stats_.bank_misses += 1;
            } else {
                for (auto it = cache_.begin(); it != cache_.end(); ++it)
                    if (*it == e) {
                        cache_.splice(cache_.begin(), cache_, it);  // most recently used first
                        break;
                    }
                stats_.bank_hits += 1;
                r->bank_hit = true;
            }
            e->pins += 1;
            e->uses += 1;
            e->last_use = ++clock_;
            *entry = e;
            r->t_start = now_ms();
            (void)s;
            return r;
        }
        cv_.wait(lk);
    }
}

void Server::worker(Slot* s) {
    (void)hipSetDevice(device_);
    while (true) {
        std::shared_ptr<CacheEntry> e;
        bool builder = false;
        ServedRequest* r = admit(s, &e, &builder);
        if (!r) return;
        std::unique_lock<std::mutex> prof_lock(s->prof_mu);
        ProfilerBinding bind(s->prof);
        try {
            std::string why = unsupported_reason(r->params, true);
            // the served spectra are kept for the frame's own grid
            if (!why.empty()) throw Error{LDH_ERR_UNSUPPORTED, why};
            if (builder) {
                try {
                    debug_failure_point("served-bank");
                    std::shared_ptr<DeviceBank> bank;
                    std::vector<std::shared_ptr<DeviceBank>> zones;
                    if (has_field_phase(r->params))
                        zones = build_banks_batched(s->stream, fft_, s->bw, zone_requests(r->params), LDH_BATCH_GROUPED,
                                                    nullptr);
                    else
                        bank = std::shared_ptr<DeviceBank>(
                            build_bank(s->stream, fft_, s->bw, r->params, served_sets(r->params)).release());
                    if (spectral_fullkernel(r->params)) {
                        // the precombined RGB kernels of the bank, or of each zone (zone z at
                        // z * 6 * K * K; the zones share their kernel size)
                        std::vector<const DeviceBank*> sets;
                        if (bank) sets.push_back(bank.get());
                        for (const auto& z : zones) sets.push_back(z.get());
                        const bool sty = with_style(r->params);
                        const size_t per = size_t(sets.front()->kernel) * sets.front()->kernel * (sty ? 6 : 3);
                        e->combined.reset(sizeof(float) * per * sets.size());
                        for (size_t k = 0; k < sets.size(); ++k) {
                            const DeviceBank& b = *sets[k];
                            const SpectrumConfig cfg = build_spectrum_config(r->params, b.wavelengths);
                            combine_kernels(s->stream, b.kernel_ptr(LDH_KERNEL_FULL, 0), b.scale_ptr(LDH_KERNEL_FULL, 0),
                                            b.bins, b.kernel, cfg, sty, e->combined.as<float>() + per * k);
                        }
                        e->combined_style = sty;
                        e->has_combined = true;
                    }
                    LDH_HIP(hipStreamSynchronize(s->stream), "served-bank");
                    std::lock_guard<std::mutex> lk(mu_);
                    e->bank = bank;
                    e->zones = zones;
                    e->ready = true;
                } catch (const Error& err) {
                    std::lock_guard<std::mutex> lk(mu_);
                    e->failed = true;
                    e->error = err.tag;
                    cache_.remove(e);
                    cv_.notify_all();
                    throw;
                }
                cv_.notify_all();
            } else {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return e->ready || e->failed || stopping_; });
                if (!e->ready) throw Error{LDH_ERR_DEVICE, e->failed ? e->error : "server-stopping"};
            }
            // The spectra record of this entry for key sk: found (waited for until it is ready) or
            // created and built here by `build`.
            auto spectra_for = [&](const SpectraKey& sk,
                                   const std::function<void(CacheEntry::Spectra&)>& build) {
                std::shared_ptr<CacheEntry::Spectra> spec;
                bool mine = false;
                {
                    std::unique_lock<std::mutex> lk(mu_);
                    for (auto& sp : e->spectra)
                        if (sp->key == sk) spec = sp;
                    if (!spec) {
                        spec = std::make_shared<CacheEntry::Spectra>();
                        spec->key = sk;
                        e->spectra.push_back(spec);
                        mine = true;
                        stats_.spectra_misses += 1;
                    } else {
                        stats_.spectra_hits += 1;
                        r->spectra_hit = true;
                        cv_.wait(lk, [&] { return spec->ready || spec->failed || stopping_; });
                        if (!spec->ready) throw Error{LDH_ERR_DEVICE, spec->failed ? spec->error : "server-stopping"};
                    }
                }
                if (mine) {
                    try {
                        debug_failure_point("served-spectra");
                        build(*spec);
                        LDH_HIP(hipStreamSynchronize(s->stream), "served-spectra");
                    } catch (const Error& err) {
                        // the record leaves the entry (a later request builds it again); the
                        // requests already waiting on it take this error
                        std::lock_guard<std::mutex> lk(mu_);
                        spec->failed = true;
                        spec->error = err.tag;
                        e->spectra.erase(std::remove(e->spectra.begin(), e->spectra.end(), spec), e->spectra.end());
                        cv_.notify_all();
                        throw;
                    }
                    std::lock_guard<std::mutex> lk(mu_);
                    spec->ready = true;
                    cv_.notify_all();
                }
                return spec;
            };
            const int nk = cached_kernel_count(r->params);
            // precombined spectra need the entry's combined kernels, which exist when the request
            // that built it had a spectral full-kernel look with the same style
            const bool cacheable = nk > 0 && (nk == 1 || (e->has_combined && e->combined_style == with_style(r->params)));
            const bool fp16 = r->params.spectra_fp16 != 0;
            // a resolution-aware frame is convolved at its working size
            int gw = 0, gh = 0;
            served_frame_size(r->params, r->w, r->h, gw, gh);
            if (!e->zones.empty()) {
                if (is_static_view(r->params.debug_view)) {
                    // zone (1, 1) holds the global terms; the edge views show zone (2, 2)
                    const bool edge =
                        r->params.debug_view == LDH_VIEW_PHASE_EDGE || r->params.debug_view == LDH_VIEW_FIELD_PSF;
                    const ldh_params q = edge ? field_zone_params(r->params, 2, 2) : r->params;
                    render_static_view(s->stream, s->fw, q, r->params.debug_view, *e->zones[edge ? 8 : 4], r->dst,
                                       r->w, r->h);
                } else {
                    std::vector<const DeviceBank*> zb;
                    for (const auto& z : e->zones) zb.push_back(z.get());
                    // a spectral full-kernel look convolves each zone with its pre-combined RGB kernels
                    ldh_params p = r->params;
                    const bool pre = spectral_fullkernel(p);
                    p.conv_mode = pre ? LDH_CONV_PRECOMBINED : LDH_CONV_PER_BIN;
                    std::shared_ptr<CacheEntry::Spectra> spec;
                    std::vector<const KernelSpectra*> zc;
                    if (cacheable) {
                        // every zone's kernel spectra on the zones' common grid
                        const Grid g = zone_grid(gw, gh, zb);
                        spec = spectra_for(SpectraKey{g.pw, g.ph, nk, fp16 ? 1 : 0}, [&](CacheEntry::Spectra& sp) {
                            sp.zones.resize(size_t(kFieldZones));
                            for (int z = 0; z < kFieldZones; ++z) {
                                const DeviceBank& zbank = *zb[size_t(z)];
                                const size_t kk = size_t(zbank.kernel) * zbank.kernel;
                                std::vector<KernelRef> ks;
                                if (nk == 1) {
                                    ks.push_back({zbank.kernel_ptr(LDH_KERNEL_FULL, 0), zbank.scale_ptr(LDH_KERNEL_FULL, 0)});
                                } else {
                                    const float* comb = e->combined.as<float>() + kk * nk * size_t(z);
                                    for (int i = 0; i < nk; ++i) ks.push_back({comb + kk * i, nullptr});
                                }
                                // pre-combined kernels as half spectra (the real-input path)
                                build_kernel_spectra(s->stream, fft_, ks, zbank.kernel, zbank.support, g, fp16,
                                                     sp.zones[size_t(z)], s->tmp, nk > 1);
                            }
                        });
                        for (const auto& z : spec->zones) zc.push_back(&z);
                    }
                    render_frame_zones(s->stream, fft_, s->fw, p, zb, r->src, r->dst, r->w, r->h,
                                       pre ? EffectMode::Precombined : EffectMode::PerBin, spec ? &zc : nullptr);
                }
            } else if (is_static_view(r->params.debug_view)) {
                // drawn from the resident bank or the pupil
                render_static_view(s->stream, s->fw, r->params, r->params.debug_view, *e->bank, r->dst, r->w, r->h);
            } else {
                const DeviceBank& bank = *e->bank;
                std::shared_ptr<CacheEntry::Spectra> spec;
                if (cacheable) {
                    const Grid g = make_grid(gw, gh, bank.support);
                    spec = spectra_for(SpectraKey{g.pw, g.ph, nk, fp16 ? 1 : 0}, [&](CacheEntry::Spectra& sp) {
                        std::vector<KernelRef> ks;
                        if (nk == 1) {
                            ks.push_back({bank.kernel_ptr(LDH_KERNEL_FULL, 0), bank.scale_ptr(LDH_KERNEL_FULL, 0)});
                        } else {
                            const size_t kk = size_t(bank.kernel) * bank.kernel;
                            for (int i = 0; i < nk; ++i) ks.push_back({e->combined.as<float>() + kk * i, nullptr});
                        }
                        build_kernel_spectra(s->stream, fft_, ks, bank.kernel, bank.support, g, fp16, sp.spec, s->tmp,
                                             nk > 1);
                    });
                }
                ldh_params p = r->params;
                if (spec && nk > 1) {
                    p.conv_mode = LDH_CONV_PRECOMBINED;
                    s->fw.combined.reset(0);  // cached spectra carry the combined kernels
                }
                render_frame(s->stream, fft_, s->fw, p, bank, r->src, r->dst, r->w, r->h,
                             spec && nk > 1 ? EffectMode::Precombined : EffectMode::PerBin,
                             spec ? &spec->spec : nullptr);
            }
            LDH_HIP(hipStreamSynchronize(s->stream), "served-frame");
            r->status = LDH_OK;
        } catch (const Error& err) {
            r->status = err.status;
            r->error = err.tag;
            (void)hipGetLastError();
        } catch (...) {
            r->status = LDH_ERR_DEVICE;
            r->error = "internal-error";
        }
        prof_lock.unlock();
        std::lock_guard<std::mutex> lk(mu_);
        if (builder && !e->ready && !e->failed) {
            // the entry this request was to build will never be ready: release its waiters
            e->failed = true;
            e->error = r->error;
            cache_.remove(e);
        }
        r->t_done = now_ms();
        r->done = true;
        e->pins -= 1;
        stats_.completed += 1;
        cv_.notify_all();
    }
}

}  // namespace ldh

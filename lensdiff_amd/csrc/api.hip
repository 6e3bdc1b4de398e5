// The extern "C" boundary of liblensdiff_hip.so (include/lensdiff_hip.h).
//
// Contexts own one HIP stream, the FFT tables, the bank workspace and frame scratch, and a
// single-entry bank cache keyed like LensDiffPsfBankCache (LensDiffTypes.h:216-290, rebuilt on
// any key change, LensDiffCpuReference.cpp:1704-1706). Every entry point is serialised on the
// context mutex and converts failures to a status plus a kebab-case tag.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "autotune.h"
#include "bank.h"
#include "fft_io.h"
#include "frame.h"
#include "host_params.h"
#include "profiler.h"
#include "server.h"

namespace ldh {
std::mutex g_debug_mu;
std::map<std::string, int> g_debug_failures;

void debug_failure_point(const char* site) {
    std::lock_guard<std::mutex> lk(g_debug_mu);
    auto it = g_debug_failures.find(site);
    if (it == g_debug_failures.end()) return;
    if (--it->second <= 0) g_debug_failures.erase(it);
    throw Error{LDH_ERR_DEVICE, std::string("injected-") + site};
}
}  // namespace ldh

struct ldh_bank {
    std::shared_ptr<ldh::DeviceBank> bank;
    ldh::BankKey key;
    int refs = 1;
};

struct ldh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::string last_error;
    uint64_t budget = 0;
    ldh::FftTables fft;
    ldh::BankWorkspace bw;
    ldh::BatchLanes lanes;  // extra branches of graph batches
    ldh::FrameWork fw;
    ldh::Profiler prof;
    // single-entry cache of the plain path
    bool cached = false;
    ldh::BankKey cached_key;
    std::shared_ptr<ldh::DeviceBank> cached_bank;
    // the nine zone banks of the last field-phase render, keyed by their zones' bank keys
    std::vector<ldh::BankKey> zone_keys;
    std::vector<std::shared_ptr<ldh::DeviceBank>> zone_banks;
    // host staging
    float* pinned = nullptr;
    size_t pinned_bytes = 0;
    ldh::DeviceBuffer dsrc, ddst, scratch, scratch2;
    // served request loop (created by ldh_server_start)
    std::unique_ptr<ldh::Server> server;
};

namespace {

template <class F>
ldh_status guarded(ldh_context* ctx, F&& f) {
    if (!ctx) return LDH_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    try {
        if (hipSetDevice(ctx->device) != hipSuccess) throw ldh::Error{LDH_ERR_DEVICE, "set-device-failed"};
        ctx->last_error.clear();
        ldh::ProfilerBinding bind(ctx->prof);
        f();
        return LDH_OK;
    } catch (const ldh::Error& e) {
        ctx->last_error = e.tag;
        return e.status;
    } catch (const std::bad_alloc&) {
        ctx->last_error = "host-allocation-failed";
        return LDH_ERR_OUT_OF_MEMORY;
    } catch (...) {
        ctx->last_error = "internal-error";
        return LDH_ERR_DEVICE;
    }
}

void require(bool ok, ldh_status s, const char* tag) {
    if (!ok) throw ldh::Error{s, tag};
}

void sync(ldh_context* ctx, const char* tag) { LDH_HIP(hipStreamSynchronize(ctx->stream), tag); }

void h2d(ldh_context* ctx, void* dst, const void* src, size_t bytes) {
    LDH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "copy-to-device");
}
void d2h(ldh_context* ctx, void* dst, const void* src, size_t bytes) {
    LDH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "copy-to-host");
    sync(ctx, "copy-to-host");
}

float* pinned(ldh_context* ctx, size_t bytes) {
    if (bytes > ctx->pinned_bytes) {
        if (ctx->pinned) (void)hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
        void* p = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            throw ldh::Error{LDH_ERR_OUT_OF_MEMORY, "pinned-allocation-failed"};
        }
        ctx->pinned = static_cast<float*>(p);
        ctx->pinned_bytes = bytes;
    }
    return ctx->pinned;
}

// The bank of p holding at least its first `sets` kernel sets: the cached one, or built in its place.
std::shared_ptr<ldh::DeviceBank> ensure_bank(ldh_context* ctx, const ldh_params& p, int sets) {
    const ldh::BankKey key = ldh::make_bank_key(p);
    if (ctx->cached && ctx->cached_key == key && ctx->cached_bank->sets >= sets) return ctx->cached_bank;
    ldh::ProfileScope ps("bank-build", 0.0);
    ctx->cached = false;
    ctx->cached_bank.reset();
    std::shared_ptr<ldh::DeviceBank> b(ldh::build_bank(ctx->stream, ctx->fft, ctx->bw, p, sets).release());
    ctx->cached_bank = b;
    ctx->cached_key = key;
    ctx->cached = true;
    return b;
}

// The zone banks of a field-phase render (EnsureLensDiffPsfBank's zone caches,
// LensDiffCpuReference.cpp:1756-1790), built together in shared launches (the zones have equal
// sizes) and kept until the zones' parameters change.
std::vector<const ldh::DeviceBank*> ensure_zone_banks(ldh_context* ctx, const ldh_params& p) {
    std::vector<ldh_params> zones;
    std::vector<ldh::BankKey> keys;
    for (int z = 0; z < ldh::kFieldZones; ++z) {
        zones.push_back(ldh::field_zone_params(p, z % 3, z / 3));
        // a zone's bank is built from its resolved global terms alone (buildPsfBankGlobalOnly)
        zones.back().phase_field_strength = 0.0;
        keys.push_back(ldh::make_bank_key(zones.back()));
    }
    if (keys != ctx->zone_keys || ctx->zone_banks.size() != size_t(ldh::kFieldZones)) {
        ldh::ProfileScope ps("bank-build", 0.0);
        ctx->zone_keys.clear();
        ctx->zone_banks.clear();
        ctx->zone_banks = ldh::build_banks_batched(ctx->stream, ctx->fft, ctx->bw, zones, LDH_BATCH_GROUPED, nullptr);
        ctx->zone_keys = keys;
    }
    std::vector<const ldh::DeviceBank*> out;
    for (const auto& b : ctx->zone_banks) out.push_back(b.get());
    return out;
}

// Renders the frame at d_src into d_dst with the bank or, for field-dependent phase, the zone banks.
void render_with_banks(ldh_context* ctx, const ldh_params& p, const float4* d_src, float4* d_dst, int w, int h) {
    const ldh::EffectMode mode =
        p.conv_mode == LDH_CONV_PRECOMBINED ? ldh::EffectMode::Precombined : ldh::EffectMode::PerBin;
    if (ldh::is_static_view(p.debug_view)) {
        // the global bank, or the edge zone (2, 2) for the phase-edge and field-PSF views of a
        // field-dependent phase (BuildLensDiffStaticDebugRgba's edgeZone)
        const bool edge = ldh::has_field_phase(p) &&
                          (p.debug_view == LDH_VIEW_PHASE_EDGE || p.debug_view == LDH_VIEW_FIELD_PSF);
        ldh_params q = edge ? ldh::field_zone_params(p, 2, 2) : p;
        q.phase_field_strength = 0.0;
        auto bank = ensure_bank(ctx, q, 1);  // the static views read the full kernels only
        ldh::render_static_view(ctx->stream, ctx->fw, q, p.debug_view, *bank, d_dst, w, h);
        return;
    }
    if (ldh::has_field_phase(p)) {
        const auto banks = ensure_zone_banks(ctx, p);
        ldh::ProfileScope ps("frame", 0.0);
        ldh::render_frame_zones(ctx->stream, ctx->fft, ctx->fw, p, banks, d_src, d_dst, w, h, mode);
        return;
    }
    auto bank = ensure_bank(ctx, p, ldh::kernel_sets_needed(p));
    ldh::ProfileScope ps("frame", 0.0);
    ldh::render_frame(ctx->stream, ctx->fft, ctx->fw, p, *bank, d_src, d_dst, w, h, mode, nullptr);
}

void check_params(const ldh_params* p, bool allow_field_zones = false) {
    require(p != nullptr, LDH_ERR_INVALID_ARGUMENT, "missing-params");
    const std::string why = ldh::unsupported_reason(*p, allow_field_zones);
    if (!why.empty()) throw ldh::Error{LDH_ERR_UNSUPPORTED, why};
}

__global__ void permute_rows_kernel(const float2* __restrict__ in, float2* __restrict__ out, int w, int h,
                                    ldh::ColSplit sp, int to_storage) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(w) * h) return;
    const int x = int(i % w), k = int(i / w);
    const int sr = ldh::col_storage_row(sp, k);
    if (to_storage) out[int64_t(sr) * w + x] = in[i];
    else out[i] = in[int64_t(sr) * w + x];
}

__global__ void scale_c64_kernel(float2* __restrict__ d, int64_t n, float s) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) d[i] = make_float2(d[i].x * s, d[i].y * s);
}

}  // namespace

extern "C" {

const char* ldh_status_string(ldh_status s) {
    switch (s) {
        case LDH_OK: return "ok";
        case LDH_ERR_INVALID_ARGUMENT: return "invalid-argument";
        case LDH_ERR_MISSING_BUFFER: return "missing-source-or-destination-buffer";
        case LDH_ERR_INVALID_BOUNDS: return "invalid-source-bounds";
        case LDH_ERR_UNSUPPORTED: return "unsupported";
        case LDH_ERR_OUT_OF_MEMORY: return "out-of-memory";
        case LDH_ERR_DEVICE: return "device-error";
        case LDH_ERR_NOT_READY: return "not-ready";
        case LDH_ERR_TOO_LARGE: return "too-large";
        default: return "unknown-status";
    }
}

ldh_status ldh_context_create(int32_t device, uint64_t budget, ldh_context** out) {
    if (!out) return LDH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        (void)hipGetLastError();
        return LDH_ERR_DEVICE;
    }
    auto* ctx = new (std::nothrow) ldh_context();
    if (!ctx) return LDH_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    ctx->budget = budget ? budget : (uint64_t(64) << 30);
    ldh_status st = guarded(ctx, [&] {
        LDH_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "stream-create");
        ldh::pool_register_stream(ctx->device, ctx->stream);
        ldh::noalloc_scratch_init(ctx->device);
        ctx->prof.stream = ctx->stream;
        ctx->fft.init(ctx->stream);
    });
    if (st != LDH_OK) {
        ldh_context_destroy(ctx);
        return st;
    }
    *out = ctx;
    return LDH_OK;
}

ldh_status ldh_context_destroy(ldh_context* ctx) {
    if (!ctx) return LDH_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipSetDevice(ctx->device);
        ctx->server.reset();
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        if (ctx->pinned) (void)hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
    }
    hipStream_t s = ctx->stream;
    const int device = ctx->device;
    delete ctx;  // device buffers go back to the pool here
    if (s) {
        ldh::pool_unregister_stream(device, s);
        (void)hipStreamDestroy(s);
    }
    ldh::pool_trim(device);
    return LDH_OK;
}

const char* ldh_last_error(const ldh_context* ctx) { return ctx ? ctx->last_error.c_str() : "no-context"; }

ldh_status ldh_synchronize(ldh_context* ctx) {
    return guarded(ctx, [&] { sync(ctx, "synchronize"); });
}

ldh_status ldh_resolve_sizes(const ldh_params* p, ldh_sizes* out) {
    if (!p || !out) return LDH_ERR_INVALID_ARGUMENT;
    const ldh::Sizes s = ldh::resolve_sizes(*p);
    out->pupil_size = s.pupil;
    out->raw_psf_size = s.raw;
    out->max_kernel_radius_px = s.max_radius;
    out->optics_short_side_px = s.optics_short_side;
    out->diffraction_scale_px = s.diffraction_scale_px;
    return LDH_OK;
}

ldh_status ldh_pupil_amplitude(ldh_context* ctx, const ldh_params* p, float* out) {
    return guarded(ctx, [&] {
        require(p && out, LDH_ERR_INVALID_ARGUMENT, "missing-argument");
        require(p->aperture_mode != LDH_APERTURE_CUSTOM, LDH_ERR_UNSUPPORTED, "custom-aperture-not-supported");
        const ldh::PupilGeom g = ldh::make_pupil_geom(*p);
        const size_t bytes = sizeof(float) * g.size * g.size;
        ctx->scratch.reset(bytes);
        ldh::pupil_amplitude(ctx->stream, g, ctx->scratch.as<float>(), false);
        d2h(ctx, out, ctx->scratch.get(), bytes);
    });
}

ldh_status ldh_pupil_phase(ldh_context* ctx, const ldh_params* p, float* out) {
    return guarded(ctx, [&] {
        require(p && out, LDH_ERR_INVALID_ARGUMENT, "missing-argument");
        const int n = ldh::resolve_sizes(*p).pupil;
        const ldh::PhaseCoef c = ldh::make_phase_coef(*p);
        ctx->scratch2.reset(sizeof(c));
        h2d(ctx, ctx->scratch2.get(), &c, sizeof(c));
        const size_t bytes = sizeof(float) * n * n;
        ctx->scratch.reset(bytes);
        ldh::pupil_phase(ctx->stream, ctx->scratch2.as<ldh::PhaseCoef>(), n, 1, ctx->scratch.as<float>(), false);
        d2h(ctx, out, ctx->scratch.get(), bytes);
    });
}

ldh_status ldh_raw_psf(ldh_context* ctx, const float* amp, const float* phase, int32_t P, int32_t R, float* out) {
    return guarded(ctx, [&] {
        require(amp && out && P > 0 && R >= P && R >= 16, LDH_ERR_INVALID_ARGUMENT, "invalid-psf-arguments");
        require(R <= ldh::kMaxRawPsf, LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
        ldh::DeviceBuffer damp(sizeof(float) * P * P), dph(sizeof(float) * P * P), dout(sizeof(float) * size_t(R) * R),
            dtot(sizeof(float));
        // the transform takes the pupil transposed ([x][y])
        auto transposed = [P](const float* m) {
            std::vector<float> t(size_t(P) * P);
            for (int y = 0; y < P; ++y)
                for (int x = 0; x < P; ++x) t[size_t(x) * P + y] = m[size_t(y) * P + x];
            return t;
        };
        // host copies stay alive until the final synchronising download
        const std::vector<float> amp_t = transposed(amp);
        const std::vector<float> phase_t = phase ? transposed(phase) : std::vector<float>();
        h2d(ctx, damp.get(), amp_t.data(), sizeof(float) * P * P);
        if (phase) {
            h2d(ctx, dph.get(), phase_t.data(), sizeof(float) * P * P);
        } else {
            LDH_HIP(hipMemsetAsync(dph.get(), 0, sizeof(float) * P * P, ctx->stream), "raw-psf");
        }
        ctx->bw.plan.prepare(ctx->stream, P, R);
        ctx->bw.plan.raw_psfs(ctx->stream, ctx->fft, damp.as<float>(), dph.as<float>(), 1, dout.as<float>(),
                              dtot.as<float>(), ctx->bw.seqsum, ctx->bw.psf);
        d2h(ctx, out, dout.get(), sizeof(float) * size_t(R) * R);
    });
}

ldh_status ldh_raw_psf_window(ldh_context* ctx, const float* amp, const float* phase, int32_t P, int32_t R,
                              int32_t h, float* out) {
    return guarded(ctx, [&] {
        require(amp && out && P > 0 && R >= P && R >= 16 && h > 0, LDH_ERR_INVALID_ARGUMENT, "invalid-psf-arguments");
        require(R <= ldh::kMaxRawPsf, LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
        ldh::DeviceBuffer damp(sizeof(float) * P * P), dph(sizeof(float) * P * P);
        auto transposed = [P](const float* m) {
            std::vector<float> t(size_t(P) * P);
            for (int y = 0; y < P; ++y)
                for (int x = 0; x < P; ++x) t[size_t(x) * P + y] = m[size_t(y) * P + x];
            return t;
        };
        const std::vector<float> amp_t = transposed(amp);
        const std::vector<float> phase_t = phase ? transposed(phase) : std::vector<float>();
        h2d(ctx, damp.get(), amp_t.data(), sizeof(float) * P * P);
        if (phase) {
            h2d(ctx, dph.get(), phase_t.data(), sizeof(float) * P * P);
        } else {
            LDH_HIP(hipMemsetAsync(dph.get(), 0, sizeof(float) * P * P, ctx->stream), "raw-psf-window");
        }
        ldh::PsfPlan& plan = ctx->bw.plan;
        plan.prepare(ctx->stream, P, R);
        require(plan.prepare_window(ctx->stream, h), LDH_ERR_UNSUPPORTED, "raw-psf-window-not-applicable");
        const size_t n = size_t(2 * h + 1);
        ldh::DeviceBuffer dout(sizeof(float) * n * n);
        plan.window_psfs(ctx->stream, ctx->fft, damp.as<float>(), dph.as<float>(), 1, dout.as<float>(), ctx->bw.psf);
        d2h(ctx, out, dout.get(), sizeof(float) * n * n);
    });
}

ldh_status ldh_first_minimum_radius(ldh_context* ctx, const float* psf, int32_t R, float* out) {
    return guarded(ctx, [&] {
        require(psf && out && R >= 4, LDH_ERR_INVALID_ARGUMENT, "invalid-argument");
        ldh::DeviceBuffer d(sizeof(float) * size_t(R) * R);
        h2d(ctx, d.get(), psf, sizeof(float) * size_t(R) * R);
        *out = ldh::first_minimum_from_profile(ldh::radial_profile(ctx->stream, ctx->bw, d.as<float>(), R), R);
    });
}

ldh_status ldh_resample_kernel(ldh_context* ctx, const float* psf, int32_t R, float scale, int32_t support, float* out) {
    return guarded(ctx, [&] {
        require(psf && out && R >= 4 && support >= 0, LDH_ERR_INVALID_ARGUMENT, "invalid-argument");
        const int K = 2 * support + 1;
        ldh::DeviceBuffer d(sizeof(float) * size_t(R) * R), k(sizeof(float) * size_t(K) * K), sc(sizeof(float)),
            tot(sizeof(float));
        h2d(ctx, d.get(), psf, sizeof(float) * size_t(R) * R);
        h2d(ctx, sc.get(), &scale, sizeof(float));
        ldh::resample_kernels(ctx->stream, d.as<float>(), R, sc.as<float>(), nullptr, support, 1, k.as<float>());
        ctx->bw.seqsum.normalize(ctx->stream, k.as<float>(), int64_t(K) * K, int64_t(K) * K, 1, tot.as<float>());
        d2h(ctx, out, k.get(), sizeof(float) * size_t(K) * K);
    });
}

ldh_status ldh_shape_kernels(ldh_context* ctx, const ldh_params* p, const float* base, int32_t K, float* full,
                             float* core, float* structure, int32_t* out_size) {
    return guarded(ctx, [&] {
        require(p && base && full && core && structure && out_size && K >= 3 && (K & 1), LDH_ERR_INVALID_ARGUMENT,
                "invalid-argument");
        const int max_r = ldh::resolve_sizes(*p).max_radius;
        const int64_t kk = int64_t(K) * K;
        ldh::DeviceBuffer db(sizeof(float) * kk), d3(sizeof(float) * kk * 3), dr(sizeof(int)), dsc(sizeof(float) * 3);
        h2d(ctx, db.get(), base, sizeof(float) * kk);
        ldh::shape_kernels(ctx->stream, ctx->bw, db.as<float>(), nullptr, K, 1, p->anisotropy_emphasis, max_r, d3.as<float>(),
                           dr.as<int>(), dsc.as<float>());
        ldh::apply_kernel_scales(ctx->stream, d3.as<float>(), K, 1, dr.as<int>(), dsc.as<float>());
        std::vector<float> h(size_t(kk) * 3);
        int r = 0;
        LDH_HIP(hipMemcpyAsync(&r, dr.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream), "shape");
        d2h(ctx, h.data(), d3.get(), sizeof(float) * kk * 3);
        const int rc = std::min(std::max(r, 1), K / 2);
        const int n = 2 * rc + 1, c = K / 2;
        float* outs[3] = {full, core, structure};
        for (int w = 0; w < 3; ++w)
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x)
                    outs[w][size_t(y) * n + x] = h[size_t(w) * kk + size_t(c - rc + y) * K + size_t(c - rc + x)];
        *out_size = n;
    });
}

ldh_status ldh_sequential_sum(ldh_context* ctx, const float* x, int64_t n, float* out) {
    return guarded(ctx, [&] {
        require(x && out && n >= 0, LDH_ERR_INVALID_ARGUMENT, "invalid-argument");
        if (n == 0) { *out = 0.0f; return; }
        ldh::DeviceBuffer d(sizeof(float) * n), t(sizeof(float));
        h2d(ctx, d.get(), x, sizeof(float) * n);
        ctx->bw.seqsum.run(ctx->stream, d.as<float>(), n, n, 1, t.as<float>());
        d2h(ctx, out, t.get(), sizeof(float));
    });
}

ldh_status ldh_sequential_sums(ldh_context* ctx, const float* x, int64_t n, int32_t count, int32_t block,
                               float* out) {
    return guarded(ctx, [&] {
        require(x && out && n >= 0 && count > 0, LDH_ERR_INVALID_ARGUMENT, "invalid-argument");
        require(block == 0 || (block >= ldh::kSeqSumMinBlock && block <= ldh::kSeqSumBlock), LDH_ERR_INVALID_ARGUMENT,
                "seqsum-block");
        if (n == 0) {
            for (int i = 0; i < count; ++i) out[i] = 0.0f;
            return;
        }
        ldh::DeviceBuffer d(sizeof(float) * size_t(n) * count), t(sizeof(float) * count);
        h2d(ctx, d.get(), x, sizeof(float) * size_t(n) * count);
        ldh::SeqSegments g;
        g.array[0] = d.as<float>();
        g.per_array = count;
        g.stride = n;
        ctx->bw.seqsum.run(ctx->stream, g, n, count, t.as<float>(), nullptr, block == 0 ? ldh::kSeqSumBlock : block);
        d2h(ctx, out, t.get(), sizeof(float) * count);
    });
}

ldh_status ldh_sequential_sums_noalloc(ldh_context* ctx, const float* x, int64_t n, int32_t count, float* totals,
                                       void* stream) {
    return guarded(ctx, [&] {
        require(x && totals && n >= 0 && count > 0, LDH_ERR_INVALID_ARGUMENT, "invalid-argument");
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        ldh::sequential_sums_noalloc(s, ctx->device, x, n, count, totals);
    });
}

ldh_status ldh_fft2d(ldh_context* ctx, float* data, int32_t w, int32_t h, int32_t inverse) {
    return guarded(ctx, [&] {
        require(data && ldh::grid_length_ok(w) && w >= 16 && ldh::grid_length_ok(h), LDH_ERR_INVALID_ARGUMENT,
                "fft-size-unsupported");
        const int64_t n = int64_t(w) * h;
        ldh::DeviceBuffer a(sizeof(float2) * n), b(sizeof(float2) * n);
        h2d(ctx, a.get(), data, sizeof(float2) * n);
        const dim3 grid(unsigned(ldh::div_up(n, 256)));
        const ldh::ColSplit sp = ldh::col_plan(ctx->stream, ctx->fft.twiddle.as<float2>(), w, h);
        if (!inverse) {
            ldh::fft2d_forward(ctx->stream, ctx->fft, a.as<float2>(), w, h);
            hipLaunchKernelGGL(permute_rows_kernel, grid, dim3(256), 0, ctx->stream, a.as<float2>(), b.as<float2>(), w, h,
                               sp, 0);
        } else {
            hipLaunchKernelGGL(permute_rows_kernel, grid, dim3(256), 0, ctx->stream, a.as<float2>(), b.as<float2>(), w, h,
                               sp, 1);
            ldh::fft2d_inverse(ctx->stream, ctx->fft, b.as<float2>(), w, h);
            hipLaunchKernelGGL(scale_c64_kernel, grid, dim3(256), 0, ctx->stream, b.as<float2>(), n, 1.0f / float(n));
        }
        LDH_CHECK_LAUNCH("fft2d");
        d2h(ctx, data, b.get(), sizeof(float2) * n);
    });
}

ldh_status ldh_bank_acquire(ldh_context* ctx, const ldh_params* p, ldh_bank** out) {
    return guarded(ctx, [&] {
        require(out != nullptr, LDH_ERR_INVALID_ARGUMENT, "missing-argument");
        check_params(p, true);
        // the bins of LensDiffPsfBankCache: the global bank, also for a field-dependent phase
        // (buildPsfBankGlobalOnly, LensDiffCpuReference.cpp:1769)
        ldh_params q = *p;
        q.phase_field_strength = 0.0;
        auto* b = new ldh_bank();
        b->bank = ensure_bank(ctx, q, 3);
        b->key = ctx->cached_key;
        *out = b;
    });
}

ldh_status ldh_bank_build_batch(ldh_context* ctx, const ldh_params* params, int32_t count, int32_t mode,
                                ldh_bank** banks, ldh_batch_stats* stats) {
    return guarded(ctx, [&] {
        require(params && banks && count > 0 && mode >= LDH_BATCH_PER_REQUEST && mode <= LDH_BATCH_GRAPH,
                LDH_ERR_INVALID_ARGUMENT, "invalid-argument");
        for (int32_t i = 0; i < count; ++i) banks[i] = nullptr;
        std::vector<ldh_params> reqs(params, params + count);
        // a field-dependent phase keeps its global bank here, as the reference's cache.bins
        // (buildPsfBankGlobalOnly); its zone banks live with the renders
        for (ldh_params& q : reqs) q.phase_field_strength = 0.0;
        ldh::BatchTiming tm;
        // a capture must not record timing events: bind a disabled profiler
        ldh::Profiler none;
        ldh::ProfilerBinding unbind(none);
        auto built = ldh::build_banks_batched(ctx->stream, ctx->fft, ctx->bw, reqs, mode, &tm, &ctx->lanes);
        for (int32_t i = 0; i < count; ++i) {
            auto* b = new ldh_bank();
            b->bank = built[size_t(i)];
            banks[i] = b;
        }
        if (stats) {
            stats->exec_ms = tm.exec_ms;
            stats->capture_ms = tm.capture_ms;
            stats->groups = tm.groups;
            stats->psfs = tm.psfs;
            stats->graph_nodes = tm.graph_nodes;
            stats->mode = tm.mode;
            stats->branches = std::max(1, tm.branches);
        }
    });
}

ldh_status ldh_bank_release(ldh_context* ctx, ldh_bank* bank) {
    if (!bank) return LDH_ERR_INVALID_ARGUMENT;
    return guarded(ctx, [&] {
        if (--bank->refs <= 0) delete bank;
    });
}

ldh_status ldh_bank_info_get(ldh_context* ctx, const ldh_bank* bank, ldh_bank_info* out) {
    return guarded(ctx, [&] {
        require(bank && out, LDH_ERR_INVALID_ARGUMENT, "missing-argument");
        const ldh::DeviceBank& b = *bank->bank;
        std::memset(out, 0, sizeof(*out));
        out->bin_count = b.bins;
        out->support_radius_px = b.support;
        out->pupil_size = b.pupil;
        out->raw_psf_size = b.raw;
        out->raw_window_px = b.raw_window;
        for (int i = 0; i < b.bins && i < LDH_MAX_WAVELENGTHS; ++i) {
            out->wavelengths_nm[i] = b.wavelengths[size_t(i)];
            out->kernel_size[i] = 2 * std::min(std::max(b.radius[size_t(i)], 1), b.kernel / 2) + 1;
        }
    });
}

ldh_status ldh_bank_download(ldh_context* ctx, const ldh_bank* bank, int32_t bin, int32_t which, float* out,
                             int64_t capacity) {
    return guarded(ctx, [&] {
        require(bank && out, LDH_ERR_INVALID_ARGUMENT, "missing-argument");
        const ldh::DeviceBank& b = *bank->bank;
        require(bin >= 0 && bin < b.bins && which >= 0 && which <= 2, LDH_ERR_INVALID_ARGUMENT, "invalid-bin");
        const int K = b.kernel;
        const int rc = std::min(std::max(b.radius[size_t(bin)], 1), K / 2);
        const int n = 2 * rc + 1, c = K / 2;
        require(capacity >= int64_t(n) * n, LDH_ERR_INVALID_ARGUMENT, "capacity-too-small");
        std::vector<float> h(size_t(K) * K);
        d2h(ctx, h.data(), b.kernel_ptr(which, bin), sizeof(float) * size_t(K) * K);
        // the stored texels times the kernel's unit-sum scale: the normalised kernel
        const float sc = b.host_scale(which, bin);
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) out[size_t(y) * n + x] = h[size_t(c - rc + y) * K + size_t(c - rc + x)] * sc;
    });
}

ldh_status ldh_render_device(ldh_context* ctx, const ldh_params* p, const float* d_src, float* d_dst, int32_t w,
                             int32_t h) {
    return guarded(ctx, [&] {
        require(d_src && d_dst, LDH_ERR_MISSING_BUFFER, "missing-source-or-destination-buffer");
        require(w > 0 && h > 0, LDH_ERR_INVALID_BOUNDS, "invalid-source-bounds");
        check_params(p, true);
        render_with_banks(ctx, *p, reinterpret_cast<const float4*>(d_src), reinterpret_cast<float4*>(d_dst), w, h);
    });
}

// RunLensDiffCpuReference's I/O contract on host views (LensDiffCpuReference.cpp:2077-2085,
// :128-192, :2235-2258): decode from the strided source view, render, write the render window
// intersected with the destination bounds; pixels outside the source read as (0,0,0,1).
ldh_status ldh_render(ldh_context* ctx, const ldh_render_request* req, const ldh_params* p) {
    return guarded(ctx, [&] {
        require(req != nullptr, LDH_ERR_INVALID_ARGUMENT, "missing-request");
        require(req->src.data && req->dst.data, LDH_ERR_MISSING_BUFFER, "missing-source-or-destination-buffer");
        const ldh_rect sb = req->src.bounds;
        const int w = sb.x2 - sb.x1, h = sb.y2 - sb.y1;
        require(w > 0 && h > 0, LDH_ERR_INVALID_BOUNDS, "invalid-source-bounds");
        require(req->src.bytes_per_component == 4 && req->src.components >= 3 && req->dst.bytes_per_component == 4 &&
                    req->dst.components >= 3, LDH_ERR_UNSUPPORTED, "unsupported-pixel-format");
        check_params(p, true);
        const size_t bytes = sizeof(float) * 4 * size_t(w) * h;
        float* host = pinned(ctx, bytes);
        const auto* sbase = static_cast<const uint8_t*>(req->src.data);
        const int64_t spix = int64_t(req->src.components) * 4;
        for (int y = 0; y < h; ++y) {
            const uint8_t* row = sbase + int64_t(y) * req->src.row_bytes;
            float* o = host + size_t(y) * w * 4;
            if (req->src.components == 4) {
                std::memcpy(o, row, sizeof(float) * 4 * w);
            } else {
                for (int x = 0; x < w; ++x) {
                    const float* q = reinterpret_cast<const float*>(row + x * spix);
                    o[4 * x] = q[0]; o[4 * x + 1] = q[1]; o[4 * x + 2] = q[2]; o[4 * x + 3] = 1.0f;
                }
            }
        }
        ctx->dsrc.reset(bytes);
        ctx->ddst.reset(bytes);
        h2d(ctx, ctx->dsrc.get(), host, bytes);
        render_with_banks(ctx, *p, ctx->dsrc.as<float4>(), ctx->ddst.as<float4>(), w, h);
        d2h(ctx, host, ctx->ddst.get(), bytes);
        const ldh_rect db = req->dst.bounds, rw = req->render_window;
        const int x1 = std::max(rw.x1, db.x1), y1 = std::max(rw.y1, db.y1);
        const int x2 = std::max(x1, std::min(rw.x2, db.x2)), y2 = std::max(y1, std::min(rw.y2, db.y2));
        auto* dbase = static_cast<uint8_t*>(req->dst.data);
        const int64_t dpix = int64_t(req->dst.components) * 4;
        for (int y = y1; y < y2; ++y)
            for (int x = x1; x < x2; ++x) {
                const int sx = x - sb.x1, sy = y - sb.y1;
                float rgba[4] = {0.0f, 0.0f, 0.0f, 1.0f};
                if (sx >= 0 && sx < w && sy >= 0 && sy < h) std::memcpy(rgba, host + (size_t(sy) * w + sx) * 4, 16);
                float* d = reinterpret_cast<float*>(dbase + int64_t(y - db.y1) * req->dst.row_bytes + int64_t(x - db.x1) * dpix);
                d[0] = rgba[0]; d[1] = rgba[1]; d[2] = rgba[2];
                if (req->dst.components > 3) d[3] = rgba[3];
            }
    });
}

ldh_status ldh_server_start(ldh_context* ctx, int32_t slots, uint64_t budget) {
    return guarded(ctx, [&] {
        require(slots >= 1 && slots <= 16, LDH_ERR_INVALID_ARGUMENT, "invalid-slot-count");
        require(!ctx->server, LDH_ERR_INVALID_ARGUMENT, "server-already-running");
        ctx->server = std::make_unique<ldh::Server>(ctx->fft, ctx->device, slots, budget ? budget : ctx->budget);
    });
}

ldh_status ldh_server_stop(ldh_context* ctx) {
    return guarded(ctx, [&] { ctx->server.reset(); });
}

// The served calls take the server pointer under the context lock and then wait without it, so
// submissions and polls from other threads are not blocked by a waiting caller.
static ldh::Server* server_of(ldh_context* ctx) {
    std::lock_guard<std::mutex> lock(ctx->mu);
    return ctx->server.get();
}

ldh_status ldh_submit(ldh_context* ctx, const ldh_params* p, const float* d_src, float* d_dst, int32_t w, int32_t h,
                      int64_t* ticket) {
    if (!ctx || !p || !ticket) return LDH_ERR_INVALID_ARGUMENT;
    if (!d_src || !d_dst) return LDH_ERR_MISSING_BUFFER;
    if (w <= 0 || h <= 0) return LDH_ERR_INVALID_BOUNDS;
    ldh::Server* s = server_of(ctx);
    if (!s) return LDH_ERR_INVALID_ARGUMENT;
    *ticket = s->submit(*p, reinterpret_cast<const float4*>(d_src), reinterpret_cast<float4*>(d_dst), w, h);
    return LDH_OK;
}

static ldh_status poll_impl(ldh_context* ctx, int64_t ticket, ldh_request_stats* out, bool wait) {
    if (!ctx) return LDH_ERR_INVALID_ARGUMENT;
    ldh::Server* s = server_of(ctx);
    if (!s) return LDH_ERR_INVALID_ARGUMENT;
    ldh::ServedRequest r;
    try {
        if (!s->poll(ticket, &r, wait)) return LDH_ERR_NOT_READY;
    } catch (const ldh::Error& e) {
        return e.status;
    }
    if (out) {
        std::memset(out, 0, sizeof(*out));
        out->latency_ms = r.t_done - r.t_submit;
        out->queue_ms = r.t_start - r.t_submit;
        out->service_ms = r.t_done - r.t_start;
        out->bank_hit = r.bank_hit ? 1 : 0;
        out->spectra_hit = r.spectra_hit ? 1 : 0;
        out->status = r.status;
        std::strncpy(out->error, r.error.c_str(), sizeof(out->error) - 1);
    }
    return r.status;
}

ldh_status ldh_poll(ldh_context* ctx, int64_t ticket, ldh_request_stats* out) { return poll_impl(ctx, ticket, out, false); }
ldh_status ldh_wait(ldh_context* ctx, int64_t ticket, ldh_request_stats* out) { return poll_impl(ctx, ticket, out, true); }

ldh_status ldh_server_stats_get(ldh_context* ctx, ldh_server_stats* out) {
    if (!ctx || !out) return LDH_ERR_INVALID_ARGUMENT;
    ldh::Server* s = server_of(ctx);
    if (!s) return LDH_ERR_INVALID_ARGUMENT;
    *out = s->stats();
    return LDH_OK;
}

ldh_status ldh_device_alloc(ldh_context* ctx, uint64_t bytes, void** out) {
    return guarded(ctx, [&] {
        require(out != nullptr, LDH_ERR_INVALID_ARGUMENT, "missing-argument");
        if (hipMalloc(out, bytes) != hipSuccess) {
            (void)hipGetLastError();
            throw ldh::Error{LDH_ERR_OUT_OF_MEMORY, "device-allocation-failed"};
        }
    });
}
ldh_status ldh_device_free(ldh_context* ctx, void* ptr) {
    return guarded(ctx, [&] { LDH_HIP(hipFree(ptr), "device-free"); });
}
ldh_status ldh_memcpy_h2d(ldh_context* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded(ctx, [&] {
        h2d(ctx, dst, src, bytes);
        sync(ctx, "copy-to-device");
    });
}
ldh_status ldh_memcpy_d2h(ldh_context* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded(ctx, [&] { d2h(ctx, dst, src, bytes); });
}

ldh_status ldh_profiling_enable(ldh_context* ctx, int32_t enable) {
    return guarded(ctx, [&] {
        ctx->prof.enabled = enable != 0;
        if (ctx->server) ctx->server->set_profiling(enable != 0);
    });
}
int32_t ldh_profiling_read(ldh_context* ctx, ldh_kernel_stat* out, int32_t capacity) {
    int32_t n = 0;
    guarded(ctx, [&] {
        sync(ctx, "profiling-read");
        ctx->prof.resolve();
        if (ctx->server) ctx->server->merge_profiles(ctx->prof.stats);
        n = ctx->prof.read(out, capacity);
    });
    return n;
}
ldh_status ldh_profiling_reset(ldh_context* ctx) {
    return guarded(ctx, [&] {
        sync(ctx, "profiling-reset");
        ctx->prof.reset();
    });
}

int32_t ldh_tuning_report(char* buf, int32_t capacity) {
    const std::string r = ldh::Autotuner::instance().report() + ldh::col_plan_report();
    if (buf != nullptr && capacity > 0) {
        const size_t n = std::min(r.size(), size_t(capacity - 1));
        std::memcpy(buf, r.data(), n);
        buf[n] = '\0';
    }
    return int32_t(r.size());
}

void ldh_tuning_force(int32_t variant) { ldh::Autotuner::instance().force(int(variant)); }

void ldh_debug_inject_failure(const char* site, int32_t count) {
    if (site == nullptr) return;
    std::lock_guard<std::mutex> lk(ldh::g_debug_mu);
    if (count <= 0)
        ldh::g_debug_failures.erase(site);
    else
        ldh::g_debug_failures[site] = int(count);
}

}  // extern "C"

// Batched bank synthesis: many lenses with uneven pupil sizes and wavelength counts built
// together. Requests with equal (pupil P, raw size R, kernel radius r) form a group; the group's
// (request, wavelength) units go through every stage of the single-bank build
// (buildPsfBankGlobalOnly + populatePsfBankCacheFromBaseKernels, LensDiffCpuReference.cpp:
// 1616-1656, 1702-1734) in shared launches, each unit reading its own amplitude, phase
// coefficients, resample scale and anisotropy gain. Everything that synchronises with the host
// (clear-circle first minima, Bluestein filter tables, parameter uploads, buffer sizing) happens
// before the launches, so the whole schedule can be captured into one HIP graph and launched once.
// In the graph the groups form up to kGraphBranches parallel branches (balanced by their texel
// counts, each branch with its own workspace), so that one group's latency-bound steps (the
// sequential-sum walks, one wavefront per sum) overlap another's bandwidth-bound ones.
// The banks of a group share one kernel store, [which][unit][K*K].
#include <algorithm>
#include <chrono>
#include <map>
#include <tuple>

#include "bank.h"
#include "host_params.h"

namespace ldh {

namespace {

struct Group {
    int P = 0, R = 0, r = 0, K = 0;
    std::vector<int> requests;   // batch indices, in batch order
    std::vector<int> unit0;      // first unit of each request
    int units = 0;
    std::vector<PupilGeom> geoms;
    std::vector<int> amp_index;
    std::vector<PhaseCoef> coefs;
    std::vector<float> scales, gains;
    PsfPlan plan;
    DeviceBuffer d_geoms, d_amp_index, d_coefs, d_scales, d_gains, d_unit0, amp, phase, radius;
    bool defocus_only = true;  // every request's coefficient sets differ only in defocus
    int window = 0;            // raw PSFs synthesised on their centred (2 window + 1)^2 block
    int form = LDH_RAW_PSF_AUTO;
    std::shared_ptr<DeviceBuffer> store, kernel_scales;  // [which][unit] unit-sum multipliers
    int branch = 0;                                       // graph branch (its workspace and stream)
    double cost = 0.0;                                    // texels synthesised and shaped
};

template <class T>
void upload(DeviceBuffer& d, const std::vector<T>& h, hipStream_t s) {
    d.reset(sizeof(T) * h.size());
    LDH_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s), "batch-upload");
}

// The launches of one group (no host synchronisation, no allocation once the workspace holds
// the batch's largest group).
void enqueue_group(hipStream_t s, const FftTables& t, BankWorkspace& w, Group& g) {
    const int P = g.P, R = g.R, U = g.units, K = g.K;
    pupil_amplitudes(s, g.d_geoms.as<PupilGeom>(), P, int(g.requests.size()), g.amp.as<float>(), true);
    // raw PSFs (their centred windows, or whole with unit sums applied inside the resample), base
    // kernels, shaping
    const bool win = g.window > 0;
    const float* phase = nullptr;
    const int* amp_index = nullptr;
    const PhaseCoef* dcoefs = g.d_coefs.as<PhaseCoef>();
    const int* unit0 = g.d_unit0.as<int>();
    if (!g.defocus_only) {
        pupil_phase(s, g.d_coefs.as<PhaseCoef>(), P, U, g.phase.as<float>(), true);
        phase = g.phase.as<float>();
        amp_index = g.d_amp_index.as<int>();
        dcoefs = nullptr;
        unit0 = nullptr;
    }
    const int pupils = g.defocus_only ? int(g.requests.size()) : 1;
    if (win)
        g.plan.window_psfs(s, t, g.amp.as<float>(), phase, U, w.raw.as<float>(), w.psf, amp_index, dcoefs, unit0,
                           pupils);
    else
        g.plan.raw_psfs(s, t, g.amp.as<float>(), phase, U, w.raw.as<float>(), w.totals.as<float>(), w.seqsum, w.psf,
                        false, amp_index, dcoefs, unit0, pupils);
    base_kernels(s, w, w.raw.as<float>(), R, g.d_scales.as<float>(), win ? nullptr : w.totals.as<float>(), g.r, U,
                 w.base.as<float>(), w.base_totals.as<float>(), g.window);
    shape_kernels(s, w, w.base.as<float>(), w.base_totals.as<float>(), K, U, 0.0, g.r, g.store->as<float>(),
                  g.radius.as<int>(), g.kernel_scales->as<float>(), g.d_gains.as<float>());
}

}  // namespace

void BatchLanes::ensure(int extra) {
    while (int(streams.size()) < extra) {
        hipStream_t st = nullptr;
        LDH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "batch-stream");
        streams.push_back(st);
        work.push_back(std::make_unique<BankWorkspace>());
    }
}

BatchLanes::~BatchLanes() {
    work.clear();
    for (hipStream_t st : streams) (void)hipStreamDestroy(st);
}

std::vector<std::shared_ptr<DeviceBank>> build_banks_batched(hipStream_t s, const FftTables& t, BankWorkspace& w,
                                                             const std::vector<ldh_params>& requests, int mode,
                                                             BatchTiming* timing, BatchLanes* lanes) {
    const int n = int(requests.size());
    std::vector<std::shared_ptr<DeviceBank>> banks(static_cast<size_t>(n));
    BatchTiming tm;
    tm.mode = mode;
    hipEvent_t e0, e1;
    LDH_HIP(hipEventCreate(&e0), "batch-event");
    LDH_HIP(hipEventCreate(&e1), "batch-event");
    struct Events {
        hipEvent_t a, b;
        ~Events() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } ev{e0, e1};

    for (const ldh_params& p : requests) {
        const std::string why = unsupported_reason(p);
        if (!why.empty()) fail(LDH_ERR_UNSUPPORTED, why);
    }

    if (mode == LDH_BATCH_PER_REQUEST) {
        // one launch sequence per request: the single-bank build, in order
        for (int i = 0; i < n; ++i) (void)clear_circle_first_zero(s, t, w, resolve_sizes(requests[size_t(i)]).pupil,
                                                                     resolve_sizes(requests[size_t(i)]).raw);
        LDH_HIP(hipEventRecord(e0, s), "batch-event");
        for (int i = 0; i < n; ++i) {
            banks[size_t(i)] = std::shared_ptr<DeviceBank>(build_bank(s, t, w, requests[size_t(i)]).release());
            tm.psfs += banks[size_t(i)]->bins;
        }
        LDH_HIP(hipEventRecord(e1, s), "batch-event");
        LDH_HIP(hipEventSynchronize(e1), "batch-event");
        float ms = 0.0f;
        LDH_HIP(hipEventElapsedTime(&ms, e0, e1), "batch-event");
        tm.exec_ms = ms;
        tm.groups = n;
        if (timing) *timing = tm;
        return banks;
    }

    // ---- plan: groups by (P, R, r), units in request order ----
    std::map<std::tuple<int, int, int, int>, int> index;
    std::vector<std::unique_ptr<Group>> groups;
    std::vector<std::vector<float>> wls(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        const ldh_params& p = requests[size_t(i)];
        const Sizes sz = resolve_sizes(p);
        if (sz.max_radius > kMaxKernelRadius) fail(LDH_ERR_TOO_LARGE, "kernel-radius-unsupported");
        if (sz.raw > kMaxRawPsf) fail(LDH_ERR_TOO_LARGE, "raw-psf-size-unsupported");
        const auto key = std::make_tuple(sz.pupil, sz.raw, sz.max_radius, p.raw_psf_form);
        auto it = index.find(key);
        if (it == index.end()) {
            it = index.emplace(key, int(groups.size())).first;
            auto g = std::make_unique<Group>();
            g->P = sz.pupil;
            g->R = sz.raw;
            g->r = sz.max_radius;
            g->K = 2 * sz.max_radius + 1;
            g->form = p.raw_psf_form;
            groups.push_back(std::move(g));
        }
        Group& g = *groups[size_t(it->second)];
        g.requests.push_back(i);
        wls[size_t(i)] = wavelength_list(p);
    }
    // ---- graph branches: groups dealt to the least loaded branch, largest first ----
    int branches = 1;
    if (mode == LDH_BATCH_GRAPH && lanes != nullptr && groups.size() > 1) {
        branches = std::min(kGraphBranches, int(groups.size()));
        lanes->ensure(branches - 1);
    }
    auto work_of = [&](const Group& g) -> BankWorkspace& { return g.branch == 0 ? w : *lanes->work[size_t(g.branch - 1)]; };
    // ---- host-side set-up: first minima, plans, per-unit parameters, buffers ----
    for (auto& gp : groups) {
        Group& g = *gp;
        const float first_zero = clear_circle_first_zero(s, t, w, g.P, g.R);
        for (size_t k = 0; k < g.requests.size(); ++k) {
            const int i = g.requests[k];
            const ldh_params& p = requests[size_t(i)];
            std::vector<PhaseCoef> coefs;
            std::vector<float> scales;
            bank_wavelength_params(p, resolve_sizes(p), wls[size_t(i)], first_zero, coefs, scales);
            g.unit0.push_back(g.units);
            g.defocus_only = g.defocus_only && defocus_only_coefs(coefs);
            g.geoms.push_back(make_pupil_geom(p));
            const float gain = 1.0f + std::max(0.0f, float(p.anisotropy_emphasis)) * 4.0f;
            for (size_t b = 0; b < coefs.size(); ++b) {
                g.amp_index.push_back(int(k));
                g.coefs.push_back(coefs[b]);
                g.scales.push_back(scales[b]);
                g.gains.push_back(gain);
            }
            g.units += int(coefs.size());
        }
        g.plan.prepare(s, g.P, g.R);
        const int h = resample_window_half_width(g.scales, g.r);
        g.window = g.form == LDH_RAW_PSF_AUTO && g.plan.prepare_window(s, h) ? h : 0;
        const double side = g.window > 0 ? 2.0 * g.window + 1.0 : double(g.R);
        g.cost = double(g.units) * (side * side + double(g.K) * g.K);
    }
    if (branches > 1) {
        std::vector<Group*> order;
        for (auto& gp : groups) order.push_back(gp.get());
        std::stable_sort(order.begin(), order.end(), [](const Group* a, const Group* b) { return a->cost > b->cost; });
        std::vector<double> load(static_cast<size_t>(branches), 0.0);
        for (Group* g : order) {
            const size_t k = size_t(std::min_element(load.begin(), load.end()) - load.begin());
            g->branch = int(k);
            load[k] += g->cost;
        }
    }
    for (auto& gp : groups) {
        Group& g = *gp;
        BankWorkspace& gw = work_of(g);
        upload(g.d_geoms, g.geoms, s);
        upload(g.d_amp_index, g.amp_index, s);
        upload(g.d_coefs, g.coefs, s);
        std::vector<int> bounds = g.unit0;
        bounds.push_back(g.units);
        upload(g.d_unit0, bounds, s);
        upload(g.d_scales, g.scales, s);
        upload(g.d_gains, g.gains, s);
        const int64_t kk = int64_t(g.K) * g.K;
        g.amp.reset(sizeof(float) * size_t(g.P) * g.P * g.requests.size());
        g.phase.reset(sizeof(float) * size_t(g.P) * g.P * g.units);
        g.radius.reset(sizeof(int) * g.units);
        g.store = std::make_shared<DeviceBuffer>(sizeof(float) * kk * 3 * g.units);
        g.kernel_scales = std::make_shared<DeviceBuffer>(sizeof(float) * 3 * g.units);
        // each workspace grows to its branch's largest group before any launch
        const int side = g.window > 0 ? 2 * g.window + 1 : g.R;
        gw.raw.reset(sizeof(float) * size_t(side) * side * g.units);
        gw.totals.reset(sizeof(float) * g.units);
        gw.base.reset(sizeof(float) * kk * g.units);
        gw.base_totals.reset(sizeof(float) * g.units);
        gw.psf.reserve(g.P, side, g.units);
        if (g.window == 0) gw.seqsum.reserve(int64_t(g.R) * g.R, g.units);
        gw.reserve_shaping(s, g.K, g.units, g.r);
        tm.psfs += g.units;
    }
    tm.branches = branches;
    LDH_HIP(hipStreamSynchronize(s), "batch-upload");
    tm.groups = int(groups.size());

    // ---- launches: captured into one graph, or issued directly ----
    hipGraphExec_t exec = nullptr;
    if (mode == LDH_BATCH_GRAPH) {
        const auto c0 = std::chrono::steady_clock::now();
        hipGraph_t graph = nullptr;
        // branch b > 0 joins the capture on its own stream after a fork event on s and is joined
        // back before the capture ends
        std::vector<hipEvent_t> marks(static_cast<size_t>(branches), nullptr);
        struct Marks {
            std::vector<hipEvent_t>& v;
            ~Marks() {
                for (hipEvent_t e : v)
                    if (e) (void)hipEventDestroy(e);
            }
        } marks_guard{marks};
        for (auto& e : marks) LDH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), "batch-event");
        auto stream_of = [&](int b) { return b == 0 ? s : lanes->streams[size_t(b - 1)]; };
        LDH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "batch-capture");
        try {
            if (branches > 1) {
                LDH_HIP(hipEventRecord(marks[0], s), "batch-fork");
                for (int b = 1; b < branches; ++b) LDH_HIP(hipStreamWaitEvent(stream_of(b), marks[0], 0), "batch-fork");
            }
            for (auto& g : groups) enqueue_group(stream_of(g->branch), t, work_of(*g), *g);
            for (int b = 1; b < branches; ++b) {
                LDH_HIP(hipEventRecord(marks[size_t(b)], stream_of(b)), "batch-join");
                LDH_HIP(hipStreamWaitEvent(s, marks[size_t(b)], 0), "batch-join");
            }
        } catch (...) {
            hipGraph_t broken = nullptr;
            (void)hipStreamEndCapture(s, &broken);
            if (broken) (void)hipGraphDestroy(broken);
            throw;
        }
        LDH_HIP(hipStreamEndCapture(s, &graph), "batch-capture");
        size_t nodes = 0;
        (void)hipGraphGetNodes(graph, nullptr, &nodes);
        tm.graph_nodes = int(nodes);
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        LDH_HIP(ie, "batch-instantiate");
        tm.capture_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
        LDH_HIP(hipEventRecord(e0, s), "batch-event");
        const hipError_t le = hipGraphLaunch(exec, s);
        if (le != hipSuccess) {
            (void)hipGraphExecDestroy(exec);
            LDH_HIP(le, "batch-launch");
        }
    } else {
        LDH_HIP(hipEventRecord(e0, s), "batch-event");
        for (auto& g : groups) enqueue_group(s, t, w, *g);
    }
    LDH_HIP(hipEventRecord(e1, s), "batch-event");
    const hipError_t se = hipEventSynchronize(e1);
    if (exec) (void)hipGraphExecDestroy(exec);
    LDH_HIP(se, "batch-run");
    float ms = 0.0f;
    LDH_HIP(hipEventElapsedTime(&ms, e0, e1), "batch-event");
    tm.exec_ms = ms;

    // ---- banks: views into their group's store ----
    for (auto& gp : groups) {
        Group& g = *gp;
        std::vector<int> radius(static_cast<size_t>(g.units));
        std::vector<float> kscales(static_cast<size_t>(3 * g.units));
        LDH_HIP(hipMemcpy(radius.data(), g.radius.get(), sizeof(int) * g.units, hipMemcpyDeviceToHost), "batch-radius");
        LDH_HIP(hipMemcpy(kscales.data(), g.kernel_scales->get(), sizeof(float) * 3 * g.units, hipMemcpyDeviceToHost),
                "batch-scales");
        const int64_t kk = int64_t(g.K) * g.K;
        for (size_t k = 0; k < g.requests.size(); ++k) {
            const int i = g.requests[k];
            auto bank = std::make_shared<DeviceBank>();
            bank->bins = int(wls[size_t(i)].size());
            bank->pupil = g.P;
            bank->raw = g.R;
            bank->max_radius = g.r;
            bank->kernel = g.K;
            bank->raw_window = g.window;
            bank->wavelengths = wls[size_t(i)];
            bank->radius.assign(radius.begin() + g.unit0[k], radius.begin() + g.unit0[k] + bank->bins);
            bank->support = 4;
            for (int v : bank->radius) bank->support = std::max(bank->support, v);
            bank->shared = g.store;
            bank->store = g.store->as<float>() + int64_t(g.unit0[k]) * kk;
            bank->which_stride = kk * g.units;
            bank->shared_scales = g.kernel_scales;
            bank->scales = g.kernel_scales->as<float>() + g.unit0[k];
            bank->scales_which_stride = g.units;
            bank->host_scales.resize(size_t(3) * bank->bins);
            for (int which = 0; which < 3; ++which)
                for (int b = 0; b < bank->bins; ++b)
                    bank->host_scales[size_t(which) * bank->bins + b] = kscales[size_t(which) * g.units + g.unit0[k] + b];
            banks[size_t(i)] = bank;
        }
    }
    if (timing) *timing = tm;
    return banks;
}

}  // namespace ldh

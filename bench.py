"""Benchmark: UHD frames/s with 31-wavelength PSF banks on one MI355X, served vs plain.

Workload: 3840x2160 RGBA float32 frames (0.05% highlights, log-uniform [1, 5000]); lenses shaped
like the precision config (P=1024 7-blade pupil, Noll 2-21 N(0, 0.2) waves, 31 wavelengths
400-700 nm, R=4092 raw PSFs, K=2047 kernels, 6144x4608 FFT grid where the reference pads to
8192x8192). The request stream draws 70%
of its lenses from 8 resident lens sets and 30% fresh ones (a new bank each). Inputs are resident
in HBM before timing. A step is one frame (one request).

  plain   the per-frame path with the reference's semantics: single-entry bank cache, kernel
          spectra recomputed per frame, one convolution per wavelength bin (ldh_render_device)
  served  the request loop: resident banks and pre-combined kernel spectra keyed by lens, 4 slots
          on their own streams, admission under a 64 GB budget (ldh_submit / ldh_wait)

`value` is served frames/s with the whole stream submitted at once; p50/p99 latency of both paths
is measured under Poisson arrivals at 0.8x the plain path's sustainable rate. `roofline` is the
kernel family with the most device time in the plain timed region (HIP events on the context
stream, where kernels run one at a time); `roofline_served` is the same for the served region,
whose overlapping slot streams stretch each launch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "UHD frames/s @31λ bank + p99 render latency, served vs plain; rel L2 ≤1e-4"


def lens(seed):
    from lensdiff_amd import workloads
    return workloads.spectral_params(2160, 1024, 7, range(2, 22), 0.2, seed, 31, 47.36)


def stream(n, rng, fresh_base):
    """(lens seed, frame index) per request: 70% one of 8 resident lenses, 30% fresh."""
    out = []
    for i in range(n):
        seed = int(rng.integers(10, 18)) if rng.random() < 0.7 else fresh_base + i
        out.append((seed, i % 2))
    return out


def cpu_baseline(p, w, h, bins):
    """One bin's convolution and one raw PSF of the workload on one host core, scaled to a frame."""
    import ctypes as C
    from oracle.oracle import Reference, reference_available
    if reference_available():
        ref = Reference()
        t_conv = ref.lib.ref_time_convolve(w, h, 2047)
        t_psf = ref.lib.ref_time_raw_psf(C.byref(p))
        kind = "reference"
    else:
        from oracle.oracle import Oracle
        o = Oracle()
        plane = np.zeros((h, w), np.float32)
        plane.ravel()[::997] = 1.0
        ker = np.full((2047, 2047), 1.0 / 2047 ** 2, np.float32)
        t0 = time.perf_counter()
        o.convolve(plane, ker)
        t_conv = time.perf_counter() - t0
        amp = o.pupil_amplitude(p)
        t0 = time.perf_counter()
        o.raw_psf(amp, o.pupil_phase(p), o.sizes(p).raw_psf_size)
        t_psf = time.perf_counter() - t0
        kind = "port"
    frame_s = bins * (t_conv + t_psf)
    return {"value": 1.0 / frame_s, "unit": "frames/s", "cores": 1, "kind": kind,
            "sample": "1 of %d bin convolutions (8192x8192 grid, %.1f s) + 1 of %d raw PSFs (R=4092, %.1f s) "
                      "of one UHD frame with a fresh 31-wavelength bank, scaled x%d (%.0f s/frame); needed work "
                      "only" % (bins, t_conv, bins, t_psf, bins, frame_s)}


def fifo_latency(arrivals, service):
    done, out = 0.0, []
    for a, s in zip(arrivals, service):
        done = max(done, a) + s
        out.append(done - a)
    return np.array(out)


def traffic_of(name):
    """HBM bytes per launch of a kernel family measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), as recorded in profiles/traffic.json; None if absent."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get("families", {}).get(name)
    return None if rec is None else rec["hbm_bytes_per_launch"]


def dominant(stats):
    """Roofline of the kernel family with the most device time that has a byte model."""
    fams = {k: v for k, v in stats.items() if v[0] > 0 and v[2] > 0}
    name, (launches, total_ms, nbytes) = max(fams.items(), key=lambda kv: kv[1][1])
    avg_s = total_ms / 1e3 / launches
    achieved = nbytes / launches / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_of(name), "kernel": name,
            "launches": launches, "avg_launch_ms": round(avg_s * 1e3, 4), "bytes_per_launch": nbytes / launches}


def families(stats):
    return {k: {"launches": v[0], "ms": round(v[1], 3),
                "GB/s": round(v[2] / (v[1] / 1e3) / 1e9, 1) if v[1] > 0 and v[2] > 0 else None}
            for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])}


def batched_form(ctx, progress):
    """64 PSF-synthesis requests of uneven pupils (128-1024) and 8-31 wavelengths: one bank build after
    another, shared launches per size group, and the groups captured into one HIP graph. A first pass
    of each mode warms caches and the autotuner; the second is reported."""
    from lensdiff_amd import _abi, workloads
    reqs = workloads.batch_requests()
    res = {"requests": len(reqs)}
    for mode, name in ((_abi.LDH_BATCH_PER_REQUEST, "per_request"), (_abi.LDH_BATCH_GROUPED, "grouped"),
                       (_abi.LDH_BATCH_GRAPH, "graph")):
        for _ in range(2):
            handles, st = ctx.build_banks(reqs, mode)
            for h in handles:
                ctx.bank_release(h)
        res[name] = {"ms": round(st.exec_ms, 2), "psfs_per_s": round(st.psfs / (st.exec_ms / 1e3), 1),
                     "groups": st.groups}
        if mode == _abi.LDH_BATCH_GRAPH:
            res[name].update({"capture_ms": round(st.capture_ms, 2), "graph_nodes": st.graph_nodes})
        res["psfs"] = st.psfs
        progress("batched %s: %.1f ms, %.0f PSFs/s" % (name, st.exec_ms, res[name]["psfs_per_s"]))
    res["graph_vs_per_request"] = round(res["per_request"]["ms"] / res["graph"]["ms"], 2)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="served requests in the timed stream")
    ap.add_argument("--warmup", type=int, default=8, help="resident lens sets built before timing")
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-form measurement")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device-to-device 3 times (HBM counter calibration) and exit")
    ap.add_argument("--bank-probe", action="store_true",
                    help="only build two fresh 31-wavelength UHD banks (profiling the PSF path) and exit")
    ap.add_argument("--batch-probe", action="store_true", help="only run the batched form and exit")
    ap.add_argument("--field-probe", action="store_true",
                    help="only time a UHD frame with field-dependent phase (nine zone banks) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4, "copies": 3}))
        return

    from lensdiff_amd import workloads
    from lensdiff_amd.backend import Context
    if args.batch_probe:
        ctx = Context(0)
        print(json.dumps(batched_form(ctx, lambda m: print(m, file=sys.stderr, flush=True))))
        ctx.close()
        return
    if args.bank_probe:
        ctx = Context(0)
        ms = [round(ctx.bank_build_seconds(lens(5000 + i)) * 1e3, 2) for i in range(2)]
        ctx.profile_reset()
        ctx.profiling(True)
        ms.append(round(ctx.bank_build_seconds(lens(5002)) * 1e3, 2))
        ctx.profiling(False)
        fams = families(ctx.profile_read())
        print(json.dumps({"bank_build_ms": ms, "kernel_sides": ctx.bank_kernel_sizes(lens(5002)),
                          "autotune": ctx.tuning_report(), "kernel_families": fams}))
        ctx.close()
        return
    if args.field_probe:
        # one UHD frame with field-dependent phase: nine zone banks (31 wavelengths each) built in
        # shared launches, nine zone convolutions blended; then the same frame with the zone banks
        # cached, and the frame without field phase for comparison
        ctx = Context(0)
        f = workloads.c2_frame()
        h, w = f.shape[:2]
        d_src, d_dst = ctx.alloc(f.nbytes), ctx.alloc(f.nbytes)
        ctx.upload(d_src, f)
        p = lens(7000)
        p.phase_mode = 1
        p.phase_field_strength, p.phase_field_edge_bias = 1.0, 0.3
        p.phase_field_coma_radial, p.phase_field_astig_tangential, p.phase_field_defocus = 0.3, 0.2, 0.1

        def timed(q):
            t0 = time.perf_counter()
            ctx.render_device(q, d_src, d_dst, w, h)
            ctx.synchronize()
            return round((time.perf_counter() - t0) * 1e3, 2)
        out = {"field_zone_frame_first_ms": timed(p), "field_zone_frame_cached_ms": [timed(p) for _ in range(3)]}
        q = lens(7000)
        timed(q)
        out["global_frame_cached_ms"] = [timed(q) for _ in range(3)]
        print(json.dumps(out))
        ctx.close()
        return
    from oracle.oracle import rel_l2

    t_bench = time.perf_counter()
    rng = np.random.default_rng(4)
    frames = [workloads.c2_frame(), workloads.frame(3840, 2160, 0.0005, seed=1, log_uniform=(1.0, 5000.0))]
    h, w = frames[0].shape[:2]
    ctx = Context(0)
    d_src = []
    for f in frames:
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        d_src.append(d)
    nbytes = frames[0].nbytes
    lenses = {}

    def L(seed):
        if seed not in lenses:
            lenses[seed] = lens(seed)
        return lenses[seed]

    def progress(msg):
        print("[bench %.1fs] %s" % (time.perf_counter() - t_bench, msg), file=sys.stderr, flush=True)

    progress("inputs resident")
    # ---- plain: one request at a time, single-entry cache, per-bin convolution ----
    plain_reqs = stream(args.plain_steps, rng, 1000)
    d_plain = ctx.alloc(nbytes)
    ctx.render_device(L(10), d_src[0], d_plain, w, h)  # warm: kernels loaded, one bank built
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profiling(True)
    plain_service = []
    for seed, fi in plain_reqs:
        t0 = time.perf_counter()
        ctx.render_device(L(seed), d_src[fi], d_plain, w, h)
        ctx.synchronize()
        plain_service.append(time.perf_counter() - t0)
    ctx.profiling(False)
    plain_stats = ctx.profile_read()
    plain_fps = len(plain_service) / sum(plain_service)

    progress("plain: %d frames, %.2f frames/s" % (len(plain_service), plain_fps))
    # ---- breakdown: fresh 31-wavelength banks alone, and a plain frame whose bank is cached ----
    build_s = sorted(ctx.bank_build_seconds(L(5000 + i)) for i in range(3))
    ctx.render_device(L(10), d_src[0], d_plain, w, h)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.render_device(L(10), d_src[0], d_plain, w, h)
    ctx.synchronize()
    plain_cached_ms = (time.perf_counter() - t0) * 1e3
    breakdown = {"bank_build_ms": round(build_s[1] * 1e3, 2), "psfs_per_s": round(31 / build_s[1], 1),
                 "plain_frame_cached_bank_ms": round(plain_cached_ms, 2)}
    progress("breakdown: %s" % breakdown)
    # ---- served ----
    ctx.server_start(args.slots, 64 << 30)
    d_dst = [ctx.alloc(nbytes) for _ in range(args.steps)]
    for k, seed in enumerate(range(10, 10 + args.warmup)):  # resident lens sets
        ctx.wait(ctx.submit(L(seed), d_src[0], d_dst[k % len(d_dst)], w, h))
    served_reqs = stream(args.steps, rng, 2000)
    ctx.profile_reset()
    ctx.profiling(True)
    t_start = time.perf_counter()
    tickets = [ctx.submit(L(seed), d_src[fi], d_dst[i], w, h) for i, (seed, fi) in enumerate(served_reqs)]
    sat = [ctx.wait(t) for t in tickets]
    elapsed = time.perf_counter() - t_start
    ctx.profiling(False)
    served_stats = ctx.profile_read()
    served_fps = len(tickets) / elapsed

    progress("served: %d frames, %.2f frames/s" % (len(tickets), served_fps))
    # parity spot check: a served frame against the plain path for the same request
    seed, fi = served_reqs[0]
    ctx.render_device(L(seed), d_src[fi], d_plain, w, h)
    a = np.empty_like(frames[0])
    b = np.empty_like(frames[0])
    ctx.download(d_plain, a)
    ctx.download(d_dst[0], b)
    served_vs_plain = rel_l2(b, a)

    # ---- latency under Poisson arrivals at 0.8x the plain sustainable rate ----
    rate = 0.8 * plain_fps
    lat_reqs = stream(args.steps, rng, 3000)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(lat_reqs)))
    t0 = time.perf_counter()
    tickets = []
    for (seed, fi), at in zip(lat_reqs, arrivals):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tickets.append(ctx.submit(L(seed), d_src[fi], d_dst[len(tickets)], w, h))
    served_lat = np.array([ctx.wait(t).latency_ms for t in tickets])
    sstats = ctx.server_stats()
    ctx.server_stop()
    service = [plain_service[i % len(plain_service)] for i in range(len(arrivals))]
    plain_lat = fifo_latency(arrivals, service) * 1e3

    out = {
        "metric": METRIC,
        "value": round(served_fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "uhd-31wl served stream: 3840x2160 RGBA f32, 31 wavelengths 400-700 nm, P=1024, "
                               "R=4092, K=2047, 6144x4608 FFT grid; 70% of requests on 8 resident lens sets, 30% fresh",
                   "slots": args.slots, "budget_gb": 64},
        "plain": {"frames_per_s": round(plain_fps, 3), "requests": len(plain_service),
                  "service_ms_p50": round(float(np.median(plain_service)) * 1e3, 2)},
        "speedup_served_vs_plain": round(served_fps / plain_fps, 2),
        "latency_ms": {"arrival_rate_per_s": round(rate, 3),
                       "served_p50": round(float(np.percentile(served_lat, 50)), 2),
                       "served_p99": round(float(np.percentile(served_lat, 99)), 2),
                       "plain_p50": round(float(np.percentile(plain_lat, 50)), 2),
                       "plain_p99": round(float(np.percentile(plain_lat, 99)), 2)},
        "served_cache": {"bank_hits": sstats.bank_hits, "bank_misses": sstats.bank_misses,
                         "spectra_hits": sstats.spectra_hits, "resident_gb": round(sstats.resident_bytes / 2 ** 30, 2),
                         "evictions": sstats.evictions},
        "served_vs_plain_rel_l2": served_vs_plain,
        "breakdown": breakdown,
        "autotune": ctx.tuning_report(),
        # plain timed pass: one stream, every kernel runs alone
        "roofline": dominant(plain_stats),
        # served timed pass: slot streams overlap, so launch durations include sharing the device
        "roofline_served": dominant(served_stats),
        "kernel_families_plain": families(plain_stats),
        "kernel_families_served": families(served_stats),
    }
    progress("latency pass done")
    if not args.no_batched:
        out["batched"] = batched_form(ctx, progress)
    if not args.no_cpu_baseline:
        progress("cpu baseline (reference, one core)")
        out["cpu_baseline"] = cpu_baseline(L(10), w, h, 31)
        progress("cpu baseline done")
    for d in d_src + d_dst + [d_plain]:
        ctx.free(d)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Benchmark: UHD frames/s with 31-wavelength PSF banks on one MI355X, served vs plain.

Workload: 3840x2160 RGBA float32 frames (0.05% highlights, log-uniform [1, 5000]); lenses shaped
like the precision config (P=1024 7-blade pupil, Noll 2-21 N(0, 0.2) waves, 31 wavelengths
400-700 nm, R=4092 raw PSFs, K=2047 kernels, 6144x4608 FFT grid where the reference pads to
8192x8192). The request stream draws 70% of its lenses from 8 resident lens sets and 30% fresh
ones (a new bank each). Inputs are resident in HBM before timing; destinations are a ring of
buffers. A step is one frame (one request).

  plain   the per-frame path with the reference's semantics: single-entry bank cache, kernel
          spectra recomputed per frame, one convolution per wavelength bin (ldh_render_device)
  served  the request loop: resident banks and pre-combined kernel spectra keyed by lens, 4 slots
          on their own streams, admission under a 64 GB budget (ldh_submit / ldh_wait)

Before the served timed stream all 8 resident lens sets are built (always) and --warmup more
requests run untimed. `value` is served frames/s with the whole stream submitted at once. p50/p99
latency of both paths is measured under the same Poisson arrivals at 0.8x the plain path's
sustainable rate: the served requests are submitted at their arrival times, the plain ones run
through the plain path in arrival order on one worker (latency = completion - arrival). `c4_mixed`
repeats served and plain throughput and latency over the served config's frame-size mix. `roofline`
is the kernel family with the most device time in the plain timed region (HIP events on the
context stream, where kernels run one at a time); `roofline_served` the same for the served
region, whose overlapping slot streams stretch each launch; `roofline_served_alone` the same for
the served stream on one slot (--alone-steps requests after the resident lenses), where each
launch runs alone.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "UHD frames/s @31λ bank + p99 render latency, served vs plain; rel L2 ≤1e-4"
RESIDENT = list(range(10, 18))   # seeds of the 8 resident lens sets
C4_SIZES = [(1280, 720), (1920, 1080), (2560, 1440), (3840, 2160), (4096, 2160)]
N_DST = 16                       # destination ring


def lens(seed, n_wavelengths=31):
    from lensdiff_amd import workloads
    return workloads.spectral_params(2160, 1024, 7, range(2, 22), 0.2, seed, n_wavelengths, 47.36)


def stream(n, rng, fresh_base, sizes=1):
    """(lens seed, frame index) per request: 70% one of the 8 resident lenses, 30% fresh."""
    out = []
    for i in range(n):
        seed = int(rng.choice(RESIDENT)) if rng.random() < 0.7 else fresh_base + i
        out.append((seed, int(rng.integers(0, sizes))))
    return out


# The CPU baseline: the oracle's FP32 restatement of the reference CPU path (bit-identical to the
# reference on every fixture), one bin convolution on the reference's 8192x8192 grid and one
# R=4092 raw PSF, timed on host cores. Run in a child process so it can also run k at once.
_CPU_SAMPLE = r"""
import sys, time, json
import numpy as np
sys.path.insert(0, %(root)r)
from oracle.oracle import Oracle
from lensdiff_amd import workloads
o = Oracle()
p = workloads.spectral_params(2160, 1024, 7, range(2, 22), 0.2, 10, 31, 47.36)
plane = np.zeros((2160, 3840), np.float32)
plane.ravel()[::997] = 1.0
ker = np.full((2047, 2047), 1.0 / 2047 ** 2, np.float32)
t0 = time.perf_counter(); o.convolve(plane, ker); t_conv = time.perf_counter() - t0
amp = o.pupil_amplitude(p)
t0 = time.perf_counter(); o.raw_psf(amp, o.pupil_phase(p), o.sizes(p).raw_psf_size); t_psf = time.perf_counter() - t0
print(json.dumps({"t_conv": t_conv, "t_psf": t_psf}))
"""


def cpu_baseline(bins, progress):
    """UHD frame with a fresh 31-wavelength bank on host cores: one core, and k = host cores
    running independent frames at once (socket throughput). "needed work only" counts one
    convolution per bin; "as written" adds the core and structure convolutions the reference's
    Physical look also runs per bin (LensDiffCpuReference.cpp:2147-2151)."""
    env = dict(os.environ, OMP_NUM_THREADS="1")
    code = _CPU_SAMPLE % {"root": ROOT}

    def run(k):
        procs = [subprocess.Popen([sys.executable, "-c", code], stdout=subprocess.PIPE, env=env) for _ in range(k)]
        outs = [json.loads(p.communicate()[0].decode().strip().splitlines()[-1]) for p in procs]
        if any(p.returncode != 0 for p in procs):
            raise RuntimeError("cpu baseline sample failed")
        return outs
    one = run(1)[0]
    progress("cpu baseline one core: conv %.1f s, psf %.1f s" % (one["t_conv"], one["t_psf"]))
    k = min(os.cpu_count() or 1, 16)
    many = run(k)
    progress("cpu baseline %d cores" % k)
    needed = bins * (one["t_conv"] + one["t_psf"])
    written = bins * (3 * one["t_conv"] + one["t_psf"])
    sock = sum(1.0 / (bins * (m["t_conv"] + m["t_psf"])) for m in many)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(1.0 / needed, 8), "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": "oracle FP32 restatement of the reference CPU path: 1 of %d bin convolutions (8192x8192 grid, "
                      "%.1f s) + 1 of %d raw PSFs (R=4092, %.1f s) of one UHD frame with a fresh %d-wavelength bank, "
                      "scaled x%d (%.0f s/frame); needed work only" % (bins, one["t_conv"], bins, one["t_psf"], bins,
                                                                       bins, needed),
            "as_written_frames_per_s": round(1.0 / written, 8),
            "as_written_s_per_frame": round(written, 1),
            "socket": {"cores": k, "frames_per_s": round(sock, 8),
                       "note": "k independent frames at once, each on one core"},
            "cpu_model": model}


def poisson(n, rate, rng):
    return np.cumsum(rng.exponential(1.0 / rate, n))


def traffic_of(name):
    """HBM bytes per launch of a kernel family measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), as recorded in profiles/traffic.json; None if absent."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get("families", {}).get(name)
    return None if rec is None else rec["hbm_bytes_per_launch"]


def dominant(stats, with_traffic=True):
    """Roofline of the kernel family with the most device time that has a byte model. `traffic`
    (profiles/traffic.json) is recorded for the plain pass's launch shapes; other passes launch the
    same family at other shapes and report it as null (their measured bytes are in
    profiles/uhd31_hbm_traffic_pmc.md)."""
    fams = {k: v for k, v in stats.items() if v[0] > 0 and v[2] > 0}
    name, (launches, total_ms, nbytes) = max(fams.items(), key=lambda kv: kv[1][1])
    avg_s = total_ms / 1e3 / launches
    achieved = nbytes / launches / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_of(name) if with_traffic else None,
            "kernel": name, "launches": launches, "avg_launch_ms": round(avg_s * 1e3, 4),
            "bytes_per_launch": nbytes / launches}


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
            res[name].update({"capture_ms": round(st.capture_ms, 2), "graph_nodes": st.graph_nodes,
                              "branches": st.branches})
        res["psfs"] = st.psfs
        progress("batched %s: %.1f ms, %.0f PSFs/s" % (name, st.exec_ms, res[name]["psfs_per_s"]))
    res["graph_vs_per_request"] = round(res["per_request"]["ms"] / res["graph"]["ms"], 2)
    res["graph_vs_grouped"] = round(res["grouped"]["ms"] / res["graph"]["ms"], 2)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240, help="served requests in the timed stream")
    ap.add_argument("--warmup", type=int, default=8,
                    help="untimed served requests after the 8 resident lens sets are built")
    ap.add_argument("--plain-steps", type=int, default=60)
    ap.add_argument("--latency-steps", type=int, default=240, help="requests of each latency stream")
    ap.add_argument("--mixed-steps", type=int, default=200, help="requests of each c4 frame-size-mix stream")
    ap.add_argument("--alone-steps", type=int, default=48,
                    help="requests of the one-slot served stream timed per kernel family")
    ap.add_argument("--no-mixed", action="store_true", help="skip the c4 frame-size mix")
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
    ap.add_argument("--plain-probe", action="store_true",
                    help="only render plain UHD frames with the bank cached (the launch shapes of the plain "
                         "pass, for counter runs) and exit")
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
    if args.plain_probe:
        ctx = Context(0)
        f = workloads.c2_frame()
        h, w = f.shape[:2]
        d_src, d_dst = ctx.alloc(f.nbytes), ctx.alloc(f.nbytes)
        ctx.upload(d_src, f)
        p = lens(RESIDENT[0])
        ctx.render_device(p, d_src, d_dst, w, h)  # the bank built and cached
        ctx.synchronize()
        ctx.profile_reset()
        ctx.profiling(True)
        ms = []
        for _ in range(4):
            t0 = time.perf_counter()
            ctx.render_device(p, d_src, d_dst, w, h)
            ctx.synchronize()
            ms.append(round((time.perf_counter() - t0) * 1e3, 2))
        ctx.profiling(False)
        stats = ctx.profile_read()
        print(json.dumps({"plain_frame_cached_bank_ms": ms, "roofline": dominant(stats),
                          "kernel_families": families(stats)}))
        ctx.close()
        return
    if args.field_probe:
        # one UHD frame with field-dependent phase: nine zone banks (31 wavelengths each) built in
        # shared launches, nine zone convolutions blended; then the same frame with the zone banks
        # cached, a second lens with field phase (its zone banks built into memory the first one
        # already drew from the pool), and the frame without field phase for comparison
        ctx = Context(0)
        f = workloads.c2_frame()
        h, w = f.shape[:2]
        d_src, d_dst = ctx.alloc(f.nbytes), ctx.alloc(f.nbytes)
        ctx.upload(d_src, f)

        def field_lens(seed):
            q = lens(seed)
            q.phase_mode = 1
            q.phase_field_strength, q.phase_field_edge_bias = 1.0, 0.3
            q.phase_field_coma_radial, q.phase_field_astig_tangential, q.phase_field_defocus = 0.3, 0.2, 0.1
            return q
        p = field_lens(7000)

        def timed(q):
            t0 = time.perf_counter()
            ctx.render_device(q, d_src, d_dst, w, h)
            ctx.synchronize()
            return round((time.perf_counter() - t0) * 1e3, 2)
        out = {"field_zone_frame_first_ms": timed(p), "field_zone_frame_cached_ms": [timed(p) for _ in range(3)],
               "field_zone_frame_next_lens_ms": timed(field_lens(7001))}
        q = lens(7000)
        timed(q)
        out["global_frame_cached_ms"] = [timed(q) for _ in range(3)]
        # served: the zone banks, their pre-combined kernels and their spectra stay resident
        ctx.server_start(1, 64 << 30)
        lat = [round(ctx.wait(ctx.submit(r, d_src, d_dst, w, h)).latency_ms, 2)
               for r in (p, p, p, p, field_lens(7002), field_lens(7003))]
        ctx.server_stop()
        out["served_field_zone_frame_ms"] = {"first": lat[0], "cached": lat[1:4], "next_lenses": lat[4:]}
        print(json.dumps(out))
        ctx.close()
        return
    from oracle.oracle import rel_l2

    t_bench = time.perf_counter()

    def progress(msg):
        print("[bench %.1fs] %s" % (time.perf_counter() - t_bench, msg), file=sys.stderr, flush=True)

    rng = np.random.default_rng(4)
    ctx = Context(0)
    uhd = [workloads.c2_frame(), workloads.frame(3840, 2160, 0.0005, seed=1, log_uniform=(1.0, 5000.0))]
    h, w = uhd[0].shape[:2]
    nbytes = uhd[0].nbytes
    d_src = []
    for f in uhd:
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        d_src.append(d)
    # destinations hold the largest frame of either stream (4096x2160 RGBA f32 of the c4 mix)
    dst_bytes = max(nbytes, max(w_ * h_ for w_, h_ in C4_SIZES) * 16)
    d_dst = [ctx.alloc(dst_bytes) for _ in range(N_DST)]
    lenses = {}

    def L(seed):
        if seed not in lenses:
            lenses[seed] = lens(seed)
        return lenses[seed]

    progress("inputs resident")
    # ---- plain: one request at a time, single-entry cache, per-bin convolution ----
    plain_reqs = stream(args.plain_steps, rng, 1000, len(uhd))
    ctx.render_device(L(RESIDENT[0]), d_src[0], d_dst[0], w, h)  # warm: kernels loaded, one bank built
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profiling(True)
    plain_service = []
    for i, (seed, fi) in enumerate(plain_reqs):
        t0 = time.perf_counter()
        ctx.render_device(L(seed), d_src[fi], d_dst[i % N_DST], w, h)
        ctx.synchronize()
        plain_service.append(time.perf_counter() - t0)
    ctx.profiling(False)
    plain_stats = ctx.profile_read()
    plain_fps = len(plain_service) / sum(plain_service)
    progress("plain: %d frames, %.2f frames/s" % (len(plain_service), plain_fps))

    # ---- breakdown: fresh 31-wavelength banks alone, and a plain frame whose bank is cached ----
    build_s = sorted(ctx.bank_build_seconds(L(5000 + i)) for i in range(3))
    ctx.render_device(L(RESIDENT[0]), d_src[0], d_dst[0], w, h)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.render_device(L(RESIDENT[0]), d_src[0], d_dst[0], w, h)
    ctx.synchronize()
    plain_cached_ms = (time.perf_counter() - t0) * 1e3
    breakdown = {"bank_build_ms": round(build_s[1] * 1e3, 2), "psfs_per_s": round(31 / build_s[1], 1),
                 "plain_frame_cached_bank_ms": round(plain_cached_ms, 2)}
    progress("breakdown: %s" % breakdown)

    # ---- served: all resident lens sets built, warm-up requests, then the timed stream ----
    ctx.server_start(args.slots, 64 << 30)
    for k, seed in enumerate(RESIDENT):
        ctx.wait(ctx.submit(L(seed), d_src[0], d_dst[k % N_DST], w, h))
    for i, (seed, fi) in enumerate(stream(args.warmup, rng, 1500, len(uhd))):
        ctx.wait(ctx.submit(L(seed), d_src[fi], d_dst[i % N_DST], w, h))
    served_reqs = stream(args.steps, rng, 2000, len(uhd))
    st0 = ctx.server_stats()
    ctx.profile_reset()
    ctx.profiling(True)
    t_start = time.perf_counter()
    tickets = [ctx.submit(L(seed), d_src[fi], d_dst[i % N_DST], w, h) for i, (seed, fi) in enumerate(served_reqs)]
    for t in tickets:
        ctx.wait(t)
    elapsed = time.perf_counter() - t_start
    ctx.profiling(False)
    served_stats = ctx.profile_read()
    st1 = ctx.server_stats()
    served_fps = len(tickets) / elapsed
    progress("served: %d frames, %.2f frames/s" % (len(tickets), served_fps))

    # parity spot check: a served frame against the plain path for the same request (own buffers)
    seed, fi = served_reqs[0]
    d_a, d_b = ctx.alloc(nbytes), ctx.alloc(nbytes)
    ctx.wait(ctx.submit(L(seed), d_src[fi], d_b, w, h))
    ctx.render_device(L(seed), d_src[fi], d_a, w, h)
    a = np.empty_like(uhd[0])
    b = np.empty_like(uhd[0])
    ctx.download(d_a, a)
    ctx.download(d_b, b)
    served_vs_plain = rel_l2(b, a)

    # ---- latency: the same Poisson arrivals through both paths ----
    rate = 0.8 * plain_fps
    lat_reqs = stream(args.latency_steps, rng, 3000, len(uhd))
    arrivals = poisson(len(lat_reqs), rate, rng)
    served_lat, served_q, served_svc = latency_served(ctx, [(L(sd), d_src[fi], w, h) for sd, fi in lat_reqs],
                                                     arrivals, d_dst)
    st2 = ctx.server_stats()
    ctx.server_stop()
    plain_lat, plain_svc = latency_plain(ctx, [(L(sd), d_src[fi], w, h) for sd, fi in lat_reqs], arrivals, d_dst)
    progress("latency pass done")

    # ---- the served stream on one slot: each launch runs alone (kernel families vs roofline) ----
    ctx.server_start(1, 64 << 30)
    for k, seed in enumerate(RESIDENT):
        ctx.wait(ctx.submit(L(seed), d_src[0], d_dst[k % N_DST], w, h))
    alone_reqs = stream(args.alone_steps, np.random.default_rng(5), 4000, len(uhd))
    ctx.profile_reset()
    ctx.profiling(True)
    tickets = [ctx.submit(L(seed), d_src[fi], d_dst[i % N_DST], w, h) for i, (seed, fi) in enumerate(alone_reqs)]
    for t in tickets:
        ctx.wait(t)
    ctx.profiling(False)
    alone_stats = ctx.profile_read()
    ctx.server_stop()
    progress("one-slot served pass done")

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
                   "slots": args.slots, "budget_gb": 64, "resident_lens_sets_prebuilt": len(RESIDENT),
                   "destination_ring": N_DST},
        "plain": {"frames_per_s": round(plain_fps, 3), "requests": len(plain_service),
                  "service_ms_p50": round(float(np.median(plain_service)) * 1e3, 2)},
        "speedup_served_vs_plain": round(served_fps / plain_fps, 2),
        "latency_ms": {"requests": len(lat_reqs), "arrival_rate_per_s": round(rate, 3),
                       "served_p50": pct(served_lat, 50), "served_p99": pct(served_lat, 99),
                       "plain_p50": pct(plain_lat, 50), "plain_p99": pct(plain_lat, 99),
                       "served_queue_p99": pct(served_q, 99), "served_service_p50": pct(served_svc, 50),
                       "served_service_p99": pct(served_svc, 99), "plain_service_p50": pct(plain_svc, 50),
                       "plain_service_p99": pct(plain_svc, 99),
                       "plain": "measured: the stream run through the plain path in arrival order"},
        "served_cache": {"bank_hits": st1.bank_hits - st0.bank_hits, "bank_misses": st1.bank_misses - st0.bank_misses,
                         "spectra_hits": st1.spectra_hits - st0.spectra_hits,
                         "resident_gb": round(st2.resident_bytes / 2 ** 30, 2), "evictions": st2.evictions},
        "served_vs_plain_rel_l2": served_vs_plain,
        "breakdown": breakdown,
        "autotune": ctx.tuning_report(),
        # plain timed pass: one stream, every kernel runs alone
        "roofline": dominant(plain_stats),
        # served timed pass: slot streams overlap, so launch durations include sharing the device
        "roofline_served": dominant(served_stats, False),
        # the served stream on one slot: every launch runs alone
        "roofline_served_alone": dominant(alone_stats, False),
        "kernel_families_plain": families(plain_stats),
        "kernel_families_served": families(served_stats),
        "kernel_families_served_alone": families(alone_stats),
    }
    for d in d_src + [d_a, d_b]:
        ctx.free(d)
    if not args.no_mixed:
        out["c4_mixed"] = c4_mixed(ctx, args, d_dst, dst_bytes, rng, progress)
    if not args.no_batched:
        out["batched"] = batched_form(ctx, progress)
    for d in d_dst:
        ctx.free(d)
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(31, progress)
    print(json.dumps(out))


def latency_served(ctx, reqs, arrivals, d_dst):
    """Requests submitted at their arrival times: latency = completion - submission, and its
    queue (admission wait) and service parts, per request (ms)."""
    t0 = time.perf_counter()
    tickets = []
    for i, ((p, src, w, h), at) in enumerate(zip(reqs, arrivals)):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tickets.append(ctx.submit(p, src, d_dst[i % len(d_dst)], w, h))
    st = [ctx.wait(t) for t in tickets]
    return (np.array([s.latency_ms for s in st]), np.array([s.queue_ms for s in st]),
            np.array([s.service_ms for s in st]))


def latency_plain(ctx, reqs, arrivals, d_dst):
    """The same arrivals through the plain path on one worker, in arrival order: a request starts
    when it has arrived and the previous one is done. Latency and service per request (ms)."""
    t0 = time.perf_counter()
    lat, svc = [], []
    for i, ((p, src, w, h), at) in enumerate(zip(reqs, arrivals)):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        t1 = time.perf_counter()
        ctx.render_device(p, src, d_dst[i % len(d_dst)], w, h)
        ctx.synchronize()
        t2 = time.perf_counter()
        svc.append((t2 - t1) * 1e3)
        lat.append((t2 - t0 - at) * 1e3)
    return np.array(lat), np.array(svc)


def pct(a, q):
    return round(float(np.percentile(a, q)), 2)


def c4_mixed(ctx, args, d_dst, dst_bytes, rng, progress):
    """The served config's frame-size mix (1280x720 .. 4096x2160 RGBA f32, uniform), lenses as in
    the UHD stream (70% on the 8 resident sets, each kernel sized for a 2160-line optics short side,
    so one bank serves every frame size; 30% fresh with 16-31 wavelengths): served and plain
    throughput, then both under the same Poisson arrivals at 0.8x the plain rate."""
    from lensdiff_amd import workloads
    srcs = []
    for k, (w, h) in enumerate(C4_SIZES):
        if w * h * 16 > dst_bytes:
            raise ValueError("destination buffers hold %d bytes, a %dx%d frame needs %d" % (dst_bytes, w, h, w * h * 16))
        f = workloads.frame(w, h, 0.0005, seed=40 + k, log_uniform=(1.0, 5000.0))
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        srcs.append((d, w, h))
    fresh_wl = {}

    def lens_of(seed):
        if seed in RESIDENT:
            return lens(seed)
        if seed not in fresh_wl:
            fresh_wl[seed] = int(rng.integers(16, 32))
        return lens(seed, fresh_wl[seed])

    def reqs(n, base):
        return [(lens_of(sd), srcs[fi][0], srcs[fi][1], srcs[fi][2]) for sd, fi in stream(n, rng, base, len(C4_SIZES))]
    n = args.mixed_steps
    plain_reqs = reqs(n, 6000)
    t0 = time.perf_counter()
    for i, (p, src, w, h) in enumerate(plain_reqs):
        ctx.render_device(p, src, d_dst[i % len(d_dst)], w, h)
    ctx.synchronize()
    plain_fps = n / (time.perf_counter() - t0)
    progress("c4 mixed plain: %.2f frames/s" % plain_fps)
    ctx.server_start(args.slots, 64 << 30)
    for k, seed in enumerate(RESIDENT):
        ctx.wait(ctx.submit(lens(seed), srcs[k % len(srcs)][0], d_dst[k % len(d_dst)], *srcs[k % len(srcs)][1:]))
    served_reqs = reqs(n, 7000)
    t0 = time.perf_counter()
    tickets = [ctx.submit(p, src, d_dst[i % len(d_dst)], w, h) for i, (p, src, w, h) in enumerate(served_reqs)]
    for t in tickets:
        ctx.wait(t)
    served_fps = n / (time.perf_counter() - t0)
    progress("c4 mixed served: %.2f frames/s" % served_fps)
    rate = 0.8 * plain_fps
    lat_reqs = reqs(n, 8000)
    arrivals = poisson(n, rate, rng)
    s_lat, _, _ = latency_served(ctx, lat_reqs, arrivals, d_dst)
    st = ctx.server_stats()
    ctx.server_stop()
    p_lat, p_svc = latency_plain(ctx, lat_reqs, arrivals, d_dst)
    for d, _, _ in srcs:
        ctx.free(d)
    return {"requests": n, "frame_sizes": ["%dx%d" % wh for wh in C4_SIZES],
            "served_frames_per_s": round(served_fps, 3), "plain_frames_per_s": round(plain_fps, 3),
            "speedup": round(served_fps / plain_fps, 2), "arrival_rate_per_s": round(rate, 3),
            "served_p50_ms": round(float(np.percentile(s_lat, 50)), 2),
            "served_p99_ms": round(float(np.percentile(s_lat, 99)), 2),
            "plain_p50_ms": round(float(np.percentile(p_lat, 50)), 2),
            "plain_p99_ms": round(float(np.percentile(p_lat, 99)), 2),
            "plain_service_p50_ms": pct(p_svc, 50),
            "evictions": st.evictions, "resident_gb": round(st.resident_bytes / 2 ** 30, 2)}


if __name__ == "__main__":
    main()

"""Parity at the benchmark configurations' full sizes: the HIP path against the oracle (the FP32
restatement of the reference CPU path, bit-identical to it on every golden fixture), each result
also scored against the FP64 oracle and that error printed.

  c1  1920x1080 RGBA, P=512 7 blades, Noll 2-15 N(0, 0.15), 16 wavelengths, R=2044, K=1023
  c2  its 31-wavelength bank (P=1024, R=4092, K=2047) every bin; a 3840x2160 frame with the c2
      lens at the tristimulus wavelengths (plain FP32, served FP32 and FP16 spectra)
  c3  one batched request per pupil size 128 .. 1024 (r = P/2, R = 2P)
  c4  one request at each served frame size, alone = in the mix, and against the oracle

The oracle spreads its FFT lines over OpenMP threads; these tests take a few seconds to a minute
each on the host cores."""
import numpy as np
import pytest

from lensdiff_amd import _abi, workloads
from oracle.oracle import rel_l2
from tests.kernel_compare import kernel_rel_l2, kernel_rel_l2_ring

pytestmark = pytest.mark.gpu

TOL = 1e-4


def max_abs(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).max())


def report(name, got, want32, want64):
    print("%-44s rel L2 vs FP32 oracle %.3g, vs FP64 %.3g (max abs %.3g)"
          % (name, rel_l2(got, want32), rel_l2(got, want64), max_abs(got, want64)))


def served_outputs(ctx, reqs, slots=1, budget=16 << 30, mixed=False):
    """Frames of (params, image) requests through the served loop: one at a time, or (mixed) all
    submitted at once."""
    bufs = []
    for p, img in reqs:
        s, d = ctx.alloc(img.nbytes), ctx.alloc(img.nbytes)
        ctx.upload(s, img)
        bufs.append((s, d))
    ctx.server_start(slots, budget)
    try:
        if mixed:
            tickets = [ctx.submit(p, s, d, img.shape[1], img.shape[0]) for (p, img), (s, d) in zip(reqs, bufs)]
            for t in tickets:
                ctx.wait(t)
        else:
            for (p, img), (s, d) in zip(reqs, bufs):
                ctx.wait(ctx.submit(p, s, d, img.shape[1], img.shape[0]))
    finally:
        ctx.server_stop()
    outs = []
    for (p, img), (s, d) in zip(reqs, bufs):
        o = np.empty_like(img)
        ctx.download(d, o)
        outs.append(o)
        ctx.free(s)
        ctx.free(d)
    return outs


# Kernels of R = 4092 raw PSFs. The reference normalises each raw PSF and then each kernel by
# sequential FP32 sums (LensDiffCpuReference.cpp:436-449) of 16.7M and 4.2M values that drift by
# 0.25% and 0.6% from their exact sums; where those sums run through their binades depends on the
# values' overall scale, and for some bins a 0.3% change of the raw PSF's scale moves the
# reference's own kernels by ~4e-4. The whole form (LDH_RAW_PSF_WHOLE) normalises as the reference
# does and meets 1e-4; the window form (LDH_RAW_PSF_AUTO) skips the raw PSF's unit-sum scale, which
# cancels in exact arithmetic, and its kernels follow the reference's within that sensitivity.
# The structure kernel, max(0, full - mean) renormalised, carries the errors of both and is
# smaller than either: its bound is twice theirs.
BANK_TOL = {_abi.RAW_PSF_WHOLE: {"full": 1e-4, "core": 1e-4, "structure": 2e-4},
            _abi.RAW_PSF_AUTO: {"full": 5e-4, "core": 5e-4, "structure": 5e-4}}


@pytest.fixture(scope="module")
def c2_oracle_bank(oracle):
    p = workloads.c2_params()
    want = oracle.bank(p)
    # FP64 oracle for the first, middle and last bins (each bin depends on its wavelength only)
    q = p.copy()
    q.set_wavelengths([want["wavelengths"][i] for i in (0, 15, 30)])
    return want, oracle.bank(q, fp64=True)


@pytest.mark.parametrize("form", [_abi.RAW_PSF_AUTO, _abi.RAW_PSF_WHOLE], ids=["window", "whole"])
def test_c2_bank_every_bin(ctx, oracle, c2_oracle_bank, form):
    """31 wavelengths, P=1024, R=4092, K=2047: every bin's full, core and structure kernels against
    the oracle, every crop radius equal; FP64 errors of both for the first, middle and last bins."""
    p = workloads.c2_params()
    p.raw_psf_form = form
    got = ctx.bank(p)
    want, want64 = c2_oracle_bank
    assert (got["raw_window"] > 0) == (form == _abi.RAW_PSF_AUTO)
    assert len(got["full"]) == len(want["full"]) == 31
    flips = [b for b in range(31) if got["full"][b].shape != want["full"][b].shape]
    err = {name: [kernel_rel_l2(got[name][b], want[name][b]) if b not in flips else 0.0 for b in range(31)]
           for name in ("full", "core", "structure")}
    for j, b in enumerate((0, 15, 30)):
        for name in ("full", "core", "structure"):
            g, w32, w64 = got[name][b], want[name][b], want64[name][j]
            if g.shape == w64.shape == w32.shape:
                print("c2 bin %2d %-9s rel L2 vs FP32 oracle %.3g; vs FP64: HIP %.3g, reference FP32 %.3g"
                      % (b, name, err[name][b], rel_l2(g, w64), rel_l2(w32, w64)))
    print("c2 bank (%s): raw window %d, %d of 31 crop radii differ; worst rel L2 vs FP32 oracle: full %.3g, "
          "core %.3g, structure %.3g" % ("window" if form == _abi.RAW_PSF_AUTO else "whole", got["raw_window"],
                                         len(flips), max(err["full"]), max(err["core"]), max(err["structure"])))
    assert not flips, flips
    for name in ("full", "core", "structure"):
        assert max(err[name]) < BANK_TOL[form][name], (name, err[name])


def tristimulus_c2():
    p = workloads.c2_params()
    p.set_wavelengths([])
    p.spectral_mode = _abi.SPECTRAL_TRISTIMULUS
    return p


def test_c2_uhd_frame_fp32_and_fp16_spectra(ctx, oracle):
    """3840x2160 frame, c2 lens at 460/550/610 nm (K=2047; the reference pads to 8192x8192, the
    device grid is 6144x4608): plain FP32, served FP32 and FP16 spectra."""
    p = tristimulus_c2()
    img = workloads.c2_frame()
    want32 = oracle.render_image(p, img)
    want64 = oracle.render_image(p, img, fp64=True)
    plain = ctx.render_image(p, img)
    p16 = p.copy()
    p16.spectra_fp16 = 1
    fp32, fp16 = served_outputs(ctx, [(p, img), (p16, img)])
    report("reference FP32 (oracle restatement)", want32, want32, want64)
    report("HIP plain per-bin FP32", plain, want32, want64)
    report("HIP served pre-combined FP32 spectra", fp32, want32, want64)
    report("HIP served pre-combined FP16 spectra", fp16, want32, want64)
    assert rel_l2(plain, want32) <= TOL
    assert rel_l2(fp32, want32) <= TOL
    assert rel_l2(fp16, want32) <= 5e-4
    assert rel_l2(fp16, want64) <= rel_l2(fp32, want64) + 5e-4


def test_c2_uhd_31_wavelengths_fp16_against_fp32(ctx):
    """The full c2 request (31 wavelengths, UHD): served FP16 spectra against served FP32 spectra
    and the plain per-bin path, all on the device (its oracle comparison is the bank test above
    and the tristimulus frame)."""
    p = workloads.c2_params()
    img = workloads.c2_frame()
    plain = ctx.render_image(p, img)
    p16 = p.copy()
    p16.spectra_fp16 = 1
    fp32, fp16 = served_outputs(ctx, [(p, img), (p16, img)])
    print("c2 31 wavelengths UHD: served FP32 vs plain %.3g, FP16 vs FP32 spectra %.3g"
          % (rel_l2(fp32, plain), rel_l2(fp16, fp32)))
    assert rel_l2(fp32, plain) <= 1e-5
    assert 0.0 < rel_l2(fp16, fp32) <= 5e-4


def test_c1_frame_16_wavelengths(ctx, oracle):
    """1920x1080, 16 wavelengths, R=2044 (Bluestein), K=1023; the reference pads to 4096x4096."""
    p = workloads.c1_params()
    img = workloads.c1_frame()
    assert oracle.sizes(p).raw_psf_size == 2044 and oracle.sizes(p).max_kernel_radius_px == 511
    want32 = oracle.render_image(p, img)
    want64 = oracle.render_image(p, img, fp64=True)
    plain = ctx.render_image(p, img)
    report("c1 reference FP32", want32, want32, want64)
    report("c1 HIP plain per-bin FP32", plain, want32, want64)
    assert rel_l2(plain, want32) <= TOL


def test_c3_batched_request_per_pupil_size(ctx, oracle):
    """One request per pupil size 128, 256, 384, 512, 768, 1024 (r = P/2, R = 2P: 768 and 1536
    are Bluestein sizes), built together in one captured graph, each bank against the oracle's."""
    reqs = []
    for k, P in enumerate((128, 256, 384, 512, 768, 1024)):
        q = workloads.batch_requests(n=1, seed=30 + k, pupils=(P,), wl_range=(8, 8))[0]
        reqs.append(q)
    handles, st = ctx.build_banks(reqs, _abi.LDH_BATCH_GRAPH)
    try:
        banks = [ctx.bank_arrays(h) for h in handles]
    finally:
        for h in handles:
            ctx.bank_release(h)
    assert st.groups == 6
    flips = 0
    for q, got in zip(reqs, banks):
        want = oracle.bank(q)
        worst = 0.0
        for b in range(len(want["full"])):
            flipped = got["full"][b].shape != want["full"][b].shape
            flips += flipped
            for name in ("full", "core", "structure"):
                e = (kernel_rel_l2_ring if flipped else kernel_rel_l2)(got[name][b], want[name][b])
                worst = max(worst, e)
                assert e < TOL, (q.pupil_resolution, b, name, e)
        print("c3 P=%4d R=%4d window %3d: worst kernel rel L2 %.3g"
              % (q.pupil_resolution, got["raw_size"], got["raw_window"], worst))
    # the crop radius thresholds float sums of PSFs that differ from the reference's FP32 ones in
    # their last bits: a bin sitting on a threshold can land one ring apart (reported)
    print("c3: %d of 48 crop radii differ" % flips)
    assert flips <= 2


C4_SIZES = [(1280, 720), (1920, 1080), (2560, 1440), (3840, 2160), (4096, 2160)]


def c4_lens(seed, n_wavelengths, radius_percent):
    p = workloads.spectral_params(2160, 512, 5 + seed % 5, range(2, 16), 0.15, seed, n_wavelengths,
                                  radius_percent)
    p.rotation_deg = 37.0 * seed
    return p


def test_c4_served_sizes_alone_mixed_and_oracle(ctx, oracle):
    """One request at each served frame size with 16- and 31-wavelength lenses: each result is the
    same alone and in the mix (<= 1e-6); a tristimulus lens at every size against the oracle."""
    reqs = []
    for k, (w, h) in enumerate(C4_SIZES):
        img = workloads.frame(w, h, 0.0005, seed=60 + k, log_uniform=(1.0, 5000.0))
        reqs.append((c4_lens(1 + k % 2, 16 if k % 2 == 0 else 31, 20.0), img))
    alone = served_outputs(ctx, reqs, slots=4)
    mixed = served_outputs(ctx, reqs, slots=4, mixed=True)
    for (p, img), a, m in zip(reqs, alone, mixed):
        assert rel_l2(m, a) <= 1e-6, img.shape
    tri = c4_lens(7, 3, 10.0)
    tri.set_wavelengths([])
    tri.spectral_mode = _abi.SPECTRAL_TRISTIMULUS
    outs = served_outputs(ctx, [(tri, img) for _, img in reqs], slots=4, mixed=True)
    for (_, img), got in zip(reqs, outs):
        want32 = oracle.render_image(tri, img)
        e = rel_l2(got, want32)
        print("c4 %dx%d tristimulus served vs oracle FP32: rel L2 %.3g" % (img.shape[1], img.shape[0], e))
        assert e <= TOL

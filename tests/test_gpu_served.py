"""Served form: results through the request loop equal the same requests served alone, and the
plain per-frame path, within tolerance; the cache hits when lens parameters repeat."""
import numpy as np
import pytest

from lensdiff_amd import _abi, workloads
from oracle.oracle import rel_l2

pytestmark = pytest.mark.gpu


def lens(seed, spectral=True, n=8, style=False):
    p = workloads.base_params(96)
    p.pupil_resolution = 64
    p.blade_count = 5 + seed % 4
    p.max_kernel_radius_px = 18.0
    p.set_noll(workloads.noll_coefficients(range(2, 16), 0.15, seed))
    if spectral:
        p.set_wavelengths(np.linspace(400, 700, n))
        p.chromatic_focus = 0.3
    if style:
        p.spectrum_style = _abi.STYLE_CYAN_MAGENTA
        p.spectrum_force = 0.5
    return p


def test_served_mix_matches_alone_and_plain(ctx):
    lenses = [lens(1), lens(2, spectral=False), lens(3, n=12, style=True)]
    sizes = [(160, 96), (200, 120)]
    reqs = []
    rng = np.random.default_rng(0)
    for i in range(10):
        w, h = sizes[i % 2]
        img = workloads.frame(w, h, 0.004, seed=100 + i)
        reqs.append((lenses[int(rng.integers(0, 3))], img))
    bufs = []
    for p, img in reqs:
        s = ctx.alloc(img.nbytes)
        d = ctx.alloc(img.nbytes)
        ctx.upload(s, img)
        bufs.append((s, d))
    # plain per-frame path, one request at a time
    plain = []
    for (p, img), (s, d) in zip(reqs, bufs):
        h, w = img.shape[:2]
        ctx.render_device(p, s, d, w, h)
        out = np.empty_like(img)
        ctx.download(d, out)
        plain.append(out)
    ctx.server_start(3, 4 << 30)
    try:
        alone = []
        for (p, img), (s, d) in zip(reqs, bufs):
            h, w = img.shape[:2]
            ctx.wait(ctx.submit(p, s, d, w, h))
            out = np.empty_like(img)
            ctx.download(d, out)
            alone.append(out)
        tickets = [ctx.submit(p, s, d, img.shape[1], img.shape[0]) for (p, img), (s, d) in zip(reqs, bufs)]
        stats = [ctx.wait(t) for t in tickets]
        mixed = []
        for (p, img), (s, d) in zip(reqs, bufs):
            out = np.empty_like(img)
            ctx.download(d, out)
            mixed.append(out)
        st = ctx.server_stats()
    finally:
        ctx.server_stop()
        for s, d in bufs:
            ctx.free(s)
            ctx.free(d)
    for a, m, pl in zip(alone, mixed, plain):
        assert rel_l2(m, a) <= 1e-6
        assert rel_l2(a, pl) <= 1e-5   # pre-combined spectra vs per-bin convolution
    assert all(s.bank_hit for s in stats)          # the mixed pass reuses the resident banks
    assert st.bank_misses == 3 and st.completed == 20


def test_bank_kernel_sets_follow_the_look(ctx):
    """A render that convolves only the full kernels (a physical look) builds its bank with that
    set alone; the bank interface and a split look get all three. The full kernels are the same
    bit for bit either way (every sum is the sequential FP32 sum, whatever else is summed beside
    it), so a frame is identical whether its bank holds one set or three. Served, a physical and a
    split request of one lens are two cache entries."""
    p = lens(5)
    q = type(p).from_buffer_copy(p)
    q.look_mode = _abi.LOOK_SPLIT
    img = workloads.frame(160, 96, 0.004, seed=7)
    h, w = img.shape[:2]
    one_set = ctx.render_image(p, img)      # bank with the full kernels alone
    sets3 = ctx.bank(p)                      # the bank interface: all three sets, cached for p
    three_sets = ctx.render_image(p, img)   # the cached three-set bank
    assert np.array_equal(one_set, three_sets)
    assert all(np.abs(k).sum() > 0 for k in sets3["core"] + sets3["structure"])
    split = ctx.render_image(q, img)
    s = ctx.alloc(img.nbytes)
    d = ctx.alloc(img.nbytes)
    ctx.upload(s, img)
    ctx.server_start(2, 4 << 30)
    try:
        served = []
        for r in (p, q, p):
            ctx.wait(ctx.submit(r, s, d, w, h))
            out = np.empty_like(img)
            ctx.download(d, out)
            served.append(out)
        st = ctx.server_stats()
    finally:
        ctx.server_stop()
        ctx.free(s)
        ctx.free(d)
    assert rel_l2(served[0], one_set) <= 1e-5   # pre-combined spectra vs per-bin convolution
    assert rel_l2(served[1], split) <= 1e-6
    assert np.array_equal(served[2], served[0])
    assert st.bank_misses == 2 and st.bank_hits == 1


@pytest.mark.parametrize("w,h,style", [(400, 200, False), (1920, 1080, False), (1920, 1080, True), (97, 63, False)],
                         ids=["grid256-fused", "grid1152", "grid1152-styled", "grid128-one-pass"])
def test_served_column_passes_per_grid_height(ctx, w, h, style):
    """A served pre-combined frame's column work takes one of three forms by grid height and its
    tuned split: the driver's last forward pass fused with every channel's first inverse pass
    (256 rows: 16 x 16), the passes apart after F1 where the split's second factor has no fused
    tile (1152 rows when the tuner keeps 32 x 36; it may take 48 x 24, fused), and one pass each
    way (128 rows). Each against the plain per-bin path."""
    p = lens(11, n=8, style=style)
    img = workloads.frame(w, h, 0.002, seed=60)
    plain = ctx.render_image(p, img)
    s = ctx.alloc(img.nbytes)
    d = ctx.alloc(img.nbytes)
    ctx.upload(s, img)
    ctx.server_start(1, 4 << 30)
    try:
        ctx.wait(ctx.submit(p, s, d, w, h))
        out = np.empty_like(img)
        ctx.download(d, out)
    finally:
        ctx.server_stop()
        ctx.free(s)
        ctx.free(d)
    # the split the tuner kept (it keeps the balanced one unless another is more than 5% faster)
    for key, rec in ctx.tuning_report().items():
        if key.startswith("fft-cols/"):
            cands = key[key.index("[") + 1:key.index("]")].split(",")
            print("%s: %s" % (key.split("[")[0], cands[rec["variant"]]))
    assert rel_l2(out, plain) <= 1e-5


def test_served_rejects_unsupported_and_reports(ctx):
    p = lens(4)
    p.aperture_mode = _abi.APERTURE_CUSTOM
    img = workloads.frame(64, 48, 0.01, seed=1)
    s = ctx.alloc(img.nbytes)
    d = ctx.alloc(img.nbytes)
    ctx.server_start(1, 1 << 30)
    try:
        with pytest.raises(Exception) as e:
            ctx.wait(ctx.submit(p, s, d, 64, 48))
        assert "custom-aperture-not-supported" in str(e.value)
    finally:
        ctx.server_stop()
        ctx.free(s)
        ctx.free(d)


def test_served_static_views_and_working_scale_match_plain(ctx):
    """A served static view is drawn from the resident bank like the plain one; a served
    resolution-aware frame convolves at its working size with cached spectra."""
    img = workloads.frame(96, 64, 0.01, seed=3)
    cases = []
    for view in (_abi.VIEW_PSF, _abi.VIEW_OTF, _abi.VIEW_PUPIL):
        p = workloads.base_params(64)
        p.pupil_resolution = 64
        p.max_kernel_radius_px = 20.0
        p.spectral_mode = 1
        p.debug_view = view
        cases.append(p)
    for mode in (0, 1):  # mono (one cached kernel) and tristimulus (pre-combined spectra)
        p = workloads.base_params(64)
        p.pupil_resolution = 64
        p.resolution_aware = 1
        p.frame_short_side_px = 1350
        p.max_kernel_radius_px = 0.8
        p.diffraction_scale_px = 0.1
        p.spectral_mode = mode
        cases.append(p)
    plain = [ctx.render_image(p, img) for p in cases]
    s = ctx.alloc(img.nbytes)
    d = ctx.alloc(img.nbytes)
    ctx.upload(s, img)
    ctx.server_start(1, 1 << 30)
    try:
        for p, want in zip(cases, plain):
            ctx.wait(ctx.submit(p, s, d, 96, 64))
            got = np.empty_like(img)
            ctx.download(d, got)
            assert rel_l2(got, want) <= 1e-5, (p.debug_view, p.spectral_mode)
    finally:
        ctx.server_stop()
        ctx.free(s)
        ctx.free(d)


def test_served_field_zones_match_plain(ctx):
    """A served request with field-dependent phase keeps its nine zone banks resident and renders
    the zone blend like the plain path (pre-combined kernels for the spectral look); its edge views
    come from zone (2, 2)."""
    img = workloads.frame(96, 64, 0.01, seed=8)
    cases = []
    for view in (_abi.VIEW_FINAL, _abi.VIEW_FIELD_PSF, _abi.VIEW_PHASE_EDGE):
        p = workloads.base_params(64)
        p.pupil_resolution = 64
        p.max_kernel_radius_px = 20.0
        p.spectral_mode = 1
        p.set_noll({4: 0.25, 11: -0.15})
        p.phase_field_strength, p.phase_field_edge_bias = 1.0, 0.3
        p.phase_field_coma_radial, p.phase_field_astig_tangential = 0.3, 0.2
        p.debug_view = view
        cases.append(p)
    plain = [ctx.render_image(p, img) for p in cases]
    s = ctx.alloc(img.nbytes)
    d = ctx.alloc(img.nbytes)
    ctx.upload(s, img)
    ctx.server_start(1, 1 << 30)
    try:
        for i, (p, want) in enumerate(zip(cases, plain)):
            st = ctx.wait(ctx.submit(p, s, d, 96, 64))
            assert st.bank_hit == (i > 0)  # one entry holds the zone banks for every view
            got = np.empty_like(img)
            ctx.download(d, got)
            # the served zones convolve with pre-combined kernels, the plain ones per bin
            assert rel_l2(got, want) <= 1e-5, p.debug_view
    finally:
        ctx.server_stop()
        ctx.free(s)
        ctx.free(d)


def test_served_field_zone_spectra_cached(ctx):
    """The nine zones' kernel spectra of a field-dependent phase are kept with the entry (one set per
    grid): a repeated request reuses them, and every result equals the plain render."""
    img = workloads.frame(96, 64, 0.01, seed=9)
    lenses = []
    for spectral in (1, 0):
        p = workloads.base_params(64)
        p.pupil_resolution = 64
        p.max_kernel_radius_px = 20.0
        p.spectral_mode = spectral
        p.set_noll({4: 0.2, 8: -0.1})
        p.phase_field_strength, p.phase_field_edge_bias = 1.0, 0.3
        p.phase_field_coma_radial, p.phase_field_defocus = 0.3, 0.15
        lenses.append(p)
    plain = [ctx.render_image(p, img) for p in lenses]
    s = ctx.alloc(img.nbytes)
    d = ctx.alloc(img.nbytes)
    ctx.upload(s, img)
    ctx.server_start(1, 1 << 30)
    try:
        for p, want in zip(lenses, plain):
            for k in range(3):
                st = ctx.wait(ctx.submit(p, s, d, 96, 64))
                assert st.spectra_hit == (k > 0)
                got = np.empty_like(img)
                ctx.download(d, got)
                assert rel_l2(got, want) <= 1e-5, (p.spectral_mode, k)
        stats = ctx.server_stats()
        assert stats.spectra_misses == 2 and stats.spectra_hits == 4
    finally:
        ctx.server_stop()
        ctx.free(s)
        ctx.free(d)

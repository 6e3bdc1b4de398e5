"""Host-side logic without a GPU: size resolution (LensDiffTypes.h:292-320,
LensDiffCpuReference.cpp:104-115), the workloads' shapes, parameter defaults and the Python seam's
error behaviour on invalid requests."""
import math

import numpy as np
import pytest

from lensdiff_amd import _abi, workloads


def sizes(oracle, **kw):
    p = _abi.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return oracle.sizes(p)


@pytest.mark.parametrize("req,P", [(10, 64), (64, 64), (600, 600), (1024, 1024), (5000, 1024)])
def test_pupil_resolution_clamp(oracle, req, P):
    assert sizes(oracle, pupil_resolution=req).pupil_size == P   # no power-of-two snap in the core


@pytest.mark.parametrize("P,S,pct,R,r", [(256, 512, 50.0, 1024, 256), (512, 1080, 47.3, 2044, 511),
                                         (1024, 2160, 47.36, 4092, 1023), (64, 100, 1.0, 512, 4),
                                         (1024, 2160, 200.0, 16384, 4320)])
def test_raw_size_rule(oracle, P, S, pct, R, r):
    s = sizes(oracle, pupil_resolution=P, frame_short_side_px=S, max_kernel_radius_px=pct)
    assert (s.raw_psf_size, s.max_kernel_radius_px) == (R, r)
    assert s.raw_psf_size == min(16384, max(128, max(2 * P, 512, 4 * r)))


def test_resolution_aware_uses_2160(oracle):
    s = sizes(oracle, frame_short_side_px=1080, resolution_aware=1, diffraction_scale_px=2.0)
    assert s.optics_short_side_px == 2160 and math.isclose(s.diffraction_scale_px, 43.2)


def test_workload_shapes(oracle):
    c0, c1, c2 = workloads.c0_params(), workloads.c1_params(), workloads.c2_params()
    for p, (P, R, r) in ((c0, (256, 1024, 256)), (c1, (512, 2044, 511)), (c2, (1024, 4092, 1023))):
        s = oracle.sizes(p)
        assert (s.pupil_size, s.raw_psf_size, s.max_kernel_radius_px) == (P, R, r)
    assert list(oracle.wavelengths(c1)) == list(np.linspace(400, 700, 16).astype(np.float32))
    assert len(oracle.wavelengths(c2)) == 31
    f = workloads.c2_frame()
    hl = (f[..., 0] > 1.0).sum()
    assert f.shape == (2160, 3840, 4) and hl == round(3840 * 2160 * 0.0005)
    assert f[..., :3].max() <= 5000.0 and f[f[..., 0] > 1.0, 0].min() >= 1.0
    f0 = workloads.c0_frame()
    assert (f0[..., 0] > 1.0).sum() == round(512 * 512 * 0.001) and f0[..., 0].max() <= 1000.0


def test_fixed_wavelength_lists(oracle):
    want = {0: [550], 1: [460, 550, 610], 2: [440, 490, 540, 590, 640],
            3: [420, 450, 480, 510, 540, 570, 600, 630, 660]}
    for mode, wl in want.items():
        p = _abi.default_params()
        p.spectral_mode = mode
        assert list(oracle.wavelengths(p)) == wl


def test_noll_mapping_matches_reference_names():
    p = _abi.default_params()
    p.set_noll({4: 1.0, 5: 2.0, 6: 3.0, 7: 4.0, 8: 5.0, 11: 6.0, 2: 7.0, 21: 8.0})
    assert (p.phase_defocus, p.phase_astigmatism45, p.phase_astigmatism0, p.phase_coma_y, p.phase_coma_x,
            p.phase_spherical) == (1.0, 2.0, 3.0, 4.0, 5.0, 6.0)
    assert p.zernike_extra[0] == 7.0 and p.zernike_extra[5] == 8.0 and p.phase_mode == _abi.PHASE_ENABLED
    with pytest.raises(ValueError):
        p.set_noll({22: 1.0})


def test_oracle_render_error_tags(oracle):
    p = _abi.default_params()
    req = _abi.RenderRequest()
    ok, err = oracle.render(p, req)
    assert not ok and err == "missing-source-or-destination-buffer"
    img = np.zeros((4, 4, 4), np.float32)
    req.src.data = img.ctypes.data
    req.dst.data = img.ctypes.data
    ok, err = oracle.render(p, req)
    assert not ok and err == "invalid-source-bounds"


def test_reference_render_error_tags(reference):
    p = _abi.default_params()
    ok, err = reference.render(p, _abi.RenderRequest())
    assert not ok and err == "missing-source-or-destination-buffer"


@pytest.mark.parametrize("n", [2, 8])
def test_bench_refuses_more_than_one_gpu(n):
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says so, before any
    device work (nothing is printed on stdout, so no result line can be mistaken for one)."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert r.stdout == ""
    assert "one GPU" in r.stderr

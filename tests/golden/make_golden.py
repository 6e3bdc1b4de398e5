"""Writes tests/golden/reference_vectors.npz from the compiled reference CPU path.

Run in a container that has the reference sources (oracle/Makefile target `ref` builds
oracle/_ref/liblensdiff_ref.so from them):  python tests/golden/make_golden.py
The fixtures are inputs and the reference's outputs only; the CPU tests check the oracle
against them where the reference itself is absent.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from lensdiff_amd import _abi, workloads  # noqa: E402
from oracle.oracle import Reference  # noqa: E402

OUT = os.path.join(HERE, "reference_vectors.npz")


def params(**kw):
    p = _abi.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


FIELD = dict(phase_field_strength=1.0, phase_field_edge_bias=0.3, phase_field_defocus=0.15,
             phase_field_coma_radial=0.3, phase_field_astig_tangential=0.2, phase_field_trefoil_radial=0.1,
             phase_field_quadrafoil_tangential=0.05, phase_field_secondary_coma_radial=-0.07)


def pack_params(p) -> np.ndarray:
    return np.frombuffer(bytes(p), np.uint8).copy()


def main():
    ref = Reference()
    out = {}
    rng = np.random.default_rng(2024)
    # pupil amplitude per aperture mode, phase with every reference term
    for mode in (0, 1, 2, 3, 5, 6, 7):
        p = params(aperture_mode=mode, pupil_resolution=64, blade_count=7, roundness=0.3, rotation_deg=17.0,
                   central_obstruction=0.2, vane_count=3, vane_thickness=0.02, apodization_mode=mode % 3)
        out["pupil_%d_params" % mode] = pack_params(p)
        out["pupil_%d" % mode] = ref.pupil_amplitude(p, 64)
    p = params(pupil_resolution=64, rotation_deg=11.0, pupil_decenter_x=0.03, pupil_decenter_y=-0.02)
    p.set_noll({j: float(rng.uniform(-0.3, 0.3)) for j in range(4, 18)})
    out["phase_params"] = pack_params(p)
    out["phase"] = ref.pupil_phase(p, 64)
    # raw PSFs: power-of-two and Bluestein sizes, flat and aberrated
    amp = ref.pupil_amplitude(params(aperture_mode=1, pupil_resolution=64, blade_count=6), 64)
    out["psf_amp"] = amp
    out["psf_phase"] = out["phase"]
    for raw in (256, 300):
        out["psf_flat_%d" % raw] = ref.raw_psf(amp, None, raw)
        out["psf_phase_%d" % raw] = ref.raw_psf(amp, out["phase"], raw)
    # 1-D FFTs
    for n in (37, 64, 100):
        z = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        out["fft_in_%d" % n] = z
        out["fft_fwd_%d" % n] = ref.fft1d(z).astype(np.complex64)
        out["fft_inv_%d" % n] = ref.fft1d(z, inverse=True).astype(np.complex64)
    z = (rng.standard_normal((12, 20)) + 1j * rng.standard_normal((12, 20))).astype(np.complex64)
    out["fft2_in"] = z
    out["fft2_fwd"] = ref.fft2d(z).astype(np.complex64)
    # first minimum of the clear-circle PSF
    circ = ref.pupil_amplitude(params(aperture_mode=0, roundness=1.0, pupil_resolution=64), 64)
    cpsf = ref.raw_psf(circ, None, 512)
    out["first_min_psf"] = cpsf
    out["first_min"] = np.float32(ref.first_minimum_radius(cpsf))
    # resampling and shaping
    psf = out["psf_phase_300"]
    for i, (scale, support) in enumerate(((0.37, 20), (2.5, 15), (1.0, 12))):
        out["resample_%d" % i] = ref.resample_kernel(psf, scale, support)
        out["resample_%d_args" % i] = np.array([scale, support], np.float64)
    base = out["resample_0"]
    for aniso in (0.0, 0.3):
        p = params(pupil_resolution=64, frame_short_side_px=200, max_kernel_radius_px=8.0, anisotropy_emphasis=aniso)
        f, c, s, r = ref.shape_kernels(p, base)
        key = "shape_%d" % int(aniso * 10)
        out[key + "_params"] = pack_params(p)
        out[key + "_full"], out[key + "_core"], out[key + "_structure"] = f, c, s
        out[key + "_radius"] = np.int32(r)
    # banks for the four spectral modes
    for mode in range(4):
        p = workloads.base_params(200)
        p.pupil_resolution = 64
        p.blade_count = 7
        p.max_kernel_radius_px = 16.0
        p.spectral_mode = mode
        p.chromatic_focus = 0.2
        p.set_noll({4: 0.2, 11: -0.1})
        b = ref.bank(p)
        key = "bank_%d" % mode
        out[key + "_params"] = pack_params(p)
        out[key + "_support"] = np.int32(b["support"])
        for i in range(len(b["full"])):
            out["%s_%d_full" % (key, i)] = b["full"][i]
            out["%s_%d_core" % (key, i)] = b["core"][i]
            out["%s_%d_structure" % (key, i)] = b["structure"][i]
    # spectrum matrices
    for mode, wl in ((1, [460, 550, 610]), (2, [440, 490, 540, 590, 640]),
                     (3, [420, 450, 480, 510, 540, 570, 600, 630, 660]), (0, [500, 620])):
        for style in range(3):
            p = params(spectral_mode=mode, spectrum_style=style)
            nat, sty = ref.spectrum_config(p, np.array(wl, np.float32))
            out["spectrum_%d_%d_nat" % (mode, style)] = nat
            out["spectrum_%d_%d_sty" % (mode, style)] = sty
            out["spectrum_%d_%d_wl" % (mode, style)] = np.array(wl, np.float32)
    # selection, convolution
    img = (rng.random((24, 32, 4)) * 3).astype(np.float32)
    p = params(threshold=1.0, softness_stops=0.5, point_emphasis=0.3)
    out["selection_img"] = img
    out["selection_params"] = pack_params(p)
    m, r_, d = ref.selection(p, img)
    out["selection_mask"], out["selection_redist"], out["selection_driver"] = m, r_, d
    plane = rng.random((20, 27)).astype(np.float32)
    ker = rng.random((7, 7)).astype(np.float32)
    out["conv_plane"], out["conv_kernel"] = plane, ker
    out["conv_out"] = ref.convolve(plane, ker)
    # end-to-end frames
    cases = {
        "mono_physical_augment": dict(),
        "mono_split_preserve": dict(look_mode=1, energy_mode=0, core_preserve=0.1, core_shoulder=2.0),
        "tristimulus": dict(spectral_mode=1, chromatic_focus=0.3),
        "spectral9_split": dict(spectral_mode=3, look_mode=1),
        # field-dependent phase: nine zone banks blended with tent weights
        "field_zones_split": dict(look_mode=1, **FIELD),
        "field_zones_tristimulus": dict(spectral_mode=1, chromatic_focus=0.3, energy_mode=1, **FIELD),
        # post passes on the effect: Gaussian scatter and creative fringe, and their views
        "scatter_fringe_split": dict(look_mode=1, scatter_amount=0.3, scatter_radius=2.5, creative_fringe=1.5),
        "scatter_view": dict(scatter_amount=0.5, scatter_radius=3.0, debug_view=_abi.VIEW_SCATTER),
        "fringe_view_preserve": dict(creative_fringe=2.0, energy_mode=_abi.ENERGY_PRESERVE,
                                     debug_view=_abi.VIEW_CREATIVE_FRINGE),
    }
    frame = workloads.frame(48, 40, 0.01, seed=7)
    out["frame_in"] = frame
    for name, kw in cases.items():
        p = workloads.base_params(40)
        p.pupil_resolution = 64
        p.max_kernel_radius_px = 20.0
        p.set_noll({4: 0.25, 11: -0.15})
        for k, v in kw.items():
            setattr(p, k, v)
        out["frame_%s_params" % name] = pack_params(p)
        out["frame_%s" % name] = ref.render_image(p, frame)
    # resolution-aware: the frame resampled to the working size (optics short side 2160 over the
    # frame's) and the display images resampled back
    for name, fss, kw in (("ra_up", 1080, dict(spectral_mode=1)),
                          ("ra_down_split_preserve", 4320, dict(look_mode=1, energy_mode=0)),
                          ("ra_selection", 1350, dict(debug_view=_abi.VIEW_SELECTION)),
                          ("ra_scatter_fringe", 1080, dict(scatter_amount=0.4, scatter_radius=0.1,
                                                           creative_fringe=0.05, input_transfer=1))):
        p = workloads.base_params(40)
        p.pupil_resolution = 64
        p.resolution_aware = 1
        p.frame_short_side_px = fss
        p.max_kernel_radius_px = 0.8
        p.diffraction_scale_px = 0.1
        p.set_noll({4: 0.25, 11: -0.15})
        for k, v in kw.items():
            setattr(p, k, v)
        img = workloads.davinci_intermediate_encode(frame) if p.input_transfer == 1 else frame
        out["frame_%s_params" % name] = pack_params(p)
        out["frame_%s_in" % name] = img
        out["frame_%s" % name] = ref.render_image(p, img)
    # static debug views (pupil, PSF, phase, edge-zone phase and PSF, chromatic split)
    for name, kw in (("view_pupil", dict(debug_view=_abi.VIEW_PUPIL)),
                     ("view_psf_spectral", dict(debug_view=_abi.VIEW_PSF, spectral_mode=2)),
                     ("view_otf", dict(debug_view=_abi.VIEW_OTF)),
                     ("view_phase", dict(debug_view=_abi.VIEW_PHASE)),
                     ("view_phase_edge_field", dict(debug_view=_abi.VIEW_PHASE_EDGE, **FIELD)),
                     ("view_field_psf_field", dict(debug_view=_abi.VIEW_FIELD_PSF, **FIELD)),
                     ("view_chromatic_split", dict(debug_view=_abi.VIEW_CHROMATIC_SPLIT, spectral_mode=3))):
        p = workloads.base_params(40)
        p.pupil_resolution = 64
        p.max_kernel_radius_px = 20.0
        p.set_noll({4: 0.25, 11: -0.15, 7: 0.1})
        for k, v in kw.items():
            setattr(p, k, v)
        out["frame_%s_params" % name] = pack_params(p)
        out["frame_%s" % name] = ref.render_image(p, frame)
    np.savez_compressed(OUT, **out)
    print("wrote %s (%d arrays, %.1f KiB)" % (OUT, len(out), os.path.getsize(OUT) / 1024))


if __name__ == "__main__":
    main()

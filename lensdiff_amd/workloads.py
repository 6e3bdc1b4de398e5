"""Seeded synthetic workloads.

c0  reference CPU case: 512x512 RGB, 0.1% highlights, P=256 6-blade polygon, Noll 4/11, mono 550 nm
c1  1920x1080 RGBA, P=512 7 blades, Noll 2-15 N(0, 0.15), 16 wavelengths 400-700 nm, K=1023
c2  3840x2160 RGBA, 0.05% highlights log-uniform [1, 5000], P=1024, Noll 2-21 N(0, 0.2),
    31 wavelengths 400-700 nm, K=2047
Build-chosen constants (recorded here): roundness 0, rotation 0, diffraction scale 2% of the
short side, threshold log2(1/0.18) stops (highlights are > 1.0 linear, background <= 0.5),
softness 0.01 stops, core preserve 0, Augment energy, core compensation 0, Physical look,
chromatic focus 0.3 waves for spectral banks.
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi

THRESHOLD_STOPS = math.log2(1.0 / 0.18)


def frame(width: int, height: int, density: float, seed: int, log_uniform=None, alpha=1.0) -> np.ndarray:
    """RGBA float32 frame: background U[0, 0.5) per channel, `density` of the pixels set to
    white highlights. Highlights: 10 + Exp(mean 100) resampled to <= 1000 (default), or
    log-uniform in `log_uniform` = (lo, hi)."""
    rng = np.random.default_rng(seed)
    img = np.empty((height, width, 4), np.float32)
    img[..., :3] = rng.random((height, width, 3), dtype=np.float32) * 0.5
    img[..., 3] = alpha
    n = max(1, int(round(width * height * density)))
    idx = rng.choice(width * height, n, replace=False)
    if log_uniform is None:
        v = 10.0 + rng.exponential(100.0, n)
        while (v > 1000.0).any():
            bad = v > 1000.0
            v[bad] = 10.0 + rng.exponential(100.0, int(bad.sum()))
    else:
        lo, hi = log_uniform
        v = np.exp(rng.uniform(math.log(lo), math.log(hi), n))
    flat = img.reshape(-1, 4)
    flat[idx, 0] = v
    flat[idx, 1] = v
    flat[idx, 2] = v
    return img


def davinci_intermediate_encode(rgba: np.ndarray) -> np.ndarray:
    """Linear -> DaVinci Intermediate on RGB (LensDiffTransfer.cpp:9-21); alpha unchanged."""
    out = rgba.copy()
    x = rgba[..., :3].astype(np.float64)
    enc = np.where(x <= 0.00262409, x * 10.44426855,
                   (np.log2(np.maximum(x, 0.0) + 0.0075) + 7.0) * 0.07329248)
    out[..., :3] = enc.astype(np.float32)
    return out


def base_params(short_side: int) -> _abi.Params:
    p = _abi.default_params()
    p.frame_short_side_px = short_side
    p.aperture_mode = _abi.APERTURE_POLYGON
    p.roundness = 0.0
    p.rotation_deg = 0.0
    p.diffraction_scale_px = 2.0
    p.threshold = THRESHOLD_STOPS
    p.softness_stops = 0.01
    p.point_emphasis = 0.0
    p.core_preserve = 0.0
    p.energy_mode = _abi.ENERGY_AUGMENT
    p.core_compensation = 0.0
    p.effect_gain = 1.0
    p.look_mode = _abi.LOOK_PHYSICAL
    p.input_transfer = _abi.TRANSFER_LINEAR
    return p


def noll_coefficients(indices, sigma, seed) -> dict:
    rng = np.random.default_rng(seed)
    return {j: float(rng.normal(0.0, sigma)) for j in indices}


def c0_params() -> _abi.Params:
    p = base_params(512)
    p.pupil_resolution = 256
    p.blade_count = 6
    p.max_kernel_radius_px = 50.0
    rng = np.random.default_rng(1)
    p.set_noll({4: float(rng.uniform(-0.5, 0.5)), 11: float(rng.uniform(-0.5, 0.5))})
    p.spectral_mode = _abi.SPECTRAL_MONO
    return p


def c0_frame() -> np.ndarray:
    return frame(512, 512, 0.001, seed=0)


def spectral_params(short_side, pupil, blades, noll, sigma, seed, n_wavelengths, max_radius_percent) -> _abi.Params:
    p = base_params(short_side)
    p.pupil_resolution = pupil
    p.blade_count = blades
    p.max_kernel_radius_px = max_radius_percent
    p.set_noll(noll_coefficients(noll, sigma, seed))
    p.set_wavelengths(np.linspace(400.0, 700.0, n_wavelengths))
    p.chromatic_focus = 0.3
    return p


def c1_params() -> _abi.Params:
    # r = ceil(47.3% of 1080) = 511 -> K = 1023, R = max(1024, 512, 2044) = 2044
    return spectral_params(1080, 512, 7, range(2, 16), 0.15, 1, 16, 47.3)


def c1_frame() -> np.ndarray:
    return frame(1920, 1080, 0.001, seed=0)


def c2_params() -> _abi.Params:
    # r = ceil(47.36% of 2160) = 1023 -> K = 2047, R = 4092
    return spectral_params(2160, 1024, 7, range(2, 22), 0.2, 2, 31, 47.36)


def c2_frame() -> np.ndarray:
    return frame(3840, 2160, 0.0005, seed=0, log_uniform=(1.0, 5000.0))


def scaled_spectral_params(short_side, pupil, n_wavelengths, radius_px, seed=5, blades=7,
                           noll=range(2, 16), sigma=0.15) -> _abi.Params:
    """A spectral workload shaped like c1/c2 at a smaller size (for parity tests)."""
    percent = 100.0 * (radius_px - 0.5) / short_side
    return spectral_params(short_side, pupil, blades, noll, sigma, seed, n_wavelengths, percent)


def batch_requests(n=64, seed=3, pupils=(128, 256, 384, 512, 768, 1024), wl_range=(8, 31), noll_sigma=0.15):
    """PSF-synthesis requests of the batched form: pupil sizes drawn from `pupils`, 5-9 blades,
    rotation uniform over the circle, Noll 2-15 N(0, noll_sigma) waves, 8-31 wavelengths evenly
    over 400-700 nm. Each lens is set on a frame whose short side equals its pupil size with a
    50% kernel radius, so r = P/2 and R = 2P (Bluestein for P = 384 and 768)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        P = int(rng.choice(pupils))
        p = base_params(P)
        p.pupil_resolution = P
        p.blade_count = int(rng.integers(5, 10))
        p.rotation_deg = float(rng.uniform(0.0, 360.0))
        p.max_kernel_radius_px = 50.0
        p.set_noll(noll_coefficients(range(2, 16), noll_sigma, seed * 1000 + i))
        p.set_wavelengths(np.linspace(400.0, 700.0, int(rng.integers(wl_range[0], wl_range[1] + 1))))
        p.chromatic_focus = 0.3
        out.append(p)
    return out

"""ctypes access to the CPU oracle (liblensdiff_oracle.so) and, where it was built, the compiled
reference CPU path (_ref/liblensdiff_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline, as the checker. The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from lensdiff_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liblensdiff_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "liblensdiff_ref.so")
REF_SOURCES = "/root/reference/source/src/core/LensDiffCpuReference.cpp"

_F = C.POINTER(C.c_float)
_Dp = C.POINTER(C.c_double)
_P = C.POINTER(_abi.Params)


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference shim when the reference sources are present)."""
    args = ["make", "-C", HERE]
    if force:
        args.append("-B")
    subprocess.check_call(args + ["all"])
    if os.path.exists(REF_SOURCES):
        subprocess.check_call(args + ["ref"])


def _ptr(a: np.ndarray, dtype):
    assert a.dtype == dtype and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_F if dtype == np.float32 else _Dp)


class Oracle:
    """Both instantiations of the restatement: fp64=False is FP32 faithful, True is FP64."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError("oracle library not built: %s (run make -C oracle)" % path)
        self.lib = C.CDLL(path)
        L = self.lib
        L.orc_params_default.argtypes = [_P]
        L.orc_resolve_sizes.argtypes = [_P, C.POINTER(_abi.Sizes)]
        L.orc_wavelengths.argtypes = [_P, _F]
        L.orc_wavelengths.restype = C.c_int
        L.orc_sequential_sum_f32.argtypes = [_F, C.c_long]
        L.orc_sequential_sum_f32.restype = C.c_float
        for s, T, TP in (("f32", C.c_float, _F), ("f64", C.c_double, _Dp)):
            getattr(L, "orc_pupil_amplitude_" + s).argtypes = [_P, TP]
            getattr(L, "orc_pupil_phase_" + s).argtypes = [_P, TP]
            getattr(L, "orc_raw_psf_" + s).argtypes = [TP, TP, C.c_int, C.c_int, TP]
            getattr(L, "orc_fft1d_" + s).argtypes = [TP, C.c_int, C.c_int]
            getattr(L, "orc_fft2d_" + s).argtypes = [TP, C.c_int, C.c_int, C.c_int]
            f = getattr(L, "orc_first_minimum_radius_" + s)
            f.argtypes = [TP, C.c_int]
            f.restype = T
            getattr(L, "orc_resample_kernel_" + s).argtypes = [TP, C.c_int, T, C.c_int, TP]
            f = getattr(L, "orc_shape_kernels_" + s)
            f.argtypes = [_P, TP, C.c_int, TP, TP, TP, C.POINTER(C.c_int)]
            f.restype = C.c_int
            f = getattr(L, "orc_bank_build_" + s)
            f.argtypes = [_P]
            f.restype = C.c_void_p
            getattr(L, "orc_spectrum_config_" + s).argtypes = [_P, C.c_int, _F, TP, TP]
            getattr(L, "orc_map_bins_" + s).argtypes = [_P, C.c_int, TP, TP, TP, C.c_long, TP]
            getattr(L, "orc_selection_" + s).argtypes = [_P, _F, C.c_int, C.c_int, TP, TP, TP]
            getattr(L, "orc_convolve_" + s).argtypes = [TP, C.c_int, C.c_int, TP, C.c_int, TP]
            f = getattr(L, "orc_render_image_" + s)
            f.argtypes = [_P, _F, C.c_int, C.c_int, TP]
            f.restype = C.c_int
            getattr(L, "orc_bank_kernel_" + s).argtypes = [C.c_void_p, C.c_int, C.c_int, TP]
            getattr(L, "orc_bank_base_" + s).argtypes = [C.c_void_p, C.c_int, TP]
        for n in ("bin_count", "support_radius", "raw_size", "base_size"):
            f = getattr(L, "orc_bank_" + n)
            f.argtypes = [C.c_void_p]
            f.restype = C.c_int
        L.orc_bank_kernel_size.argtypes = [C.c_void_p, C.c_int]
        L.orc_bank_kernel_size.restype = C.c_int
        L.orc_bank_wavelength.argtypes = [C.c_void_p, C.c_int]
        L.orc_bank_wavelength.restype = C.c_float
        L.orc_bank_free.argtypes = [C.c_void_p]
        L.orc_render_f32.argtypes = [_P, C.POINTER(_abi.RenderRequest), C.c_char_p, C.c_int]
        L.orc_render_f32.restype = C.c_int

    # -- helpers -------------------------------------------------------------------------
    @staticmethod
    def _t(fp64):
        return (np.float64, "f64") if fp64 else (np.float32, "f32")

    def sizes(self, p: _abi.Params) -> _abi.Sizes:
        s = _abi.Sizes()
        self.lib.orc_resolve_sizes(C.byref(p), C.byref(s))
        return s

    def wavelengths(self, p: _abi.Params) -> np.ndarray:
        out = np.zeros(_abi.LDH_MAX_WAVELENGTHS, np.float32)
        n = self.lib.orc_wavelengths(C.byref(p), _ptr(out, np.float32))
        return out[:n].copy()

    def pupil_amplitude(self, p, fp64=False):
        dt, s = self._t(fp64)
        n = self.sizes(p).pupil_size
        out = np.zeros((n, n), dt)
        getattr(self.lib, "orc_pupil_amplitude_" + s)(C.byref(p), _ptr(out, dt))
        return out

    def pupil_phase(self, p, fp64=False):
        dt, s = self._t(fp64)
        n = self.sizes(p).pupil_size
        out = np.zeros((n, n), dt)
        getattr(self.lib, "orc_pupil_phase_" + s)(C.byref(p), _ptr(out, dt))
        return out

    def raw_psf(self, amp, phase, raw_size, fp64=False):
        dt, s = self._t(fp64)
        amp = np.ascontiguousarray(amp, dt)
        ph = None if phase is None else np.ascontiguousarray(phase, dt)
        out = np.zeros((raw_size, raw_size), dt)
        getattr(self.lib, "orc_raw_psf_" + s)(_ptr(amp, dt), None if ph is None else _ptr(ph, dt),
                                               amp.shape[0], raw_size, _ptr(out, dt))
        return out

    def fft1d(self, z: np.ndarray, inverse=False, fp64=False):
        dt, s = self._t(fp64)
        a = np.ascontiguousarray(np.stack([z.real, z.imag], -1), dt)
        getattr(self.lib, "orc_fft1d_" + s)(_ptr(a, dt), z.shape[0], int(inverse))
        return a[..., 0] + 1j * a[..., 1]

    def fft2d(self, z: np.ndarray, inverse=False, fp64=False):
        dt, s = self._t(fp64)
        h, w = z.shape
        a = np.ascontiguousarray(np.stack([z.real, z.imag], -1), dt)
        getattr(self.lib, "orc_fft2d_" + s)(_ptr(a, dt), w, h, int(inverse))
        return a[..., 0] + 1j * a[..., 1]

    def first_minimum_radius(self, raw, fp64=False):
        dt, s = self._t(fp64)
        raw = np.ascontiguousarray(raw, dt)
        return float(getattr(self.lib, "orc_first_minimum_radius_" + s)(_ptr(raw, dt), raw.shape[0]))

    def resample_kernel(self, raw, scale, support, fp64=False):
        dt, s = self._t(fp64)
        raw = np.ascontiguousarray(raw, dt)
        k = 2 * support + 1
        out = np.zeros((k, k), dt)
        getattr(self.lib, "orc_resample_kernel_" + s)(_ptr(raw, dt), raw.shape[0], scale, support,
                                                       _ptr(out, dt))
        return out

    def shape_kernels(self, p, base, fp64=False):
        dt, s = self._t(fp64)
        base = np.ascontiguousarray(base, dt)
        k = base.shape[0]
        outs = [np.zeros(k * k, dt) for _ in range(3)]
        size = C.c_int(0)
        r = getattr(self.lib, "orc_shape_kernels_" + s)(C.byref(p), _ptr(base, dt), k,
                                                         *[_ptr(o, dt) for o in outs], C.byref(size))
        n = size.value
        full, core, struct = (o[: n * n].reshape(n, n).copy() for o in outs)
        return full, core, struct, r

    def bank(self, p, fp64=False):
        """Returns dict(wavelengths, support, raw_size, full/core/structure lists, base list)."""
        dt, s = self._t(fp64)
        L = self.lib
        h = getattr(L, "orc_bank_build_" + s)(C.byref(p))
        try:
            n = L.orc_bank_bin_count(h)
            out = {"wavelengths": [], "support": L.orc_bank_support_radius(h),
                   "raw_size": L.orc_bank_raw_size(h), "full": [], "core": [], "structure": [],
                   "base": []}
            kb = L.orc_bank_base_size(h)
            for b in range(n):
                k = L.orc_bank_kernel_size(h, b)
                out["wavelengths"].append(L.orc_bank_wavelength(h, b))
                for which, name in ((0, "full"), (1, "core"), (2, "structure")):
                    a = np.zeros((k, k), dt)
                    getattr(L, "orc_bank_kernel_" + s)(h, b, which, _ptr(a, dt))
                    out[name].append(a)
                a = np.zeros((kb, kb), dt)
                getattr(L, "orc_bank_base_" + s)(h, b, _ptr(a, dt))
                out["base"].append(a)
            return out
        finally:
            L.orc_bank_free(h)

    def spectrum_config(self, p, wavelengths, fp64=False):
        dt, s = self._t(fp64)
        wl = np.ascontiguousarray(wavelengths, np.float32)
        n = wl.shape[0]
        nat = np.zeros((3, n), dt)
        sty = np.zeros((3, n), dt)
        getattr(self.lib, "orc_spectrum_config_" + s)(C.byref(p), n, _ptr(wl, np.float32),
                                                       _ptr(nat, dt), _ptr(sty, dt))
        return nat, sty

    def map_bins(self, p, nat, sty, bins, fp64=False):
        dt, s = self._t(fp64)
        bins = np.ascontiguousarray(bins, dt)
        n = bins.shape[0]
        npx = int(np.prod(bins.shape[1:]))
        nat = np.ascontiguousarray(nat, dt)
        sty = np.ascontiguousarray(sty, dt)
        rgb = np.zeros((npx, 3), dt)
        getattr(self.lib, "orc_map_bins_" + s)(C.byref(p), n, _ptr(nat, dt), _ptr(sty, dt),
                                                _ptr(bins, dt), npx, _ptr(rgb, dt))
        return rgb.reshape(bins.shape[1:] + (3,))

    def selection(self, p, rgba, fp64=False):
        dt, s = self._t(fp64)
        rgba = np.ascontiguousarray(rgba, np.float32)
        h, w = rgba.shape[:2]
        mask = np.zeros((h, w), dt)
        red = np.zeros((h, w, 4), dt)
        drv = np.zeros((h, w), dt)
        getattr(self.lib, "orc_selection_" + s)(C.byref(p), _ptr(rgba, np.float32), w, h,
                                                 _ptr(mask, dt), _ptr(red, dt), _ptr(drv, dt))
        return mask, red, drv

    def convolve(self, plane, kernel, fp64=False):
        dt, s = self._t(fp64)
        plane = np.ascontiguousarray(plane, dt)
        kernel = np.ascontiguousarray(kernel, dt)
        h, w = plane.shape
        out = np.zeros((h, w), dt)
        getattr(self.lib, "orc_convolve_" + s)(_ptr(plane, dt), w, h, _ptr(kernel, dt),
                                                kernel.shape[0], _ptr(out, dt))
        return out

    def render_image(self, p, rgba, fp64=False):
        dt, s = self._t(fp64)
        rgba = np.ascontiguousarray(rgba, np.float32)
        h, w = rgba.shape[:2]
        out = np.zeros((h, w, 4), dt)
        ok = getattr(self.lib, "orc_render_image_" + s)(C.byref(p), _ptr(rgba, np.float32), w, h,
                                                         _ptr(out, dt))
        if not ok:
            raise RuntimeError("oracle render failed")
        return out

    def render(self, p, req) -> tuple[bool, str]:
        err = C.create_string_buffer(64)
        ok = self.lib.orc_render_f32(C.byref(p), C.byref(req), err, 64)
        return bool(ok), err.value.decode()

    def sequential_sum(self, x: np.ndarray) -> float:
        x = np.ascontiguousarray(x, np.float32).ravel()
        return float(self.lib.orc_sequential_sum_f32(_ptr(x, np.float32), x.size))


class Reference:
    """The reference CPU path itself (compiled by oracle/Makefile `ref`); FP32 only."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        L = self.lib
        L.ref_pupil_amplitude.argtypes = [_P, _F]
        L.ref_pupil_phase.argtypes = [_P, _F]
        L.ref_raw_psf.argtypes = [_F, _F, C.c_int, C.c_int, _F]
        L.ref_fft1d.argtypes = [_F, C.c_int, C.c_int]
        L.ref_fft2d.argtypes = [_F, C.c_int, C.c_int, C.c_int]
        L.ref_first_minimum_radius.argtypes = [_F, C.c_int]
        L.ref_first_minimum_radius.restype = C.c_float
        L.ref_resample_kernel.argtypes = [_F, C.c_int, C.c_float, C.c_int, _F]
        L.ref_shape_kernels.argtypes = [_P, _F, C.c_int, _F, _F, _F, C.POINTER(C.c_int)]
        L.ref_shape_kernels.restype = C.c_int
        L.ref_bank_build.argtypes = [_P]
        L.ref_bank_build.restype = C.c_void_p
        for n in ("bin_count", "support_radius"):
            getattr(L, "ref_bank_" + n).argtypes = [C.c_void_p]
            getattr(L, "ref_bank_" + n).restype = C.c_int
        L.ref_bank_kernel_size.argtypes = [C.c_void_p, C.c_int]
        L.ref_bank_kernel_size.restype = C.c_int
        L.ref_bank_wavelength.argtypes = [C.c_void_p, C.c_int]
        L.ref_bank_wavelength.restype = C.c_float
        L.ref_bank_kernel.argtypes = [C.c_void_p, C.c_int, C.c_int, _F]
        L.ref_bank_free.argtypes = [C.c_void_p]
        L.ref_spectrum_config.argtypes = [_P, C.c_int, _F, _F, _F]
        L.ref_map_bins.argtypes = [_P, C.c_int, _F, _F, C.c_long, _F]
        L.ref_selection.argtypes = [_P, _F, C.c_int, C.c_int, _F, _F, _F]
        L.ref_convolve.argtypes = [_F, C.c_int, C.c_int, _F, C.c_int, _F]
        L.ref_render.argtypes = [_P, C.POINTER(_abi.RenderRequest), C.c_char_p, C.c_int]
        L.ref_render.restype = C.c_int
        L.ref_render_image.argtypes = [_P, _F, C.c_int, C.c_int, _F]
        L.ref_render_image.restype = C.c_int
        L.ref_time_raw_psf.argtypes = [_P]
        L.ref_time_raw_psf.restype = C.c_double
        L.ref_time_convolve.argtypes = [C.c_int, C.c_int, C.c_int]
        L.ref_time_convolve.restype = C.c_double

    def pupil_amplitude(self, p, n):
        out = np.zeros((n, n), np.float32)
        self.lib.ref_pupil_amplitude(C.byref(p), _ptr(out, np.float32))
        return out

    def pupil_phase(self, p, n):
        out = np.zeros((n, n), np.float32)
        self.lib.ref_pupil_phase(C.byref(p), _ptr(out, np.float32))
        return out

    def raw_psf(self, amp, phase, raw_size):
        amp = np.ascontiguousarray(amp, np.float32)
        ph = None if phase is None else np.ascontiguousarray(phase, np.float32)
        out = np.zeros((raw_size, raw_size), np.float32)
        self.lib.ref_raw_psf(_ptr(amp, np.float32), None if ph is None else _ptr(ph, np.float32),
                             amp.shape[0], raw_size, _ptr(out, np.float32))
        return out

    def fft1d(self, z, inverse=False):
        a = np.ascontiguousarray(np.stack([z.real, z.imag], -1), np.float32)
        self.lib.ref_fft1d(_ptr(a, np.float32), z.shape[0], int(inverse))
        return a[..., 0] + 1j * a[..., 1]

    def fft2d(self, z, inverse=False):
        h, w = z.shape
        a = np.ascontiguousarray(np.stack([z.real, z.imag], -1), np.float32)
        self.lib.ref_fft2d(_ptr(a, np.float32), w, h, int(inverse))
        return a[..., 0] + 1j * a[..., 1]

    def first_minimum_radius(self, raw):
        raw = np.ascontiguousarray(raw, np.float32)
        return float(self.lib.ref_first_minimum_radius(_ptr(raw, np.float32), raw.shape[0]))

    def resample_kernel(self, raw, scale, support):
        raw = np.ascontiguousarray(raw, np.float32)
        k = 2 * support + 1
        out = np.zeros((k, k), np.float32)
        self.lib.ref_resample_kernel(_ptr(raw, np.float32), raw.shape[0], scale, support,
                                     _ptr(out, np.float32))
        return out

    def shape_kernels(self, p, base):
        base = np.ascontiguousarray(base, np.float32)
        k = base.shape[0]
        outs = [np.zeros(k * k, np.float32) for _ in range(3)]
        size = C.c_int(0)
        r = self.lib.ref_shape_kernels(C.byref(p), _ptr(base, np.float32), k,
                                       *[_ptr(o, np.float32) for o in outs], C.byref(size))
        n = size.value
        full, core, struct = (o[: n * n].reshape(n, n).copy() for o in outs)
        return full, core, struct, r

    def bank(self, p):
        L = self.lib
        h = L.ref_bank_build(C.byref(p))
        try:
            n = L.ref_bank_bin_count(h)
            out = {"wavelengths": [], "support": L.ref_bank_support_radius(h), "full": [],
                   "core": [], "structure": []}
            for b in range(n):
                k = L.ref_bank_kernel_size(h, b)
                out["wavelengths"].append(L.ref_bank_wavelength(h, b))
                for which, name in ((0, "full"), (1, "core"), (2, "structure")):
                    a = np.zeros((k, k), np.float32)
                    L.ref_bank_kernel(h, b, which, _ptr(a, np.float32))
                    out[name].append(a)
            return out
        finally:
            L.ref_bank_free(h)

    def spectrum_config(self, p, wavelengths):
        wl = np.ascontiguousarray(wavelengths, np.float32)
        n = wl.shape[0]
        nat = np.zeros((3, n), np.float32)
        sty = np.zeros((3, n), np.float32)
        self.lib.ref_spectrum_config(C.byref(p), n, _ptr(wl, np.float32), _ptr(nat, np.float32),
                                     _ptr(sty, np.float32))
        return nat, sty

    def map_bins(self, p, wavelengths, bins):
        wl = np.ascontiguousarray(wavelengths, np.float32)
        bins = np.ascontiguousarray(bins, np.float32)
        npx = int(np.prod(bins.shape[1:]))
        rgb = np.zeros((npx, 3), np.float32)
        self.lib.ref_map_bins(C.byref(p), wl.shape[0], _ptr(wl, np.float32), _ptr(bins, np.float32),
                              npx, _ptr(rgb, np.float32))
        return rgb.reshape(bins.shape[1:] + (3,))

    def selection(self, p, rgba):
        rgba = np.ascontiguousarray(rgba, np.float32)
        h, w = rgba.shape[:2]
        mask = np.zeros((h, w), np.float32)
        red = np.zeros((h, w, 4), np.float32)
        drv = np.zeros((h, w), np.float32)
        self.lib.ref_selection(C.byref(p), _ptr(rgba, np.float32), w, h, _ptr(mask, np.float32),
                               _ptr(red, np.float32), _ptr(drv, np.float32))
        return mask, red, drv

    def convolve(self, plane, kernel):
        plane = np.ascontiguousarray(plane, np.float32)
        kernel = np.ascontiguousarray(kernel, np.float32)
        h, w = plane.shape
        out = np.zeros((h, w), np.float32)
        self.lib.ref_convolve(_ptr(plane, np.float32), w, h, _ptr(kernel, np.float32),
                              kernel.shape[0], _ptr(out, np.float32))
        return out

    def render_image(self, p, rgba):
        rgba = np.ascontiguousarray(rgba, np.float32)
        h, w = rgba.shape[:2]
        out = np.zeros((h, w, 4), np.float32)
        if not self.lib.ref_render_image(C.byref(p), _ptr(rgba, np.float32), w, h,
                                         _ptr(out, np.float32)):
            raise RuntimeError("reference render failed")
        return out

    def render(self, p, req) -> tuple[bool, str]:
        err = C.create_string_buffer(64)
        ok = self.lib.ref_render(C.byref(p), C.byref(req), err, 64)
        return bool(ok), err.value.decode()


def reference_available() -> bool:
    return os.path.exists(REF_SO)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return float(num / den) if den > 0 else float(num)

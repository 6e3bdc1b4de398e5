"""Host-side mirror of the plugin's render seam for the HIP backend.

``run_lensdiff_hip(request, params, cache)`` follows the signature and error behaviour of
``RunLensDiffCpuReference`` / ``RunLensDiffCuda`` (source/src/core/LensDiffCpuReference.h:64-67,
source/src/cuda/LensDiffCuda.h:7-10): it returns ``(ok, error_tag)`` and never raises for
render failures. ``Context`` wraps an ``ldh_context`` and exposes the per-stage entry points
used by the parity tests.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _abi
from ._lib import LensDiffError, load

_F = C.POINTER(C.c_float)


def _fptr(a: np.ndarray):
    if a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float32 array")
    return a.ctypes.data_as(_F)


class Context:
    """One HIP device context (stream, FFT tables, bank cache)."""

    def __init__(self, device: int = 0, memory_budget: int = 0):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.ldh_context_create(device, memory_budget, C.byref(h))
        if st != 0:
            raise LensDiffError(st, self.lib.ldh_status_string(st).decode())
        self.h = h

    def close(self):
        if self.h:
            self.lib.ldh_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int):
        if st != 0:
            raise LensDiffError(st, self.lib.ldh_last_error(self.h).decode())

    # ---- sizes --------------------------------------------------------------------------
    def sizes(self, p: _abi.Params) -> _abi.Sizes:
        s = _abi.Sizes()
        self._check(self.lib.ldh_resolve_sizes(C.byref(p), C.byref(s)))
        return s

    # ---- PSF stages -----------------------------------------------------------------------
    def pupil_amplitude(self, p) -> np.ndarray:
        n = self.sizes(p).pupil_size
        out = np.zeros((n, n), np.float32)
        self._check(self.lib.ldh_pupil_amplitude(self.h, C.byref(p), _fptr(out)))
        return out

    def pupil_phase(self, p) -> np.ndarray:
        n = self.sizes(p).pupil_size
        out = np.zeros((n, n), np.float32)
        self._check(self.lib.ldh_pupil_phase(self.h, C.byref(p), _fptr(out)))
        return out

    def raw_psf(self, amp, phase, raw_size) -> np.ndarray:
        amp = np.ascontiguousarray(amp, np.float32)
        ph = None if phase is None else np.ascontiguousarray(phase, np.float32)
        out = np.zeros((raw_size, raw_size), np.float32)
        self._check(self.lib.ldh_raw_psf(self.h, _fptr(amp), None if ph is None else _fptr(ph),
                                         amp.shape[0], raw_size, _fptr(out)))
        return out

    def raw_psf_window(self, amp, phase, raw_size, half_width) -> np.ndarray:
        """Centred (2h+1)^2 block of the raw PSF synthesised alone (unnormalised)."""
        amp = np.ascontiguousarray(amp, np.float32)
        p = amp.shape[0]
        n = 2 * half_width + 1
        out = np.zeros((n, n), np.float32)
        ph = None if phase is None else np.ascontiguousarray(phase, np.float32)
        self._check(self.lib.ldh_raw_psf_window(self.h, _fptr(amp), None if ph is None else _fptr(ph), p, raw_size,
                                                half_width, _fptr(out)))
        return out

    def first_minimum_radius(self, psf) -> float:
        psf = np.ascontiguousarray(psf, np.float32)
        r = C.c_float()
        self._check(self.lib.ldh_first_minimum_radius(self.h, _fptr(psf), psf.shape[0], C.byref(r)))
        return float(r.value)

    def resample_kernel(self, psf, scale, support) -> np.ndarray:
        psf = np.ascontiguousarray(psf, np.float32)
        k = 2 * support + 1
        out = np.zeros((k, k), np.float32)
        self._check(self.lib.ldh_resample_kernel(self.h, _fptr(psf), psf.shape[0], scale, support, _fptr(out)))
        return out

    def shape_kernels(self, p, base):
        base = np.ascontiguousarray(base, np.float32)
        k = base.shape[0]
        outs = [np.zeros(k * k, np.float32) for _ in range(3)]
        n = C.c_int32()
        self._check(self.lib.ldh_shape_kernels(self.h, C.byref(p), _fptr(base), k, *[_fptr(o) for o in outs],
                                               C.byref(n)))
        m = n.value
        return tuple(o[: m * m].reshape(m, m).copy() for o in outs)

    def sequential_sum(self, x: np.ndarray) -> float:
        x = np.ascontiguousarray(x, np.float32).ravel()
        out = C.c_float()
        self._check(self.lib.ldh_sequential_sum(self.h, _fptr(x), x.size, C.byref(out)))
        return float(out.value)

    def sequential_sums(self, x: np.ndarray, block: int = 0) -> np.ndarray:
        """Sequential FP32 sums of the rows of a 2-D array, over device blocks of `block`."""
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(x.shape[0], np.float32)
        self._check(self.lib.ldh_sequential_sums(self.h, _fptr(x), x.shape[1], x.shape[0], block, _fptr(out)))
        return out

    def fft2d(self, z: np.ndarray, inverse=False) -> np.ndarray:
        h, w = z.shape
        a = np.ascontiguousarray(np.stack([z.real, z.imag], -1), np.float32)
        self._check(self.lib.ldh_fft2d(self.h, _fptr(a), w, h, int(inverse)))
        return a[..., 0] + 1j * a[..., 1]

    # ---- banks ------------------------------------------------------------------------------
    def bank(self, p) -> dict:
        """Bank for `p` as host arrays: wavelengths, support, full/core/structure kernels."""
        b = C.c_void_p()
        self._check(self.lib.ldh_bank_acquire(self.h, C.byref(p), C.byref(b)))
        try:
            return self.bank_arrays(b)
        finally:
            self.lib.ldh_bank_release(self.h, b)

    def build_banks(self, params: list, mode: int):
        """ldh_bank_build_batch: (bank handles, BatchStats). Release each with bank_release."""
        n = len(params)
        arr = (_abi.Params * n)(*params)
        handles = (C.c_void_p * n)()
        st = _abi.BatchStats()
        self._check(self.lib.ldh_bank_build_batch(self.h, arr, n, mode, handles, C.byref(st)))
        return [C.c_void_p(h) for h in handles], st

    def bank_release(self, b):
        self._check(self.lib.ldh_bank_release(self.h, b))

    def bank_arrays(self, b) -> dict:
        """Host copy of a bank handle's kernels."""
        info = _abi.BankInfo()
        self._check(self.lib.ldh_bank_info_get(self.h, b, C.byref(info)))
        out = {"wavelengths": [], "support": info.support_radius_px, "raw_size": info.raw_psf_size,
               "raw_window": info.raw_window_px,
               "full": [], "core": [], "structure": []}
        for i in range(info.bin_count):
            k = info.kernel_size[i]
            out["wavelengths"].append(info.wavelengths_nm[i])
            for which, name in ((0, "full"), (1, "core"), (2, "structure")):
                a = np.zeros((k, k), np.float32)
                self._check(self.lib.ldh_bank_download(self.h, b, i, which, _fptr(a), a.size))
                out[name].append(a)
        return out

    def bank_kernel_sizes(self, p) -> list:
        """Cropped kernel side per bin of the bank for `p`."""
        b = C.c_void_p()
        self._check(self.lib.ldh_bank_acquire(self.h, C.byref(p), C.byref(b)))
        try:
            info = _abi.BankInfo()
            self._check(self.lib.ldh_bank_info_get(self.h, b, C.byref(info)))
            return [int(info.kernel_size[i]) for i in range(info.bin_count)]
        finally:
            self.lib.ldh_bank_release(self.h, b)

    def bank_build_seconds(self, p) -> float:
        """Wall time of building the bank for `p` on the device (synchronised), bank released."""
        b = C.c_void_p()
        self.synchronize()
        t0 = time.perf_counter()
        self._check(self.lib.ldh_bank_acquire(self.h, C.byref(p), C.byref(b)))
        self.synchronize()
        dt = time.perf_counter() - t0
        self.lib.ldh_bank_release(self.h, b)
        return dt

    # ---- render ------------------------------------------------------------------------------
    def render_image(self, p, rgba: np.ndarray) -> np.ndarray:
        """Whole-frame render of a packed RGBA float32 image through the host-view entry."""
        src = np.ascontiguousarray(rgba, np.float32)
        h, w = src.shape[:2]
        dst = np.zeros_like(src)
        req = make_request(src, dst)
        self._check(self.lib.ldh_render(self.h, C.byref(req), C.byref(p)))
        return dst

    def render(self, req: _abi.RenderRequest, p) -> None:
        self._check(self.lib.ldh_render(self.h, C.byref(req), C.byref(p)))

    # ---- device buffers (bench) -----------------------------------------------------------
    def alloc(self, nbytes: int) -> int:
        ptr = C.c_void_p()
        self._check(self.lib.ldh_device_alloc(self.h, nbytes, C.byref(ptr)))
        return ptr.value

    def free(self, ptr: int):
        self._check(self.lib.ldh_device_free(self.h, C.c_void_p(ptr)))

    def upload(self, ptr: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self._check(self.lib.ldh_memcpy_h2d(self.h, C.c_void_p(ptr), a.ctypes.data_as(C.c_void_p), a.nbytes))

    def download(self, ptr: int, a: np.ndarray):
        self._check(self.lib.ldh_memcpy_d2h(self.h, a.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), a.nbytes))

    def render_device(self, p, d_src: int, d_dst: int, w: int, h: int):
        self._check(self.lib.ldh_render_device(self.h, C.byref(p), C.c_void_p(d_src), C.c_void_p(d_dst), w, h))

    # ---- served form --------------------------------------------------------------------------
    def server_start(self, slots: int = 4, budget: int = 64 << 30):
        self._check(self.lib.ldh_server_start(self.h, slots, budget))

    def server_stop(self):
        self._check(self.lib.ldh_server_stop(self.h))

    def submit(self, p, d_src: int, d_dst: int, w: int, h: int) -> int:
        t = C.c_int64()
        st = self.lib.ldh_submit(self.h, C.byref(p), C.c_void_p(d_src), C.c_void_p(d_dst), w, h, C.byref(t))
        if st != 0:
            raise LensDiffError(st, self.lib.ldh_status_string(st).decode())
        return t.value

    def wait(self, ticket: int) -> _abi.RequestStats:
        s = _abi.RequestStats()
        st = self.lib.ldh_wait(self.h, ticket, C.byref(s))
        if st != 0:
            raise LensDiffError(st, s.error.decode() or self.lib.ldh_status_string(st).decode())
        return s

    def poll(self, ticket: int):
        s = _abi.RequestStats()
        st = self.lib.ldh_poll(self.h, ticket, C.byref(s))
        if st == _abi.LDH_ERR_NOT_READY:
            return None
        if st != 0:
            raise LensDiffError(st, s.error.decode() or self.lib.ldh_status_string(st).decode())
        return s

    def server_stats(self) -> _abi.ServerStats:
        s = _abi.ServerStats()
        self._check(self.lib.ldh_server_stats_get(self.h, C.byref(s)))
        return s

    def synchronize(self):
        self._check(self.lib.ldh_synchronize(self.h))

    def profiling(self, enable: bool):
        self._check(self.lib.ldh_profiling_enable(self.h, int(enable)))

    def profile_read(self) -> dict:
        stats = (_abi.KernelStat * 128)()
        n = self.lib.ldh_profiling_read(self.h, stats, 128)
        return {stats[i].name.decode(): (stats[i].launches, stats[i].total_ms, stats[i].algorithmic_bytes)
                for i in range(n)}

    def profile_reset(self):
        self._check(self.lib.ldh_profiling_reset(self.h))

    def tuning_force(self, variant: int):
        """Pins every autotuned kernel family to `variant` (-1: tune again on first use)."""
        self.lib.ldh_tuning_force(int(variant))

    def tuning_report(self) -> dict:
        """The autotuner's choices: {key: {"variant": k, "ms": [per-variant times]}}."""
        n = self.lib.ldh_tuning_report(None, 0)
        buf = C.create_string_buffer(n + 1)
        self.lib.ldh_tuning_report(buf, n + 1)
        out = {}
        for line in buf.value.decode().splitlines():
            key, rest = line.split(": ", 1)
            parts = rest.split()
            out[key] = {"variant": int(parts[0][1:]), "ms": [float(x) for x in parts[1:-1]]}
        return out


def make_request(src: np.ndarray, dst: np.ndarray, render_window=None, src_origin=(0, 0),
                 dst_origin=(0, 0)) -> _abi.RenderRequest:
    """A render request over packed RGBA arrays (row bytes = width * 16)."""
    h, w = src.shape[:2]
    dh, dw = dst.shape[:2]
    req = _abi.RenderRequest()
    sx, sy = src_origin
    dx, dy = dst_origin
    req.src.data = src.ctypes.data
    req.src.row_bytes = src.strides[0]
    req.src.bounds = _abi.Rect(sx, sy, sx + w, sy + h)
    req.src.components = src.shape[2]
    req.src.bytes_per_component = 4
    req.dst.data = dst.ctypes.data
    req.dst.row_bytes = dst.strides[0]
    req.dst.bounds = _abi.Rect(dx, dy, dx + dw, dy + dh)
    req.dst.components = dst.shape[2]
    req.dst.bytes_per_component = 4
    rw = render_window or (dx, dy, dx + dw, dy + dh)
    req.render_window = _abi.Rect(*rw)
    req.frame_bounds = _abi.Rect(sx, sy, sx + w, sy + h)
    req.frame_short_side_px = min(w, h)
    return req


def run_lensdiff_hip(ctx: Context, request: _abi.RenderRequest, params: _abi.Params) -> tuple[bool, str]:
    """The seam: render `request` with `params`; returns (ok, kebab-case error tag)."""
    st = ctx.lib.ldh_render(ctx.h, C.byref(request), C.byref(params))
    if st != 0:
        return False, ctx.lib.ldh_last_error(ctx.h).decode()
    return True, ""

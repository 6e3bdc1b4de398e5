"""The C-ABI library: loads without a GPU, exports every function include/lensdiff_hip.h declares,
and the ctypes mirrors have the C layout (checked against gcc's sizeof/offsetof)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from lensdiff_amd import _abi, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lensdiff_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ldh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("ldh_context_create", "ldh_render", "ldh_render_device", "ldh_bank_acquire",
                     "ldh_bank_download", "ldh_raw_psf", "ldh_status_string", "ldh_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_status_strings_and_defaults_without_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("HIP library not built")
    lib = _lib.load()
    assert lib.ldh_status_string(2) == b"missing-source-or-destination-buffer"
    assert lib.ldh_status_string(3) == b"invalid-source-bounds"
    p = _abi.Params()
    lib.ldh_params_default(ctypes.byref(p))
    q = _abi.default_params()
    assert bytes(p) == bytes(q)
    s = _abi.Sizes()
    assert lib.ldh_resolve_sizes(ctypes.byref(q), ctypes.byref(s)) == 0
    # LensDiffParams defaults: P=256, r=ceil(64% of 2160)=1383, R=clamp(max(512,512,5532))=5532
    assert (s.pupil_size, s.max_kernel_radius_px, s.raw_psf_size) == (256, 1383, 5532)


C_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "lensdiff_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(ldh_params), sizeof(ldh_render_request), sizeof(ldh_image_view),
         sizeof(ldh_sizes), sizeof(ldh_bank_info), sizeof(ldh_kernel_stat));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(ldh_params, threshold), offsetof(ldh_params, max_kernel_radius_px),
         offsetof(ldh_params, phase_defocus), offsetof(ldh_params, wavelength_count),
         offsetof(ldh_params, zernike_extra), offsetof(ldh_params, spectra_fp16));
  printf("%zu %zu\n", offsetof(ldh_render_request, src), offsetof(ldh_render_request, frame_short_side_px));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(ldh_batch_stats), offsetof(ldh_batch_stats, groups),
         sizeof(ldh_request_stats), sizeof(ldh_server_stats), offsetof(ldh_batch_stats, mode),
         offsetof(ldh_batch_stats, capture_ms));
  return 0;
}
"""


def test_ctypes_layout_matches_c():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(src, "w").write(C_PROBE)
        subprocess.check_call(["gcc", "-I" + os.path.join(ROOT, "include"), "-o", exe, src])
        lines = subprocess.check_output([exe]).decode().split("\n")
    sizes = [int(v) for v in lines[0].split()]
    assert sizes == [ctypes.sizeof(t) for t in (_abi.Params, _abi.RenderRequest, _abi.ImageView, _abi.Sizes,
                                                 _abi.BankInfo, _abi.KernelStat)]
    offs = [int(v) for v in lines[1].split()]
    P = _abi.Params
    assert offs == [P.threshold.offset, P.max_kernel_radius_px.offset, P.phase_defocus.offset,
                    P.wavelength_count.offset, P.zernike_extra.offset, P.spectra_fp16.offset]
    offs = [int(v) for v in lines[2].split()]
    assert offs == [_abi.RenderRequest.src.offset, _abi.RenderRequest.frame_short_side_px.offset]
    B = _abi.BatchStats
    assert [int(v) for v in lines[3].split()] == [ctypes.sizeof(B), B.groups.offset, ctypes.sizeof(_abi.RequestStats),
                                                  ctypes.sizeof(_abi.ServerStats), B.mode.offset, B.capture_ms.offset]


def test_context_create_fails_cleanly_without_device():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("HIP library not built")
    lib = _lib.load()
    h = ctypes.c_void_p()
    st = lib.ldh_context_create(99, 0, ctypes.byref(h))
    assert st == _abi.LDH_ERR_DEVICE and not h.value

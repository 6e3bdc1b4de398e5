"""ctypes mirror of the C ABI declared in include/lensdiff_hip.h.

The field order and types must match the C structs exactly; tests check sizeof against the
library (``ldh_abi_sizes``).
"""
from __future__ import annotations

import ctypes as C

LDH_MAX_WAVELENGTHS = 64

# status codes
LDH_OK = 0
LDH_ERR_INVALID_ARGUMENT = 1
LDH_ERR_MISSING_BUFFER = 2
LDH_ERR_INVALID_BOUNDS = 3
LDH_ERR_UNSUPPORTED = 4
LDH_ERR_OUT_OF_MEMORY = 5
LDH_ERR_DEVICE = 6
LDH_ERR_NOT_READY = 7
LDH_ERR_TOO_LARGE = 8

# enums (LensDiffTypes.h:14-91)
EXTRACT_MAX_RGB, EXTRACT_LUMA = 0, 1
(APERTURE_CIRCLE, APERTURE_POLYGON, APERTURE_STAR, APERTURE_SPIRAL, APERTURE_CUSTOM,
 APERTURE_HEXAGON, APERTURE_SQUARE_GRID, APERTURE_SNOWFLAKE) = range(8)
APOD_FLAT, APOD_COSINE, APOD_GAUSSIAN = range(3)
SPECTRAL_MONO, SPECTRAL_TRISTIMULUS, SPECTRAL_5, SPECTRAL_9 = range(4)
STYLE_NATURAL, STYLE_CYAN_MAGENTA, STYLE_WARM_COOL = range(3)
LOOK_PHYSICAL, LOOK_SPLIT = 0, 1
ENERGY_PRESERVE, ENERGY_AUGMENT = 0, 1
(VIEW_FINAL, VIEW_SELECTION, VIEW_PUPIL, VIEW_PSF, VIEW_OTF, VIEW_CORE, VIEW_STRUCTURE,
 VIEW_EFFECT, VIEW_PHASE, VIEW_PHASE_EDGE, VIEW_FIELD_PSF, VIEW_CHROMATIC_SPLIT,
 VIEW_CREATIVE_FRINGE, VIEW_SCATTER) = range(14)
TRANSFER_LINEAR, TRANSFER_DAVINCI_INTERMEDIATE = 0, 1
PHASE_OFF, PHASE_ENABLED = 0, 1
KERNEL_FULL, KERNEL_CORE, KERNEL_STRUCTURE = range(3)
CONV_PER_BIN, CONV_PRECOMBINED = 0, 1
RAW_PSF_AUTO, RAW_PSF_WHOLE = 0, 1


class Rect(C.Structure):
    _fields_ = [("x1", C.c_int32), ("y1", C.c_int32), ("x2", C.c_int32), ("y2", C.c_int32)]


class ImageView(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("row_bytes", C.c_int64),
        ("bounds", Rect),
        ("origin_x", C.c_int32),
        ("origin_y", C.c_int32),
        ("components", C.c_int32),
        ("bytes_per_component", C.c_int32),
    ]


class RenderRequest(C.Structure):
    _fields_ = [
        ("render_window", Rect),
        ("frame_bounds", Rect),
        ("src", ImageView),
        ("dst", ImageView),
        ("frame_short_side_px", C.c_int32),
    ]


_I = C.c_int32
_D = C.c_double

_PARAM_FIELDS = [
    ("extraction_mode", _I), ("aperture_mode", _I), ("apodization_mode", _I), ("spectral_mode", _I),
    ("spectrum_style", _I), ("look_mode", _I), ("energy_mode", _I), ("debug_view", _I),
    ("input_transfer", _I), ("resolution_aware", _I),
    ("threshold", _D), ("softness_stops", _D), ("point_emphasis", _D), ("core_preserve", _D),
    ("blade_count", _I), ("vane_count", _I), ("pupil_resolution", _I), ("frame_short_side_px", _I),
    ("roundness", _D), ("rotation_deg", _D), ("central_obstruction", _D), ("vane_thickness", _D),
    ("diffraction_scale_px", _D), ("max_kernel_radius_px", _D),
    ("phase_mode", _I), ("chromatic_affects_luma", _I),
    ("phase_defocus", _D), ("phase_astigmatism0", _D), ("phase_astigmatism45", _D),
    ("phase_coma_x", _D), ("phase_coma_y", _D),
    ("phase_spherical", _D), ("phase_trefoil_x", _D), ("phase_trefoil_y", _D),
    ("phase_secondary_astigmatism0", _D), ("phase_secondary_astigmatism45", _D),
    ("phase_quadrafoil0", _D), ("phase_quadrafoil45", _D), ("phase_secondary_coma_x", _D),
    ("phase_secondary_coma_y", _D),
    ("pupil_decenter_x", _D), ("pupil_decenter_y", _D),
    ("phase_field_strength", _D), ("phase_field_edge_bias", _D), ("phase_field_defocus", _D),
    ("phase_field_astig_radial", _D), ("phase_field_astig_tangential", _D),
    ("phase_field_coma_radial", _D), ("phase_field_coma_tangential", _D), ("phase_field_spherical", _D),
    ("phase_field_trefoil_radial", _D), ("phase_field_trefoil_tangential", _D),
    ("phase_field_secondary_astig_radial", _D), ("phase_field_secondary_astig_tangential", _D),
    ("phase_field_quadrafoil_radial", _D), ("phase_field_quadrafoil_tangential", _D),
    ("phase_field_secondary_coma_radial", _D), ("phase_field_secondary_coma_tangential", _D),
    ("spectrum_force", _D), ("spectrum_saturation", _D),
    ("chromatic_focus", _D), ("chromatic_spread", _D), ("creative_fringe", _D),
    ("scatter_amount", _D), ("scatter_radius", _D),
    ("effect_gain", _D), ("core_compensation", _D), ("anisotropy_emphasis", _D),
    ("core_gain", _D), ("structure_gain", _D), ("core_shoulder", _D), ("structure_shoulder", _D),
    ("wavelength_count", _I), ("conv_mode", _I),
    ("wavelengths_nm", C.c_float * LDH_MAX_WAVELENGTHS),
    ("zernike_extra", _D * 6),
    ("raw_psf_form", _I), ("spectra_fp16", _I),
]


class Params(C.Structure):
    """ldh_params: LensDiffParams (LensDiffTypes.h:128-208) plus this backend's extensions."""

    _fields_ = _PARAM_FIELDS

    def copy(self) -> "Params":
        out = Params()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(Params))
        return out

    def set_wavelengths(self, wl) -> None:
        wl = list(wl)
        if len(wl) > LDH_MAX_WAVELENGTHS:
            raise ValueError("at most %d wavelengths" % LDH_MAX_WAVELENGTHS)
        self.wavelength_count = len(wl)
        for i, v in enumerate(wl):
            self.wavelengths_nm[i] = float(v)

    def set_noll(self, coeffs: dict) -> None:
        """Set Zernike coefficients (waves) by Noll index 2..21.

        Noll 4..17 are the reference's named terms (LensDiffPhase.cpp:182-197); 2, 3, 18..21
        go to the extension slots. Enables the phase.
        """
        named = {4: "phase_defocus", 5: "phase_astigmatism45", 6: "phase_astigmatism0",
                 7: "phase_coma_y", 8: "phase_coma_x", 9: "phase_trefoil_y", 10: "phase_trefoil_x",
                 11: "phase_spherical", 12: "phase_secondary_astigmatism0",
                 13: "phase_secondary_astigmatism45", 14: "phase_quadrafoil0",
                 15: "phase_quadrafoil45", 16: "phase_secondary_coma_x",
                 17: "phase_secondary_coma_y"}
        extra = {2: 0, 3: 1, 18: 2, 19: 3, 20: 4, 21: 5}
        for j, c in coeffs.items():
            if j in named:
                setattr(self, named[j], float(c))
            elif j in extra:
                self.zernike_extra[extra[j]] = float(c)
            else:
                raise ValueError("Noll index %d outside 2..21" % j)
        self.phase_mode = PHASE_ENABLED


class Sizes(C.Structure):
    _fields_ = [
        ("pupil_size", C.c_int32),
        ("raw_psf_size", C.c_int32),
        ("max_kernel_radius_px", C.c_int32),
        ("optics_short_side_px", C.c_int32),
        ("diffraction_scale_px", C.c_double),
    ]


class BankInfo(C.Structure):
    _fields_ = [
        ("bin_count", C.c_int32),
        ("support_radius_px", C.c_int32),
        ("pupil_size", C.c_int32),
        ("raw_psf_size", C.c_int32),
        ("wavelengths_nm", C.c_float * LDH_MAX_WAVELENGTHS),
        ("kernel_size", C.c_int32 * LDH_MAX_WAVELENGTHS),
        ("raw_window_px", C.c_int32),
    ]


LDH_BATCH_PER_REQUEST, LDH_BATCH_GROUPED, LDH_BATCH_GRAPH = 0, 1, 2


class BatchStats(C.Structure):
    _fields_ = [
        ("exec_ms", C.c_double),
        ("capture_ms", C.c_double),
        ("groups", C.c_int32),
        ("psfs", C.c_int32),
        ("graph_nodes", C.c_int32),
        ("mode", C.c_int32),
        ("branches", C.c_int32),
    ]


class RequestStats(C.Structure):
    _fields_ = [
        ("latency_ms", C.c_double),
        ("queue_ms", C.c_double),
        ("service_ms", C.c_double),
        ("bank_hit", C.c_int32),
        ("spectra_hit", C.c_int32),
        ("status", C.c_int32),
        ("error", C.c_char * 48),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("bank_hits", C.c_int64),
        ("bank_misses", C.c_int64), ("spectra_hits", C.c_int64), ("spectra_misses", C.c_int64),
        ("evictions", C.c_int64), ("resident_bytes", C.c_uint64), ("budget_bytes", C.c_uint64),
        ("slots", C.c_int32), ("cached_banks", C.c_int32),
    ]


class KernelStat(C.Structure):
    _fields_ = [
        ("name", C.c_char * 48),
        ("launches", C.c_int64),
        ("total_ms", C.c_double),
        ("algorithmic_bytes", C.c_double),
    ]


def default_params() -> Params:
    """LensDiffParams' member defaults (LensDiffTypes.h:128-208)."""
    p = Params()
    p.extraction_mode = EXTRACT_MAX_RGB
    p.aperture_mode = APERTURE_CIRCLE
    p.apodization_mode = APOD_FLAT
    p.spectral_mode = SPECTRAL_MONO
    p.spectrum_style = STYLE_NATURAL
    p.look_mode = LOOK_SPLIT
    p.energy_mode = ENERGY_PRESERVE
    p.debug_view = VIEW_FINAL
    p.input_transfer = TRANSFER_LINEAR
    p.threshold = 4.0
    p.softness_stops = 1.0
    p.core_preserve = 0.1
    p.blade_count = 6
    p.roundness = 0.5
    p.vane_thickness = 0.01
    p.diffraction_scale_px = 24.0
    p.pupil_resolution = 256
    p.max_kernel_radius_px = 64.0
    p.frame_short_side_px = 2160
    p.phase_mode = PHASE_OFF
    p.spectrum_saturation = 1.0
    p.effect_gain = 1.0
    p.core_compensation = 1.0
    p.core_gain = 1.0
    p.structure_gain = 1.0
    return p

"""Comparison of cropped bank kernels.

The adaptive support radius (estimateAdaptiveSupportRadius, LensDiffCpuReference.cpp:984-1036) is
a hard threshold on the captured-energy fraction and the outside peak at each ring. The device forms
the ring sums exactly as the reference does (float, raster order), so kernels shaped from the same
input have the same radius and are compared side for side (`kernel_rel_l2`). Kernels shaped from
inputs that differ in their last bits (a device PSF against the oracle's) can still cross a
threshold differently when a bin sits on it: those comparisons centre both in the larger square and
allow the radius to move by the step such a flip makes (`kernel_rel_l2_ring`); the callers count
and report such flips."""
import numpy as np

from oracle.oracle import rel_l2


def centred(a, n):
    k = a.shape[0]
    if k == n:
        return a
    out = np.zeros((n, n), a.dtype)
    o = (n - k) // 2
    out[o:o + k, o:o + k] = a
    return out


def kernel_rel_l2(got, want):
    """rel L2 of two odd square kernels of the same side (asserted)."""
    assert got.shape == want.shape, (got.shape, want.shape)
    return rel_l2(got, want)


def kernel_rel_l2_ring(got, want):
    """rel L2 of two odd square kernels centred in the larger one; asserts the radii differ by at
    most two: the estimate by one ring, and with it possibly its padding (max(6, min(24,
    ceil(5% of the estimate))), LensDiffCpuReference.cpp:1031-1034) by one more."""
    assert got.ndim == 2 and want.ndim == 2 and got.shape[0] == got.shape[1] and want.shape[0] == want.shape[1]
    assert abs(got.shape[0] - want.shape[0]) <= 4, (got.shape, want.shape)
    n = max(got.shape[0], want.shape[0])
    return rel_l2(centred(got, n), centred(want, n))

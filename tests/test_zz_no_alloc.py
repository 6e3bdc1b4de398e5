"""The allocation-free entry point ldh_sequential_sums_noalloc against ldh_sequential_sums (the
existing entry point on host buffers): the same sequential FP32 sums bit for bit, on two calls in
a row with the second smaller, plus the size limit of its static scratch."""
import ctypes as C

import numpy as np
import pytest

from lensdiff_amd import _abi

pytestmark = pytest.mark.gpu

LDH_ERR_TOO_LARGE = 8


def arrays(count, n, seed):
    rng = np.random.default_rng(seed)
    rows = [rng.random(n).astype(np.float32),
            (rng.exponential(1e-6, n)).astype(np.float32),
            np.ones(n, np.float32),  # every addend past 2^24 is a rounding tie
            np.ldexp(rng.random(n), -rng.integers(0, 50, n)).astype(np.float32)]
    return np.ascontiguousarray(np.stack([rows[i % len(rows)] for i in range(count)]))


def noalloc_sums(ctx, x):
    count, n = x.shape
    d_x, d_t = ctx.alloc(x.nbytes), ctx.alloc(4 * count)
    try:
        ctx.upload(d_x, x)
        st = ctx.lib.ldh_sequential_sums_noalloc(ctx.h, C.c_void_p(d_x), n, count, C.c_void_p(d_t), None)
        assert st == 0, ctx.lib.ldh_last_error(ctx.h)
        ctx.synchronize()
        out = np.empty(count, np.float32)
        ctx.download(d_t, out)
        return out
    finally:
        ctx.free(d_x)
        ctx.free(d_t)


def test_noalloc_sums_match_the_allocating_entry_point(ctx):
    # the second call is smaller: the scratch keeps the first call's tables, which must not leak in
    for count, n, seed in ((4, 5_000_011, 1), (3, 100_003, 2)):
        x = arrays(count, n, seed)
        want = ctx.sequential_sums(x)
        got = noalloc_sums(ctx, x)
        assert np.array_equal(got, want), (count, n, got, want)


def test_noalloc_rejects_sizes_beyond_its_scratch(ctx):
    d = ctx.alloc(4096)
    try:
        # 65537 blocks of 4096: rejected before anything is read or launched
        st = ctx.lib.ldh_sequential_sums_noalloc(ctx.h, C.c_void_p(d), 4096 * 65537, 1, C.c_void_p(d), None)
        assert st == LDH_ERR_TOO_LARGE
        assert ctx.lib.ldh_last_error(ctx.h).decode() == "noalloc-scratch-too-small"
    finally:
        ctx.free(d)

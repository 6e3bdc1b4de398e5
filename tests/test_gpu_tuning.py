"""The autotuned schedule: the column split of a 2-D FFT height is timed once per process and
reported; transforms on such heights match numpy whatever it picks."""
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def test_column_split_is_tuned_and_reported(ctx):
    # 384 rows (3 * 2^7) need two column passes; every split of 384 into tile lengths is a candidate
    rng = np.random.default_rng(7)
    z = (rng.standard_normal((384, 64)) + 1j * rng.standard_normal((384, 64))).astype(np.complex64)
    got = ctx.fft2d(z)
    want = np.fft.fft2(z.astype(np.complex128))
    assert crel(got, want) < 2e-6
    assert crel(ctx.fft2d(got, inverse=True), z) < 2e-6
    rep = ctx.tuning_report()
    keys = [k for k in rep if k.startswith("fft-cols/384[")]
    assert len(keys) == 1, rep
    cands = re.search(r"\[(.*)\]", keys[0]).group(1).split(",")
    assert len(cands) >= 2
    for c in cands:
        a, b = (int(x) for x in c.split("x"))
        assert a * b == 384
    entry = rep[keys[0]]
    assert 0 <= entry["variant"] < len(cands)
    # timed on first use (no graph capture here): one time per candidate
    assert len(entry["ms"]) == len(cands) and all(t > 0 for t in entry["ms"])
    # fixed for the process: a second transform of the same height keeps the choice
    ctx.fft2d(z)
    assert ctx.tuning_report()[keys[0]]["variant"] == entry["variant"]

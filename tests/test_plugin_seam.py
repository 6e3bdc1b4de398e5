"""The plugin-side seam RunLensDiffHip, compiled as the LensDiff plugin would compile it.

tests/seam/Makefile builds lensdiff_amd/plugin/LensDiffHip.cpp against the plugin's headers and
core translation units (LensDiffTypes.h, LensDiffCpuReference.h/.cpp) where they lie, links it
to the built liblensdiff_hip.so and runs tests/seam/seam_check.cpp. Needs the plugin sources,
so it runs only where they are present; the program holds compiled plugin code and is never
built anywhere else.

The tags follow the plugin's backends: RunLensDiffCuda's early exits
(LensDiffCuda.cu:4642-4657) and the error-string convention of LensDiff.cpp:3830-3838.
"""
import fcntl
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEAM = os.path.join(ROOT, "tests", "seam")
PLUGIN_CORE = "/root/reference/source/src/core/LensDiffCpuReference.cpp"
LIB = os.path.join(ROOT, "lensdiff_amd", "liblensdiff_hip.so")

pytestmark = pytest.mark.skipif(
    not (os.path.exists(PLUGIN_CORE) and os.path.exists(LIB)),
    reason="needs the plugin sources and the built library")


@pytest.fixture(scope="module")
def seam_check():
    # parallel test workers each run this fixture: one make at a time into the shared _build/
    os.makedirs(os.path.join(SEAM, "_build"), exist_ok=True)
    with open(os.path.join(SEAM, "_build", ".lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        subprocess.run(["make", "-s", "-j8", "-C", SEAM], check=True, capture_output=True, timeout=600)
    return os.path.join(SEAM, "_build", "seam_check")


def _run(exe, what):
    out = subprocess.run([exe, what], capture_output=True, text=True, timeout=120)
    lines = {}
    for line in out.stdout.splitlines():
        name, _, rest = line.partition(" ")
        lines[name] = rest
    return out.returncode, lines, out.stdout


def _device_present():
    import torch
    return torch.cuda.is_available()


def test_seam_tags(seam_check):
    rc, lines, text = _run(seam_check, "tags")
    assert rc == 0, text
    assert "error=missing-source-or-destination-buffer" in lines["null-source"], text
    assert "error=missing-source-or-destination-buffer" in lines["null-destination"], text
    assert "error=invalid-source-bounds" in lines["empty-source-width"], text
    assert "error=invalid-source-bounds" in lines["negative-source-height"], text
    for case in ("null-source", "null-destination", "empty-source-width", "negative-source-height"):
        assert lines[case].startswith("ok"), text
        assert "untouched=1" in lines[case], text
    well = lines["well-formed"]
    if _device_present():
        assert well.startswith("rendered") and "cache_valid=1" in well, text
    else:
        # no HIP device: the shared context cannot be made; nothing is written, nothing cached
        assert well == "failed error=hip-device-unavailable untouched=1 cache_valid=0", text
    assert lines["well-formed-repeat"].startswith("ok"), text


def test_seam_parameter_translation(seam_check):
    rc, lines, text = _run(seam_check, "params")
    assert rc == 0, text
    failed = [k for k, v in lines.items() if v.startswith("FAIL")]
    assert not failed, failed
    # every forwarded LensDiffParams member was checked, plus the sizes sweep and the bank key
    assert len(lines) >= 72, text
    assert lines["sizes"].startswith("ok 1386 cases, 0 differ"), text
    assert lines["bank-key"].startswith("ok"), text

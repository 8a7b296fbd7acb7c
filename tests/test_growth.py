"""The growth recurrence advanced in binade-sized jumps equals the step-by-step recurrence.

The kernels advance a cell's half-diagonal over the steps since it was last looked at with
exaca_amd/csrc/growth.hpp (the constant-increment plateau of the growth table crossed one binade
at a time). tests/growth_check.cpp compares that, bit for bit, with one float rounding per step
(src/CAupdate.cpp:257-261) on random values, values next to powers of two, the 0.045 cap, and
increments that make every step a rounding tie.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("hipcc") is None and not Path("/opt/rocm/bin/hipcc").exists(),
                    reason="hipcc not installed")
def test_jumped_growth_matches_stepwise(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "growth_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
                    "-I", str(REPO / "exaca_amd" / "csrc"), str(REPO / "tests" / "growth_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout


@pytest.mark.skipif(shutil.which("hipcc") is None and not Path("/opt/rocm/bin/hipcc").exists(),
                    reason="hipcc not installed")
def test_threshold_note_bound(tmp_path):
    """The claim pass's threshold note (exaca_amd/csrc/threshold.hpp): its bound never exceeds the
    next-least critical length, so a wake-up it lets skip the row reaches nothing; a copy that
    rounds the gap up fails on the same cases (tests/threshold_check.cpp)."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "threshold_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
                    "-I", str(REPO / "exaca_amd" / "csrc"), str(REPO / "tests" / "threshold_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures 0 loose 0" in r.stdout

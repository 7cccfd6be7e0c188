"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every
exported function of oracle/mc_oracle.cpp on small grids (the reference's
unit-test inputs, ragged random fields with origin and spacing, degenerate
grids, the adjoints, the OpenMP pipeline), built and run natively by
`make -C oracle sanitize`.  Any sanitizer report or failed known answer
fails the run."""
import shutil
import subprocess
from pathlib import Path

import pytest

ORACLE = Path(__file__).resolve().parent.parent / "oracle"


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                    reason="needs g++ and make")
def test_oracle_under_asan_ubsan():
    p = subprocess.run(["make", "-s", "-B", "-C", str(ORACLE), "sanitize"], capture_output=True,
                       text=True, timeout=600)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-4000:]
    assert "sanitize ok" in out
    assert "runtime error" not in out and "AddressSanitizer" not in out

"""The C++ tests of the Dax API slice (tests/cpp, built by
__graft_entry__.build() with hipcc against libdaxhip.so), run on the GPU."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "cpp" / "bin"
TESTS = ["UnitTestDeviceAdapterHip", "UnitTestWorkletMarchingCubesHip",
         "UnitTestWorkletCellGradientHip", "UnitTestWorkletSliceHip",
         "UnitTestWorkletMarchingTetrahedraHip", "UnitTestWorkletThresholdHip"]


@pytest.mark.parametrize("name", TESTS)
def test_cpp(name):
    exe = BIN / name
    assert exe.exists(), f"{exe} not built (run __graft_entry__.build())"
    print(f"[case] cpp {name}", flush=True)
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=240)
    print(r.stdout[-4000:])
    assert r.returncode == 0 and "PASSED" in r.stdout

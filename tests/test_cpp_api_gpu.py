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


def magnitude_field(n):
    """|p| on the grid 0..n-1 per axis in float32, summed from x as
    dax::math::Magnitude does (dax/math/VectorAnalysis.h:95-130)."""
    import numpy as np
    ax = np.arange(n, dtype=np.float32)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    return np.sqrt((x * x + y * y) + z * z).astype(np.float32)


@pytest.mark.parametrize("pipeline", [1, 2])
def test_benchmark_pipeline(pipeline):
    """The reference benchmark's pipeline (Benchmarks/MarchingCubes/
    Pipeline.h:100-132: Magnitude through DispatcherMapField, then
    MarchingCubesCount and MarchingCubesGenerate through the two
    dispatchers, iso 100) built against include/dax: its output sizes equal
    the oracle's on the same field, with (pipeline 2) and without merging."""
    import json

    from oracle import oracle as orc
    exe = BIN / "BenchmarkPipelineMarchingCubesHip"
    assert exe.exists(), f"{exe} not built (run __graft_entry__.build())"
    n = 96
    print(f"[case] cpp benchmark pipeline {pipeline}", flush=True)
    r = subprocess.run([str(exe), f"--size={n}", f"--pipeline={pipeline}", "--reps=2"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout[-4000:])
    assert r.returncode == 0 and "PASSED" in r.stdout
    line = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
    ref = orc.generate(magnitude_field(n), 100.0, dedup=(pipeline == 2))
    assert line["triangles"] == ref["n_tris"] > 0
    assert line["points"] == ref["n_pts"]
    assert line["cells"] == (n - 1) ** 3 and len(line["seconds"]) == 2

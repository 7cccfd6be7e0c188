"""Edge cases of the unstructured contour and threshold paths against the
oracle: no cells, no output, values exactly at the isovalue or the range
bounds (ties: `>` for the case, `>=`/`<=` for the range), NaN fields, and an
empty range (min > max)."""
import numpy as np
import pytest
import torch

from daxtoolkit_amd import mc, topology
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("shape", ["tets", "hexes"])
def test_contour_ties_empty_and_full(shape):
    builder = orc.freudenthal_tets if shape == "tets" else orc.voxel_hexes
    run = mc.tetrahedra if shape == "tets" else mc.hexahedra
    ref_fn = orc.mt_generate if shape == "tets" else orc.mc_hex_generate
    cells, pts = builder((11, 9, 7))
    rng = np.random.default_rng(12)
    # quantised values: many points exactly at the isovalue 0.5
    f = (rng.integers(0, 5, pts.shape[0]) / 4.0).astype(np.float32)
    for iso, label in ((0.5, "ties"), (2.0, "all below"), (-1.0, "all above")):
        print(f"[case] {shape} {label}", flush=True)
        out = run(_cuda(cells), _cuda(f), _cuda(pts), iso)
        ref = ref_fn(cells, f, pts, iso)
        assert out.num_triangles == ref["n_tris"] and out.num_points == ref["n_pts"]
        assert np.array_equal(out.connectivity.cpu().numpy().reshape(-1), ref["conn"])
        assert np.array_equal(out.vertices.cpu().numpy().view(np.uint32),
                              ref["coords"].view(np.uint32))
        if label != "ties":
            assert out.num_triangles == 0
    print(f"[case] {shape} no cells", flush=True)
    nv = cells.shape[1]
    out = run(torch.empty((0, nv), dtype=torch.int32, device="cuda"), _cuda(f), _cuda(pts), 0.5)
    assert out.num_triangles == 0 and out.num_points == 0


def test_threshold_bounds_nan_and_empty():
    dims = (68, 20, 9)  # nx % 4 == 0: the vector count kernel and the bit-word layout
    rng = np.random.default_rng(13)
    f = (rng.integers(0, 9, dims[::-1]) / 8.0).astype(np.float32)  # values on the bounds
    f[3, 4, 5] = np.nan
    f[7, 10, 60] = np.nan
    cases = ((0.25, 0.75, "bounds"), (0.0, 1.0, "all but NaN cells"), (0.8, 0.2, "empty range"),
             (2.0, 3.0, "nothing in range"))
    for lo, hi, label in cases:
        print(f"[case] threshold {label}", flush=True)
        ft = _cuda(f)
        out = topology.threshold(ft, lo, hi)
        ref = orc.threshold(f, lo, hi)
        assert out.num_cells == ref["n_cells"] and out.num_points == ref["n_points"]
        assert np.array_equal(out.counts.cpu().numpy(), ref["counts"])
        assert np.array_equal(out.connectivity.cpu().numpy(), ref["conn"])
        assert np.array_equal(out.points.cpu().numpy(), ref["coords"])
        assert np.array_equal(out.point_mask.cpu().numpy(), ref["mask"])
        vals = out.compact_point_field(ft.reshape(-1)).cpu().numpy()
        assert np.array_equal(vals, ref["field"])
        if label in ("empty range", "nothing in range"):
            assert out.num_cells == 0 and out.num_points == 0


def test_threshold_unstructured_nothing_kept():
    tets, pts = orc.freudenthal_tets((6, 5, 4))
    f = np.full(pts.shape[0], 5.0, np.float32)
    print("[case] threshold tets nothing kept", flush=True)
    out = topology.threshold(_cuda(f), 0.0, 1.0, cells=_cuda(tets), points=_cuda(pts))
    assert out.num_cells == 0 and out.num_points == 0
    assert int(out.point_mask.sum()) == 0

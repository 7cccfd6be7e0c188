"""The oracle's Threshold (ThresholdCount / ThresholdTopology through
DispatcherMapCell and DispatcherGenerateTopology, dax/worklet/Threshold.h:
30-110, dax/cont/DispatcherGenerateTopology.h:146-256): the reference's unit
test (dax/worklet/testing/UnitTestWorkletThreshold.cxx:38-243: 26^3,
f = x+y+z, [70, 82]) with its checks and the closed-form cell count, and an
independent numpy restatement (masks and ranks by np.unique instead of the
reference's sort/unique/lower-bounds sequence) on random uniform and
unstructured grids."""
import numpy as np
import pytest

from oracle import oracle as orc

CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def _uniform_cells(dims):
    hexes, pts = orc.voxel_hexes(dims)
    return hexes, pts


def _numpy_threshold(cells, f, pts, lo, hi, dedup):
    keep = np.all((f[cells] >= lo) & (f[cells] <= hi), axis=1)
    conn = cells[keep]
    out = {"counts": keep.astype(np.int32), "n_cells": int(keep.sum())}
    if dedup:
        used = np.unique(conn)
        out["n_points"] = used.size
        out["coords"] = pts[used]
        out["conn"] = np.searchsorted(used, conn).astype(np.int32)
        mask = np.zeros(f.size, np.int32)
        mask[used] = 1
        out["mask"] = mask
        out["field"] = f[used]
    else:
        out["conn"] = conn
    return out


def _same(a, b, dedup):
    assert a["n_cells"] == b["n_cells"]
    assert np.array_equal(a["counts"], b["counts"])
    assert np.array_equal(a["conn"], b["conn"])
    if dedup:
        assert a["n_points"] == b["n_points"]
        assert np.array_equal(a["coords"].view(np.uint32), b["coords"].view(np.uint32))
        assert np.array_equal(a["mask"], b["mask"])
        assert np.array_equal(a["field"], b["field"])


def test_reference_unit_test_26():
    n = 26
    f = np.add.outer(np.add.outer(np.arange(n), np.arange(n)), np.arange(n)).astype(np.float32)
    r = orc.threshold(f, 70, 82)
    # closed form: voxel (i,j,k) passes iff its corner sums i+j+k .. i+j+k+3 lie in [70, 82]
    expect = sum(1 for i in range(n - 1) for j in range(n - 1) for k in range(n - 1)
                 if 70 <= i + j + k and i + j + k + 3 <= 82)
    assert r["n_cells"] == expect == 10
    # the reference's checks: compacted values in range, connections sized,
    # each output cell's points are the input cell's coordinates in order
    assert np.all((r["field"] >= 70) & (r["field"] <= 82))
    assert r["field"].size == r["n_points"]
    hexes, pts = _uniform_cells((n, n, n))
    kept = hexes[r["counts"] != 0]
    assert r["conn"].shape == (r["n_cells"], 8)
    assert np.array_equal(r["coords"][r["conn"]], pts[kept])


@pytest.mark.parametrize("dedup", [True, False])
@pytest.mark.parametrize("dims", [(9, 7, 6), (17, 5, 12), (2, 2, 2)])
def test_uniform_matches_numpy(dims, dedup):
    rng = np.random.default_rng(sum(dims))
    nx, ny, nz = dims
    f = rng.random((nz, ny, nx)).astype(np.float32)
    hexes, pts = _uniform_cells(dims)
    origin, spacing = (0.5, -1.0, 2.0), (0.25, 1.5, 1.0)
    _, pts = orc.voxel_hexes(dims, origin, spacing)
    a = orc.threshold(f, 0.1, 0.9, dedup=dedup, origin=origin, spacing=spacing)
    b = _numpy_threshold(hexes, f.reshape(-1), pts, 0.1, 0.9, dedup)
    _same(a, b, dedup)


@pytest.mark.parametrize("dedup", [True, False])
def test_unstructured_shuffled_tets_match_numpy(dedup):
    rng = np.random.default_rng(3)
    tets, pts = orc.freudenthal_tets((10, 9, 8))
    perm = rng.permutation(pts.shape[0]).astype(np.int32)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size, dtype=np.int32)
    tets, pts = inv[tets], pts[perm]
    f = rng.random(pts.shape[0]).astype(np.float32)
    a = orc.threshold(f, 0.2, 0.95, dedup=dedup, cells=tets, points=pts)
    b = _numpy_threshold(tets, f, pts, 0.2, 0.95, dedup)
    _same(a, b, dedup)


def test_hexahedra_of_uniform_grid_match_uniform():
    dims = (11, 8, 9)
    rng = np.random.default_rng(8)
    f = rng.random(dims[::-1]).astype(np.float32)
    hexes, pts = orc.voxel_hexes(dims)
    a = orc.threshold(f, 0.05, 0.97)
    b = orc.threshold(f.reshape(-1), 0.05, 0.97, cells=hexes, points=pts)
    for k in ("counts", "conn", "mask", "field"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["coords"], b["coords"])


def test_nan_fails_the_range():
    f = np.ones((3, 3, 3), np.float32)
    f[1, 1, 1] = np.nan
    r = orc.threshold(f, 0.0, 2.0)
    assert r["n_cells"] == 0

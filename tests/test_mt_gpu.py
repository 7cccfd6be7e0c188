"""GPU marching tetrahedra (MarchingTetrahedraCount / MarchingTetrahedraGenerate,
dax/worklet/MarchingTetrahedra.h:29-143): the reference's unit test, every
tetrahedron case, and bit-exact parity with the oracle (connectivity,
interpolation records, coordinates) on Freudenthal grids with random fields,
with the point numbering shuffled, in both dedup modes."""
import numpy as np
import pytest
import torch

from daxtoolkit_amd import _lib, mc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _compare(tets, f, pts, iso, dedup, label):
    print(f"[case] mt {label} tets={tets.shape[0]} dedup={dedup}", flush=True)
    out = mc.tetrahedra(torch.from_numpy(tets).cuda(), torch.from_numpy(f).cuda(),
                        torch.from_numpy(pts).cuda(), iso, dedup=dedup)
    ref = orc.mt_generate(tets, f, pts, iso, dedup=dedup)
    assert out.num_triangles == ref["n_tris"] and out.num_points == ref["n_pts"]
    assert np.array_equal(out.connectivity.cpu().numpy().reshape(-1), ref["conn"])
    assert np.array_equal(out.interp_first.cpu().numpy(), ref["first"])
    assert np.array_equal(out.interp_second.cpu().numpy(), ref["second"])
    w = out.interp_weight.cpu().numpy()
    assert np.array_equal(w.view(np.uint32), ref["weight"].view(np.uint32))
    v = out.vertices.cpu().numpy()
    assert np.array_equal(v.view(np.uint32), ref["coords"].view(np.uint32))
    return out, ref


def test_reference_unit_test_26():
    tets, pts = orc.freudenthal_tets((26, 26, 26))
    f = (pts[:, 0] + pts[:, 1] + pts[:, 2]).astype(np.float32)
    out0, _ = _compare(tets, f, pts, 70.0, False, "unit26")
    out1, _ = _compare(tets, f, pts, 70.0, True, "unit26")
    assert out0.num_points == 3 * out0.num_triangles
    assert out1.num_points == 151


def test_every_case():
    """One tetrahedron per corner sign pattern, its ids in every order of a
    few permutations (the endpoint order follows the ids)."""
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    rng = np.random.default_rng(0)
    for perm in ([0, 1, 2, 3], [3, 2, 1, 0], [1, 3, 0, 2]):
        for c in range(16):
            f = np.array([0.9 if (c >> v) & 1 else 0.1 for v in range(4)], np.float32)
            f += rng.random(4).astype(np.float32) * 0.05
            tets = np.array([perm], np.int32)
            for dedup in (True, False):
                _compare(tets, f, pts, 0.5, dedup, f"case {c} perm {perm}")


@pytest.mark.parametrize("dims", [(9, 7, 6), (33, 17, 5), (48, 48, 48)])
def test_random_fields_shuffled_ids(dims):
    rng = np.random.default_rng(sum(dims))
    tets, pts = orc.freudenthal_tets(dims, origin=(-1.0, 0.5, 2.0), spacing=(0.5, 1.25, 0.75))
    f = rng.standard_normal(pts.shape[0]).astype(np.float32)
    for dedup in (True, False):
        _compare(tets, f, pts, 0.1, dedup, f"random {dims}")
    # renumber the points at random: endpoint order and dedup order change
    perm = rng.permutation(pts.shape[0]).astype(np.int32)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size, dtype=np.int32)
    tets2 = inv[tets]
    _compare(tets2, f[perm], pts[perm], 0.1, True, f"random {dims} shuffled")


def test_bad_ids_are_rejected():
    pts = torch.zeros((4, 3), device="cuda")
    f = torch.zeros(4, device="cuda")
    tets = torch.tensor([[0, 1, 2, 4]], dtype=torch.int32, device="cuda")
    with pytest.raises(_lib.ErrorControlBadValue):
        mc.tetrahedra(tets, f, pts, 0.5)

"""GPU marching cubes on an unstructured hexahedral grid (MarchingCubesCount /
MarchingCubesGenerate on UnstructuredGrid<Hexahedron>,
dax/worklet/MarchingCubes.h:53-164): the reference's unit test, every cube
case, bit-exact parity with the oracle (connectivity, interpolation records,
coordinates) on random fields with origin and spacing and with the point
numbering shuffled, in both dedup modes; and at 256^3 the hexahedra of the
uniform grid give the uniform-grid forward's output bit for bit."""
import numpy as np
import pytest
import torch

from daxtoolkit_amd import _lib, fields, mc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _compare(hexes, f, pts, iso, dedup, label):
    print(f"[case] hex {label} hexes={hexes.shape[0]} dedup={dedup}", flush=True)
    out = mc.hexahedra(torch.from_numpy(hexes).cuda(), torch.from_numpy(f).cuda(),
                       torch.from_numpy(pts).cuda(), iso, dedup=dedup)
    ref = orc.mc_hex_generate(hexes, f, pts, iso, dedup=dedup)
    assert out.num_triangles == ref["n_tris"] and out.num_points == ref["n_pts"]
    assert np.array_equal(out.connectivity.cpu().numpy().reshape(-1), ref["conn"])
    assert np.array_equal(out.interp_first.cpu().numpy(), ref["first"])
    assert np.array_equal(out.interp_second.cpu().numpy(), ref["second"])
    w = out.interp_weight.cpu().numpy()
    assert np.array_equal(w.view(np.uint32), ref["weight"].view(np.uint32))
    v = out.vertices.cpu().numpy()
    assert np.array_equal(v.view(np.uint32), ref["coords"].view(np.uint32))
    return out


def test_reference_unit_test_26():
    hexes, pts = orc.voxel_hexes((26, 26, 26))
    f = (pts[:, 0] + pts[:, 1] + pts[:, 2]).astype(np.float32)
    out0 = _compare(hexes, f, pts, 70.0, False, "unit26")
    out1 = _compare(hexes, f, pts, 70.0, True, "unit26")
    assert out0.num_points == 3 * out0.num_triangles
    assert out1.num_points == 45


def test_every_case():
    """One hexahedron per corner sign pattern (256), in two vertex
    numberings (the endpoint order follows the ids)."""
    corner = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1),
                       (1, 1, 1), (0, 1, 1)], np.float32)
    rng = np.random.default_rng(1)
    for ids in ([0, 1, 2, 3, 4, 5, 6, 7], [6, 2, 7, 0, 5, 1, 3, 4]):
        ids = np.array(ids, np.int32)
        hexes = np.repeat(ids[None, :], 256, axis=0) + 8 * np.arange(256, dtype=np.int32)[:, None]
        pts = np.zeros((256 * 8, 3), np.float32)
        f = np.zeros(256 * 8, np.float32)
        for c in range(256):
            for v in range(8):
                p = hexes[c, v]
                pts[p] = corner[v] + np.float32(2 * c)
                f[p] = (0.9 if (c >> v) & 1 else 0.1) + np.float32(rng.random() * 0.05)
        for dedup in (True, False):
            _compare(hexes, f, pts, 0.5, dedup, f"all cases ids {ids.tolist()}")


@pytest.mark.parametrize("dims", [(9, 7, 6), (33, 17, 5), (40, 40, 40)])
def test_random_fields_shuffled_ids(dims):
    rng = np.random.default_rng(sum(dims) + 1)
    hexes, pts = orc.voxel_hexes(dims, origin=(-1.0, 0.5, 2.0), spacing=(0.5, 1.25, 0.75))
    f = rng.standard_normal(pts.shape[0]).astype(np.float32)
    for dedup in (True, False):
        _compare(hexes, f, pts, 0.1, dedup, f"random {dims}")
    perm = rng.permutation(pts.shape[0]).astype(np.int32)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size, dtype=np.int32)
    _compare(inv[hexes], f[perm], pts[perm], 0.1, True, f"random {dims} shuffled")


def test_unaligned_connectivity():
    """Connectivity not 16-byte aligned takes the scalar-load kernels."""
    hexes, pts = orc.voxel_hexes((12, 11, 10))
    rng = np.random.default_rng(9)
    f = rng.standard_normal(pts.shape[0]).astype(np.float32)
    flat = torch.from_numpy(np.concatenate([[0], hexes.reshape(-1)]).astype(np.int32)).cuda()
    view = flat[1:].view(-1, 8)
    assert view.data_ptr() % 16 != 0
    print("[case] hex unaligned", flush=True)
    out = mc.hexahedra(view, torch.from_numpy(f).cuda(), torch.from_numpy(pts).cuda(), 0.2)
    ref = orc.mc_hex_generate(hexes, f, pts, 0.2)
    assert np.array_equal(out.connectivity.cpu().numpy().reshape(-1), ref["conn"])
    assert np.array_equal(out.vertices.cpu().numpy().view(np.uint32), ref["coords"].view(np.uint32))


def test_256_equals_uniform_forward():
    """The 16-Gaussian 256^3 field: the hexahedra of the uniform grid give
    the uniform forward's triangles, records and coordinate bits."""
    n = 256
    dims = (n, n, n)
    print("[case] hex 256^3 vs uniform forward", flush=True)
    field = fields.gaussian_mixture_torch(dims, fields.gaussian_mixture_params(0, dims))
    hexes, pts = fields.voxel_cells_torch(dims, "hexahedron")
    out = mc.hexahedra(hexes, field.reshape(-1), pts, 0.5)
    ref = mc.forward(field, 0.5, normals=False, interp=True)
    assert out.num_triangles == ref.num_triangles and out.num_points == ref.num_points
    assert out.num_triangles > 100_000
    assert torch.equal(out.connectivity.reshape(-1), ref.connectivity.reshape(-1))
    assert torch.equal(out.interp_first, ref.interp_first)
    assert torch.equal(out.interp_second, ref.interp_second)
    assert torch.equal(out.interp_weight.view(torch.int32), ref.interp_weight.view(torch.int32))
    assert torch.equal(out.vertices.view(torch.int32), ref.vertices.view(torch.int32))
    del hexes, pts, field
    torch.cuda.empty_cache()


def test_bad_ids_are_rejected():
    pts = torch.zeros((8, 3), device="cuda")
    f = torch.zeros(8, device="cuda")
    hexes = torch.tensor([[0, 1, 2, 3, 4, 5, 6, -1]], dtype=torch.int32, device="cuda")
    with pytest.raises(_lib.ErrorControlBadValue):
        mc.hexahedra(hexes, f, pts, 0.5)

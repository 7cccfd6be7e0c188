"""The oracle's marching cubes on an unstructured hexahedral grid
(dax/worklet/MarchingCubes.h:53-164 on UnstructuredGrid<Hexahedron>): the
reference's unit test on its hexahedral test grid
(dax/worklet/testing/UnitTestWorkletMarchingCubes.cxx:211-216 runs it on
both grid types), equality with the uniform-grid oracle when the hexahedra
are the voxels of a uniform grid (an independent code path: the uniform
oracle walks the grid's implicit topology), and a brute-force count of the
mesh's sign-change edges when the point ids are permuted."""
import numpy as np
import pytest

from oracle import oracle as orc

HEX_EDGES = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7),
             (0, 4), (1, 5), (2, 6), (3, 7)]


def test_reference_unit_test_26():
    """26^3 hexahedra, f = x+y+z, iso 70: 3T points unmerged, 45 merged."""
    hexes, pts = orc.voxel_hexes((26, 26, 26))
    f = (pts[:, 0] + pts[:, 1] + pts[:, 2]).astype(np.float32)
    r0 = orc.mc_hex_generate(hexes, f, pts, 70.0, dedup=False)
    r1 = orc.mc_hex_generate(hexes, f, pts, 70.0, dedup=True)
    assert r0["n_pts"] == 3 * r0["n_tris"] and r0["n_tris"] == r1["n_tris"]
    assert r1["n_pts"] == 45


@pytest.mark.parametrize("dims,origin,spacing,dedup", [
    ((9, 7, 6), (0, 0, 0), (1, 1, 1), True),
    ((13, 5, 11), (-1.5, 2.0, 0.25), (0.5, 1.25, 2.0), True),
    ((13, 5, 11), (-1.5, 2.0, 0.25), (0.5, 1.25, 2.0), False),
    ((2, 2, 2), (0, 0, 0), (1, 1, 1), True),
])
def test_equals_uniform_oracle(dims, origin, spacing, dedup):
    rng = np.random.default_rng(11)
    nx, ny, nz = dims
    field = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    hexes, pts = orc.voxel_hexes(dims, origin, spacing)
    a = orc.mc_hex_generate(hexes, field.reshape(-1), pts, 0.1, dedup=dedup)
    b = orc.generate(field, 0.1, dedup=dedup, origin=origin, spacing=spacing)
    assert a["n_tris"] == b["n_tris"] and a["n_pts"] == b["n_pts"]
    for k in ("conn", "first", "second"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("weight", "coords"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


def test_permuted_ids_dedup_points_are_the_sign_change_edges():
    rng = np.random.default_rng(5)
    hexes, pts = orc.voxel_hexes((8, 9, 7))
    perm = rng.permutation(pts.shape[0]).astype(np.int32)  # old id -> new id
    hexes = perm[hexes]
    pts2 = np.empty_like(pts)
    pts2[perm] = pts
    f = rng.standard_normal(pts.shape[0]).astype(np.float32)
    r = orc.mc_hex_generate(hexes, f, pts2, -0.3, dedup=True)
    e = np.concatenate([hexes[:, [a, b]] for a, b in HEX_EDGES])
    e = np.unique(np.sort(e, axis=1), axis=0)
    cut = e[(f[e[:, 0]] > -0.3) != (f[e[:, 1]] > -0.3)]
    assert r["n_pts"] == cut.shape[0]
    assert np.array_equal(np.stack([r["first"], r["second"]], axis=1), cut)
    # every triangle's corners are points of one hexahedron's edges
    nf, _ = orc.tables()
    cls = np.zeros(hexes.shape[0], np.int64)
    for v in range(8):
        cls |= (f[hexes[:, v]] > -0.3).astype(np.int64) << v
    assert r["n_tris"] == int(nf[cls].sum())


@pytest.mark.parametrize("shape,builder", [("tetrahedron", orc.freudenthal_tets),
                                           ("hexahedron", orc.voxel_hexes)])
def test_package_voxel_cells_match_oracle_builder(shape, builder):
    """The package's device mesh builder (used by the benchmark) and the
    oracle's numpy one give the same cells and the same coordinate bits."""
    from daxtoolkit_amd import fields
    dims, origin, spacing = (7, 5, 4), (-1.5, 2.0, 0.25), (0.5, 1.25, 2.0)
    c, p = fields.voxel_cells_torch(dims, shape, origin, spacing, device="cpu")
    c2, p2 = builder(dims, origin, spacing)
    assert np.array_equal(c.numpy(), c2)
    assert np.array_equal(p.numpy().view(np.uint32), p2.view(np.uint32))

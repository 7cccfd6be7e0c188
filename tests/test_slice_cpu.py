"""The oracle's slice (Dax's SliceCount / SliceGenerate on a uniform grid,
dax/worklet/Slice.h:28-173): its point values against an independent numpy
float32 restatement of the same operation order, the reference's unit test
(dax/worklet/testing/UnitTestWorkletSlice.cxx:39-121) and a closed form."""
import numpy as np

from oracle import oracle as orc


def _numpy_slice_field(dims, po, pn, origin, spacing):
    nx, ny, nz = dims
    f32 = np.float32
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = [f32(origin[d]) + f32(spacing[d]) * a.astype(f32) for d, a in enumerate((x, y, z))]
    r = f32(pn[0]) * (c[0] - f32(po[0]))
    r = r + f32(pn[1]) * (c[1] - f32(po[1]))
    r = r + f32(pn[2]) * (c[2] - f32(po[2]))
    iso = f32(pn[0]) * f32(po[0])
    iso = iso + f32(pn[1]) * f32(po[1])
    iso = iso + f32(pn[2]) * f32(po[2])
    return r.astype(f32), f32(iso)


def test_slice_field_matches_numpy_restatement():
    for dims, po, pn, org, sp in (((8, 8, 8), (3, 2, 1), (1, 0, 0), (0, 0, 0), (1, 1, 1)),
                                  ((17, 9, 13), (0.3, -1.25, 2.5), (0.48, -0.6, 0.64),
                                   (-1.0, 0.5, 2.0), (0.25, 0.5, 0.75)),
                                  ((31, 5, 7), (1.0, 1.0, 1.0), (0.2, 0.3, -0.9), (0, 0, 0),
                                   (1.5, 1.0, 0.5))):
        f, iso = orc.slice_field(dims, po, pn, origin=org, spacing=sp)
        g, giso = _numpy_slice_field(dims, po, pn, org, sp)
        assert f.shape == (dims[2], dims[1], dims[0])
        assert np.array_equal(f.view(np.uint32), g.view(np.uint32))
        assert np.float32(iso).view(np.uint32) == giso.view(np.uint32)


def test_slice_reference_unit_test():
    """DIM 8, ORIGIN (3,2,1), NORMAL (1,0,0): 2 (DIM-1)^2 triangles, 3 points
    each unmerged, DIM^2 merged."""
    f, iso = orc.slice_field((8, 8, 8), (3, 2, 1), (1, 0, 0))
    r0 = orc.generate(f, iso, dedup=False)
    r1 = orc.generate(f, iso, dedup=True)
    assert r0["n_tris"] == r1["n_tris"] == 2 * 7 * 7
    assert r0["n_pts"] == 3 * 98
    assert r1["n_pts"] == 64


def test_slice_axis_plane_closed_form():
    """The plane y = 2 * 2.25 = 4.5 (the reference's value dot(n, c - o)
    against dot(n, o)) cuts every cell between y = 4 and 5: 2 triangles per
    cell, one point per y-edge, each at weight 1/2."""
    dims = (13, 11, 9)
    f, iso = orc.slice_field(dims, (0.0, 2.25, 0.0), (0.0, 1.0, 0.0))
    r = orc.generate(f, iso, dedup=True)
    assert r["n_tris"] == 2 * 12 * 8
    assert r["n_pts"] == 13 * 9
    assert np.all(r["weight"] == np.float32(0.5))
    assert np.all(r["coords"][:, 1] == np.float32(4.5))

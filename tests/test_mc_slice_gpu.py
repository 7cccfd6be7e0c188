"""GPU slice (SliceCount / SliceGenerate, dax/worklet/Slice.h:28-173, with
the plane distance computed in the kernels): the reference's unit test,
bit-exact parity with the oracle on tilted planes over ragged grids, and
equality bit for bit with the marching-cubes forward over the materialised
field at 256^3."""
import numpy as np
import pytest
import torch

from daxtoolkit_amd import mc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _compare(dims, po, pn, origin=(0, 0, 0), spacing=(1, 1, 1), dedup=True):
    print(f"[case] slice dims={dims} plane={po},{pn} dedup={dedup}", flush=True)
    out = mc.slice(dims, po, pn, origin=origin, spacing=spacing, dedup=dedup, normals=True,
                   interp=True)
    torch.cuda.synchronize()
    f, iso = orc.slice_field(dims, po, pn, origin=origin, spacing=spacing)
    ref = orc.generate(f, iso, dedup=dedup, origin=origin, spacing=spacing)
    assert out.num_triangles == ref["n_tris"] and out.num_points == ref["n_pts"]
    assert np.array_equal(out.connectivity.cpu().numpy().reshape(-1), ref["conn"])
    assert np.array_equal(out.interp_first.cpu().numpy(), ref["first"])
    assert np.array_equal(out.interp_second.cpu().numpy(), ref["second"])
    w = out.interp_weight.cpu().numpy()
    assert np.array_equal(w.view(np.uint32), ref["weight"].view(np.uint32))
    v = out.vertices.cpu().numpy()
    assert np.array_equal(v.view(np.uint32), ref["coords"].view(np.uint32))
    nref = orc.normals(f, ref["first"], ref["second"], ref["weight"], spacing=spacing)
    assert np.array_equal(out.normals.cpu().numpy().view(np.uint32), nref.view(np.uint32))
    return out, ref


def test_slice_reference_unit_test():
    """UnitTestWorkletSlice.cxx:39-121: DIM 8, ORIGIN (3,2,1), NORMAL
    (1,0,0) gives 98 triangles, 294 points unmerged and 64 merged."""
    out0, _ = _compare((8, 8, 8), (3, 2, 1), (1, 0, 0), dedup=False)
    out1, _ = _compare((8, 8, 8), (3, 2, 1), (1, 0, 0), dedup=True)
    assert out0.num_triangles == out1.num_triangles == 98
    assert out0.num_points == 294 and out1.num_points == 64


@pytest.mark.parametrize("dims,po,pn,org,sp", [
    ((17, 9, 13), (0.3, -1.25, 2.5), (0.48, -0.6, 0.64), (-1.0, 0.5, 2.0), (0.25, 0.5, 0.75)),
    ((130, 33, 17), (10.0, 5.0, 3.0), (0.2, 0.3, -0.9), (0, 0, 0), (1.5, 1.0, 0.5)),
    ((64, 64, 64), (16.0, 16.0, 16.0), (1.0, 1.0, 1.0), (0, 0, 0), (1, 1, 1)),
    ((200, 2, 2), (50.0, 0.0, 0.0), (1.0, 0.01, 0.0), (0, 0, 0), (1, 1, 1))])
def test_slice_matches_oracle(dims, po, pn, org, sp):
    for dedup in (True, False):
        _compare(dims, po, pn, origin=org, spacing=sp, dedup=dedup)


def test_slice_equals_forward_on_materialised_field():
    """256^3, a tilted plane: the slice (no field read) and the forward over
    the oracle's materialised values give the same outputs bit for bit.  (The
    reference's surface is dot(n, c) = 2 dot(n, o): here near the centre.)"""
    dims = (256, 256, 256)
    po, pn = (64.0, 64.0, 50.0), (0.36, 0.48, 0.8)
    print("[case] slice 256^3 vs forward", flush=True)
    s = mc.slice(dims, po, pn, normals=True, interp=True)
    f, iso = orc.slice_field(dims, po, pn)
    m = mc.forward(torch.from_numpy(f).cuda(), float(iso), normals=True, edges=True, interp=True)
    torch.cuda.synchronize()
    assert s.num_triangles == m.num_triangles > 100_000
    assert s.num_points == m.num_points
    for a, b in ((s.connectivity, m.connectivity), (s.vertices, m.vertices),
                 (s.normals, m.normals), (s.edges, m.edges), (s.interp_weight, m.interp_weight)):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))

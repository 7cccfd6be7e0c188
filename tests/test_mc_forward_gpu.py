"""GPU parity of the marching-cubes forward against the oracle.

Topology (case index, per-cell count, scan offsets, connectivity, vertex
order) must be bit-exact; vertex coordinates and normals within 1e-5 relative
to the vector magnitude (floor 1 grid spacing for coordinates).  Because
both sides compile with FP contraction off and IEEE division/sqrt, the FP32
outputs are in fact expected to be bit-identical, which is asserted as well.
"""
import numpy as np
import pytest
import torch

from daxtoolkit_amd import fields, mc
from helpers import dax_test_equal, rel_err_vec
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _run_and_compare(f_host, iso, dedup=True, origin=(0, 0, 0), spacing=(1, 1, 1), label=""):
    print(f"[case] {label} dims={f_host.shape[::-1]} iso={iso} dedup={dedup}", flush=True)
    field = torch.from_numpy(np.ascontiguousarray(f_host)).cuda()
    out = mc.forward(field, iso, origin=origin, spacing=spacing, dedup=dedup, normals=True,
                     edges=dedup and f_host.size <= 2**30, interp=True)
    torch.cuda.synchronize()
    ref = orc.generate(f_host, iso, dedup=dedup, origin=origin, spacing=spacing)
    assert out.num_triangles == ref["n_tris"]
    assert out.num_points == ref["n_pts"]
    conn = out.connectivity.cpu().numpy().reshape(-1)
    assert np.array_equal(conn, ref["conn"]), "connectivity"
    first = out.interp_first.cpu().numpy()
    second = out.interp_second.cpu().numpy()
    w = out.interp_weight.cpu().numpy()
    assert np.array_equal(first, ref["first"]) and np.array_equal(second, ref["second"])
    assert np.array_equal(w.view(np.uint32), ref["weight"].view(np.uint32)), "weights"
    verts = out.vertices.cpu().numpy()
    assert rel_err_vec(verts, ref["coords"]) <= TOL
    assert np.array_equal(verts.view(np.uint32), ref["coords"].view(np.uint32))
    nref = orc.normals(f_host, ref["first"], ref["second"], ref["weight"], spacing=spacing)
    nrm = out.normals.cpu().numpy()
    assert rel_err_vec(nrm, nref, floor=1e-30) <= TOL
    assert np.array_equal(nrm.view(np.uint32), nref.view(np.uint32))
    if dedup and out.edges is not None:
        e = out.edges.cpu().numpy().view(np.uint32)
        axis = e >> 30
        pa = (e & ((1 << 30) - 1)).astype(np.int64)
        nz, ny, nx = f_host.shape
        stride = np.array([1, nx, nx * ny], np.int64)
        assert np.array_equal(pa, ref["first"])
        assert np.array_equal(pa + stride[axis], ref["second"])
    return out, ref


def test_reference_unit_test_linear_26():
    """UnitTestWorkletMarchingCubes.cxx:39-200: 26^3, f = x+y+z, iso 70."""
    f = fields.linear(26, 26, 26)
    out0, _ = _run_and_compare(f, 70.0, dedup=False, label="unit26 nodedup")
    assert out0.num_points == 3 * out0.num_triangles
    out1, _ = _run_and_compare(f, 70.0, dedup=True, label="unit26 dedup")
    assert out1.num_points == 45
    # secondary linear field interpolates exactly (test_equal, 1e-4)
    sec = fields.linear(26, 26, 26, (-0.5, 1.0, 2.0))
    for out in (out0, out1):
        vals = mc.interpolate_scalar(torch.from_numpy(sec).cuda().reshape(-1), out).cpu().numpy()
        xyz = out.vertices.cpu().numpy().astype(np.float64)
        expect = xyz @ np.array([-0.5, 1.0, 2.0])
        assert dax_test_equal(vals, expect)


@pytest.mark.parametrize("dedup", [True, False])
def test_sphere_128(dedup):
    """128^3 sphere distance, iso 40 (the CPU-runnable configuration)."""
    f = fields.sphere_distance(128)
    out, ref = _run_and_compare(f, 40.0, dedup=dedup, label="sphere128")
    assert out.num_triangles == 60284
    if dedup:
        assert out.num_points == 30144


def test_sphere_128_cell_outputs():
    """Per-cell case, count and exclusive offsets, bit-exact."""
    f = fields.sphere_distance(128)
    field = torch.from_numpy(f).cuda()
    ws, cnt = mc.classify(field, 40.0)
    cases, count, offs = mc.cell_outputs(field, ws)
    T, U = mc.read_counts(cnt)
    rc, rn = orc.count(f, 40.0)
    ro, total = orc.scan_exclusive(rn)
    assert np.array_equal(cases.cpu().numpy(), rc)
    assert np.array_equal(count.cpu().numpy(), rn)
    assert np.array_equal(offs.cpu().numpy(), ro)
    assert T == total == 60284 and U == 30144


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 5, 7), (65, 3, 4), (64, 64, 3), (127, 9, 11),
                                  (130, 33, 17), (200, 2, 2), (1, 5, 5), (5, 1, 5)])
def test_random_fields(dims):
    rng = np.random.default_rng(hash(dims) % 2**32)
    nx, ny, nz = dims
    f = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    for dedup in (True, False):
        _run_and_compare(f, 0.1, dedup=dedup, origin=(-1.5, 2.0, 0.25), spacing=(0.5, 2.0, 1.25),
                         label=f"random {dims}")


def test_all_256_cases():
    """Every corner sign pattern, one cell each, against the oracle and the
    case table's triangle count."""
    cells = []
    for c in range(256):
        cube = np.zeros((2, 2, 2), np.float32)
        corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1),
                   (0, 1, 1)]
        for v, (x, y, z) in enumerate(corners):
            cube[z, y, x] = 1.0 if (c >> v) & 1 else -1.0
        cells.append(cube)
    nf, _ = orc.tables()
    for c, cube in enumerate(cells):
        out, ref = _run_and_compare(cube, 0.0, label=f"case {c}")
        assert out.num_triangles == nf[c]


def test_ties_at_iso():
    """Quantised field with many values equal to the iso (ties are 'below')."""
    rng = np.random.default_rng(7)
    f = (rng.integers(0, 11, size=(40, 50, 70)) / 10.0).astype(np.float32)
    _run_and_compare(f, 0.5, label="ties")


def test_empty_and_full():
    f = np.zeros((20, 20, 20), np.float32)
    for iso in (1.0, -1.0):
        out, _ = _run_and_compare(f, iso, label="empty")
        assert out.num_triangles == 0


def test_gaussian_mixture_96():
    dims = (96, 80, 72)
    p = fields.gaussian_mixture_params(3, dims, n_gauss=8)
    f = fields.gaussian_mixture_numpy(dims, p)
    _run_and_compare(f, 0.5, label="gm96")


def test_gaussian_mixture_512_full():
    """The 512^3 benchmark field (seed 0), full parity with the oracle."""
    dims = (512, 512, 512)
    p = fields.gaussian_mixture_params(0, dims)
    field = fields.gaussian_mixture_torch(dims, p)
    f_host = field.cpu().numpy()
    print("[case] gm512 generated", flush=True)
    _run_and_compare(f_host, 0.5, label="gm512")


def test_point_and_cell_gradient():
    dims = (33, 21, 17)
    p = fields.gaussian_mixture_params(5, dims, n_gauss=4, sigma=(4.0, 9.0))
    f = fields.gaussian_mixture_numpy(dims, p)
    field = torch.from_numpy(f).cuda()
    sp, org = (0.5, 1.5, 2.0), (1.0, -2.0, 3.0)
    g = mc.point_gradient(field, spacing=sp).cpu().numpy()
    gr = orc.point_gradient(f, spacing=sp)
    assert np.array_equal(g.view(np.uint32), gr.view(np.uint32))
    cg = mc.cell_gradient(field, origin=org, spacing=sp).cpu().numpy()
    cr = orc.cell_gradient(f, origin=org, spacing=sp)
    assert np.array_equal(cg.view(np.uint32), cr.view(np.uint32))


def test_cell_gradient_sphere_128():
    """CellGradient (dax/worklet/CellGradient.h:28-48) on the 128^3 point
    sphere-distance field of the CPU workload, bit for bit against the
    oracle over all 127^3 cells.  The field is |p - c|, so away from the
    centre (cell centres at distance >= 3) each gradient is within 3% of the
    unit vector (p - c)/|p - c| in length and within 1e-4 in direction."""
    f = fields.sphere_distance(128)
    print("[case] cell gradient sphere 128", flush=True)
    field = torch.from_numpy(f).cuda()
    cg = mc.cell_gradient(field).cpu().numpy()
    cr = orc.cell_gradient(f)
    assert cg.shape == (127 ** 3, 3)
    assert np.array_equal(cg.view(np.uint32), cr.view(np.uint32))
    ax = np.arange(127) + 0.5 - 63.5
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z).reshape(-1)
    far = r >= 3.0
    g = cg.astype(np.float64)[far]
    norm = np.linalg.norm(g, axis=1)
    assert np.all(np.abs(norm - 1.0) < 0.03)
    unit = np.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1)], 1)[far] / r[far, None]
    assert np.all(np.einsum("ij,ij->i", g, unit) / norm > 1.0 - 1e-4)


@pytest.mark.parametrize("shape_kind", ["aligned", "ragged"])
def test_every_tuning_variant_is_bit_identical(shape_kind):
    """Each classify / scan / generate variant gives identical outputs."""
    if shape_kind == "aligned":
        f = fields.sphere_distance(128)
        iso = 40.0
    else:
        dims = (100, 70, 45)
        p = fields.gaussian_mixture_params(9, dims, n_gauss=6, sigma=(6.0, 15.0))
        f = fields.gaussian_mixture_numpy(dims, p)
        iso = 0.5
    nz, ny, nx = f.shape
    field = torch.from_numpy(f).cuda()
    default = mc.get_tuning((0, 0, 0))
    variants = ([[u, 1, 4, 512] for u in (8, 16, 32)] + [[u, 4, 4, 512] for u in (2, 4, 8)] +
                [[16, 1, s, 512] for s in (1, 2, 4, 8)] +
                [[16, 1, 4, t, x] for t in (64, 128, 256, 512, 1024) for x in (0, 1)])
    ref = None
    try:
        for v in variants:
            print(f"[case] tuning {shape_kind} {v}", flush=True)
            mc.set_tuning((nx, ny, nz), v)
            out = mc.forward(field, iso, edges=True, interp=True)
            got = [out.vertices.cpu().numpy(), out.normals.cpu().numpy(),
                   out.connectivity.cpu().numpy(), out.edges.cpu().numpy()]
            if ref is None:
                ref = got
            else:
                for a, b in zip(got, ref):
                    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    finally:
        mc.set_tuning((nx, ny, nz), default)

"""1024^3 Mandelbulb (power 8, <= 10 iterations, bailout 2, [-1.2, 1.2]^3,
iso 0.5): the oracle does not run at this size, so the forward and the
adjoint are checked through properties the mathematics fixes, computed
independently with torch on the same field:
- U = the number of lattice edges whose end points differ in `f > iso`;
- T = the sum over cells of NumFaces[case] (case from the 8 corner bits in
  the reference's corner order, NumFaces from the oracle's table);
- every vertex lies on its recorded edge, between the edge's end points, and
  every normal has unit length or is zero;
- every triangle's ids are < U and its three edges lie in one unit cube;
- the adjoint is finite, deterministic, and zero wherever the 5x5x5 box
  around a point holds no inside/outside change."""
import numpy as np
import pytest
import torch

from daxtoolkit_amd import fields, mc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

N = 1024
ISO = 0.5
CHUNK = 64  # planes per torch chunk


@pytest.fixture(scope="module")
def bulb():
    print(f"[case] mandelbulb {N}^3", flush=True)
    f = fields.mandelbulb_torch(N)
    out = mc.forward(f, ISO, edges=True)
    torch.cuda.synchronize()
    yield f, out
    del f, out
    torch.cuda.empty_cache()


def _inside(f, z0, z1):
    return f[z0:z1] > ISO


def test_counts_match_edges_and_cases(bulb):
    f, out = bulb
    nf, _ = orc.tables()
    num_faces = torch.from_numpy(nf.astype(np.int64)).cuda()
    edges = 0
    tris = 0
    for z0 in range(0, N, CHUNK):
        z1 = min(N, z0 + CHUNK)
        b = _inside(f, z0, min(N, z1 + 1))  # one plane of halo
        own = b[: z1 - z0]
        edges += int((own[:, :, 1:] != own[:, :, :-1]).sum())
        edges += int((own[:, 1:, :] != own[:, :-1, :]).sum())
        edges += int((b[1:] != b[:-1]).sum())  # z edges with lower plane in z0 .. z1-1
        # cells of planes z0 .. min(z1, N-1) - 1
        nzc = min(z1, N - 1) - z0
        if nzc <= 0:
            continue
        c = torch.zeros((nzc, N - 1, N - 1), dtype=torch.int64, device="cuda")
        corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
        for v, (dx, dy, dz) in enumerate(corners):
            c |= b[dz:dz + nzc, dy:dy + N - 1, dx:dx + N - 1].to(torch.int64) << v
        tris += int(num_faces[c].sum())
        del c, b
    print(f"[info] U={out.num_points} edges={edges} T={out.num_triangles} faces={tris}", flush=True)
    assert out.num_points == edges
    assert out.num_triangles == tris
    assert out.num_triangles > 10_000_000


def test_vertices_normals_triangles(bulb):
    f, out = bulb
    e = out.edges.to(torch.int64) & 0xFFFFFFFF
    axis = e >> 30
    a = e & 0x3FFFFFFF
    ax = torch.stack([a % N, (a // N) % N, a // (N * N)], dim=1)  # lower end point (x, y, z)
    assert int(axis.max()) <= 2
    v = out.vertices.to(torch.float64)
    lo = ax.to(torch.float64)
    onehot = torch.nn.functional.one_hot(axis, 3).to(torch.float64)
    # off-axis coordinates are exact, the on-axis one is in [a, a+1]
    assert torch.equal(v * (1 - onehot), lo * (1 - onehot))
    t = ((v - lo) * onehot).sum(dim=1)
    assert float(t.min()) >= 0.0 and float(t.max()) <= 1.0
    n = out.normals.to(torch.float64).norm(dim=1)
    unit = (n - 1.0).abs() <= 1e-5
    assert bool((unit | (n == 0.0)).all())
    conn = out.connectivity.to(torch.int64)
    assert int(conn.min()) >= 0 and int(conn.max()) < out.num_points
    corner = ax[conn]  # [T, 3, 3]
    cmin = corner.min(dim=1).values
    d = corner - cmin[:, None, :]
    assert int(d.max()) <= 1 and int(d.min()) >= 0


def test_adjoint_deterministic_and_local(bulb):
    f, out = bulb
    U = out.num_points
    gen = torch.Generator(device="cuda").manual_seed(3)
    dx = torch.randn((U, 3), generator=gen, device="cuda")
    dn = torch.randn((U, 3), generator=gen, device="cuda")
    print(f"[case] mandelbulb adjoint U={U}", flush=True)
    g1 = mc.backward(f, ISO, out, dx, dn)
    g2 = mc.backward(f, ISO, out, dx, dn)
    assert torch.equal(g1, g2)
    assert bool(torch.isfinite(g1).all())
    assert int((g1 != 0).sum()) > 0
    for z0 in range(0, N, CHUNK):
        z1 = min(N, z0 + CHUNK)
        h0, h1 = max(0, z0 - 2), min(N, z1 + 2)
        b = _inside(f, h0, h1).to(torch.float16)[None, None]
        hi = torch.nn.functional.max_pool3d(b, 5, stride=1, padding=2)
        lo = -torch.nn.functional.max_pool3d(-b, 5, stride=1, padding=2)
        uniform = (hi == lo)[0, 0, z0 - h0: z0 - h0 + (z1 - z0)]
        assert int((g1[z0:z1][uniform] != 0).sum()) == 0
        del b, hi, lo, uniform


@pytest.mark.parametrize("tuning", [[4, 4, 8, 128, 0], [4, 4, 2, 256, 1]])
def test_small_generate_tiles_same_output(bulb, tuning):
    """Generate tiles of 128 and 256 words: 131072 and 65536 tiles, so each of
    the scan block's 1024 threads orders more than 32 tiles (the vector
    chunk path of the dispatch order).  The output is the default tiling's,
    bit for bit, in the same order."""
    f, out = bulb
    shape = (N, N, N)
    saved = mc.get_tuning(shape)
    print(f"[case] mandelbulb forward tuning {tuning}", flush=True)
    try:
        mc.set_tuning(shape, tuning)
        o2 = mc.forward(f, ISO, edges=True)
        torch.cuda.synchronize()
    finally:
        mc.set_tuning(shape, saved)
    assert (o2.num_points, o2.num_triangles) == (out.num_points, out.num_triangles)
    assert torch.equal(o2.edges, out.edges)
    assert torch.equal(o2.connectivity, out.connectivity)
    assert torch.equal(o2.vertices.view(torch.int32), out.vertices.view(torch.int32))
    assert torch.equal(o2.normals.view(torch.int32), out.normals.view(torch.int32))
    del o2

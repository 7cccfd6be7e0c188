"""GPU adjoint: bit for bit against the oracle's FP32 restatement of the
device operation order, against the float64 torch autograd restatement and
the oracle's FP64 adjoint (every point), against FP64 host central
differences (256^3 workload, 4096 points), and deterministic across runs."""
import numpy as np
import pytest
import torch

from daxtoolkit_amd import _lib, fields, mc
from mc_autograd import MarchingCubesFunction, torch_reference_loss
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _bit_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


def _close(got, ref, rtol, floor_frac):
    """|got-ref| <= rtol * max(|ref|, floor_frac * max|ref|) elementwise."""
    scale = max(np.abs(ref).max(), 1e-30)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), floor_frac * scale)
    return float(err.max())


@pytest.mark.parametrize("dims,org,sp", [((20, 17, 15), (1.0, -2.0, 0.5), (0.5, 1.0, 1.5)),
                                         ((70, 33, 29), (0, 0, 0), (1, 1, 1))])
def test_autograd_harness_vs_torch_and_oracle(dims, org, sp):
    print(f"[case] adjoint small {dims}", flush=True)
    p = fields.gaussian_mixture_params(21, dims, n_gauss=4, sigma=(3.0, 8.0))
    f = fields.gaussian_mixture_numpy(dims, p)
    iso = 0.45
    field = torch.from_numpy(f).cuda().requires_grad_(True)
    verts, nrms = MarchingCubesFunction.apply(field, iso, org, sp)
    U = verts.shape[0]
    assert U > 50
    gen = torch.Generator().manual_seed(1)
    dldx = torch.randn((U, 3), generator=gen, dtype=torch.float32)
    dldn = torch.randn((U, 3), generator=gen, dtype=torch.float32)
    L = (verts * dldx.cuda()).sum() + (nrms * dldn.cuda()).sum()
    L.backward()
    got = field.grad.cpu().numpy().reshape(-1).astype(np.float64)

    r = orc.generate(f, iso, origin=org, spacing=sp)
    assert r["n_pts"] == U
    ref = orc.backward_f64(f, iso, r["first"], r["second"], dldx.numpy(), dldn.numpy(),
                           origin=org, spacing=sp)
    f64 = torch.tensor(f, dtype=torch.float64, requires_grad=True)
    torch_reference_loss(f64, r["first"], r["second"], iso, dldx.numpy(), dldn.numpy(), org,
                         sp).backward()
    tref = f64.grad.numpy().reshape(-1)
    assert np.abs(tref - ref).max() <= 1e-9 * np.abs(ref).max()
    assert np.array_equal(got == 0.0, ref == 0.0) or np.abs(got[ref == 0.0]).max() == 0.0
    assert _close(got, ref, 1e-4, 1e-3) <= 1e-4
    f32 = orc.backward_f32(f, iso, r["first"], r["second"], dldx.numpy(), dldn.numpy(),
                           origin=org, spacing=sp)
    assert _bit_equal(field.grad.cpu().numpy().reshape(-1), f32)


def test_backward_deterministic_and_fd_256():
    """256^3, 16 Gaussians (seed 0), iso 0.5, upstream N(0,1).

    1. Two runs are bitwise identical, and equal bit for bit to the
       oracle's FP32 restatement of the device operation order.
    2. Every point of dL/dfield matches the oracle's FP64 analytic adjoint
       (itself pinned to 1e-9 by float64 torch autograd and by converged
       finite differences) within 1e-4 relative (floor 1e-3 of the max).
    3. FP64 host central differences on 4096 points (half uniform over the
       grid, half drawn from active-edge endpoints, all farther than 1e-3 from
       the isovalue).  The prescribed step 1e-3 is reported but is too coarse
       for this field: its +-0.01 noise makes edge differences |fB-fA| ~ 1e-2,
       so the step-1e-3 truncation error of t = (iso-fA)/(fB-fA) is ~1e-2
       relative.  The assertion uses step 1e-5 on the points where the
       difference quotient has converged (steps 1e-4 and 1e-5 agree to 1e-4).
    """
    n = 256
    dims = (n, n, n)
    p = fields.gaussian_mixture_params(0, dims)
    field = fields.gaussian_mixture_torch(dims, p)
    iso = 0.5
    out = mc.forward(field, iso, edges=True, interp=True)
    U = out.num_points
    print(f"[case] adjoint 256^3 U={U}", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(3)
    dldx = torch.randn((U, 3), generator=gen, device="cuda", dtype=torch.float32)
    dldn = torch.randn((U, 3), generator=gen, device="cuda", dtype=torch.float32)
    g1 = mc.backward(field, iso, out, dldx, dldn)
    g2 = mc.backward(field, iso, out, dldx, dldn)
    torch.cuda.synchronize()
    assert torch.equal(g1, g2)
    f = field.cpu().numpy()
    got = g1.cpu().numpy().reshape(-1).astype(np.float64)
    first = out.interp_first.cpu().numpy()
    second = out.interp_second.cpu().numpy()
    dx, dn = dldx.cpu().numpy(), dldn.cpu().numpy()

    assert _bit_equal(g1.cpu().numpy().reshape(-1), orc.backward_f32(f, iso, first, second, dx, dn))
    ref = orc.backward_f64(f, iso, first, second, dx, dn)
    e_all = _close(got, ref, 1e-4, 1e-3)
    print(f"vs FP64 analytic adjoint: max rel err {e_all:.3e}", flush=True)
    assert e_all <= 1e-4

    rng = np.random.default_rng(5)
    flat = f.reshape(-1)
    ok = np.abs(flat - np.float32(iso)) > 1e-3
    uni = rng.choice(np.flatnonzero(ok), size=2048, replace=False)
    ends = np.unique(np.concatenate([first, second]))
    ends = ends[ok[ends]]
    near = rng.choice(ends, size=2048, replace=False)
    pts = np.concatenate([uni, near])
    fd3 = orc.fd_f64(f, iso, first, second, dx, dn, pts, h=1e-3)
    fd4 = orc.fd_f64(f, iso, first, second, dx, dn, pts, h=1e-4)
    fd5 = orc.fd_f64(f, iso, first, second, dx, dn, pts, h=1e-5)
    scale = np.abs(fd5).max()
    rel = lambda a, b: np.abs(a - b) / np.maximum(np.abs(b), 1e-3 * scale)
    e3 = rel(got[pts], fd3)
    conv = rel(fd4, fd5) <= 1e-4
    e5 = rel(got[pts], fd5)
    print(f"FD step 1e-3: max {e3.max():.3e} median {np.median(e3):.3e} "
          f"frac<=1e-3 {np.mean(e3 <= 1e-3):.3f}; converged points {conv.mean():.3f}, "
          f"max rel err there (step 1e-5) {e5[conv].max():.3e}", flush=True)
    # This replaces the workload's FD check at step 1e-3 (its statistics are
    # only printed above): at that step the truncation error of this noisy
    # field is ~1e-2 relative.  The tolerance stays 1e-3, asserted at step
    # 1e-5 where the difference quotient has converged; the every-point
    # FP64 analytic check above is the stronger one.
    assert conv.mean() >= 0.5
    assert e5[conv].max() <= 1e-3
    # away from the surface the gradient is exactly zero
    assert np.all(got[uni][fd5[:2048] == 0.0] == 0.0)


def test_backward_dense_surface():
    """A field with surface everywhere (sum of short-period sines, 64^3 and a
    ragged 70x45x38): most gather tiles then hold more stencil endpoints than
    their LDS record space and take the direct per-point path, the others the
    shared per-endpoint path.  Checked against the oracle's FP64 adjoint and
    for run-to-run determinism."""
    for dims in ((64, 64, 64), (70, 45, 38)):
        nx, ny, nz = dims
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        f = (np.sin(2 * np.pi * x / 6.0) + np.sin(2 * np.pi * y / 6.5) +
             np.sin(2 * np.pi * z / 7.0)).astype(np.float32)
        iso = 0.3
        print(f"[case] adjoint dense surface {dims}", flush=True)
        field = torch.from_numpy(f).cuda()
        out = mc.forward(field, iso, edges=True)
        U = out.num_points
        # endpoints (points with an active incident edge) per 16384 points
        b = f > iso
        ends = np.zeros_like(b)
        for ax in range(3):
            d = np.diff(b, axis=ax) != 0
            sl_lo = [slice(None)] * 3
            sl_hi = [slice(None)] * 3
            sl_lo[ax] = slice(0, -1)
            sl_hi[ax] = slice(1, None)
            ends[tuple(sl_lo)] |= d
            ends[tuple(sl_hi)] |= d
        assert ends.sum() / f.size * 16384 > 2048
        gen = torch.Generator().manual_seed(7)
        dx = torch.randn((U, 3), generator=gen, dtype=torch.float32)
        dn = torch.randn((U, 3), generator=gen, dtype=torch.float32)
        g1 = mc.backward(field, iso, out, dx.cuda(), dn.cuda())
        g2 = mc.backward(field, iso, out, dx.cuda(), dn.cuda())
        assert torch.equal(g1, g2)
        r = orc.generate(f, iso)
        assert r["n_pts"] == U
        ref = orc.backward_f64(f, iso, r["first"], r["second"], dx.numpy(), dn.numpy())
        assert _bit_equal(g1.cpu().numpy().reshape(-1),
                          orc.backward_f32(f, iso, r["first"], r["second"], dx.numpy(), dn.numpy()))
        got = g1.cpu().numpy().reshape(-1).astype(np.float64)
        assert np.array_equal(got == 0.0, ref == 0.0) or np.abs(got[ref == 0.0]).max() == 0.0
        assert _close(got, ref, 1e-4, 1e-3) <= 1e-4


def _gather_tile_endpoints(f, iso, rows):
    """Endpoint slots of a gather tile of `rows` full rows (one plane) and of
    its 2^lg spans: the endpoints (points with an active incident edge) of
    its rows, of the row before and after it, and of its rows in the planes
    below and above."""
    nz, ny, _ = f.shape
    b = f > iso
    ends = np.zeros_like(b)
    for ax in range(3):
        a = np.moveaxis(b, ax, 0)
        e = np.moveaxis(ends, ax, 0)
        d = a[1:] != a[:-1]
        e[1:] |= d
        e[:-1] |= d
    rc = ends.sum(axis=2)

    def slots(k, j0, j1):
        s = rc[k, j0:j1].sum() + (rc[k, j0 - 1] if j0 > 0 else 0) + (rc[k, j1] if j1 < ny else 0)
        return int(s + (rc[k - 1, j0:j1].sum() if k > 0 else 0) +
                   (rc[k + 1, j0:j1].sum() if k + 1 < nz else 0))

    tiles = []
    for k in range(nz):
        for j0 in range(0, ny, rows):
            n = slots(k, j0, j0 + rows)
            lg = 1
            while lg < 4 and (1536 << lg) < n:
                lg += 1
            sub = rows >> lg
            tiles.append((n, lg, [slots(k, j0 + s * sub, j0 + (s + 1) * sub) for s in range(1 << lg)]))
    return tiles


@pytest.mark.parametrize("tiles_per_cu", [16, 0])
def test_backward_dense_tiles_in_spans(tiles_per_cu):
    """Gather tiles (256 row words: here one 64 x 256 plane) with more
    endpoint slots than their 2048 records.  With few tiles per CU every
    such tile is listed for the span pass; with threshold 0 a tile of up to
    6144 slots runs as 2 or 4 spans in its own block, listing only a span
    that still does not fit, and a denser tile is listed whole.  The field
    (zero, with a noisy strip, a noisy half plane and a box face) holds
    every kind; the adjoint is checked bit for bit against the oracle and
    for determinism."""
    nx, ny, nz = 256, 64, 12
    rng = np.random.default_rng(3)
    f = np.zeros((nz, ny, nx), np.float32)
    f[3, 20:30, :] = rng.integers(0, 2, (10, nx))
    f[5, :, :128] = rng.integers(0, 2, (ny, 128))
    f[8, :, :20] = 1.0
    iso = 0.5
    tiles = _gather_tile_endpoints(f, iso, 64)
    in_block = [t for t in tiles if 2048 < t[0] <= 6144]
    assert any(t[1] == 1 and max(t[2]) <= 2048 for t in in_block)
    assert any(t[1] == 2 and max(t[2]) <= 2048 for t in in_block)
    assert any(max(t[2]) > 2048 for t in in_block)  # a span listed from its block
    assert any(t[0] > 6144 for t in tiles)          # tiles listed whole
    print(f"[case] adjoint dense tiles in spans {(nx, ny, nz)} threshold {tiles_per_cu}",
          flush=True)
    field = torch.from_numpy(f).cuda()
    out = mc.forward(field, iso, edges=True)
    U = out.num_points
    gen = torch.Generator().manual_seed(11)
    dx = torch.randn((U, 3), generator=gen, dtype=torch.float32)
    dn = torch.randn((U, 3), generator=gen, dtype=torch.float32)
    lib = _lib.load()
    assert lib.dax_hip_mc_backward_set_span_threshold(tiles_per_cu) == 0
    try:
        g1 = mc.backward(field, iso, out, dx.cuda(), dn.cuda())
        g2 = mc.backward(field, iso, out, dx.cuda(), dn.cuda())
        torch.cuda.synchronize()
    finally:
        assert lib.dax_hip_mc_backward_set_span_threshold(16) == 0
    assert torch.equal(g1, g2)
    r = orc.generate(f, iso)
    assert r["n_pts"] == U
    assert _bit_equal(g1.cpu().numpy().reshape(-1),
                      orc.backward_f32(f, iso, r["first"], r["second"], dx.numpy(), dn.numpy()))


@pytest.mark.parametrize("dims", [(4000, 3, 4), (15872, 2, 3)])
def test_backward_wide_rows(dims):
    """Rows of 63 and 248 words (the widest the adjoint takes): a gather tile then
    spans few rows and its LDS exceeds 64 KiB.  A smooth field; bit for bit
    against the oracle's FP32 restatement, and against its FP64 adjoint within
    1e-3 (near-tangent crossings, with active-edge differences down to 4e-3,
    put the FP32 rounding itself at up to 3e-4 relative here)."""
    nx, ny, nz = dims
    print(f"[case] adjoint wide rows {dims}", flush=True)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = (0.5 + 0.4 * np.sin(2 * np.pi * x / 23.7) *
         np.cos(2 * np.pi * (y + 1.3 * z) / 4.1)).astype(np.float32)
    iso = 0.5
    field = torch.from_numpy(f).cuda()
    out = mc.forward(field, iso, edges=True)
    U = out.num_points
    r = orc.generate(f, iso)
    assert r["n_pts"] == U and U > 1000
    gen = torch.Generator().manual_seed(9)
    dx = torch.randn((U, 3), generator=gen, dtype=torch.float32)
    dn = torch.randn((U, 3), generator=gen, dtype=torch.float32)
    got = mc.backward(field, iso, out, dx.cuda(), dn.cuda()).cpu().numpy().reshape(-1)
    assert _bit_equal(got, orc.backward_f32(f, iso, r["first"], r["second"], dx.numpy(),
                                            dn.numpy()))
    ref = orc.backward_f64(f, iso, r["first"], r["second"], dx.numpy(), dn.numpy())
    got = got.astype(np.float64)
    assert np.all(got[ref == 0.0] == 0.0)
    assert _close(got, ref, 1e-3, 1e-3) <= 1e-3


def test_backward_rejects_rows_beyond_limit():
    nx = 15872 + 64
    field = torch.rand((2, 2, nx), device="cuda")
    out = mc.forward(field, 0.5, edges=True)
    dx = torch.zeros((out.num_points, 3), device="cuda")
    with pytest.raises(RuntimeError, match="unsupported grid"):
        mc.backward(field, 0.5, out, dx, dx.clone())

"""Test harness only: marching cubes as a torch.autograd.Function whose
forward and backward are the HIP kernels (through the C-ABI), and a float64
torch restatement of the forward for autograd reference gradients."""
from __future__ import annotations

import numpy as np
import torch

from daxtoolkit_amd import mc


class MarchingCubesFunction(torch.autograd.Function):
    """(field) -> (vertices [U,3], normals [U,3]); topology held fixed."""

    @staticmethod
    def forward(ctx, field, iso, origin, spacing):
        out = mc.forward(field.detach(), iso, origin=origin, spacing=spacing, dedup=True,
                         normals=True, edges=True, interp=True)
        ctx.iso, ctx.origin, ctx.spacing = iso, origin, spacing
        ctx.contour = out
        ctx.save_for_backward(field)
        return out.vertices.clone(), out.normals.clone()

    @staticmethod
    def backward(ctx, dl_dv, dl_dn):
        (field,) = ctx.saved_tensors
        c = ctx.contour
        dl_dv = torch.zeros_like(c.vertices) if dl_dv is None else dl_dv.contiguous().float()
        dl_dn = torch.zeros_like(c.normals) if dl_dn is None else dl_dn.contiguous().float()
        g = mc.backward(field.detach(), ctx.iso, c, dl_dv, dl_dn, origin=ctx.origin,
                        spacing=ctx.spacing)
        return g, None, None, None


def torch_reference_loss(field64, first, second, iso, dldx, dldn, origin=(0, 0, 0),
                         spacing=(1, 1, 1)):
    """Float64 restatement of vertices (Lerp of corner coordinates) and
    normals (normalised Lerp of central-difference gradients) for a fixed
    edge list, and L = sum dldx.x + dldn.n, differentiable by autograd."""
    nz, ny, nx = field64.shape
    f = field64.reshape(-1)
    dev = f.device
    first = torch.as_tensor(first, dtype=torch.int64, device=dev)
    second = torch.as_tensor(second, dtype=torch.int64, device=dev)
    fA, fB = f[first], f[second]
    iso = float(np.float32(iso))  # the kernels compare and interpolate with the f32 isovalue
    t = (iso - fA) / (fB - fA)
    org = torch.tensor(origin, dtype=torch.float64, device=dev)
    sp = torch.tensor(spacing, dtype=torch.float64, device=dev)

    def ijk(p):
        return torch.stack([p % nx, (p // nx) % ny, p // (nx * ny)], 1).to(torch.float64)

    cA = org + sp * ijk(first)
    cB = org + sp * ijk(second)
    x = cA + t[:, None] * (cB - cA)
    grads = []
    for axis, n in enumerate((nx, ny, nz)):
        dim = 2 - axis
        h = float(spacing[axis])
        if n < 2:
            grads.append(torch.zeros_like(field64))
            continue
        g = torch.empty_like(field64)
        sl = lambda a, b: tuple(slice(a, b) if d == dim else slice(None) for d in range(3))
        g_int = (field64[sl(2, None)] - field64[sl(None, -2)]) / (2.0 * h)
        g_lo = (field64[sl(1, 2)] - field64[sl(0, 1)]) / h
        g_hi = (field64[sl(n - 1, n)] - field64[sl(n - 2, n - 1)]) / h
        g = torch.cat([g_lo, g_int, g_hi], dim=dim) if n > 2 else torch.cat([g_lo, g_hi], dim=dim)
        grads.append(g)
    G = torch.stack([g.reshape(-1) for g in grads], 1)
    gA, gB = G[first], G[second]
    m = gA + t[:, None] * (gB - gA)
    nrm = m / torch.linalg.norm(m, dim=1, keepdim=True)
    dldx = torch.as_tensor(dldx, dtype=torch.float64, device=dev)
    dldn = torch.as_tensor(dldn, dtype=torch.float64, device=dev)
    return (dldx * x).sum() + (dldn * nrm).sum()

"""Synthetic scalar fields of the benchmark workloads (seeded, reproducible).

All fields are float32 arrays of shape (nz, ny, nx) with point (i, j, k) at
coordinates origin + spacing*(i, j, k).  Parameters are drawn on the host from
numpy's PCG64 ``default_rng(seed)``; the values are computed in float64 and
rounded once to float32.  Device generation (torch) and host generation
(numpy) follow the same formulas; tests that compare two implementations
always feed both the same float32 bits.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def sphere_distance(n: int = 128, center: float = 63.5) -> np.ndarray:
    """|p - (c, c, c)| on an n^3 grid, spacing 1, origin 0 (float64 -> float32)."""
    ax = np.arange(n, dtype=np.float64) - center
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    return np.sqrt(x * x + y * y + z * z).astype(np.float32)


def linear(nx: int, ny: int, nz: int, g=(1.0, 1.0, 1.0)) -> np.ndarray:
    """dot(p, g): the reference's MC test field (x+y+z for g=(1,1,1))."""
    x = np.arange(nx, dtype=np.float64)
    y = np.arange(ny, dtype=np.float64)
    z = np.arange(nz, dtype=np.float64)
    return (g[0] * x[None, None, :] + g[1] * y[None, :, None] + g[2] * z[:, None, None]).astype(
        np.float32)


@dataclass
class GaussianMixture:
    centers: np.ndarray    # [G, 3] (x, y, z) in voxels
    sigmas: np.ndarray     # [G]
    amps: np.ndarray       # [G]
    noise_seed: int
    noise: float = 0.01


def gaussian_mixture_params(seed: int, dims, n_gauss: int = 16, sigma=(20.0, 60.0),
                            amp=(0.5, 1.5), noise: float = 0.01) -> GaussianMixture:
    """Centres uniform in the box [0, n-1] per axis, sigma ~ U[20,60] voxels,
    amplitude ~ U[0.5,1.5], plus uniform noise in [-noise, noise]."""
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    hi = np.array([nx - 1, ny - 1, nz - 1], dtype=np.float64)
    centers = rng.uniform(0.0, 1.0, size=(n_gauss, 3)) * hi
    sigmas = rng.uniform(sigma[0], sigma[1], size=n_gauss)
    amps = rng.uniform(amp[0], amp[1], size=n_gauss)
    noise_seed = int(rng.integers(0, 2**31 - 1))
    return GaussianMixture(centers, sigmas, amps, noise_seed, noise)


def gaussian_mixture_numpy(dims, p: GaussianMixture) -> np.ndarray:
    nx, ny, nz = dims
    x = np.arange(nx, dtype=np.float64)
    y = np.arange(ny, dtype=np.float64)
    z = np.arange(nz, dtype=np.float64)
    f = np.zeros((nz, ny, nx), dtype=np.float64)
    for c, s, a in zip(p.centers, p.sigmas, p.amps):
        k = -0.5 / (s * s)
        ex = np.exp(k * (x - c[0]) ** 2)
        ey = np.exp(k * (y - c[1]) ** 2)
        ez = a * np.exp(k * (z - c[2]) ** 2)
        f += ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
    rng = np.random.default_rng(p.noise_seed)
    f += rng.uniform(-p.noise, p.noise, size=f.shape)
    return f.astype(np.float32)


def gaussian_mixture_torch(dims, p: GaussianMixture, device="cuda"):
    """Same mixture generated on the device (noise from torch's generator)."""
    import torch
    nx, ny, nz = dims
    dev = torch.device(device)
    x = torch.arange(nx, dtype=torch.float64, device=dev)
    y = torch.arange(ny, dtype=torch.float64, device=dev)
    f = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(p.noise_seed)
    exs = [torch.exp((-0.5 / (s * s)) * (x - c[0]) ** 2) for c, s in zip(p.centers, p.sigmas)]
    eys = [torch.exp((-0.5 / (s * s)) * (y - c[1]) ** 2) for c, s in zip(p.centers, p.sigmas)]
    chunk = max(1, (1 << 26) // (nx * ny))
    for z0 in range(0, nz, chunk):
        z1 = min(nz, z0 + chunk)
        z = torch.arange(z0, z1, dtype=torch.float64, device=dev)
        acc = torch.zeros((z1 - z0, ny, nx), dtype=torch.float64, device=dev)
        for c, s, a, ex, ey in zip(p.centers, p.sigmas, p.amps, exs, eys):
            ez = a * torch.exp((-0.5 / (s * s)) * (z - c[2]) ** 2)
            acc += ez[:, None, None] * (ey[:, None] * ex[None, :])[None]
        noise = torch.rand((z1 - z0, ny, nx), dtype=torch.float64, device=dev, generator=gen)
        acc += (2.0 * noise - 1.0) * p.noise
        f[z0:z1] = acc.to(torch.float32)
    return f


def mandelbulb_torch(n: int, power: float = 8.0, max_iter: int = 10, bailout: float = 2.0,
                     extent: float = 1.2, device="cuda"):
    """Escape-time Mandelbulb (White-Nylander, z0 = c) over [-extent, extent]^3.

    z <- z^power + c in spherical form (r^p, p*theta, p*phi with
    theta = acos(z_z/r), phi = atan2(z_y, z_x)); the value is the number of
    iterations done while |z| <= bailout, at most max_iter, divided by
    max_iter.  Point (i, j, k) is c = -extent + 2*extent*(i, j, k)/(n-1).
    """
    import torch
    dev = torch.device(device)
    out = torch.empty((n, n, n), dtype=torch.float32, device=dev)
    ax = torch.linspace(-extent, extent, n, dtype=torch.float64, device=dev)
    chunk = max(1, (1 << 24) // (n * n))
    for z0 in range(0, n, chunk):
        z1 = min(n, z0 + chunk)
        cz, cy, cx = torch.meshgrid(ax[z0:z1], ax, ax, indexing="ij")
        zx, zy, zz = cx.clone(), cy.clone(), cz.clone()
        it = torch.zeros_like(cx)
        alive = torch.ones_like(cx, dtype=torch.bool)
        for _ in range(max_iter):
            r = torch.sqrt(zx * zx + zy * zy + zz * zz)
            alive = alive & (r <= bailout)
            it += alive.to(torch.float64)
            theta = torch.acos(torch.clamp(zz / torch.clamp(r, min=1e-300), -1.0, 1.0))
            phi = torch.atan2(zy, zx)
            rp = r ** power
            nzx = rp * torch.sin(power * theta) * torch.cos(power * phi) + cx
            nzy = rp * torch.sin(power * theta) * torch.sin(power * phi) + cy
            nzz = rp * torch.cos(power * theta) + cz
            zx = torch.where(alive, nzx, zx)
            zy = torch.where(alive, nzy, zy)
            zz = torch.where(alive, nzz, zz)
        out[z0:z1] = (it / max_iter).to(torch.float32)
    return out


@dataclass
class VolumeBatch:
    dims: list            # [(nx, ny, nz)]
    isos: list            # per-volume isovalue
    params: list          # per-volume GaussianMixture


def volume_batch(n: int, lo: int = 32, hi: int = 160, seed: int = 0, gauss=(4, 16),
                 sigma=(20.0, 60.0), iso=(0.3, 0.7)) -> VolumeBatch:
    """n independent volumes: per-axis point extents ~ U{lo..hi}, each a
    seeded mixture of U{gauss} Gaussians (sigma ~ U[sigma] voxels, amplitude
    ~ U[0.5,1.5], noise U(-0.01,0.01)), isovalue ~ U[iso]."""
    rng = np.random.default_rng(seed)
    dims, isos, params = [], [], []
    for v in range(n):
        d = tuple(int(x) for x in rng.integers(lo, hi + 1, size=3))
        g = int(rng.integers(gauss[0], gauss[1] + 1))
        dims.append(d)
        isos.append(float(np.float32(rng.uniform(iso[0], iso[1]))))
        params.append(gaussian_mixture_params(int(rng.integers(0, 2**31 - 1)), d, n_gauss=g,
                                              sigma=sigma))
    return VolumeBatch(dims, isos, params)


def volume_batch_torch(b: VolumeBatch, device="cuda"):
    """The batch's fields back to back in one flat float32 device tensor."""
    import torch
    sizes = [d[0] * d[1] * d[2] for d in b.dims]
    flat = torch.empty(sum(sizes), dtype=torch.float32, device=device)
    off = 0
    for d, p, s in zip(b.dims, b.params, sizes):
        flat[off:off + s] = gaussian_mixture_torch(d, p, device).reshape(-1)
        off += s
    return flat


# Vertex lists of the cells a uniform grid's voxel is split into, in the
# voxel's corner order (0..7 = (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1)
# (1,0,1) (1,1,1) (0,1,1)), as the reference's test grids build them
# (dax/cont/testing/TestingGridGenerator.h:229-257 (corners), :266-295 (one
# hexahedron per voxel), :300-339 (6 tetrahedra per voxel, Freudenthal)).
VOXEL_SPLITS = {
    "hexahedron": [[0, 1, 2, 3, 4, 5, 6, 7]],
    "tetrahedron": [[0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6], [0, 5, 6, 4], [1, 5, 6, 0],
                    [1, 6, 2, 0]],
}


def voxel_cells_torch(dims, shape: str, origin=(0, 0, 0), spacing=(1, 1, 1), device="cuda"):
    """An unstructured grid of the voxels of the uniform grid `dims` (points
    x-fastest, id i + nx*(j + ny*k)): cells [C, V] int32 (voxels x-fastest,
    each voxel's cells in VOXEL_SPLITS order) and points [P, 3] float32 at
    origin + spacing*(i, j, k) (one float32 multiply and add per axis)."""
    import torch
    nx, ny, nz = (int(v) for v in dims)
    split = torch.tensor(VOXEL_SPLITS[shape], dtype=torch.int64, device=device)   # [S, V]
    corner = torch.tensor([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1],
                           [1, 1, 1], [0, 1, 1]], dtype=torch.int64, device=device)
    coff = corner[:, 0] + nx * (corner[:, 1] + ny * corner[:, 2])                  # [8]
    k, j, i = torch.meshgrid(torch.arange(nz - 1, device=device),
                             torch.arange(ny - 1, device=device),
                             torch.arange(nx - 1, device=device), indexing="ij")
    base = (i + nx * (j + ny * k)).reshape(-1)                                     # [C]
    cells = (base[:, None, None] + coff[split][None]).reshape(-1, split.shape[1])
    k, j, i = torch.meshgrid(torch.arange(nz, device=device), torch.arange(ny, device=device),
                             torch.arange(nx, device=device), indexing="ij")
    o = torch.tensor(origin, dtype=torch.float32, device=device)
    s = torch.tensor(spacing, dtype=torch.float32, device=device)
    pts = torch.stack([o[0] + s[0] * i.reshape(-1).float(), o[1] + s[1] * j.reshape(-1).float(),
                       o[2] + s[2] * k.reshape(-1).float()], dim=1).contiguous()
    return cells.to(torch.int32).contiguous(), pts

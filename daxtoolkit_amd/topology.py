"""Threshold and topology generation on the GPU (the reference's
ThresholdCount / ThresholdTopology, dax/worklet/Threshold.h:30-110, through
DispatcherMapCell and DispatcherGenerateTopology,
dax/cont/DispatcherGenerateTopology.h:146-256).

A uniform grid is given by its field of shape (nz, ny, nx) (and origin,
spacing); an unstructured grid by `cells` [C, V] int32 point ids and `points`
[P, 3] float32.  Every call is synchronous where the reference reads a size
back (the output cell and point counts)."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import torch

from . import _lib
from ._lib import MC_DEDUP, TOPOLOGY_MASK_OPTIONAL


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


@dataclass
class Topology:
    """Output of DispatcherGenerateTopology: connectivity [T, V] (ids of the
    input's points, or with duplicate removal their ranks among the used
    points), the used points' coordinates [U, 3] and the point mask [P]
    (duplicate removal only)."""
    counts: torch.Tensor
    connectivity: torch.Tensor
    points: Optional[torch.Tensor]
    point_mask: Optional[torch.Tensor]
    workspace: Optional[torch.Tensor] = None   # the prepare step's (ranks of the used points)
    num_input_cells: int = 0
    _mask_buffer: Optional[torch.Tensor] = None  # what prepare was given for the mask

    @property
    def num_cells(self) -> int:
        return int(self.connectivity.shape[0])

    @property
    def num_points(self) -> int:
        return 0 if self.points is None else int(self.points.shape[0])

    def compact_point_field(self, values: torch.Tensor) -> torch.Tensor:
        """CompactPointField (:97-110): the used points' values in point
        order (a StreamCompact by the point mask); needs duplicate removal."""
        if self._mask_buffer is None:
            raise ValueError("compact_point_field needs duplicate removal")
        P = self._mask_buffer.numel()
        v = values.reshape(P, -1).contiguous()
        if v.element_size() * v.shape[1] % 4:
            raise ValueError("point values of a multiple of 4 bytes")
        out = torch.empty((max(self.num_points, 1), v.shape[1]), dtype=v.dtype, device=v.device)
        _lib.call("dax_hip_topology_compact_point_field", v.data_ptr(),
                  v.element_size() * v.shape[1], self.num_input_cells, P,
                  self._mask_buffer.data_ptr(), self.workspace.data_ptr(), self.workspace.numel(),
                  out.data_ptr(), _stream())
        res = out[: self.num_points]
        return res.reshape(-1) if values.dim() == 1 else res.reshape((self.num_points,) +
                                                                     tuple(values.shape[1:]))


def _check_cells(cells: torch.Tensor, n_points: int):
    if (cells.dtype != torch.int32 or cells.dim() != 2 or not 1 <= cells.shape[1] <= 8 or
            not cells.is_contiguous()):
        raise ValueError("cells must be a contiguous int32 [C, V] tensor with 1 <= V <= 8")
    if n_points < 1:
        raise ValueError("an unstructured grid needs points")


def threshold_count(field: torch.Tensor, lo: float, hi: float,
                    cells: Optional[torch.Tensor] = None) -> torch.Tensor:
    """ThresholdCount through DispatcherMapCell: 1 per cell whose point
    values all lie in [lo, hi], else 0 (int32 [C])."""
    if field.dtype != torch.float32 or not field.is_contiguous():
        raise ValueError("field must be a contiguous float32 tensor")
    if cells is None:
        if field.dim() != 3:
            raise ValueError("a uniform grid's field has shape (nz, ny, nx)")
        nz, ny, nx = field.shape
        counts = torch.empty(max((nx - 1) * (ny - 1) * (nz - 1), 1), dtype=torch.int32,
                             device=field.device)
        _lib.call("dax_hip_threshold_count_uniform", field.data_ptr(), nx, ny, nz, float(lo),
                  float(hi), counts.data_ptr(), _stream())
        return counts
    _check_cells(cells, field.numel())
    C, V = cells.shape
    counts = torch.empty(max(C, 1), dtype=torch.int32, device=field.device)
    _lib.call("dax_hip_threshold_count_cells", cells.data_ptr(), V, C, field.data_ptr(),
              field.numel(), float(lo), float(hi), counts.data_ptr(), _stream())
    return counts[:C]


def _generate(counts, dims, cells, points, origin, spacing, dedup, ws, mask, T, U,
              mask_valid=True) -> Topology:
    dev = counts.device
    if cells is None:
        nx, ny, nz = (int(v) for v in dims)
        C, P, V = (nx - 1) * (ny - 1) * (nz - 1), nx * ny * nz, 8
    else:
        C, V = cells.shape
        P = mask.numel() if mask is not None else int(points.shape[0])
    flags = MC_DEDUP if dedup else 0
    mptr = mask.data_ptr() if dedup else None
    conn = torch.empty((max(T, 1), V), dtype=torch.int32, device=dev)
    coords = torch.empty((max(U, 1), 3), dtype=torch.float32, device=dev) if dedup else None
    cptr = coords.data_ptr() if dedup else None
    if T > 0:
        if cells is None:
            _lib.call("dax_hip_topology_generate_uniform", nx, ny, nz, _f3(origin), _f3(spacing),
                      counts.data_ptr(), flags, ws.data_ptr(), ws.numel(), mptr,
                      conn.data_ptr(), cptr, _stream())
        else:
            _lib.call("dax_hip_topology_generate_cells", cells.data_ptr(), V, C,
                      points.data_ptr() if points is not None else None, P, counts.data_ptr(),
                      flags, ws.data_ptr(), ws.numel(), mptr, conn.data_ptr(), cptr, _stream())
    return Topology(counts=counts[:C], connectivity=conn[:T],
                    points=coords[:U] if dedup else None,
                    point_mask=mask if (dedup and mask_valid) else None,
                    workspace=ws, num_input_cells=C, _mask_buffer=mask)


def generate_topology(counts: torch.Tensor, dims=None, cells: Optional[torch.Tensor] = None,
                      points: Optional[torch.Tensor] = None, origin=(0, 0, 0), spacing=(1, 1, 1),
                      dedup: bool = True) -> Topology:
    """DispatcherGenerateTopology with ThresholdTopology: each cell repeated
    counts[c] times with its vertices (uniform grids: the voxel's corners in
    the hexahedron's order)."""
    if counts.dtype != torch.int32 or not counts.is_contiguous():
        raise ValueError("counts must be a contiguous int32 tensor")
    dev = counts.device
    lib = _lib.load()
    t, u = ctypes.c_int64(), ctypes.c_int64()
    flags = MC_DEDUP if dedup else 0
    if cells is None:
        nx, ny, nz = (int(v) for v in dims)
        C, P, V = (nx - 1) * (ny - 1) * (nz - 1), nx * ny * nz, 8
    else:
        P = int(points.shape[0]) if points is not None else int(cells.max()) + 1
        _check_cells(cells, P)
        if points is not None and (points.dtype != torch.float32 or points.shape != (P, 3) or
                                   not points.is_contiguous()):
            raise ValueError("points must be a contiguous float32 [P, 3] tensor")
        C, V = cells.shape
    if counts.numel() < C:
        raise ValueError("one count per cell")
    ws = torch.empty(int(lib.dax_hip_topology_workspace_bytes(C, P)), dtype=torch.uint8,
                     device=dev)
    mask = torch.empty(P, dtype=torch.int32, device=dev) if dedup else None
    mptr = mask.data_ptr() if dedup else None
    if cells is None:
        _lib.call("dax_hip_topology_prepare_uniform", nx, ny, nz, counts.data_ptr(), flags,
                  ws.data_ptr(), ws.numel(), mptr, ctypes.byref(t), ctypes.byref(u), _stream())
    else:
        _lib.call("dax_hip_topology_prepare_cells", cells.data_ptr(), V, C, P, counts.data_ptr(),
                  flags, ws.data_ptr(), ws.numel(), mptr, ctypes.byref(t), ctypes.byref(u),
                  _stream())
    return _generate(counts, dims, cells, points, origin, spacing, dedup, ws, mask,
                     int(t.value), int(u.value))


def threshold(field: torch.Tensor, lo: float, hi: float, cells: Optional[torch.Tensor] = None,
              points: Optional[torch.Tensor] = None, origin=(0, 0, 0), spacing=(1, 1, 1),
              dedup: bool = True, point_mask: bool = True) -> Topology:
    """The reference's Threshold pipeline (Benchmarks/Threshold/Pipeline.h:
    163-190): ThresholdCount, then DispatcherGenerateTopology.  On a uniform
    grid both run in one call (dax_hip_threshold_prepare_uniform); with
    point_mask=False the int32 point mask is not kept (CompactPointField
    does not need it)."""
    if cells is not None or field.dim() != 3:
        counts = threshold_count(field, lo, hi, cells)
        return generate_topology(counts, cells=cells, points=points, origin=origin,
                                 spacing=spacing, dedup=dedup)
    if field.dtype != torch.float32 or not field.is_contiguous():
        raise ValueError("field must be a contiguous float32 tensor")
    nz, ny, nx = field.shape
    C, P = (nx - 1) * (ny - 1) * (nz - 1), nx * ny * nz
    dev = field.device
    counts = torch.empty(max(C, 1), dtype=torch.int32, device=dev)
    ws = torch.empty(int(_lib.load().dax_hip_topology_workspace_bytes(C, P)), dtype=torch.uint8,
                     device=dev)
    mask = torch.empty(P, dtype=torch.int32, device=dev) if dedup else None
    flags = (MC_DEDUP if dedup else 0) | (0 if point_mask else TOPOLOGY_MASK_OPTIONAL)
    t, u = ctypes.c_int64(), ctypes.c_int64()
    _lib.call("dax_hip_threshold_prepare_uniform", field.data_ptr(), nx, ny, nz, float(lo),
              float(hi), flags, counts.data_ptr(), ws.data_ptr(), ws.numel(),
              mask.data_ptr() if dedup else None, ctypes.byref(t), ctypes.byref(u), _stream())
    return _generate(counts, (nx, ny, nz), None, None, origin, spacing, dedup, ws, mask,
                     int(t.value), int(u.value), mask_valid=point_mask)

"""Python front end of the marching-cubes C-ABI over torch device tensors.

torch only provides device memory and the stream; all computation is in
libdaxhip.so.  Field tensors are float32, C-contiguous, shape (nz, ny, nx) on
a HIP device, so the flat point index is i + nx*(j + ny*k) as in Dax
(dax/Extent.h:104-112).

The call sequence mirrors the reference's two dispatchers
(Benchmarks/MarchingCubes/Pipeline.h:115-130):
``DispatcherMapCell<MarchingCubesCount>`` -> :func:`classify` and
``DispatcherGenerateInterpolatedCells<MarchingCubesGenerate>`` ->
:func:`forward`.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import torch

from . import _lib
from ._lib import MC_DEDUP, MC_NORMALS


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _f3(v: Sequence[float]):
    import ctypes
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def _check_field(field: torch.Tensor) -> tuple[int, int, int]:
    if field.dim() != 3:
        raise ValueError("field must have shape (nz, ny, nx)")
    if field.dtype != torch.float32 or not field.is_cuda or not field.is_contiguous():
        raise ValueError("field must be a contiguous float32 tensor on a HIP device")
    nz, ny, nx = field.shape
    return int(nx), int(ny), int(nz)


def check_tensor(t: Optional[torch.Tensor], name: str, dtype: torch.dtype, min_numel: int,
                 device=None, exact: bool = False) -> None:
    """Reject a tensor the kernels would index out of bounds or misread:
    wrong dtype, not contiguous, fewer than `min_numel` elements (exactly
    `min_numel` when `exact`), or not on `device` (a HIP device when None).
    The kernels take raw pointers and index them by grid or contour ids, so
    these checks run before every call."""
    if t is None:
        raise ValueError(f"{name} is required")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    n = t.numel()
    if (n != min_numel) if exact else (n < min_numel):
        rel = "exactly" if exact else "at least"
        raise ValueError(f"{name} must hold {rel} {min_numel} elements (got {n})")
    if device is not None:
        if t.device != torch.device(device):
            raise ValueError(f"{name} must be on {device} (got {t.device})")
    elif not t.is_cuda:
        raise ValueError(f"{name} must be on a HIP device (got {t.device})")


def workspace_bytes(nx: int, ny: int, nz: int) -> int:
    return int(_lib.load().dax_hip_mc_workspace_bytes(nx, ny, nz))


def new_workspace(nx: int, ny: int, nz: int, device=None) -> torch.Tensor:
    return torch.empty(workspace_bytes(nx, ny, nz), dtype=torch.uint8, device=device or "cuda")


@dataclass
class Contour:
    """Output of one forward: an UnstructuredGrid<CellTagTriangle> plus extras."""
    vertices: torch.Tensor            # [U, 3] f32
    connectivity: torch.Tensor        # [T, 3] i32
    normals: Optional[torch.Tensor]   # [U, 3] f32
    edges: Optional[torch.Tensor]     # [U] u32 as int32 bits: (axis << 30) | lower point id
    interp_first: Optional[torch.Tensor]
    interp_second: Optional[torch.Tensor]
    interp_weight: Optional[torch.Tensor]
    workspace: Optional[torch.Tensor] = None   # inside-bits + vertex bases (for backward)
    counts: Optional[torch.Tensor] = None      # device {T, U}
    grid_points: Optional[int] = None          # points of the input grid (interp ids index it)

    @property
    def num_triangles(self) -> int:
        return int(self.connectivity.shape[0])

    @property
    def num_points(self) -> int:
        return int(self.vertices.shape[0])


def classify(field: torch.Tensor, iso: float, workspace: Optional[torch.Tensor] = None,
             counts: Optional[torch.Tensor] = None):
    """Classify + count/scan.  Returns (workspace, counts[2] = {T, U})."""
    nx, ny, nz = _check_field(field)
    ws = workspace if workspace is not None else new_workspace(nx, ny, nz, field.device)
    cnt = counts if counts is not None else torch.empty(2, dtype=torch.int64, device=field.device)
    _lib.call("dax_hip_mc_classify", field.data_ptr(), nx, ny, nz, float(iso), ws.data_ptr(),
              ws.numel(), cnt.data_ptr(), _stream())
    return ws, cnt


def read_counts(counts: torch.Tensor, cap_points: int = 2**62, cap_tris: int = 2**62):
    import ctypes
    t, u = ctypes.c_int64(), ctypes.c_int64()
    _lib.call("dax_hip_mc_read_counts", counts.data_ptr(), cap_points, cap_tris,
              ctypes.byref(t), ctypes.byref(u), _stream())
    return int(t.value), int(u.value)


def cell_outputs(field: torch.Tensor, workspace: torch.Tensor):
    """Per-cell case (u8), NumFaces (i32) and exclusive triangle offset (i32)."""
    nx, ny, nz = _check_field(field)
    check_tensor(workspace, "workspace", torch.uint8, workspace_bytes(nx, ny, nz), field.device)
    C = max(nx - 1, 0) * max(ny - 1, 0) * max(nz - 1, 0)
    dev = field.device
    cases = torch.empty(C, dtype=torch.uint8, device=dev)
    count = torch.empty(C, dtype=torch.int32, device=dev)
    offs = torch.empty(C, dtype=torch.int32, device=dev)
    _lib.call("dax_hip_mc_cell_outputs", nx, ny, nz, workspace.data_ptr(), workspace.numel(),
              cases.data_ptr(), count.data_ptr(), offs.data_ptr(), _stream())
    return cases, count, offs


class ForwardBuffers:
    """Preallocated outputs of :func:`forward_into` (for timed, graph-friendly runs)."""

    def __init__(self, nx, ny, nz, cap_points, cap_tris, device="cuda", normals=True,
                 edges=False, interp=False):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.cap_points, self.cap_tris = int(cap_points), int(cap_tris)
        self.workspace = new_workspace(nx, ny, nz, device)
        self.vertices = torch.empty((self.cap_points, 3), dtype=torch.float32, device=device)
        self.normals = (torch.empty((self.cap_points, 3), dtype=torch.float32, device=device)
                        if normals else None)
        self.conn = torch.empty((self.cap_tris, 3), dtype=torch.int32, device=device)
        self.edges = torch.empty(self.cap_points, dtype=torch.int32, device=device) if edges else None
        if interp:
            self.interp_first = torch.empty(self.cap_points, dtype=torch.int32, device=device)
            self.interp_second = torch.empty(self.cap_points, dtype=torch.int32, device=device)
            self.interp_weight = torch.empty(self.cap_points, dtype=torch.float32, device=device)
        else:
            self.interp_first = self.interp_second = self.interp_weight = None
        self.counts = torch.zeros(2, dtype=torch.int64, device=device)


def _check_buf(buf: "ForwardBuffers", dims: tuple[int, int, int], device) -> None:
    """The buffers were made for this grid and live on the field's device."""
    if (buf.nx, buf.ny, buf.nz) != tuple(dims):
        raise ValueError(f"buffers were made for a {buf.nx}x{buf.ny}x{buf.nz} grid, "
                         f"not {dims[0]}x{dims[1]}x{dims[2]}")
    check_tensor(buf.workspace, "workspace", torch.uint8, workspace_bytes(*dims), device)
    check_tensor(buf.vertices, "vertices", torch.float32, 3 * buf.cap_points, device)
    check_tensor(buf.conn, "connectivity", torch.int32, 3 * buf.cap_tris, device)
    check_tensor(buf.counts, "counts", torch.int64, 2, device)


def forward_into(field: torch.Tensor, iso: float, buf: ForwardBuffers, origin=(0, 0, 0),
                 spacing=(1, 1, 1), dedup: bool = True, normals: bool = True) -> None:
    """Enqueue one forward into preallocated buffers (no host synchronisation)."""
    nx, ny, nz = _check_field(field)
    _check_buf(buf, (nx, ny, nz), field.device)
    flags = (MC_DEDUP if dedup else 0) | (MC_NORMALS if (normals and buf.normals is not None) else 0)
    _lib.call("dax_hip_mc_forward", field.data_ptr(), nx, ny, nz, _f3(origin), _f3(spacing),
              float(iso), flags, buf.workspace.data_ptr(), buf.workspace.numel(),
              buf.vertices.data_ptr(), _ptr(buf.normals), buf.conn.data_ptr(),
              _ptr(buf.edges) if dedup else None, _ptr(buf.interp_first),
              _ptr(buf.interp_second), _ptr(buf.interp_weight), buf.cap_points, buf.cap_tris,
              buf.counts.data_ptr(), _stream())


def classify_into(field: torch.Tensor, iso: float, buf: ForwardBuffers) -> None:
    """The forward's first kernel alone: inside-bits into buf.workspace (and
    the count/scan state cleared)."""
    nx, ny, nz = _check_field(field)
    _check_buf(buf, (nx, ny, nz), field.device)
    _lib.call("dax_hip_mc_bits", field.data_ptr(), nx, ny, nz, float(iso), buf.workspace.data_ptr(),
              buf.workspace.numel(), _stream())


def count_scan_into(buf: ForwardBuffers) -> None:
    """The forward's second kernel alone (after :func:`classify_into`)."""
    _lib.call("dax_hip_mc_scan", buf.nx, buf.ny, buf.nz, buf.workspace.data_ptr(),
              buf.workspace.numel(), _stream())


def generate_into(field: torch.Tensor, iso: float, buf: ForwardBuffers, origin=(0, 0, 0),
                  spacing=(1, 1, 1), dedup: bool = True, normals: bool = True) -> None:
    """The forward's third kernel alone (after classify and count/scan)."""
    nx, ny, nz = _check_field(field)
    _check_buf(buf, (nx, ny, nz), field.device)
    flags = (MC_DEDUP if dedup else 0) | (MC_NORMALS if (normals and buf.normals is not None) else 0)
    _lib.call("dax_hip_mc_generate", field.data_ptr(), nx, ny, nz, _f3(origin), _f3(spacing),
              float(iso), flags, buf.workspace.data_ptr(), buf.workspace.numel(),
              buf.vertices.data_ptr(), _ptr(buf.normals), buf.conn.data_ptr(),
              _ptr(buf.edges) if dedup else None, _ptr(buf.interp_first),
              _ptr(buf.interp_second), _ptr(buf.interp_weight), buf.cap_points, buf.cap_tris,
              buf.counts.data_ptr(), _stream())


def forward_phases(field: torch.Tensor, iso: float, buf: ForwardBuffers, origin=(0, 0, 0),
                   spacing=(1, 1, 1), dedup: bool = True, normals: bool = True,
                   events=None) -> None:
    """The same forward as three calls (one kernel each); if `events` (four
    torch.cuda.Event) is given they are recorded before, between and after
    the kernels on the current stream."""
    nx, ny, nz = _check_field(field)
    _check_buf(buf, (nx, ny, nz), field.device)
    flags = (MC_DEDUP if dedup else 0) | (MC_NORMALS if (normals and buf.normals is not None) else 0)
    ws = buf.workspace
    s = _stream()
    if events:
        events[0].record()
    _lib.call("dax_hip_mc_bits", field.data_ptr(), nx, ny, nz, float(iso), ws.data_ptr(),
              ws.numel(), s)
    if events:
        events[1].record()
    _lib.call("dax_hip_mc_scan", nx, ny, nz, ws.data_ptr(), ws.numel(), s)
    if events:
        events[2].record()
    _lib.call("dax_hip_mc_generate", field.data_ptr(), nx, ny, nz, _f3(origin), _f3(spacing),
              float(iso), flags, ws.data_ptr(), ws.numel(), buf.vertices.data_ptr(),
              _ptr(buf.normals), buf.conn.data_ptr(), _ptr(buf.edges) if dedup else None,
              _ptr(buf.interp_first), _ptr(buf.interp_second), _ptr(buf.interp_weight),
              buf.cap_points, buf.cap_tris, buf.counts.data_ptr(), s)
    if events:
        events[3].record()


def forward(field: torch.Tensor, iso: float, origin=(0, 0, 0), spacing=(1, 1, 1),
            dedup: bool = True, normals: bool = True, edges: bool = False,
            interp: bool = False) -> Contour:
    """Classify, size the outputs, then run the full forward."""
    nx, ny, nz = _check_field(field)
    ws, cnt = classify(field, iso)
    T, U = read_counts(cnt)
    if not dedup:
        U = 3 * T
    buf = ForwardBuffers(nx, ny, nz, max(U, 1), max(T, 1), field.device, normals=normals,
                         edges=edges and dedup, interp=interp)
    forward_into(field, iso, buf, origin, spacing, dedup, normals)
    T2, U2 = read_counts(buf.counts, buf.cap_points, buf.cap_tris)
    assert (T2, U2) == (T, U), (T2, U2, T, U)
    return Contour(
        vertices=buf.vertices[:U], connectivity=buf.conn[:T],
        normals=None if buf.normals is None else buf.normals[:U],
        edges=None if buf.edges is None else buf.edges[:U],
        interp_first=None if buf.interp_first is None else buf.interp_first[:U],
        interp_second=None if buf.interp_second is None else buf.interp_second[:U],
        interp_weight=None if buf.interp_weight is None else buf.interp_weight[:U],
        workspace=buf.workspace, counts=buf.counts, grid_points=nx * ny * nz)


def slice_into(dims, plane_origin, plane_normal, buf: ForwardBuffers, origin=(0, 0, 0),
               spacing=(1, 1, 1), dedup: bool = True, normals: bool = False) -> None:
    """Enqueue one slice (the marching-cubes surface of the plane distance the
    reference's SliceCount / SliceGenerate classify, dax/worklet/Slice.h:28-173,
    computed in the kernels) into preallocated buffers; dims = (nx, ny, nz)."""
    nx, ny, nz = (int(v) for v in dims)
    flags = (MC_DEDUP if dedup else 0) | (MC_NORMALS if (normals and buf.normals is not None) else 0)
    _lib.call("dax_hip_mc_slice", nx, ny, nz, _f3(origin), _f3(spacing), _f3(plane_origin),
              _f3(plane_normal), flags, buf.workspace.data_ptr(), buf.workspace.numel(),
              buf.vertices.data_ptr(), _ptr(buf.normals), buf.conn.data_ptr(),
              _ptr(buf.edges) if dedup else None, _ptr(buf.interp_first),
              _ptr(buf.interp_second), _ptr(buf.interp_weight), buf.cap_points, buf.cap_tris,
              buf.counts.data_ptr(), _stream())


def slice(dims, plane_origin, plane_normal, origin=(0, 0, 0), spacing=(1, 1, 1),
          dedup: bool = True, normals: bool = False, interp: bool = False,
          device="cuda") -> Contour:
    """Slice of the uniform grid dims = (nx, ny, nz) by the plane through
    plane_origin with normal plane_normal, as Dax's SliceCount then
    SliceGenerate (dax/worklet/Slice.h:28-173) produce it: count, size the
    outputs, generate.  No field is read."""
    nx, ny, nz = (int(v) for v in dims)
    ws = new_workspace(nx, ny, nz, device)
    cnt = torch.empty(2, dtype=torch.int64, device=device)
    _lib.call("dax_hip_mc_slice_classify", nx, ny, nz, _f3(origin), _f3(spacing),
              _f3(plane_origin), _f3(plane_normal), ws.data_ptr(), ws.numel(), cnt.data_ptr(),
              _stream())
    T, U = read_counts(cnt)
    if not dedup:
        U = 3 * T
    buf = ForwardBuffers(nx, ny, nz, max(U, 1), max(T, 1), device, normals=normals,
                         edges=dedup, interp=interp)
    slice_into((nx, ny, nz), plane_origin, plane_normal, buf, origin, spacing, dedup, normals)
    T2, U2 = read_counts(buf.counts, buf.cap_points, buf.cap_tris)
    assert (T2, U2) == (T, U), (T2, U2, T, U)
    return Contour(
        vertices=buf.vertices[:U], connectivity=buf.conn[:T],
        normals=None if buf.normals is None else buf.normals[:U],
        edges=None if buf.edges is None else buf.edges[:U],
        interp_first=None if buf.interp_first is None else buf.interp_first[:U],
        interp_second=None if buf.interp_second is None else buf.interp_second[:U],
        interp_weight=None if buf.interp_weight is None else buf.interp_weight[:U],
        workspace=buf.workspace, counts=buf.counts, grid_points=nx * ny * nz)


def _cells_contour(prefix: str, nv: int, cells: torch.Tensor, field: torch.Tensor,
                   points: torch.Tensor, iso: float, dedup: bool) -> Contour:
    import ctypes
    if (cells.dtype != torch.int32 or cells.dim() != 2 or cells.shape[1] != nv or
            not cells.is_contiguous()):
        raise ValueError(f"cells must be a contiguous int32 [C, {nv}] tensor")
    if field.dtype != torch.float32 or not field.is_contiguous():
        raise ValueError("field must be a contiguous float32 tensor")
    P = field.numel()
    if points.dtype != torch.float32 or points.shape != (P, 3) or not points.is_contiguous():
        raise ValueError("points must be a contiguous float32 [P, 3] tensor")
    dev = field.device
    C = int(cells.shape[0])
    counts = torch.empty(max(C, 1), dtype=torch.int32, device=dev)
    offsets = torch.empty(max(C, 1), dtype=torch.int32, device=dev)
    t = ctypes.c_int64()
    _lib.call(f"{prefix}_count", cells.data_ptr(), C, field.data_ptr(), P, float(iso),
              counts.data_ptr(), offsets.data_ptr(), ctypes.byref(t), _stream())
    T = int(t.value)
    n = max(3 * T, 1)
    verts = torch.empty((n, 3), dtype=torch.float32, device=dev)
    conn = torch.empty(n, dtype=torch.int32, device=dev)
    first = torch.empty(n, dtype=torch.int32, device=dev)
    second = torch.empty(n, dtype=torch.int32, device=dev)
    weight = torch.empty(n, dtype=torch.float32, device=dev)
    nbytes = int(getattr(_lib.load(), f"{prefix}_scratch_bytes")(T)) if dedup else 0
    scratch = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
    u = ctypes.c_int64()
    _lib.call(f"{prefix}_generate", cells.data_ptr(), C, field.data_ptr(), points.data_ptr(), P,
              float(iso), MC_DEDUP if dedup else 0, counts.data_ptr(), offsets.data_ptr(), T,
              scratch.data_ptr(), scratch.numel(), verts.data_ptr(), conn.data_ptr(),
              first.data_ptr(), second.data_ptr(), weight.data_ptr(), ctypes.byref(u), _stream())
    U = int(u.value)
    return Contour(vertices=verts[:U], connectivity=conn[:3 * T].view(T, 3), normals=None,
                   edges=None, interp_first=first[:U], interp_second=second[:U],
                   interp_weight=weight[:U], grid_points=P)


def tetrahedra(tets: torch.Tensor, field: torch.Tensor, points: torch.Tensor, iso: float,
               dedup: bool = True) -> Contour:
    """Marching tetrahedra on an unstructured tetrahedral grid (the
    reference's MarchingTetrahedraCount / MarchingTetrahedraGenerate,
    dax/worklet/MarchingTetrahedra.h:29-143): tets [C, 4] int32 point ids,
    field [P] float32, points [P, 3] float32 coordinates.  Synchronous."""
    return _cells_contour("dax_hip_mt", 4, tets, field, points, iso, dedup)


def hexahedra(hexes: torch.Tensor, field: torch.Tensor, points: torch.Tensor, iso: float,
              dedup: bool = True) -> Contour:
    """Marching cubes on an unstructured hexahedral grid (the reference's
    MarchingCubesCount / MarchingCubesGenerate on an
    UnstructuredGrid<Hexahedron>, dax/worklet/MarchingCubes.h:53-164): hexes
    [C, 8] int32 point ids in the hexahedron's vertex order, field [P]
    float32, points [P, 3] float32 coordinates.  No normals.  Synchronous."""
    return _cells_contour("dax_hip_mc_hex", 8, hexes, field, points, iso, dedup)


def backward(field: torch.Tensor, iso: float, contour: Contour, dl_dvertices: torch.Tensor,
             dl_dnormals: torch.Tensor, origin=(0, 0, 0), spacing=(1, 1, 1),
             out: Optional[torch.Tensor] = None, scratch: Optional[torch.Tensor] = None):
    """dL/dfield for upstream dL/dvertices, dL/dnormals ([U, 3] f32) of a
    forward run with edges=True on the same field and iso."""
    nx, ny, nz = _check_field(field)
    if contour.edges is None or contour.workspace is None:
        raise ValueError("backward needs a forward with edges=True (dedup)")
    U = contour.num_points
    for t in (dl_dvertices, dl_dnormals):
        if t.shape != (U, 3) or t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("upstream gradients must be contiguous float32 [U, 3]")
    check_tensor(contour.edges, "edges", torch.int32, U, field.device)
    check_tensor(contour.workspace, "workspace", torch.uint8, workspace_bytes(nx, ny, nz),
                 field.device)
    check_tensor(contour.counts, "counts", torch.int64, 2, field.device)
    lib = _lib.load()
    nbytes = int(lib.dax_hip_mc_backward_scratch_bytes(U))
    if scratch is None or scratch.numel() < nbytes:
        scratch = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=field.device)
    check_tensor(scratch, "scratch", torch.uint8, nbytes, field.device)
    if out is None:
        out = torch.empty_like(field)
    check_tensor(out, "out", torch.float32, nx * ny * nz, field.device, exact=True)
    _lib.call("dax_hip_mc_backward", field.data_ptr(), nx, ny, nz, _f3(origin), _f3(spacing),
              float(iso), contour.workspace.data_ptr(), contour.workspace.numel(),
              contour.edges.data_ptr(), contour.counts.data_ptr(), U, dl_dvertices.data_ptr(),
              dl_dnormals.data_ptr(), scratch.data_ptr(), scratch.numel(), out.data_ptr(),
              _stream())
    return out


def backward_into(field: torch.Tensor, iso: float, buf: ForwardBuffers,
                  dl_dvertices: torch.Tensor, dl_dnormals: torch.Tensor, scratch: torch.Tensor,
                  out: torch.Tensor, origin=(0, 0, 0), spacing=(1, 1, 1)) -> None:
    """Enqueue the backward of the last :func:`forward_into` on `buf` (run
    with edges) with preallocated scratch and output (no synchronisation).
    Upstream gradients cover at least the forward's U points."""
    nx, ny, nz = _check_field(field)
    _check_buf(buf, (nx, ny, nz), field.device)
    if buf.edges is None:
        raise ValueError("backward needs forward buffers with edges")
    dev = field.device
    check_tensor(buf.edges, "edges", torch.int32, buf.cap_points, dev)
    check_tensor(dl_dvertices, "dl_dvertices", torch.float32, 3 * buf.cap_points, dev)
    check_tensor(dl_dnormals, "dl_dnormals", torch.float32, 3 * buf.cap_points, dev)
    check_tensor(scratch, "scratch", torch.uint8,
                 int(_lib.load().dax_hip_mc_backward_scratch_bytes(buf.cap_points)), dev)
    check_tensor(out, "out", torch.float32, nx * ny * nz, dev, exact=True)
    _lib.call("dax_hip_mc_backward", field.data_ptr(), nx, ny, nz, _f3(origin), _f3(spacing),
              float(iso), buf.workspace.data_ptr(), buf.workspace.numel(), buf.edges.data_ptr(),
              buf.counts.data_ptr(), buf.cap_points, dl_dvertices.data_ptr(),
              dl_dnormals.data_ptr(), scratch.data_ptr(), scratch.numel(), out.data_ptr(),
              _stream())


def autotune(field: torch.Tensor, iso: float, buf: ForwardBuffers, reps: int = 10,
             dedup: bool = True, normals: bool = True):
    """Run the native autotuner for this shape; returns (params5, ms3)."""
    import ctypes
    nx, ny, nz = _check_field(field)
    flags = (MC_DEDUP if dedup else 0) | (MC_NORMALS if (normals and buf.normals is not None) else 0)
    best = (ctypes.c_int * 5)()
    ms = (ctypes.c_float * 3)()
    _lib.call("dax_hip_mc_autotune", field.data_ptr(), nx, ny, nz, float(iso), flags,
              buf.workspace.data_ptr(), buf.workspace.numel(), buf.vertices.data_ptr(),
              _ptr(buf.normals), buf.conn.data_ptr(), _ptr(buf.edges) if dedup else None,
              buf.cap_points, buf.cap_tris, buf.counts.data_ptr(), int(reps), best, ms, _stream())
    return list(best), list(ms)


def set_tuning(shape, params) -> None:
    """params: classify_unroll, classify_vec, scan_sub, gen_tile[, gen_xcd=0]."""
    import ctypes
    nx, ny, nz = shape
    p = list(params) + [0] * (5 - len(params))
    arr = (ctypes.c_int * 5)(*p)
    _lib.call("dax_hip_mc_set_tuning", nx, ny, nz, arr)


def get_tuning(shape):
    import ctypes
    nx, ny, nz = shape
    arr = (ctypes.c_int * 5)()
    _lib.call("dax_hip_mc_get_tuning", nx, ny, nz, arr)
    return list(arr)


def interpolate_scalar(values: torch.Tensor, contour: Contour) -> torch.Tensor:
    """CompactPointField of a scalar point field onto the contour points."""
    if contour.interp_first is None:
        raise ValueError("forward(..., interp=True) is needed to interpolate point fields")
    if contour.grid_points is None:
        raise ValueError("the contour does not record its grid's point count")
    n = contour.num_points
    dev = contour.interp_first.device
    check_tensor(values, "values", torch.float32, contour.grid_points, dev, exact=True)
    for name in ("interp_first", "interp_second"):
        check_tensor(getattr(contour, name), name, torch.int32, n, dev)
    check_tensor(contour.interp_weight, "interp_weight", torch.float32, n, dev)
    out = torch.empty(n, dtype=torch.float32, device=values.device)
    _lib.call("dax_hip_mc_interpolate_scalar", values.data_ptr(), contour.interp_first.data_ptr(),
              contour.interp_second.data_ptr(), contour.interp_weight.data_ptr(), n,
              out.data_ptr(), _stream())
    return out


def point_gradient(field: torch.Tensor, spacing=(1, 1, 1)) -> torch.Tensor:
    nx, ny, nz = _check_field(field)
    out = torch.empty((nx * ny * nz, 3), dtype=torch.float32, device=field.device)
    _lib.call("dax_hip_point_gradient", field.data_ptr(), nx, ny, nz, _f3(spacing),
              out.data_ptr(), _stream())
    return out


def cell_gradient(field: torch.Tensor, origin=(0, 0, 0), spacing=(1, 1, 1)) -> torch.Tensor:
    nx, ny, nz = _check_field(field)
    C = max(nx - 1, 0) * max(ny - 1, 0) * max(nz - 1, 0)
    out = torch.empty((C, 3), dtype=torch.float32, device=field.device)
    _lib.call("dax_hip_cell_gradient", field.data_ptr(), nx, ny, nz, _f3(origin), _f3(spacing),
              out.data_ptr(), _stream())
    return out


class Batch:
    """Many independent volumes run as one batched forward (optionally as a
    captured hipGraph).  `fields` is one flat float32 device tensor holding
    the volumes back to back (volume v at ``field_offsets[v]``)."""

    def __init__(self, dims, isos, origins=None, spacings=None, device="cuda"):
        import ctypes

        import numpy as np
        self.n = len(dims)
        self.dims = np.ascontiguousarray(np.asarray(dims, dtype=np.int64).reshape(self.n, 3))
        sizes = self.dims.prod(axis=1)
        self.field_offsets = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)[:-1]]),
                                                  dtype=np.int64)
        self.total_points = int(sizes.sum())
        self.isos = np.ascontiguousarray(np.asarray(isos, dtype=np.float32).reshape(self.n))
        self.origins = (None if origins is None else
                        np.ascontiguousarray(np.asarray(origins, np.float32).reshape(self.n, 3)))
        self.spacings = (None if spacings is None else
                         np.ascontiguousarray(np.asarray(spacings, np.float32).reshape(self.n, 3)))
        lib = _lib.load()
        p = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
        nbytes = int(lib.dax_hip_mc_batch_workspace_bytes(self.n, p(self.dims)))
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.header = ctypes.create_string_buffer(int(lib.dax_hip_mc_batch_header_bytes()))
        _lib.call("dax_hip_mc_batch_prepare", self.n, p(self.dims), p(self.field_offsets),
                  p(self.isos), p(self.origins), p(self.spacings), self.workspace.data_ptr(),
                  self.workspace.numel(), self.header, _stream())
        self.device = device
        self.counts = torch.zeros(2, dtype=torch.int64, device=device)
        self.vol_offsets = torch.zeros((self.n, 2), dtype=torch.int64, device=device)

    def _call(self, name, fields, bufs, flags, extra=()):
        return _lib.call(name, fields.data_ptr(), self.header, self.workspace.data_ptr(),
                         self.workspace.numel(), flags, bufs["vertices"].data_ptr(),
                         _ptr(bufs.get("normals")), bufs["conn"].data_ptr(), _ptr(bufs.get("edges")),
                         bufs["cap_points"], bufs["cap_tris"], self.counts.data_ptr(),
                         self.vol_offsets.data_ptr(), _stream(), *extra)

    def alloc(self, cap_points, cap_tris, normals=True, edges=True):
        dev = self.device
        return {"vertices": torch.empty((max(cap_points, 1), 3), dtype=torch.float32, device=dev),
                "normals": (torch.empty((max(cap_points, 1), 3), dtype=torch.float32, device=dev)
                            if normals else None),
                "conn": torch.empty((max(cap_tris, 1), 3), dtype=torch.int32, device=dev),
                "edges": torch.empty(max(cap_points, 1), dtype=torch.int32, device=dev) if edges else None,
                "cap_points": int(cap_points), "cap_tris": int(cap_tris)}

    def set_tuning(self, params5):
        """Launch shapes {classify unroll, classify blocks/CU, scan threads,
        generate blocks/CU, generate XCD order} for this batch."""
        import ctypes
        arr = (ctypes.c_int * 5)(*[int(x) for x in params5])
        _lib.call("dax_hip_mc_batch_set_tuning", self.header, arr)

    def get_tuning(self):
        import ctypes
        arr = (ctypes.c_int * 5)()
        _lib.call("dax_hip_mc_batch_get_tuning", self.header, arr)
        return list(arr)

    def autotune(self, fields, bufs, reps=5, dedup=True, normals=True):
        """Time every launch shape of each phase on this batch and keep the
        fastest (also cached for later batches of the same shape class);
        returns (params5, ms3).  Synchronous."""
        import ctypes
        flags = (MC_DEDUP if dedup else 0) | (MC_NORMALS if (normals and bufs.get("normals") is not None) else 0)
        self.check_buffers(fields, bufs)
        best = (ctypes.c_int * 5)()
        ms = (ctypes.c_float * 3)()
        _lib.call("dax_hip_mc_batch_autotune", fields.data_ptr(), self.header,
                  self.workspace.data_ptr(), self.workspace.numel(), flags,
                  bufs["vertices"].data_ptr(), _ptr(bufs.get("normals")), bufs["conn"].data_ptr(),
                  _ptr(bufs.get("edges")), bufs["cap_points"], bufs["cap_tris"],
                  self.counts.data_ptr(), int(reps), best, ms, _stream())
        return list(best), list(ms)

    def sizes(self, fields, dedup=True):
        """Run once without room for output to learn {T, U}."""
        probe = self.alloc(0, 0, normals=False, edges=False)
        self.check_buffers(fields, probe)
        self._call("dax_hip_mc_batch_forward", fields, probe, MC_DEDUP if dedup else 0)
        return read_counts(self.counts)

    def check_buffers(self, fields, bufs, need_edges=False):
        """The flat fields hold every volume and the outputs their capacities."""
        dev = torch.device(self.device) if not isinstance(self.device, torch.device) else self.device
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        check_tensor(fields, "fields", torch.float32, self.total_points, dev)
        cp, ct = int(bufs["cap_points"]), int(bufs["cap_tris"])
        check_tensor(bufs["vertices"], "vertices", torch.float32, 3 * cp, dev)
        check_tensor(bufs["conn"], "conn", torch.int32, 3 * ct, dev)
        if bufs.get("normals") is not None:
            check_tensor(bufs["normals"], "normals", torch.float32, 3 * cp, dev)
        if bufs.get("edges") is not None or need_edges:
            check_tensor(bufs.get("edges"), "edges", torch.int32, cp, dev)
        return dev

    def forward_into(self, fields, bufs, dedup=True, normals=True):
        self.check_buffers(fields, bufs)
        flags = (MC_DEDUP if dedup else 0) | (MC_NORMALS if (normals and bufs.get("normals") is not None) else 0)
        if not dedup and bufs.get("edges") is not None:
            bufs = dict(bufs, edges=None)
        self._call("dax_hip_mc_batch_forward", fields, bufs, flags)

    def capture(self, fields, bufs, dedup=True, normals=True, stream=None):
        """Capture the batch forward into a graph on `stream` (a non-default
        torch stream); returns the graph handle."""
        import ctypes
        flags = (MC_DEDUP if dedup else 0) | (MC_NORMALS if (normals and bufs.get("normals") is not None) else 0)
        self.check_buffers(fields, bufs)
        h = ctypes.c_void_p()
        with torch.cuda.stream(stream):
            self._call("dax_hip_mc_batch_graph_create", fields, bufs, flags, extra=(ctypes.byref(h),))
        return h

    @staticmethod
    def launch(graph, stream=None):
        s = (stream or torch.cuda.current_stream()).cuda_stream
        _lib.call("dax_hip_graph_launch", graph, s)

    @staticmethod
    def destroy(graph):
        _lib.call("dax_hip_graph_destroy", graph)

    def scratch_bytes(self, n_points: int) -> int:
        return int(_lib.load().dax_hip_mc_backward_scratch_bytes(int(n_points)))

    def backward(self, fields, bufs, dl_dvertices, dl_dnormals, scratch, out):
        """dL/dfields (flat, like `fields`) of the last batch forward into
        `bufs` (run with edges) for upstream [U, 3] gradients."""
        U = bufs["cap_points"]
        if bufs.get("edges") is None:
            raise ValueError("the batch backward needs a forward with the edge list")
        dev = self.check_buffers(fields, bufs, need_edges=True)
        check_tensor(dl_dvertices, "dl_dvertices", torch.float32, 3 * U, dev)
        check_tensor(dl_dnormals, "dl_dnormals", torch.float32, 3 * U, dev)
        check_tensor(scratch, "scratch", torch.uint8, self.scratch_bytes(U), dev)
        check_tensor(out, "out", torch.float32, self.total_points, dev)
        _lib.call("dax_hip_mc_batch_backward", fields.data_ptr(), self.header,
                  self.workspace.data_ptr(), self.workspace.numel(), bufs["edges"].data_ptr(),
                  self.counts.data_ptr(), U, dl_dvertices.data_ptr(), dl_dnormals.data_ptr(),
                  scratch.data_ptr(), scratch.numel(), out.data_ptr(), _stream())

    def loop_workspace(self):
        lib = _lib.load()
        need = max(int(lib.dax_hip_mc_workspace_bytes(*map(int, d))) for d in self.dims)
        return torch.empty(need, dtype=torch.uint8, device=self.device)

    def loop(self, fields, bufs, vol_offsets_host, workspace, counts_scratch, dedup=True,
             normals=True, grads=None):
        """Baseline: one single-volume forward (and, with grads =
        (dl_dvertices, dl_dnormals, scratch, dl_dfields), one single-volume
        backward) per volume, as a native launch loop, into the same
        concatenated outputs."""
        import ctypes

        import numpy as np
        p = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
        dev = self.check_buffers(fields, bufs)
        need = max(int(_lib.load().dax_hip_mc_workspace_bytes(*map(int, d))) for d in self.dims)
        check_tensor(workspace, "workspace", torch.uint8, need, dev)
        check_tensor(counts_scratch, "counts_scratch", torch.int64, 2, dev)
        offs = np.ascontiguousarray(vol_offsets_host, dtype=np.int64).reshape(self.n, 2)
        flags = (MC_DEDUP if dedup else 0) | (MC_NORMALS if (normals and bufs.get("normals") is not None) else 0)
        dv = dn = sc = df = None
        sc_bytes = 0
        if grads is not None:
            dv, dn, scr, dfo = grads
            if bufs.get("edges") is None:
                raise ValueError("the loop backward needs forward buffers with the edge list")
            U = int(bufs["cap_points"])
            check_tensor(dv, "dl_dvertices", torch.float32, 3 * U, dev)
            check_tensor(dn, "dl_dnormals", torch.float32, 3 * U, dev)
            check_tensor(scr, "scratch", torch.uint8, self.scratch_bytes(U), dev)
            check_tensor(dfo, "dl_dfields", torch.float32, self.total_points, dev)
            dv, dn, sc, df, sc_bytes = dv.data_ptr(), dn.data_ptr(), scr.data_ptr(), dfo.data_ptr(), scr.numel()
        _lib.call("dax_hip_mc_volume_loop", self.n, p(self.dims), p(self.field_offsets),
                  p(self.isos), p(self.origins), p(self.spacings), fields.data_ptr(), flags,
                  workspace.data_ptr(), workspace.numel(), p(offs), bufs["cap_points"],
                  bufs["cap_tris"], bufs["vertices"].data_ptr(), _ptr(bufs.get("normals")),
                  bufs["conn"].data_ptr(), _ptr(bufs.get("edges")) if dedup else None,
                  counts_scratch.data_ptr(), dv, dn, sc, sc_bytes, df, _stream())

    def split(self, bufs, dedup=True):
        """Per-volume views (vertices, normals, connectivity) of the outputs."""
        T, U = read_counts(self.counts, bufs["cap_points"], bufs["cap_tris"])
        offs = self.vol_offsets.cpu().numpy()
        out = []
        for v in range(self.n):
            t0, p0 = int(offs[v, 0]), int(offs[v, 1])
            t1 = int(offs[v + 1, 0]) if v + 1 < self.n else T
            p1 = int(offs[v + 1, 1]) if v + 1 < self.n else U
            out.append((bufs["vertices"][p0:p1],
                        None if bufs.get("normals") is None else bufs["normals"][p0:p1],
                        bufs["conn"][t0:t1]))
        return out


def capture(fn, stream: torch.cuda.Stream):
    """Capture every launch `fn()` makes through this library on `stream`
    (a non-default torch stream) into one executable graph; returns its
    handle for :meth:`Batch.launch` / :meth:`Batch.destroy`.  `fn` must not
    allocate or synchronise."""
    import ctypes
    h = ctypes.c_void_p()
    with torch.cuda.stream(stream):
        _lib.call("dax_hip_graph_begin_capture", stream.cuda_stream)
        try:
            fn()
        finally:
            _lib.call("dax_hip_graph_end_capture", stream.cuda_stream, ctypes.byref(h))
    return h


def launch_graph(graph, stream: torch.cuda.Stream) -> None:
    """Replay a graph made by :func:`capture` on `stream`."""
    _lib.call("dax_hip_graph_launch", graph, stream.cuda_stream)


def destroy_graph(graph) -> None:
    _lib.call("dax_hip_graph_destroy", graph)

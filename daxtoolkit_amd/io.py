"""VTK legacy ASCII output of the isosurface and topology results (the
format of the reference benchmarks' PrintContentsToStream,
Benchmarks/MarchingCubes/Pipeline.h:51-98 and
Benchmarks/Threshold/Pipeline.h:99-150), written by the library's
multi-threaded host writer."""
from __future__ import annotations

import numpy as np

from . import _lib

# VTK cell type ids of the cell shapes this package produces
VTK_TRIANGLE, VTK_TETRA, VTK_VOXEL, VTK_HEXAHEDRON = 5, 10, 11, 12


def _host(a, dtype):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(a, dtype)


def write_vtk_polydata(path, points, connectivity) -> None:
    """DATASET POLYDATA with POLYGONS: points [U, 3], connectivity [T, V]
    (a marching-cubes Contour's vertices and connectivity)."""
    pts = _host(points, np.float32).reshape(-1, 3)
    conn = _host(connectivity, np.int32)
    nv = conn.shape[1] if conn.ndim == 2 else 3
    _lib.call("dax_hip_write_vtk_polydata", str(path).encode(), pts.ctypes.data, pts.shape[0],
              conn.ctypes.data, conn.size // nv, nv)


def write_vtk_unstructured(path, points, connectivity, cell_type: int = VTK_HEXAHEDRON) -> None:
    """DATASET UNSTRUCTURED_GRID with CELLS and CELL_TYPES: points [U, 3],
    connectivity [T, V] (a Threshold result's points and connectivity)."""
    pts = _host(points, np.float32).reshape(-1, 3)
    conn = _host(connectivity, np.int32)
    nv = conn.shape[1] if conn.ndim == 2 else 8
    _lib.call("dax_hip_write_vtk_unstructured", str(path).encode(), pts.ctypes.data,
              pts.shape[0], conn.ctypes.data, conn.size // nv, nv, int(cell_type))

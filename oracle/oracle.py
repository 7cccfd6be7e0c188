"""ctypes wrapper of the oracle library (test infrastructure only).

The oracle is a CPU restatement of Dax's Serial-adapter marching-cubes path
(see mc_oracle.cpp for the reference lines each function follows).  Only the
tests, the smoke check and the CPU-baseline leg of bench.py use it, always as
the checker or the baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_SO = _HERE / "liboracle_mc.so"
_lib = None

_P = np.ctypeslib.ndpointer
_i64p = C.POINTER(C.c_int64)


def build() -> None:
    subprocess.run(["make", "-C", str(_HERE)], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not _SO.exists():
            build()
        L = C.CDLL(str(_SO))
        L.oracle_tables_expand.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_mc_count.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p]
        L.oracle_mc_count.restype = C.c_int
        for n in ("oracle_scan_inclusive_i32", "oracle_scan_exclusive_i32"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
            getattr(L, n).restype = C.c_int64
        for n in ("oracle_lower_bounds_i32", "oracle_upper_bounds_i32"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p]
        L.oracle_mc_generate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                         C.c_int, C.c_int64, C.c_int64, _i64p, _i64p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
        L.oracle_mc_generate.restype = C.c_int
        L.oracle_interpolate_scalar.argtypes = [C.c_void_p] * 4 + [C.c_int64, C.c_void_p]
        L.oracle_point_gradient.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_mc_normals.argtypes = [C.c_void_p] * 6 + [C.c_int64, C.c_void_p]
        L.oracle_cell_gradient.argtypes = [C.c_void_p] * 5
        L.oracle_mc_backward_f64.argtypes = ([C.c_void_p] * 4 + [C.c_float] + [C.c_void_p] * 2 +
                                             [C.c_int64] + [C.c_void_p] * 3)
        L.oracle_mc_backward_f32.argtypes = L.oracle_mc_backward_f64.argtypes
        L.oracle_slice_field.argtypes = [C.c_void_p] * 6
        L.oracle_slice_field.restype = C.c_float
        L.oracle_mt_generate.argtypes = ([C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                          C.c_float, C.c_int, C.c_int64, _i64p, _i64p] +
                                         [C.c_void_p] * 5)
        L.oracle_mt_generate.restype = C.c_int
        L.oracle_threshold.argtypes = ([C.c_void_p, C.c_int, C.c_int64] + [C.c_void_p] * 4 +
                                       [C.c_int64, C.c_void_p, C.c_float, C.c_float, C.c_int,
                                        C.c_int64, C.c_void_p, _i64p, _i64p] + [C.c_void_p] * 3)
        L.oracle_threshold.restype = C.c_int
        L.oracle_write_vtk.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                                       C.c_int64, C.c_int, C.c_int]
        L.oracle_write_vtk.restype = C.c_int
        L.oracle_mc_hex_generate.argtypes = L.oracle_mt_generate.argtypes
        L.oracle_mc_hex_generate.restype = C.c_int
        L.oracle_tet_tables.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_mc_fd_f64.argtypes = ([C.c_void_p] * 4 + [C.c_float] + [C.c_void_p] * 2 +
                                       [C.c_int64] + [C.c_void_p] * 3 + [C.c_int64, C.c_double,
                                                                          C.c_void_p])
        L.oracle_mc_pipeline_omp.argtypes = [C.c_void_p] * 4 + [C.c_float, C.c_int, C.c_int,
                                                                 _i64p, _i64p]
        L.oracle_mc_pipeline_omp.restype = C.c_double
        L.oracle_unstructured_pipeline_omp.argtypes = [C.c_void_p, C.c_int, C.c_int64,
                                                       C.c_void_p, C.c_void_p, C.c_float, C.c_int,
                                                       C.c_int, _i64p, _i64p]
        L.oracle_unstructured_pipeline_omp.restype = C.c_double
        L.oracle_threshold_pipeline_omp.argtypes = [C.c_void_p] * 4 + [C.c_float, C.c_float,
                                                                        C.c_int, C.c_int, _i64p,
                                                                        _i64p]
        L.oracle_threshold_pipeline_omp.restype = C.c_double
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _dims(field: np.ndarray) -> np.ndarray:
    nz, ny, nx = field.shape
    return np.array([nx, ny, nz], dtype=np.int64)


def _f3(v) -> np.ndarray:
    return np.asarray(v, dtype=np.float32).reshape(3)


def _field(field) -> np.ndarray:
    f = np.ascontiguousarray(field, dtype=np.float32)
    assert f.ndim == 3
    return f


def tables():
    nf = np.zeros(256, np.uint8)
    tri = np.zeros(4096, np.uint8)
    lib().oracle_tables_expand(_p(nf), _p(tri))
    return nf, tri.reshape(256, 16)


def count(field, iso):
    """Per-cell case (u8) and triangle count (i32) in flat cell order."""
    f = _field(field)
    nz, ny, nx = f.shape
    C_ = max(nx - 1, 0) * max(ny - 1, 0) * max(nz - 1, 0)
    cases = np.zeros(C_, np.uint8)
    cnt = np.zeros(C_, np.int32)
    lib().oracle_mc_count(_p(f), _p(_dims(f)), float(iso), _p(cases), _p(cnt))
    return cases, cnt


def scan_inclusive(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    out = np.zeros_like(a)
    total = lib().oracle_scan_inclusive_i32(_p(a), _p(out), a.size)
    return out, int(total)


def scan_exclusive(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    out = np.zeros_like(a)
    total = lib().oracle_scan_exclusive_i32(_p(a), _p(out), a.size)
    return out, int(total)


def lower_bounds(sorted_, vals):
    s = np.ascontiguousarray(sorted_, dtype=np.int32)
    v = np.ascontiguousarray(vals, dtype=np.int32)
    out = np.zeros(v.size, np.int32)
    lib().oracle_lower_bounds_i32(_p(s), s.size, _p(v), v.size, _p(out))
    return out


def upper_bounds(sorted_, vals):
    s = np.ascontiguousarray(sorted_, dtype=np.int32)
    v = np.ascontiguousarray(vals, dtype=np.int32)
    out = np.zeros(v.size, np.int32)
    lib().oracle_upper_bounds_i32(_p(s), s.size, _p(v), v.size, _p(out))
    return out


def generate(field, iso, dedup=True, origin=(0, 0, 0), spacing=(1, 1, 1)):
    """The reference's generate step; returns dict with n_tris, n_pts, conn,
    first, second, weight (interpolation records) and coords."""
    f = _field(field)
    dims = _dims(f)
    o, s = _f3(origin), _f3(spacing)
    _, cnt = count(f, iso)
    T = int(cnt.sum(dtype=np.int64))
    cap_p = max(3 * T, 1)
    conn = np.zeros(max(3 * T, 1), np.int32)
    first = np.zeros(cap_p, np.int32)
    second = np.zeros(cap_p, np.int32)
    weight = np.zeros(cap_p, np.float32)
    coords = np.zeros((cap_p, 3), np.float32)
    nt, npt = C.c_int64(), C.c_int64()
    rc = lib().oracle_mc_generate(_p(f), _p(dims), _p(o), _p(s), float(iso), int(dedup), T, cap_p,
                                  C.byref(nt), C.byref(npt), _p(conn), _p(first), _p(second),
                                  _p(weight), _p(coords))
    assert rc == 0
    T, U = int(nt.value), int(npt.value)
    return dict(n_tris=T, n_pts=U, conn=conn[:3 * T], first=first[:U], second=second[:U],
                weight=weight[:U], coords=coords[:U])


def interpolate_scalar(values, first, second, weight):
    v = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
    first = np.ascontiguousarray(first, np.int32)
    second = np.ascontiguousarray(second, np.int32)
    weight = np.ascontiguousarray(weight, np.float32)
    out = np.zeros(first.size, np.float32)
    lib().oracle_interpolate_scalar(_p(v), _p(first), _p(second), _p(weight), first.size, _p(out))
    return out


def point_gradient(field, spacing=(1, 1, 1)):
    f = _field(field)
    g = np.zeros((f.size, 3), np.float32)
    lib().oracle_point_gradient(_p(f), _p(_dims(f)), _p(_f3(spacing)), _p(g))
    return g


def normals(field, first, second, weight, spacing=(1, 1, 1)):
    f = _field(field)
    first = np.ascontiguousarray(first, np.int32)
    second = np.ascontiguousarray(second, np.int32)
    weight = np.ascontiguousarray(weight, np.float32)
    out = np.zeros((first.size, 3), np.float32)
    lib().oracle_mc_normals(_p(f), _p(_dims(f)), _p(_f3(spacing)), _p(first), _p(second),
                            _p(weight), first.size, _p(out))
    return out


def cell_gradient(field, origin=(0, 0, 0), spacing=(1, 1, 1)):
    f = _field(field)
    nz, ny, nx = f.shape
    C_ = max(nx - 1, 0) * max(ny - 1, 0) * max(nz - 1, 0)
    out = np.zeros((C_, 3), np.float32)
    lib().oracle_cell_gradient(_p(f), _p(_dims(f)), _p(_f3(origin)), _p(_f3(spacing)), _p(out))
    return out


def slice_field(dims, plane_origin, plane_normal, origin=(0, 0, 0), spacing=(1, 1, 1)):
    """The point values and isovalue Dax's SliceCount/SliceGenerate classify
    and interpolate (see mc_oracle.cpp); dims = (nx, ny, nz).  Returns
    (field [nz, ny, nx] float32, iso float32); oracle.generate on them is
    the slice."""
    nx, ny, nz = (int(v) for v in dims)
    out = np.zeros((nz, ny, nx), np.float32)
    d = np.array([nx, ny, nz], np.int64)
    iso = lib().oracle_slice_field(_p(d), _p(_f3(origin)), _p(_f3(spacing)),
                                   _p(_f3(plane_origin)), _p(_f3(plane_normal)), _p(out))
    return out, np.float32(iso)


def _voxel_cells(dims, split, origin, spacing):
    nx, ny, nz = (int(v) for v in dims)
    corner = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                       (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], np.int64)
    k, j, i = np.meshgrid(np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
    cell = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)          # x-fastest cells
    pts_ijk = cell[:, None, None, :] + corner[split][None, :, :, :]     # [C, S, V, 3]
    ids = pts_ijk[..., 0] + nx * (pts_ijk[..., 1] + ny * pts_ijk[..., 2])
    cells = ids.reshape(-1, split.shape[1]).astype(np.int32)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    o = np.asarray(origin, np.float32)
    sp = np.asarray(spacing, np.float32)
    pts = np.stack([o[0] + sp[0] * i.ravel().astype(np.float32),
                    o[1] + sp[1] * j.ravel().astype(np.float32),
                    o[2] + sp[2] * k.ravel().astype(np.float32)], axis=1).astype(np.float32)
    return cells, pts


def freudenthal_tets(dims, origin=(0, 0, 0), spacing=(1, 1, 1)):
    """The tetrahedral grid the reference's tests build from a uniform grid
    (dax/cont/testing/TestingGridGenerator.h:229-257,300-339): each voxel
    split into 6 tetrahedra (Freudenthal), voxel corners numbered as the
    hexahedron's.  Returns (tets [6C, 4] int32, points [P, 3] float32)."""
    split = np.array([[0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6], [0, 5, 6, 4],
                      [1, 5, 6, 0], [1, 6, 2, 0]], np.int64)
    return _voxel_cells(dims, split, origin, spacing)


def voxel_hexes(dims, origin=(0, 0, 0), spacing=(1, 1, 1)):
    """The hexahedral grid the reference's tests build from a uniform grid
    (dax/cont/testing/TestingGridGenerator.h:229-257,266-295): one
    hexahedron per voxel, vertices in the voxel's corner order.  Returns
    (hexes [C, 8] int32, points [P, 3] float32)."""
    return _voxel_cells(dims, np.arange(8, dtype=np.int64)[None, :], origin, spacing)


def tet_tables():
    """(NumFaces uint8[16], TriTable uint8[16, 8]) in the reference's layout."""
    nf = np.zeros(16, np.uint8)
    tri = np.zeros((16, 8), np.uint8)
    lib().oracle_tet_tables(_p(nf), _p(tri))
    return nf, tri


def _unstructured(fn, nv, cells, field, points, iso, dedup):
    cells = np.ascontiguousarray(cells, np.int32).reshape(-1, nv)
    field = np.ascontiguousarray(field, np.float32).reshape(-1)
    points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    nt, npt = C.c_int64(), C.c_int64()
    cap = 3 * 5 * max(cells.shape[0], 1)
    conn = np.zeros(cap, np.int32)
    first = np.zeros(cap, np.int32)
    second = np.zeros(cap, np.int32)
    weight = np.zeros(cap, np.float32)
    coords = np.zeros((cap, 3), np.float32)
    rc = fn(_p(cells), cells.shape[0], _p(field), _p(points), float(iso), int(dedup), cap,
            C.byref(nt), C.byref(npt), _p(conn), _p(first), _p(second), _p(weight), _p(coords))
    assert rc == 0
    T, U = nt.value, npt.value
    return {"n_tris": T, "n_pts": U, "conn": conn[:3 * T], "first": first[:U],
            "second": second[:U], "weight": weight[:U], "coords": coords[:U]}


def mt_generate(tets, field, points, iso, dedup=True):
    """Marching tetrahedra, the Serial-adapter sequence (see mc_oracle.cpp)."""
    return _unstructured(lib().oracle_mt_generate, 4, tets, field, points, iso, dedup)


def mc_hex_generate(hexes, field, points, iso, dedup=True):
    """Marching cubes on an unstructured hexahedral grid, the Serial-adapter
    sequence (see mc_oracle.cpp)."""
    return _unstructured(lib().oracle_mc_hex_generate, 8, hexes, field, points, iso, dedup)


def threshold(field, lo, hi, dedup=True, cells=None, points=None, origin=(0, 0, 0),
              spacing=(1, 1, 1)):
    """Threshold through the Serial-adapter sequence (see mc_oracle.cpp).
    Uniform grid: field of shape (nz, ny, nx); unstructured: cells [C, V]
    int32 and points [P, 3].  Returns counts, n_cells, n_points, conn
    [n_cells, V], coords [n_points, 3] and mask (dedup), and the compacted
    field (CompactPointField = StreamCompact of the field by the mask)."""
    if cells is None:
        f = _field(field)
        dims = _dims(f)
        nv, C_ = 8, int((dims[0] - 1) * (dims[1] - 1) * (dims[2] - 1))
        P = int(f.size)
        cptr, pptr = None, None
        f = f.reshape(-1)
    else:
        cells = np.ascontiguousarray(cells, np.int32)
        nv, C_ = cells.shape[1], cells.shape[0]
        f = np.ascontiguousarray(field, np.float32).reshape(-1)
        points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        P = int(f.size)
        dims = np.zeros(3, np.int64)
        cptr, pptr = _p(cells), _p(points)
    counts = np.zeros(max(C_, 1), np.int32)
    conn = np.zeros(max(C_, 1) * nv, np.int32)
    coords = np.zeros((max(P, 1), 3), np.float32)
    mask = np.zeros(max(P, 1), np.int32)
    nt, npt = C.c_int64(), C.c_int64()
    rc = lib().oracle_threshold(cptr, nv, C_, _p(dims), _p(_f3(origin)), _p(_f3(spacing)), pptr, P,
                                _p(f), float(lo), float(hi), int(dedup), C_, _p(counts),
                                C.byref(nt), C.byref(npt), _p(conn), _p(coords), _p(mask))
    assert rc == 0
    T, U = nt.value, npt.value
    out = {"counts": counts[:C_], "n_cells": T, "n_points": U, "conn": conn[:T * nv].reshape(T, nv)}
    if dedup:
        out["coords"] = coords[:U]
        out["mask"] = mask[:P]
        out["field"] = f[mask[:P] != 0]
    return out


def write_vtk(path, points, conn, polydata=True, cell_type=12):
    """The reference benchmarks' ASCII VTK writers (see mc_oracle.cpp)."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    conn = np.ascontiguousarray(conn, np.int32)
    nv = conn.shape[1] if conn.ndim == 2 else 3
    n_cells = conn.size // nv
    rc = lib().oracle_write_vtk(str(path).encode(), int(polydata), _p(pts), pts.shape[0], _p(conn),
                                n_cells, nv, cell_type)
    assert rc == 0


def backward_f64(field, iso, first, second, dldx, dldn, origin=(0, 0, 0), spacing=(1, 1, 1)):
    f = _field(field)
    first = np.ascontiguousarray(first, np.int32)
    second = np.ascontiguousarray(second, np.int32)
    dldx = np.ascontiguousarray(dldx, np.float32)
    dldn = np.ascontiguousarray(dldn, np.float32)
    out = np.zeros(f.size, np.float64)
    lib().oracle_mc_backward_f64(_p(f), _p(_dims(f)), _p(_f3(origin)), _p(_f3(spacing)),
                                 float(iso), _p(first), _p(second), first.size, _p(dldx),
                                 _p(dldn), _p(out))
    return out


def backward_f32(field, iso, first, second, dldx, dldn, origin=(0, 0, 0), spacing=(1, 1, 1)):
    """dL/dfield in FP32 with the device adjoint's operation order (bit-level
    checker; backward_f64 is the mathematical one)."""
    f = _field(field)
    first = np.ascontiguousarray(first, np.int32)
    second = np.ascontiguousarray(second, np.int32)
    dldx = np.ascontiguousarray(dldx, np.float32)
    dldn = np.ascontiguousarray(dldn, np.float32)
    out = np.zeros(f.size, np.float32)
    lib().oracle_mc_backward_f32(_p(f), _p(_dims(f)), _p(_f3(origin)), _p(_f3(spacing)),
                                 float(iso), _p(first), _p(second), first.size, _p(dldx),
                                 _p(dldn), _p(out))
    return out


def fd_f64(field, iso, first, second, dldx, dldn, points, h=1e-3, origin=(0, 0, 0),
           spacing=(1, 1, 1)):
    f = _field(field)
    first = np.ascontiguousarray(first, np.int32)
    second = np.ascontiguousarray(second, np.int32)
    dldx = np.ascontiguousarray(dldx, np.float32)
    dldn = np.ascontiguousarray(dldn, np.float32)
    pts = np.ascontiguousarray(points, np.int64)
    out = np.zeros(pts.size, np.float64)
    lib().oracle_mc_fd_f64(_p(f), _p(_dims(f)), _p(_f3(origin)), _p(_f3(spacing)), float(iso),
                           _p(first), _p(second), first.size, _p(dldx), _p(dldn), _p(pts),
                           pts.size, float(h), _p(out))
    return out


def pipeline_omp(field, iso, dedup=True, threads=0, origin=(0, 0, 0), spacing=(1, 1, 1)):
    """Time the OpenMP restatement of Dax's OpenMP-adapter MC pipeline.
    Returns (seconds, T, U)."""
    f = _field(field)
    nt, npt = C.c_int64(), C.c_int64()
    sec = lib().oracle_mc_pipeline_omp(_p(f), _p(_dims(f)), _p(_f3(origin)), _p(_f3(spacing)),
                                       float(iso), int(dedup), int(threads), C.byref(nt),
                                       C.byref(npt))
    return float(sec), int(nt.value), int(npt.value)


def unstructured_pipeline_omp(cells, field, points, iso, dedup=True, threads=0):
    """Time the OpenMP restatement of the unstructured-grid contour (cells
    [C, 4]: marching tetrahedra, [C, 8]: marching cubes on hexahedra).
    Returns (seconds, T, U)."""
    cells = np.ascontiguousarray(cells, np.int32)
    nv = cells.shape[1]
    assert nv in (4, 8)
    field = np.ascontiguousarray(field, np.float32).reshape(-1)
    points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    nt, npt = C.c_int64(), C.c_int64()
    sec = lib().oracle_unstructured_pipeline_omp(_p(cells), nv, cells.shape[0], _p(field),
                                                 _p(points), float(iso), int(dedup), int(threads),
                                                 C.byref(nt), C.byref(npt))
    return float(sec), int(nt.value), int(npt.value)


def threshold_pipeline_omp(field, lo, hi, dedup=True, threads=0, origin=(0, 0, 0),
                           spacing=(1, 1, 1)):
    """Time the OpenMP restatement of threshold on a uniform grid (field of
    shape (nz, ny, nx)).  Returns (seconds, output cells, output points)."""
    f = _field(field)
    nc, npt = C.c_int64(), C.c_int64()
    sec = lib().oracle_threshold_pipeline_omp(_p(_dims(f)), _p(_f3(origin)), _p(_f3(spacing)),
                                              _p(f), float(lo), float(hi), int(dedup),
                                              int(threads), C.byref(nc), C.byref(npt))
    return float(sec), int(nc.value), int(npt.value)

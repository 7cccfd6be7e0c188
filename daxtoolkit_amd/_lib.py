"""ctypes binding of ``libdaxhip.so`` (the C-ABI declared in include/dax_hip.h).

The library is required: there is no fallback path.  Loading fails loudly if
it has not been built (``python -c "import __graft_entry__ as g; g.build()"``).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libdaxhip.so"
_lib = None

OK, OOM, EXEC_ERROR, BAD_VALUE, CAPACITY = 0, 1, 2, 3, 4
MC_DEDUP = 1
MC_NORMALS = 2
TOPOLOGY_MASK_OPTIONAL = 4


class DaxHipError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class ErrorControlOutOfMemory(DaxHipError):
    pass


class ErrorExecution(DaxHipError):
    pass


class ErrorControlBadValue(DaxHipError):
    pass


class ErrorCapacity(DaxHipError):
    pass


_ERRORS = {OOM: ErrorControlOutOfMemory, EXEC_ERROR: ErrorExecution,
           BAD_VALUE: ErrorControlBadValue, CAPACITY: ErrorCapacity}

vp, i64, u32, f32, sz = C.c_void_p, C.c_int64, C.c_uint32, C.c_float, C.c_size_t
P_i64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
_SIGNATURES = {
    "dax_hip_last_error": (C.c_char_p, []),
    "dax_hip_version": (C.c_char_p, []),
    "dax_hip_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "dax_hip_malloc": (C.c_int, [C.POINTER(vp), sz]),
    "dax_hip_free": (C.c_int, [vp]),
    "dax_hip_memcpy_h2d": (C.c_int, [vp, vp, sz, vp]),
    "dax_hip_memcpy_d2h": (C.c_int, [vp, vp, sz, vp]),
    "dax_hip_memcpy_d2d": (C.c_int, [vp, vp, sz, vp]),
    "dax_hip_memset": (C.c_int, [vp, C.c_int, sz, vp]),
    "dax_hip_synchronize": (C.c_int, [vp]),
    "dax_hip_copy": (C.c_int, [vp, vp, sz, vp]),
    "dax_hip_magnitude_vec3": (C.c_int, [vp, i64, vp, vp]),
    "dax_hip_reduce_sum_i32": (C.c_int, [vp, i64, vp, vp]),
    "dax_hip_scan_inclusive_i32": (C.c_int, [vp, vp, i64, vp, vp]),
    "dax_hip_scan_exclusive_i32": (C.c_int, [vp, vp, i64, vp, vp]),
    "dax_hip_lower_bounds_i32": (C.c_int, [vp, i64, vp, i64, C.c_int, vp, vp, vp]),
    "dax_hip_upper_bounds_i32": (C.c_int, [vp, i64, vp, i64, C.c_int, vp, vp, vp]),
    "dax_hip_scan_inclusive_i64": (C.c_int, [vp, vp, i64, vp, vp]),
    "dax_hip_scan_exclusive_i64": (C.c_int, [vp, vp, i64, vp, vp]),
    "dax_hip_lower_bounds_i64": (C.c_int, [vp, i64, vp, i64, C.c_int, vp, vp, vp]),
    "dax_hip_upper_bounds_i64": (C.c_int, [vp, i64, vp, i64, C.c_int, vp, vp, vp]),
    "dax_hip_scan_inclusive_f32": (C.c_int, [vp, vp, i64, vp, vp]),
    "dax_hip_scan_exclusive_f32": (C.c_int, [vp, vp, i64, vp, vp]),
    "dax_hip_lower_bounds_f32": (C.c_int, [vp, i64, vp, i64, C.c_int, vp, vp, vp]),
    "dax_hip_upper_bounds_f32": (C.c_int, [vp, i64, vp, i64, C.c_int, vp, vp, vp]),
    "dax_hip_scan_inclusive_f64": (C.c_int, [vp, vp, i64, vp, vp]),
    "dax_hip_scan_exclusive_f64": (C.c_int, [vp, vp, i64, vp, vp]),
    "dax_hip_lower_bounds_f64": (C.c_int, [vp, i64, vp, i64, C.c_int, vp, vp, vp]),
    "dax_hip_upper_bounds_f64": (C.c_int, [vp, i64, vp, i64, C.c_int, vp, vp, vp]),
    "dax_hip_sort_i32": (C.c_int, [vp, i64, C.c_int, vp]),
    "dax_hip_sort_by_key_i32": (C.c_int, [vp, vp, sz, i64, C.c_int, vp]),
    "dax_hip_unique_i32": (C.c_int, [vp, i64, P_i64, vp]),
    "dax_hip_sort_i64": (C.c_int, [vp, i64, C.c_int, vp]),
    "dax_hip_sort_by_key_i64": (C.c_int, [vp, vp, sz, i64, C.c_int, vp]),
    "dax_hip_unique_i64": (C.c_int, [vp, i64, P_i64, vp]),
    "dax_hip_sort_u32": (C.c_int, [vp, i64, C.c_int, vp]),
    "dax_hip_sort_by_key_u32": (C.c_int, [vp, vp, sz, i64, C.c_int, vp]),
    "dax_hip_unique_u32": (C.c_int, [vp, i64, P_i64, vp]),
    "dax_hip_sort_u64": (C.c_int, [vp, i64, C.c_int, vp]),
    "dax_hip_sort_by_key_u64": (C.c_int, [vp, vp, sz, i64, C.c_int, vp]),
    "dax_hip_unique_u64": (C.c_int, [vp, i64, P_i64, vp]),
    "dax_hip_sort_f32": (C.c_int, [vp, i64, C.c_int, vp]),
    "dax_hip_sort_by_key_f32": (C.c_int, [vp, vp, sz, i64, C.c_int, vp]),
    "dax_hip_unique_f32": (C.c_int, [vp, i64, P_i64, vp]),
    "dax_hip_sort_f64": (C.c_int, [vp, i64, C.c_int, vp]),
    "dax_hip_sort_by_key_f64": (C.c_int, [vp, vp, sz, i64, C.c_int, vp]),
    "dax_hip_unique_f64": (C.c_int, [vp, i64, P_i64, vp]),
    "dax_hip_stream_compact": (C.c_int, [vp, sz, vp, C.c_int, i64, vp, vp, vp, P_i64, vp]),
    "dax_hip_timer_create": (C.c_int, [C.POINTER(vp)]),
    "dax_hip_timer_destroy": (C.c_int, [vp]),
    "dax_hip_timer_reset": (C.c_int, [vp, vp]),
    "dax_hip_timer_elapsed": (C.c_int, [vp, vp, C.POINTER(C.c_double)]),
    "dax_hip_mc_workspace_bytes": (sz, [i64, i64, i64]),
    "dax_hip_mc_forward": (C.c_int, [vp, i64, i64, i64, vp, vp, f32, C.c_uint, vp, sz,
                                     vp, vp, vp, vp, vp, vp, vp, i64, i64, vp, vp]),
    "dax_hip_mc_forward_noalloc": (C.c_int, [vp, i64, i64, i64, vp, vp, f32, C.c_uint,
                                             vp, vp, vp, vp, vp, vp, vp, i64, i64, vp, vp]),
    "dax_hip_mc_read_counts": (C.c_int, [vp, i64, i64, P_i64, P_i64, vp]),
    "dax_hip_mc_classify": (C.c_int, [vp, i64, i64, i64, f32, vp, sz, vp, vp]),
    "dax_hip_mc_slice_classify": (C.c_int, [i64, i64, i64, vp, vp, vp, vp, vp, sz, vp, vp]),
    "dax_hip_mt_count": (C.c_int, [vp, i64, vp, i64, f32, vp, vp, P_i64, vp]),
    "dax_hip_mt_scratch_bytes": (sz, [i64]),
    "dax_hip_mt_generate": (C.c_int, [vp, i64, vp, vp, i64, f32, C.c_uint, vp, vp, i64, vp, sz,
                                      vp, vp, vp, vp, vp, P_i64, vp]),
    "dax_hip_mc_hex_count": (C.c_int, [vp, i64, vp, i64, f32, vp, vp, P_i64, vp]),
    "dax_hip_mc_hex_scratch_bytes": (sz, [i64]),
    "dax_hip_mc_hex_generate": (C.c_int, [vp, i64, vp, vp, i64, f32, C.c_uint, vp, vp, i64, vp,
                                          sz, vp, vp, vp, vp, vp, P_i64, vp]),
    "dax_hip_threshold_count_uniform": (C.c_int, [vp, i64, i64, i64, f32, f32, vp, vp]),
    "dax_hip_threshold_count_cells": (C.c_int, [vp, C.c_int, i64, vp, i64, f32, f32, vp, vp]),
    "dax_hip_topology_workspace_bytes": (sz, [i64, i64]),
    "dax_hip_threshold_prepare_uniform": (C.c_int, [vp, i64, i64, i64, f32, f32, C.c_uint, vp, vp,
                                                    sz, vp, P_i64, P_i64, vp]),
    "dax_hip_topology_prepare_uniform": (C.c_int, [i64, i64, i64, vp, C.c_uint, vp, sz, vp, P_i64,
                                                   P_i64, vp]),
    "dax_hip_topology_generate_uniform": (C.c_int, [i64, i64, i64, vp, vp, vp, C.c_uint, vp, sz,
                                                    vp, vp, vp, vp]),
    "dax_hip_topology_prepare_cells": (C.c_int, [vp, C.c_int, i64, i64, vp, C.c_uint, vp, sz, vp,
                                                 P_i64, P_i64, vp]),
    "dax_hip_topology_generate_cells": (C.c_int, [vp, C.c_int, i64, vp, i64, vp, C.c_uint, vp, sz,
                                                  vp, vp, vp, vp]),
    "dax_hip_topology_compact_point_field": (C.c_int, [vp, sz, i64, i64, vp, vp, sz, vp, vp]),
    "dax_hip_write_vtk_polydata": (C.c_int, [C.c_char_p, vp, i64, vp, i64, C.c_int]),
    "dax_hip_write_vtk_unstructured": (C.c_int, [C.c_char_p, vp, i64, vp, i64, C.c_int,
                                                 C.c_int]),
    "dax_hip_mc_slice": (C.c_int, [i64, i64, i64, vp, vp, vp, vp, C.c_uint, vp, sz, vp, vp, vp,
                                   vp, vp, vp, vp, i64, i64, vp, vp]),
    "dax_hip_mc_bits": (C.c_int, [vp, i64, i64, i64, f32, vp, sz, vp]),
    "dax_hip_mc_scan": (C.c_int, [i64, i64, i64, vp, sz, vp]),
    "dax_hip_mc_generate": (C.c_int, [vp, i64, i64, i64, vp, vp, f32, C.c_uint, vp, sz,
                                      vp, vp, vp, vp, vp, vp, vp, i64, i64, vp, vp]),
    "dax_hip_mc_cell_outputs": (C.c_int, [i64, i64, i64, vp, sz, vp, vp, vp, vp]),
    "dax_hip_mc_interpolate_scalar": (C.c_int, [vp, vp, vp, vp, i64, vp, vp]),
    "dax_hip_mc_set_tuning": (C.c_int, [i64, i64, i64, vp]),
    "dax_hip_mc_get_tuning": (C.c_int, [i64, i64, i64, vp]),
    "dax_hip_mc_autotune": (C.c_int, [vp, i64, i64, i64, f32, C.c_uint, vp, sz, vp, vp, vp, vp,
                                      i64, i64, vp, C.c_int, vp, vp, vp]),
    "dax_hip_mc_tuning_save": (C.c_int, [C.c_char_p]),
    "dax_hip_mc_tuning_load": (C.c_int, [C.c_char_p]),
    "dax_hip_mc_batch_workspace_bytes": (sz, [i64, vp]),
    "dax_hip_mc_batch_header_bytes": (sz, []),
    "dax_hip_mc_batch_prepare": (C.c_int, [i64, vp, vp, vp, vp, vp, vp, sz, vp, vp]),
    "dax_hip_mc_batch_forward": (C.c_int, [vp, vp, vp, sz, C.c_uint, vp, vp, vp, vp, i64, i64,
                                           vp, vp, vp]),
    "dax_hip_mc_batch_graph_create": (C.c_int, [vp, vp, vp, sz, C.c_uint, vp, vp, vp, vp, i64,
                                                i64, vp, vp, vp, C.POINTER(vp)]),
    "dax_hip_mc_batch_backward": (C.c_int, [vp, vp, vp, sz, vp, vp, i64, vp, vp, vp, sz, vp, vp]),
    "dax_hip_mc_batch_set_tuning": (C.c_int, [vp, vp]),
    "dax_hip_mc_batch_get_tuning": (C.c_int, [vp, vp]),
    "dax_hip_mc_batch_autotune": (C.c_int, [vp, vp, vp, sz, C.c_uint, vp, vp, vp, vp, i64, i64,
                                            vp, C.c_int, vp, vp, vp]),
    "dax_hip_mc_volume_loop": (C.c_int, [i64, vp, vp, vp, vp, vp, vp, C.c_uint, vp, sz, vp, i64,
                                         i64, vp, vp, vp, vp, vp, vp, vp, vp, sz, vp, vp]),
    "dax_hip_graph_begin_capture": (C.c_int, [vp]),
    "dax_hip_graph_end_capture": (C.c_int, [vp, vp]),
    "dax_hip_graph_launch": (C.c_int, [vp, vp]),
    "dax_hip_graph_destroy": (C.c_int, [vp]),
    "dax_hip_mc_backward_scratch_bytes": (sz, [i64]),
    "dax_hip_mc_backward_set_span_threshold": (C.c_int, [i64]),
    "dax_hip_mc_backward": (C.c_int, [vp, i64, i64, i64, vp, vp, f32, vp, sz, vp, vp, i64,
                                      vp, vp, vp, sz, vp, vp]),
    "dax_hip_mc_interpolate_vec3": (C.c_int, [vp, vp, vp, vp, i64, vp, vp]),
    "dax_hip_point_gradient": (C.c_int, [vp, i64, i64, i64, vp, vp, vp]),
    "dax_hip_cell_gradient": (C.c_int, [vp, i64, i64, i64, vp, vp, vp, vp]),
}

# Optional groups of entry points, bound when the library exports them.
EXTRA_SIGNATURES: dict = {}


def library_path() -> Path:
    return _LIB_PATH


def load() -> C.CDLL:
    """Load libdaxhip.so (once) and bind the C-ABI signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise RuntimeError(
            f"daxtoolkit_amd: {_LIB_PATH} is not built; run "
            "`python -c 'import __graft_entry__ as g; g.build()'` first")
    lib = C.CDLL(str(_LIB_PATH))
    for name, (res, args) in {**_SIGNATURES, **EXTRA_SIGNATURES}.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols() -> list[str]:
    return sorted({**_SIGNATURES, **EXTRA_SIGNATURES})


def check(rc: int) -> None:
    if rc != OK:
        msg = load().dax_hip_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, DaxHipError)(rc, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))

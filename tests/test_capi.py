"""The C-ABI library loads and exports every symbol include/*.h declares
(no device calls; runs without a GPU)."""
import re
from pathlib import Path

import daxtoolkit_amd
from daxtoolkit_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_in_headers():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(dax_hip_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = daxtoolkit_amd.load_library()
    names = declared_in_headers()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    names = declared_in_headers()
    bound = set(_lib.declared_symbols())
    assert names == bound, (sorted(names - bound), sorted(bound - names))


def test_version_and_error_string():
    lib = daxtoolkit_amd.load_library()
    assert lib.dax_hip_version().startswith(b"daxhip")
    assert isinstance(lib.dax_hip_last_error(), bytes)


def test_workspace_size_is_host_only():
    lib = daxtoolkit_amd.load_library()
    assert lib.dax_hip_mc_workspace_bytes(0, 1, 1) == 0
    b = lib.dax_hip_mc_workspace_bytes(512, 512, 512)
    # bit words (P/8) + two 32-bit bases per word + tile states
    assert 512**3 // 8 <= b < 512**3 // 8 * 3


def test_span_threshold_setter_is_host_only():
    lib = daxtoolkit_amd.load_library()
    assert lib.dax_hip_mc_backward_set_span_threshold(-1) != 0
    assert b"negative" in lib.dax_hip_last_error()
    assert lib.dax_hip_mc_backward_set_span_threshold(0) == 0
    assert lib.dax_hip_mc_backward_set_span_threshold(16) == 0

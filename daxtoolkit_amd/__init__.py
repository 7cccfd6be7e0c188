"""daxtoolkit_amd -- MI355X (gfx950) device adapter for Dax's marching-cubes path.

Native library: ``libdaxhip.so`` (C-ABI in ``include/dax_hip.h``), built from
``daxtoolkit_amd/csrc``.  The C++ host API mirroring Dax (``include/dax``)
and this Python front end both call that C-ABI.
"""
from . import _lib  # noqa: F401
from ._lib import (DaxHipError, ErrorCapacity, ErrorControlBadValue,  # noqa: F401
                   ErrorControlOutOfMemory, ErrorExecution)

__all__ = ["_lib", "mc", "fields", "DaxHipError", "ErrorCapacity", "ErrorControlBadValue",
           "ErrorControlOutOfMemory", "ErrorExecution"]


def load_library():
    """Load libdaxhip.so; raises RuntimeError if it has not been built."""
    return _lib.load()

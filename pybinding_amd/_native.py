"""Loader for the engine's native libraries.

``_kpmx`` is the pybind11 host module (C++ core + HIP compute backend); ``libkpmx.so`` is the
HIP engine behind the C-ABI declared in ``include/kpmx.h``.  Both are built in-tree by
``__graft_entry__.build()`` (``make -C pybinding_amd/csrc``).  There is no fallback: if the
libraries are missing, importing the engine raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkpmx.so")


class NativeLibraryMissing(ImportError):
    pass


def load_module():
    """Import the compiled host module `_kpmx` (raises NativeLibraryMissing if it is not built)."""
    try:
        from . import _kpmx
    except ImportError as e:  # pragma: no cover - depends on the build state
        raise NativeLibraryMissing(
            "pybinding_amd: the native KPM engine is not built (run `make -C pybinding_amd/csrc` or "
            "`python -c 'import __graft_entry__ as g; g.build()'`): %s" % e) from e
    return _kpmx


def load_abi():
    """ctypes handle on libkpmx.so, for callers that bind the C-ABI directly."""
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing("pybinding_amd: %s is not built" % LIB_PATH)
    return ctypes.CDLL(LIB_PATH)

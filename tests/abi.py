"""Re-export of the engine's ctypes C-ABI binding for the tests."""
from pybinding_amd.abi import *  # noqa: F401,F403
from pybinding_amd.abi import LIB, HEADER, declared_functions, lib  # noqa: F401

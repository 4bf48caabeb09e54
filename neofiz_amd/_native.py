"""Loads the HIP library, libneofiz_hip.so, built in-tree under neofiz_amd/_build/.

There is no fallback: a missing or unloadable library raises, so no compute call can
silently run anywhere but on the GPU.
"""
import ctypes
import os

from ._abi import Binding

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(_HERE, "_build", "libneofiz_hip.so")
_binding = None


def library_path():
    return LIBRARY


def binding():
    global _binding
    if _binding is None:
        if not os.path.exists(LIBRARY):
            raise ImportError(
                f"{LIBRARY} is not built; run __graft_entry__.build() (hipcc, gfx950) first")
        _binding = Binding(ctypes.CDLL(LIBRARY), "nf_")
    return _binding

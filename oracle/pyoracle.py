"""TEST INFRASTRUCTURE -- Python binding of the oracle (libneofiz_oracle.so, prefix nfo_).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this, and only to
check results against. It exposes the same ExplicitSolver surface as the product over the oracle's
C ABI.
"""
import ctypes
import os
import subprocess

from neofiz_amd._abi import Binding
from neofiz_amd.solver import ExplicitSolver

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "_build", "libneofiz_oracle.so")
_binding = None


def build():
    """Brings the oracle up to date; concurrent test workers take turns (a file lock), so none
    loads the library while another is relinking it."""
    import fcntl
    with open(os.path.join(HERE, ".build.lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        subprocess.run(["make", "-s", "-C", HERE], check=True)


def binding():
    global _binding
    if _binding is None:
        if not os.path.exists(LIBRARY):
            build()
        _binding = Binding(ctypes.CDLL(LIBRARY), "nfo_")
    return _binding


def OracleSolver(*args, **kwargs):
    kwargs["binding"] = binding()
    return ExplicitSolver(*args, **kwargs)


def OracleDefectField(*args, **kwargs):
    """The oracle's DefectField (oracle/defect_field.cpp) behind the product's Python class."""
    from neofiz_amd.defects import DefectField
    kwargs["binding"] = binding()
    return DefectField(*args, **kwargs)

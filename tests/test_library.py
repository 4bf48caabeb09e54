"""The HIP library builds for gfx950, loads on any host, and exports every entry point the C
header declares (no compute call is made here: there may be no GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def declared_symbols():
    with open(os.path.join(ROOT, "include", "neofiz.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(nf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_solver_surface():
    names = declared_symbols()
    for must in ("nf_solver_create", "nf_run_until", "nf_set_rigid_wall", "nf_set_damage",
                 "nf_get_node_field", "nf_get_audits", "nf_point_path"):
        assert must in names


def test_hip_library_exports_every_declared_symbol():
    from neofiz_amd import _native
    path = _native.library_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_hip_library_targets_gfx950():
    from neofiz_amd import _native
    with open(_native.library_path(), "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob


# Entry points with no counterpart in the reference (timing, the copy-bandwidth probe, and
# the step loop's derivative, which is checked against finite differences of the GPU
# forward solver instead).
GPU_ONLY = ("nf_profile", "nf_profile_read", "nf_copy_bandwidth", "nf_enable_sensitivity",
            "nf_get_node_sensitivity", "nf_get_step_sensitivity", "nf_defect_sample_onto_device",
            "nf_contact_set_limits", "nf_defect_sample_onto_static", "nf_adjoint_enable",
            "nf_adjoint_gradient", "nf_adjoint_info")


def test_oracle_exports_the_same_surface(oracle):
    lib = oracle.binding().lib
    missing = [n for n in declared_symbols()
               if n not in GPU_ONLY and not hasattr(lib, "nfo_" + n[3:])]
    assert not missing, missing

"""Drives single material points through J2.update via nf_point_path / nfo_point_path."""
import ctypes as C

import numpy as np

from neofiz_amd import _abi

FIELDS = ["s_rr", "s_zz", "s_tt", "s_rz", "b_rr", "b_zz", "b_tt", "b_rz", "eps_p", "work",
          "dgamma", "strength", "softening", "width"]


def run(binding, material, increments, damage=None, formulation=0):
    """increments: (paths, steps, 5) {dRR, dZZ, dTT, dGamRZ, dt}; returns (paths, steps, 14)."""
    inc = np.ascontiguousarray(increments, dtype=np.float64)
    paths, steps = inc.shape[0], inc.shape[1]
    out = np.empty((paths, steps, 14))
    dm = None
    if damage is not None:
        dm = np.ascontiguousarray(damage, dtype=np.float64).reshape(paths, 7)
    mat = material.to_c()
    dp = _abi.c_double_p
    rc = binding.point_path(C.byref(mat), formulation, paths, steps, inc.ctypes.data_as(dp),
                            dm.ctypes.data_as(dp) if dm is not None else None,
                            out.ctypes.data_as(dp))
    if rc != 0:
        raise RuntimeError(binding.last_error().decode())
    return out


def mises(o):
    p = (o[..., 0] + o[..., 1] + o[..., 2]) / 3.0
    a, b, c, d = o[..., 0] - p, o[..., 1] - p, o[..., 2] - p, o[..., 3]
    return np.sqrt(1.5 * (a * a + b * b + c * c + 2.0 * d * d))


def pressure(o):
    return -(o[..., 0] + o[..., 1] + o[..., 2]) / 3.0

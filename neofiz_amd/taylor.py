"""The Taylor impact harness: validate/TaylorImpactCase.java:214-331 over this library.

`run` builds the solid-cylinder mesh, pins the axis, places the anvil, fires the rod and steps
`while time < duration`, then measures final length, mushroom diameter and the audits exactly
as the reference harness does. The solver class is a parameter so that the same harness drives
the HIP library and, in the tests, the oracle.

The two workloads below are the rod the benchmarks are quoted on: a 25.4 mm x 7.62 mm copper-like
rod at 190 m/s, axisymmetric, reduced integration with hourglass control, CFL 0.5, 80 us.
"""
import math
import time as _time
from dataclasses import dataclass

import numpy as np

from .solver import (AXISYMMETRIC, FINITE_STRAIN, REDUCED, ExplicitSolver, JohnsonCook,
                     Material, QuadMesh, RigidWall, IllegalArgumentException)

ROD_LENGTH = 25.4e-3
ROD_DIAMETER = 7.62e-3
ROD_SPEED = 190.0
ROD_DURATION = 80e-6

# Linear isotropic hardening rod: E 117 GPa, nu 0.35, rho 8930, sy 400 MPa, H 100 MPa.
ROD_LINEAR = Material("Cu (linear)", 117.0e9, 0.35, 8930.0).yielding(400.0e6, 100.0e6)
# The same elastic constants with the OFHC copper Johnson-Cook law
# (A 90 MPa, B 292 MPa, n 0.31, C 0.025, m 1.09, Tm 1356 K, beta 0.9).
ROD_JC = Material("Cu (Johnson-Cook)", 117.0e9, 0.35, 8930.0).withJohnsonCook(
    JohnsonCook.COPPER_OFHC)
ROD_JC_FRICTION = 0.09

# The reference harness's own case (validate/TaylorImpactCase.java:69-77): 4340 steel,
# 37.97 mm x 7.595 mm at 181 m/s, Johnson-Cook.
STEEL_LENGTH = 37.97e-3
STEEL_DIAMETER = 7.595e-3
STEEL_SPEED = 181.0


def johnsonCook4340():
    return Material.STEEL_4340.withJohnsonCook(JohnsonCook.STEEL_4340)


@dataclass
class Result:
    nr: int
    nz: int
    elementCount: int
    nodeCount: int
    steps: int
    time: float
    finalLength: float
    lengthRatio: float
    mushroomDiameter: float
    maxPlasticStrain: float
    momentumBalanceError: float
    energyBalanceError: float
    plasticFraction: float
    hourglassFraction: float
    wallClockSeconds: float
    solver: object = None


def java_round(x):
    """Math.round(double): floor(x + 0.5)."""
    return int(math.floor(x + 0.5))


def build(material, length, diameter, speed, nr, nz=None, integration=REDUCED, friction=0.0,
          solver_cls=ExplicitSolver):
    """Mesh, solver, axis pins, anvil and initial velocity (TaylorImpactCase.java:225-249).

    nz=None keeps elements square as the reference harness does (:227-230); the rod
    workloads pass nz explicitly (8 x 80 is not square).
    """
    if material.isElastic():
        raise IllegalArgumentException(
            "the Taylor test measures permanent deformation; the material must yield")
    radius = 0.5 * diameter
    if nz is None:
        h = radius / nr
        nz = max(1, java_round(length / h))
    mesh = QuadMesh.solidCylinder(radius, length, nr, nz)
    solver = solver_cls(mesh, material, AXISYMMETRIC, integration, FINITE_STRAIN, 1.0, 0.5)
    for n in QuadMesh.axisNodes(nr, nz):
        solver.fixRadial(n)
    solver.setRigidWall(RigidWall.atZ(0.0, friction))
    solver.setUniformVelocity(0.0, -speed)
    return mesh, solver, nz


def measure(mesh, solver, length, momentum0, kinetic0):
    """Length, mushroom and audits (TaylorImpactCase.java:264-311)."""
    ur = solver.radialDisplacement()
    uz = solver.axialDisplacement()
    zc = mesh.z + uz
    z_max = -math.inf
    z_min = math.inf
    for v in zc:  # Math.max / Math.min in node order
        z_max = max(z_max, v)
        z_min = min(z_min, v)
    final_length = z_max - z_min
    mushroom_radius = 0.0
    for v in mesh.r + ur:
        mushroom_radius = max(mushroom_radius, v)
    a = solver.audits()
    st = solver.state()
    momentum_error = abs((a.axial_momentum - momentum0) - st.wall_impulse) / abs(momentum0)
    accounted = (a.plastic_dissipation + a.kinetic_energy + a.strain_energy
                 + st.wall_energy_loss + a.hourglass_energy + st.wall_friction_dissipation)
    return dict(finalLength=final_length, lengthRatio=final_length / length,
                mushroomDiameter=2.0 * mushroom_radius, maxPlasticStrain=a.max_plastic_strain,
                momentumBalanceError=momentum_error, energyBalanceError=accounted / kinetic0 - 1.0,
                plasticFraction=a.plastic_dissipation / kinetic0,
                hourglassFraction=a.hourglass_energy / kinetic0)


def run(material, length, diameter, speed, nr, nz=None, integration=REDUCED, friction=0.0,
        duration=ROD_DURATION, solver_cls=ExplicitSolver, keep_solver=False):
    """TaylorImpactCase.run (validate/TaylorImpactCase.java:214-331)."""
    mesh, solver, nz = build(material, length, diameter, speed, nr, nz, integration, friction,
                             solver_cls)
    a0 = solver.audits()
    momentum0, kinetic0 = a0.axial_momentum, a0.kinetic_energy
    t0 = _time.perf_counter()
    steps = solver.runUntil(duration)
    wall = _time.perf_counter() - t0
    m = measure(mesh, solver, length, momentum0, kinetic0)
    st = solver.state()
    res = Result(nr=nr, nz=nz, elementCount=mesh.elementCount, nodeCount=mesh.nodeCount,
                 steps=steps, time=st.time, wallClockSeconds=wall,
                 solver=solver if keep_solver else None, **m)
    if not keep_solver:
        solver.shutdown()
    return res


def rod(nr, nz, johnson_cook=True, solver_cls=ExplicitSolver, duration=ROD_DURATION,
        keep_solver=False):
    """The benchmark rod at nr x nz: Johnson-Cook with mu = 0.09, or linear frictionless."""
    if johnson_cook:
        return run(ROD_JC, ROD_LENGTH, ROD_DIAMETER, ROD_SPEED, nr, nz, REDUCED,
                   ROD_JC_FRICTION, duration, solver_cls, keep_solver)
    return run(ROD_LINEAR, ROD_LENGTH, ROD_DIAMETER, ROD_SPEED, nr, nz, REDUCED, 0.0,
               duration, solver_cls, keep_solver)

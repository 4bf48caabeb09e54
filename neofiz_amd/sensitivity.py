"""Gradients of the Taylor rod's final length and mushroom diameter with respect to the
anvil friction and the Johnson-Cook A, B, n, C, m.

The derivative is the solver's own (ExplicitSolver.enableSensitivity): the exact derivative
of the discrete step loop, carried forward in six directions alongside the state, so one
run gives all twelve partial derivatives. The objectives follow TaylorImpactCase.java:
272-311: L/L0 from the axial extent of the nodes, the mushroom diameter from the largest
radius, each extreme taken at the first node that attains it (Math.max in node order); a
max's derivative is its selected node's.

Checking: `finite_differences` reruns the forward solver at p (1 +- h) for each parameter
(as one batch of twelve rods) and takes central differences with the parameter steps that
were actually represented.
"""
import math
import time
from dataclasses import dataclass, replace

import numpy as np

from . import taylor
from .solver import Batch, ExplicitSolver

PARAMETERS = ExplicitSolver.SENSITIVITY_PARAMETERS


@dataclass(frozen=True)
class Point:
    """One parameter point: mesh, impact speed, friction and the Johnson-Cook constants."""
    nr: int
    nz: int
    speed: float
    friction: float
    a: float
    b: float
    n: float
    c: float
    m: float

    def values(self):
        return (self.friction, self.a, self.b, self.n, self.c, self.m)

    def with_value(self, k, v):
        return replace(self, **{("friction", "a", "b", "n", "c", "m")[k]: v})


def from_instance(inst):
    """A point from an ensemble draw (the ensemble module's JC scaling)."""
    jc = taylor.ROD_JC.johnsonCook
    sa, sb, sn, sc, sm = inst.scale
    return Point(inst.nr, inst.nz, inst.speed, inst.friction, jc.a * sa, jc.b * sb, jc.n * sn,
                 jc.c * sc, jc.m * sm)


def build(point, solver_cls=ExplicitSolver):
    jc = taylor.ROD_JC.johnsonCook
    law = replace(jc, a=point.a, b=point.b, n=point.n, c=point.c, m=point.m)
    mat = taylor.ROD_JC.withJohnsonCook(law)
    mesh, solver, _ = taylor.build(mat, taylor.ROD_LENGTH, taylor.ROD_DIAMETER, point.speed,
                                   point.nr, point.nz, friction=point.friction,
                                   solver_cls=solver_cls)
    return mesh, solver


def objectives(mesh, solver):
    """(L/L0, mushroom diameter) and the nodes that set them: (i_zmax, i_zmin, i_rmax)."""
    zc = mesh.z + solver.axialDisplacement()
    rc = mesh.r + solver.radialDisplacement()
    i_hi = i_lo = 0
    for i in range(1, zc.size):  # Math.max / Math.min keep the first extreme
        if zc[i] > zc[i_hi]:
            i_hi = i
        if zc[i] < zc[i_lo]:
            i_lo = i
    i_r = 0
    for i in range(1, rc.size):
        if rc[i] > rc[i_r]:
            i_r = i
    return ((zc[i_hi] - zc[i_lo]) / taylor.ROD_LENGTH, 2.0 * rc[i_r]), (i_hi, i_lo, i_r)


def gradients(point, end_time=taylor.ROD_DURATION, max_steps=0):
    """Runs one rod with sensitivities on. Returns (objectives, gradient [2][6], steps,
    seconds of the run)."""
    mesh, solver = build(point)
    solver.enableSensitivity()
    t0 = time.perf_counter()
    solver.runUntil(end_time, max_steps)
    seconds = time.perf_counter() - t0
    obj, (i_hi, i_lo, i_r) = objectives(mesh, solver)
    grad = np.zeros((2, len(PARAMETERS)))
    for k in range(len(PARAMETERS)):
        uz = solver.nodeSensitivity("uz", k)
        ur = solver.nodeSensitivity("ur", k)
        grad[0, k] = (uz[i_hi] - uz[i_lo]) / taylor.ROD_LENGTH
        grad[1, k] = 2.0 * ur[i_r]
    steps = solver.steps()
    solver.shutdown()
    return obj, grad, steps, seconds


def forward(point, end_time=taylor.ROD_DURATION, max_steps=0):
    """The plain forward run of the same rod: (objectives, steps, seconds)."""
    mesh, solver = build(point)
    t0 = time.perf_counter()
    solver.runUntil(end_time, max_steps)
    seconds = time.perf_counter() - t0
    obj, _ = objectives(mesh, solver)
    steps = solver.steps()
    solver.shutdown()
    return obj, steps, seconds


def finite_differences(point, end_time=taylor.ROD_DURATION, max_steps=0, rel=1e-6):
    """Central differences of the forward solver, all twelve perturbed rods in one batch.
    Returns (gradient [2][6], step counts of the perturbed runs)."""
    pts = []
    steps = []
    for k, v in enumerate(point.values()):
        h = rel * abs(v) if v != 0.0 else rel
        pts.append(point.with_value(k, v + h))
        pts.append(point.with_value(k, v - h))
    built = [build(p) for p in pts]
    batch = Batch([s for _, s in built])
    batch.runUntil(end_time, max_steps)
    batch.scatter()
    batch.close()
    grad = np.zeros((2, len(PARAMETERS)))
    for k in range(len(PARAMETERS)):
        (mp, sp), (mm, sm) = built[2 * k], built[2 * k + 1]
        op, _ = objectives(mp, sp)
        om, _ = objectives(mm, sm)
        dp = pts[2 * k].values()[k] - pts[2 * k + 1].values()[k]
        for j in range(2):
            grad[j, k] = (op[j] - om[j]) / dp
        steps += [sp.steps(), sm.steps()]
    for _, s in built:
        s.shutdown()
    return grad, steps

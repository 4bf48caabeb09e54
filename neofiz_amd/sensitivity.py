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


def objective_seeds(mesh, extremes):
    """d(L/L0, mushroom diameter)/d(UR, UZ, VR, VZ) at the final state, [2][4][nodes], for
    the extreme nodes (i_zmax, i_zmin, i_rmax) that set them."""
    i_hi, i_lo, i_r = extremes
    seeds = np.zeros((2, 4, mesh.nodeCount))
    seeds[0, 1, i_hi] += 1.0 / taylor.ROD_LENGTH
    seeds[0, 1, i_lo] -= 1.0 / taylor.ROD_LENGTH
    seeds[1, 0, i_r] = 2.0
    return seeds


def adjoint_gradients(point, end_time=taylor.ROD_DURATION, max_steps=0, interval=64):
    """The rod run with the discrete adjoint: a forward run recording checkpoints every
    `interval` steps, then one reverse sweep carrying both objectives. Returns (objectives,
    gradient [2][6], steps, forward seconds, reverse seconds, adjoint info)."""
    mesh, solver = build(point)
    solver.enableAdjoint(interval)
    t0 = time.perf_counter()
    solver.runUntil(end_time, max_steps)
    fwd = time.perf_counter() - t0
    obj, extremes = objectives(mesh, solver)
    t0 = time.perf_counter()
    grad = solver.adjointGradient(objective_seeds(mesh, extremes))
    rev = time.perf_counter() - t0
    steps = solver.steps()
    info = solver.adjointInfo()
    solver.shutdown()
    return obj, grad, steps, fwd, rev, info


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


def study(points=16, seed=4, nr=32, nz=320, end_time=taylor.ROD_DURATION, max_steps=20000, rel=1e-6):
    """The derivative against central differences on seeded parameter points drawn as the
    ensemble draws them (mesh fixed at nr x nz), with the objectives' extreme nodes held at
    the unperturbed run's. For each point: the worst and median relative disagreement over the
    twelve derivatives, the cost ratio, and the asymmetry of the one-sided differences of the
    nodal positions, max|(u+ - u0) + (u- - u0)| / max|u+ - u-| over the nodes: of order h on a
    smooth branch, of order one where a perturbed run crossed a discrete event (a node landing on the anvil a step
    earlier or later zeroes its velocity, a jump of the discrete map), across which a central
    difference is not a derivative."""
    from . import ensemble
    rows = []
    for k, inst in enumerate(ensemble.draw(points, seed=seed)):
        p0 = from_instance(inst)
        pt = Point(nr, nz, p0.speed, p0.friction, p0.a, p0.b, p0.n, p0.c, p0.m)
        mesh, s = build(pt)
        s.enableSensitivity()
        t0 = time.perf_counter()
        s.runUntil(end_time, max_steps)
        sec = time.perf_counter() - t0
        _, (ih, il, ir) = objectives(mesh, s)
        tang = np.array([[(s.nodeSensitivity("uz", q)[ih] - s.nodeSensitivity("uz", q)[il]) / taylor.ROD_LENGTH
                          for q in range(len(PARAMETERS))],
                         [2.0 * s.nodeSensitivity("ur", q)[ir] for q in range(len(PARAMETERS))]])
        z0 = mesh.z + s.axialDisplacement()
        r0 = mesh.r + s.radialDisplacement()
        steps = s.steps()
        s.shutdown()
        _, _, fsec = forward(pt, end_time, max_steps)
        pts = []
        for q, v in enumerate(pt.values()):
            h = rel * abs(v) if v != 0.0 else rel
            pts += [pt.with_value(q, v + h), pt.with_value(q, v - h)]
        built = [build(p) for p in pts]
        batch = Batch([x for _, x in built])
        batch.runUntil(end_time, max_steps)
        batch.scatter()
        batch.close()
        fd = np.zeros_like(tang)
        asym = 0.0
        for q in range(len(PARAMETERS)):
            (mp, sp), (mm, sm) = built[2 * q], built[2 * q + 1]
            dp = pts[2 * q].values()[q] - pts[2 * q + 1].values()[q]
            zp, zm = mesh.z + sp.axialDisplacement(), mesh.z + sm.axialDisplacement()
            rp, rm = mesh.r + sp.radialDisplacement(), mesh.r + sm.radialDisplacement()
            fd[0, q] = ((zp[ih] - zp[il]) - (zm[ih] - zm[il])) / taylor.ROD_LENGTH / dp
            fd[1, q] = 2.0 * (rp[ir] - rm[ir]) / dp
            # Symmetry of the two one-sided differences over every node: second order in h
            # on a smooth branch, of the order of the difference itself across a jump.
            for up, um, u0 in ((zp, zm, z0), (rp, rm, r0)):
                spread = float(np.abs(up - um).max())
                if spread > 0.0:
                    asym = max(asym, float(np.abs((up - u0) + (um - u0)).max()) / spread)
        for _, x in built:
            x.shutdown()
        err = np.abs(tang - fd) / np.maximum(np.abs(fd), 1e-300)
        rows.append({"point": k, "speed": pt.speed, "friction": pt.friction, "steps": steps,
                     "max_rel": float(err.max()), "median_rel": float(np.median(err)),
                     "asymmetry": asym, "cost_ratio": sec / fsec})
    return rows


if __name__ == "__main__":
    import json
    for row in study():
        print(json.dumps(row), flush=True)

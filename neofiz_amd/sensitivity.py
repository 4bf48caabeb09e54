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


def central_differences(point, steps, extremes, rel=1e-6):
    """Central differences of the forward solver at a fixed step count (the derivative the
    adjoint and the tangent compute is that of the state after `steps` steps; a perturbed run
    stopped by the end time instead may take a step more or less, a jump of the map). All
    twelve perturbed rods run as one batch. Returns (fd [2][6], objectives of the + and - runs
    [2][2][6])."""
    ih, il, ir = extremes
    pts = []
    for q, v in enumerate(point.values()):
        h = rel * abs(v) if v != 0.0 else rel
        pts += [point.with_value(q, v + h), point.with_value(q, v - h)]
    built = [build(p) for p in pts]
    batch = Batch([x for _, x in built])
    batch.runUntil(1.0, steps)  # the step count binds, not the time
    batch.scatter()
    batch.close()
    fd = np.zeros((2, len(PARAMETERS)))
    obj = np.zeros((2, 2, len(PARAMETERS)))
    for q in range(len(PARAMETERS)):
        dp = pts[2 * q].values()[q] - pts[2 * q + 1].values()[q]
        for side in range(2):
            mesh, sv = built[2 * q + side]
            assert sv.steps() == steps
            z = mesh.z + sv.axialDisplacement()
            r = mesh.r + sv.radialDisplacement()
            obj[side, 0, q] = (z[ih] - z[il]) / taylor.ROD_LENGTH
            obj[side, 1, q] = 2.0 * r[ir]
        fd[:, q] = (obj[0, :, q] - obj[1, :, q]) / dp
    for _, x in built:
        x.shutdown()
    return fd, obj


def study(points=16, seed=4, nr=32, nz=320, end_time=taylor.ROD_DURATION, max_steps=20000, rel=1e-6):
    """The discrete adjoint against central differences on seeded parameter points drawn as
    the ensemble draws them (mesh fixed at nr x nz), the differences at the adjoint run's step
    count with the objectives' extreme nodes held at the unperturbed run's. For each point and
    each of the twelve derivatives: the relative disagreement and the asymmetry of the two
    one-sided differences of the objective, |J+ + J- - 2 J0| / |J+ - J-|, which is of order h
    where the map is smooth on the scale of the step and of order one where a perturbed run
    crossed a discrete event (a node landing, leaving the anvil or switching between stick and
    slip a step earlier or later; a point yielding or not); and the difference's resolution,
    |FD(h) - FD(h/2)| / |FD(h)|, how far the central difference itself moves when its step is
    halved (of order h^2 on a smooth map, the map's roughness at that scale otherwise)."""
    from . import ensemble
    rows = []
    for k, inst in enumerate(ensemble.draw(points, seed=seed)):
        p0 = from_instance(inst)
        pt = Point(nr, nz, p0.speed, p0.friction, p0.a, p0.b, p0.n, p0.c, p0.m)
        mesh, s = build(pt)
        s.enableAdjoint(64)
        t0 = time.perf_counter()
        s.runUntil(end_time, max_steps)
        rec = time.perf_counter() - t0
        obj0, ext = objectives(mesh, s)
        t0 = time.perf_counter()
        grad = s.adjointGradient(objective_seeds(mesh, ext))
        rev = time.perf_counter() - t0
        steps = s.steps()
        s.shutdown()
        _, _, fsec = forward(pt, end_time, max_steps)
        j0 = np.array(obj0)[:, None]
        diffs = []
        for h in (rel, 0.5 * rel):
            fd, obj = central_differences(pt, steps, ext, h)
            spread = np.abs(obj[0] - obj[1])
            diffs.append((fd, np.abs(obj[0] + obj[1] - 2.0 * j0) / np.maximum(spread, 1e-300)))
        (fd, asym), (fd2, asym2) = diffs
        err = np.abs(grad - fd) / np.maximum(np.abs(fd), 1e-300)
        # the difference's own resolution: how far it moves when the step is halved
        resolution = np.abs(fd - fd2) / np.maximum(np.abs(fd), 1e-300)
        rows.append({"point": k, "speed": pt.speed, "friction": pt.friction, "steps": steps,
                     "grad": grad.tolist(), "fd": fd.tolist(), "fd_half_step": fd2.tolist(),
                     "rel_err": err.tolist(), "resolution": resolution.tolist(),
                     "asymmetry": np.maximum(asym, asym2).tolist(),
                     "cost_ratio": (rec + rev) / fsec})
    return rows


if __name__ == "__main__":
    import json
    for row in study():
        print(json.dumps(row), flush=True)

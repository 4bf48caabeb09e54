"""The batched Taylor ensemble: many rods of uneven meshes and parameters on one GPU.

The reference runs ensembles as independent solvers, one per thread (M0Cylinder.java:870-907,
sweep/Sweep.java:77-106); here they are one `Batch`, stepped together with a step size per
instance. This module draws the ensemble, builds its solvers through the Taylor harness,
tunes the batch's tile sizes, and measures batched against one-by-one.

Ensemble draw (a seeded numpy PCG64 generator; the reference has no generator for this):
mesh uniformly from {8x80, 16x160, 32x320, 64x640}, impact speed U[150, 250] m/s, anvil
friction U[0, 0.2], and each of the Johnson-Cook A, B, n, C, m scaled by U[0.9, 1.1]
independently; everything else is the copper rod of the Johnson-Cook Taylor case.
"""
import time
from dataclasses import dataclass, replace

import numpy as np

from . import taylor
from .solver import Batch, ExplicitSolver

MESHES = ((8, 80), (16, 160), (32, 320), (64, 640))
TILES = (64, 128, 256)


@dataclass(frozen=True)
class Instance:
    nr: int
    nz: int
    speed: float
    friction: float
    scale: tuple  # multipliers of A, B, n, C, m

    @property
    def elements(self):
        return self.nr * self.nz


def draw(count=256, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nr, nz = MESHES[int(rng.integers(0, len(MESHES)))]
        speed = float(rng.uniform(150.0, 250.0))
        mu = float(rng.uniform(0.0, 0.2))
        scale = tuple(float(x) for x in rng.uniform(0.9, 1.1, 5))
        out.append(Instance(nr, nz, speed, mu, scale))
    return out


def material(inst):
    jc = taylor.ROD_JC.johnsonCook
    sa, sb, sn, sc, sm = inst.scale
    law = replace(jc, a=jc.a * sa, b=jc.b * sb, n=jc.n * sn, c=jc.c * sc, m=jc.m * sm)
    return taylor.ROD_JC.withJohnsonCook(law)


def build(inst, solver_cls=ExplicitSolver):
    """One rod of the ensemble through the Taylor harness (TaylorImpactCase.java:225-249)."""
    mesh, solver, _ = taylor.build(material(inst), taylor.ROD_LENGTH, taylor.ROD_DIAMETER,
                                   inst.speed, inst.nr, inst.nz, friction=inst.friction,
                                   solver_cls=solver_cls)
    return mesh, solver


def results(mesh, solver):
    """L/L0 and mushroom diameter of a finished rod (TaylorImpactCase.java:272-311)."""
    zc = mesh.z + solver.axialDisplacement()
    rc = mesh.r + solver.radialDisplacement()
    z_max, z_min, radius = -np.inf, np.inf, 0.0
    for v in zc:  # Math.max / Math.min in node order
        z_max = max(z_max, v)
        z_min = min(z_min, v)
    for v in rc:
        radius = max(radius, v)
    return (z_max - z_min) / taylor.ROD_LENGTH, 2.0 * radius


def predicted_cost(inst):
    """Element-steps to the end time, up to a factor common to the ensemble: the stable step
    scales with the cell size, the rod's diameter over nr, so the step count scales with nr."""
    return inst.elements * inst.nr


def shard(instances, devices):
    """Assigns each instance to one of `devices` GPUs, which then run their shares as
    independent batches with no communication: costliest first (predicted_cost), each to the
    device with the least cost so far, ties to the lowest device. Returns one ascending list
    of instance indices per device. The heaviest device carries at most the mean load plus
    one instance (the longest-processing-time bound)."""
    if devices < 1:
        raise ValueError("at least one device")
    load = [0] * devices
    out = [[] for _ in range(devices)]
    order = sorted(range(len(instances)), key=lambda i: (-predicted_cost(instances[i]), i))
    for i in order:
        d = min(range(devices), key=lambda k: (load[k], k))
        load[d] += predicted_cost(instances[i])
        out[d].append(i)
    return [sorted(x) for x in out]


def tune(batch, instances, end_time, trial_steps=24):
    """The batch's autotuner: instances are bucketed by mesh size, and each bucket's tile
    size (threads per tile, so how many of its tiles share a workgroup) is chosen in turn
    from TILES by timing `trial_steps` real steps of the whole batch with every other bucket
    held at its best so far. The trial steps are part of the run (they advance every
    instance exactly as any other steps would). Returns ({elements: tile}, timings, steps
    taken by every instance still running)."""
    buckets = sorted({i.elements for i in instances})
    choice = {b: 256 for b in buckets}

    def apply():
        for k, inst in enumerate(instances):
            batch.setTile(k, choice[inst.elements])

    timings = {}
    taken = 0
    for b in buckets:
        best, best_t = choice[b], None
        for ts in TILES:
            choice[b] = ts
            apply()
            batch.runUntil(end_time, 2)  # capture and warm the new tables
            t0 = time.perf_counter()
            batch.runUntil(end_time, trial_steps)
            t = time.perf_counter() - t0
            taken += 2 + trial_steps
            timings[(b, ts)] = t
            if best_t is None or t < best_t:
                best, best_t = ts, t
        choice[b] = best
    apply()
    return choice, timings, taken


def run_batched(instances, end_time=taylor.ROD_DURATION, max_steps=0, autotune=True):
    """Builds every instance and runs them as one batch (each while its time is below
    end_time, for at most max_steps steps when max_steps > 0). Returns (meshes, solvers,
    seconds, tile choice); the solvers hold the final states. The timed region includes the
    tuner's steps, which are part of the run."""
    built = [build(i) for i in instances]
    meshes = [m for m, _ in built]
    solvers = [s for _, s in built]
    batch = Batch(solvers)
    t0 = time.perf_counter()
    choice = None
    taken = 0
    if autotune:
        choice, _, taken = tune(batch, instances, end_time)
    if max_steps <= 0:
        batch.runUntil(end_time, 0)  # no cap: every instance to end_time
    elif max_steps > taken:
        batch.runUntil(end_time, max_steps - taken)
    seconds = time.perf_counter() - t0
    batch.scatter()
    batch.close()
    return meshes, solvers, seconds, choice


def run_one_by_one(instances, end_time=taylor.ROD_DURATION, max_steps=0):
    """The same instances through the single-instance GPU path, one after another. Returns
    per instance (steps, seconds, (L/L0, mushroom diameter))."""
    out = []
    for inst in instances:
        mesh, solver = build(inst)
        t0 = time.perf_counter()
        solver.runUntil(end_time, max_steps)
        seconds = time.perf_counter() - t0
        out.append((solver.steps(), seconds, results(mesh, solver)))
        solver.shutdown()
    return out

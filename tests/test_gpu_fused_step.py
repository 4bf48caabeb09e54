"""The fused step kernel (one launch per step) against the two-launch step, on the GPU.

Without penalty contact, erosion or pressure edges the library enqueues each step after a run
call's first as one launch: the node update that opens the step, computed per patch of 64
elements with the patch's border nodes recomputed by each neighbour, followed by the element
pass on corners staged in LDS. The two-launch step (NEOFIZ_FUSED=0 at solver creation) is the
same arithmetic in another order of launches, so every state array must be bit-identical; the
wall audits, which the two forms sum over different workgroups, agree to rounding. Run calls
are split at awkward places (one step, two steps, an open-ended run) because each call opens
and closes a run of fused launches. The oracle parity tests run the same configurations
through the fused form (it is the default) against the oracle.
"""
import math
import os

import numpy as np
import pytest

from conftest import first_difference
from neofiz_amd import taylor
from neofiz_amd.solver import (AXISYMMETRIC, FINITE_STRAIN, FULL, PLANE_STRAIN, REDUCED,
                               SMALL_STRAIN, ExplicitSolver, Material, QuadMesh)

pytestmark = pytest.mark.gpu

NODE = ("ur", "uz", "vr", "vz", "fr", "fz", "mass", "on_wall")
POINT = ("stress", "back", "plastic_strain", "plastic_work")
DAMAGE = ("damage_strength", "damage_softening", "damage_width", "damage")


def build(fn, fused):
    old = os.environ.get("NEOFIZ_FUSED")
    os.environ["NEOFIZ_FUSED"] = "1" if fused else "0"
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["NEOFIZ_FUSED"]
        else:
            os.environ["NEOFIZ_FUSED"] = old


def compare(a, b, label):
    for f in NODE:
        d = first_difference(a.nodeField(f), b.nodeField(f))
        assert d is None, f"{label}: {f}: {d}"
    for f in POINT + (DAMAGE if getattr(a, "_damaged", False) else ()):
        d = first_difference(a.pointField(f), b.pointField(f))
        assert d is None, f"{label}: {f}: {d}"
    sa, sb = a.state(), b.state()
    assert (sa.steps, sa.time, sa.dt) == (sb.steps, sb.time, sb.dt), label
    assert sa.wall_contact_count == sb.wall_contact_count, label
    assert sa.wall_stuck_count == sb.wall_stuck_count, label
    for f in ("wall_impulse", "wall_energy_loss", "wall_friction_dissipation"):
        x, y = getattr(sa, f), getattr(sb, f)
        assert abs(x - y) <= 1e-12 * max(abs(y), 1e-300), (label, f, x, y)


def rod(nr, nz, friction, integration=REDUCED):
    def make():
        _, s, _ = taylor.build(taylor.ROD_JC, taylor.ROD_LENGTH, taylor.ROD_DIAMETER, taylor.ROD_SPEED,
                               nr, nz, integration, friction, ExplicitSolver)
        return s
    return make


def block(integration, kinematics, formulation, kinematic_hardening=False, gravity=False, damping=False):
    def make():
        mesh = QuadMesh.solidCylinder(0.01, 0.02, 5, 10)
        mat = Material("steel", 200e9, 0.3, 7800.0).yielding(3e8, 1e9, 5e8 if kinematic_hardening else 0.0)
        s = ExplicitSolver(mesh, mat, formulation, integration, kinematics, 0.01, 0.5)
        if gravity:
            s.setGravity(0.0, -9.81)
        if damping:
            s.setRelaxationDamping(0.05, 2e5)
        rng = np.random.default_rng(7)
        s.putNodeField("vr", rng.uniform(-50, 50, mesh.nodeCount))
        s.putNodeField("vz", rng.uniform(-50, 50, mesh.nodeCount))
        return s
    return make


def fan(integration):
    """Six quads around one centre node (six force slots there): the fused step's node-major
    slots hold at most four a node, so this mesh runs the two-launch step even when the fused
    one is asked for."""
    def make():
        ring = 6
        r, z = [0.05], [0.05]
        for k in range(ring):
            a = 2 * math.pi * k / ring
            r += [0.05 + 0.01 * math.cos(a), 0.05 + 0.012 * math.cos(a + math.pi / ring)]
            z += [0.05 + 0.01 * math.sin(a), 0.05 + 0.012 * math.sin(a + math.pi / ring)]
        conn = []
        for k in range(ring):
            spoke, mid, nxt = 1 + 2 * k, 2 + 2 * k, 1 + 2 * ((k + 1) % ring)
            conn += [0, spoke, mid, nxt]
        mesh = QuadMesh(np.array(r), np.array(z), np.array(conn, dtype=np.int32))
        mat = Material("steel", 200e9, 0.3, 7800.0).yielding(3e8, 1e9, 0.0)
        s = ExplicitSolver(mesh, mat, PLANE_STRAIN, integration, FINITE_STRAIN, 0.01, 0.5)
        rng = np.random.default_rng(11)
        s.putNodeField("vr", rng.uniform(-50, 50, mesh.nodeCount))
        s.putNodeField("vz", rng.uniform(-50, 50, mesh.nodeCount))
        return s
    return make


def damaged_rod(integration):
    """The Johnson-Cook rod with damage and no erosion (the fused step carries damage; erosion
    leaves the solver on the two-launch step): failure strains low enough that points soften
    within the run."""
    def make():
        mesh, s, _ = taylor.build(taylor.ROD_JC, taylor.ROD_LENGTH, taylor.ROD_DIAMETER, taylor.ROD_SPEED,
                                  8, 80, integration, 0.09, ExplicitSolver)
        rng = np.random.default_rng(3)
        s.setDamage(np.full(mesh.elementCount, 2.0e3), rng.uniform(0.02, 0.08, mesh.elementCount))
        s._damaged = True
        return s
    return make


CASES = {
    "jc-rod-friction": rod(16, 160, 0.09),
    "jc-rod-welded": rod(8, 80, math.inf),
    "jc-rod-full": rod(8, 80, 0.2, FULL),
    "kinematic-gravity-damping": block(REDUCED, FINITE_STRAIN, AXISYMMETRIC, True, True, True),
    "plane-strain-full-small": block(FULL, SMALL_STRAIN, PLANE_STRAIN),
    "fan-six-slot-node": fan(REDUCED),
    "jc-rod-damage": damaged_rod(REDUCED),
    "jc-rod-damage-full": damaged_rod(FULL),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_fused_equals_two_launch(case):
    make = CASES[case]
    fused, plain = build(make, True), build(make, False)
    done = 0
    for n in (1, 2, 1, 37, 300):  # the first call has no fused launch, the second one
        fused.run(n)
        plain.run(n)
        done += n
        compare(fused, plain, f"{case} after {done} steps")
    t = fused.state().time + 40 * fused.state().dt
    a, b = fused.runUntil(t), plain.runUntil(t)  # open-ended: blocks of look-ahead, no-op tails
    assert a == b
    compare(fused, plain, f"{case} after runUntil")
    fused.shutdown()
    plain.shutdown()

"""Pins for the follower pressure edges and the relaxation damping, on the oracle (and, marked
gpu, the HIP library bit for bit against it).

1. test:validate/ThickWallCylinderTest.java restated in full over validate/
   ThickWallCylinderCase.java:71-147: a thick-walled cylinder (A = 25 mm, B = 30 mm,
   P = 50 MPa, AISI 4340) loaded through pressure edges on its bore with a smoothstep ramp and
   relaxed by mass-proportional damping to a static state, against Lame's closed form
   (validate/Lame.java:23-64) -- both end conditions under both quadrature rules.
2. Closed forms that fix what the Lame gate cannot see:
   - the axisymmetric edge weights 2 pi L (ra/3 + rb/6, ra/6 + rb/3): the bore edges have
     ra = rb, so a swap of the two weights would pass the gate. On a pressurised end face the
     first moment of the nodal forces, sum F_i r_i = -2 pi p integral r^2 dr = -2 pi p R^3/3,
     distinguishes them (a swap gives the mean of r^2 weighted the wrong way);
   - the direction of the normal (the total force sign) and the plane-strain halves;
   - the damped central difference v' = decay v + gain f/m with half = damping dtAvg / 2:
     under gravity from rest every node follows v_n = -(g/c)(1 - decay^n), whose limit is
     the continuous terminal speed g/c, so a wrong half or gain changes every step.
"""
import math

import numpy as np
import pytest

import thick_wall as T
from neofiz_amd.mesh import Outline
from neofiz_amd.solver import (AXISYMMETRIC, ExplicitSolver, FULL, Material, PLANE_STRAIN,
                               QuadMesh, REDUCED, SMALL_STRAIN)

from conftest import first_difference, same_bits

A, B, P = 0.025, 0.030, 50.0e6
STEEL = Material.STEEL_4340
RULES = [pytest.param(FULL, id="FULL"), pytest.param(REDUCED, id="REDUCED")]
ENDS = [T.RESTRAINED, T.OPEN]


@pytest.fixture
def cls(oracle):
    return oracle.OracleSolver


# ------------------------------------------------------------------ ThickWallCylinderTest

def test_lame_satisfies_its_boundary_conditions():
    """lameBoundaryConditions (:36-46)."""
    assert abs(T.radial_stress(A, A, B, P) + P) <= P * 1e-12
    assert abs(T.radial_stress(B, A, B, P)) <= P * 1e-12
    assert T.hoop_stress(A, A, B, P) > T.hoop_stress(B, A, B, P)


def test_end_conditions_share_stresses_but_not_displacements():
    """endConditionsShareStressesButNotDisplacements (:48-54)."""
    strain = T.radial_displacement_plane_strain(A, A, B, P, 205e9, 0.29)
    stress = T.radial_displacement_plane_stress(A, A, B, P, 205e9, 0.29)
    assert abs(stress - strain) / strain > 0.05


@pytest.mark.parametrize("rule", RULES)
@pytest.mark.parametrize("end", ENDS)
def test_matches_lame_at_production_resolution(cls, end, rule):
    """GATE 1 (:62-80): settled to KE/SE < 1e-9, displacement and hoop stress within 1 %."""
    r = T.run(cls, STEEL, A, B, P, 10, end, rule)
    assert r.kineticOverStrainEnergy < 1e-9
    assert r.maxDisplacementErrorPercent < 1.0
    assert r.maxHoopStressErrorPercent < 1.0


@pytest.mark.parametrize("end", ENDS)
def test_reduced_tracks_full_integration(cls, end):
    """reducedTracksFullIntegration (:82-96)."""
    full = T.run(cls, STEEL, A, B, P, 10, end, FULL)
    reduced = T.run(cls, STEEL, A, B, P, 10, end, REDUCED)
    rel = abs(reduced.boreDisplacementFem - full.boreDisplacementFem) / abs(full.boreDisplacementFem)
    assert rel < 1e-3


@pytest.mark.parametrize("end", ENDS)
def test_hourglass_energy_stays_negligible(cls, end):
    """hourglassEnergyStaysNegligible (:98-108)."""
    r = T.run(cls, STEEL, A, B, P, 10, end, REDUCED)
    assert r.hourglassOverStrainEnergy < 1e-4


def test_reduced_integration_cures_volumetric_locking(cls):
    """reducedIntegrationCuresVolumetricLocking (:110-136), nu = 0.499."""
    soft = Material("nu -> 0.5", 205e9, 0.499, 7850.0)
    full = T.run(cls, soft, A, B, P, 3, T.RESTRAINED, FULL)
    reduced = T.run(cls, soft, A, B, P, 3, T.RESTRAINED, REDUCED)
    assert full.kineticOverStrainEnergy < 1e-9 and reduced.kineticOverStrainEnergy < 1e-9
    assert full.maxDisplacementErrorPercent > 5.0
    assert reduced.maxDisplacementErrorPercent < 1.0


def test_answer_is_insensitive_to_the_hourglass_coefficient(cls):
    """answerIsInsensitiveToTheHourglassCoefficient (:138-158)."""
    bores = [T.bore_with_hourglass_coefficient(cls, STEEL, A, B, P, q) for q in (0.01, 0.05, 0.10)]
    for b in bores[1:]:
        assert abs(b - bores[0]) / abs(bores[0]) < 1e-4


def test_converges_under_refinement(cls):
    """convergesUnderRefinement (:160-170)."""
    coarse = T.run(cls, STEEL, A, B, P, 2, T.RESTRAINED)
    fine = T.run(cls, STEEL, A, B, P, 16, T.RESTRAINED)
    assert fine.maxHoopStressErrorPercent < coarse.maxHoopStressErrorPercent


def test_timestep_follows_cfl(cls):
    """timestepFollowsCfl (:172-180)."""
    coarse = T.run(cls, STEEL, A, B, P, 5, T.RESTRAINED)
    fine = T.run(cls, STEEL, A, B, P, 10, T.RESTRAINED)
    assert abs(coarse.timestep / fine.timestep - 2.0) <= 1e-9


# ------------------------------------------------------------------ closed forms

def _end_face_disc(nr, nz, radius=0.01, height=0.02):
    """A solid cylinder whose top face carries pressure edges, ordered with r increasing so
    that the rotated tangent (tz, -tr)/len points down into the body."""
    base = QuadMesh.solidCylinder(radius, height, nr, nz)
    top = nz * (nr + 1)
    edges = np.array([[top + i, top + i + 1] for i in range(nr)], dtype=np.int32)
    return QuadMesh(base.r, base.z, base.conn, edges)


@pytest.mark.parametrize("nr", [1, 3, 8])
def test_axisymmetric_edge_weights_integrate_r_dS_exactly(cls, nr):
    """On the end face of a disc of radius R under pressure p the consistent loads satisfy
    sum Fz = -p pi R^2 and sum Fz r = -2 pi p R^3 / 3 (integral of N_i 2 pi r against r is
    exact for linear r); the element forces are zero at rest, so advance() leaves exactly the
    pressure loads in forceZ. Swapped weights change the moment by 2 pi p L (ra - rb)^2 / 6
    per edge, a flipped normal changes both signs."""
    R, p = 0.01, 7.0e7
    mesh = _end_face_disc(nr, 4, R)
    s = cls(mesh, STEEL, AXISYMMETRIC, REDUCED, SMALL_STRAIN, 1.0, 0.5)
    s.setPressureRamp(p, 0.0)
    s.advance()
    fz, fr = s.forceZ(), s.forceR()
    assert np.all(fr == 0.0)  # the normal of a horizontal edge has no radial part
    total = float(np.sum(fz))
    moment = float(np.sum(fz * mesh.r))
    assert abs(total + p * math.pi * R * R) <= 1e-12 * p * math.pi * R * R
    assert abs(moment + 2.0 * math.pi * p * R ** 3 / 3.0) <= 1e-12 * 2.0 * math.pi * p * R ** 3 / 3.0
    # node by node against the closed form of each edge's two loads
    want = np.zeros(mesh.nodeCount)
    for a, b in mesh.pressureEdges.reshape(-1, 2):
        ra, rb = mesh.r[a], mesh.r[b]
        L = rb - ra
        want[a] -= p * 2.0 * math.pi * L * (ra / 3.0 + rb / 6.0)
        want[b] -= p * 2.0 * math.pi * L * (ra / 6.0 + rb / 3.0)
    assert np.allclose(fz, want, rtol=1e-13, atol=1e-9 * p * R * R)
    assert s.borePressure() == p
    s.shutdown()


def test_plane_strain_edge_loads_are_halves(cls):
    """Plane strain: each end of an edge takes thickness L / 2 of the load."""
    p, t = 3.0e7, 0.004
    base = Outline.rectangle(0.0, 0.0, 6.0e-3, 2.0e-3).mesh(1.0e-3)
    top = np.nonzero(base.z == base.z.max())[0]
    top = top[np.argsort(base.r[top])]
    edges = np.stack([top[:-1], top[1:]], axis=1).astype(np.int32)
    mesh = QuadMesh(base.r, base.z, base.conn, edges)
    s = cls(mesh, STEEL, PLANE_STRAIN, REDUCED, SMALL_STRAIN, t, 0.5)
    s.setPressureRamp(p, 0.0)
    s.advance()
    fz = s.forceZ()
    want = np.zeros(mesh.nodeCount)
    for a, b in edges:
        half = t * (base.r[b] - base.r[a]) * 0.5
        want[a] -= p * half
        want[b] -= p * half
    assert np.allclose(fz, want, rtol=1e-14, atol=0.0)
    assert abs(float(fz.sum()) + p * t * 6.0e-3) <= 1e-12 * p * t * 6.0e-3
    s.shutdown()


def test_pressure_ramp_is_the_smoothstep(cls):
    """borePressure() follows p s^2 (3 - 2 s), s = min(1, t / ramp) (:1203-1207), and is the
    load in the last assembly: the end-face moment scales with it."""
    R, p = 0.01, 7.0e7
    mesh = _end_face_disc(2, 4, R)
    s = cls(mesh, STEEL, AXISYMMETRIC, REDUCED, SMALL_STRAIN, 1.0, 0.5)
    ramp = 40.0 * s.timestep()
    s.setPressureRamp(p, ramp)
    for _ in range(6):
        s.run(13)
        x = min(1.0, s.time() / ramp)
        assert abs(s.borePressure() - p * x * x * (3.0 - 2.0 * x)) <= 1e-15 * p
    s.run(60)
    assert s.borePressure() == p
    s.shutdown()


@pytest.mark.parametrize("zeta", [0.05, 0.5])
def test_damped_free_fall_follows_the_closed_form(cls, zeta):
    """Relaxation damping c = 2 zeta omega under gravity from rest: the nodal update
    v' = decay v + gain g with half = c dtAvg / 2, decay = (1 - half)/(1 + half),
    gain = dtAvg/(1 + half) gives v_n = -(g/c)(1 - decay^n) at every node, tending to the
    continuous terminal speed g/c."""
    g = 9.81
    mesh = Outline.rectangle(0.0, 0.0, 6.0e-3, 4.0e-3).mesh(1.0e-3)
    s = cls(mesh, STEEL, PLANE_STRAIN, REDUCED, SMALL_STRAIN, 0.01, 0.5)
    s.setGravity(0.0, -g)
    dt = s.timestep()
    omega = 0.2 / dt  # c dt = 0.4 zeta: strong enough that a wrong factor shows at once
    c = 2.0 * zeta * omega
    s.setRelaxationDamping(zeta, omega)
    half = 0.5 * c * dt
    decay = (1.0 - half) / (1.0 + half)
    n = 0
    for chunk in (1, 1, 3, 25, 200):
        s.run(chunk)
        n += chunk
        want = -(g / c) * (1.0 - decay ** n)
        vz = s.velocityZ()
        assert np.all(np.abs(vz - want) <= 1e-11 * abs(want)), (n, vz[0], want)
    assert abs(s.velocityZ()[0] + g / c) <= 1e-9 * g / c or zeta < 0.1
    s.shutdown()


# ------------------------------------------------------------------ the HIP library

@pytest.mark.gpu
@pytest.mark.parametrize("rule", RULES)
@pytest.mark.parametrize("end", ENDS)
def test_thick_wall_cylinder_bitwise_on_gpu(oracle, end, rule):
    """The Lame gate run on the HIP library equals the oracle's run bit for bit: the pressure
    edges' slots, the ramp and the damped nodal update of every one of its ~10,000 steps."""
    runs = {}
    for name, c in (("hip", ExplicitSolver), ("oracle", oracle.OracleSolver)):
        r = T.run(c, STEEL, A, B, P, 10, end, rule, keep_solver=True)
        s = r.solver
        runs[name] = (r, {k: s.nodeField(k) for k in ("ur", "uz", "vr", "vz", "fr", "fz")},
                      s.stress(), s.time(), s.borePressure())
        s.shutdown()
    (rh, fh, sh, th, ph), (ro, fo, so, to, po) = runs["hip"], runs["oracle"]
    assert rh.steps == ro.steps and th == to and ph == po
    for k in fh:
        assert same_bits(fh[k], fo[k]), (k, first_difference(fh[k], fo[k]))
    assert same_bits(sh, so), first_difference(sh, so)
    assert rh.maxDisplacementErrorPercent < 1.0 and rh.kineticOverStrainEnergy < 1e-9


@pytest.mark.gpu
def test_end_face_pressure_and_damped_fall_bitwise_on_gpu(oracle):
    """The two closed-form scenes above, HIP against oracle bit for bit."""
    out = []
    for c in (ExplicitSolver, oracle.OracleSolver):
        mesh = _end_face_disc(3, 4)
        s = c(mesh, STEEL, AXISYMMETRIC, REDUCED, SMALL_STRAIN, 1.0, 0.5)
        s.setPressureRamp(7.0e7, 50.0 * s.timestep())
        s.setRelaxationDamping(0.3, 0.05 / s.timestep())
        s.setGravity(0.0, -9.81)
        s.run(137)
        out.append([s.forceZ(), s.forceR(), s.velocityZ(), s.velocityR(), s.axialDisplacement()])
        s.shutdown()
    for a, b in zip(*out):
        assert same_bits(a, b), first_difference(a, b)

"""Argument and state errors surface where the reference throws them (solver/
ExplicitSolver.java:268-276, :387, :451, :467, :550-558, :602-609, :730-755; RigidWall.java:
110-112; Contact.java:336-364; Material.java:21-37; JohnsonCook.java:73-89)."""
import math

import numpy as np
import pytest

from neofiz_amd.solver import (PLANE_STRESS, REDUCED, FINITE_STRAIN, AXISYMMETRIC,
                               IllegalArgumentException, IllegalStateException, JohnsonCook,
                               Material, QuadMesh, RigidWall, Contact)


def make(oracle, material=None, formulation=AXISYMMETRIC):
    mesh = QuadMesh.solidCylinder(0.01, 0.02, 2, 4)
    material = material or Material("s", 200e9, 0.3, 7800.0).yielding(3e8, 1e9)
    return oracle.OracleSolver(mesh, material, formulation, REDUCED, FINITE_STRAIN, 1.0, 0.5)


def test_plane_stress_plasticity_is_refused(oracle):
    with pytest.raises(IllegalArgumentException):
        make(oracle, formulation=PLANE_STRESS)
    make(oracle, Material("el", 200e9, 0.3, 7800.0), PLANE_STRESS).shutdown()


@pytest.mark.parametrize("bad", [
    dict(youngsModulus=-1.0), dict(poissonRatio=0.5), dict(density=0.0), dict(yieldStress=-1.0),
    dict(isotropicHardening=-1.0)])
def test_material_validation(oracle, bad):
    m = Material("s", 200e9, 0.3, 7800.0).yielding(3e8, 1e9)
    from dataclasses import replace
    with pytest.raises(IllegalArgumentException):
        make(oracle, replace(m, **bad))


def test_johnson_cook_validation(oracle):
    from dataclasses import replace
    law = replace(JohnsonCook.COPPER_OFHC, n=1.5)
    with pytest.raises(IllegalArgumentException):
        make(oracle, Material("c", 117e9, 0.35, 8930.0).withJohnsonCook(law))
    m = Material("c", 117e9, 0.35, 8930.0).withJohnsonCook(JohnsonCook.COPPER_OFHC)
    with pytest.raises(IllegalArgumentException):
        make(oracle, replace(m, yieldStress=1e8))  # disagrees with A


def test_setter_validation(oracle):
    s = make(oracle)
    for call in (lambda: s.setThreads(0), lambda: s.setRelaxationDamping(0.1, 0.0),
                 lambda: s.setHourglassCoefficient(-0.01), lambda: s.setMassScaling(0.0),
                 lambda: s.setMassScaling(s.timestep() * 0.5),
                 lambda: s.setDamage(1e3, np.ones(3))):
        with pytest.raises(IllegalArgumentException):
            call()
    with pytest.raises(IllegalStateException):
        s.setErosion(1.0)  # no damage yet
    s.setDamage(1e3, np.full(8, 0.1))
    for t in (0.0, 1.5, math.nan):
        with pytest.raises(IllegalArgumentException):
            s.setErosion(t)
    s.setErosion(1.0)
    s.shutdown()


def test_wall_and_contact_validation():
    with pytest.raises(IllegalArgumentException):
        RigidWall.atZ(0.0, -0.1)
    with pytest.raises(IllegalArgumentException):
        RigidWall.atZ(0.0, math.nan)
    with pytest.raises(IllegalArgumentException):
        Contact().withFriction(-1.0)
    with pytest.raises(IllegalArgumentException):
        Contact().withDamping(1.5)
    with pytest.raises(IllegalArgumentException):
        Contact().withStiffnessScale(2.0)


def test_inverted_element_is_refused(oracle):
    with pytest.raises(IllegalStateException):
        oracle.OracleSolver(QuadMesh([0, 1, 1, 0], [0, 0, 1, 1], [0, 3, 2, 1]),
                            Material("el", 200e9, 0.3, 7800.0))


def test_mass_scaling_scales_the_step(oracle):
    """MassScalingTest: the step reaches the target, mass grows by (target/dt)^2, dt0 follows."""
    s = make(oracle)
    dt = s.timestep()
    m0 = s.nodeField("mass").sum()
    s.setMassScaling(4.0 * dt)
    assert s.timestep() == 4.0 * dt
    assert s.massScale() == 16.0
    assert abs(s.nodeField("mass").sum() / m0 - 16.0) < 1e-12
    assert s.state().dt0 == 4.0 * dt
    s.shutdown()


@pytest.mark.gpu
def test_sensitivity_refuses_contact_damage_and_erosion_set_later():
    """A solver carrying sensitivities has derivative kernels for the Taylor step only, so
    contact, damage and erosion are refused after enableSensitivity as well as before it."""
    from neofiz_amd import taylor
    mesh, s, _ = taylor.build(taylor.ROD_JC, taylor.ROD_LENGTH, taylor.ROD_DIAMETER,
                              taylor.ROD_SPEED, 2, 8, friction=taylor.ROD_JC_FRICTION)
    s.enableSensitivity()
    with pytest.raises(IllegalStateException):
        s.setContact(Contact(mesh).withFriction(0.2))
    with pytest.raises(IllegalStateException):
        s.setDamage(1e4, np.full(mesh.elementCount, 0.5))
    with pytest.raises(IllegalStateException):
        s.setErosion(1.0)
    s.setContact(None)  # switching contact off is harmless
    s.run(5)
    assert s.steps() == 5
    s.shutdown()

"""ExplicitSolver's single-value queries and the damage() view (solver/ExplicitSolver.java:
713-717, :794, :826, :1778-1794, :1832-1840, :1867-1877, :1928-1930; solver/Damage.java:
266-412), restated from the solver's arrays on the oracle, and, marked gpu, equal between the
HIP library and the oracle on the same run."""
import math

import numpy as np
import pytest

from conftest import same_bits
from neofiz_amd.mesh import Outline
from neofiz_amd.solver import (ExplicitSolver, FINITE_STRAIN, IllegalArgumentException,
                               IllegalStateException, JohnsonCook, Material, PLANE_STRAIN, REDUCED)

MM = 1.0e-3
COPPER = Material.COPPER_OFHC.withJohnsonCook(JohnsonCook.COPPER_OFHC)


def pulled_bar(cls, uniform_gf=False, steps=2500):
    """ErosionTest's bar pulled apart at 80 m/s, damage on, no erosion."""
    mesh = Outline.rectangle(0.0, 0.0, 20.0 * MM, 4.0 * MM).mesh(MM)
    s = cls(mesh, COPPER, PLANE_STRAIN, REDUCED, FINITE_STRAIN, 10.0 * MM, 0.5)
    eps = np.full(mesh.elementCount, 0.05)
    eps[::7] = 0.04  # some elements fail first
    if uniform_gf:
        s.setDamage(6.0e3, eps)
    else:
        s.setDamage(np.full(mesh.elementCount, 6.0e3), eps)
    for i in range(mesh.nodeCount):
        s.setVelocity(i, -80.0 if mesh.r[i] < 10.0 * MM else 80.0, 0.0)
    s.run(steps)
    return mesh, s


def expected_queries(mesh, s):
    """The reference's loops over the solver's arrays."""
    ppe = s.state().points_per_element
    ep = s.plasticStrain()
    w = s.plasticWorkDensity()
    jc = COPPER.johnsonCook
    temp = jc.roomTemperature + jc.taylorQuinney * w / (COPPER.density * jc.specificHeat)
    dmg = s.damage()
    worst, top = -1, 0.0
    for e in range(mesh.elementCount):
        for g in range(ppe):
            d = dmg.at(e * ppe + g, ep[e * ppe + g])
            if d > top:
                top, worst = d, e
    front = 0.0
    for e in range(mesh.elementCount):
        if any(ep[e * ppe + g] > 0.0 for g in range(ppe)):
            c = mesh.conn[4 * e:4 * e + 4]
            front = max(front, max(mesh.r[c]))
    return temp, worst, front


def test_queries_follow_the_reference_loops(oracle):
    mesh, s = pulled_bar(oracle.OracleSolver)
    temp, worst, front = expected_queries(mesh, s)
    assert same_bits(s.pointField("temperature"), temp)
    for p in (0, 7, mesh.elementCount - 1):
        assert s.temperature(p) == temp[p]
    assert s.mostDamagedElement() == worst and worst >= 0
    assert s.plasticFrontRadius() == front and front > 0.0
    assert all(s.meshZ(i) == mesh.z[i] for i in (0, 5, mesh.nodeCount - 1))
    assert s.borePressure() == 0.0
    dmg = s.damage()
    ep = s.plasticStrain()
    assert [dmg.at(p, ep[p]) for p in range(dmg.points())] == list(s.pointField("damage"))
    assert dmg.maximum(ep) == s.maxDamage()
    assert dmg.begunFraction() == s.audits().damaged_fraction
    assert dmg.failedFraction(ep) == s.audits().failed_fraction
    assert dmg.fractureEnergy(3) == 6.0e3 and dmg.points() == mesh.elementCount
    assert dmg.hasBegun(worst) and dmg.strengthAtOnset(worst) > 0.0 and dmg.softeningModulus(worst) < 0.0
    assert dmg.bandWidth(worst) > 0.0 and dmg.clampedPoints() == 0
    with pytest.raises(IllegalArgumentException):
        s.temperature(-1)
    with pytest.raises(IllegalArgumentException):
        s.meshZ(mesh.nodeCount)
    s.shutdown()


def test_uniform_fracture_energy_is_the_array_form(oracle):
    """setDamage(double, double[]) :713-717 fills the array and calls the array form."""
    _, a = pulled_bar(oracle.OracleSolver, uniform_gf=False, steps=800)
    _, b = pulled_bar(oracle.OracleSolver, uniform_gf=True, steps=800)
    for k in ("plastic_strain", "damage_strength", "damage_softening", "damage"):
        assert same_bits(a.pointField(k), b.pointField(k)), k
    a.shutdown()
    b.shutdown()


def test_temperature_needs_a_thermal_model(oracle):
    """temperature(int) throws IllegalStateException without a Johnson-Cook law (:1834-1837);
    the bulk field reads NaN there; no damage model means damage() is None and
    mostDamagedElement() is -1."""
    mesh = Outline.rectangle(0.0, 0.0, 4.0 * MM, 2.0 * MM).mesh(MM)
    s = oracle.OracleSolver(mesh, Material.STEEL_4340, PLANE_STRAIN, REDUCED, FINITE_STRAIN, 1.0, 0.5)
    with pytest.raises(IllegalStateException):
        s.temperature(0)
    assert np.all(np.isnan(s.pointField("temperature")))
    assert s.damage() is None and s.mostDamagedElement() == -1
    assert s.plasticFrontRadius() == 0.0
    s.shutdown()


@pytest.mark.gpu
def test_queries_equal_the_oracle_on_gpu(oracle):
    out = []
    for cls in (ExplicitSolver, oracle.OracleSolver):
        mesh, s = pulled_bar(cls, uniform_gf=True)
        fields = [s.pointField(k) for k in ("temperature", "damage", "damage_failure_strain",
                                             "damage_fracture_energy", "damage_width")]
        out.append((fields, s.temperature(11), s.mostDamagedElement(), s.plasticFrontRadius(),
                    s.borePressure(), s.meshZ(9)))
        s.shutdown()
    (fh, *qh), (fo, *qo) = out
    for a, b in zip(fh, fo):
        assert same_bits(a, b)
    assert qh == qo

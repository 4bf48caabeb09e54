"""Bit-for-bit determinism across thread counts (src/test/java/org/neofiz/solver/
ThreadingTest.java:49-198): the element pass writes private slots and the gather sums each
node in a fixed order, so the answer never depends on the schedule."""
import numpy as np

from conftest import same_bits
from neofiz_amd import taylor


def run(oracle, threads, steps=600, change_at=None):
    mesh, s, _ = taylor.build(taylor.ROD_JC, taylor.ROD_LENGTH, taylor.ROD_DIAMETER,
                              taylor.ROD_SPEED, 8, 80, friction=0.09,
                              solver_cls=oracle.OracleSolver)
    s.setThreads(threads)
    if change_at is None:
        s.run(steps)
    else:
        s.run(change_at)
        s.setThreads(3)
        s.run(steps - change_at)
    out = [s.nodeField(k) for k in ("ur", "uz", "vr", "vz", "fr", "fz")]
    out.append(s.stress())
    out.append(s.plasticStrain())
    s.shutdown()
    return out


def test_any_thread_count_is_bit_identical(oracle):
    ref = run(oracle, 1)
    for t in (2, 5, 8):
        got = run(oracle, t)
        assert all(same_bits(a, b) for a, b in zip(got, ref)), t


def test_changing_threads_mid_run_is_bit_identical(oracle):
    assert all(same_bits(a, b) for a, b in zip(run(oracle, 2, change_at=250), run(oracle, 1)))


def test_more_threads_than_elements(oracle):
    from neofiz_amd.solver import Material, QuadMesh, PLANE_STRAIN, REDUCED, FINITE_STRAIN
    mesh = QuadMesh.solidCylinder(0.01, 0.01, 2, 2)
    outs = []
    for t in (1, 8):
        s = oracle.OracleSolver(mesh, Material("s", 200e9, 0.3, 7800.0).yielding(3e8, 1e9),
                                PLANE_STRAIN, REDUCED, FINITE_STRAIN, 0.01, 0.5)
        s.setThreads(t)
        s.putNodeField("vz", np.linspace(-10, 10, mesh.nodeCount))
        s.run(50)
        outs.append(s.stress())
        s.shutdown()
    assert same_bits(outs[0], outs[1])

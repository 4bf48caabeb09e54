"""Film frames captured on the device against the oracle's, bit for bit: positions and every
built-in element quantity after a damaged, eroding run, the whole film's JSON text, and a frame
of the 128 x 1280 Taylor rod."""
import time

import numpy as np
import pytest

from conftest import first_difference, same_bits
from neofiz_amd import taylor
from neofiz_amd.film import Film
from neofiz_amd.solver import (ExplicitSolver, FINITE_STRAIN, FULL, JohnsonCook, Material, PLANE_STRAIN,
                               QuadMesh)

pytestmark = pytest.mark.gpu

COPPER = Material.COPPER_OFHC.withJohnsonCook(JohnsonCook.COPPER_OFHC.quasiStatic())
BUILTINS = [Film.PLASTIC_STRAIN, Film.DAMAGE, Film.SPEED, Film.MISES, Film.STRESS_RR, Film.STRESS_ZZ,
            Film.STRESS_TT, Film.STRESS_RZ, Film.CENTROID_Z, Film.ALIVE]


def scene(cls):
    mesh = QuadMesh.cylinderWall(0.0, 0.004, 0.012, 16, 48)
    s = cls(mesh, COPPER, PLANE_STRAIN, FULL, FINITE_STRAIN, 0.01, 0.5)
    s.setDamage(np.full(mesh.elementCount, 5.0e3), np.full(mesh.elementCount, 0.05))
    s.setErosion(1.0)
    v = np.zeros(mesh.nodeCount)
    v[mesh.z > 0.0119] = 300.0
    s.putNodeField("vz", v)
    return mesh, s


def film_of(mesh):
    f = Film(mesh).planar()
    for b in BUILTINS:
        f.colourBy(b.name, "", Film.AUTO, Film.AUTO, b)
    f.colourBy("von Mises", "MPa", Film.AUTO, Film.AUTO, Film.MISES, transform=lambda v: v / 1e6)
    return (f.aliveIf(Film.ALIVE).readout("time, us", lambda s: 1e6 * s.time())
            .readout("elements gone", lambda s: s.erodedElements()))


def test_frames_match_the_oracle_bitwise():
    from oracle.pyoracle import OracleSolver
    mesh, gpu = scene(ExplicitSolver)
    _, ref = scene(OracleSolver)
    fg, fr = film_of(mesh), film_of(mesh)
    for chunk in range(6):
        print(f"film chunk {chunk}", flush=True)
        fg.capture(gpu)
        fr.capture(ref)
        gpu.run(150)
        ref.run(150)
    fg.capture(gpu)
    fr.capture(ref)
    assert ref.state().eroded_elements > 0
    for k, (a, b) in enumerate(zip(fg.frames, fr.frames)):
        assert same_bits(a.r, b.r) and same_bits(a.z, b.z), k
        for f in range(len(BUILTINS) + 1):
            assert same_bits(a.values[f], b.values[f]), (k, f, first_difference(a.values[f], b.values[f]))
        assert a.live == b.live and a.scalars == b.scalars
    assert fg.json("scene") == fr.json("scene")
    gpu.shutdown()
    ref.shutdown()


def test_taylor_rod_frame_matches_the_oracle():
    from oracle.pyoracle import OracleSolver
    out = []
    for cls in (ExplicitSolver, OracleSolver):
        mesh, s, _ = taylor.build(taylor.ROD_JC, taylor.ROD_LENGTH, taylor.ROD_DIAMETER, taylor.ROD_SPEED,
                                  128, 1280, friction=taylor.ROD_JC_FRICTION, solver_cls=cls)
        if cls is OracleSolver:
            s.setThreads(8)
        s.run(200)
        f = Film(mesh)
        for b in BUILTINS:
            f.colourBy(b.name, "", Film.AUTO, Film.AUTO, b)
        t0 = time.perf_counter()
        f.capture(s)
        print(f"{cls.__name__}: capture of {mesh.elementCount} elements {1e3 * (time.perf_counter() - t0):.1f} ms",
              flush=True)
        out.append(f.frames[0])
        s.shutdown()
    a, b = out
    assert same_bits(a.r, b.r) and same_bits(a.z, b.z)
    for f in range(len(BUILTINS)):
        assert same_bits(a.values[f], b.values[f]), (f, first_difference(a.values[f], b.values[f]))

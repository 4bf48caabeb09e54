"""Film frames on the oracle: the reference's FilmTest (src/test/java/org/neofiz/render/
FilmTest.java) restated, the number format of Film.number, and the built-in element
quantities against the solver's own accessors."""
import json
import math
import re

import numpy as np
import pytest

from neofiz_amd.film import Film, number
from neofiz_amd.solver import (AXISYMMETRIC, FINITE_STRAIN, PLANE_STRAIN, REDUCED, FULL, JohnsonCook,
                               Material, QuadMesh)

COPPER = Material.COPPER_OFHC.withJohnsonCook(JohnsonCook.COPPER_OFHC.quasiStatic())  # BurstCase.java:143-144


def solver(oracle, mesh):  # FilmTest.java:38-41
    return oracle.OracleSolver(mesh, COPPER, AXISYMMETRIC, REDUCED, FINITE_STRAIN, 1.0, 0.5)


def wall():  # :44-46
    return QuadMesh.cylinderWall(24.0e-3, 26.0e-3, 3.0e-3, 4, 3)


def body(text, key, occurrence=0):
    at = -1
    for _ in range(occurrence + 1):
        at = text.index(key, at + 1)
    start = at + len(key)
    return text[start:text.index("]", start)]


def test_boundary_is_the_free_surface(oracle):  # :49-72
    film = Film(wall())
    s = solver(oracle, wall())
    film.capture(s)
    text = film.json("t")
    b = [int(x) for x in body(text, '"boundary":[').split(",")]
    assert len(b) == 2 * 14
    seen = {(min(b[i], b[i + 1]), max(b[i], b[i + 1])) for i in range(0, len(b), 2)}
    assert len(seen) == 14
    s.shutdown()


def test_captures_deformed_geometry(oracle):  # :75-98
    mesh = wall()
    s = solver(oracle, mesh)
    film = Film(mesh)
    film.capture(s)
    s.setPressureRamp(40.0e6, 1.0e-6)
    s.run(400)
    film.capture(s)
    text = film.json("t")
    assert text.count('{"t":') == 2
    first = np.array([float(x) for x in body(text, '"r":[').split(",")])
    second = np.array([float(x) for x in body(text, '"r":[', 1).split(",")])
    assert np.any(np.abs(second - first) > 1e-9)
    s.shutdown()


def test_auto_range_fits_the_data(oracle):  # :101-122
    mesh = wall()
    s = solver(oracle, mesh)
    film = (Film(mesh).colourBy("fixed", "", 0.0, 10.0, lambda solver, e: e)
            .colourBy("fitted", "", Film.AUTO, Film.AUTO, lambda solver, e: e))
    film.capture(s)
    text = film.json("t")
    assert '"name":"fixed","units":"","low":0,"high":10,"peak":10,"auto":false' in text
    assert '"name":"fitted","units":"","low":0,"high":11,"peak":11,"auto":true' in text
    s.shutdown()


def test_auto_range_clips_outliers(oracle):  # :125-147
    mesh = QuadMesh.cylinderWall(24.0e-3, 26.0e-3, 3.0e-3, 40, 10)
    s = solver(oracle, mesh)
    film = Film(mesh).colourBy("spiky", "", Film.AUTO, Film.AUTO, lambda v, e: 500.0 if e == 0 else e)
    for _ in range(10):
        film.capture(s)
    text = film.json("t")
    tail = text[text.index('"name":"spiky"'):]
    high = float(re.search(r'"high":([^,}\]]+)', tail).group(1))
    peak = float(re.search(r'"peak":([^,}\]]+)', tail).group(1))
    assert peak == 500.0
    assert 380.0 <= high < 450.0
    s.shutdown()


def test_flat_field_is_survivable(oracle):  # :150-160
    mesh = wall()
    film = Film(mesh).colourBy("flat", "", Film.AUTO, Film.AUTO, lambda s, e: 0.0)
    s = solver(oracle, mesh)
    film.capture(s)
    text = film.json("t")
    assert "NaN" not in text and "Infinity" not in text
    s.shutdown()


def test_order_is_declaration_order(oracle):  # :163-177
    mesh = wall()
    film = (Film(mesh).colourBy("alpha", "", 0.0, 1.0, lambda s, e: 0.0)
            .colourBy("beta", "MPa", 0.0, 1.0, lambda s, e: 0.0)
            .readout("first", lambda s: 1.0).readout("second", lambda s: 2.0))
    s = solver(oracle, mesh)
    film.capture(s)
    text = film.json("t")
    assert text.index('"alpha"') < text.index('"beta"')
    assert text.index('"first"') < text.index('"second"')
    assert '"s":[1,2]' in text
    s.shutdown()


def test_escapes_titles(oracle):  # :180-185
    film = Film(wall())
    s = solver(oracle, wall())
    film.capture(s)
    assert film.json('a "quoted" title').startswith('{"title":"a \\"quoted\\" title"')
    s.shutdown()


def test_the_film_is_the_viewers_json(oracle, tmp_path):
    """The written file is the two-line script the viewer loads, and its JSON parses."""
    mesh = wall()
    s = solver(oracle, mesh)
    film = Film(mesh).planar().colourBy("plastic strain", "", Film.AUTO, Film.AUTO, Film.PLASTIC_STRAIN)
    film.aliveIf(Film.ALIVE).readout("time, us", lambda v: 1e6 * v.time())
    film.capture(s)
    path = film.write(str(tmp_path / "wall.js"), "wall")
    text = open(path).read()
    head, tail = "window.NEOFIZ_FILM = ", ";\n(window.NEOFIZ_FILMS = window.NEOFIZ_FILMS || []).push(window.NEOFIZ_FILM);\n"
    assert text.startswith(head) and text.endswith(tail)
    doc = json.loads(text[len(head):-len(tail)])
    assert doc["axisymmetric"] is False and doc["nodes"] == mesh.nodeCount and len(doc["frames"]) == 1
    assert doc["frames"][0]["live"] == "1" * mesh.elementCount
    s.shutdown()


@pytest.mark.parametrize("value,text", [
    (0.0, "0"), (-0.0, "0"), (math.nan, "0"), (math.inf, "0"), (10.0, "10"), (11.0, "11"), (500.0, "500"),
    (0.5, "0.5"), (1.0 / 3.0, "0.333333"), (2.0 / 3.0, "0.666667"), (1e-5, "1e-5"), (-2.5e-7, "-2.5e-7"),
    (0.0001, "0.0001"), (0.000123456789, "0.000123457"), (123456789.0, "1.23457e8"), (999999.0, "999999"),
    (999999.5, "1e6"), (100000.5, "100001"), (0.125, "0.125"), (1.000005, "1.00001"), (1.0000005, "1"), (6.02214076e23, "6.02214e23"),
    (1.5e300, "1.5e300"), (2.2250738585072014e-308, "2.22507e-308"), (-40.0e6, "-4e7"), (25.99, "25.99"), (0.0004500005, "0.000450001"), (0.00045000049, "0.00045"),
])
def test_number_is_javas_general_format(value, text):
    """Film.number (Film.java:295-314): %.6g rounded half-up on the shortest digits, trimmed.
    1.000005 is the shortest form of its double, so half-up on those digits gives 1.00001,
    whereas the exact binary value (1.00000499999999994...) would round down; Java's
    Formatter takes the former."""
    assert number(value) == text


def test_builtin_fields_are_the_accessors(oracle):
    """Each built-in quantity equals the solver accessor it stands for, element by element,
    after a damaged, eroding plane-strain run with full integration (four points)."""
    mesh = QuadMesh.cylinderWall(0.0, 0.004, 0.012, 8, 24)
    s = oracle.OracleSolver(mesh, COPPER, PLANE_STRAIN, FULL, FINITE_STRAIN, 0.01, 0.5)
    s.setDamage(np.full(mesh.elementCount, 5.0e3), np.full(mesh.elementCount, 0.05))
    s.setErosion(1.0)
    v = np.zeros(mesh.nodeCount)
    v[mesh.z > 0.0119] = 300.0
    s.putNodeField("vz", v)
    s.run(600)
    assert s.state().eroded_elements > 0
    builtins = [Film.PLASTIC_STRAIN, Film.DAMAGE, Film.SPEED, Film.MISES, Film.STRESS_RR, Film.STRESS_ZZ,
                Film.STRESS_TT, Film.STRESS_RZ, Film.CENTROID_Z, Film.ALIVE]
    film = Film(mesh)
    for bi in builtins:
        film.colourBy(bi.name, "", Film.AUTO, Film.AUTO, bi)
    film.capture(s)
    fr = film.frames[0]
    assert np.array_equal(fr.r, mesh.r + s.radialDisplacement())
    assert np.array_equal(fr.z, mesh.z + s.axialDisplacement())
    vr, vz = s.velocityR(), s.velocityZ()
    eroded = s.elementField("eroded")
    for e in range(mesh.elementCount):
        c = mesh.conn[4 * e:4 * e + 4]
        sig = s.centroidStress(e)
        a, b, cc = sig[0] - sig[1], sig[1] - sig[2], sig[2] - sig[0]
        speed = 0.0
        height = 0.0
        for n in c:
            speed += math.sqrt(vr[n] * vr[n] + vz[n] * vz[n])
            height += mesh.z[n] + s.axialDisplacement()[n]
        want = [s.elementPlasticStrain(e), s.elementDamage(e), 0.25 * speed,
                math.sqrt(0.5 * (a * a + b * b + cc * cc) + 3.0 * sig[3] * sig[3]), *sig, 0.25 * height,
                0.0 if eroded[e] else 1.0]
        got = [fr.values[k][e] for k in range(len(builtins))]
        assert got == want, e
    s.shutdown()

"""Device-side defect-field seeding against the oracle, bit for bit, and the reference's
statistical properties of the field (src/test/java/org/neofiz/core/DefectFieldTest.java)
measured on the device at the reference's sizes.

Bitwise: sampleOnto on ring and slab meshes (lattice, graded-and-jittered, axis-touching,
multi-body), memoised and hashed; the point functions; Normal's erfc, erf, cdf and survival;
the device-pointer entry. Properties: size effect over 3000 parts per tube, the body
distribution against the closed form, the weakest link under a 1024-fold refinement.
"""
import math

import numpy as np
import pytest

from conftest import first_difference, same_bits
from neofiz_amd import perforation
from neofiz_amd.defects import DefectField, Normal, Weibull
from neofiz_amd.mesh import Outline
from neofiz_amd.solver import IllegalArgumentException, QuadMesh

pytestmark = pytest.mark.gpu

BASE = Weibull(0.42, 20.0)
V0 = 1e-9
LENGTH = 1e-3


def jittered(nr, nz, seed):
    """A cylinder wall whose interior nodes are moved by up to a fifth of a cell, then graded
    radially by r = 0.05 + 0.05 s^1.5 (a monotone map, so every element keeps its orientation)."""
    m = QuadMesh.cylinderWall(0.05, 0.10, 0.02, nr, nz)
    rng = np.random.default_rng(seed)
    s = (m.r - 0.05) / 0.05
    z = m.z.copy()
    inner = (s > 0.0) & (s < 1.0) & (m.z > 0.0) & (m.z < 0.02)
    s[inner] += rng.uniform(-0.2, 0.2, inner.sum()) / nr
    z[inner] += rng.uniform(-0.2, 0.2, inner.sum()) * (0.02 / nz)
    return QuadMesh(0.05 + 0.05 * s ** 1.5, z, m.conn)


MESHES = {
    "wall64x32": lambda: QuadMesh.cylinderWall(0.05, 0.10, 0.02, 64, 32),
    "solid8x32": lambda: QuadMesh.solidCylinder(0.0038, 0.0254, 8, 32),
    "jittered": lambda: jittered(48, 24, 3),
    "rings1x400": lambda: QuadMesh.cylinderWall(0.05, 0.10, 0.05, 1, 400),
    "bodies": lambda: Outline.assemble(0.4e-3, *perforation.outlines(3)[0], perforation.outlines(3)[1]),
}


@pytest.fixture(scope="module")
def fields(oracle):
    made = []

    def make(seed, base=BASE, v0=V0, length=LENGTH):
        dev = DefectField(base, v0, length, seed)
        ref = oracle.OracleDefectField(base, v0, length, seed)
        made.extend([dev, ref])
        return dev, ref
    yield make
    for f in made:
        f.close()


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("depth", [None, 0.02])
def test_sample_onto_matches_oracle_bitwise(fields, name, depth):
    mesh = MESHES[name]()
    dev, ref = fields(42)
    seeds = np.array([42, -7, 2**40 + 3], dtype=np.int64)
    print(f"sampleOnto {name} depth={depth} elements={mesh.elementCount}", flush=True)
    got = dev.sampleOnto(mesh, seeds=seeds, depth=depth)
    want = ref.sampleOnto(mesh, seeds=seeds, depth=depth)
    assert same_bits(got, want), first_difference(got.reshape(-1), want.reshape(-1))
    hashed = dev.sampleOnto(mesh, seeds=seeds, depth=depth, cached=False)
    assert same_bits(hashed, want), first_difference(hashed.reshape(-1), want.reshape(-1))
    assert same_bits(dev.sampleOnto(mesh, depth=depth), want[0])


def test_point_functions_match_oracle_bitwise(fields):
    dev, ref = fields(20260919)
    rng = np.random.default_rng(5)
    n = 5000
    r = rng.uniform(-0.02, 0.2, n)
    z = rng.uniform(-0.1, 0.1, n)
    area = np.exp(rng.uniform(math.log(1e-10), math.log(1e-4), n))
    vol = np.exp(rng.uniform(math.log(1e-15), math.log(1.0), n))
    g = rng.normal(0.0, 3.0, n)
    for name, a, b in (("gaussianAt", dev.gaussianAt(r, z), ref.gaussianAt(r, z)),
                       ("failureStrainAt", dev.failureStrainAt(r, z, area), ref.failureStrainAt(r, z, area)),
                       ("failureStrainForVolume", dev.failureStrainForVolume(r, z, vol),
                        ref.failureStrainForVolume(r, z, vol)),
                       ("effectiveVolume", dev.effectiveVolume(r, area), ref.effectiveVolume(r, area)),
                       ("transform", dev.transform(g, vol), ref.transform(g, vol))):
        assert same_bits(a, b), (name, first_difference(a, b))
    i = rng.integers(-2**31, 2**31 - 1, n).astype(np.float64)
    j = rng.integers(-2**31, 2**31 - 1, n).astype(np.float64)
    assert same_bits(dev.white(i, j), ref.white(i, j))


def test_normal_matches_oracle_and_published_values(oracle):
    b = oracle.binding()
    x = np.concatenate([np.linspace(-40.0, 40.0, 40001), [0.46875, 4.0, -0.0, 1e-300]])
    for fn in ("erfc", "erf", "cdf", "survival"):
        a, c = getattr(Normal, fn)(x), getattr(Normal, fn)(x, binding=b)
        assert same_bits(a, c), (fn, first_difference(a, c))
    for xv, expected in ((0.4, 0.5716076449533316), (1.0, 0.15729920705028516),
                         (4.0, 1.541725790028002e-8), (7.0, 4.183825607779415e-23)):
        assert abs(Normal.erfc(xv) - expected) <= abs(expected) * 1e-13


class DeviceArrays:
    """Device buffers from the HIP runtime the library itself links (torch ships a runtime of
    its own, which cannot open the GPU once this one has)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), max(a.nbytes, 16)) == 0
        self.held.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p.value

    def get(self, ptr, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.held:
            self.hip.hipFree(p)


def test_device_pointer_entry(fields):
    mesh = MESHES["jittered"]()
    dev, ref = fields(11)
    seeds = np.array([11, 12], dtype=np.int64)
    mem = DeviceArrays()
    try:
        r, z, conn = mem.put(mesh.r), mem.put(mesh.z), mem.put(mesh.conn.astype(np.int32))
        out = mem.put(np.zeros((2, mesh.elementCount)))
        dev.sampleOntoDevice(mesh.nodeCount, r, z, mesh.elementCount, conn, out, seeds=seeds)
        got = mem.get(out, (2, mesh.elementCount))
        assert same_bits(got, ref.sampleOnto(mesh, seeds=seeds))
        ms = dev.timing()
        assert len(ms) == 3 and all(v >= 0.0 for v in ms) and ms[2] > 0.0
        bad = mesh.conn.astype(np.int32).copy()
        bad[7] = -1
        with pytest.raises(IllegalArgumentException):   # checked on the device before any read
            dev.sampleOntoDevice(mesh.nodeCount, r, z, mesh.elementCount, mem.put(bad), out)
    finally:
        mem.free()


def test_errors_are_the_references(fields):
    dev, _ = fields(1)
    with pytest.raises(IllegalArgumentException):
        dev.failureStrainForVolume(0.05, 0.01, -1.0)       # Weibull.forVolume: volume must be positive
    brittle, ref = fields(1, base=Weibull(1.0, 0.01))
    for f in (brittle, ref):                                # the weakest-link scale underflows to 0
        with pytest.raises(IllegalArgumentException, match="scale must be positive"):
            f.failureStrainForVolume(0.05, 0.01, 1e10 * V0)
    mesh = MESHES["wall64x32"]()
    mesh.conn[5] = mesh.nodeCount + 3
    with pytest.raises(IllegalArgumentException):
        dev.sampleOnto(mesh)
    for v0, length in ((0.0, LENGTH), (V0, math.nan), (V0, math.inf)):
        with pytest.raises(IllegalArgumentException):
            DefectField(BASE, v0, length, 1)


def test_large_multibody_mesh_matches_oracle(fields):
    """The 200k-element perforation assembly, slab rule, one seed: device == oracle."""
    cell = perforation.cell_for(200_000)
    wall, slug = perforation.outlines()
    mesh = Outline.assemble(cell, *wall, slug)
    dev, ref = fields(5, base=Weibull(1.0, 10.0))
    print(f"perforation mesh {mesh.elementCount} elements", flush=True)
    got = dev.sampleOnto(mesh, depth=perforation.DEPTH)
    assert same_bits(got, dev.sampleOnto(mesh, depth=perforation.DEPTH, cached=False))
    want = ref.sampleOnto(mesh, depth=perforation.DEPTH)
    assert same_bits(got, want), first_difference(got, want)


# ------------------------------------------------------------ the field's statistics, on the device

def test_size_effect_is_measured_on_device(fields):  # DefectFieldTest.java:259-300
    dev, _ = fields(9000)
    seeds = 9000 + np.arange(3000)
    volumes, means = [], []
    for ri in (0.02, 0.04, 0.08, 0.16):
        mesh = QuadMesh.cylinderWall(ri, ri + 0.005, 0.02, 8, 16)
        means.append(dev.sampleOnto(mesh, seeds=seeds).min(axis=1).sum() / seeds.size)
        volumes.append(mesh.totalRingVolume())
    for i in range(1, 4):
        ratio = volumes[i] / volumes[0]
        assert abs(means[i] / means[0] - ratio ** (-1.0 / BASE.modulus)) <= 0.005
    assert means[3] < 0.94 * means[0]


def test_body_matches_the_closed_form_on_device(fields):  # :222-256
    mesh = QuadMesh.cylinderWall(0.05, 0.10, 0.02, 64, 32)
    dev, _ = fields(7000)
    weakest = dev.sampleOnto(mesh, seeds=7000 + np.arange(300)).min(axis=1)
    mean = weakest.mean()
    sd = math.sqrt(((weakest - mean) ** 2).sum() / (weakest.size - 1))
    predicted = BASE.forVolume(mesh.totalRingVolume(), V0)
    assert abs(mean - predicted.mean()) <= predicted.mean() * 0.04
    assert abs(sd - predicted.standardDeviation()) <= predicted.standardDeviation() * 0.25


def test_refinement_does_not_move_the_body_on_device(fields):  # :181-219
    seeds = np.arange(1, 9)
    dev, _ = fields(1)
    weakest = []
    for refine in (2, 8, 32, 64):
        mesh = QuadMesh.cylinderWall(0.05, 0.10, 0.02, 4 * refine, 2 * refine)
        weakest.append(dev.sampleOnto(mesh, seeds=seeds).min(axis=1))
    drift = np.abs(weakest[3] - weakest[2]) / weakest[2]
    assert np.all(drift < 0.03), drift
    assert np.all((weakest[3] > 0.6 * weakest[0]) & (weakest[3] < 1.4 * weakest[0]))

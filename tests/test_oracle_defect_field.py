"""The oracle's defect field (oracle/defect_field.cpp) and the host Weibull record, pinned.

Restated from the reference's own tests (src/test/java/org/neofiz/core/): NormalTest (erfc and
cdf values, branch joins, tails, monotonicity), WeibullTest (the weakest-link algebra) and
DefectFieldTest (the field as a random field: standard normal marginal, correlation at the
physical length, mesh independence bit for bit, resolution floor, weakest link under
refinement, closed-form body distribution, size and ring effects, marginal KS, no degenerate
values), with the reference's sizes and tolerances. Beyond them, the pieces are pinned by what
the mathematics fixes:
  - erfc against mpmath's correctly rounded values over [-6, 26];
  - the SplitMix64 finaliser against the generator's published outputs for seed 1234567;
  - white(i, j) against Box-Muller on those hashes with correctly rounded log and cos;
  - the smoothing against a brute-force 20 x 20 sum at scattered points;
  - the copula transform against scipy's normal and Weibull quantiles.
"""
import math

import mpmath
import numpy as np
import pytest
import scipy.special as sp
import scipy.stats as st

from neofiz_amd.defects import DefectField, Normal, Weibull
from neofiz_amd.solver import IllegalArgumentException, QuadMesh

mpmath.mp.prec = 120

BASE = Weibull(0.42, 20.0)    # DefectFieldTest.java:34-38
V0 = 1e-9
LENGTH = 1e-3
M64 = (1 << 64) - 1


@pytest.fixture(scope="module")
def b(oracle):
    return oracle.binding()


@pytest.fixture(scope="module")
def field(oracle):
    made = []

    def make(seed, base=BASE, v0=V0, length=LENGTH):
        f = oracle.OracleDefectField(base, v0, length, seed)
        made.append(f)
        return f
    yield make
    for f in made:
        f.close()


# ------------------------------------------------------------------ Normal (NormalTest.java)

def test_erfc_closed_forms(b):  # :23-29
    assert Normal.erf(0.0, binding=b) == 0.0
    assert abs(Normal.erfc(0.0, binding=b) - 1.0) <= 1e-16
    assert abs(Normal.erfc(-100.0, binding=b) - 2.0) <= 1e-15
    assert Normal.erfc(100.0, binding=b) == 0.0


@pytest.mark.parametrize("x,expected", [(0.4, 0.5716076449533316), (0.5, 0.4795001221869535),
                                        (1.0, 0.15729920705028516), (2.0, 0.0046777349810472645),
                                        (4.0, 1.541725790028002e-8), (5.0, 1.537459794428035e-12),
                                        (7.0, 4.183825607779415e-23)])
def test_erfc_published_values(b, x, expected):  # :39-55
    assert abs(Normal.erfc(x, binding=b) - expected) <= abs(expected) * 1e-13


def test_erfc_against_correctly_rounded(b):
    """Cody's approximation is within a few parts in 1e16 of the exact erfc everywhere it is
    used (the reference's bar is 1e-13)."""
    x = np.linspace(-6.0, 26.0, 3201)
    got = Normal.erfc(x, binding=b)
    exact = np.array([float(mpmath.erfc(mpmath.mpf(float(v)))) for v in x])
    assert np.max(np.abs(got - exact) / exact) < 2e-15
    e = Normal.erf(x[x < 6.0], binding=b)
    exact_erf = np.array([float(mpmath.erf(mpmath.mpf(float(v)))) for v in x[x < 6.0]])
    assert np.max(np.abs(e - exact_erf)) < 4e-16


def test_erfc_branches_join(b):  # :58-72
    for x in (0.46875, 4.0):
        below = Normal.erfc(np.nextafter(x, 0.0), binding=b)
        above = Normal.erfc(np.nextafter(x, 9.0), binding=b)
        assert abs(below - above) <= abs(below) * 1e-13


def test_cdf_values_and_symmetry(b):  # :75-90
    assert abs(Normal.cdf(0.0, binding=b) - 0.5) <= 1e-16
    for x, v in ((1.0, 0.8413447460685429), (2.0, 0.9772498680518208), (3.0, 0.9986501019683699),
                 (1.96, 0.9750021048517795), (1.959963984540054, 0.975)):
        assert abs(Normal.cdf(x, binding=b) - v) <= 1e-15
    x = np.arange(0.0, 8.0, 0.25)
    assert np.max(np.abs(Normal.cdf(x, binding=b) + Normal.cdf(-x, binding=b) - 1.0)) <= 1e-15


def test_tail_does_not_cancel(b):  # :93-107
    assert 1.0 - Normal.cdf(9.0, binding=b) == 0.0
    s9 = Normal.survival(9.0, binding=b)
    assert s9 > 0.0 and abs(s9 - 1.1285884059538297e-19) <= 1.1285884059538297e-19 * 1e-13
    s30 = Normal.survival(30.0, binding=b)
    assert s30 > 0.0 and abs(s30 - 4.9067139271480216e-198) <= 4.9067139271480216e-198 * 1e-13


def test_tail_amplifies_the_argument(b):  # :110-127
    by_multiply = Normal.erfc(30.0 * 0.70710678118654752440, binding=b)
    by_divide = Normal.erfc(30.0 / math.sqrt(2.0), binding=b)
    assert by_multiply != by_divide
    assert abs(by_divide - by_multiply) <= abs(by_divide) * 1e-12
    assert abs(by_divide - by_multiply) > abs(by_divide) * 1e-14


def test_cdf_monotone(b):  # :134-142
    x = -8.0 + 0.0009765625 * np.arange(16385)
    p = Normal.cdf(x, binding=b)
    assert np.all(np.diff(p) >= 0.0) and p[0] >= 0.0


# ------------------------------------------------------------------ Weibull (WeibullTest.java)

STEEL = Weibull(0.42, 30.0)
IRON = Weibull(0.012, 8.0)


def test_weibull_scale_is_characteristic():  # :27-33
    assert abs(STEEL.survival(STEEL.scale) - math.exp(-1.0)) <= 1e-15
    assert abs(IRON.survival(IRON.scale) - math.exp(-1.0)) <= 1e-15
    assert abs(STEEL.cumulativeHazard(STEEL.scale) - 1.0) <= 1e-15


def test_weibull_survival_cdf_and_quantiles():  # :36-71
    x = 0.01
    while x < 1.0:
        assert abs(STEEL.survival(x) + STEEL.cdf(x) - 1.0) <= 1e-15
        x *= 1.3
    assert abs(Weibull(1.0, 1.0).cdf(1e-14) - 1e-14) <= 1e-28
    p = 1e-12
    while p < 1.0:
        assert abs(STEEL.cdf(STEEL.quantile(p)) - p) <= p * 1e-12
        p *= 3.0
    s = 1e-15
    while s < 0.5:
        assert abs(STEEL.survival(STEEL.quantileFromSurvival(s)) - s) <= s * 1e-12
        s *= 3.0
    s = 1e-9
    while s < 0.5:
        assert abs(STEEL.quantileFromSurvival(s) - STEEL.quantile(1.0 - s)) <= STEEL.scale * 1e-9
        s *= 3.0
    assert abs(STEEL.quantileFromSurvival(1e-15) - STEEL.quantile(1.0 - 1e-15)) > STEEL.scale * 1e-8


def test_weibull_minimum_of_n():  # :74-107
    for w in (STEEL, IRON):
        for n in (1.0, 2.0, 7.5, 1000.0, 1e6):
            weakest = w.minimumOf(n)
            assert weakest.modulus == w.modulus
            for x in (0.2 * w.scale, w.scale, 1.5 * w.scale):
                expected = n * w.cumulativeHazard(x)
                assert abs(weakest.cumulativeHazard(x) - expected) <= expected * 1e-13
    for n in (1.0, 2.0, 7.5, 1000.0, 1e6):
        x = 0.9 * STEEL.scale
        expected = STEEL.survival(x) ** n
        assert abs(STEEL.minimumOf(n).survival(x) - expected) <= expected * 1e-8
    assert STEEL.minimumOf(1.0) == STEEL
    once, twice = STEEL.minimumOf(60.0), STEEL.minimumOf(12.0).minimumOf(5.0)
    assert abs(once.scale - twice.scale) <= once.scale * 1e-14
    assert STEEL.minimumOf(8.0).scale == STEEL.forVolume(8e-6, 1e-6).scale


def test_weibull_headline_numbers():  # :122-131
    w = Weibull(1.0, 10.0)
    assert abs(w.sizeFactor(8.0, 1.0) - 8.0 ** -0.1) <= 1e-15
    assert abs(100.0 * (1.0 - w.sizeFactor(8.0, 1.0)) - 18.775) <= 5e-3
    assert abs(100.0 * (1.0 - STEEL.sizeFactor(8.0, 1.0)) - 6.697) <= 5e-3
    assert abs(100.0 * (1.0 - IRON.sizeFactor(8.0, 1.0)) - 22.889) <= 5e-3


def test_weibull_hazard_additive_and_trap():  # :134-182
    rng = np.random.default_rng(20260919)
    v0, body = 1e-9, 3.7e-4
    for w in (STEEL, IRON):
        for parts in (2, 17, 500, 20000):
            cut = rng.uniform(0.1, 1.0, parts)
            total = float(cut.sum())
            x = 0.8 * w.scale
            summed = sum(w.forVolume(body * c / total, v0).cumulativeHazard(x) for c in cut)
            whole = w.forVolume(body, v0).cumulativeHazard(x)
            assert abs(whole - summed) <= whole * 1e-11
    assert STEEL.minimumOf(100000).scale < 0.8 * STEEL.minimumOf(100).scale
    coarse = STEEL.forVolume(body / 100, v0).minimumOf(100).scale
    fine = STEEL.forVolume(body / 100000, v0).minimumOf(100000).scale
    assert abs(coarse - fine) <= coarse * 1e-14
    assert abs(STEEL.forVolume(body, v0).scale - coarse) <= coarse * 1e-14


@pytest.mark.parametrize("m", [1.0, 2.0, 5.0, 8.0, 30.0, 50.0])
def test_weibull_mean_matches_integrated_survival(m):  # :185-196
    w = Weibull(0.42, m)
    n = 2_000_000
    top = w.scale * max(8.0, math.pow(27.6, 1.0 / m))
    h = top / n
    xs = np.arange(1, n) * h
    inner = np.exp(-np.power(xs / w.scale, m)).sum()
    area = 0.5 * (w.survival(0.0) + w.survival(top)) + inner
    assert abs(area * h - w.mean()) <= w.mean() * 2e-6


def test_weibull_closed_forms_and_gamma():  # :199-236
    e = Weibull(3.0, 1.0)
    assert abs(e.mean() - 3.0) <= 1e-14 and abs(e.standardDeviation() - 3.0) <= 1e-13
    assert abs(e.coefficientOfVariation() - 1.0) <= 1e-13
    ray = Weibull(3.0, 2.0)
    assert abs(ray.mean() - 3.0 * math.sqrt(math.pi) / 2.0) <= 1e-14
    assert abs(ray.standardDeviation() - 3.0 * math.sqrt(1.0 - math.pi / 4.0)) <= 1e-14
    assert abs(Weibull(1.0, 8.0).coefficientOfVariation() - Weibull(1e6, 8.0).coefficientOfVariation()) <= 1e-15
    assert abs(IRON.coefficientOfVariation() - 0.148369) <= 1e-5
    assert abs(STEEL.coefficientOfVariation() - 0.041769) <= 1e-5
    for x, v, tol in ((1.0, 1.0, 1e-14), (2.0, 1.0, 1e-14), (4.0, 6.0, 1e-13), (10.0, 362880.0, 1e-8),
                      (0.5, math.sqrt(math.pi), 1e-14), (1.5, math.sqrt(math.pi) / 2.0, 1e-14),
                      (-0.5, -2.0 * math.sqrt(math.pi), 1e-13)):
        assert abs(Weibull.gamma(x) - v) <= tol


def test_weibull_validation():  # :239-251
    for args in ((0.0, 10.0), (-1.0, 10.0), (1.0, 0.0), (1.0, math.nan), (math.nan, 10.0)):
        with pytest.raises(IllegalArgumentException):
            Weibull(*args)
    with pytest.raises(IllegalArgumentException):
        STEEL.minimumOf(0.0)
    with pytest.raises(IllegalArgumentException):
        STEEL.forVolume(1.0, 0.0)
    with pytest.raises(IllegalArgumentException):
        STEEL.quantile(1.5)


# ------------------------------------------------------------------ the pieces of the field

def mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_splitmix64_published_outputs():
    """mix(seed + k golden) is SplitMix64's k-th output; these five are the generator's
    published values for seed 1234567."""
    s, out = 1234567, []
    for _ in range(5):
        s = (s + 0x9E3779B97F4A7C15) & M64
        out.append(mix(s))
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def white_exact(seed, i, j):
    """DefectField.white (:325-332) with correctly rounded log, cos and sqrt."""
    h = mix((seed + 0x9E3779B97F4A7C15) & M64)
    h = mix(h ^ ((i * 0xBF58476D1CE4E5B9) & M64))
    h = mix(h ^ ((j * 0x94D049BB133111EB) & M64))
    u1 = ((h >> 11) + 0.5) * 2.0 ** -53
    u2 = ((mix(h ^ 0xD6E8FEB86659FD93) >> 11) + 0.5) * 2.0 ** -53
    t = 2.0 * math.pi * u2
    return math.sqrt(-2.0 * float(mpmath.log(mpmath.mpf(u1)))) * float(mpmath.cos(mpmath.mpf(t)))


def test_white_noise_is_box_muller_on_the_hash(field):
    f = field(20260919)
    rng = np.random.default_rng(1)
    i = rng.integers(-2**31, 2**31 - 1, 300).astype(np.float64)
    j = rng.integers(-2**31, 2**31 - 1, 300).astype(np.float64)
    i[:3], j[:3] = [0, -1, 2**31 - 1], [0, 5, -2**31]
    got = f.white(i, j)
    exact = np.array([white_exact(20260919, int(a), int(c)) for a, c in zip(i, j)])
    assert np.max(np.abs(got - exact)) <= 4e-16 * np.maximum(1.0, np.abs(exact)).max()
    with pytest.raises(IllegalArgumentException):
        f.white(0.5, 0.0)


def test_white_noise_is_standard_normal(field):
    f = field(7)
    a, c = np.meshgrid(np.arange(-150, 150), np.arange(-150, 150))
    w = f.white(a.astype(float), c.astype(float)).reshape(-1)
    assert abs(w.mean()) < 0.01 and abs(w.var() - 1.0) < 0.01
    assert st.kstest(w, "norm").pvalue > 1e-3


def smooth_brute(seed, r, z, cell, sigma_cells, radius=9):
    """DefectField.smooth (:289-319) as a plain double sum with math.exp weights."""
    gr, gz = r / cell, z / cell
    i0, j0 = math.floor(gr) - radius, math.floor(gz) - radius
    taps = 2 * radius + 2
    two = 2.0 * sigma_cells * sigma_cells
    wr = [math.exp(-(gr - (i0 + k)) ** 2 / two) for k in range(taps)]
    wz = [math.exp(-(gz - (j0 + k)) ** 2 / two) for k in range(taps)]
    total = 0.0
    for bb in range(taps):
        row = 0.0
        for a in range(taps):
            row += wr[a] * white_exact(seed, i0 + a, j0 + bb)
        total += wz[bb] * row
    return total / math.sqrt(sum(v * v for v in wr) * sum(v * v for v in wz))


def test_smoothing_against_brute_force(field):
    seed = 42
    f = field(seed)
    cell = LENGTH / 3
    sigma = (LENGTH / math.sqrt(2.0)) / cell
    rng = np.random.default_rng(3)
    r = np.concatenate([rng.uniform(-0.01, 0.2, 6), [0.0, 0.05]])
    z = np.concatenate([rng.uniform(-0.05, 0.05, 6), [0.0, 0.02]])
    got = f.gaussianAt(r, z)
    brute = np.array([smooth_brute(seed, a, c, cell, sigma) for a, c in zip(r, z)])
    assert np.max(np.abs(got - brute)) < 1e-13


def test_transform_is_the_weibull_quantile_of_the_normal_cdf(field):
    f = field(1)
    g = np.concatenate([np.linspace(-9.0, 9.0, 181), [-12.0, 12.0, 0.0, -0.0]])
    for volume in (1e-12, 1e-9, 3.7e-7, 1.0):
        got = f.transform(g, np.full(g.size, volume))
        local = BASE.forVolume(volume, V0)
        c = np.clip(g, -9.0, 9.0)
        expected = np.where(c <= 0.0,
                            st.weibull_min.ppf(st.norm.cdf(c), BASE.modulus, scale=local.scale),
                            st.weibull_min.isf(st.norm.sf(c), BASE.modulus, scale=local.scale))
        assert np.max(np.abs(got - expected) / expected) < 1e-12, volume
    assert np.all(np.isfinite(got)) and np.all(got > 0.0)
    with pytest.raises(IllegalArgumentException):
        f.transform(0.0, 0.0)
    with pytest.raises(IllegalArgumentException):
        f.failureStrainForVolume(0.05, 0.01, -1.0)


# ------------------------------------------------------------------ the field (DefectFieldTest.java)

@pytest.mark.parametrize("seed", [1, 7, 20260919])
def test_gaussian_marginal_is_standard(field, seed):  # :42-66
    f = field(seed)
    i, j = np.meshgrid(np.arange(300), np.arange(300), indexing="ij")
    g = f.gaussianAt(0.05 + i * 0.37e-3, 0.01 + j * 0.41e-3).reshape(-1)
    mean = g.sum() / g.size
    assert abs(mean) <= 0.05
    assert abs((g * g).sum() / g.size - mean * mean - 1.0) <= 0.07


def test_correlation_follows_the_kernel(field):  # :69-93
    f = field(11)
    i, j = np.meshgrid(np.arange(260), np.arange(260), indexing="ij")
    r = 0.05 + i * 0.29e-3
    z = 0.01 + j * 0.31e-3
    a = f.gaussianAt(r, z)
    for lengths in (0.25, 0.5, 1.0, 1.5, 2.0):
        d = lengths * LENGTH
        bv = f.gaussianAt(r + d, z)
        measured = (a * bv).sum() / math.sqrt((a * a).sum() * (bv * bv).sum())
        assert abs(measured - f.correlation(d)) <= 0.03, lengths


def test_correlation_endpoints(field):  # :96-107
    f = field(3)
    assert f.correlation(0.0) == 1.0 and f.correlation(5.0 * LENGTH) < 1e-5
    a = f.gaussianAt(0.06, 0.02)
    assert f.gaussianAt(0.06, 0.02) == a
    assert abs(a - f.gaussianAt(0.06 + 0.02 * LENGTH, 0.02)) < 0.05


def test_two_meshes_see_the_same_field(field):  # :112-145
    f, again, other = field(42), field(42), field(43)
    coarse = QuadMesh.cylinderWall(0.05, 0.10, 0.02, 4, 2)
    fine = QuadMesh.cylinderWall(0.05, 0.10, 0.02, 64, 32)
    assert abs(coarse.totalRingVolume() - fine.totalRingVolume()) <= coarse.totalRingVolume() * 1e-12
    r, z = np.meshgrid(np.arange(0.051, 0.0999, 0.0037), np.arange(0.001, 0.0199, 0.0031), indexing="ij")
    direct = f.gaussianAt(r, z)
    assert np.array_equal(direct, again.gaussianAt(r, z))
    assert np.all(direct != other.gaussianAt(r, z))
    sampled = f.sampleOnto(coarse)
    rc = np.array([coarse.centroidRadius(e) for e in range(coarse.elementCount)])
    zc = np.array([coarse.centroidZ(e) for e in range(coarse.elementCount)])
    assert np.array_equal(f.failureStrainAt(rc, zc, coarse.signedAreas()), sampled)
    # the memoised grid and hashing every tap give the same bits, on both meshes
    assert np.array_equal(f.sampleOnto(fine), f.sampleOnto(fine, cached=False))
    assert np.array_equal(sampled, f.sampleOnto(coarse, cached=False))


def test_resolution_floor(field):  # :148-178
    f = field(1)
    r = 0.075
    coarse = 4.0 * LENGTH * LENGTH
    v = f.effectiveVolume(r, coarse)
    assert abs(v - 2.0 * math.pi * r * coarse) <= 2.0 * math.pi * r * coarse * 1e-12
    fine = 0.25 * LENGTH * LENGTH
    assert abs(f.effectiveVolume(r, fine) - 2.0 * math.pi * r * LENGTH * LENGTH) <= 2.0 * math.pi * r * LENGTH ** 2 * 1e-12
    assert f.effectiveVolume(r, fine) == f.effectiveVolume(r, fine / 1000.0)
    assert abs(v / (LENGTH * coarse) - 2.0 * math.pi * r / LENGTH) <= 1e-9
    assert abs(f.effectiveVolume(1e-9, coarse) - 4.0 * LENGTH ** 3) <= 4e-18


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6, 7, 8])
def test_refinement_does_not_move_the_body(field, seed):  # :181-219
    f = field(seed)
    weakest, counts = [], []
    for refine in (2, 8, 32, 64):
        mesh = QuadMesh.cylinderWall(0.05, 0.10, 0.02, 4 * refine, 2 * refine)
        weakest.append(f.sampleOnto(mesh).min())
        counts.append(mesh.elementCount)
    assert counts[0] == 32 and counts[3] == 32768
    assert abs(weakest[3] - weakest[2]) / weakest[2] < 0.03
    assert 0.6 * weakest[0] < weakest[3] < 1.4 * weakest[0]


def test_body_matches_the_closed_form(field):  # :222-256
    mesh = QuadMesh.cylinderWall(0.05, 0.10, 0.02, 64, 32)
    f = field(7000)
    weakest = f.sampleOnto(mesh, seeds=7000 + np.arange(300)).min(axis=1)
    mean = weakest.mean()
    sd = math.sqrt(((weakest - mean) ** 2).sum() / (weakest.size - 1))
    predicted = BASE.forVolume(mesh.totalRingVolume(), V0)
    assert abs(mean - predicted.mean()) <= predicted.mean() * 0.04
    assert abs(sd - predicted.standardDeviation()) <= predicted.standardDeviation() * 0.25


def test_size_effect_is_measured(field):  # :259-300
    f = field(9000)
    seeds = 9000 + np.arange(3000)
    volumes, means = [], []
    for ri in (0.02, 0.04, 0.08, 0.16):
        mesh = QuadMesh.cylinderWall(ri, ri + 0.005, 0.02, 8, 16)
        means.append(f.sampleOnto(mesh, seeds=seeds).min(axis=1).sum() / seeds.size)
        volumes.append(mesh.totalRingVolume())
    for i in range(1, 4):
        ratio = volumes[i] / volumes[0]
        measured = means[i] / means[0]
        assert abs(measured - ratio ** (-1.0 / BASE.modulus)) <= 0.005
        assert measured < 1.0
    assert means[3] < 0.94 * means[0]


def test_unnormalised_field_collapses(field):  # :303-323
    f = field(5)

    def weakest(mesh):
        rc = np.array([mesh.centroidRadius(e) for e in range(mesh.elementCount)])
        zc = np.array([mesh.centroidZ(e) for e in range(mesh.elementCount)])
        return f.transform(f.gaussianAt(rc, zc), np.full(rc.size, V0)).min()
    coarse = weakest(QuadMesh.cylinderWall(0.05, 0.10, 0.02, 8, 4))
    fine = weakest(QuadMesh.cylinderWall(0.05, 0.10, 0.02, 128, 64))
    assert fine < 0.8 * coarse


def test_marginal_is_weibull(field):  # :328-368
    mesh = QuadMesh.cylinderWall(0.05, 0.10, 0.05, 1, 400)
    f = field(99)
    v = f.effectiveVolume(mesh.centroidRadius(0), mesh.signedArea(0))
    assert abs(v - mesh.ringVolume(0)) <= v * 1e-12
    rc = np.array([mesh.centroidRadius(e) for e in range(mesh.elementCount)])
    assert np.all(np.abs(f.effectiveVolume(rc, mesh.signedAreas()) - v) <= v * 1e-12)
    strain = f.sampleOnto(mesh, seeds=1000 + np.arange(200))
    pooled = np.sort(strain[:, ::33][:, :12].reshape(-1))
    n = pooled.size
    expected = BASE.forVolume(v, V0)
    cdf = np.array([expected.cdf(x) for x in pooled])
    i = np.arange(n)
    ks = max(np.max(cdf - i / n), np.max((i + 1) / n - cdf))
    assert ks < 1.63 / math.sqrt(n)


def test_ring_effect(field):  # :371-398
    narrow = QuadMesh.cylinderWall(0.030, 0.035, 0.02, 4, 16)
    wide = QuadMesh.cylinderWall(0.060, 0.065, 0.02, 4, 16)
    ratio = wide.totalRingVolume() / narrow.totalRingVolume()
    assert abs(ratio - 1.9231) <= 1e-3
    f = field(77)
    nb, wb = f.forBody(narrow.totalRingVolume()), f.forBody(wide.totalRingVolume())
    assert abs(wb.scale / nb.scale - ratio ** (-1.0 / BASE.modulus)) <= 1e-12
    assert abs(100.0 * (1.0 - wb.scale / nb.scale) - 3.21) <= 0.02


def test_repeated_runs_scatter(field):  # :401-425
    mesh = QuadMesh.cylinderWall(0.05, 0.10, 0.02, 16, 8)
    weakest = field(5000).sampleOnto(mesh, seeds=5000 + np.arange(400)).min(axis=1)
    mean = weakest.mean()
    spread = math.sqrt(((weakest - mean) ** 2).sum() / (weakest.size - 1)) / mean
    assert 0.005 < spread < 0.10
    assert weakest.min() < 0.9 * mean


def test_no_degenerate_values(field):  # :430-447
    mesh = QuadMesh.cylinderWall(0.05, 0.10, 0.02, 32, 16)
    x = field(0).sampleOnto(mesh, seeds=np.arange(40))
    assert np.all(np.isfinite(x)) and np.all(x > 0.0)
    f = field(1)
    assert f.failureStrainForVolume(0.05, 0.01, 1.0) > 0.0
    assert math.isfinite(f.failureStrainForVolume(0.05, 0.01, 1e-15))


def test_axis_elements_are_fine(field):  # :450-461
    solid = QuadMesh.solidCylinder(0.0038, 0.0254, 8, 32)
    assert abs(solid.totalRingVolume() - math.pi * 0.0038 ** 2 * 0.0254) <= 1e-12
    x = field(9).sampleOnto(solid)
    assert np.all(np.isfinite(x)) and np.all(x > 0.0)


def test_validation(oracle):  # :464-470
    for v0, length in ((0.0, LENGTH), (V0, 0.0), (V0, math.nan), (V0, math.inf)):
        with pytest.raises(IllegalArgumentException):
            oracle.OracleDefectField(BASE, v0, length, 1)


def test_slab_rule_is_the_plane_strain_count(field):
    """The plane-strain rule (no counterpart in the reference): l^3 max(1, A/l^2)
    max(1, depth/l) through failureStrainForVolume, element by element."""
    mesh = QuadMesh.cylinderWall(0.0, 0.02, 0.01, 40, 20)
    f = field(5)
    depth = 0.02
    got = f.sampleOnto(mesh, depth=depth)
    rc = np.array([mesh.centroidRadius(e) for e in range(mesh.elementCount)])
    zc = np.array([mesh.centroidZ(e) for e in range(mesh.elementCount)])
    a = mesh.signedAreas()
    vol = LENGTH ** 3 * np.maximum(1.0, a / LENGTH ** 2) * max(1.0, depth / LENGTH)
    assert np.array_equal(got, f.failureStrainForVolume(rc, zc, vol))
    with pytest.raises(IllegalArgumentException):
        f.sampleOnto(mesh, depth=-1.0)

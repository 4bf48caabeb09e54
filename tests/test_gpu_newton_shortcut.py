"""The Johnson-Cook solve's proven bisections, on the GPU.

The element kernels skip the leading iterations of the bracketed Newton solve
(solver/J2.java:332-375) whose decisions -- g <= 0, |g| above the tolerance, the Newton step
leaving the bracket through zero -- are proven by a single-precision estimate with a 2^-14
margin, because the state such an iteration leaves does not depend on g's value. Checked
here: the hardware's single-precision log2, exp2 and reciprocal that the estimate uses are
well inside that margin over the arguments the solve can give them; the solve with the
shortcut returns the same bits as the solve that evaluates every iteration, over a million
solves that sweep the material, the plastic strain, the thermal factor, the time step and the
overstress from 1e-12 (points that barely yield and hit the 25-iteration cap) to 1; and the
shortcut is actually taken. The rod, batch, adjoint and point-path parity tests run the
kernels with the shortcut against the oracle, which evaluates every iteration.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

dp = C.POINTER(C.c_double)


def evaluate(hip, fn, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.size if fn < 21 else x.size // 5
    out = np.empty(n)
    yy = np.zeros(n) if y is None else np.ascontiguousarray(y, dtype=np.float64)
    assert hip.math_eval(fn, x.ctypes.data_as(dp), yy.ctypes.data_as(dp), out.ctypes.data_as(dp), n) == 0
    return out


def test_single_precision_hardware_functions_within_the_margin(hip):
    """log2 to 2^-21 (absolute, times max(1, |log2 x|)), exp2 and the reciprocal to 2^-21
    relative, over the arguments the estimate gives them: the shortcut's error analysis
    assumes errors near 1e-6 at most against its 2^-14 margin."""
    rng = np.random.default_rng(7)
    x = np.exp(rng.uniform(np.log(1e-13), np.log(1e16), 400_000))
    xf = x.astype(np.float32).astype(np.float64)
    want = np.log2(xf)
    got = evaluate(hip, 18, x)
    err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
    print(f"log2: max scaled error {err.max():.2e}")
    assert err.max() <= 2.0 ** -21

    z = rng.uniform(-60.0, 10.0, 400_000)
    zf = z.astype(np.float32).astype(np.float64)
    got = evaluate(hip, 19, z)
    err = np.abs(got - np.exp2(zf)) / np.exp2(zf)
    print(f"exp2: max relative error {err.max():.2e}")
    assert err.max() <= 2.0 ** -21

    got = evaluate(hip, 20, x)
    err = np.abs(got * xf - 1.0)
    print(f"rcp: max relative error {err.max():.2e}")
    assert err.max() <= 2.0 ** -21


def solves(rng, n, eps0):
    """n solve inputs and one material: a Johnson-Cook law drawn around OFHC copper and 4340
    steel, a shear modulus of 40-80 GPa, plastic strain 0 or log-uniform in [1e-9, 8], a
    thermal factor in [0, 1] (0 and 1 included), dt 0 or log-uniform in [1e-12, 1e-6], and
    norm = sqrt(2/3) sigma_y (1 + overstress) with the overstress log-uniform in [1e-12, 1]
    (sigma_y the quasi-static flow stress at that plastic strain and thermal factor, so every
    solve yields)."""
    a = rng.uniform(50e6, 800e6)
    b = rng.uniform(0.0, 600e6)
    nexp = rng.uniform(0.05, 1.0)
    c = rng.choice([0.0, rng.uniform(0.0, 0.1)])
    mu = rng.uniform(40e9, 80e9, n)
    stiffness = 2.0 * mu
    ps = np.where(rng.random(n) < 0.2, 0.0, np.exp(rng.uniform(np.log(1e-9), np.log(8.0), n)))
    thermal = rng.uniform(0.0, 1.0, n)
    thermal[:16] = 1.0
    thermal[16:32] = 0.0
    dt = np.where(rng.random(n) < 0.05, 0.0, np.exp(rng.uniform(np.log(1e-12), np.log(1e-6), n)))
    flow = (a + b * ps ** nexp) * thermal
    over = np.exp(rng.uniform(np.log(1e-12), 0.0, n))
    norm = np.sqrt(2.0 / 3.0) * flow * (1.0 + over)
    norm = np.where(norm > 0.0, norm, rng.uniform(1e6, 1e8, n))
    x = np.stack([norm, stiffness, ps, thermal, dt], axis=1)
    y = np.array([a, b, nexp, c, eps0])
    return x, y


@pytest.mark.parametrize("eps0", [1.0, 1e-3, 10.0])
def test_shortcut_solve_equals_every_iteration(hip, eps0):
    rng = np.random.default_rng(int(eps0 * 1000) + 11)
    total = proven_total = capped = 0
    for trial in range(8):
        x, y = solves(rng, 131_072, eps0)
        fast = evaluate(hip, 21, x, y)
        slow = evaluate(hip, 22, x, y)
        proven = evaluate(hip, 23, x, y)
        bad = np.nonzero(fast.view(np.uint64) != slow.view(np.uint64))[0]
        assert bad.size == 0, (trial, y, x[bad[0]], fast[bad[0]], slow[bad[0]])
        total += fast.size
        proven_total += proven.sum()
        capped += int((proven == 25).sum())
        assert proven.min() >= 0 and proven.max() <= 25
    print(f"eps0 {eps0}: {total} solves bit-identical, {proven_total / total:.2f} bisections proven "
          f"per solve, {capped} solves proven to the cap")
    assert proven_total / total > 2.0

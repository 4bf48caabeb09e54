"""The shared deterministic transcendentals (neofiz_amd/csrc/nf_math.h) against correctly
rounded values from mpmath: log, exp and pow within 0.6 ulp and correctly rounded in the vast
majority of cases; the trigonometric functions used only at damage onset within a few ulp."""
import ctypes as C
import math

import mpmath
import numpy as np
import pytest

mpmath.mp.prec = 200

CASES = {
    # fn: (name, generator, reference, max ulp, min fraction correctly rounded)
    0: ("log", lambda r: (np.exp(r.uniform(-700, 700, 4000)), None), mpmath.log, 0.52, 0.99),
    1: ("exp", lambda r: (r.uniform(-700, 700, 4000), None), mpmath.exp, 0.56, 0.98),
    2: ("pow", lambda r: (np.exp(r.uniform(-30, 3, 4000)), r.uniform(0.05, 1.5, 4000)),
        lambda x, y: mpmath.power(x, y), 0.56, 0.98),
    3: ("atan2", lambda r: (r.uniform(-2, 2, 4000), r.uniform(-2, 2, 4000)), mpmath.atan2, 4.0, 0.5),
    4: ("sin", lambda r: (r.uniform(-1.6, 1.6, 4000), None), mpmath.sin, 1.5, 0.8),
    5: ("cos", lambda r: (r.uniform(-1.6, 1.6, 4000), None), mpmath.cos, 1.5, 0.8),
    6: ("hypot", lambda r: (r.uniform(-2, 2, 4000), r.uniform(-2, 2, 4000)),
        lambda a, b: mpmath.sqrt(a * a + b * b), 2.0, 0.5),
    # The defect field's Box-Muller angle 2 pi u spans the whole turn, and its copula takes
    # log1p(-p) for p in (0, 1/2] (core/DefectField.java:331, core/Weibull.java:118).
    10: ("cos, whole turn", lambda r: (r.uniform(0.0, 2.0 * math.pi, 4000), None), mpmath.cos, 1.5, 0.8),
    9: ("log1p", lambda r: (-np.exp(r.uniform(-60, math.log(0.5), 4000)), None), mpmath.log1p, 0.51, 0.99),
}
FN = {10: 5}


def evaluate(binding, fn, x, y):
    dp = C.POINTER(C.c_double)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), np.float64)
    out = np.empty_like(x)
    assert binding.math_eval(fn, x.ctypes.data_as(dp), y.ctypes.data_as(dp), out.ctypes.data_as(dp), x.size) == 0
    return out


@pytest.mark.parametrize("fn", sorted(CASES))
def test_against_correctly_rounded(oracle, fn):
    name, gen, ref, max_ulp, min_cr = CASES[fn]
    rng = np.random.default_rng(fn)
    x, y = gen(rng)
    got = evaluate(oracle.binding(), FN.get(fn, fn), x, y)
    worst, cr = 0.0, 0
    for i in range(x.size):
        args = (mpmath.mpf(float(x[i])),) if y is None else (mpmath.mpf(float(x[i])), mpmath.mpf(float(y[i])))
        exact = ref(*args)
        e = float(exact)
        err = float(abs(mpmath.mpf(float(got[i])) - exact) / math.ulp(e)) if e != 0 else abs(got[i])
        worst = max(worst, err)
        cr += got[i] == e
    assert worst <= max_ulp, (name, worst)
    assert cr / x.size >= min_cr, (name, cr / x.size)


def test_special_values(oracle):
    b = oracle.binding()
    one = np.array([1.0])
    assert evaluate(b, 0, one, None)[0] == 0.0
    assert evaluate(b, 1, np.array([0.0]), None)[0] == 1.0
    x = np.array([0.37, 2.5, 1e-300])
    assert np.array_equal(evaluate(b, 2, x, np.ones(3)), x)          # pow(x, 1) = x
    assert np.array_equal(evaluate(b, 2, x, np.zeros(3)), np.ones(3))  # pow(x, 0) = 1
    assert math.isnan(evaluate(b, 0, np.array([-1.0]), None)[0])
    assert evaluate(b, 0, np.array([0.0]), None)[0] == -math.inf
    assert evaluate(b, 1, np.array([800.0]), None)[0] == math.inf
    assert evaluate(b, 1, np.array([-800.0]), None)[0] == 0.0
    tiny = np.array([0.0, -0.0, 1e-300, -3e-17])
    assert np.array_equal(np.signbit(evaluate(b, 9, tiny, None)), np.signbit(tiny))
    assert np.array_equal(evaluate(b, 9, tiny, None), tiny)      # log1p x = x below 2^-54
    assert evaluate(b, 9, np.array([-1.0]), None)[0] == -math.inf
    assert math.isnan(evaluate(b, 9, np.array([-1.5]), None)[0])


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_ordinary_input_variants_are_pow_and_log(oracle):
    """pow_ordinary / log_ordinary / pow_log_ordinary (functions 10-13: the Newton loop's pow
    and log with the special inputs out of line, and the derivative pass's pow with its shared
    log) return pow's and log's own bits on every input: ordinary ones over many binades, and
    every special base and exponent."""
    b = oracle.binding()
    rng = np.random.default_rng(11)
    specials = np.array([0.0, -0.0, 1.0, -1.0, 2.0, 0.5, 5e-324, 2.2250738585072014e-308, 1e-310,
                         1.7976931348623157e308, math.inf, -math.inf, math.nan, -2.5])
    xs = np.concatenate([np.exp(rng.uniform(-740, 709, 3000)), rng.uniform(0.5, 2.0, 1000), specials])
    ys = np.concatenate([rng.uniform(-3, 3, 2000), rng.uniform(0.05, 1.5, 2000), specials])
    x = np.repeat(xs, 4)[: 4 * xs.size]
    y = np.resize(ys, x.size)
    xx, yy = np.meshgrid(specials, specials)
    x = np.concatenate([x, xx.ravel()])
    y = np.concatenate([y, yy.ravel()])
    assert np.array_equal(_bits(evaluate(b, 10, x, y)), _bits(evaluate(b, 2, x, y)))
    assert np.array_equal(_bits(evaluate(b, 11, xs, None)), _bits(evaluate(b, 0, xs, None)))
    # the derivative pass's pow with the logarithm it shares (functions 12, 13)
    assert np.array_equal(_bits(evaluate(b, 12, x, y)), _bits(evaluate(b, 2, x, y)))
    assert np.array_equal(_bits(evaluate(b, 13, x, y)), _bits(evaluate(b, 0, x, None)))
    # the Newton loop's pow(x, y) and log(y) side by side (functions 14, 15), and the yield
    # check's pow(x, y) and pow(y, x) side by side (16, 17)
    assert np.array_equal(_bits(evaluate(b, 14, x, y)), _bits(evaluate(b, 2, x, y)))
    assert np.array_equal(_bits(evaluate(b, 15, x, y)), _bits(evaluate(b, 0, y, None)))
    assert np.array_equal(_bits(evaluate(b, 16, x, y)), _bits(evaluate(b, 2, x, y)))
    assert np.array_equal(_bits(evaluate(b, 17, x, y)), _bits(evaluate(b, 2, y, x)))
    # the Newton loop's pair for an exponent in (0, 1), without exp's range selects (24, 25)
    inside = (y > 0.0) & (y < 1.0)
    assert inside.sum() > 1000
    assert np.array_equal(_bits(evaluate(b, 24, x[inside], y[inside])), _bits(evaluate(b, 2, x[inside], y[inside])))
    assert np.array_equal(_bits(evaluate(b, 25, x[inside], y[inside])), _bits(evaluate(b, 0, y[inside], None)))

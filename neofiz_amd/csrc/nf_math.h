// Deterministic double-precision transcendentals, identical bit-for-bit on the host (g++,
// x86-64 SSE2) and on gfx950.
//
// Why these exist: the Johnson-Cook return map calls pow and log inside its Newton loop
// (core/JohnsonCook.java:139, :182, :196) and damage onset calls atan2, sin, cos and hypot
// (solver/Damage.java:315, :321-322). Library implementations differ between glibc and the
// AMD device library in the last bit, and a plastic run amplifies a last-bit difference over
// tens of thousands of steps. Both the CPU restatement and the HIP kernels therefore call the
// functions below, which use only IEEE-754 add, subtract, multiply, divide, square root and
// scaleB by a power of two (each correctly rounded on both targets), integer bit manipulation
// and comparisons. The file must be compiled with floating-point contraction off
// (-ffp-contract=off) on every target, so that no multiply-add is fused on one side and not
// the other.
//
// Methods (all from the mathematics; the tables in nf_math_tables.h are correctly rounded
// values computed by gen_math_tables.py with mpmath):
//  - log: Tang's table method (128 points per octave around 1, Taylor series of log1p to
//    degree 8 on |r| <= 2^-7.5), double-double recombination with Knuth's two-sum and an
//    exact product by Veltkamp splitting. No division: the dependency chain is short, which
//    is what the Johnson-Cook Newton loop waits on.
//  - log1p: log of the rounded 1 + x as a double-double, plus the two-sum error of 1 + x
//    divided by it.
//  - exp: table of 2^(j/64) and Cody-Waite reduction by ln2/64 (split constant), Taylor
//    series of expm1 to degree 7 on |r| <= ln2/128.
//  - pow(x, y) for x > 0: exp(y log x) with log x carried as a double-double.
//  - atan: reduction atan t = pi/4 + atan((t-1)/(t+1)) and one half-angle step, then the
//    Gregory series; atan2 by quadrant.
//  - sin, cos: reduction by pi/2 (two-part constant) to |r| <= pi/4, Taylor series.
//  - hypot: scaled sqrt(a^2 + b^2).
// Accuracy against correctly rounded values (mpmath) is checked in the test suite.
#pragma once

#include <stdint.h>

#include "nf_math_tables.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define NF_HD __host__ __device__ __forceinline__
#else
#define NF_HD static inline
#endif

namespace nfm {

NF_HD uint64_t bits_of(double x) {
    uint64_t u;
    __builtin_memcpy(&u, &x, sizeof u);
    return u;
}

NF_HD double from_bits(uint64_t u) {
    double x;
    __builtin_memcpy(&x, &u, sizeof x);
    return x;
}

NF_HD double inf() { return from_bits(0x7FF0000000000000ULL); }
NF_HD double qnan() { return from_bits(0x7FF8000000000000ULL); }
NF_HD bool is_nan(double x) { return x != x; }

// Error-free transforms. Contraction must be off for these to be exact.
NF_HD void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

NF_HD void fast_two_sum(double a, double b, double& s, double& e) {  // requires |a| >= |b|
    s = a + b;
    e = b - (s - a);
}

NF_HD void split(double a, double& hi, double& lo) {
    const double t = 134217729.0 * a;  // 2^27 + 1
    hi = t - (t - a);
    lo = a - hi;
}

// Exact product a b = p + e. The error of a rounded product is itself a double, and
// fma(a, b, -p) computes it exactly: a fused multiply-add used deliberately, as an exact
// operation (the compiler never contracts anything implicitly here), which gives the same
// bits on every IEEE target -- the GPU's v_fma_f64 and the host's correctly rounded fma().
NF_HD void two_prod(double a, double b, double& p, double& e) {
    p = a * b;
    e = __builtin_fma(a, b, -p);
}

// log(x) as an unevaluated sum hi + lo (|lo| <= ulp(hi)) for positive finite x.
// Tang's table method: x = 2^k m, m in [3/4, 3/2); F = j/128 nearest to m (j = 96..192);
// f = m - F exactly; r = f/F carried as r + r_lo (r = f * (1/F), r_lo from the exact
// remainder f - r F); log(1 + r) = r + r^2 (-1/2 + r/3 - ... - r^6/8) by Estrin's scheme,
// |r| <= 2^-7.5; log x = k ln2 + log F + log(1 + r), summed as a double-double. No
// division.
// The tables are passed in so that a GPU kernel can read them from LDS (the same values,
// copied there once per workgroup); the overloads below use the constant arrays directly.
struct Tables {
    const double (*log)[3];
    const double (*exp)[2];
};

// MAY_BE_SUB = false is the same computation for a normal x, without the subnormal select.
template <bool MAY_BE_SUB>
NF_HD void log_dd_impl(double x, double& hi, double& lo, const double (*LT)[3]) {
    // Branch-free: every range decision is a select, so the compiler can interleave this
    // chain with independent work (the Newton loop evaluates a pow and a log side by side).
    const uint64_t b0 = bits_of(x);
    const bool sub = MAY_BE_SUB && b0 < 0x0010000000000000ULL;  // subnormal (or zero)
    const double xs = sub ? x * 0x1p54 : x;
    const uint64_t b = bits_of(xs);
    // x = 2^k m, m in [3/4, 3/2), in exact integer arithmetic on the positive normal xs:
    // subtracting the bits of 3/4 leaves k in the exponent field (a mantissa of 1.5 or more
    // carries into it, which is the case m = mantissa / 2), and taking k back out of the
    // exponent leaves m.
    // The low word of the bits does not take part.
    const uint32_t bh = (uint32_t)(b >> 32);
    const uint32_t th = bh - 0x3FE80000u;
    const int k = ((int32_t)th >> 20) - (sub ? 54 : 0);
    const double m = from_bits(((uint64_t)(bh - (th & 0xFFF00000u)) << 32) | (b & 0xFFFFFFFFULL));
    const double jd = __builtin_rint(m * 128.0);  // nearest integer (ties to even), 96..192
    // The byte offset of F = jd / 128's row, 24 (jd - 96), in double arithmetic on small
    // integers (exact), which the device computes at full rate, unlike an integer multiply.
    const int off = (int)(jd * 24.0 - 2304.0);
    const double* row = (const double*)((const char*)&LT[0][0] + off);
    const double F = jd * 0x1p-7;
    const double f = m - F;  // exact (Sterbenz)
    // r = f (1/F) and the remainder f - r F to full precision: r F = p + e exactly (e by an
    // exact fma), f - p exactly (Sterbenz), so r_lo = ((f - p) - e) / F carries the rest.
    const double invF = row[0];
    const double r = f * invF;
    double p, e;
    two_prod(r, F, p, e);
    const double r_lo = ((f - p) - e) * invF;
    // log(1 + r) - r, Estrin
    const double r2 = r * r;
    const double r4 = r2 * r2;
    const double a = -0.5 + r * (1.0 / 3.0);
    const double bq = -0.25 + r * (1.0 / 5.0);
    const double c = -1.0 / 6.0 + r * (1.0 / 7.0);
    const double q = (a + bq * r2) + (c + (-0.125) * r2) * r4;
    const double P = r2 * q;
    const double kd = (double)k;
    double s1, e1, s2, e2;
    // Both sums are exact with the three-operation form: |k ln2| >= 0.69 > 0.41 >= |log F|
    // when k != 0, and for k = 0 and F != 1, |log F| >= log(128/127) > 2^-7.5 >= |r| (for
    // F = 1 the first operand is 0).
    fast_two_sum(kd * NFM_LN2_HI, row[1], s1, e1);
    fast_two_sum(s1, r, s2, e2);
    const double l = (e1 + e2) + ((kd * NFM_LN2_LO + row[2]) + (r_lo + P));
    fast_two_sum(s2, l, hi, lo);
}

NF_HD void log_dd(double x, double& hi, double& lo, const double (*LT)[3]) {
    log_dd_impl<true>(x, hi, lo, LT);
}

NF_HD double log(double x, const Tables& T) {
    double hi, lo;
    // the main path on a harmless stand-in, then the special values by selection
    const bool ok = x > 0.0 && x < inf();
    log_dd(ok ? x : 1.0, hi, lo, T.log);
    const double special = x == 0.0 ? -inf() : (x == inf() ? x : qnan());
    return ok ? hi + lo : special;
}

NF_HD double log(double x) { return log(x, Tables{NFM_LOG_TBL, NFM_EXP_TBL}); }

// log(1 + x). For |x| >= 2^-20: u = 1 + x rounds, with an exact error c (Knuth's two-sum), so
// log(1 + x) = log(u) + log(1 + c/u) = log(u) + c/u, the dropped terms and the rounding of
// c/u below 2^-80 of the result; log u is the double-double above. For |x| < 2^-20 the series
// x - x^2/2 + x^3/3 - x^4/4, whose truncation is below 2^-80 relative; x*x underflows
// harmlessly, and a zero keeps its sign.
NF_HD double log1p(double x) {
    double u, c;
    two_sum(1.0, x, u, c);
    const bool ok = u > 0.0 && u < inf();
    double hi, lo;
    log_dd(ok ? u : 1.0, hi, lo, NFM_LOG_TBL);
    const double v = hi + (lo + c / u);
    const double series = x + (x * x) * (-0.5 + x * (1.0 / 3.0 + x * -0.25));
    const double special = is_nan(x) ? x : (u == 0.0 ? -inf() : (u == inf() ? u : qnan()));
    const double ax = x < 0.0 ? -x : x;
    return ax < 0x1p-20 ? series : (ok ? v : special);
}

// exp(a + b) with |b| <= ulp(a). Table method: a + b = (64 k + j) ln2/64 + r, |r| <= ln2/128;
// exp = 2^k 2^(j/64) (1 + r + r^2 (1/2 + r/6 + ... + r^5/5040)), the polynomial by Estrin's
// scheme and the table value carried as hi + lo. Branch-free: the scaling by 2^k is IEEE's
// scaleB (ldexp: exact, or one rounding into the subnormal range), and the out-of-range
// results are selected at the end.
// The main path of exp_dd for a in [-745.2, 709.782712893384].
NF_HD double exp_dd_main(double a, double b, const double (*ET)[2]) {
    const double nd = __builtin_rint(a * NFM_INV_L64);  // nearest integer (ties to even)
    const int n = (int)nd;
    const int j = n & 63;
    const int k = (n - j) / 64;
    const double t = a - nd * NFM_L64_HI;  // exact
    double r, rl;
    two_sum(t, b - nd * NFM_L64_LO, r, rl);
    const double r2 = r * r;
    const double qa = 0.5 + r * (1.0 / 6.0);
    const double qb = 1.0 / 24.0 + r * (1.0 / 120.0);
    const double qc = 1.0 / 720.0 + r * (1.0 / 5040.0);
    const double Q = (qa + qb * r2) + qc * (r2 * r2);
    const double E = rl + r2 * Q;  // exp(r + rl) = 1 + r + E
    const double p = r + E;
    const double th = ET[j][0], tl = ET[j][1];
    const double y = th + (tl + (th * r + (th * E + tl * p)));
    // y 2^k with one correct rounding (IEEE scaleB: exact unless the result is subnormal)
    return __builtin_ldexp(y, k);
}

NF_HD double exp_dd(double a0, double b, const double (*ET)[2]) {
    const bool nan = is_nan(a0);
    const bool over = a0 > 709.782712893384;
    const bool under = a0 < -745.2;
    const double a = (nan || over || under) ? 0.0 : a0;
    const double v = exp_dd_main(a, b, ET);
    return nan ? a0 : (over ? inf() : (under ? 0.0 : v));
}

NF_HD double exp(double x) { return exp_dd(x, 0.0, NFM_EXP_TBL); }

// The main path of pow below: exp(y log x) with log x as a double-double.
NF_HD double pow_main(double x, double y, const Tables& T) {
    double l_hi, l_lo;
    log_dd(x, l_hi, l_lo, T.log);
    double p_hi, p_lo;
    two_prod(y, l_hi, p_hi, p_lo);
    p_lo = p_lo + y * l_lo;
    double t_hi, t_lo;
    fast_two_sum(p_hi, p_lo, t_hi, t_lo);
    return exp_dd(t_hi, t_lo, T.exp);
}

// x^y for the cases the solver needs: x > 0 finite, y finite. Special values follow IEEE pow
// for x = 0, x = 1, y = 0, y = 1 (exactly x), NaN; a negative base yields NaN. Branch-free.
NF_HD double pow(double x, double y, const Tables& T) {
    const bool regular = x > 0.0 && x < inf() && !is_nan(y);
    const double v = pow_main(regular ? x : 1.0, y, T);
    double special;
    if (is_nan(x) || is_nan(y)) special = qnan();
    else if (x == 0.0) special = y > 0.0 ? 0.0 : inf();
    else if (x < 0.0) special = qnan();
    else special = y > 0.0 ? inf() : 0.0;  // x = +inf
    const double out = regular ? v : special;
    return y == 0.0 ? 1.0 : ((y == 1.0 || x == 1.0) ? (x == 1.0 ? 1.0 : x) : out);
}

NF_HD double pow(double x, double y) { return pow(x, y, Tables{NFM_LOG_TBL, NFM_EXP_TBL}); }

// pow and log with their special inputs out of line, for loops that almost never meet one
// (the Johnson-Cook Newton solve). An ordinary input takes the main path alone, with none of
// the selects that pick the special values: x normal, finite and (for pow) not 1; y neither
// NaN, 0 nor 1; and y log x inside exp's range. Any other input calls pow / log / exp_dd
// themselves. For every input the result is pow's or log's own, bit for bit: on ordinary
// inputs those return exactly the main path.
NF_HD double pow_ordinary(double x, double y, const Tables& T) {
    if (x >= 0x1p-1022 && x < inf() && x != 1.0 && !is_nan(y) && y != 0.0 && y != 1.0) {
        double l_hi, l_lo;
        log_dd_impl<false>(x, l_hi, l_lo, T.log);
        double p_hi, p_lo;
        two_prod(y, l_hi, p_hi, p_lo);
        p_lo = p_lo + y * l_lo;
        double t_hi, t_lo;
        fast_two_sum(p_hi, p_lo, t_hi, t_lo);
        if (t_hi >= -745.2 && t_hi <= 709.782712893384) return exp_dd_main(t_hi, t_lo, T.exp);
        return exp_dd(t_hi, t_lo, T.exp);
    }
    return pow(x, y, T);
}

// pow_ordinary(x, y) together with log(x) (into lx), sharing the logarithm's double-double
// when the input is ordinary: lx is log's own bits either way.
NF_HD double pow_log_ordinary(double x, double y, const Tables& T, double& lx) {
    if (x >= 0x1p-1022 && x < inf() && x != 1.0 && !is_nan(y) && y != 0.0 && y != 1.0) {
        double l_hi, l_lo;
        log_dd_impl<false>(x, l_hi, l_lo, T.log);
        lx = l_hi + l_lo;
        double p_hi, p_lo;
        two_prod(y, l_hi, p_hi, p_lo);
        p_lo = p_lo + y * l_lo;
        double t_hi, t_lo;
        fast_two_sum(p_hi, p_lo, t_hi, t_lo);
        if (t_hi >= -745.2 && t_hi <= 709.782712893384) return exp_dd_main(t_hi, t_lo, T.exp);
        return exp_dd(t_hi, t_lo, T.exp);
    }
    lx = log(x, T);
    return pow(x, y, T);
}

NF_HD double log_ordinary(double x, const Tables& T) {
    if (x >= 0x1p-1022 && x < inf()) {
        double hi, lo;
        log_dd_impl<false>(x, hi, lo, T.log);
        return hi + lo;
    }
    return log(x, T);
}

// Two of the ordinary-input functions side by side, for a dependency chain that waits on
// both (the Johnson-Cook Newton loop: the strain term's pow and the rate term's log; the yield
// check: the strain and thermal terms' pows). When every input is ordinary, both main paths
// run as one straight-line block, so that the two chains interleave; exp's out-of-range
// results come from exp_dd's selects rather than a branch (inside exp's range exp_dd is
// exp_dd_main exactly). Otherwise each function runs on its own. Either way the results are
// pow_ordinary's and log_ordinary's own bits, hence pow's and log's.
NF_HD bool pow_input_ordinary(double x, double y) {
    return x >= 0x1p-1022 && x < inf() && x != 1.0 && !is_nan(y) && y != 0.0 && y != 1.0;
}

NF_HD double pow_from_log(double l_hi, double l_lo, double y, const Tables& T) {
    double p_hi, p_lo;
    two_prod(y, l_hi, p_hi, p_lo);
    p_lo = p_lo + y * l_lo;
    double t_hi, t_lo;
    fast_two_sum(p_hi, p_lo, t_hi, t_lo);
    return exp_dd(t_hi, t_lo, T.exp);
}

// pw = pow_ordinary(x, y), lg = log_ordinary(z)
NF_HD void pow_and_log_ordinary(double x, double y, double z, const Tables& T, double& pw, double& lg) {
    if (pow_input_ordinary(x, y) && z >= 0x1p-1022 && z < inf()) {
        double l_hi, l_lo, m_hi, m_lo;
        log_dd_impl<false>(x, l_hi, l_lo, T.log);
        log_dd_impl<false>(z, m_hi, m_lo, T.log);
        lg = m_hi + m_lo;
        pw = pow_from_log(l_hi, l_lo, y, T);
    } else {
        pw = pow_ordinary(x, y, T);
        lg = log_ordinary(z, T);
    }
}

// pow_and_log_ordinary for 0 < y < 1 (the Johnson-Cook strain exponent): with x ordinary,
// log x lies in [-708.4, 709.79] and y log x, rounded as a double-double, in the same interval,
// inside exp's range [-745.2, 709.782712893384], where exp_dd is exp_dd_main exactly; so the
// range selects are left out and the results are pow_and_log_ordinary's, bit for bit.
NF_HD void pow_and_log_ordinary_y01(double x, double y, double z, const Tables& T, double& pw, double& lg) {
    if (pow_input_ordinary(x, y) && z >= 0x1p-1022 && z < inf()) {
        double l_hi, l_lo, m_hi, m_lo;
        log_dd_impl<false>(x, l_hi, l_lo, T.log);
        log_dd_impl<false>(z, m_hi, m_lo, T.log);
        lg = m_hi + m_lo;
        double p_hi, p_lo;
        two_prod(y, l_hi, p_hi, p_lo);
        p_lo = p_lo + y * l_lo;
        double t_hi, t_lo;
        fast_two_sum(p_hi, p_lo, t_hi, t_lo);
        pw = exp_dd_main(t_hi, t_lo, T.exp);
    } else {
        pw = pow_ordinary(x, y, T);
        lg = log_ordinary(z, T);
    }
}

// p1 = pow_ordinary(x1, y1), p2 = pow_ordinary(x2, y2)
NF_HD void pow_pair_ordinary(double x1, double y1, double x2, double y2, const Tables& T, double& p1,
                             double& p2) {
    if (pow_input_ordinary(x1, y1) && pow_input_ordinary(x2, y2)) {
        double a_hi, a_lo, b_hi, b_lo;
        log_dd_impl<false>(x1, a_hi, a_lo, T.log);
        log_dd_impl<false>(x2, b_hi, b_lo, T.log);
        p1 = pow_from_log(a_hi, a_lo, y1, T);
        p2 = pow_from_log(b_hi, b_lo, y2, T);
    } else {
        p1 = pow_ordinary(x1, y1, T);
        p2 = pow_ordinary(x2, y2, T);
    }
}

#define NFM_PI 3.141592653589793
#define NFM_PI_2_HI 0x1.921fb54442d18p0    // pi/2 rounded
#define NFM_PI_2_LO 0x1.1a62633145c07p-54  // pi/2 - NFM_PI_2_HI
#define NFM_PI_4_HI 0x1.921fb54442d18p-1
#define NFM_PI_4_LO 0x1.1a62633145c07p-55

// atan(t) for 0 <= t <= 1.
NF_HD double atan_unit(double t) {
    double base_hi = 0.0, base_lo = 0.0;
    double u = t;
    if (t > 0.41421356237309503) {  // tan(pi/8)
        u = (t - 1.0) / (t + 1.0);
        base_hi = NFM_PI_4_HI;
        base_lo = NFM_PI_4_LO;
    }
    // Half-angle step: atan u = 2 atan(u / (1 + sqrt(1 + u^2))), |u'| <= tan(pi/16).
    const double w = u / (1.0 + __builtin_sqrt(1.0 + u * u));
    const double z = w * w;
    double p = -1.0 / 27.0;
    p = 1.0 / 25.0 + z * p;
    p = -1.0 / 23.0 + z * p;
    p = 1.0 / 21.0 + z * p;
    p = -1.0 / 19.0 + z * p;
    p = 1.0 / 17.0 + z * p;
    p = -1.0 / 15.0 + z * p;
    p = 1.0 / 13.0 + z * p;
    p = -1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = -1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = -1.0 / 3.0 + z * p;
    const double a = 2.0 * (w + (w * z) * p);
    return base_hi + (a + base_lo);
}

NF_HD double atan2(double y, double x) {
    if (is_nan(x) || is_nan(y)) return qnan();
    const double ay = y < 0.0 ? -y : y;
    const double ax = x < 0.0 ? -x : x;
    const bool neg_y = __builtin_signbit(y) != 0;
    if (ay == 0.0) {
        // atan2(+-0, x): +-0 for x > 0 or x = +0, +-pi for x < 0 or x = -0.
        const bool neg_x = __builtin_signbit(x) != 0;
        const double r = neg_x ? NFM_PI : 0.0;
        return neg_y ? -r : r;
    }
    if (ax == inf() || ay == inf()) {
        double r;
        if (ax == inf() && ay == inf()) r = x > 0.0 ? 0.25 * NFM_PI : 0.75 * NFM_PI;
        else if (ax == inf()) r = x > 0.0 ? 0.0 : NFM_PI;
        else r = 0.5 * NFM_PI;
        return neg_y ? -r : r;
    }
    double a;
    if (ay <= ax) {
        a = atan_unit(ay / ax);
    } else {
        a = NFM_PI_2_HI + (NFM_PI_2_LO - atan_unit(ax / ay));
    }
    if (x < 0.0) a = (NFM_PI_2_HI - a) + NFM_PI_2_HI + 2.0 * NFM_PI_2_LO;
    return neg_y ? -a : a;
}

// sin and cos of r with |r| <= pi/4 + a little.
NF_HD double sin_kernel(double r) {
    const double z = r * r;
    double p = -1.0 / 121645100408832000.0;  // -1/19!
    p = 1.0 / 355687428096000.0 + z * p;   // 1/17!
    p = -1.0 / 1307674368000.0 + z * p;    // -1/15!
    p = 1.0 / 6227020800.0 + z * p;        // 1/13!
    p = -1.0 / 39916800.0 + z * p;         // -1/11!
    p = 1.0 / 362880.0 + z * p;            // 1/9!
    p = -1.0 / 5040.0 + z * p;             // -1/7!
    p = 1.0 / 120.0 + z * p;               // 1/5!
    p = -1.0 / 6.0 + z * p;                // -1/3!
    return r + (r * z) * p;
}

NF_HD double cos_kernel(double r) {
    const double z = r * r;
    double p = -1.0 / 6402373705728000.0;  // -1/18!
    p = 1.0 / 20922789888000.0 + z * p;    // 1/16!
    p = -1.0 / 87178291200.0 + z * p;      // -1/14!
    p = 1.0 / 479001600.0 + z * p;         // 1/12!
    p = -1.0 / 3628800.0 + z * p;          // -1/10!
    p = 1.0 / 40320.0 + z * p;             // 1/8!
    p = -1.0 / 720.0 + z * p;              // -1/6!
    p = 1.0 / 24.0 + z * p;                // 1/4!
    // cos r = 1 - z/2 + z^2 p, summed so the leading 1 - z/2 keeps its low bits.
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * z) * p);
}

// Quadrant reduction for |x| up to a few thousand. k (pi/2)_hi = p + e exactly (two_prod),
// and x - p is exact (Sterbenz: p lies within a factor of two of x whenever k != 0), so
// r = ((x - p) - e) - k (pi/2)_lo carries the reduced argument to about 2^-100 absolute. For
// |k| <= 1 the product is exact (e = 0) and r is (x - k hi) - k lo.
NF_HD void reduce_pi_2(double x, double& r, int& q) {
    const double shifter = 0x1.8p52;
    const double kd = (x * 0.6366197723675814 + shifter) - shifter;  // 2/pi
    q = (int)kd;
    double p, e;
    two_prod(kd, NFM_PI_2_HI, p, e);
    r = ((x - p) - e) - kd * NFM_PI_2_LO;
}

NF_HD double sin(double x) {
    if (is_nan(x) || x == inf() || x == -inf()) return qnan();
    double r;
    int q;
    reduce_pi_2(x, r, q);
    switch (q & 3) {
        case 0: return sin_kernel(r);
        case 1: return cos_kernel(r);
        case 2: return -sin_kernel(r);
        default: return -cos_kernel(r);
    }
}

NF_HD double cos(double x) {
    if (is_nan(x) || x == inf() || x == -inf()) return qnan();
    double r;
    int q;
    reduce_pi_2(x, r, q);
    switch (q & 3) {
        case 0: return cos_kernel(r);
        case 1: return -sin_kernel(r);
        case 2: return -cos_kernel(r);
        default: return sin_kernel(r);
    }
}

NF_HD double hypot(double a, double b) {
    if (a < 0.0) a = -a;
    if (b < 0.0) b = -b;
    if (a == inf() || b == inf()) return inf();
    if (is_nan(a) || is_nan(b)) return qnan();
    double big = a > b ? a : b;
    double small = a > b ? b : a;
    if (big == 0.0) return 0.0;
    const double ratio = small / big;
    return big * __builtin_sqrt(1.0 + ratio * ratio);
}

}  // namespace nfm

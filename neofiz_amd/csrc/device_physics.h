// Device-side constitutive and element arithmetic for the gfx950 kernels.
//
// Each function restates one method of the Java reference operation for operation, in the
// same evaluation order, so that the kernels reproduce the reference's arithmetic exactly
// (compiled with -ffp-contract=off; IEEE division and square root). State is passed in
// registers rather than through the reference's strided arrays.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nf_math.h"

namespace nfd {

#define NFD_MAX_VALUE 1.7976931348623157e308
#define NFD_PI 3.141592653589793

// sqrt(2/3) and 2/3 as Java computes them (solver/J2.java:82-83).
#define NFD_SQRT_2_3 0x1.a20bd700c2c3ep-1
#define NFD_TWO_THIRDS (2.0 / 3.0)

// Math.min / Math.max with Java's NaN and signed-zero rules.
__device__ __forceinline__ double jmin(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return __builtin_signbit(a) ? a : b;
    return a <= b ? a : b;
}
__device__ __forceinline__ double jmax(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return __builtin_signbit(a) ? b : a;
    return a >= b ? a : b;
}
__device__ __forceinline__ double jabs(double a) { return __builtin_fabs(a); }

// Per-material constants, unrolled once on the host (solver/ExplicitSolver.java:289-299,
// solver/J2.java:105-113, core/JohnsonCook.java:68-71).
struct Mat {
    double lambda, mu, bulk, density;
    double sy0, hIso, hKin, thermalCoefficient;
    double a, b, n, c, m, eps0, troom, tmelt, cp, beta;
    double youngs;
    int32_t hasJC;
    int32_t pad;
};

// core/JohnsonCook.java:138-197
// T: the log/exp tables, read from LDS in the kernels (nfm::Tables).
__device__ __forceinline__ double jc_strainTerm(const Mat& j, double e, const nfm::Tables& T) {
    return e <= 0.0 ? j.a : j.a + j.b * nfm::pow_ordinary(e, j.n, T);
}
__device__ __forceinline__ double jc_rateTerm(const Mat& j, double rate, const nfm::Tables& T) {
    if (rate <= j.eps0) return 1.0;
    // x / 1.0 is x exactly, so the division is skipped at the usual reference rate of 1/s.
    return 1.0 + j.c * nfm::log_ordinary(j.eps0 == 1.0 ? rate : rate / j.eps0, T);
}
__device__ __forceinline__ double jc_thermalTerm(const Mat& j, double t, const nfm::Tables& T) {
    const double homologous = (t - j.troom) / (j.tmelt - j.troom);
    if (homologous <= 0.0) return 1.0;
    if (homologous >= 1.0) return 0.0;
    return 1.0 - nfm::pow_ordinary(homologous, j.m, T);
}
// strainTerm(e) and thermalTerm(t) together, their two pows side by side
// (nfm::pow_pair_ordinary). A term that takes no pow gives its pow a stand-in input of 0.5,
// whose result is not used; each result is the single-term function's own. A wave none of
// whose points takes a pow (points that never yielded: no plastic strain, no plastic work,
// so no heating) skips the pows altogether.
__device__ __forceinline__ void jc_terms(const Mat& j, double e, double t, const nfm::Tables& T,
                                         double& strain, double& thermal) {
    const double homologous = (t - j.troom) / (j.tmelt - j.troom);
    const bool strainPow = !(e <= 0.0), thermalPow = !(homologous <= 0.0) && !(homologous >= 1.0);
    double ps = 0.0, pt = 0.0;
    if (__any(strainPow || thermalPow))
        nfm::pow_pair_ordinary(strainPow ? e : 0.5, j.n, thermalPow ? homologous : 0.5, j.m, T, ps, pt);
    strain = e <= 0.0 ? j.a : j.a + j.b * ps;
    thermal = homologous <= 0.0 ? 1.0 : (homologous >= 1.0 ? 0.0 : 1.0 - pt);
}

// ---- The solve's leading bisection steps, proven instead of evaluated.
//
// The solve starts at the top of its bracket, dGamma = hi = norm / stiffness, which for a
// point that barely yields lies far above the root. Its first iterations are then all alike:
// g <= 0 (so hi = dGamma), |g| > tolerance, and a Newton step that leaves the bracket through
// zero (so dGamma = 0.5 (lo + hi) with lo still +0). The state after such an iteration,
// (lo, hi, dGamma) = (0, d, 0.5 (0 + d)), depends on its decisions alone, not on the values
// of g and the slope. An iteration whose decisions are proven can therefore be skipped, and
// every later value is bit-identical to the reference's.
//
// Proof for the iterate d: g < -tolerance (hence g <= 0 and |g| > tolerance) and
// |g| >= d |slope| (so that g / slope >= d in real arithmetic, its rounded quotient Q >= d
// because rounding is monotone and d is a double, and next = d - Q <= 0 is outside the
// bracket (0, d) -- or the slope is not negative and next >= d or NaN, outside it as well).
// g and d |slope| are estimated with single-precision hardware log2, exp2 and reciprocal for
// the strain term's power, the rate term's logarithm and 1/eps (relative errors near 1e-6 at
// most for the arguments allowed: eps > 1e-12, rate / eps0 < 1e15; the device self-check
// measures them), and the exact double-precision evaluation is bounded by the
// estimate plus FF_MARGIN = 2^-14 times the sum of the magnitudes of the terms. All terms are
// nonnegative (A > 0, B >= 0, n in (0, 1], C >= 0 by the reference's JohnsonCook constructor,
// core/JohnsonCook.java:74-77; thermal factor in [0, 1]; stiffness > 0), so relative errors
// add up. (strain - A), which the exact slope forms by a rounded subtraction, is bounded by
// B pw plus 2^-50 of the strain term. A non-finite value anywhere fails the comparisons, and
// the iteration is then evaluated exactly.
#define NFD_FF_MARGIN 0x1p-14

struct FFSolve {
    double plasticStrain, norm, stiffness, thermal, tolerance;
    double invRateScale;  // 1 / (dt eps0), or 0 when dt <= 0 (rate 0, on its clamp)
    double a, b, n, c;
    float nf;
};

__device__ __forceinline__ bool ff_bisects(double d, const FFSolve& f) {
    const double dEps = NFD_SQRT_2_3 * d;
    const double eps = f.plasticStrain + dEps;  // the iteration's own eps, bit for bit
    const float ef = (float)eps;
    const double pw = (double)__builtin_amdgcn_exp2f(f.nf * __builtin_amdgcn_logf(ef));
    const double bpw = f.b * pw;
    const double strain = f.a + bpw;
    const double arg = dEps * f.invRateScale;  // rate / eps0
    const double lr = arg > 1.0 ? (double)__builtin_amdgcn_logf((float)arg) * 0x1.62e42fefa39efp-1 : 0.0;
    const double rateFactor = 1.0 + f.c * lr;
    const double flow = NFD_SQRT_2_3 * strain * rateFactor * f.thermal;
    const double sd = f.stiffness * d;
    const double gUp = f.norm - sd - flow + NFD_FF_MARGIN * (f.norm + sd + flow);
    const double dOverEps = d * (double)__builtin_amdgcn_rcpf(ef);
    const double dSlope = sd + (NFD_TWO_THIRDS * f.n * (bpw + 0x1p-50 * strain) * dOverEps * rateFactor * f.thermal +
                                NFD_SQRT_2_3 * strain * f.c * f.thermal) * (1.0 + NFD_FF_MARGIN);
    return eps > 1e-12 && arg < 1e15 && gUp < -f.tolerance && -gUp >= dSlope * (1.0 + 0x1p-40);
}

// The number of leading iterations proven to bisect toward zero, at most 25, with the state
// they leave: hi and dGamma updated as the reference's loop would, lo still +0. Candidates are
// tested two at a time (the second is the first's bisection), so that their chains overlap.
__device__ __forceinline__ int ff_leading_bisections(const FFSolve& f, double lo, double& hi, double& dGamma) {
    int first = 0;
    while (first < 25) {
        const double d0 = dGamma;
        const double d1 = 0.5 * (lo + d0);
        const bool p0 = ff_bisects(d0, f), p1 = ff_bisects(d1, f);
        if (!p0) break;
        hi = d0;
        dGamma = d1;
        first++;
        if (!p1 || first == 25) break;
        hi = d1;
        dGamma = 0.5 * (lo + d1);
        first++;
    }
    return first;
}

// solver/J2.java:332-375, the bracketed Newton solve for the Johnson-Cook multiplier.
// `thermal` is thermalTerm(temperature), which the caller has already evaluated for the
// yield check on the same temperature (J2.java:221 and :334 compute the same value twice).
// UNIT_RATE: the reference strain rate is 1/s, where rate / eps0 is rate itself; the loop is
// instantiated for it so that no iteration evaluates a division whose result is discarded.
// REC (the adjoint's recording pass only): *status becomes 2 when the solve stops on its
// iteration cap rather than its tolerance. The plain instantiation has no such code.
// `iterations` (optional) is increased by the solve's cost: the iterations it evaluated, plus
// a quarter of one for each leading bisection it proved (ff_leading_bisections); the element
// pass schedules its waves by it.
// FF = false evaluates every iteration (the device self-check compares the two forms).
// Y01: the strain exponent n lies in (0, 1), so the loop's pow needs no range selects
// (nfm::pow_and_log_ordinary_y01, the same bits).
template <bool UNIT_RATE, bool REC = false, bool FF = true, bool Y01 = false>
__device__ __forceinline__ double j2_solve_loop(double norm, double stiffness, double plasticStrain,
                                                double thermal, double dt, const Mat& j,
                                                const nfm::Tables& T, int* status = nullptr,
                                                int* iterations = nullptr, int* proven = nullptr) {
    const double tolerance = 1.0e-10 * j.a;
    double lo = 0.0;
    double hi = norm / stiffness;
    double dGamma = hi;
    FFSolve f;
    f.plasticStrain = plasticStrain;
    f.norm = norm;
    f.stiffness = stiffness;
    f.thermal = thermal;
    f.tolerance = tolerance;
    f.invRateScale = dt > 0.0 ? 1.0 / (dt * j.eps0) : 0.0;
    f.a = j.a;
    f.b = j.b;
    f.n = j.n;
    f.c = j.c;
    f.nf = (float)j.n;
    const int first = FF ? ff_leading_bisections(f, lo, hi, dGamma) : 0;
    if (proven) *proven = first;
    // The scheduling cost: exact iterations, plus a quarter of one per proven bisection.
    if (iterations) *iterations += (first + 2) / 4 - first;
    for (int iteration = first; iteration < 25; iteration++) {
        // The body is written with selects rather than branches (the values are the
        // reference's): the strain term's pow and the rate term's log are independent, and
        // without branches between them the compiler interleaves the two chains.
        const double dEps = NFD_SQRT_2_3 * dGamma;
        const double eps = plasticStrain + dEps;
        const double rate = dt > 0.0 ? dEps / dt : 0.0;
        double pw, lg;
        if constexpr (Y01) nfm::pow_and_log_ordinary_y01(eps, j.n, UNIT_RATE ? rate : rate / j.eps0, T, pw, lg);
        else nfm::pow_and_log_ordinary(eps, j.n, UNIT_RATE ? rate : rate / j.eps0, T, pw, lg);
        const double strain = eps <= 0.0 ? j.a : j.a + j.b * pw;  // strainTerm
        const bool clamped = rate <= j.eps0;
        const double rateFactor = clamped ? 1.0 : 1.0 + j.c * lg;  // rateTerm
        const double g = norm - stiffness * dGamma - NFD_SQRT_2_3 * strain * rateFactor * thermal;
        lo = g > 0.0 ? dGamma : lo;
        hi = g > 0.0 ? hi : dGamma;
        if (jabs(g) <= tolerance) {
            if (iterations) *iterations += iteration + 1;
            return dGamma;
        }
        const double slope0 = -stiffness - NFD_TWO_THIRDS * (j.n * (strain - j.a) / eps) * rateFactor * thermal;
        const double slope = clamped ? slope0 : slope0 - NFD_SQRT_2_3 * strain * j.c * thermal / dGamma;
        const double next = dGamma - g / slope;
        dGamma = (next > lo && next < hi) ? next : 0.5 * (lo + hi);
    }
    if (REC) *status = 2;
    if (iterations) *iterations += 25;
    return dGamma;
}

template <bool REC = false>
__device__ __forceinline__ double j2_solve(double norm, double stiffness, double plasticStrain,
                                           double thermal, double dt, const Mat& j,
                                           const nfm::Tables& T, int* status = nullptr,
                                           int* iterations = nullptr) {
    if (j.eps0 != 1.0)
        return j2_solve_loop<false, REC>(norm, stiffness, plasticStrain, thermal, dt, j, T, status, iterations);
    return j.n > 0.0 && j.n < 1.0
               ? j2_solve_loop<true, REC, true, true>(norm, stiffness, plasticStrain, thermal, dt, j, T, status, iterations)
               : j2_solve_loop<true, REC>(norm, stiffness, plasticStrain, thermal, dt, j, T, status, iterations);
}

// Damage state of one Gauss point (solver/Damage.java:156-192).
struct DamagePoint {
    double strength, softening, width;  // mutable: written at onset
    double failureStrain, fractureEnergy, youngs;
    double area, uR, uZ, vR, vZ;
};

// solver/Damage.java:306-336
__device__ __forceinline__ void damage_begin(DamagePoint& d, double flowStress, double nRR, double nZZ,
                             double nRZ) {
    const double diff = nRR - nZZ;
    const double cross = 2.0 * nRZ;
    const double anisotropy = nfm::hypot(diff, cross);
    double h;
    if (anisotropy < 1.0e-9) {
        h = __builtin_sqrt(d.area);
    } else {
        const double phi = 0.5 * nfm::atan2(cross, diff);
        const double tR = -nfm::sin(phi), tZ = nfm::cos(phi);
        const double line = jabs(d.uR * tR + d.uZ * tZ) + jabs(d.vR * tR + d.vZ * tZ);
        h = d.area / line;
    }
    d.width = h;
    d.strength = flowStress;
    const double slope = -flowStress * flowStress * h / (2.0 * d.fractureEnergy);
    const double floor = -d.youngs;
    d.softening = slope < floor ? floor : slope;
}

// solver/Damage.java:342-347
__device__ __forceinline__ double damage_at(double strength, double softening,
                                            double failureStrain, double plasticStrain) {
    const double sf = strength;
    if (sf <= 0.0) return 0.0;
    const double d = -softening * (plasticStrain - failureStrain) / sf;
    return d <= 0.0 ? 0.0 : (d >= 1.0 ? 1.0 : d);
}

// Copies the log/exp tables into the workgroup's LDS (call from every thread, before use).
__device__ __forceinline__ nfm::Tables stage_tables(double (*sLog)[3], double (*sExp)[2]) {
    for (int i = threadIdx.x; i < 97 * 3; i += blockDim.x) (&sLog[0][0])[i] = (&NFM_LOG_TBL[0][0])[i];
    for (int i = threadIdx.x; i < 64 * 2; i += blockDim.x) (&sExp[0][0])[i] = (&NFM_EXP_TBL[0][0])[i];
    __syncthreads();
    return nfm::Tables{sLog, sExp};
}

// Point state carried through the update.
struct PointState {
    double s0, s1, s2, s3;  // stress rr, zz, tt, rz
    double b0, b1, b2, b3;  // back stress
    double epsP, work;
};

// solver/J2.java:163-296. `dmg` is null when the solver has no damage model. KIN is false
// when no material has kinematic hardening: the back stress is then +0 forever (its
// co-rotation and its update only ever add signed zeros to +0), so it is neither loaded nor
// stored, and subtracting it or adding alpha : n changes no bit of any result.
// REC (the adjoint's recording pass only): *status receives 0 for an elastic step, 1 for a
// plastic one whose multiplier met the solve's tolerance, 2 where the solve hit its cap.
// `iterations` (optional) accumulates the Johnson-Cook solve's iteration count.
template <bool DAMAGE, bool KIN, bool REC = false>
__device__ __forceinline__ double j2_update(PointState& q, double dRR, double dZZ, double dTT,
                                            double dGamRZ, const Mat& fl, double strainDt,
                                            DamagePoint* dmg, const nfm::Tables& T,
                                            int* status = nullptr, int* iterations = nullptr) {
    const double lambda = fl.lambda, mu = fl.mu;
    const double lt = lambda * (dRR + dZZ + dTT);
    const double tRR = q.s0 + lt + 2.0 * mu * dRR;
    const double tZZ = q.s1 + lt + 2.0 * mu * dZZ;
    const double tTT = q.s2 + lt + 2.0 * mu * dTT;
    const double tRZ = q.s3 + mu * dGamRZ;

    const double p = (tRR + tZZ + tTT) / 3.0;
    const double xRR = KIN ? tRR - p - q.b0 : tRR - p;
    const double xZZ = KIN ? tZZ - p - q.b1 : tZZ - p;
    const double xTT = KIN ? tTT - p - q.b2 : tTT - p;
    const double xRZ = KIN ? tRZ - q.b3 : tRZ;
    const double norm = __builtin_sqrt(xRR * xRR + xZZ * xZZ + xTT * xTT + 2.0 * xRZ * xRZ);

    const bool jc = fl.hasJC != 0;
    const double temperature = jc ? fl.troom + fl.thermalCoefficient * q.work : 0.0;

    const bool softening = DAMAGE && dmg->strength > 0.0;
    double hardening = 0.0;
    double flowStress;
    double thermal = 1.0;
    if (softening) {
        flowStress = dmg->strength + dmg->softening * (q.epsP - dmg->failureStrain);
        if (flowStress > 0.0) {
            hardening = dmg->softening;
        } else {
            flowStress = 0.0;
        }
    } else if (!jc) {
        flowStress = fl.sy0 + fl.hIso * q.epsP;
        hardening = fl.hIso;
    } else {
        double strain;
        jc_terms(fl, q.epsP, temperature, T, strain, thermal);
        flowStress = strain * thermal;  // quasiStaticFlowStress
    }
    const double radius = NFD_SQRT_2_3 * flowStress;

    if (!(norm > radius)) {
        q.s0 = tRR;
        q.s1 = tZZ;
        q.s2 = tTT;
        q.s3 = tRZ;
        if (REC) *status = 0;
        return 0.0;
    }
    if (REC) *status = 1;

    const double stiffness = 2.0 * mu + NFD_TWO_THIRDS * fl.hKin;
    double dGamma;
    if (!jc || softening) {
        dGamma = (norm - radius) / (stiffness + NFD_TWO_THIRDS * hardening);
        if (softening && flowStress + hardening * NFD_SQRT_2_3 * dGamma < 0.0) {
            dGamma = norm / stiffness;
        }
    } else {
        dGamma = j2_solve<REC>(norm, stiffness, q.epsP, thermal, strainDt, fl, T, status, iterations);
    }
    const double inv = 1.0 / norm;
    const double nRR = xRR * inv, nZZ = xZZ * inv, nTT = xTT * inv, nRZ = xRZ * inv;

    const double pull = 2.0 * mu * dGamma;
    q.s0 = tRR - pull * nRR;
    q.s1 = tZZ - pull * nZZ;
    q.s2 = tTT - pull * nTT;
    q.s3 = tRZ - pull * nRZ;

    if (KIN) {
        const double alphaDotN = q.b0 * nRR + q.b1 * nZZ + q.b2 * nTT + 2.0 * q.b3 * nRZ;
        q.work += dGamma * (norm - pull + alphaDotN);
        const double shift = NFD_TWO_THIRDS * fl.hKin * dGamma;
        q.b0 += shift * nRR;
        q.b1 += shift * nZZ;
        q.b2 += shift * nTT;
        q.b3 += shift * nRZ;
    } else {
        q.work += dGamma * (norm - pull);
    }

    q.epsP += NFD_SQRT_2_3 * dGamma;

    if (DAMAGE && !softening && q.epsP >= dmg->failureStrain) {
        // the onset flow stress: the same temperature, so the same thermal factor
        damage_begin(*dmg, !jc ? fl.sy0 + fl.hIso * q.epsP : jc_strainTerm(fl, q.epsP, T) * thermal,
                     nRR, nZZ, nRZ);
    }
    return dGamma;
}

// solver/Corotational.java:59-97, applied to one tensor held in registers.
__device__ __forceinline__ void corot_rotate(double& rr, double& zz, double& rz, double cos,
                                             double sin) {
    const double cc = cos * cos, ss = sin * sin, cs = cos * sin;
    const double r0 = rr, z0 = zz, x0 = rz;
    rr = cc * r0 - 2.0 * cs * x0 + ss * z0;
    zz = ss * r0 + 2.0 * cs * x0 + cc * z0;
    rz = cs * (r0 - z0) + (cc - ss) * x0;
}
__device__ __forceinline__ double corot_cos(double spinDt) {
    const double h = 0.5 * spinDt;
    return (1.0 - h * h) / (1.0 + h * h);
}
__device__ __forceinline__ double corot_sin(double spinDt) {
    const double h = 0.5 * spinDt;
    return 2.0 * h / (1.0 + h * h);
}

__device__ __forceinline__ double sq(double a, double b) { return a * a + b * b; }

// solver/ExplicitSolver.java:1191-1197
__device__ __forceinline__ double shortestEdge(double running, double r0, double r1, double r2,
                                               double r3, double z0, double z1, double z2,
                                               double z3) {
    return jmin(running, jmin(jmin(sq(r1 - r0, z1 - z0), sq(r2 - r1, z2 - z1)),
                              jmin(sq(r3 - r2, z3 - z2), sq(r0 - r3, z0 - z3))));
}

}  // namespace nfd

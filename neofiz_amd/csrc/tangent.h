// Parameter sensitivities of the step loop: the exact derivative of the discrete algorithm
// with respect to (friction, A, B, n, C, m), carried forward alongside the state.
//
// Included by neofiz_hip.hip after the single-solver kernels.
//
// Every quantity of the reduced-integration Johnson-Cook step that depends on the
// parameters is a Dual: its value and its six directional derivatives. The value part of
// every Dual operation is the same IEEE operation, in the same order, as the plain kernels
// perform, so a solver with sensitivities on follows the plain trajectory bit for bit; the
// derivative parts follow the chain rule through the same operations. Two places are not
// differentiated operation by operation:
//   - the Johnson-Cook Newton solve (solver/J2.java:332-375) is run in plain doubles, and
//     the derivative of its root comes from the implicit function theorem at the root:
//     dGamma' = -g'|dGamma fixed / (dg/dGamma), with dg/dGamma the solve's own slope
//     expression. The solve stops at |g| <= 1e-10 A, so this is the derivative of the
//     exact root, which the truncated iterate matches to that tolerance;
//   - min and max (the CFL step's shortest edge, Math.min/max in the flow rules, the
//     objectives' extreme nodes) take the derivative of the argument they select, the
//     first one on ties, as the plain selection does.
// Discrete events (a node landing on the wall, a stick/slip switch, the yield test) are
// fixed by the plain trajectory; the derivative is that of the branch taken.
#pragma once

constexpr int NP = 6;  // friction, A, B, n, C, m
enum { SP_FRICTION = 0, SP_A, SP_B, SP_N, SP_C, SP_M };
// Each thread carries NL of the NP directions: two threads share an element (or a node),
// one with directions 0-2, one with 3-5. Both compute the value parts, which are identical;
// only the first stores them. The derivative of one direction never reads another's, so the
// split changes no bit of either.
constexpr int NL = 3;

namespace nft {

struct D {
    double v;
    double t[NL];
};

__device__ __forceinline__ D cst(double v) {
    D r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = 0.0;
    return r;
}
// v seeded in direction k, as seen by a thread that carries directions dir0 .. dir0 + NL - 1.
__device__ __forceinline__ D seed(double v, int k, int dir0) {
    D r;
    r.v = v;
#pragma unroll
    for (int m = 0; m < NL; m++) r.t[m] = m == k - dir0 ? 1.0 : 0.0;
    return r;
}

__device__ __forceinline__ D operator+(const D& a, const D& b) {
    D r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = a.t[k] + b.t[k];
    return r;
}
__device__ __forceinline__ D operator-(const D& a, const D& b) {
    D r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = a.t[k] - b.t[k];
    return r;
}
__device__ __forceinline__ D operator-(const D& a) {
    D r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = -a.t[k];
    return r;
}
__device__ __forceinline__ D operator*(const D& a, const D& b) {
    D r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = a.t[k] * b.v + a.v * b.t[k];
    return r;
}
__device__ __forceinline__ D operator/(const D& a, const D& b) {
    D r;
    r.v = a.v / b.v;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = (a.t[k] - r.v * b.t[k]) / b.v;
    return r;
}
__device__ __forceinline__ D operator+(const D& a, double b) {
    D r = a;
    r.v = a.v + b;
    return r;
}
__device__ __forceinline__ D operator+(double a, const D& b) {
    D r = b;
    r.v = a + b.v;
    return r;
}
__device__ __forceinline__ D operator-(const D& a, double b) {
    D r = a;
    r.v = a.v - b;
    return r;
}
__device__ __forceinline__ D operator-(double a, const D& b) {
    D r;
    r.v = a - b.v;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = -b.t[k];
    return r;
}
__device__ __forceinline__ D operator*(const D& a, double b) {
    D r;
    r.v = a.v * b;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = a.t[k] * b;
    return r;
}
__device__ __forceinline__ D operator*(double a, const D& b) {
    D r;
    r.v = a * b.v;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = a * b.t[k];
    return r;
}
__device__ __forceinline__ D operator/(const D& a, double b) {
    D r;
    r.v = a.v / b;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = a.t[k] / b;
    return r;
}
__device__ __forceinline__ D operator/(double a, const D& b) {
    D r;
    r.v = a / b.v;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = -r.v * b.t[k] / b.v;
    return r;
}
__device__ __forceinline__ D sqrt(const D& a) {
    D r;
    r.v = __builtin_sqrt(a.v);
    const double s = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = a.t[k] * s;
    return r;
}
// log and pow: the value from the shared table method, the derivative from calculus.
__device__ __forceinline__ D log(const D& a, const nfm::Tables& T) {
    D r;
    r.v = nfm::log_ordinary(a.v, T);
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = a.t[k] / a.v;
    return r;
}
__device__ __forceinline__ D pow(const D& x, const D& y, const nfm::Tables& T) {
    D r;
    double logx;
    r.v = nfm::pow_log_ordinary(x.v, y.v, T, logx);
    const double lx = x.v > 0.0 ? logx : 0.0;
    const double dx = x.v != 0.0 ? r.v * y.v / x.v : 0.0;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = dx * x.t[k] + r.v * lx * y.t[k];
    return r;
}
// Math.min / Math.max: the selected argument, value and derivative.
__device__ __forceinline__ D jmin(const D& a, const D& b) {
    const double m = nfd::jmin(a.v, b.v);
    return nfm::bits_of(m) == nfm::bits_of(a.v) ? a : b;
}
__device__ __forceinline__ D jmax(const D& a, const D& b) {
    const double m = nfd::jmax(a.v, b.v);
    return nfm::bits_of(m) == nfm::bits_of(a.v) ? a : b;
}

// The differentiated material: JC constants as Duals seeded in their directions.
struct DMat {
    double lambda, mu, eps0, troom, tmelt, thermalCoefficient;
    D a, b, n, c, m;
};

__device__ __forceinline__ DMat dmat(const Mat& f, int dir0) {
    DMat j;
    j.lambda = f.lambda;
    j.mu = f.mu;
    j.eps0 = f.eps0;
    j.troom = f.troom;
    j.tmelt = f.tmelt;
    j.thermalCoefficient = f.thermalCoefficient;
    j.a = seed(f.a, SP_A, dir0);
    j.b = seed(f.b, SP_B, dir0);
    j.n = seed(f.n, SP_N, dir0);
    j.c = seed(f.c, SP_C, dir0);
    j.m = seed(f.m, SP_M, dir0);
    return j;
}

// core/JohnsonCook.java:138-197 over Duals (values as nfd::jc_*).
__device__ __forceinline__ D strainTerm(const DMat& j, const D& e, const nfm::Tables& T) {
    return e.v <= 0.0 ? j.a : j.a + j.b * pow(e, j.n, T);
}
__device__ __forceinline__ D rateTerm(const DMat& j, const D& rate, const nfm::Tables& T) {
    if (rate.v <= j.eps0) return cst(1.0);
    return 1.0 + j.c * log(j.eps0 == 1.0 ? rate : rate / j.eps0, T);
}
__device__ __forceinline__ D thermalTerm(const DMat& j, const D& t, const nfm::Tables& T) {
    const D homologous = (t - j.troom) / (j.tmelt - j.troom);
    if (homologous.v <= 0.0) return cst(1.0);
    if (homologous.v >= 1.0) return cst(0.0);
    return 1.0 - pow(homologous, j.m, T);
}

struct DPoint {
    D s0, s1, s2, s3, epsP, work;
};

// J2.solve (solver/J2.java:332-375) over Duals, iteration by iteration: the derivative of
// the iterate the solve actually returns. Used where the returned multiplier is not a root
// to the solve's tolerance (the iteration cap was reached), so that the implicit function
// theorem does not apply. Values: nfd::j2_solve's operations exactly.
__device__ __forceinline__ D j2_solve(const D& norm, double stiffness, const D& plasticStrain,
                                      const D& thermal, const D& dt, const DMat& j,
                                      const nfm::Tables& T) {
    const double tolerance = 1.0e-10 * j.a.v;
    D lo = cst(0.0);
    D hi = norm / stiffness;
    D dGamma = hi;
    for (int iteration = 0; iteration < 25; iteration++) {
        const D dEps = NFD_SQRT_2_3 * dGamma;
        const D eps = plasticStrain + dEps;
        const D rate = dt.v > 0.0 ? dEps / dt : cst(0.0);
        const D strain = eps.v <= 0.0 ? j.a : j.a + j.b * pow(eps, j.n, T);
        const bool clamped = rate.v <= j.eps0;
        const D rateFactor = clamped ? cst(1.0) : 1.0 + j.c * log(j.eps0 == 1.0 ? rate : rate / j.eps0, T);
        const D g = norm - stiffness * dGamma - NFD_SQRT_2_3 * strain * rateFactor * thermal;
        if (g.v > 0.0) lo = dGamma;
        else hi = dGamma;
        if (nfd::jabs(g.v) <= tolerance) return dGamma;
        const D slope0 = -stiffness - NFD_TWO_THIRDS * (j.n * (strain - j.a) / eps) * rateFactor * thermal;
        const D slope = clamped ? slope0 : slope0 - NFD_SQRT_2_3 * strain * j.c * thermal / dGamma;
        const D next = dGamma - g / slope;
        dGamma = (next.v > lo.v && next.v < hi.v) ? next : 0.5 * (lo + hi);
    }
    return dGamma;
}

// J2.update (solver/J2.java:163-296) for a Johnson-Cook material without kinematic
// hardening or damage, over Duals; the multiplier's derivative by the implicit function
// theorem at the Newton root.
__device__ __forceinline__ void j2_update(DPoint& q, const D& dRR, const D& dZZ, const D& dTT,
                                          const D& dGamRZ, const DMat& j, const Mat& fl,
                                          const D& strainDt, const nfm::Tables& T, int* unconverged) {
    const double lambda = j.lambda, mu = j.mu;
    const D lt = lambda * (dRR + dZZ + dTT);
    const D tRR = q.s0 + lt + 2.0 * mu * dRR;
    const D tZZ = q.s1 + lt + 2.0 * mu * dZZ;
    const D tTT = q.s2 + lt + 2.0 * mu * dTT;
    const D tRZ = q.s3 + mu * dGamRZ;
    const D p = (tRR + tZZ + tTT) / 3.0;
    const D xRR = tRR - p, xZZ = tZZ - p, xTT = tTT - p, xRZ = tRZ;
    const D norm = sqrt(xRR * xRR + xZZ * xZZ + xTT * xTT + 2.0 * xRZ * xRZ);
    const D temperature = j.troom + j.thermalCoefficient * q.work;
    const D thermal = thermalTerm(j, temperature, T);
    const D flowStress = strainTerm(j, q.epsP, T) * thermal;
    const D radius = NFD_SQRT_2_3 * flowStress;
    if (!(norm.v > radius.v)) {
        q.s0 = tRR;
        q.s1 = tZZ;
        q.s2 = tTT;
        q.s3 = tRZ;
        return;
    }
    const double stiffness = 2.0 * mu + NFD_TWO_THIRDS * fl.hKin;
    const double g0 = nfd::j2_solve(norm.v, stiffness, q.epsP.v, thermal.v, strainDt.v, fl, T);
    // g(dGamma) = norm - stiffness dGamma - sqrt(2/3) S(eps) R(rate) Theta, differentiated
    // with dGamma held: its derivative over the slope is -dGamma'.
    const double dEpsV = NFD_SQRT_2_3 * g0;
    const D eps = q.epsP + dEpsV;
    const D rate = strainDt.v > 0.0 ? dEpsV / strainDt : cst(0.0);
    const D strain = strainTerm(j, eps, T);
    const D rateFactor = rateTerm(j, rate, T);
    const D g = norm - stiffness * g0 - NFD_SQRT_2_3 * strain * rateFactor * thermal;
    const bool clamped = rate.v <= j.eps0;
    const double slope0 = -stiffness - NFD_TWO_THIRDS * (j.n.v * (strain.v - j.a.v) / eps.v) * rateFactor.v * thermal.v;
    const double slope = clamped ? slope0 : slope0 - NFD_SQRT_2_3 * strain.v * j.c.v * thermal.v / g0;
    D dGamma;
    if (nfd::jabs(g.v) <= 1.0e-10 * j.a.v) {
        dGamma.v = g0;
#pragma unroll
        for (int k = 0; k < NL; k++) dGamma.t[k] = -g.t[k] / slope;
    } else {
        dGamma = j2_solve(norm, stiffness, q.epsP, thermal, strainDt, j, T);
        if (unconverged) atomicAdd(unconverged, 1);
    }

    const D inv = 1.0 / norm;
    const D nRR = xRR * inv, nZZ = xZZ * inv, nTT = xTT * inv, nRZ = xRZ * inv;
    const D pull = 2.0 * mu * dGamma;
    q.s0 = tRR - pull * nRR;
    q.s1 = tZZ - pull * nZZ;
    q.s2 = tTT - pull * nTT;
    q.s3 = tRZ - pull * nRZ;
    q.work = q.work + dGamma * (norm - pull);
    q.epsP = q.epsP + NFD_SQRT_2_3 * dGamma;
}

__device__ __forceinline__ void corot_rotate(D& rr, D& zz, D& rz, const D& cos, const D& sin) {
    const D cc = cos * cos, ss = sin * sin, cs = cos * sin;
    const D r0 = rr, z0 = zz, x0 = rz;
    rr = cc * r0 - 2.0 * cs * x0 + ss * z0;
    zz = ss * r0 + 2.0 * cs * x0 + cc * z0;
    rz = cs * (r0 - z0) + (cc - ss) * x0;
}
__device__ __forceinline__ D sq(const D& a, const D& b) { return a * a + b * b; }

}  // namespace nft

// Derivative state, [direction][item] in every array.
struct TDev {
    double2* tu;     // [NP][N]
    double2* tv;     // [NP][N]
    double2* tf;     // [NP][N]
    double2* tslot;  // [NP][4E]
    double* tsig;    // [NP][4][P]
    double* tep;     // [NP][P]
    double* twork;   // [NP][P]
    double* thg;     // [NP][2][E]
    double* tedge;   // [NP][E]: the derivative of each element's shortest squared edge
    double* edgeVal; // [E]: its value
    double* tstep;   // [2][2][NP]: per control buffer, derivatives of dt and dtPrev
    int* argmin;     // [3]: per control buffer, the first element holding the shortest edge;
                     // [2] counts multipliers differentiated through the iteration
};

namespace nft {

__device__ __forceinline__ D load(const double* a, size_t stride, size_t i, double v) {
    D r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < NL; k++) r.t[k] = a[k * stride + i];
    return r;
}
__device__ __forceinline__ void store(double* a, size_t stride, size_t i, const D& x) {
#pragma unroll
    for (int k = 0; k < NL; k++) a[k * stride + i] = x.t[k];
}
__device__ __forceinline__ void load2(const double2* a, size_t stride, size_t i, double2 v, D& x, D& y) {
    x.v = v.x;
    y.v = v.y;
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const double2 w = a[k * stride + i];
        x.t[k] = w.x;
        y.t[k] = w.y;
    }
}
__device__ __forceinline__ void store2(double2* a, size_t stride, size_t i, const D& x, const D& y) {
#pragma unroll
    for (int k = 0; k < NL; k++) a[k * stride + i] = make_double2(x.t[k], y.t[k]);
}

// The derivative arrays as a thread carrying directions dir0 .. dir0 + NL - 1 sees them:
// every [direction][item] base advanced by dir0 directions.
__device__ __forceinline__ TDev shifted(const TDev& t, int dir0, size_t N, size_t E, size_t P) {
    TDev s = t;
    s.tu += dir0 * N;
    s.tv += dir0 * N;
    s.tf += dir0 * N;
    s.tslot += dir0 * 4 * E;
    s.tsig += dir0 * 4 * P;
    s.tep += dir0 * P;
    s.twork += dir0 * P;
    s.thg += dir0 * 2 * E;
    s.tedge += dir0 * E;
    s.tstep += dir0;
    return s;
}

}  // namespace nft

// The value results of one element, stored by the first of its two threads after both have
// read the state they replace.
struct TElementOut {
    double sig0, sig1, sig2, sig3, epsP, work, hgR, hgZ, hgK, edge;
    double2 slot[4];
};

// One element under reduced integration (solver/ExplicitSolver.java:1373-1504) over Duals:
// the value parts are reduced_element's operations exactly. Stores this thread's derivative
// directions (t is the shifted view) and returns the values in o.
__device__ void tangent_element(const Dev& d, const TDev& t, int e, int4 n, const nft::D& strainDt,
                                const nfm::Tables& T, int dir0, int* unconverged, TElementOut& o) {
    using nft::D;
    const Mat fl = d.mat0;
    const nft::DMat j = nft::dmat(fl, dir0);
    const double mu = fl.mu;
    const size_t N = d.N, E = d.E, P = d.P;
    const D mid = -0.5 * strainDt;
    const int idx[4] = {n.x, n.y, n.z, n.w};
    D r[4], z[4], a[4], c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int i = idx[k];
        const double2 p = d.x0[i];
        D ur, uz, vr, vz;
        nft::load2(t.tu, N, i, d.u[i], ur, uz);
        nft::load2(t.tv, N, i, d.v[i], vr, vz);
        if (d.finite) {
            r[k] = p.x + ur + mid * vr;
            z[k] = p.y + uz + mid * vz;
        } else {
            r[k] = nft::cst(p.x);
            z[k] = nft::cst(p.y);
        }
        a[k] = vr * strainDt;
        c[k] = vz * strainDt;
    }
    D minEdge = nft::cst(NFD_MAX_VALUE);
    if (d.finite)
        minEdge = nft::jmin(minEdge, nft::jmin(nft::jmin(nft::sq(r[1] - r[0], z[1] - z[0]), nft::sq(r[2] - r[1], z[2] - z[1])),
                                               nft::jmin(nft::sq(r[3] - r[2], z[3] - z[2]), nft::sq(r[0] - r[3], z[0] - z[3]))));

    const D J00 = 0.25 * (-r[0] + r[1] + r[2] - r[3]);
    const D J01 = 0.25 * (-z[0] + z[1] + z[2] - z[3]);
    const D J10 = 0.25 * (-r[0] - r[1] + r[2] + r[3]);
    const D J11 = 0.25 * (-z[0] - z[1] + z[2] + z[3]);
    const D det = J00 * J11 - J01 * J10;
    const D inv = 1.0 / det;
    const D d0r = 0.25 * (J01 - J11) * inv;
    const D d1r = 0.25 * (J01 + J11) * inv;
    const D d2r = -d0r;
    const D d3r = -d1r;
    const D d0z = 0.25 * (J10 - J00) * inv;
    const D d1z = -0.25 * (J10 + J00) * inv;
    const D d2z = -d0z;
    const D d3z = -d1z;
    const D dEpsR = d0r * a[0] + d1r * a[1] + d2r * a[2] + d3r * a[3];
    const D dEpsZ = d0z * c[0] + d1z * c[1] + d2z * c[2] + d3z * c[3];
    const D dRdZ = d0z * a[0] + d1z * a[1] + d2z * a[2] + d3z * a[3];
    const D dZdR = d0r * c[0] + d1r * c[1] + d2r * c[2] + d3r * c[3];
    const D dGamRZ = dRdZ + dZdR;
    D rg = nft::cst(0.0), dEpsT = nft::cst(0.0);
    if (d.axi) {
        rg = 0.25 * (r[0] + r[1] + r[2] + r[3]);
        dEpsT = 0.25 * (a[0] + a[1] + a[2] + a[3]) / rg;
    }

    nft::DPoint q;
    q.s1 = nft::load(t.tsig + P, 4 * P, e, d.sig1[e]);
    q.s2 = nft::load(t.tsig + 2 * P, 4 * P, e, d.sig2[e]);
    q.s3 = nft::load(t.tsig + 3 * P, 4 * P, e, d.sig3[e]);
    q.s0 = nft::load(t.tsig, 4 * P, e, d.sig0[e]);
    q.epsP = nft::load(t.tep, P, e, d.epsP[e]);
    q.work = nft::load(t.twork, P, e, d.work[e]);
    if (d.finite) {
        const D spinDt = 0.5 * (dZdR - dRdZ);
        if (spinDt.v != 0.0) {
            const D h = 0.5 * spinDt;
            const D cs = (1.0 - h * h) / (1.0 + h * h);
            const D sn = 2.0 * h / (1.0 + h * h);
            nft::corot_rotate(q.s0, q.s1, q.s3, cs, sn);
        }
    }
    nft::j2_update(q, dEpsR, dEpsZ, dEpsT, dGamRZ, j, fl, strainDt, T, unconverged);
    o.sig0 = q.s0.v;
    o.sig1 = q.s1.v;
    o.sig2 = q.s2.v;
    o.sig3 = q.s3.v;
    o.epsP = q.epsP.v;
    o.work = q.work.v;
    nft::store(t.tsig, 4 * P, e, q.s0);
    nft::store(t.tsig + P, 4 * P, e, q.s1);
    nft::store(t.tsig + 2 * P, 4 * P, e, q.s2);
    nft::store(t.tsig + 3 * P, 4 * P, e, q.s3);
    nft::store(t.tep, P, e, q.epsP);
    nft::store(t.twork, P, e, q.work);

    const D sR = q.s0, sZ = q.s1, sT = q.s2, sRZ = q.s3;
    const D dV = 4.0 * (d.axi ? 2.0 * NFD_PI * rg * det : d.thickness * det);
    D g0r, g1r, g2r, g3r;
    if (d.axi) {
        const D h = sT / rg * dV * 0.25;
        g0r = -((d0r * sR + d0z * sRZ) * dV + h);
        g1r = -((d1r * sR + d1z * sRZ) * dV + h);
        g2r = -((d2r * sR + d2z * sRZ) * dV + h);
        g3r = -((d3r * sR + d3z * sRZ) * dV + h);
    } else {
        g0r = -(d0r * sR + d0z * sRZ) * dV;
        g1r = -(d1r * sR + d1z * sRZ) * dV;
        g2r = -(d2r * sR + d2z * sRZ) * dV;
        g3r = -(d3r * sR + d3z * sRZ) * dV;
    }
    D g0z = -(d0z * sZ + d0r * sRZ) * dV;
    D g1z = -(d1z * sZ + d1r * sRZ) * dV;
    D g2z = -(d2z * sZ + d2r * sRZ) * dV;
    D g3z = -(d3z * sZ + d3r * sRZ) * dV;
    if (d.qhg > 0.0) {
        const D hr = r[0] - r[1] + r[2] - r[3];
        const D hz = z[0] - z[1] + z[2] - z[3];
        const D y0 = 1.0 - hr * d0r - hz * d0z;
        const D y1 = -1.0 - hr * d1r - hz * d1z;
        const D y2 = 1.0 - hr * d2r - hz * d2z;
        const D y3 = -1.0 - hr * d3r - hz * d3z;
        const D k = d.qhg * mu * dV *
                    (d0r * d0r + d1r * d1r + d2r * d2r + d3r * d3r + d0z * d0z + d1z * d1z + d2z * d2z + d3z * d3z);
        const D hgR = nft::load(t.thg, 2 * E, e, d.hgR[e]);
        const D hgZ = nft::load(t.thg + E, 2 * E, e, d.hgZ[e]);
        const D qr = hgR + k * (y0 * a[0] + y1 * a[1] + y2 * a[2] + y3 * a[3]);
        const D qz = hgZ + k * (y0 * c[0] + y1 * c[1] + y2 * c[2] + y3 * c[3]);
        o.hgR = qr.v;
        o.hgZ = qz.v;
        nft::store(t.thg, 2 * E, e, qr);
        nft::store(t.thg + E, 2 * E, e, qz);
        g0r = g0r - y0 * qr; g1r = g1r - y1 * qr; g2r = g2r - y2 * qr; g3r = g3r - y3 * qr;
        g0z = g0z - y0 * qz; g1z = g1z - y1 * qz; g2z = g2z - y2 * qz; g3z = g3z - y3 * qz;
        o.hgK = k.v;
    }
    const D gr[4] = {g0r, g1r, g2r, g3r}, gz[4] = {g0z, g1z, g2z, g3z};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        o.slot[k] = make_double2(gr[k].v, gz[k].v);
        nft::store2(t.tslot, 4 * E, 4 * (size_t)e + k, gr[k], gz[k]);
    }
    o.edge = minEdge.v;
    nft::store(t.tedge, E, e, minEdge);
}

// Two threads per element, in the two halves of the workgroup: the first carries directions
// 0-2, the second 3-5. Both read the element's value state and compute the same values; after
// a barrier, which every thread reaches, only the first stores them.
__global__ void __launch_bounds__(BLOCK) k_t_element(Dev d, TDev t, int q, int mode) {
    Ctl* c = d.ctl;
    Step& st = c->s[q];
    if (mode == MODE_STEP && !st.go) return;
    __shared__ double sLog[97][3];
    __shared__ double sExp[64][2];
    const nfm::Tables T = nfd::stage_tables(sLog, sExp);
    const int half = threadIdx.x / (BLOCK / 2);
    const int dir0 = half * NL;
    const TDev th = nft::shifted(t, dir0, d.N, d.E, d.P);
    nft::D strainDt = nft::cst(0.0);
    if (mode != MODE_INITIAL) strainDt = nft::load(th.tstep + (size_t)q * 2 * NP, 1, 0, st.dt);
    const int e = blockIdx.x * (BLOCK / 2) + threadIdx.x % (BLOCK / 2);
    const bool active = e < d.E;
    bool sawNaN = false;
    double edge = NFD_MAX_VALUE;
    TElementOut o{};
    if (active) {
        tangent_element(d, th, e, d.conn[e], strainDt, T, dir0, half == 0 ? t.argmin + 2 : nullptr, o);
        edge = o.edge;
        sawNaN = edge != edge;
    }
    __syncthreads();
    if (active && half == 0) {
        d.sig0[e] = o.sig0;
        d.sig1[e] = o.sig1;
        d.sig2[e] = o.sig2;
        d.sig3[e] = o.sig3;
        d.epsP[e] = o.epsP;
        d.work[e] = o.work;
        if (d.qhg > 0.0) {
            d.hgR[e] = o.hgR;
            d.hgZ[e] = o.hgZ;
            d.hgK[e] = o.hgK;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) d.slot[4 * e + k] = o.slot[k];
        t.edgeVal[e] = o.edge;
    }
    unsigned long long bits = sawNaN ? MAX_BITS : (unsigned long long)nfm::bits_of(edge);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o2 = __shfl_xor(bits, off, 64);
        bits = o2 < bits ? o2 : bits;
    }
    const unsigned long long anyNaN = __ballot(sawNaN);
    if ((threadIdx.x & 63) == 0) {
        if (bits != MAX_BITS) atomicMin(&st.minEdgeBits, bits);
        if (anyNaN) atomicOr(&st.nanEdge, 1);
    }
}

// The first element whose shortest edge is the step's minimum: the CFL step's derivative
// is that element's.
__global__ void __launch_bounds__(BLOCK) k_t_argmin(Dev d, TDev t, int q) {
    const Step& st = d.ctl->s[q];
    if (!st.go) return;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e < d.E && (unsigned long long)nfm::bits_of(t.edgeVal[e]) == st.minEdgeBits) atomicMin(&t.argmin[q], e);
}

// k_gn (PART_ALL) over Duals, without the wall's audit partials.
// Two threads per node, as in k_t_element: directions 0-2 and 3-5. The node's value state is
// read by both and written by the first after a barrier.
template <bool WALL>
__global__ void __launch_bounds__(BLOCK) k_t_gn(Dev d, TDev t, int q, int phase, int gatherFirst) {
    using nft::D;
    Ctl* c = d.ctl;
    const Step& in = c->s[q];
    Step& out = c->s[q ^ 1];
    const int half = threadIdx.x / (BLOCK / 2);
    const int dir0 = half * NL;
    const TDev th = nft::shifted(t, dir0, d.N, d.E, d.P);
    const bool lead = blockIdx.x == 0 && threadIdx.x % (BLOCK / 2) == 0;  // one per half
    const double* tin = th.tstep + (size_t)q * 2 * NP;
    double* tout = th.tstep + (size_t)(q ^ 1) * 2 * NP;
    if (phase == PHASE_STEP && !in.go) {
        if (lead) {
            if (half == 0) {
                out = in;
                t.argmin[q ^ 1] = 0x7FFFFFFF;
            }
            for (int k = 0; k < NL; k++) {
                tout[k] = tin[k];
                tout[NP + k] = tin[NP + k];
            }
        }
        return;
    }
    const bool gather = phase == PHASE_STEP || gatherFirst;
    D dt = nft::load(tin, 1, 0, in.dt);
    D dtPrev = nft::load(tin + NP, 1, 0, in.dtPrev);
    bool next = true;
    if (phase == PHASE_STEP) {
        if (d.finite) {
            D minEdge = nft::cst(NFD_MAX_VALUE);
            if (!in.nanEdge && in.minEdgeBits != MAX_BITS) {
                const int am = t.argmin[q];
                minEdge = nft::load(th.tedge, d.E, am, nfm::from_bits(in.minEdgeBits));
            }
            dtPrev = dt;
            dt = d.cfl * nft::sqrt(minEdge * d.massScale) / d.fastestWave;
        }
        next = in.time < in.endTime && in.steps < in.stepLimit;
    } else if (phase == PHASE_ASSEMBLE_END) {
        next = false;
    }
    if (lead) {
        if (half == 0) {
            Step o = in;
            o.dt = dt.v;
            o.dtPrev = dtPrev.v;
            o.go = next ? 1 : 0;
            if (next) {
                o.time = in.time + dt.v;
                o.steps = in.steps + 1;
            }
            o.nanEdge = 0;
            o.minEdgeBits = MAX_BITS;
            out = o;
            t.argmin[q ^ 1] = 0x7FFFFFFF;
        }
        for (int k = 0; k < NL; k++) {
            tout[k] = dt.t[k];
            tout[NP + k] = dtPrev.t[k];
        }
    }
    const int i = blockIdx.x * (BLOCK / 2) + threadIdx.x % (BLOCK / 2);
    const bool active = i < d.N;
    const size_t N = d.N, E4 = 4 * (size_t)d.E;
    D fr = nft::cst(0.0), fz = nft::cst(0.0);
    if (active) {
        if (gather) {
            for (int k = d.slotStart[i]; k < d.slotStart[i + 1]; k++) {
                const int s = d.slotOf[k];
                D sr, sz;
                nft::load2(th.tslot, E4, s, d.slot[s], sr, sz);
                fr = fr + sr;
                fz = fz + sz;
            }
            const double m = d.mass[i];
            fr = fr + m * d.gravityR;
            fz = fz + m * d.gravityZ;
            if (half == 0) d.f[i] = make_double2(fr.v, fz.v);  // read by no other thread here
            nft::store2(th.tf, N, i, fr, fz);
        } else {
            nft::load2(th.tf, N, i, d.f[i], fr, fz);
        }
    }
    if (!next) return;  // uniform over the grid
    const D dtAvg = 0.5 * (dtPrev + dt);
    const D half_ = 0.5 * d.damping * dtAvg;
    const D decay = (1.0 - half_) / (1.0 + half_);
    const D gain = dtAvg / (1.0 + half_);

    // nodal_node over Duals (solver/ExplicitSolver.java:998-1062, solver/RigidWall.java:163-279)
    D vr = nft::cst(0.0), vz = nft::cst(0.0), ur = nft::cst(0.0), uz = nft::cst(0.0);
    uint8_t fl = 0;
    if (active) {
        const double m = d.mass[i];
        const double invM = 1.0 / m;
        fl = d.flags[i];
        nft::load2(th.tv, N, i, d.v[i], vr, vz);
        nft::load2(th.tu, N, i, d.u[i], ur, uz);
        const bool resting = WALL && (fl & ON_WALL) && fz.v <= 0.0;
        if (fl & FIXED_R) {
            vr = nft::cst(0.0);
        } else if (WALL && resting && d.wallFrictional) {
            const D normal = -fz;
            const D trial = -m * decay * vr / gain - fr;
            D limit = nft::cst(HUGE_VAL);
            if (!__builtin_isinf(d.friction)) limit = nft::seed(d.friction, SP_FRICTION, dir0) * normal;
            if (nfd::jabs(trial.v) <= limit.v) {
                vr = nft::cst(0.0);
            } else {
                D tangential = limit;
                tangential.v = __builtin_copysign(limit.v, trial.v);
                if (__builtin_signbit(trial.v))
                    for (int k = 0; k < NL; k++) tangential.t[k] = -limit.t[k];
                vr = decay * vr + gain * (fr + tangential) / m;
            }
            ur = ur + vr * dt;
        } else {
            vr = decay * vr + gain * fr * invM;
            ur = ur + vr * dt;
        }
        if (fl & FIXED_Z) {
            vz = nft::cst(0.0);
        } else if (WALL && resting) {
            // held: nothing moves
        } else {
            const D trial = decay * vz + gain * fz * invM;
            const D zc = d.z0[i] + uz;
            if (WALL && zc.v + trial.v * dt.v < d.plane) {
                uz = uz + (d.plane - zc);
                vz = nft::cst(0.0);
                fl |= ON_WALL;
            } else {
                vz = trial;
                uz = uz + vz * dt;
                if (WALL) fl &= (uint8_t)~ON_WALL;
            }
        }
        nft::store2(th.tv, N, i, vr, vz);
        nft::store2(th.tu, N, i, ur, uz);
    }
    __syncthreads();  // both threads of a node have read its values
    if (active && half == 0) {
        d.v[i] = make_double2(vr.v, vz.v);
        d.u[i] = make_double2(ur.v, uz.v);
        if (WALL) d.flags[i] = fl;
    }
}

// Discrete adjoint (reverse mode) of the step loop: gradients of objectives of the final
// nodal state with respect to (anvil friction, Johnson-Cook A, B, n, C, m), for the Taylor
// case the parameter sensitivities cover (reduced integration of one Johnson-Cook material,
// no kinematic hardening, damage, erosion, contact or pressure edges).
//
// Included by neofiz_hip.hip after tangent.h.
//
// The forward run records checkpoints of the whole solver state every K steps. The reverse
// sweep walks the checkpoints from the last to the first; for each it restores the state,
// reruns the K steps with the plain kernels while recording what each step read (the
// element state before its pass, the nodes after their update, the gathered forces, the
// step's dt and its shortest edge's element) and what the return map decided (the
// multiplier and whether the Newton solve met its tolerance), then runs the K steps
// backwards. One reverse step is three launches:
//   k_adj_reduce   one workgroup: the adjoints of the step's dt and previous dt from the
//                  per-block partials of the step after it, in a fixed order; the running
//                  parameter gradient; the adjoint of the shortest squared edge that set dt;
//   k_adj_element  per element: its forward values recomputed from the record (the element
//                  pass's own operations; the multiplier is read, not solved again), then
//                  the adjoint of every operation in reverse order: hourglass, internal
//                  force with the hoop term, the J2 return (the multiplier through the
//                  implicit function theorem at the Newton root, or, where the solve stopped
//                  on its iteration cap, through the Jacobian of the iterate it returned),
//                  Cayley rotation, strain increments, B matrix and mid-step coordinates;
//                  writes the adjoints of its four corners' displacement and velocity to
//                  per-corner slots, the adjoint of its own previous state in place;
//   k_adj_node     per node: its corners' adjoints gathered in the forward gather's slot
//                  order, then the adjoint of the central-difference update with the rigid
//                  wall (the branch the forward took: free, held, landed, stuck, slid on the
//                  Coulomb cone); writes the adjoint of the force it read.
// No atomic touches an adjoint value: the element-to-node transpose is the forward gather
// read backwards (each corner slot belongs to one element), and every sum over elements or
// nodes goes through per-block partials reduced in a fixed order. The gradients are
// therefore deterministic; they equal the forward-mode derivatives of tangent.h to rounding.
//
// Several objectives are carried at once: each element and node has one thread per
// objective, and each thread recomputes the forward values it needs (the recomputation is
// cheap next to the Newton solves it does not repeat).
#pragma once

constexpr int AW_MAX = 2;        // objectives carried by one reverse sweep
constexpr int A_EP = 6;          // per-element-block partials: dt, A, B, n, C, m
constexpr int A_NP = 4;          // per-node-block partials: dt, decay, gain, friction
enum { AP_DT = 0, AP_A, AP_B, AP_N, AP_C, AP_M };
enum { AN_DT = 0, AN_DECAY, AN_GAIN, AN_MU };
constexpr int CAP_DIRS = 8;      // inputs of the Newton solve: norm, epsP, thermal, dt, A, B, n, C

// Reverse-sweep control, one per objective.
struct AdjCtl {
    double grad[AW_MAX][NP];     // running gradient (friction, A, B, n, C, m)
    double dtBar[AW_MAX];        // adjoint of the dt of the step being reversed (after reduce)
    double dtPrevBar[AW_MAX];    // adjoint of dtPrev carried from the step after
    double edgeBar[AW_MAX];      // adjoint of the shortest squared edge of the step being reversed
};

// Per-step record of one segment (index j = steps since the segment's checkpoint).
struct ARec {
    int K;                      // steps a segment holds
    // nodes, before the element pass of step j (j = 0: at the checkpoint):
    // u, v and flags after step j's nodal update, f gathered in step j - 1
    double2* u;                 // [(K + 1) N]
    double2* v;                 // [(K + 1) N]
    double2* f;                 // [(K + 1) N]
    uint8_t* flags;             // [(K + 1) N]
    // elements, j = 1..K: the state the element pass of step j read
    double* sig;                // [K][4][E]
    double* ep;                 // [K][E]
    double* work;               // [K][E]
    double* hg;                 // [K][2][E]
    double* dGamma;             // [K][E] the return map's multiplier
    uint8_t* status;            // [K][E] 0 elastic, 1 plastic (tolerance met), 2 plastic (cap)
    double* capJac;             // [K][CAP_DIRS][E] d multiplier / d solve inputs, status 2 only
    double* edge;               // [E] shortest squared edge of each element, last pass
    double* stepDt;             // [K + 1]: dt, dtPrev of step j
    double* stepDtPrev;
    unsigned long long* minEdgeBits;  // [K + 1]
    int* argmin;                // [K + 1] first element holding the shortest edge
    int* capList;               // [K E] record items (j - 1) E + e whose solve hit its cap
    int* capCount;              // [1]
};

// Adjoint state, [objective][item] in every array.
struct ADev {
    ARec rec;
    int W;                      // objectives this sweep carries (1 or AW_MAX)
    double2* uBar;              // [W][N] adjoint of u after the nodal update being reversed
    double2* vBar;              // [W][N]
    double2* fBar;              // [W][N] adjoint of the gathered force
    double4* cornerBar;         // [W][4E] (u_r, u_z, v_r, v_z) adjoints of each element corner
    double* sigBar;             // [W][4][E]
    double* epBar;              // [W][E]
    double* workBar;            // [W][E]
    double* hgBar;              // [W][2][E]
    double* elemPartial;        // [W][elementBlocks][A_EP]
    double* nodePartial;        // [W][nodeBlocks][A_NP]
    int elemBlocks, nodeBlocks;
    AdjCtl* ctl;
};

namespace nfa {

// ------------------------------------------------------------------ recording pass

// Copies what the element pass of step j is about to read, and the nodes' state, into the
// segment's record. Launched for j = 0 (the checkpoint's nodes alone); for j >= 1,
// k_element_rec does the same copies itself.
__global__ void __launch_bounds__(BLOCK) k_adj_save(Dev d, ARec a, int j, int q) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const size_t N = d.N, E = d.E;
    if (i < d.N) {
        a.u[j * N + i] = d.u[i];
        a.v[j * N + i] = d.v[i];
        a.f[j * N + i] = d.f[i];
        a.flags[j * N + i] = d.flags[i];
    }
    if (j > 0 && i < d.E) {
        const size_t b = (size_t)(j - 1);
        a.sig[(b * 4 + 0) * E + i] = d.sig0[i];
        a.sig[(b * 4 + 1) * E + i] = d.sig1[i];
        a.sig[(b * 4 + 2) * E + i] = d.sig2[i];
        a.sig[(b * 4 + 3) * E + i] = d.sig3[i];
        a.ep[b * E + i] = d.epsP[i];
        a.work[b * E + i] = d.work[i];
        a.hg[(b * 2 + 0) * E + i] = d.hgR[i];
        a.hg[(b * 2 + 1) * E + i] = d.hgZ[i];
    }
    if (i == 0) {
        const Step& st = d.ctl->s[q];
        a.stepDt[j] = st.dt;
        a.stepDtPrev[j] = st.dtPrev;
        a.argmin[j] = 0x7FFFFFFF;
    }
}

// k_element for the recorded Taylor configuration, plus the return map's multiplier and
// outcome and each element's shortest edge. The element arithmetic is reduced_element's.
// The launch covers max(N, E) threads and also does what k_adj_save (step j) and
// k_adj_argmin (step j - 1) would do in launches of their own: thread i first copies node i
// and element i into the record as this pass is about to read them, and, for j > 1, takes
// part in the first-argmin of step j - 1's shortest edge from the edges that step's pass left
// in a.edge (which this pass then overwrites, each thread its own element).
__global__ void __launch_bounds__(BLOCK, 3) k_element_rec(Dev d, ARec a, int j, int q) {
    Ctl* c = d.ctl;
    Step& st = c->s[q];
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    {
        const size_t N = d.N, E = d.E;
        if (e < d.N) {
            a.u[j * N + e] = d.u[e];
            a.v[j * N + e] = d.v[e];
            a.f[j * N + e] = d.f[e];
            a.flags[j * N + e] = d.flags[e];
        }
        if (e < d.E) {
            const size_t b = (size_t)(j - 1);
            a.sig[(b * 4 + 0) * E + e] = d.sig0[e];
            a.sig[(b * 4 + 1) * E + e] = d.sig1[e];
            a.sig[(b * 4 + 2) * E + e] = d.sig2[e];
            a.sig[(b * 4 + 3) * E + e] = d.sig3[e];
            a.ep[b * E + e] = d.epsP[e];
            a.work[b * E + e] = d.work[e];
            a.hg[(b * 2 + 0) * E + e] = d.hgR[e];
            a.hg[(b * 2 + 1) * E + e] = d.hgZ[e];
        }
        if (e == 0) {
            a.stepDt[j] = st.dt;
            a.stepDtPrev[j] = st.dtPrev;
            a.argmin[j] = 0x7FFFFFFF;
        }
        if (j > 1) {  // step j - 1 read Step q ^ 1, which still holds its shortest edge
            const Step& sp = c->s[q ^ 1];
            if (e == 0) a.minEdgeBits[j - 1] = sp.nanEdge ? MAX_BITS : sp.minEdgeBits;
            if (e < d.E && !sp.nanEdge && (unsigned long long)nfm::bits_of(a.edge[e]) == sp.minEdgeBits)
                atomicMin(&a.argmin[j - 1], e);
        }
    }
    if (!st.go) return;
    __shared__ double sLog[97][3];
    __shared__ double sExp[64][2];
    const nfm::Tables T = nfd::stage_tables(sLog, sExp);
    double edge = NFD_MAX_VALUE;
    bool sawNaN = false;
    if (e < d.E) {
        double dg = 0.0;
        int status = 0;
        edge = reduced_element<false, false, true>(d, e, d.conn[e], st.dt, T, &dg, &status);
        sawNaN = edge != edge;
        const size_t b = (size_t)(j - 1) * d.E + e;
        a.dGamma[b] = dg;
        a.status[b] = (uint8_t)status;
        a.edge[e] = edge;
        if (status == 2) a.capList[atomicAdd(a.capCount, 1)] = (int)b;
    }
    unsigned long long bits = sawNaN ? MAX_BITS : (unsigned long long)nfm::bits_of(edge);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o < bits ? o : bits;
    }
    const unsigned long long anyNaN = __ballot(sawNaN);
    __shared__ unsigned long long ldsMin[BLOCK / 64];
    __shared__ int ldsNaN;
    if (threadIdx.x == 0) ldsNaN = 0;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        ldsMin[threadIdx.x >> 6] = bits;
        if (anyNaN) ldsNaN = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = ldsMin[0];
        for (int w = 1; w < BLOCK / 64; w++) m = ldsMin[w] < m ? ldsMin[w] : m;
        if (m != MAX_BITS) atomicMin(&st.minEdgeBits, m);
        if (ldsNaN) atomicOr(&st.nanEdge, 1);
    }
    if (d.wall && blockIdx.x == 0) reduce_wall_partials(d, d.wallPartial, d.wallCountPartial, d.nodalBlocks);
}

// The step's shortest squared edge and the first element holding it (the CFL step's
// derivative is that element's), after the element pass of step j: launched for a segment's
// last step (k_element_rec of step j + 1 does it for the others).
__global__ void __launch_bounds__(BLOCK) k_adj_argmin(Dev d, ARec a, int j, int q) {
    const Step& st = d.ctl->s[q];
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e == 0) a.minEdgeBits[j] = st.nanEdge ? MAX_BITS : st.minEdgeBits;
    if (e < d.E && !st.nanEdge && (unsigned long long)nfm::bits_of(a.edge[e]) == st.minEdgeBits)
        atomicMin(&a.argmin[j], e);
}

// ------------------------------------------------------------------ the Newton solve's Jacobian

// A value with one derivative, for the solve's Jacobian where it stopped on its cap: the
// value part of every operation is the plain solve's IEEE operation, so the iterates and
// every branch are the plain solve's.
struct D1 {
    double v, t;
};
__device__ __forceinline__ D1 mk(double v, double t = 0.0) { return D1{v, t}; }
__device__ __forceinline__ D1 operator+(D1 a, D1 b) { return D1{a.v + b.v, a.t + b.t}; }
__device__ __forceinline__ D1 operator-(D1 a, D1 b) { return D1{a.v - b.v, a.t - b.t}; }
__device__ __forceinline__ D1 operator*(D1 a, D1 b) { return D1{a.v * b.v, a.t * b.v + a.v * b.t}; }
__device__ __forceinline__ D1 operator/(D1 a, D1 b) {
    const double r = a.v / b.v;
    return D1{r, (a.t - r * b.t) / b.v};
}
__device__ __forceinline__ D1 operator*(double a, D1 b) { return D1{a * b.v, a * b.t}; }
__device__ __forceinline__ D1 operator*(D1 a, double b) { return D1{a.v * b, a.t * b}; }
__device__ __forceinline__ D1 operator/(D1 a, double b) { return D1{a.v / b, a.t / b}; }
__device__ __forceinline__ D1 operator+(double a, D1 b) { return D1{a + b.v, b.t}; }
__device__ __forceinline__ D1 operator-(double a, D1 b) { return D1{a - b.v, -b.t}; }

// The solve of nfd::j2_solve_loop over D1 (solver/J2.java:332-375): the derivative of the
// iterate it returns. Inputs are seeded by the caller.
__device__ __forceinline__ D1 solve_d1(D1 norm, double stiffness, D1 plasticStrain, D1 thermal, D1 dt,
                                       D1 A, D1 B, D1 n, D1 C, const Mat& j, const nfm::Tables& T) {
    const double tolerance = 1.0e-10 * A.v;
    D1 lo = mk(0.0);
    D1 hi = norm / stiffness;
    D1 dGamma = hi;
    for (int iteration = 0; iteration < 25; iteration++) {
        const D1 dEps = NFD_SQRT_2_3 * dGamma;
        const D1 eps = plasticStrain + dEps;
        const D1 rate = dt.v > 0.0 ? dEps / dt : mk(0.0);
        D1 strain;
        if (eps.v <= 0.0) {
            strain = A;
        } else {
            double lx;
            const double pw = nfm::pow_log_ordinary(eps.v, n.v, T, lx);
            const D1 p = D1{pw, pw * n.v / eps.v * eps.t + pw * lx * n.t};
            strain = A + B * p;
        }
        const bool clamped = rate.v <= j.eps0;
        D1 rateFactor = mk(1.0);
        if (!clamped) {
            const D1 x = j.eps0 == 1.0 ? rate : rate / j.eps0;
            const D1 lg = D1{nfm::log_ordinary(x.v, T), x.t / x.v};
            rateFactor = 1.0 + C * lg;
        }
        const D1 g = norm - stiffness * dGamma - NFD_SQRT_2_3 * strain * rateFactor * thermal;
        if (g.v > 0.0) lo = dGamma;
        else hi = dGamma;
        if (nfd::jabs(g.v) <= tolerance) return dGamma;
        const D1 slope0 = (-stiffness) - NFD_TWO_THIRDS * (n * (strain - A) / eps) * rateFactor * thermal;
        const D1 slope = clamped ? slope0 : slope0 - NFD_SQRT_2_3 * strain * C * thermal / dGamma;
        const D1 next = dGamma - g / slope;
        dGamma = (next.v > lo.v && next.v < hi.v) ? next : 0.5 * (lo + hi);
    }
    return dGamma;
}

// ------------------------------------------------------------------ element forward values

// The element pass's values, recomputed from the record (reduced_element's operations).
struct EF {
    double mid, r[4], z[4], a[4], c[4];
    double J00, J01, J10, J11, det, inv, d0r, d1r, d0z, d1z;
    double dEpsR, dEpsZ, dRdZ, dZdR, dGamRZ, rg, dEpsT;
    bool rotated;
    double h, cs, sn, cc, ss, csn;
    double s0, s1, s2, s3;      // stress as stored
    double q0, q1, q3;          // after the rotation
    double tRR, tZZ, tTT, tRZ, xRR, xZZ, xTT, xRZ, norm;
    double work, epsP;
    int thermalCase;            // 0 homologous <= 0, 1 homologous >= 1, 2 between
    double hom, thermal, homPow, homLog;
    int status;
    double dGamma, invNorm, nRR, nZZ, nTT, nRZ, pull;
    double sR, sZ, sT, sRZ, dV;
    double hr, hz, y[4], D2, k, Sa, Sc, qr, qz;
};

__device__ __forceinline__ void element_forward(const Dev& d, const ARec& a, int j, int e, int4 n,
                                                const Mat& fl, const nfm::Tables& T, EF& f) {
    const size_t N = d.N, E = d.E;
    const double dt = a.stepDt[j];
    f.mid = -0.5 * dt;
    const int idx[4] = {n.x, n.y, n.z, n.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double2 p = d.x0[idx[k]];
        const double2 u = a.u[j * N + idx[k]];
        const double2 v = a.v[j * N + idx[k]];
        if (d.finite) {
            f.r[k] = p.x + u.x + f.mid * v.x;
            f.z[k] = p.y + u.y + f.mid * v.y;
        } else {
            f.r[k] = p.x;
            f.z[k] = p.y;
        }
        f.a[k] = v.x * dt;
        f.c[k] = v.y * dt;
    }
    f.J00 = 0.25 * (-f.r[0] + f.r[1] + f.r[2] - f.r[3]);
    f.J01 = 0.25 * (-f.z[0] + f.z[1] + f.z[2] - f.z[3]);
    f.J10 = 0.25 * (-f.r[0] - f.r[1] + f.r[2] + f.r[3]);
    f.J11 = 0.25 * (-f.z[0] - f.z[1] + f.z[2] + f.z[3]);
    f.det = f.J00 * f.J11 - f.J01 * f.J10;
    f.inv = 1.0 / f.det;
    f.d0r = 0.25 * (f.J01 - f.J11) * f.inv;
    f.d1r = 0.25 * (f.J01 + f.J11) * f.inv;
    f.d0z = 0.25 * (f.J10 - f.J00) * f.inv;
    f.d1z = -0.25 * (f.J10 + f.J00) * f.inv;
    const double dr[4] = {f.d0r, f.d1r, -f.d0r, -f.d1r}, dz[4] = {f.d0z, f.d1z, -f.d0z, -f.d1z};
    f.dEpsR = dr[0] * f.a[0] + dr[1] * f.a[1] + dr[2] * f.a[2] + dr[3] * f.a[3];
    f.dEpsZ = dz[0] * f.c[0] + dz[1] * f.c[1] + dz[2] * f.c[2] + dz[3] * f.c[3];
    f.dRdZ = dz[0] * f.a[0] + dz[1] * f.a[1] + dz[2] * f.a[2] + dz[3] * f.a[3];
    f.dZdR = dr[0] * f.c[0] + dr[1] * f.c[1] + dr[2] * f.c[2] + dr[3] * f.c[3];
    f.dGamRZ = f.dRdZ + f.dZdR;
    f.rg = 0.0;
    f.dEpsT = 0.0;
    if (d.axi) {
        f.rg = 0.25 * (f.r[0] + f.r[1] + f.r[2] + f.r[3]);
        f.dEpsT = 0.25 * (f.a[0] + f.a[1] + f.a[2] + f.a[3]) / f.rg;
    }
    const size_t b = (size_t)(j - 1);
    f.s0 = a.sig[(b * 4 + 0) * E + e];
    f.s1 = a.sig[(b * 4 + 1) * E + e];
    f.s2 = a.sig[(b * 4 + 2) * E + e];
    f.s3 = a.sig[(b * 4 + 3) * E + e];
    f.epsP = a.ep[b * E + e];
    f.work = a.work[b * E + e];
    f.q0 = f.s0;
    f.q1 = f.s1;
    f.q3 = f.s3;
    f.rotated = false;
    if (d.finite) {
        const double spinDt = 0.5 * (f.dZdR - f.dRdZ);
        if (spinDt != 0.0) {
            f.rotated = true;
            f.h = 0.5 * spinDt;
            f.cs = (1.0 - f.h * f.h) / (1.0 + f.h * f.h);
            f.sn = 2.0 * f.h / (1.0 + f.h * f.h);
            f.cc = f.cs * f.cs;
            f.ss = f.sn * f.sn;
            f.csn = f.cs * f.sn;
            f.q0 = f.cc * f.s0 - 2.0 * f.csn * f.s3 + f.ss * f.s1;
            f.q1 = f.ss * f.s0 + 2.0 * f.csn * f.s3 + f.cc * f.s1;
            f.q3 = f.csn * (f.s0 - f.s1) + (f.cc - f.ss) * f.s3;
        }
    }
    // J2 (solver/J2.java:163-296), values only
    const double lt = fl.lambda * (f.dEpsR + f.dEpsZ + f.dEpsT);
    f.tRR = f.q0 + lt + 2.0 * fl.mu * f.dEpsR;
    f.tZZ = f.q1 + lt + 2.0 * fl.mu * f.dEpsZ;
    f.tTT = f.s2 + lt + 2.0 * fl.mu * f.dEpsT;
    f.tRZ = f.q3 + fl.mu * f.dGamRZ;
    const double p = (f.tRR + f.tZZ + f.tTT) / 3.0;
    f.xRR = f.tRR - p;
    f.xZZ = f.tZZ - p;
    f.xTT = f.tTT - p;
    f.xRZ = f.tRZ;
    f.norm = __builtin_sqrt(f.xRR * f.xRR + f.xZZ * f.xZZ + f.xTT * f.xTT + 2.0 * f.xRZ * f.xRZ);
    f.status = a.status[b * E + e];
    f.dGamma = a.dGamma[b * E + e];
    if (f.status == 0) {
        f.sR = f.tRR;
        f.sZ = f.tZZ;
        f.sT = f.tTT;
        f.sRZ = f.tRZ;
    } else {
        const double temperature = fl.troom + fl.thermalCoefficient * f.work;
        f.hom = (temperature - fl.troom) / (fl.tmelt - fl.troom);
        f.thermalCase = f.hom <= 0.0 ? 0 : (f.hom >= 1.0 ? 1 : 2);
        f.thermal = 1.0;
        if (f.thermalCase == 1) f.thermal = 0.0;
        if (f.thermalCase == 2) {
            f.homPow = nfm::pow_log_ordinary(f.hom, fl.m, T, f.homLog);
            f.thermal = 1.0 - f.homPow;
        }
        f.invNorm = 1.0 / f.norm;
        f.nRR = f.xRR * f.invNorm;
        f.nZZ = f.xZZ * f.invNorm;
        f.nTT = f.xTT * f.invNorm;
        f.nRZ = f.xRZ * f.invNorm;
        f.pull = 2.0 * fl.mu * f.dGamma;
        f.sR = f.tRR - f.pull * f.nRR;
        f.sZ = f.tZZ - f.pull * f.nZZ;
        f.sT = f.tTT - f.pull * f.nTT;
        f.sRZ = f.tRZ - f.pull * f.nRZ;
    }
    f.dV = 4.0 * (d.axi ? 2.0 * NFD_PI * f.rg * f.det : d.thickness * f.det);
    if (d.qhg > 0.0) {
        f.hr = f.r[0] - f.r[1] + f.r[2] - f.r[3];
        f.hz = f.z[0] - f.z[1] + f.z[2] - f.z[3];
        const double sgn[4] = {1.0, -1.0, 1.0, -1.0};
#pragma unroll
        for (int k = 0; k < 4; k++) f.y[k] = sgn[k] - f.hr * dr[k] - f.hz * dz[k];
        f.D2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] + dr[3] * dr[3] + dz[0] * dz[0] + dz[1] * dz[1] +
               dz[2] * dz[2] + dz[3] * dz[3];
        f.k = d.qhg * fl.mu * f.dV * f.D2;
        f.Sa = f.y[0] * f.a[0] + f.y[1] * f.a[1] + f.y[2] * f.a[2] + f.y[3] * f.a[3];
        f.Sc = f.y[0] * f.c[0] + f.y[1] * f.c[1] + f.y[2] * f.c[2] + f.y[3] * f.c[3];
        f.qr = a.hg[(b * 2 + 0) * E + e] + f.k * f.Sa;
        f.qz = a.hg[(b * 2 + 1) * E + e] + f.k * f.Sc;
    }
}

// Jacobian of the capped solve's returned multiplier with respect to its inputs, one
// direction per thread (eight threads per listed item: the element passes whose solve hit
// its cap in the recorded segment).
__global__ void __launch_bounds__(BLOCK) k_adj_cap(Dev d, ARec a, int count) {
    __shared__ double sLog[97][3];
    __shared__ double sExp[64][2];
    const nfm::Tables T = nfd::stage_tables(sLog, sExp);
    const long long t = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const int dir = (int)(t % CAP_DIRS);
    const long long item = t / CAP_DIRS;
    if (item >= count) return;
    const int b = a.capList[item];
    const int j = 1 + b / d.E;
    const int e = b % d.E;
    const Mat& fl = d.mat0;
    EF f;
    element_forward(d, a, j, e, d.conn[e], fl, T, f);
    const double stiffness = 2.0 * fl.mu + NFD_TWO_THIRDS * fl.hKin;
    const D1 norm = mk(f.norm, dir == 0), epsP = mk(f.epsP, dir == 1), thermal = mk(f.thermal, dir == 2);
    const D1 dt = mk(a.stepDt[j], dir == 3);
    const D1 A = mk(fl.a, dir == 4), B = mk(fl.b, dir == 5), n = mk(fl.n, dir == 6), C = mk(fl.c, dir == 7);
    const D1 g = solve_d1(norm, stiffness, epsP, thermal, dt, A, B, n, C, fl, T);
    a.capJac[((size_t)(j - 1) * CAP_DIRS + dir) * d.E + e] = g.t;
}

// ------------------------------------------------------------------ reverse sweep

// Sums `count` partials of stride `stride` starting at p (field k) in a fixed order: each
// thread a strided share, then a wave tree and the waves in order. Call from the whole block.
__device__ double ordered_sum(const double* p, int count, int stride, int k, double* lds) {
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += BLOCK) s += p[(size_t)i * stride + k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < BLOCK / 64; w++) t += lds[w];
    __syncthreads();
    return t;
}

// Closes the reverse of step n + 1 (its partials) and opens the reverse of step n: the
// adjoints of dt_{n+1} and dtPrev_{n+1}, the gradient, and the adjoint of the shortest
// squared edge of step n, which set dt_{n+1} = cfl sqrt(edge massScale) / c.
// jNext: record index of step n + 1 (its dt, dtPrev; -1 when there is no such step),
// jEdge: record index of step n (-1 when none).
__global__ void __launch_bounds__(BLOCK) k_adj_reduce(Dev d, ADev ad, double dtNext, double dtPrevNext,
                                                      int hasNext, double minEdgeN, int hasEdge) {
    __shared__ double lds[BLOCK / 64];
    AdjCtl* c = ad.ctl;
    for (int w = 0; w < ad.W; w++) {
        const double* ep = ad.elemPartial + (size_t)w * ad.elemBlocks * A_EP;
        const double* np = ad.nodePartial + (size_t)w * ad.nodeBlocks * A_NP;
        double s[A_EP], m[A_NP];
        for (int k = 0; k < A_EP; k++) s[k] = ordered_sum(ep, ad.elemBlocks, A_EP, k, lds);
        for (int k = 0; k < A_NP; k++) m[k] = ordered_sum(np, ad.nodeBlocks, A_NP, k, lds);
        if (threadIdx.x != 0) continue;
        double dtBar = c->dtPrevBar[w], dtPrevBar = 0.0;
        if (hasNext) {
            // decay = (1 - h)/(1 + h), gain = dtAvg/(1 + h), h = damping dtAvg / 2
            const double dtAvg = 0.5 * (dtPrevNext + dtNext);
            const double h = 0.5 * d.damping * dtAvg;
            const double onePlus = 1.0 + h;
            const double dDecay = -2.0 / (onePlus * onePlus) * 0.5 * d.damping;
            const double dGain = 1.0 / onePlus - dtAvg / (onePlus * onePlus) * 0.5 * d.damping;
            const double avgBar = m[AN_DECAY] * dDecay + m[AN_GAIN] * dGain;
            dtBar += s[AP_DT] + m[AN_DT] + 0.5 * avgBar;
            dtPrevBar = 0.5 * avgBar;
            c->grad[w][SP_FRICTION] += m[AN_MU];
            c->grad[w][SP_A] += s[AP_A];
            c->grad[w][SP_B] += s[AP_B];
            c->grad[w][SP_N] += s[AP_N];
            c->grad[w][SP_C] += s[AP_C];
            c->grad[w][SP_M] += s[AP_M];
        }
        c->dtBar[w] = dtBar;
        c->dtPrevBar[w] = dtPrevBar;
        // dt_{n+1} = cfl sqrt(edge ms) / c: the edge's adjoint, for the element holding it
        double eb = 0.0;
        if (hasEdge && d.finite) {
            const double root = __builtin_sqrt(minEdgeN * d.massScale);
            eb = root > 0.0 ? dtBar * d.cfl / d.fastestWave * 0.5 / root * d.massScale : 0.0;
        }
        c->edgeBar[w] = eb;
    }
}

// The adjoint of the element pass of step n = (record j), one objective per half of the
// workgroup when two are carried.
__global__ void __launch_bounds__(BLOCK) k_adj_element(Dev d, ADev ad, int j) {
    __shared__ double sLog[97][3];
    __shared__ double sExp[64][2];
    __shared__ double ldsP[BLOCK / 64][A_EP];
    const nfm::Tables T = nfd::stage_tables(sLog, sExp);
    const ARec& a = ad.rec;
    const int per = BLOCK / ad.W;
    const int w = threadIdx.x / per;
    const int e = blockIdx.x * per + threadIdx.x % per;
    const size_t N = d.N, E = d.E;
    double part[A_EP] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (e < d.E) {
        const Mat& fl = d.mat0;
        const double mu = fl.mu;
        const int4 nn = d.conn[e];
        const int idx[4] = {nn.x, nn.y, nn.z, nn.w};
        EF f;
        element_forward(d, a, j, e, nn, fl, T, f);
        const double dr[4] = {f.d0r, f.d1r, -f.d0r, -f.d1r}, dz[4] = {f.d0z, f.d1z, -f.d0z, -f.d1z};
        // output adjoints: the force slots (the nodes' gathered-force adjoints), the new
        // state (what the next step's reverse left), and the shortest edge
        double gbr[4], gbz[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double2 fb = ad.fBar[(size_t)w * N + idx[k]];
            gbr[k] = fb.x;
            gbz[k] = fb.y;
        }
        double* sb = ad.sigBar + (size_t)w * 4 * E;
        double sbR = sb[e], sbZ = sb[E + e], sbT = sb[2 * E + e], sbRZ = sb[3 * E + e];
        double epB = ad.epBar[(size_t)w * E + e];
        double wkB = ad.workBar[(size_t)w * E + e];
        double* hb = ad.hgBar + (size_t)w * 2 * E;
        double qrB = hb[e], qzB = hb[E + e];
        double rB[4] = {0, 0, 0, 0}, zB[4] = {0, 0, 0, 0}, aB[4] = {0, 0, 0, 0}, cB[4] = {0, 0, 0, 0};
        double drB[4] = {0, 0, 0, 0}, dzB[4] = {0, 0, 0, 0};
        double dVB = 0.0, detB = 0.0, rgB = 0.0, dtB = 0.0;
        // hourglass (:1473-1497)
        double hgRB = qrB, hgZB = qzB;  // without hourglass control the forces pass through
        if (d.qhg > 0.0) {
            double yB[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                qrB -= f.y[k] * gbr[k];
                qzB -= f.y[k] * gbz[k];
                yB[k] = -f.qr * gbr[k] - f.qz * gbz[k];
            }
            hgRB = qrB;
            hgZB = qzB;
            const double kB = qrB * f.Sa + qzB * f.Sc;
            const double SaB = qrB * f.k, ScB = qzB * f.k;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                yB[k] += SaB * f.a[k] + ScB * f.c[k];
                aB[k] += SaB * f.y[k];
                cB[k] += ScB * f.y[k];
            }
            dVB += kB * d.qhg * mu * f.D2;
            const double D2B = kB * d.qhg * mu * f.dV;
            double hrB = 0.0, hzB = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                drB[k] += 2.0 * dr[k] * D2B - yB[k] * f.hr;
                dzB[k] += 2.0 * dz[k] * D2B - yB[k] * f.hz;
                hrB -= yB[k] * dr[k];
                hzB -= yB[k] * dz[k];
            }
            const double sgn[4] = {1.0, -1.0, 1.0, -1.0};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                rB[k] += sgn[k] * hrB;
                zB[k] += sgn[k] * hzB;
            }
        }
        // internal force (:1453-1471)
        double sRB = sbR, sZB = sbZ, sTB = sbT, sRZB = sbRZ;
        double hhB = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double gz = gbz[k], gr = gbr[k];
            dzB[k] -= gz * f.sZ * f.dV;
            sZB -= gz * dz[k] * f.dV;
            drB[k] -= gz * f.sRZ * f.dV;
            sRZB -= gz * dr[k] * f.dV;
            dVB -= gz * (dz[k] * f.sZ + dr[k] * f.sRZ);
            drB[k] -= gr * f.sR * f.dV;
            sRB -= gr * dr[k] * f.dV;
            dzB[k] -= gr * f.sRZ * f.dV;
            sRZB -= gr * dz[k] * f.dV;
            dVB -= gr * (dr[k] * f.sR + dz[k] * f.sRZ);
            hhB -= gr;
        }
        if (d.axi) {
            sTB += hhB * f.dV * 0.25 / f.rg;
            rgB -= hhB * f.sT * f.dV * 0.25 / (f.rg * f.rg);
            dVB += hhB * f.sT / f.rg * 0.25;
            rgB += dVB * 8.0 * NFD_PI * f.det;
            detB += dVB * 8.0 * NFD_PI * f.rg;
        } else {
            detB += dVB * 4.0 * d.thickness;
        }
        // the return map (solver/J2.java:163-296)
        double tRRB, tZZB, tTTB, tRZB;
        double epPrevB = epB, wkPrevB = wkB;
        double thetaB = 0.0;
        if (f.status == 0) {
            tRRB = sRB;
            tZZB = sZB;
            tTTB = sTB;
            tRZB = sRZB;
        } else {
            tRRB = sRB;
            tZZB = sZB;
            tTTB = sTB;
            tRZB = sRZB;
            double pullB = -(sRB * f.nRR + sZB * f.nZZ + sTB * f.nTT + sRZB * f.nRZ);
            const double nRRB = -f.pull * sRB, nZZB = -f.pull * sZB, nTTB = -f.pull * sTB, nRZB = -f.pull * sRZB;
            double dgB = wkB * (f.norm - f.pull) + NFD_SQRT_2_3 * epB;
            double normB = wkB * f.dGamma;
            pullB -= wkB * f.dGamma;
            dgB += 2.0 * mu * pullB;
            // n = x / norm
            double xRRB = nRRB * f.invNorm, xZZB = nZZB * f.invNorm, xTTB = nTTB * f.invNorm,
                   xRZB = nRZB * f.invNorm;
            const double invB = nRRB * f.xRR + nZZB * f.xZZ + nTTB * f.xTT + nRZB * f.xRZ;
            normB -= invB * f.invNorm * f.invNorm;
            // the multiplier's inputs: norm, epsP, thermal, dt, A, B, n, C
            double wB[CAP_DIRS];
            const double stiffness = 2.0 * mu + NFD_TWO_THIRDS * fl.hKin;
            if (f.status == 1) {
                // implicit function theorem at the root: dGamma' = -g_w / g_dGamma
                const double dEps = NFD_SQRT_2_3 * f.dGamma;
                const double eps = f.epsP + dEps;
                const double rate = a.stepDt[j] > 0.0 ? dEps / a.stepDt[j] : 0.0;
                double lx;
                const double pw = nfm::pow_log_ordinary(eps, fl.n, T, lx);
                const double S = fl.a + fl.b * pw;
                const bool clamped = rate <= fl.eps0;
                const double lr = clamped ? 0.0 : nfm::log_ordinary(fl.eps0 == 1.0 ? rate : rate / fl.eps0, T);
                const double R = clamped ? 1.0 : 1.0 + fl.c * lr;
                const double th = f.thermal;
                const double dSdeps = fl.b * pw * fl.n / eps;
                double slope = -stiffness - NFD_TWO_THIRDS * dSdeps * R * th;
                if (!clamped) slope -= NFD_SQRT_2_3 * S * fl.c * th / f.dGamma;
                const double k = -dgB / slope;  // times dg/dw
                wB[0] = k;
                wB[1] = k * (-NFD_SQRT_2_3 * R * th * dSdeps);
                wB[2] = k * (-NFD_SQRT_2_3 * S * R);
                wB[3] = clamped ? 0.0 : k * (NFD_SQRT_2_3 * S * th * fl.c / a.stepDt[j]);
                wB[4] = k * (-NFD_SQRT_2_3 * R * th);
                wB[5] = k * (-NFD_SQRT_2_3 * R * th * pw);
                wB[6] = k * (-NFD_SQRT_2_3 * R * th * fl.b * pw * lx);
                wB[7] = clamped ? 0.0 : k * (-NFD_SQRT_2_3 * S * th * lr);
            } else {
#pragma unroll
                for (int k = 0; k < CAP_DIRS; k++)
                    wB[k] = dgB * a.capJac[((size_t)(j - 1) * CAP_DIRS + k) * E + e];
            }
            normB += wB[0];
            epPrevB += wB[1];
            thetaB += wB[2];
            dtB += wB[3];
            part[AP_A] += wB[4];
            part[AP_B] += wB[5];
            part[AP_N] += wB[6];
            part[AP_C] += wB[7];
            // thermal = 1 - hom^m, hom = (T - T0)/(Tm - T0), T = T0 + beta W/(rho cp)
            if (f.thermalCase == 2) {
                const double homB = -thetaB * fl.m * f.homPow / f.hom;
                part[AP_M] += -thetaB * f.homPow * f.homLog;
                wkPrevB += homB / (fl.tmelt - fl.troom) * fl.thermalCoefficient;
            }
            // norm = sqrt(xRR^2 + xZZ^2 + xTT^2 + 2 xRZ^2)
            xRRB += normB * f.xRR / f.norm;
            xZZB += normB * f.xZZ / f.norm;
            xTTB += normB * f.xTT / f.norm;
            xRZB += normB * 2.0 * f.xRZ / f.norm;
            // x = t - p (deviator), p = (tRR + tZZ + tTT) / 3
            const double pB = -(xRRB + xZZB + xTTB);
            tRRB += xRRB + pB / 3.0;
            tZZB += xZZB + pB / 3.0;
            tTTB += xTTB + pB / 3.0;
            tRZB += xRZB;
        }
        if (f.status == 0) {
            // elastic: sigma = t directly (handled above), nothing else
        }
        // trial stress: t = sigma_rotated + lambda tr + 2 mu d
        const double ltB = tRRB + tZZB + tTTB;
        double dRB = 2.0 * mu * tRRB + fl.lambda * ltB;
        double dZB = 2.0 * mu * tZZB + fl.lambda * ltB;
        double dTB = 2.0 * mu * tTTB + fl.lambda * ltB;
        double dGB = mu * tRZB;
        double q0B = tRRB, q1B = tZZB, q3B = tRZB;
        double s2B = tTTB;
        // Cayley rotation (solver/Corotational.java:59-97)
        double s0B = q0B, s1B = q1B, s3B = q3B;
        double dZdRB = 0.0, dRdZB = 0.0;
        if (f.rotated) {
            s0B = f.cc * q0B + f.ss * q1B + f.csn * q3B;
            s1B = f.ss * q0B + f.cc * q1B - f.csn * q3B;
            s3B = -2.0 * f.csn * q0B + 2.0 * f.csn * q1B + (f.cc - f.ss) * q3B;
            const double ccB = q0B * f.s0 + q1B * f.s1 + q3B * f.s3;
            const double ssB = q0B * f.s1 + q1B * f.s0 - q3B * f.s3;
            const double csnB = -2.0 * q0B * f.s3 + 2.0 * q1B * f.s3 + q3B * (f.s0 - f.s1);
            const double csB = 2.0 * f.cs * ccB + f.sn * csnB;
            const double snB = 2.0 * f.sn * ssB + f.cs * csnB;
            const double hh2 = 1.0 + f.h * f.h;
            const double hB = csB * (-4.0 * f.h / (hh2 * hh2)) + snB * (2.0 * (1.0 - f.h * f.h) / (hh2 * hh2));
            const double spinB = 0.5 * hB;
            dZdRB += 0.5 * spinB;
            dRdZB -= 0.5 * spinB;
        }
        // strain increments
        dRdZB += dGB;
        dZdRB += dGB;
        if (d.axi) {
#pragma unroll
            for (int k = 0; k < 4; k++) aB[k] += 0.25 * dTB / f.rg;
            rgB -= dTB * f.dEpsT / f.rg;
#pragma unroll
            for (int k = 0; k < 4; k++) rB[k] += 0.25 * rgB;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            drB[k] += dRB * f.a[k] + dZdRB * f.c[k];
            aB[k] += dRB * dr[k] + dRdZB * dz[k];
            dzB[k] += dZB * f.c[k] + dRdZB * f.a[k];
            cB[k] += dZB * dz[k] + dZdRB * dr[k];
        }
        // B matrix (:1406-1421): d2 = -d0, d3 = -d1
        const double D0r = drB[0] - drB[2], D1r = drB[1] - drB[3];
        const double D0z = dzB[0] - dzB[2], D1z = dzB[1] - dzB[3];
        double J00B = 0.0, J01B = 0.0, J10B = 0.0, J11B = 0.0;
        J01B += 0.25 * f.inv * (D0r + D1r);
        J11B += 0.25 * f.inv * (D1r - D0r);
        J10B += 0.25 * f.inv * (D0z - D1z);
        J00B -= 0.25 * f.inv * (D0z + D1z);
        const double invB = 0.25 * (f.J01 - f.J11) * D0r + 0.25 * (f.J01 + f.J11) * D1r +
                            0.25 * (f.J10 - f.J00) * D0z - 0.25 * (f.J10 + f.J00) * D1z;
        detB -= invB * f.inv * f.inv;
        J00B += detB * f.J11;
        J11B += detB * f.J00;
        J01B -= detB * f.J10;
        J10B -= detB * f.J01;
        rB[0] += 0.25 * (-J00B - J10B);
        rB[1] += 0.25 * (J00B - J10B);
        rB[2] += 0.25 * (J00B + J10B);
        rB[3] += 0.25 * (-J00B + J10B);
        zB[0] += 0.25 * (-J01B - J11B);
        zB[1] += 0.25 * (J01B - J11B);
        zB[2] += 0.25 * (J01B + J11B);
        zB[3] += 0.25 * (-J01B + J11B);
        // the shortest edge (:1191-1197), for the element that set dt
        if (d.finite && e == a.argmin[j]) {
            const double eb = ad.ctl->edgeBar[w];
            // Math.min selection as the plain code takes it: the first argument on ties
            double es[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int kk = (k + 1) & 3;
                es[k] = nfd::sq(f.r[kk] - f.r[k], f.z[kk] - f.z[k]);
            }
            const double m01 = nfd::jmin(es[0], es[1]);
            const int i01 = nfm::bits_of(m01) == nfm::bits_of(es[0]) ? 0 : 1;
            const double m23 = nfd::jmin(es[2], es[3]);
            const int i23 = nfm::bits_of(m23) == nfm::bits_of(es[2]) ? 2 : 3;
            const double m = nfd::jmin(m01, m23);
            const int ks = nfm::bits_of(m) == nfm::bits_of(m01) ? i01 : i23;
            // edge ks runs from corner ks to corner ks + 1 (constant indices only, so the
            // corner arrays stay in registers)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (k != ks) continue;
                const int kk = (k + 1) & 3;
                const double drr = f.r[kk] - f.r[k], dzz = f.z[kk] - f.z[k];
                rB[kk] += 2.0 * drr * eb;
                rB[k] -= 2.0 * drr * eb;
                zB[kk] += 2.0 * dzz * eb;
                zB[k] -= 2.0 * dzz * eb;
            }
        }
        // mid-step coordinates and increments: r = x0 + u + mid v, a = v dt
        const double dt = a.stepDt[j];
        double midB = 0.0;
        double4* cb = ad.cornerBar + (size_t)w * 4 * E;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double2 v = a.v[j * N + idx[k]];
            double urB = 0.0, uzB = 0.0, vrB = aB[k] * dt, vzB = cB[k] * dt;
            dtB += aB[k] * v.x + cB[k] * v.y;
            if (d.finite) {
                urB = rB[k];
                uzB = zB[k];
                vrB += f.mid * rB[k];
                vzB += f.mid * zB[k];
                midB += rB[k] * v.x + zB[k] * v.y;
            }
            cb[4 * (size_t)e + k] = make_double4(urB, uzB, vrB, vzB);
        }
        dtB -= 0.5 * midB;
        part[AP_DT] += dtB;
        // the previous state's adjoint, in place
        sb[e] = s0B;
        sb[E + e] = s1B;
        sb[2 * E + e] = s2B;
        sb[3 * E + e] = s3B;
        ad.epBar[(size_t)w * E + e] = epPrevB;
        ad.workBar[(size_t)w * E + e] = wkPrevB;
        hb[e] = hgRB;
        hb[E + e] = hgZB;
    }
    // per-objective block partials (each objective owns whole waves)
#pragma unroll
    for (int k = 0; k < A_EP; k++) {
        double x = part[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((threadIdx.x & 63) == 0) ldsP[threadIdx.x >> 6][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < ad.W * A_EP) {
        const int ww = threadIdx.x / A_EP, k = threadIdx.x % A_EP;
        const int wavesPer = per / 64;
        double s = 0.0;
        for (int q = 0; q < wavesPer; q++) s += ldsP[ww * wavesPer + q][k];
        ad.elemPartial[((size_t)ww * ad.elemBlocks + blockIdx.x) * A_EP + k] = s;
    }
}

// The adjoint of the nodal update of step n = (record j): each node's corner adjoints in the
// gather's slot order, added to the adjoint carried from later steps, then the update
// reversed. jPrev = j - 1 holds the node state the update read.
template <bool WALL>
__global__ void __launch_bounds__(BLOCK) k_adj_node(Dev d, ADev ad, int j) {
    __shared__ double ldsP[BLOCK / 64][A_NP];
    const ARec& a = ad.rec;
    const int per = BLOCK / ad.W;
    const int w = threadIdx.x / per;
    const int i = blockIdx.x * per + threadIdx.x % per;
    const size_t N = d.N;
    double part[A_NP] = {0.0, 0.0, 0.0, 0.0};
    if (i < d.N) {
        // gathered corner adjoints, in the forward gather's order
        const double4* cb = ad.cornerBar + (size_t)w * 4 * d.E;
        double urB = 0.0, uzB = 0.0, vrB = 0.0, vzB = 0.0;
        for (int k = d.slotStart[i]; k < d.slotStart[i + 1]; k++) {
            const double4 c = cb[d.slotOf[k]];
            urB += c.x;
            uzB += c.y;
            vrB += c.z;
            vzB += c.w;
        }
        double2& uCarry = ad.uBar[(size_t)w * N + i];
        double2& vCarry = ad.vBar[(size_t)w * N + i];
        urB += uCarry.x;
        uzB += uCarry.y;
        vrB += vCarry.x;
        vzB += vCarry.y;
        // the update's inputs
        const double dt = a.stepDt[j], dtPrev = a.stepDtPrev[j];
        const double dtAvg = 0.5 * (dtPrev + dt);
        const double half = 0.5 * d.damping * dtAvg;
        const double decay = (1.0 - half) / (1.0 + half);
        const double gain = dtAvg / (1.0 + half);
        const double m = d.mass[i];
        const double invM = 1.0 / m;
        const uint8_t fl = a.flags[(size_t)(j - 1) * N + i];
        const double2 vi = a.v[(size_t)(j - 1) * N + i];
        const double2 ui = a.u[(size_t)(j - 1) * N + i];
        const double2 fi = a.f[(size_t)j * N + i];
        const bool resting = WALL && (fl & ON_WALL) && fi.y <= 0.0;
        double urPrevB = urB, uzPrevB = uzB, vrPrevB = 0.0, vzPrevB = 0.0;
        double frB = 0.0, fzB = 0.0;
        // radial (:1011-1034, solver/RigidWall.java:252-279)
        if (fl & FIXED_R) {
            // v = 0, u unchanged
        } else if (WALL && resting && d.wallFrictional) {
            const double normal = -fi.y;
            const double trial = -m * decay * vi.x / gain - fi.x;
            const double limit = __builtin_isinf(d.friction) ? HUGE_VAL : d.friction * normal;
            if (nfd::jabs(trial) <= limit) {
                // stuck: v = 0, u unchanged
            } else {
                const double tangential = __builtin_copysign(limit, trial);
                const double slid = decay * vi.x + gain * (fi.x + tangential) / m;
                const double vB = vrB + urB * dt;
                part[AN_DT] += urB * slid;
                vrPrevB = decay * vB;
                part[AN_DECAY] += vB * vi.x;
                part[AN_GAIN] += vB * (fi.x + tangential) / m;
                frB += gain / m * vB;
                const double tB = gain / m * vB;
                if (!__builtin_isinf(d.friction)) {
                    const double limitB = __builtin_signbit(trial) ? -tB : tB;
                    part[AN_MU] += limitB * normal;
                    fzB -= limitB * d.friction;
                }
            }
        } else {
            const double vr = decay * vi.x + gain * fi.x * invM;
            const double vB = vrB + urB * dt;
            part[AN_DT] += urB * vr;
            vrPrevB = decay * vB;
            part[AN_DECAY] += vB * vi.x;
            part[AN_GAIN] += vB * fi.x * invM;
            frB += gain * invM * vB;
        }
        // axial (:1036-1060, RigidWall hold / arrive)
        if (fl & FIXED_Z) {
            // v = 0, u unchanged
        } else if (WALL && resting) {
            vzPrevB = vzB;  // held: v and u carried over
        } else {
            const double trial = decay * vi.y + gain * fi.y * invM;
            const double zc = d.z0[i] + ui.y;
            if (WALL && zc + trial * dt < d.plane) {
                uzPrevB = 0.0;  // landed: u = plane - z0, v = 0
            } else {
                const double vB = vzB + uzB * dt;
                part[AN_DT] += uzB * trial;
                vzPrevB = decay * vB;
                part[AN_DECAY] += vB * vi.y;
                part[AN_GAIN] += vB * fi.y * invM;
                fzB += gain * invM * vB;
            }
        }
        uCarry = make_double2(urPrevB, uzPrevB);
        vCarry = make_double2(vrPrevB, vzPrevB);
        ad.fBar[(size_t)w * N + i] = make_double2(frB, fzB);
    }
#pragma unroll
    for (int k = 0; k < A_NP; k++) {
        double x = part[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((threadIdx.x & 63) == 0) ldsP[threadIdx.x >> 6][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < ad.W * A_NP) {
        const int ww = threadIdx.x / A_NP, k = threadIdx.x % A_NP;
        const int wavesPer = per / 64;
        double s = 0.0;
        for (int q = 0; q < wavesPer; q++) s += ldsP[ww * wavesPer + q][k];
        ad.nodePartial[((size_t)ww * ad.nodeBlocks + blockIdx.x) * A_NP + k] = s;
    }
}

}  // namespace nfa

// TEST INFRASTRUCTURE -- the oracle. Nothing in the product links, loads or calls this file.
//
// A CPU restatement of Neofiz's explicit step loop, written line by line from the Java
// reference (paths relative to src/main/java/org/neofiz/). It is what the HIP path is
// compared against, and -- through its threaded build -- the CPU baseline timed beside it.
// It exports the same C ABI as the product (include/neofiz.h) under the prefix nfo_.
//
// Fidelity rules followed throughout:
//  - every floating-point expression keeps Java's evaluation order and association
//    (Java evaluates left to right and never fuses a multiply-add; this file is built with
//    -ffp-contract=off and no fast-math);
//  - Math.min / Math.max keep Java's NaN and signed-zero semantics (jmin / jmax below);
//  - Math.pow / log / atan2 / sin / cos / hypot are the shared deterministic functions of
//    neofiz_amd/csrc/nf_math.h (HotSpot's are specified only to 1 ulp, so the last bit of
//    agreement with the JVM is not pinnable; agreement with the HIP path is);
//  - Math.sqrt and division are IEEE correctly rounded, as in Java.
//
// Parity pins for each part are the reference's own tests restated in tests/ (J2Test,
// RateDependentReturnMapTest, JohnsonCookTest, HourglassTest, FiniteStrainTest, GravityTest,
// DamageTest, ContactTest, ErosionTest, TaylorImpactTest, MeshConvergenceTest pinned values,
// ThreadingTest determinism).
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../include/neofiz.h"
#include "../neofiz_amd/csrc/nf_math.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

const double MAX_VALUE = 1.7976931348623157e308;  // Double.MAX_VALUE
const double PI = 3.141592653589793;              // Math.PI

// Math.min(double, double): NaN if either is NaN; -0.0 is less than +0.0.
inline double jmin(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return std::signbit(a) ? a : b;
    return a <= b ? a : b;
}

inline double jmax(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return std::signbit(a) ? b : a;
    return a >= b ? a : b;
}

inline double jabs(double a) { return std::fabs(a); }

// ------------------------------------------------------------------ JohnsonCook
// core/JohnsonCook.java:68-227
struct JohnsonCook {
    double a, b, n, c, m, eps0, troom, tmelt, cp, beta;

    double strainTerm(double e) const {  // :138-140
        return e <= 0.0 ? a : a + b * nfm::pow(e, n);
    }
    double strainSlopeFrom(double s, double e) const {  // :175-177
        return n * (s - a) / e;
    }
    double rateTerm(double rate) const {  // :180-183
        if (rate <= eps0) return 1.0;
        return 1.0 + c * nfm::log(rate / eps0);
    }
    bool rateIsClamped(double rate) const { return rate <= eps0; }  // :186-188
    double thermalTerm(double t) const {  // :191-197
        const double homologous = (t - troom) / (tmelt - troom);
        if (homologous <= 0.0) return 1.0;
        if (homologous >= 1.0) return 0.0;
        return 1.0 - nfm::pow(homologous, m);
    }
    double quasiStaticFlowStress(double e, double t) const {  // :207-209
        return strainTerm(e) * thermalTerm(t);
    }
    double temperature(double w, double rho) const {  // :218-221
        return troom + beta * w / (rho * cp);
    }
    double thermalCoefficient(double rho) const { return beta / (rho * cp); }  // :224-226
};

// ------------------------------------------------------------------ Material
// core/Material.java:17-136
struct Material {
    double E, nu, rho, sy, hiso, hkin;
    bool has_jc;
    JohnsonCook jc;

    bool isElastic() const { return std::isinf(sy); }
    double lambda() const { return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)); }  // :93-95
    double mu() const { return E / (2.0 * (1.0 + nu)); }                       // :98-100
    double dilatationalWaveSpeed() const {                                     // :106-108
        return std::sqrt((lambda() + 2.0 * mu()) / rho);
    }
};

int check_material(const nf_material* m) {
    if (!m) return fail(NF_E_ARGUMENT, "material must not be null");
    // core/Material.java:21-37
    if (m->youngs_modulus <= 0) return fail(NF_E_ARGUMENT, "E must be positive");
    if (m->poisson_ratio <= -1.0 || m->poisson_ratio >= 0.5)
        return fail(NF_E_ARGUMENT, "nu must lie in (-1, 0.5)");
    if (m->density <= 0) return fail(NF_E_ARGUMENT, "density must be positive");
    if (m->yield_stress <= 0) return fail(NF_E_ARGUMENT, "yield stress must be positive");
    if (m->isotropic_hardening < 0 || m->kinematic_hardening < 0)
        return fail(NF_E_ARGUMENT, "hardening moduli must be non-negative");
    if (m->has_johnson_cook) {
        const nf_johnson_cook& j = m->johnson_cook;  // core/JohnsonCook.java:73-89
        if (j.a <= 0) return fail(NF_E_ARGUMENT, "A must be positive");
        if (j.b < 0) return fail(NF_E_ARGUMENT, "B must be non-negative");
        if (j.n <= 0 || j.n > 1) return fail(NF_E_ARGUMENT, "n must lie in (0, 1]");
        if (j.c < 0) return fail(NF_E_ARGUMENT, "C must be non-negative");
        if (j.m <= 0) return fail(NF_E_ARGUMENT, "m must be positive");
        if (j.reference_strain_rate <= 0)
            return fail(NF_E_ARGUMENT, "reference strain rate must be positive");
        if (j.melt_temperature <= j.room_temperature)
            return fail(NF_E_ARGUMENT, "melt temperature must exceed room temperature");
        if (j.specific_heat <= 0) return fail(NF_E_ARGUMENT, "specific heat must be positive");
        if (j.taylor_quinney < 0 || j.taylor_quinney > 1)
            return fail(NF_E_ARGUMENT, "Taylor-Quinney fraction must lie in [0, 1]");
        if (m->yield_stress != j.a)
            return fail(NF_E_ARGUMENT, "yield stress disagrees with the Johnson-Cook A");
    }
    return NF_OK;
}

Material to_material(const nf_material& m) {
    Material out;
    out.E = m.youngs_modulus;
    out.nu = m.poisson_ratio;
    out.rho = m.density;
    out.sy = m.yield_stress;
    out.hiso = m.isotropic_hardening;
    out.hkin = m.kinematic_hardening;
    out.has_jc = m.has_johnson_cook != 0;
    const nf_johnson_cook& j = m.johnson_cook;
    out.jc = {j.a, j.b, j.n, j.c, j.m, j.reference_strain_rate, j.room_temperature,
              j.melt_temperature, j.specific_heat, j.taylor_quinney};
    return out;
}

// J2.Flow solver/J2.java:105-113
struct Flow {
    double sy0, hIso, hKin;
    const JohnsonCook* jc;  // null when linear
    double thermalCoefficient;
};

// ------------------------------------------------------------------ Damage
// solver/Damage.java:156-433
struct Damage {
    static const int GEOMETRY_STRIDE = 5;
    static constexpr double ISOTROPIC = 1.0e-9;
    std::vector<double> fractureEnergy, youngsModulus, failureStrain, geometry;
    std::vector<double> strength, softening, width;

    bool hasBegun(int p) const { return strength[p] > 0.0; }  // :275-277

    void begin(int p, double flowStress, double nRR, double nZZ, double nRZ) {  // :306-336
        const int g = GEOMETRY_STRIDE * p;
        const double area = geometry[g];
        const double diff = nRR - nZZ;
        const double cross = 2.0 * nRZ;
        const double anisotropy = nfm::hypot(diff, cross);
        double h;
        if (anisotropy < ISOTROPIC) {
            h = std::sqrt(area);
        } else {
            const double phi = 0.5 * nfm::atan2(cross, diff);
            const double tR = -nfm::sin(phi), tZ = nfm::cos(phi);
            const double line = jabs(geometry[g + 1] * tR + geometry[g + 2] * tZ) +
                                jabs(geometry[g + 3] * tR + geometry[g + 4] * tZ);
            h = area / line;
        }
        width[p] = h;
        strength[p] = flowStress;
        const double slope = -flowStress * flowStress * h / (2.0 * fractureEnergy[p]);
        const double floor = -youngsModulus[p];
        softening[p] = slope < floor ? floor : slope;
    }

    double at(int p, double plasticStrain) const {  // :342-347
        const double sf = strength[p];
        if (sf <= 0.0) return 0.0;
        const double d = -softening[p] * (plasticStrain - failureStrain[p]) / sf;
        return d <= 0.0 ? 0.0 : (d >= 1.0 ? 1.0 : d);
    }

    double dissipation(int p, double plasticStrain) const {  // :363-372
        const double sf = strength[p];
        if (sf <= 0.0) return 0.0;
        const double slope = softening[p];
        const double full = -sf / slope;
        double excess = plasticStrain - failureStrain[p];
        if (excess <= 0.0) return 0.0;
        if (excess > full) excess = full;
        return excess * (sf + 0.5 * slope * excess);
    }
};

// ------------------------------------------------------------------ J2
// solver/J2.java:77-407
const double SQRT_2_3 = std::sqrt(2.0 / 3.0);  // :82
const double TWO_THIRDS = 2.0 / 3.0;           // :83
const double SOLVE_TOLERANCE = 1.0e-10;        // :86
const int SOLVE_ITERATIONS = 25;               // :88

double j2_solve(double norm, double stiffness, double plasticStrain, double temperature,
                double dt, const JohnsonCook& jc) {  // :332-375
    const double thermal = jc.thermalTerm(temperature);
    const double tolerance = SOLVE_TOLERANCE * jc.a;
    double lo = 0.0;
    double hi = norm / stiffness;
    double dGamma = hi;
    for (int iteration = 0; iteration < SOLVE_ITERATIONS; iteration++) {
        const double dEps = SQRT_2_3 * dGamma;
        const double eps = plasticStrain + dEps;
        const double rate = dt > 0.0 ? dEps / dt : 0.0;
        const double strain = jc.strainTerm(eps);
        const double rateFactor = jc.rateTerm(rate);
        const double g = norm - stiffness * dGamma - SQRT_2_3 * strain * rateFactor * thermal;
        if (g > 0.0) lo = dGamma;
        else hi = dGamma;
        if (jabs(g) <= tolerance) return dGamma;
        double slope = -stiffness - TWO_THIRDS * jc.strainSlopeFrom(strain, eps) * rateFactor * thermal;
        if (!jc.rateIsClamped(rate)) {
            slope -= SQRT_2_3 * strain * jc.c * thermal / dGamma;
        }
        double next = dGamma - g / slope;
        if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
        dGamma = next;
    }
    return dGamma;
}

double j2_update(double* sig, double* back, double* epsP, double* work, int e, double dRR,
                 double dZZ, double dTT, double dGamRZ, double lambda, double mu,
                 const Flow& flow, double strainDt, Damage* damage) {  // :163-296
    const int s = 4 * e;
    const double lt = lambda * (dRR + dZZ + dTT);
    const double tRR = sig[s] + lt + 2.0 * mu * dRR;
    const double tZZ = sig[s + 1] + lt + 2.0 * mu * dZZ;
    const double tTT = sig[s + 2] + lt + 2.0 * mu * dTT;
    const double tRZ = sig[s + 3] + mu * dGamRZ;

    const double p = (tRR + tZZ + tTT) / 3.0;
    const double xRR = tRR - p - back[s];
    const double xZZ = tZZ - p - back[s + 1];
    const double xTT = tTT - p - back[s + 2];
    const double xRZ = tRZ - back[s + 3];

    const double norm = std::sqrt(xRR * xRR + xZZ * xZZ + xTT * xTT + 2.0 * xRZ * xRZ);

    const JohnsonCook* jc = flow.jc;
    const double temperature = jc == nullptr ? 0.0 : jc->troom + flow.thermalCoefficient * work[e];

    const bool softening = damage != nullptr && damage->hasBegun(e);
    double hardening = 0.0;
    double flowStress;
    if (softening) {
        flowStress = damage->strength[e] + damage->softening[e] * (epsP[e] - damage->failureStrain[e]);
        if (flowStress > 0.0) {
            hardening = damage->softening[e];
        } else {
            flowStress = 0.0;
        }
    } else if (jc == nullptr) {
        flowStress = flow.sy0 + flow.hIso * epsP[e];
        hardening = flow.hIso;
    } else {
        flowStress = jc->quasiStaticFlowStress(epsP[e], temperature);
    }
    const double radius = SQRT_2_3 * flowStress;

    if (!(norm > radius)) {
        sig[s] = tRR;
        sig[s + 1] = tZZ;
        sig[s + 2] = tTT;
        sig[s + 3] = tRZ;
        return 0.0;
    }

    const double stiffness = 2.0 * mu + TWO_THIRDS * flow.hKin;
    double dGamma;
    if (jc == nullptr || softening) {
        dGamma = (norm - radius) / (stiffness + TWO_THIRDS * hardening);
        if (softening && flowStress + hardening * SQRT_2_3 * dGamma < 0.0) {
            dGamma = norm / stiffness;
        }
    } else {
        dGamma = j2_solve(norm, stiffness, epsP[e], temperature, strainDt, *jc);
    }
    const double inv = 1.0 / norm;
    const double nRR = xRR * inv, nZZ = xZZ * inv, nTT = xTT * inv, nRZ = xRZ * inv;

    const double pull = 2.0 * mu * dGamma;
    sig[s] = tRR - pull * nRR;
    sig[s + 1] = tZZ - pull * nZZ;
    sig[s + 2] = tTT - pull * nTT;
    sig[s + 3] = tRZ - pull * nRZ;

    const double alphaDotN = back[s] * nRR + back[s + 1] * nZZ + back[s + 2] * nTT + 2.0 * back[s + 3] * nRZ;
    work[e] += dGamma * (norm - pull + alphaDotN);

    const double shift = TWO_THIRDS * flow.hKin * dGamma;
    back[s] += shift * nRR;
    back[s + 1] += shift * nZZ;
    back[s + 2] += shift * nTT;
    back[s + 3] += shift * nRZ;

    epsP[e] += SQRT_2_3 * dGamma;

    if (damage != nullptr && !softening && epsP[e] >= damage->failureStrain[e]) {
        damage->begin(e, jc == nullptr ? flow.sy0 + flow.hIso * epsP[e]
                                       : jc->quasiStaticFlowStress(epsP[e], temperature),
                      nRR, nZZ, nRZ);
    }
    return dGamma;
}

// ------------------------------------------------------------------ Corotational
// solver/Corotational.java:59-97
inline void corot_rotate(double* t, int index, double cos, double sin) {
    const int s = 4 * index;
    const double rr = t[s], zz = t[s + 1], rz = t[s + 3];
    const double cc = cos * cos, ss = sin * sin, cs = cos * sin;
    t[s] = cc * rr - 2.0 * cs * rz + ss * zz;
    t[s + 1] = ss * rr + 2.0 * cs * rz + cc * zz;
    t[s + 3] = cs * (rr - zz) + (cc - ss) * rz;
}
inline double corot_cos(double spinDt) {
    const double h = 0.5 * spinDt;
    return (1.0 - h * h) / (1.0 + h * h);
}
inline double corot_sin(double spinDt) {
    const double h = 0.5 * spinDt;
    return 2.0 * h / (1.0 + h * h);
}

// ------------------------------------------------------------------ RigidWall
// solver/RigidWall.java:95-316
struct RigidWall {
    double plane = 0.0, friction = 0.0;
    double impulse = 0.0, energyLoss = 0.0, frictionDissipation = 0.0;
    int contactCount = 0, stuckCount = 0;

    bool hasFriction() const { return friction > 0.0; }
    void beginStep() {
        contactCount = 0;
        stuckCount = 0;
    }
    bool holds(double forceZ) const { return forceZ <= 0.0; }
    void hold(double forceZ, double dtEffective) {
        impulse += -forceZ * dtEffective;
        contactCount++;
    }
    bool crosses(double zCurrent, double velocityZ, double dt) const {
        return zCurrent + velocityZ * dt < plane;
    }
    void arrive(double trialVelocity, double nodeMass) {
        impulse += -nodeMass * trialVelocity;
        energyLoss += 0.5 * nodeMass * trialVelocity * trialVelocity;
        contactCount++;
    }
    double slide(double forceR, double forceZ, double velocityR, double decay, double gain,
                 double nodeMass, double dt) {
        const double normal = -forceZ;
        const double trial = -nodeMass * decay * velocityR / gain - forceR;
        const double limit = std::isinf(friction) ? HUGE_VAL : friction * normal;
        if (jabs(trial) <= limit) {
            energyLoss += 0.5 * nodeMass * velocityR * velocityR;
            stuckCount++;
            return 0.0;
        }
        const double tangential = std::copysign(limit, trial);
        const double slid = decay * velocityR + gain * (forceR + tangential) / nodeMass;
        frictionDissipation += -tangential * slid * dt;
        return slid;
    }
};

// ------------------------------------------------------------------ QuadMesh helpers
struct Mesh {
    int nodeCount = 0, elementCount = 0;
    std::vector<double> r, z;
    std::vector<int> conn, pressureEdges;

    double signedArea(int e) const {  // mesh/QuadMesh.java:78-87
        const int b = e * 4;
        double a = 0.0;
        for (int k = 0; k < 4; k++) {
            const int i = conn[b + k];
            const int j = conn[b + (k + 1) % 4];
            a += r[i] * z[j] - r[j] * z[i];
        }
        return 0.5 * a;
    }
    double minimumEdgeLength() const {  // :61-75
        double mn = MAX_VALUE;
        for (int e = 0; e < elementCount; e++) {
            const int b = e * 4;
            for (int k = 0; k < 4; k++) {
                const int i = conn[b + k];
                const int j = conn[b + (k + 1) % 4];
                const double dr = r[j] - r[i];
                const double dz = z[j] - z[i];
                const double len = std::sqrt(dr * dr + dz * dz);
                if (len < mn) mn = len;
            }
        }
        return mn;
    }
    // surface(skip) :188-226. Emission in element order; counting by edge key.
    void surface(const std::vector<uint8_t>& skip, std::vector<int>& edges,
                 std::vector<int>& owner, std::vector<int>& nodes) const {
        std::unordered_map<uint64_t, int> count;
        count.reserve((size_t)elementCount * 4);
        auto key = [](int a, int b) {
            const int lo = a < b ? a : b, hi = a < b ? b : a;
            return ((uint64_t)(uint32_t)lo << 32) | (uint32_t)hi;
        };
        for (int e = 0; e < elementCount; e++) {
            if (!skip.empty() && skip[e]) continue;
            for (int k = 0; k < 4; k++) count[key(conn[e * 4 + k], conn[e * 4 + (k + 1) % 4])]++;
        }
        edges.clear();
        owner.clear();
        std::vector<uint8_t> on(nodeCount, 0);
        for (int e = 0; e < elementCount; e++) {
            if (!skip.empty() && skip[e]) continue;
            for (int k = 0; k < 4; k++) {
                const int a = conn[e * 4 + k];
                const int b = conn[e * 4 + (k + 1) % 4];
                if (count[key(a, b)] != 1) continue;
                edges.push_back(a);
                edges.push_back(b);
                owner.push_back(e);
                on[a] = 1;
                on[b] = 1;
            }
        }
        nodes.clear();
        for (int i = 0; i < nodeCount; i++)
            if (on[i]) nodes.push_back(i);
    }
};

// ------------------------------------------------------------------ Contact
// solver/Contact.java:132-763
struct Contact {
    static constexpr double MAX_DEPTH = 0.5;         // :153
    static constexpr double STICK_RELAXATION = 0.5;  // :171
    static constexpr double FACING = 0.25;           // :195

    const Mesh* mesh = nullptr;
    std::vector<uint8_t> dead;
    bool stale = false;
    std::vector<int> seg, owner, nodes, segStart, segOf, slot;
    double friction = 0.0, scale = 0.1, damping = 0.0;
    double step = 0.0, reference = 0.0;
    std::vector<double> cr, cz, nr, nz, length, pnr, pnz;
    std::vector<int> lastMaster, master;
    std::vector<uint8_t> dropped;
    double cellSize = 0.0, gridR0 = 0.0, gridZ0 = 0.0;
    int gridCols = 0, gridRows = 0, buckets = 0;
    std::vector<int> bucketStart, bucketOf;
    int pairs = 0;
    double energy = 0.0, dissipation = 0.0, maxPenetration = 0.0;
    int64_t escaped = 0;

    void init(const Mesh* m) {  // :273-280
        mesh = m;
        dead.assign(m->elementCount, 0);
        cr.assign(m->nodeCount, 0.0);
        cz.assign(m->nodeCount, 0.0);
        slot.assign(m->nodeCount, -1);
        adopt();
    }

    int segments() const { return (int)owner.size(); }

    void adopt() {  // :290-322
        mesh->surface(dead, seg, owner, nodes);
        std::fill(slot.begin(), slot.end(), -1);
        for (int i = 0; i < (int)nodes.size(); i++) slot[nodes[i]] = i;
        const int S = segments();
        const int nn = (int)nodes.size();
        segStart.assign(nn + 1, 0);
        for (int s = 0; s < S; s++) {
            segStart[slot[seg[2 * s]] + 1]++;
            segStart[slot[seg[2 * s + 1]] + 1]++;
        }
        for (int i = 0; i < nn; i++) segStart[i + 1] += segStart[i];
        segOf.assign(segStart[nn], 0);
        std::vector<int> fill = segStart;
        for (int s = 0; s < S; s++) {
            segOf[fill[slot[seg[2 * s]]]++] = s;
            segOf[fill[slot[seg[2 * s + 1]]]++] = s;
        }
        nr.assign(S, 0.0);
        nz.assign(S, 0.0);
        length.assign(S, 0.0);
        pnr.assign(nn, 0.0);
        pnz.assign(nn, 0.0);
        lastMaster.assign(nn, -1);
        master.assign(nn, 0);
        dropped.assign(nn, 0);
        stale = false;
    }

    void forget(int element) {  // :329-332
        dead[element] = 1;
        stale = true;
    }

    int begin(const double* ur, const double* uz, double dt, double referenceStep) {  // :424-436
        step = dt;
        reference = referenceStep;
        if (stale) adopt();
        pairs = 0;
        energy = 0.0;
        maxPenetration = 0.0;
        if (segments() == 0) return 0;
        geometry(ur, uz);
        grid();
        return (int)nodes.size();
    }

    void search(int from, int to) {  // :439-443
        for (int i = from; i < to; i++) master[i] = nearestPenetrated(i, nodes[i]);
    }

    void accumulate(const double* vr, const double* vz, const double* mass, double* fr,
                    double* fz) {  // :446-462
        for (int i = 0; i < (int)nodes.size(); i++) {
            const int n = nodes[i];
            const int best = master[i];
            if (best < 0) {
                if (lastMaster[i] >= 0 && dropped[i]) escaped++;
                lastMaster[i] = -1;
                continue;
            }
            lastMaster[i] = best;
            resolve(i, n, best, vr, vz, mass, fr, fz);
        }
    }

    void geometry(const double* ur, const double* uz) {  // :465-500
        for (int n : nodes) {
            cr[n] = mesh->r[n] + ur[n];
            cz[n] = mesh->z[n] + uz[n];
        }
        for (int s = 0; s < segments(); s++) {
            const int a = seg[2 * s], b = seg[2 * s + 1];
            const double dr = cr[b] - cr[a], dz = cz[b] - cz[a];
            const double len = std::sqrt(dr * dr + dz * dz);
            length[s] = len;
            if (len > 0.0) {
                nr[s] = dz / len;
                nz[s] = -dr / len;
            } else {
                nr[s] = 0.0;
                nz[s] = 0.0;
            }
        }
        for (int i = 0; i < (int)nodes.size(); i++) {
            double sr = 0.0, sz = 0.0;
            for (int k = segStart[i]; k < segStart[i + 1]; k++) {
                sr += nr[segOf[k]];
                sz += nz[segOf[k]];
            }
            const double len = std::sqrt(sr * sr + sz * sz);
            pnr[i] = len > 0.0 ? sr / len : 0.0;
            pnz[i] = len > 0.0 ? sz / len : 0.0;
        }
    }

    static int clamp(int v, int n) { return v < 0 ? 0 : v >= n ? n - 1 : v; }
    // Java's (int) cast of a double: truncation toward zero, saturating, NaN to 0.
    static int jint(double x) {
        if (x != x) return 0;
        if (x >= 2147483647.0) return 2147483647;
        if (x <= -2147483648.0) return (-2147483647 - 1);
        return (int)x;
    }
    int cellOf(double r, double z) const {
        const int i = clamp(jint((r - gridR0) / cellSize), gridCols);
        const int j = clamp(jint((z - gridZ0) / cellSize), gridRows);
        return j * gridCols + i;
    }

    void grid() {  // :515-551
        const int S = segments();
        double rlo = MAX_VALUE, zlo = MAX_VALUE;
        double rhi = -MAX_VALUE, zhi = -MAX_VALUE;
        double longest = 0.0;
        for (int n : nodes) {
            if (cr[n] < rlo) rlo = cr[n];
            if (cr[n] > rhi) rhi = cr[n];
            if (cz[n] < zlo) zlo = cz[n];
            if (cz[n] > zhi) zhi = cz[n];
        }
        for (int s = 0; s < S; s++)
            if (length[s] > longest) longest = length[s];
        cellSize = longest > 0.0 ? 1.5 * longest : 1.0;
        gridR0 = rlo;
        gridZ0 = zlo;
        gridCols = std::max(1, jint((rhi - rlo) / cellSize) + 1);
        gridRows = std::max(1, jint((zhi - zlo) / cellSize) + 1);
        buckets = gridCols * gridRows;
        bucketStart.assign(buckets + 1, 0);
        bucketOf.assign(2 * S, 0);
        for (int s = 0; s < S; s++) {
            bucketStart[cellOf(cr[seg[2 * s]], cz[seg[2 * s]]) + 1]++;
            bucketStart[cellOf(cr[seg[2 * s + 1]], cz[seg[2 * s + 1]]) + 1]++;
        }
        for (int c = 0; c < buckets; c++) bucketStart[c + 1] += bucketStart[c];
        std::vector<int> fill(bucketStart.begin(), bucketStart.begin() + buckets + 1);
        for (int s = 0; s < S; s++) {
            bucketOf[fill[cellOf(cr[seg[2 * s]], cz[seg[2 * s]])]++] = s;
            bucketOf[fill[cellOf(cr[seg[2 * s + 1]], cz[seg[2 * s + 1]])]++] = s;
        }
    }

    int nearestPenetrated(int i, int n) {  // :570-606
        const int ci = clamp(jint((cr[n] - gridR0) / cellSize), gridCols);
        const int cj = clamp(jint((cz[n] - gridZ0) / cellSize), gridRows);
        int best = -1;
        double deepest = 0.0;
        bool fell = false;
        for (int dj = -1; dj <= 1; dj++) {
            const int j = cj + dj;
            if (j < 0 || j >= gridRows) continue;
            for (int di = -1; di <= 1; di++) {
                const int ii = ci + di;
                if (ii < 0 || ii >= gridCols) continue;
                const int cell = j * gridCols + ii;
                for (int k = bucketStart[cell]; k < bucketStart[cell + 1]; k++) {
                    const int s = bucketOf[k];
                    if (excluded(i, n, s)) continue;
                    const double d = depth(n, s);
                    if (d <= 0.0) continue;
                    if (d > MAX_DEPTH * length[s]) {
                        if (s == lastMaster[i]) fell = true;
                        continue;
                    }
                    if (d > deepest) {
                        deepest = d;
                        best = s;
                    }
                }
            }
        }
        dropped[i] = fell;
        return best;
    }

    bool excluded(int i, int n, int s) const {  // :617-636
        const int a = seg[2 * s], b = seg[2 * s + 1];
        if (pnr[i] * nr[s] + pnz[i] * nz[s] > -FACING) return true;
        if (a == n || b == n) return true;
        for (int k = segStart[i]; k < segStart[i + 1]; k++) {
            const int own = segOf[k];
            const int p = seg[2 * own], q = seg[2 * own + 1];
            if (p == a || p == b || q == a || q == b) return true;
        }
        return false;
    }

    double depth(int n, int s) const {  // :639-649
        const int a = seg[2 * s], b = seg[2 * s + 1];
        const double len = length[s];
        if (len <= 0.0) return 0.0;
        const double tr = (cr[b] - cr[a]) / len, tz = (cz[b] - cz[a]) / len;
        const double dr = cr[n] - cr[a], dz = cz[n] - cz[a];
        const double along = dr * tr + dz * tz;
        if (along < 0.0 || along > len) return 0.0;
        const double gap = dr * nr[s] + dz * nz[s];
        return gap < 0.0 ? -gap : 0.0;
    }

    double parameter(int n, int s) const {  // :652-658
        const int a = seg[2 * s];
        const double len = length[s];
        const double tr = (cr[seg[2 * s + 1]] - cr[a]) / len;
        const double tz = (cz[seg[2 * s + 1]] - cz[a]) / len;
        return ((cr[n] - cr[a]) * tr + (cz[n] - cz[a]) * tz) / len;
    }

    void resolve(int i, int n, int s, const double* vr, const double* vz, const double* mass,
                 double* fr, double* fz) {  // :660-762
        (void)i;
        const double dt = step;
        const int a = seg[2 * s], b = seg[2 * s + 1];
        const double d = depth(n, s);
        const double t = parameter(n, s);
        const double wa = 1.0 - t, wb = t;
        const double invMaster = 0.25 / mass[a] + 0.25 / mass[b];
        if (!(invMaster > 0.0)) return;
        const double mMaster = 1.0 / invMaster;
        const double mSlave = mass[n];
        const double mStar = mSlave * mMaster / (mSlave + mMaster);
        const double k = scale * mStar / (reference * reference);
        const double nR = nr[s], nZ = nz[s];
        const double mTrue = 1.0 / (wa * wa / mass[a] + wb * wb / mass[b]);
        const double mTangent = mSlave * mTrue / (mSlave + mTrue);
        const double closing = (vr[n] - wa * vr[a] - wb * vr[b]) * nR + (vz[n] - wa * vz[a] - wb * vz[b]) * nZ;
        const double damper = 2.0 * damping * mStar * std::sqrt(scale) / reference;
        const double normal = jmax(0.0, k * d - damper * closing);

        fr[n] += normal * nR;
        fz[n] += normal * nZ;
        fr[a] -= wa * normal * nR;
        fz[a] -= wa * normal * nZ;
        fr[b] -= wb * normal * nR;
        fz[b] -= wb * normal * nZ;

        pairs++;
        energy += 0.5 * k * d * d;
        if (d > maxPenetration) maxPenetration = d;
        const double damperForce = normal - k * d;
        if (damperForce != 0.0) {
            const double closingAfter = closing + damperForce * dt / mTangent;
            dissipation -= damperForce * 0.5 * (closing + closingAfter) * dt;
        }
        if (friction <= 0.0) return;

        const double tR = -nZ, tZ = nR;
        const double slide = (vr[n] - wa * vr[a] - wb * vr[b]) * tR + (vz[n] - wa * vz[a] - wb * vz[b]) * tZ;
        double tangent = -STICK_RELAXATION * mTangent * slide / dt;
        const double cone = friction * normal;
        if (tangent > cone) tangent = cone;
        else if (tangent < -cone) tangent = -cone;

        fr[n] += tangent * tR;
        fz[n] += tangent * tZ;
        fr[a] -= wa * tangent * tR;
        fz[a] -= wa * tangent * tZ;
        fr[b] -= wb * tangent * tR;
        fz[b] -= wb * tangent * tZ;

        const double after = slide + tangent * dt / mTangent;
        dissipation += -tangent * 0.5 * (slide + after) * dt;
    }
};

// ------------------------------------------------------------------ Parallel
// solver/Parallel.java:68-203: a spin pool handing out chunks from a shared cursor, 8 chunks
// per thread, minimum grain 16. The calling thread takes chunks too.
class Pool {
  public:
    using Kernel = void (*)(void* ctx, int from, int to, int worker);
    explicit Pool(int threads) : threads_(threads) {
        for (int w = 1; w < threads; w++) workers_.emplace_back([this, w] { loop(w); });
    }
    ~Pool() {
        shutdown_.store(true, std::memory_order_release);
        ticket_.fetch_add(1, std::memory_order_acq_rel);
        for (auto& t : workers_) t.join();
    }
    int threads() const { return threads_; }
    void run(int count, Kernel k, void* ctx) {
        if (threads_ == 1) {
            k(ctx, 0, count, 0);
            return;
        }
        kernel_ = k;
        ctx_ = ctx;
        count_ = count;
        grain_ = std::max(16, count / (threads_ * 8));
        cursor_.store(0, std::memory_order_relaxed);
        completed_.store(0, std::memory_order_relaxed);
        ticket_.fetch_add(1, std::memory_order_release);
        drain(0);
        int spins = 0;
        while (completed_.load(std::memory_order_acquire) < threads_ - 1) wait(spins);
    }

  private:
    void drain(int worker) {
        while (true) {
            const int from = cursor_.fetch_add(grain_, std::memory_order_relaxed);
            if (from >= count_) return;
            kernel_(ctx_, from, std::min(from + grain_, count_), worker);
        }
    }
    void loop(int worker) {
        int seen = 0;
        int spins = 0;
        while (true) {
            const int now = ticket_.load(std::memory_order_acquire);
            if (now == seen) {
                wait(spins);
                continue;
            }
            seen = now;
            spins = 0;
            if (shutdown_.load(std::memory_order_acquire)) return;
            drain(worker);
            completed_.fetch_add(1, std::memory_order_release);
        }
    }
    static void wait(int& spins) {
        if (spins < 2000) {
            spins++;
        } else {
            std::this_thread::yield();
        }
    }
    int threads_;
    std::vector<std::thread> workers_;
    std::atomic<int> ticket_{0}, completed_{0}, cursor_{0};
    std::atomic<bool> shutdown_{false};
    Kernel kernel_ = nullptr;
    void* ctx_ = nullptr;
    int count_ = 0, grain_ = 16;
};

// ------------------------------------------------------------------ ExplicitSolver
// solver/ExplicitSolver.java:38-1966
const double G = 0.5773502691896258;  // :41
const double GP_XI[4] = {-G, G, G, -G};
const double GP_ETA[4] = {-G, -G, G, G};
const double W_REDUCED = 4.0;                          // :50
const double H[4] = {1.0, -1.0, 1.0, -1.0};           // :56
const double DEFAULT_HOURGLASS_COEFFICIENT = 0.05;    // :59
const int PAD = 8;                                     // :175

}  // namespace

struct nfo_solver {
    Mesh mesh;
    std::vector<Material> materials;  // distinct
    std::vector<int> materialOf;      // empty when uniform
    int formulation, integration, kinematics;
    double thickness, cflSafety;

    std::vector<double> lambdaOf, muOf, bulkOf, densityOf;
    std::vector<Flow> flowOf;
    double fastestWave;
    double dt0;
    Damage* damage = nullptr;
    Damage damageStore;

    std::vector<double> ur, uz, vr, vz, fr, fz, mass;
    std::vector<uint8_t> fixedR, fixedZ, onWall;
    int pointsPerElement;
    std::vector<double> sig, back, epsP, wPlastic, hgR, hgZ, hourglassStiffness;
    std::vector<double> contribution;
    int edgeSlotBase;
    std::vector<int> nodeSlotStart, nodeSlot;
    std::vector<double> minEdgeLocal;
    Pool* pool = nullptr;

    double kernelStrainDt = 0, kernelPressure = 0, nodalDecay = 0, nodalGain = 0, nodalDtAvg = 0;
    double dt, dtPrevious, minEdgeSquared, time = 0.0;
    int64_t stepCount = 0;
    double damping = 0.0, targetPressure = 0.0, rampDuration = 0.0;
    double hourglassCoefficient = DEFAULT_HOURGLASS_COEFFICIENT;
    bool hasWall = false;
    RigidWall wall;
    bool hasContact = false;
    Contact contact;
    double gravityR = 0.0, gravityZ = 0.0;
    std::vector<uint8_t> gone;  // empty when erosion is off
    double massScale = 1.0, erosionDamage = HUGE_VAL;
    int erodedCount = 0, invertedCount = 0;
    double erodedEnergy = 0.0;

    bool axi() const { return formulation == NF_AXISYMMETRIC; }
    bool finite() const { return kinematics == NF_FINITE_STRAIN; }
    const Material& materialAt(int e) const {
        return materialOf.empty() ? materials[0] : materials[materialOf[e]];
    }

    ~nfo_solver() { delete pool; }

    void buildNodeSlots() {  // :347-366
        const int n = mesh.nodeCount;
        nodeSlotStart.assign(n + 1, 0);
        nodeSlot.assign(mesh.conn.size() + mesh.pressureEdges.size(), 0);
        for (size_t k = 0; k < mesh.conn.size(); k++) nodeSlotStart[mesh.conn[k] + 1]++;
        for (size_t k = 0; k < mesh.pressureEdges.size(); k++) nodeSlotStart[mesh.pressureEdges[k] + 1]++;
        for (int i = 0; i < n; i++) nodeSlotStart[i + 1] += nodeSlotStart[i];
        std::vector<int> cursor = nodeSlotStart;
        for (int e = 0; e < mesh.elementCount; e++)
            for (int c = 0; c < 4; c++) nodeSlot[cursor[mesh.conn[e * 4 + c]]++] = e * 8 + c;
        for (size_t k = 0; k < mesh.pressureEdges.size(); k += 2) {
            const int slot = (edgeSlotBase + (int)k / 2) * 8;
            nodeSlot[cursor[mesh.pressureEdges[k]]++] = slot;
            nodeSlot[cursor[mesh.pressureEdges[k + 1]]++] = slot + 1;
        }
    }

    void lumpMass() {  // :1559-1597
        for (int e = 0; e < mesh.elementCount; e++) {
            const double rho = densityOf[e];
            const int b = e * 4;
            const int n0 = mesh.conn[b], n1 = mesh.conn[b + 1], n2 = mesh.conn[b + 2], n3 = mesh.conn[b + 3];
            const double r0 = mesh.r[n0], r1 = mesh.r[n1], r2 = mesh.r[n2], r3 = mesh.r[n3];
            const double z0 = mesh.z[n0], z1 = mesh.z[n1], z2 = mesh.z[n2], z3 = mesh.z[n3];
            for (int g = 0; g < 4; g++) {
                const double xi = GP_XI[g], eta = GP_ETA[g];
                const double xm = 1.0 - xi, xp = 1.0 + xi;
                const double em = 1.0 - eta, ep = 1.0 + eta;
                const double N0 = 0.25 * xm * em, N1 = 0.25 * xp * em;
                const double N2 = 0.25 * xp * ep, N3 = 0.25 * xm * ep;
                const double dN0x = -0.25 * em, dN1x = 0.25 * em, dN2x = 0.25 * ep, dN3x = -0.25 * ep;
                const double dN0e = -0.25 * xm, dN1e = -0.25 * xp, dN2e = 0.25 * xp, dN3e = 0.25 * xm;
                const double J00 = dN0x * r0 + dN1x * r1 + dN2x * r2 + dN3x * r3;
                const double J01 = dN0x * z0 + dN1x * z1 + dN2x * z2 + dN3x * z3;
                const double J10 = dN0e * r0 + dN1e * r1 + dN2e * r2 + dN3e * r3;
                const double J11 = dN0e * z0 + dN1e * z1 + dN2e * z2 + dN3e * z3;
                const double det = J00 * J11 - J01 * J10;
                const double rg = N0 * r0 + N1 * r1 + N2 * r2 + N3 * r3;
                const double dV = axi() ? 2.0 * PI * rg * det : thickness * det;
                mass[n0] += rho * N0 * dV;
                mass[n1] += rho * N1 * dV;
                mass[n2] += rho * N2 * dV;
                mass[n3] += rho * N3 * dV;
            }
        }
    }

    double pressureAt(double t) const {  // :1203-1207
        if (rampDuration <= 0.0) return targetPressure;
        double s = jmin(1.0, t / rampDuration);
        return targetPressure * s * s * (3.0 - 2.0 * s);
    }

    double cr(int node, double shift) const {  // :1176-1178
        return finite() ? mesh.r[node] + ur[node] + shift * vr[node] : mesh.r[node];
    }
    double cz(int node, double shift) const {
        return finite() ? mesh.z[node] + uz[node] + shift * vz[node] : mesh.z[node];
    }

    static double sq(double a, double b) { return a * a + b * b; }
    static double shortestEdge(double running, double r0, double r1, double r2, double r3,
                               double z0, double z1, double z2, double z3) {  // :1191-1197
        return jmin(running, jmin(jmin(sq(r1 - r0, z1 - z0), sq(r2 - r1, z2 - z1)),
                                  jmin(sq(r3 - r2, z3 - z2), sq(r0 - r3, z0 - z3))));
    }

    double fullIntegration(double strainDt, int from, int to) {  // :1219-1339
        const int* conn = mesh.conn.data();
        double* ef = contribution.data();
        const bool ax = axi();
        const bool fin = finite();
        const double mid = -0.5 * strainDt;
        Damage* dmg = damage;
        double minEdge = MAX_VALUE;
        for (int e = from; e < to; e++) {
            const int b = e * 4;
            if (!gone.empty() && gone[e]) {
                for (int k = 0; k < 8; k++) ef[b * 2 + k] = 0.0;
                continue;
            }
            const Flow& fl = flowOf[e];
            const double lambda = lambdaOf[e], mu = muOf[e];
            const int n0 = conn[b], n1 = conn[b + 1], n2 = conn[b + 2], n3 = conn[b + 3];
            const double r0 = cr(n0, mid), r1 = cr(n1, mid), r2 = cr(n2, mid), r3 = cr(n3, mid);
            const double z0 = cz(n0, mid), z1 = cz(n1, mid), z2 = cz(n2, mid), z3 = cz(n3, mid);
            const double a0 = vr[n0] * strainDt, a1 = vr[n1] * strainDt;
            const double a2 = vr[n2] * strainDt, a3 = vr[n3] * strainDt;
            const double c0 = vz[n0] * strainDt, c1 = vz[n1] * strainDt;
            const double c2 = vz[n2] * strainDt, c3 = vz[n3] * strainDt;
            if (fin) minEdge = shortestEdge(minEdge, r0, r1, r2, r3, z0, z1, z2, z3);

            double g0r = 0, g1r = 0, g2r = 0, g3r = 0;
            double g0z = 0, g1z = 0, g2z = 0, g3z = 0;
            for (int g = 0; g < 4; g++) {
                const double xi = GP_XI[g], eta = GP_ETA[g];
                const double xm = 1.0 - xi, xp = 1.0 + xi;
                const double em = 1.0 - eta, ep = 1.0 + eta;
                const double N0 = 0.25 * xm * em, N1 = 0.25 * xp * em;
                const double N2 = 0.25 * xp * ep, N3 = 0.25 * xm * ep;
                const double dN0x = -0.25 * em, dN1x = 0.25 * em, dN2x = 0.25 * ep, dN3x = -0.25 * ep;
                const double dN0e = -0.25 * xm, dN1e = -0.25 * xp, dN2e = 0.25 * xp, dN3e = 0.25 * xm;
                const double J00 = dN0x * r0 + dN1x * r1 + dN2x * r2 + dN3x * r3;
                const double J01 = dN0x * z0 + dN1x * z1 + dN2x * z2 + dN3x * z3;
                const double J10 = dN0e * r0 + dN1e * r1 + dN2e * r2 + dN3e * r3;
                const double J11 = dN0e * z0 + dN1e * z1 + dN2e * z2 + dN3e * z3;
                const double det = J00 * J11 - J01 * J10;
                const double inv = 1.0 / det;
                const double d0r = (J11 * dN0x - J01 * dN0e) * inv;
                const double d1r = (J11 * dN1x - J01 * dN1e) * inv;
                const double d2r = (J11 * dN2x - J01 * dN2e) * inv;
                const double d3r = (J11 * dN3x - J01 * dN3e) * inv;
                const double d0z = (-J10 * dN0x + J00 * dN0e) * inv;
                const double d1z = (-J10 * dN1x + J00 * dN1e) * inv;
                const double d2z = (-J10 * dN2x + J00 * dN2e) * inv;
                const double d3z = (-J10 * dN3x + J00 * dN3e) * inv;
                const double dEpsR = d0r * a0 + d1r * a1 + d2r * a2 + d3r * a3;
                const double dEpsZ = d0z * c0 + d1z * c1 + d2z * c2 + d3z * c3;
                const double dRdZ = d0z * a0 + d1z * a1 + d2z * a2 + d3z * a3;
                const double dZdR = d0r * c0 + d1r * c1 + d2r * c2 + d3r * c3;
                const double dGamRZ = dRdZ + dZdR;
                double rg = 0.0, dEpsT = 0.0;
                if (ax) {
                    rg = N0 * r0 + N1 * r1 + N2 * r2 + N3 * r3;
                    dEpsT = (N0 * a0 + N1 * a1 + N2 * a2 + N3 * a3) / rg;
                }
                const int p = e * 4 + g;
                if (fin) {
                    const double spinDt = 0.5 * (dZdR - dRdZ);
                    if (spinDt != 0.0) {
                        const double cs = corot_cos(spinDt), sn = corot_sin(spinDt);
                        corot_rotate(sig.data(), p, cs, sn);
                        corot_rotate(back.data(), p, cs, sn);
                    }
                }
                j2_update(sig.data(), back.data(), epsP.data(), wPlastic.data(), p, dEpsR, dEpsZ,
                          dEpsT, dGamRZ, lambda, mu, fl, strainDt, dmg);
                const int s = 4 * p;
                const double sR = sig[s], sZ = sig[s + 1], sT = sig[s + 2], sRZ = sig[s + 3];
                const double dV = ax ? 2.0 * PI * rg * det : thickness * det;
                if (ax) {
                    const double h = sT / rg * dV;
                    g0r -= (d0r * sR + d0z * sRZ) * dV + N0 * h;
                    g1r -= (d1r * sR + d1z * sRZ) * dV + N1 * h;
                    g2r -= (d2r * sR + d2z * sRZ) * dV + N2 * h;
                    g3r -= (d3r * sR + d3z * sRZ) * dV + N3 * h;
                } else {
                    g0r -= (d0r * sR + d0z * sRZ) * dV;
                    g1r -= (d1r * sR + d1z * sRZ) * dV;
                    g2r -= (d2r * sR + d2z * sRZ) * dV;
                    g3r -= (d3r * sR + d3z * sRZ) * dV;
                }
                g0z -= (d0z * sZ + d0r * sRZ) * dV;
                g1z -= (d1z * sZ + d1r * sRZ) * dV;
                g2z -= (d2z * sZ + d2r * sRZ) * dV;
                g3z -= (d3z * sZ + d3r * sRZ) * dV;
            }
            ef[b * 2] = g0r; ef[b * 2 + 1] = g1r; ef[b * 2 + 2] = g2r; ef[b * 2 + 3] = g3r;
            ef[b * 2 + 4] = g0z; ef[b * 2 + 5] = g1z; ef[b * 2 + 6] = g2z; ef[b * 2 + 7] = g3z;
        }
        return minEdge;
    }

    double reducedIntegration(double strainDt, int from, int to) {  // :1373-1504
        const int* conn = mesh.conn.data();
        double* ef = contribution.data();
        const bool ax = axi();
        const bool fin = finite();
        const double mid = -0.5 * strainDt;
        const double qhg = hourglassCoefficient;
        Damage* dmg = damage;
        double minEdge = MAX_VALUE;
        for (int e = from; e < to; e++) {
            const int b = e * 4;
            if (!gone.empty() && gone[e]) {
                for (int k = 0; k < 8; k++) ef[b * 2 + k] = 0.0;
                continue;
            }
            const Flow& fl = flowOf[e];
            const double lambda = lambdaOf[e], mu = muOf[e];
            const int n0 = conn[b], n1 = conn[b + 1], n2 = conn[b + 2], n3 = conn[b + 3];
            const double r0 = cr(n0, mid), r1 = cr(n1, mid), r2 = cr(n2, mid), r3 = cr(n3, mid);
            const double z0 = cz(n0, mid), z1 = cz(n1, mid), z2 = cz(n2, mid), z3 = cz(n3, mid);
            const double a0 = vr[n0] * strainDt, a1 = vr[n1] * strainDt;
            const double a2 = vr[n2] * strainDt, a3 = vr[n3] * strainDt;
            const double c0 = vz[n0] * strainDt, c1 = vz[n1] * strainDt;
            const double c2 = vz[n2] * strainDt, c3 = vz[n3] * strainDt;
            if (fin) minEdge = shortestEdge(minEdge, r0, r1, r2, r3, z0, z1, z2, z3);

            const double J00 = 0.25 * (-r0 + r1 + r2 - r3);
            const double J01 = 0.25 * (-z0 + z1 + z2 - z3);
            const double J10 = 0.25 * (-r0 - r1 + r2 + r3);
            const double J11 = 0.25 * (-z0 - z1 + z2 + z3);
            const double det = J00 * J11 - J01 * J10;
            const double inv = 1.0 / det;
            const double d0r = 0.25 * (J01 - J11) * inv;
            const double d1r = 0.25 * (J01 + J11) * inv;
            const double d2r = -d0r;
            const double d3r = -d1r;
            const double d0z = 0.25 * (J10 - J00) * inv;
            const double d1z = -0.25 * (J10 + J00) * inv;
            const double d2z = -d0z;
            const double d3z = -d1z;
            const double dEpsR = d0r * a0 + d1r * a1 + d2r * a2 + d3r * a3;
            const double dEpsZ = d0z * c0 + d1z * c1 + d2z * c2 + d3z * c3;
            const double dRdZ = d0z * a0 + d1z * a1 + d2z * a2 + d3z * a3;
            const double dZdR = d0r * c0 + d1r * c1 + d2r * c2 + d3r * c3;
            const double dGamRZ = dRdZ + dZdR;
            double rg = 0.0, dEpsT = 0.0;
            if (ax) {
                rg = 0.25 * (r0 + r1 + r2 + r3);
                dEpsT = 0.25 * (a0 + a1 + a2 + a3) / rg;
            }
            if (fin) {
                const double spinDt = 0.5 * (dZdR - dRdZ);
                if (spinDt != 0.0) {
                    const double cs = corot_cos(spinDt), sn = corot_sin(spinDt);
                    corot_rotate(sig.data(), e, cs, sn);
                    corot_rotate(back.data(), e, cs, sn);
                }
            }
            j2_update(sig.data(), back.data(), epsP.data(), wPlastic.data(), e, dEpsR, dEpsZ, dEpsT,
                      dGamRZ, lambda, mu, fl, strainDt, dmg);
            const int s = 4 * e;
            const double sR = sig[s], sZ = sig[s + 1], sT = sig[s + 2], sRZ = sig[s + 3];
            const double dV = W_REDUCED * (ax ? 2.0 * PI * rg * det : thickness * det);
            double g0r, g1r, g2r, g3r;
            if (ax) {
                const double h = sT / rg * dV * 0.25;
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
            double g0z = -(d0z * sZ + d0r * sRZ) * dV;
            double g1z = -(d1z * sZ + d1r * sRZ) * dV;
            double g2z = -(d2z * sZ + d2r * sRZ) * dV;
            double g3z = -(d3z * sZ + d3r * sRZ) * dV;
            if (qhg > 0.0) {
                const double hr = r0 - r1 + r2 - r3;
                const double hz = z0 - z1 + z2 - z3;
                const double y0 = H[0] - hr * d0r - hz * d0z;
                const double y1 = H[1] - hr * d1r - hz * d1z;
                const double y2 = H[2] - hr * d2r - hz * d2z;
                const double y3 = H[3] - hr * d3r - hz * d3z;
                const double k = qhg * mu * dV *
                                 (d0r * d0r + d1r * d1r + d2r * d2r + d3r * d3r + d0z * d0z +
                                  d1z * d1z + d2z * d2z + d3z * d3z);
                hgR[e] += k * (y0 * a0 + y1 * a1 + y2 * a2 + y3 * a3);
                hgZ[e] += k * (y0 * c0 + y1 * c1 + y2 * c2 + y3 * c3);
                const double qr = hgR[e], qz = hgZ[e];
                g0r -= y0 * qr; g1r -= y1 * qr; g2r -= y2 * qr; g3r -= y3 * qr;
                g0z -= y0 * qz; g1z -= y1 * qz; g2z -= y2 * qz; g3z -= y3 * qz;
                hourglassStiffness[e] = k;
            }
            ef[b * 2] = g0r; ef[b * 2 + 1] = g1r; ef[b * 2 + 2] = g2r; ef[b * 2 + 3] = g3r;
            ef[b * 2 + 4] = g0z; ef[b * 2 + 5] = g1z; ef[b * 2 + 6] = g2z; ef[b * 2 + 7] = g3z;
        }
        return minEdge;
    }

    void applyPressure(double p, int from, int to) {  // :1511-1556
        const int* edges = mesh.pressureEdges.data();
        double* c = contribution.data();
        for (int e = from; e < to; e++) {
            const int k = 2 * e;
            const int slot = (edgeSlotBase + e) * 8;
            const int a = edges[k], bnode = edges[k + 1];
            const double ra = cr(a, 0.0), rb = cr(bnode, 0.0);
            const double tr = rb - ra;
            const double tz = cz(bnode, 0.0) - cz(a, 0.0);
            const double len = std::sqrt(tr * tr + tz * tz);
            if (len == 0.0) {
                c[slot] = 0.0;
                c[slot + 1] = 0.0;
                c[slot + 4] = 0.0;
                c[slot + 5] = 0.0;
                continue;
            }
            const double nr = tz / len;
            const double nz = -tr / len;
            double wa, wb;
            if (axi()) {
                const double circumference = 2.0 * PI * len;
                wa = circumference * (ra / 3.0 + rb / 6.0);
                wb = circumference * (ra / 6.0 + rb / 3.0);
            } else {
                wa = thickness * len * 0.5;
                wb = wa;
            }
            c[slot] = p * nr * wa;
            c[slot + 1] = p * nr * wb;
            c[slot + 4] = p * nz * wa;
            c[slot + 5] = p * nz * wb;
        }
    }

    void assemble(int from, int to, int worker) {  // :1126-1137
        const int ne = edgeSlotBase;
        if (from < ne) {
            const int end = std::min(to, ne);
            const double edge = integration == NF_FULL ? fullIntegration(kernelStrainDt, from, end)
                                                       : reducedIntegration(kernelStrainDt, from, end);
            const int slot = worker * PAD;
            if (edge < minEdgeLocal[slot]) minEdgeLocal[slot] = edge;
        }
        if (to > ne) applyPressure(kernelPressure, std::max(from, ne) - ne, to - ne);
    }

    void gather(int from, int to) {  // :1147-1165
        const double* c = contribution.data();
        const int* slots = nodeSlot.data();
        for (int n = from; n < to; n++) {
            double sr = 0.0, sz = 0.0;
            const int end = nodeSlotStart[n + 1];
            for (int k = nodeSlotStart[n]; k < end; k++) {
                const int slot = slots[k];
                sr += c[slot];
                sz += c[slot + 4];
            }
            fr[n] = sr + mass[n] * gravityR;
            fz[n] = sz + mass[n] * gravityZ;
        }
    }

    static void assembleKernel(void* ctx, int from, int to, int worker) {
        static_cast<nfo_solver*>(ctx)->assemble(from, to, worker);
    }
    static void gatherKernel(void* ctx, int from, int to, int) {
        static_cast<nfo_solver*>(ctx)->gather(from, to);
    }
    static void nodalKernel(void* ctx, int from, int to, int) {
        static_cast<nfo_solver*>(ctx)->nodalUpdate(from, to);
    }
    static void contactKernel(void* ctx, int from, int to, int) {
        static_cast<nfo_solver*>(ctx)->contact.search(from, to);
    }

    void advanceMaterialAndAssemble(double strainDt) {  // :1081-1115
        kernelStrainDt = strainDt;
        kernelPressure = pressureAt(time);
        std::fill(minEdgeLocal.begin(), minEdgeLocal.end(), MAX_VALUE);
        const int sources = edgeSlotBase + (int)mesh.pressureEdges.size() / 2;
        if (pool == nullptr) {
            assemble(0, sources, 0);
            gather(0, mesh.nodeCount);
        } else {
            pool->run(sources, assembleKernel, this);
            pool->run(mesh.nodeCount, gatherKernel, this);
        }
        if (hasContact) {
            const int slaves = contact.begin(ur.data(), uz.data(), dt, dt0);
            if (slaves > 0) {
                if (pool != nullptr) pool->run(slaves, contactKernel, this);
                else contact.search(0, slaves);
                contact.accumulate(vr.data(), vz.data(), mass.data(), fr.data(), fz.data());
            }
        }
        minEdgeSquared = MAX_VALUE;
        for (size_t w = 0; w < minEdgeLocal.size(); w += PAD)
            minEdgeSquared = jmin(minEdgeSquared, minEdgeLocal[w]);
    }

    void nodalUpdate(int from, int to) {  // :998-1062
        const double decay = nodalDecay;
        const double gain = nodalGain;
        const double dtAvg = nodalDtAvg;
        for (int i = from; i < to; i++) {
            const double invM = 1.0 / mass[i];
            const bool resting = hasWall && onWall[i] && wall.holds(fz[i]);
            if (fixedR[i]) {
                vr[i] = 0.0;
            } else if (resting && wall.hasFriction()) {
                vr[i] = wall.slide(fr[i], fz[i], vr[i], decay, gain, mass[i], dt);
                ur[i] += vr[i] * dt;
            } else {
                vr[i] = decay * vr[i] + gain * fr[i] * invM;
                ur[i] += vr[i] * dt;
            }
            if (fixedZ[i]) {
                vz[i] = 0.0;
            } else if (resting) {
                wall.hold(fz[i], dtAvg);
            } else {
                const double trial = decay * vz[i] + gain * fz[i] * invM;
                const double zc = mesh.z[i] + uz[i];
                if (hasWall && wall.crosses(zc, trial, dt)) {
                    wall.arrive(trial, mass[i]);
                    uz[i] += wall.plane - zc;
                    vz[i] = 0.0;
                    onWall[i] = 1;
                } else {
                    vz[i] = trial;
                    uz[i] += vz[i] * dt;
                    if (hasWall) onWall[i] = 0;
                }
            }
        }
    }

    double currentArea(int element) const {  // :681-691
        const int b = element * 4;
        double a = 0.0;
        for (int k = 0; k < 4; k++) {
            const int i = mesh.conn[b + k];
            const int j = mesh.conn[b + (k + 1) % 4];
            a += (mesh.r[i] + ur[i]) * (mesh.z[j] + uz[j]) - (mesh.r[j] + ur[j]) * (mesh.z[i] + uz[i]);
        }
        return 0.5 * a;
    }

    double elementDamage(int element) const {  // :1764-1772
        if (damage == nullptr) return 0.0;
        double sum = 0.0;
        for (int g = 0; g < pointsPerElement; g++) {
            const int p = element * pointsPerElement + g;
            sum += damage->at(p, epsP[p]);
        }
        return sum / pointsPerElement;
    }

    double gaussVolume(int element, int g) const {  // :1939-1965
        const int b = element * 4;
        const int n0 = mesh.conn[b], n1 = mesh.conn[b + 1], n2 = mesh.conn[b + 2], n3 = mesh.conn[b + 3];
        const double r0 = mesh.r[n0], r1 = mesh.r[n1], r2 = mesh.r[n2], r3 = mesh.r[n3];
        const double z0 = mesh.z[n0], z1 = mesh.z[n1], z2 = mesh.z[n2], z3 = mesh.z[n3];
        const double xi = pointsPerElement == 1 ? 0.0 : GP_XI[g];
        const double eta = pointsPerElement == 1 ? 0.0 : GP_ETA[g];
        const double weight = pointsPerElement == 1 ? W_REDUCED : 1.0;
        const double xm = 1.0 - xi, xp = 1.0 + xi, em = 1.0 - eta, ep = 1.0 + eta;
        const double N0 = 0.25 * xm * em, N1 = 0.25 * xp * em;
        const double N2 = 0.25 * xp * ep, N3 = 0.25 * xm * ep;
        const double dN0x = -0.25 * em, dN1x = 0.25 * em, dN2x = 0.25 * ep, dN3x = -0.25 * ep;
        const double dN0e = -0.25 * xm, dN1e = -0.25 * xp, dN2e = 0.25 * xp, dN3e = 0.25 * xm;
        const double J00 = dN0x * r0 + dN1x * r1 + dN2x * r2 + dN3x * r3;
        const double J01 = dN0x * z0 + dN1x * z1 + dN2x * z2 + dN3x * z3;
        const double J10 = dN0e * r0 + dN1e * r1 + dN2e * r2 + dN3e * r3;
        const double J11 = dN0e * z0 + dN1e * z1 + dN2e * z2 + dN3e * z3;
        const double det = J00 * J11 - J01 * J10;
        const double rg = N0 * r0 + N1 * r1 + N2 * r2 + N3 * r3;
        return weight * (axi() ? 2.0 * PI * rg * det : thickness * det);
    }

    double pointStrainEnergy(int e, int g) const {  // :1666-1675
        const int p = e * pointsPerElement + g;
        const int s = 4 * p;
        const double press = (sig[s] + sig[s + 1] + sig[s + 2]) / 3.0;
        const double a = sig[s] - press, bb = sig[s + 1] - press, c = sig[s + 2] - press;
        const double d = sig[s + 3];
        const double density = (a * a + bb * bb + c * c + 2.0 * d * d) / (4.0 * muOf[e]) +
                               press * press / (2.0 * bulkOf[e]);
        return density * gaussVolume(e, g);
    }

    void erode() {  // :663-678
        for (int e = 0; e < mesh.elementCount; e++) {
            if (gone[e]) continue;
            const bool failed = damage != nullptr && elementDamage(e) >= erosionDamage;
            const bool inverted = currentArea(e) <= 0.0;
            if (!failed && !inverted) continue;
            gone[e] = 1;
            erodedCount++;
            if (inverted && !failed) invertedCount++;
            for (int g = 0; g < pointsPerElement; g++) erodedEnergy += pointStrainEnergy(e, g);
            if (hasContact) contact.forget(e);
        }
    }

    void step() {  // :940-991
        if (stepCount == 0) advanceMaterialAndAssemble(0.0);
        const double dtAvg = 0.5 * (dtPrevious + dt);
        const double half = 0.5 * damping * dtAvg;
        nodalDecay = (1.0 - half) / (1.0 + half);
        nodalGain = dtAvg / (1.0 + half);
        nodalDtAvg = dtAvg;
        if (hasWall) wall.beginStep();
        if (pool != nullptr && !hasWall) pool->run(mesh.nodeCount, nodalKernel, this);
        else nodalUpdate(0, mesh.nodeCount);
        time += dt;
        stepCount++;
        if (!gone.empty()) erode();
        advanceMaterialAndAssemble(dt);
        if (finite()) {
            dtPrevious = dt;
            dt = cflSafety * std::sqrt(minEdgeSquared * massScale) / fastestWave;
        }
    }

    void advance() {  // :1069-1075
        for (int i = 0; i < mesh.nodeCount; i++) {
            if (!fixedR[i]) ur[i] += vr[i] * dt;
            if (!fixedZ[i]) uz[i] += vz[i] * dt;
        }
        advanceMaterialAndAssemble(dt);
    }
};

// ------------------------------------------------------------------ C ABI (prefix nfo_)
// The error channel of the oracle's other translation units (defect_field.cpp).
int nfo_report(int code, const std::string& msg) { return fail(code, msg); }

extern "C" {

const char* nfo_last_error(void) { return g_error.c_str(); }

int nfo_material_check(const nf_material* m) {
    g_error.clear();
    return check_material(m);
}

int nfo_solver_create(nfo_solver** out, int32_t node_count, const double* r, const double* z,
                      int32_t element_count, const int32_t* conn, int32_t pressure_edge_count,
                      const int32_t* pressure_edges, int32_t material_count,
                      const nf_material* materials, const int32_t* material_of,
                      int32_t formulation, int32_t integration, int32_t kinematics,
                      double thickness, double cfl_safety) {
    g_error.clear();
    if (!out) return fail(NF_E_ARGUMENT, "out must not be null");
    *out = nullptr;
    if (node_count < 1 || element_count < 1 || !r || !z || !conn)
        return fail(NF_E_ARGUMENT, "mesh arrays must be non-empty");
    if (pressure_edge_count < 0 || (pressure_edge_count > 0 && !pressure_edges))
        return fail(NF_E_ARGUMENT, "bad pressure edges");
    if (material_count < 1 || !materials) return fail(NF_E_ARGUMENT, "need at least one material");
    if (!material_of && material_count != 1)
        return fail(NF_E_ARGUMENT, "a uniform mesh takes exactly one material");
    if (formulation < 0 || formulation > 2 || integration < 0 || integration > 1 ||
        kinematics < 0 || kinematics > 1)
        return fail(NF_E_ARGUMENT, "unknown formulation, integration or kinematics");
    for (int i = 0; i < 4 * element_count; i++)
        if (conn[i] < 0 || conn[i] >= node_count) return fail(NF_E_ARGUMENT, "connectivity out of range");
    for (int i = 0; i < 2 * pressure_edge_count; i++)
        if (pressure_edges[i] < 0 || pressure_edges[i] >= node_count)
            return fail(NF_E_ARGUMENT, "pressure edge out of range");
    if (material_of)
        for (int e = 0; e < element_count; e++)
            if (material_of[e] < 0 || material_of[e] >= material_count)
                return fail(NF_E_ARGUMENT, "material index out of range");
    for (int m = 0; m < material_count; m++) {
        int rc = check_material(&materials[m]);
        if (rc) return rc;
        // solver/ExplicitSolver.java:268-276
        if (formulation == NF_PLANE_STRESS && !std::isinf(materials[m].yield_stress))
            return fail(NF_E_ARGUMENT, "plane stress plasticity needs its own return map; not implemented");
    }

    nfo_solver* s = new nfo_solver();
    Mesh& mesh = s->mesh;
    mesh.nodeCount = node_count;
    mesh.elementCount = element_count;
    mesh.r.assign(r, r + node_count);
    mesh.z.assign(z, z + node_count);
    mesh.conn.assign(conn, conn + 4 * element_count);
    if (pressure_edge_count > 0) mesh.pressureEdges.assign(pressure_edges, pressure_edges + 2 * pressure_edge_count);
    for (int e = 0; e < element_count; e++) {  // mesh/QuadMesh.java:232-240
        if (mesh.signedArea(e) <= 0.0) {
            delete s;
            return fail(NF_E_STATE, "element " + std::to_string(e) +
                                        " has non-positive area; node ordering must be counter-clockwise in (r, z)");
        }
    }
    for (int m = 0; m < material_count; m++) s->materials.push_back(to_material(materials[m]));
    if (material_of) s->materialOf.assign(material_of, material_of + element_count);
    s->formulation = formulation;
    s->integration = integration;
    s->kinematics = kinematics;
    s->thickness = thickness;
    s->cflSafety = cfl_safety;

    const int E = element_count, N = node_count;
    s->lambdaOf.resize(E);
    s->muOf.resize(E);
    s->bulkOf.resize(E);
    s->densityOf.resize(E);
    s->flowOf.resize(E);
    for (int e = 0; e < E; e++) {  // :289-299
        const Material& m = s->materialAt(e);
        const double lam = formulation == NF_PLANE_STRESS ? 2.0 * m.lambda() * m.mu() / (m.lambda() + 2.0 * m.mu())
                                                          : m.lambda();
        s->lambdaOf[e] = lam;
        s->muOf[e] = m.mu();
        s->bulkOf[e] = lam + 2.0 * m.mu() / 3.0;
        s->densityOf[e] = m.rho;
        Flow f;
        f.sy0 = m.sy;
        f.hIso = m.hiso;
        f.hKin = m.hkin;
        f.jc = m.has_jc ? &m.jc : nullptr;
        f.thermalCoefficient = m.has_jc ? m.jc.thermalCoefficient(m.rho) : 0.0;
        s->flowOf[e] = f;
    }
    // Materials.fastestWave core/Materials.java:120-128
    if (s->materialOf.empty()) {
        s->fastestWave = s->materials[0].dilatationalWaveSpeed();
    } else {
        double fastest = 0.0;
        for (int e = 0; e < E; e++) {
            const double c = s->materialAt(e).dilatationalWaveSpeed();
            if (c > fastest) fastest = c;
        }
        s->fastestWave = fastest;
    }
    s->ur.assign(N, 0.0);
    s->uz.assign(N, 0.0);
    s->vr.assign(N, 0.0);
    s->vz.assign(N, 0.0);
    s->fr.assign(N, 0.0);
    s->fz.assign(N, 0.0);
    s->mass.assign(N, 0.0);
    s->fixedR.assign(N, 0);
    s->fixedZ.assign(N, 0);
    s->onWall.assign(N, 0);
    s->pointsPerElement = integration == NF_FULL ? 4 : 1;
    const int P = E * s->pointsPerElement;
    s->sig.assign(4 * P, 0.0);
    s->back.assign(4 * P, 0.0);
    s->epsP.assign(P, 0.0);
    s->wPlastic.assign(P, 0.0);
    s->hgR.assign(E, 0.0);
    s->hgZ.assign(E, 0.0);
    s->hourglassStiffness.assign(E, 0.0);
    s->edgeSlotBase = E;
    s->contribution.assign(8 * (E + pressure_edge_count), 0.0);
    s->buildNodeSlots();
    s->minEdgeLocal.assign(PAD, MAX_VALUE);
    s->lumpMass();
    const double edge = mesh.minimumEdgeLength();
    s->minEdgeSquared = edge * edge;
    s->dt = cfl_safety * edge / s->fastestWave;
    s->dt0 = s->dt;
    s->dtPrevious = s->dt;
    *out = s;
    return NF_OK;
}

int nfo_solver_destroy(nfo_solver* s) {
    delete s;
    return NF_OK;
}

int nfo_set_threads(nfo_solver* s, int32_t n) {  // :386-392
    g_error.clear();
    if (n < 1) return fail(NF_E_ARGUMENT, "need at least one thread");
    if (s->pool != nullptr && s->pool->threads() == n) return NF_OK;
    delete s->pool;
    s->pool = nullptr;
    s->minEdgeLocal.assign((size_t)PAD * n, MAX_VALUE);
    if (n > 1) s->pool = new Pool(n);
    return NF_OK;
}

int nfo_fix_radial(nfo_solver* s, int32_t node) {
    if (node < 0 || node >= s->mesh.nodeCount) return fail(NF_E_ARGUMENT, "node out of range");
    s->fixedR[node] = 1;
    return NF_OK;
}
int nfo_fix_axial(nfo_solver* s, int32_t node) {
    if (node < 0 || node >= s->mesh.nodeCount) return fail(NF_E_ARGUMENT, "node out of range");
    s->fixedZ[node] = 1;
    return NF_OK;
}
int nfo_fix_all_axial(nfo_solver* s) {
    std::fill(s->fixedZ.begin(), s->fixedZ.end(), 1);
    return NF_OK;
}
int nfo_set_pressure_ramp(nfo_solver* s, double pascals, double ramp_seconds) {
    s->targetPressure = pascals;
    s->rampDuration = ramp_seconds;
    return NF_OK;
}
int nfo_set_relaxation_damping(nfo_solver* s, double fraction, double omega) {
    if (omega <= 0.0) return fail(NF_E_ARGUMENT, "angular frequency must be positive");
    s->damping = 2.0 * fraction * omega;
    return NF_OK;
}
int nfo_set_hourglass_coefficient(nfo_solver* s, double q) {
    if (q < 0.0) return fail(NF_E_ARGUMENT, "hourglass coefficient must be >= 0");
    s->hourglassCoefficient = q;
    return NF_OK;
}
int nfo_set_rigid_wall(nfo_solver* s, int32_t enabled, double plane_z, double friction) {
    if (!enabled) {
        s->hasWall = false;
        return NF_OK;
    }
    if (friction < 0.0 || friction != friction)
        return fail(NF_E_ARGUMENT, "friction coefficient must be >= 0");
    s->hasWall = true;
    s->wall = RigidWall();
    s->wall.plane = plane_z;
    s->wall.friction = friction;
    return NF_OK;
}
int nfo_set_contact(nfo_solver* s, int32_t enabled, double friction, double damping, double scale) {
    if (!enabled) {
        s->hasContact = false;
        return NF_OK;
    }
    if (friction < 0.0) return fail(NF_E_ARGUMENT, "friction must not be negative");
    if (damping < 0.0 || damping > 1.0)
        return fail(NF_E_ARGUMENT, "contact damping is a fraction of critical, so it lies in [0, 1]");
    if (!(scale > 0.0) || scale >= 2.0) return fail(NF_E_ARGUMENT, "the stiffness scale must lie in (0, 2)");
    s->contact = Contact();
    s->contact.init(&s->mesh);
    s->contact.friction = friction;
    s->contact.damping = damping;
    s->contact.scale = scale;
    s->hasContact = true;
    return NF_OK;
}
int nfo_set_gravity(nfo_solver* s, double ar, double az) {
    s->gravityR = ar;
    s->gravityZ = az;
    return NF_OK;
}
int nfo_set_damage(nfo_solver* s, const double* gf, const double* eps) {  // :729-791
    g_error.clear();
    if (!gf || !eps) return fail(NF_E_ARGUMENT, "fracture energy and failure strain are required");
    for (const Material& m : s->materials) {
        const double worst = 2.0 * m.mu() + (2.0 / 3.0) * (m.hkin - m.E);
        if (!(worst > 0.0)) return fail(NF_E_ARGUMENT, "a fully clamped softening modulus would make the return map singular");
    }
    const int E = s->mesh.elementCount, ppe = s->pointsPerElement, P = E * ppe;
    Damage d;
    d.fractureEnergy.resize(P);
    d.youngsModulus.resize(P);
    d.failureStrain.resize(P);
    d.geometry.resize((size_t)Damage::GEOMETRY_STRIDE * P);
    const Mesh& mesh = s->mesh;
    for (int e = 0; e < E; e++) {
        const int b = e * 4;
        const int n0 = mesh.conn[b], n1 = mesh.conn[b + 1], n2 = mesh.conn[b + 2], n3 = mesh.conn[b + 3];
        const double uR = 0.5 * (mesh.r[n1] - mesh.r[n0] + mesh.r[n2] - mesh.r[n3]);
        const double uZ = 0.5 * (mesh.z[n1] - mesh.z[n0] + mesh.z[n2] - mesh.z[n3]);
        const double vR = 0.5 * (mesh.r[n3] - mesh.r[n0] + mesh.r[n2] - mesh.r[n1]);
        const double vZ = 0.5 * (mesh.z[n3] - mesh.z[n0] + mesh.z[n2] - mesh.z[n1]);
        const double area = jabs(mesh.signedArea(e));
        for (int g = 0; g < ppe; g++) {
            const int p = e * ppe + g;
            d.failureStrain[p] = eps[e];
            d.fractureEnergy[p] = gf[e];
            d.youngsModulus[p] = s->materialAt(e).E;
            const int q = Damage::GEOMETRY_STRIDE * p;
            d.geometry[q] = area;
            d.geometry[q + 1] = uR;
            d.geometry[q + 2] = uZ;
            d.geometry[q + 3] = vR;
            d.geometry[q + 4] = vZ;
        }
    }
    // Damage constructor checks solver/Damage.java:224-255
    for (int p = 0; p < P; p++) {
        if (!(d.fractureEnergy[p] > 0.0)) return fail(NF_E_ARGUMENT, "fracture energy must be positive");
        if (!(d.youngsModulus[p] > 0.0)) return fail(NF_E_ARGUMENT, "Young's modulus must be positive");
    }
    for (int p = 0; p < P; p++) {
        if (!(d.failureStrain[p] > 0.0)) return fail(NF_E_ARGUMENT, "failure strain must be positive");
        if (!(d.geometry[(size_t)Damage::GEOMETRY_STRIDE * p] > 0.0)) return fail(NF_E_ARGUMENT, "area must be positive");
    }
    d.strength.assign(P, 0.0);
    d.softening.assign(P, 0.0);
    d.width.assign(P, 0.0);
    s->damageStore = std::move(d);
    s->damage = &s->damageStore;
    return NF_OK;
}
int nfo_set_erosion(nfo_solver* s, double threshold) {  // :549-561
    if (s->damage == nullptr) return fail(NF_E_STATE, "erosion needs a failure model; call setDamage first");
    if (!(threshold > 0.0) || threshold > 1.0)
        return fail(NF_E_ARGUMENT, "the erosion threshold is a damage, so it lies in (0, 1]");
    s->erosionDamage = threshold;
    if (s->gone.empty()) s->gone.assign(s->mesh.elementCount, 0);
    return NF_OK;
}
int nfo_set_mass_scaling(nfo_solver* s, double target) {  // :601-619
    if (!(target > 0.0)) return fail(NF_E_ARGUMENT, "the target step must be positive");
    if (target < s->dt) return fail(NF_E_ARGUMENT, "mass scaling adds mass and cannot shorten a step");
    const double factor = (target / s->dt) * (target / s->dt);
    for (double& m : s->mass) m *= factor;
    s->massScale *= factor;
    s->dt = target;
    s->dtPrevious = s->dt;
    s->dt0 = s->dt;
    return NF_OK;
}
int nfo_set_velocity(nfo_solver* s, int32_t node, double vr, double vz) {
    if (node < 0 || node >= s->mesh.nodeCount) return fail(NF_E_ARGUMENT, "node out of range");
    s->vr[node] = vr;
    s->vz[node] = vz;
    return NF_OK;
}
int nfo_set_uniform_velocity(nfo_solver* s, double vr, double vz) {
    std::fill(s->vr.begin(), s->vr.end(), vr);
    std::fill(s->vz.begin(), s->vz.end(), vz);
    return NF_OK;
}
int nfo_run(nfo_solver* s, int64_t steps) {
    for (int64_t i = 0; i < steps; i++) s->step();
    return NF_OK;
}
int nfo_run_until(nfo_solver* s, double end_time, int64_t max_steps, int64_t* taken) {
    int64_t n = 0;
    while (s->time < end_time && (max_steps <= 0 || n < max_steps)) {
        s->step();
        n++;
    }
    if (taken) *taken = n;
    return NF_OK;
}
// The batch surface: on the CPU a batch is its solvers, each run to the end in turn.
struct nfo_batch {
    std::vector<nfo_solver*> solvers;
    std::vector<int> tile;
};
int nfo_batch_create(nfo_batch** out, int32_t count, nfo_solver* const* solvers) {
    if (!out || count < 1 || !solvers) return fail(NF_E_ARGUMENT, "a batch needs at least one solver");
    auto* b = new nfo_batch();
    b->solvers.assign(solvers, solvers + count);
    b->tile.assign(count, 256);
    *out = b;
    return NF_OK;
}
int nfo_batch_destroy(nfo_batch* b) {
    delete b;
    return NF_OK;
}
int nfo_batch_set_tile(nfo_batch* b, int32_t instance, int32_t threads) {
    if (instance < -1 || instance >= (int)b->tile.size()) return fail(NF_E_ARGUMENT, "instance out of range");
    if (threads != 64 && threads != 128 && threads != 256) return fail(NF_E_ARGUMENT, "a tile is 64, 128 or 256 threads");
    for (int i = 0; i < (int)b->tile.size(); i++)
        if (instance < 0 || i == instance) b->tile[i] = threads;
    return NF_OK;
}
int nfo_batch_get_tile(nfo_batch* b, int32_t instance, int32_t* threads) {
    if (instance < 0 || instance >= (int)b->tile.size() || !threads) return fail(NF_E_ARGUMENT, "instance out of range");
    *threads = b->tile[instance];
    return NF_OK;
}
int nfo_batch_run_until(nfo_batch* b, double end_time, int64_t max_steps) {
    for (nfo_solver* s : b->solvers) nfo_run_until(s, end_time, max_steps, nullptr);
    return NF_OK;
}
int nfo_batch_scatter(nfo_batch*) { return NF_OK; }

int nfo_advance(nfo_solver* s) {
    s->advance();
    return NF_OK;
}

int nfo_get_node_field(nfo_solver* s, int32_t field, double* out) {
    const int N = s->mesh.nodeCount;
    const std::vector<double>* src = nullptr;
    const std::vector<uint8_t>* flags = nullptr;
    switch (field) {
        case NF_NODE_UR: src = &s->ur; break;
        case NF_NODE_UZ: src = &s->uz; break;
        case NF_NODE_VR: src = &s->vr; break;
        case NF_NODE_VZ: src = &s->vz; break;
        case NF_NODE_FR: src = &s->fr; break;
        case NF_NODE_FZ: src = &s->fz; break;
        case NF_NODE_MASS: src = &s->mass; break;
        case NF_NODE_R: src = &s->mesh.r; break;
        case NF_NODE_Z: src = &s->mesh.z; break;
        case NF_NODE_ON_WALL: flags = &s->onWall; break;
        case NF_NODE_FIXED_R: flags = &s->fixedR; break;
        case NF_NODE_FIXED_Z: flags = &s->fixedZ; break;
        default: return fail(NF_E_ARGUMENT, "unknown node field");
    }
    if (src) std::memcpy(out, src->data(), sizeof(double) * N);
    else for (int i = 0; i < N; i++) out[i] = (*flags)[i] ? 1.0 : 0.0;
    return NF_OK;
}
int nfo_put_node_field(nfo_solver* s, int32_t field, const double* in) {
    const int N = s->mesh.nodeCount;
    std::vector<double>* dst = nullptr;
    switch (field) {
        case NF_NODE_UR: dst = &s->ur; break;
        case NF_NODE_UZ: dst = &s->uz; break;
        case NF_NODE_VR: dst = &s->vr; break;
        case NF_NODE_VZ: dst = &s->vz; break;
        default: return fail(NF_E_ARGUMENT, "only UR, UZ, VR and VZ may be written");
    }
    std::memcpy(dst->data(), in, sizeof(double) * N);
    return NF_OK;
}
int nfo_get_point_field(nfo_solver* s, int32_t field, double* out) {
    const int P = s->mesh.elementCount * s->pointsPerElement;
    switch (field) {
        case NF_POINT_STRESS: std::memcpy(out, s->sig.data(), sizeof(double) * 4 * P); return NF_OK;
        case NF_POINT_BACK: std::memcpy(out, s->back.data(), sizeof(double) * 4 * P); return NF_OK;
        case NF_POINT_PLASTIC_STRAIN: std::memcpy(out, s->epsP.data(), sizeof(double) * P); return NF_OK;
        case NF_POINT_PLASTIC_WORK: std::memcpy(out, s->wPlastic.data(), sizeof(double) * P); return NF_OK;
        case NF_POINT_DAMAGE_STRENGTH:
        case NF_POINT_DAMAGE_SOFTENING:
        case NF_POINT_DAMAGE_WIDTH:
        case NF_POINT_DAMAGE_FAILURE_STRAIN:
        case NF_POINT_DAMAGE_FRACTURE_ENERGY: {
            if (!s->damage) {
                std::memset(out, 0, sizeof(double) * P);
                return NF_OK;
            }
            const std::vector<double>& v = field == NF_POINT_DAMAGE_STRENGTH ? s->damage->strength
                                           : field == NF_POINT_DAMAGE_SOFTENING ? s->damage->softening
                                           : field == NF_POINT_DAMAGE_WIDTH ? s->damage->width
                                           : field == NF_POINT_DAMAGE_FAILURE_STRAIN ? s->damage->failureStrain
                                                                                     : s->damage->fractureEnergy;
            std::memcpy(out, v.data(), sizeof(double) * P);
            return NF_OK;
        }
        case NF_POINT_DAMAGE:  // Damage.at(p, epsP[p]) :342-347
            for (int p = 0; p < P; p++) out[p] = s->damage ? s->damage->at(p, s->epsP[p]) : 0.0;
            return NF_OK;
        case NF_POINT_TEMPERATURE:  // temperature(p) :1832-1840, NaN where it throws
            for (int p = 0; p < P; p++) {
                const Material& m = s->materialAt(p / s->pointsPerElement);
                out[p] = m.has_jc ? m.jc.temperature(s->wPlastic[p], m.rho) : NAN;
            }
            return NF_OK;
        default: return fail(NF_E_ARGUMENT, "unknown point field");
    }
}
// Film.capture (render/Film.java:159-189) with the element accessors of
// solver/ExplicitSolver.java the scenes colour by; see nf_film_capture.
int nfo_film_capture(nfo_solver* s, int32_t count, const int32_t* fields, double* r, double* z,
                     double* values) {
    if (!s) return fail(NF_E_ARGUMENT, "null handle");
    if (count < 0 || count > 16) return fail(NF_E_ARGUMENT, "a frame carries 0 to 16 fields");
    for (int f = 0; f < count; f++)
        if (fields[f] < NF_FILM_PLASTIC_STRAIN || fields[f] > NF_FILM_ALIVE)
            return fail(NF_E_ARGUMENT, "unknown film field " + std::to_string(fields[f]));
    const Mesh& m = s->mesh;
    for (int i = 0; i < m.nodeCount; i++) {  // :161-166
        r[i] = m.r[i] + s->ur[i];
        z[i] = m.z[i] + s->uz[i];
    }
    const int ppe = s->pointsPerElement;
    for (int f = 0; f < count; f++) {
        double* out = values + (size_t)f * m.elementCount;
        for (int e = 0; e < m.elementCount; e++) {
            double v = 0.0;
            switch (fields[f]) {
                case NF_FILM_PLASTIC_STRAIN: {  // elementPlasticStrain :1897-1901
                    double sum = 0.0;
                    for (int g = 0; g < ppe; g++) sum += s->epsP[e * ppe + g];
                    v = sum / ppe;
                    break;
                }
                case NF_FILM_DAMAGE: {  // elementDamage :1764-1772
                    if (s->damage == nullptr) break;
                    double sum = 0.0;
                    for (int g = 0; g < ppe; g++) {
                        const int p = e * ppe + g;
                        sum += s->damage->at(p, s->epsP[p]);
                    }
                    v = sum / ppe;
                    break;
                }
                case NF_FILM_SPEED: {  // elementSpeed :895-902
                    double sum = 0.0;
                    for (int k = 0; k < 4; k++) {
                        const int n = m.conn[e * 4 + k];
                        sum += std::sqrt(s->vr[n] * s->vr[n] + s->vz[n] * s->vz[n]);
                    }
                    v = 0.25 * sum;
                    break;
                }
                case NF_FILM_CENTROID_Z: {  // centroidZ :1918-1925
                    double sum = 0.0;
                    for (int k = 0; k < 4; k++) {
                        const int n = m.conn[e * 4 + k];
                        sum += m.z[n] + s->uz[n];
                    }
                    v = 0.25 * sum;
                    break;
                }
                case NF_FILM_ALIVE: v = (!s->gone.empty() && s->gone[e]) ? 0.0 : 1.0; break;
                default: {  // centroidStress :1886-1894; mises, Impact.java:229-232
                    double c[4] = {0.0, 0.0, 0.0, 0.0};
                    for (int g = 0; g < ppe; g++)
                        for (int k = 0; k < 4; k++) c[k] += s->sig[4 * (e * ppe + g) + k];
                    for (int k = 0; k < 4; k++) c[k] /= ppe;
                    if (fields[f] == NF_FILM_MISES) {
                        const double a = c[0] - c[1], b = c[1] - c[2], cc = c[2] - c[0];
                        v = std::sqrt(0.5 * (a * a + b * b + cc * cc) + 3.0 * c[3] * c[3]);
                    } else {
                        v = c[fields[f] - NF_FILM_STRESS_RR];
                    }
                }
            }
            out[e] = v;
        }
    }
    return NF_OK;
}

// ExplicitSolver's single-value queries.
int nfo_temperature(nfo_solver* s, int32_t point, double* out) {  // :1832-1840
    if (point < 0 || point >= (int)s->wPlastic.size()) return fail(NF_E_ARGUMENT, "point out of range");
    const Material& m = s->materialAt(point / s->pointsPerElement);
    if (!m.has_jc) return fail(NF_E_STATE, "this material has no thermal model; temperature is not defined");
    *out = m.jc.temperature(s->wPlastic[point], m.rho);
    return NF_OK;
}
int nfo_plastic_front_radius(nfo_solver* s, double* out) {  // :1867-1877
    const int ppe = s->pointsPerElement;
    double front = 0.0;
    for (int e = 0; e < s->mesh.elementCount; e++) {
        bool yielded = false;
        for (int g = 0; g < ppe; g++)
            if (s->epsP[e * ppe + g] > 0.0) yielded = true;
        if (yielded) {
            const int* c = &s->mesh.conn[4 * e];  // outerRadius :1932-1936
            const double outer = jmax(jmax(s->mesh.r[c[0]], s->mesh.r[c[1]]), jmax(s->mesh.r[c[2]], s->mesh.r[c[3]]));
            front = jmax(front, outer);
        }
    }
    *out = front;
    return NF_OK;
}
int nfo_most_damaged_element(nfo_solver* s, int32_t* out) {  // :1778-1794
    *out = -1;
    if (!s->damage) return NF_OK;
    const int ppe = s->pointsPerElement;
    double max = 0.0;
    for (int e = 0; e < s->mesh.elementCount; e++)
        for (int g = 0; g < ppe; g++) {
            const int p = e * ppe + g;
            const double d = s->damage->at(p, s->epsP[p]);
            if (d > max) {
                max = d;
                *out = e;
            }
        }
    return NF_OK;
}
int nfo_bore_pressure(nfo_solver* s, double* out) {  // :826-828
    *out = s->pressureAt(s->time);
    return NF_OK;
}
int nfo_mesh_z(nfo_solver* s, int32_t node, double* out) {  // :1928-1930
    if (node < 0 || node >= s->mesh.nodeCount) return fail(NF_E_ARGUMENT, "node out of range");
    *out = s->mesh.z[node];
    return NF_OK;
}
int nfo_set_damage_uniform(nfo_solver* s, double fracture_energy, const double* failure_strain) {  // :713-717
    const std::vector<double> gf(s->mesh.elementCount, fracture_energy);
    return nfo_set_damage(s, gf.data(), failure_strain);
}

int nfo_get_element_field(nfo_solver* s, int32_t field, double* out) {
    const int E = s->mesh.elementCount;
    switch (field) {
        case NF_ELEMENT_HG_R: std::memcpy(out, s->hgR.data(), sizeof(double) * E); return NF_OK;
        case NF_ELEMENT_HG_Z: std::memcpy(out, s->hgZ.data(), sizeof(double) * E); return NF_OK;
        case NF_ELEMENT_HG_STIFFNESS: std::memcpy(out, s->hourglassStiffness.data(), sizeof(double) * E); return NF_OK;
        case NF_ELEMENT_ERODED:
            for (int e = 0; e < E; e++) out[e] = (!s->gone.empty() && s->gone[e]) ? 1.0 : 0.0;
            return NF_OK;
        default: return fail(NF_E_ARGUMENT, "unknown element field");
    }
}
int nfo_get_state(nfo_solver* s, nf_state* o) {
    std::memset(o, 0, sizeof *o);
    o->time = s->time;
    o->dt = s->dt;
    o->dt_previous = s->dtPrevious;
    o->dt0 = s->dt0;
    o->mass_scale = s->massScale;
    o->steps = s->stepCount;
    o->node_count = s->mesh.nodeCount;
    o->element_count = s->mesh.elementCount;
    o->points_per_element = s->pointsPerElement;
    o->eroded_elements = s->erodedCount;
    o->inverted_elements = s->invertedCount;
    o->eroded_energy = s->erodedEnergy;
    if (s->hasWall) {
        o->wall_contact_count = s->wall.contactCount;
        o->wall_stuck_count = s->wall.stuckCount;
        o->wall_impulse = s->wall.impulse;
        o->wall_energy_loss = s->wall.energyLoss;
        o->wall_friction_dissipation = s->wall.frictionDissipation;
    }
    if (s->hasContact) {
        o->contact_pairs = s->contact.pairs;
        o->contact_energy = s->contact.energy;
        o->contact_dissipation = s->contact.dissipation;
        o->contact_max_penetration = s->contact.maxPenetration;
        o->contact_escaped_nodes = s->contact.escaped;
    }
    return NF_OK;
}
int nfo_get_audits(nfo_solver* s, nf_audits* o) {
    std::memset(o, 0, sizeof *o);
    const int N = s->mesh.nodeCount, E = s->mesh.elementCount, ppe = s->pointsPerElement;
    double p = 0.0;
    for (int i = 0; i < N; i++) p += s->mass[i] * s->vz[i];  // :1610-1614
    o->axial_momentum = p;
    p = 0.0;
    for (int i = 0; i < N; i++) p += s->mass[i] * s->vr[i];  // :1627-1631
    o->radial_momentum = p;
    double ke = 0.0;
    for (int i = 0; i < N; i++) ke += s->mass[i] * (s->vr[i] * s->vr[i] + s->vz[i] * s->vz[i]);
    o->kinetic_energy = 0.5 * ke;  // :1634-1640
    ke = 0.0;
    const double h = 0.5 * s->dt;  // :875-886
    for (int i = 0; i < N; i++) {
        const double invM = 1.0 / s->mass[i];
        const bool held = s->hasWall && s->onWall[i] && s->wall.holds(s->fz[i]);
        const double cvr = s->fixedR[i] ? s->vr[i] : s->vr[i] + h * s->fr[i] * invM;
        const double cvz = s->fixedZ[i] || held ? s->vz[i] : s->vz[i] + h * s->fz[i] * invM;
        ke += 0.5 * s->mass[i] * (cvr * cvr + cvz * cvz);
    }
    o->centred_kinetic_energy = ke;
    double se = 0.0;  // :1654-1663
    for (int e = 0; e < E; e++) {
        if (!s->gone.empty() && s->gone[e]) continue;
        for (int g = 0; g < ppe; g++) se += s->pointStrainEnergy(e, g);
    }
    o->strain_energy = se;
    double w = 0.0;  // :1700-1709
    for (int e = 0; e < E; e++)
        for (int g = 0; g < ppe; g++) w += s->wPlastic[e * ppe + g] * s->gaussVolume(e, g);
    o->plastic_dissipation = w;
    double hg = 0.0;  // :1807-1818
    if (!(s->integration == NF_FULL || s->hourglassCoefficient == 0.0)) {
        for (int e = 0; e < E; e++) {
            const double k = s->hourglassStiffness[e];
            if (k > 0.0) hg += 0.5 * (s->hgR[e] * s->hgR[e] + s->hgZ[e] * s->hgZ[e]) / k;
        }
    }
    o->hourglass_energy = hg;
    double fd = 0.0;  // :1730-1740
    if (s->damage) {
        for (int e = 0; e < E; e++)
            for (int g = 0; g < ppe; g++) {
                const int q = e * ppe + g;
                fd += s->damage->dissipation(q, s->epsP[q]) * s->gaussVolume(e, g);
            }
    }
    o->fracture_dissipation = fd;
    double mx = 0.0;  // :1849-1853
    int yielded = 0;
    for (double v : s->epsP) {
        if (v > mx) mx = v;
        if (v > 0.0) yielded++;
    }
    o->max_plastic_strain = mx;
    o->yielded_fraction = (double)yielded / (double)s->epsP.size();
    bool anyJc = false;  // :1832-1846
    for (const Material& m : s->materials) anyJc = anyJc || m.has_jc;
    if (anyJc) {
        double t = -HUGE_VAL;
        bool ok = true;
        for (int q = 0; q < (int)s->wPlastic.size(); q++) {
            const Material& m = s->materialAt(q / ppe);
            if (!m.has_jc) {
                ok = false;
                break;
            }
            t = jmax(t, m.jc.temperature(s->wPlastic[q], m.rho));
        }
        o->max_temperature = ok ? t : NAN;
    } else {
        o->max_temperature = NAN;
    }
    if (s->damage) {  // Damage.maximum / begunFraction / failedFraction :375-398
        double md = 0.0;
        int begun = 0, failed = 0;
        const int P = (int)s->epsP.size();
        for (int q = 0; q < P; q++) {
            const double d = s->damage->at(q, s->epsP[q]);
            if (d > md) md = d;
            if (s->damage->strength[q] > 0.0) begun++;
            if (d >= 1.0) failed++;
        }
        o->max_damage = md;
        o->damaged_fraction = (double)begun / P;
        o->failed_fraction = (double)failed / P;
    }
    return NF_OK;
}

// Contact pair set of the last step: for each surface node in surface order, its master
// segment as (node, segment end a, segment end b) or nothing; and the surface itself.
int nfo_contact_pairs(nfo_solver* s, int32_t capacity, int32_t* out, int32_t* count) {
    if (!s->hasContact) {
        *count = 0;
        return NF_OK;
    }
    const Contact& c = s->contact;
    int k = 0;
    for (int i = 0; i < (int)c.nodes.size(); i++) {
        if (c.lastMaster[i] < 0) continue;
        if (k < capacity) {
            out[3 * k] = c.nodes[i];
            out[3 * k + 1] = c.seg[2 * c.lastMaster[i]];
            out[3 * k + 2] = c.seg[2 * c.lastMaster[i] + 1];
        }
        k++;
    }
    *count = k;
    return NF_OK;
}

// The contact surface's segments (a, b) in surface order.
int nfo_contact_surface(nfo_solver* s, int32_t capacity, int32_t* out, int32_t* count) {
    if (!s->hasContact) {
        *count = 0;
        return NF_OK;
    }
    // The surface the next step will use: rebuilt from the contact's dead set, as adopt()
    // would, without touching the contact's state.
    std::vector<int> seg, owner, nodes;
    s->mesh.surface(s->contact.dead, seg, owner, nodes);
    const int S = (int)owner.size();
    for (int k = 0; k < S && k < capacity; k++) {
        out[2 * k] = seg[2 * k];
        out[2 * k + 1] = seg[2 * k + 1];
    }
    *count = S;
    return NF_OK;
}

}  // extern "C"

// The shared transcendental functions evaluated on the host, for the math self-checks.
extern "C" int nfo_math_eval(int32_t fn, const double* x, const double* y, double* out, int32_t n) {
    for (int i = 0; i < n; i++) {
        double r;
        switch (fn) {
            case 0: r = nfm::log(x[i]); break;
            case 1: r = nfm::exp(x[i]); break;
            case 2: r = nfm::pow(x[i], y[i]); break;
            case 3: r = nfm::atan2(x[i], y[i]); break;
            case 4: r = nfm::sin(x[i]); break;
            case 5: r = nfm::cos(x[i]); break;
            case 6: r = nfm::hypot(x[i], y[i]); break;
            case 7: r = std::sqrt(x[i]); break;
            case 8: r = x[i] / y[i]; break;
            case 9: r = nfm::log1p(x[i]); break;
            case 10: r = nfm::pow_ordinary(x[i], y[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}); break;
            case 11: r = nfm::log_ordinary(x[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}); break;
            case 12:
            case 13: {
                double lx;
                const double pw = nfm::pow_log_ordinary(x[i], y[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}, lx);
                r = fn == 12 ? pw : lx;
                break;
            }
            case 14:
            case 15: {  // pow_ordinary(x, y) and log_ordinary(y) side by side
                double pw, lg;
                nfm::pow_and_log_ordinary(x[i], y[i], y[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}, pw, lg);
                r = fn == 14 ? pw : lg;
                break;
            }
            case 16:
            case 17: {  // pow_ordinary(x, y) and pow_ordinary(y, x) side by side
                double p1, p2;
                nfm::pow_pair_ordinary(x[i], y[i], y[i], x[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}, p1, p2);
                r = fn == 16 ? p1 : p2;
                break;
            }
            case 24:
            case 25: {  // the Newton loop's pow(x, y) and log(y) for 0 < y < 1, no range selects
                double pw, lg;
                nfm::pow_and_log_ordinary_y01(x[i], y[i], y[i], nfm::Tables{NFM_LOG_TBL, NFM_EXP_TBL}, pw, lg);
                r = fn == 24 ? pw : lg;
                break;
            }
            default: return fail(NF_E_ARGUMENT, "unknown function");
        }
        out[i] = r;
    }
    return NF_OK;
}

// J2.update at independent material points (see nf_point_path in include/neofiz.h).
extern "C" int nfo_point_path(const nf_material* m, int32_t formulation, int32_t paths,
                              int32_t steps, const double* inc, const double* dmgIn, double* out) {
    g_error.clear();
    int rc = check_material(m);
    if (rc) return rc;
    if (paths < 1 || steps < 1 || !inc || !out) return fail(NF_E_ARGUMENT, "empty path");
    const Material mat = to_material(*m);
    const double lam = formulation == NF_PLANE_STRESS
                           ? 2.0 * mat.lambda() * mat.mu() / (mat.lambda() + 2.0 * mat.mu())
                           : mat.lambda();
    Flow fl;
    fl.sy0 = mat.sy;
    fl.hIso = mat.hiso;
    fl.hKin = mat.hkin;
    fl.jc = mat.has_jc ? &mat.jc : nullptr;
    fl.thermalCoefficient = mat.has_jc ? mat.jc.thermalCoefficient(mat.rho) : 0.0;
    Damage d;
    if (dmgIn) {
        d.fractureEnergy.resize(paths);
        d.youngsModulus.assign(paths, mat.E);
        d.failureStrain.resize(paths);
        d.geometry.resize(5 * (size_t)paths);
        for (int p = 0; p < paths; p++) {
            const double* g = dmgIn + 7 * (size_t)p;
            d.failureStrain[p] = g[0];
            d.fractureEnergy[p] = g[1];
            for (int k = 0; k < 5; k++) d.geometry[5 * (size_t)p + k] = g[2 + k];
        }
        d.strength.assign(paths, 0.0);
        d.softening.assign(paths, 0.0);
        d.width.assign(paths, 0.0);
    }
    std::vector<double> sig(4 * (size_t)paths, 0.0), back(4 * (size_t)paths, 0.0), epsP(paths, 0.0), work(paths, 0.0);
    for (int p = 0; p < paths; p++) {
        for (int k = 0; k < steps; k++) {
            const double* x = inc + 5 * ((size_t)p * steps + k);
            const double dg = j2_update(sig.data(), back.data(), epsP.data(), work.data(), p, x[0], x[1], x[2],
                                        x[3], lam, mat.mu(), fl, x[4], dmgIn ? &d : nullptr);
            double* o = out + 14 * ((size_t)p * steps + k);
            for (int c = 0; c < 4; c++) o[c] = sig[4 * (size_t)p + c];
            for (int c = 0; c < 4; c++) o[4 + c] = back[4 * (size_t)p + c];
            o[8] = epsP[p];
            o[9] = work[p];
            o[10] = dg;
            o[11] = dmgIn ? d.strength[p] : 0.0;
            o[12] = dmgIn ? d.softening[p] : 0.0;
            o[13] = dmgIn ? d.width[p] : 0.0;
        }
    }
    return NF_OK;
}

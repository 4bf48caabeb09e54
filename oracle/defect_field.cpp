// TEST INFRASTRUCTURE -- the oracle's defect field. Nothing in the product links, loads or
// calls this file.
//
// A serial CPU restatement of core/DefectField.java, core/Weibull.java (the members the field
// uses) and core/Normal.java, written from the Java in its evaluation order (paths relative
// to src/main/java/org/neofiz/). It exports the defect-field entry points of
// include/neofiz.h under the prefix nfo_. Fidelity rules are those of neofiz_oracle.cpp:
// left-to-right association, no fused multiply-adds (-ffp-contract=off), Java's Math.min /
// Math.max, and nf_math.h's log, log1p, exp, pow and cos in place of java.lang.Math's.
//
// Pins (tests/test_oracle_defect_field.py): the values of core/NormalTest.java and
// core/WeibullTest.java, scipy's erfc and Weibull quantiles, the published SplitMix64
// outputs, a brute-force restatement of the smoothing at a few points, and the statistical
// and mesh-independence properties of core/DefectFieldTest.java.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../include/neofiz.h"
#include "../neofiz_amd/csrc/nf_math.h"

int nfo_report(int code, const std::string& msg);  // neofiz_oracle.cpp

namespace {

// Math.min / Math.max on doubles: NaN wins; the signed zeros are ordered.
double jmin(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return std::signbit(a) ? a : b;
    return a <= b ? a : b;
}
double jmax(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return std::signbit(a) ? b : a;
    return a >= b ? a : b;
}
// (int) of a double: NaN to 0, saturating at the int range.
int32_t jint(double x) {
    if (x != x) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return (-2147483647 - 1);
    return (int32_t)x;
}

const double PI = 3.141592653589793;

// ---- core/Normal.java:17-130 ----
const double SQRT_HALF = 0.70710678118654752440;
const double INV_SQRT_PI = 5.6418958354775628695e-1;
const double THRESHOLD = 0.46875;
const double A[5] = {3.16112374387056560e0, 1.13864154151050156e2, 3.77485237685302021e2,
                     3.20937758913846947e3, 1.85777706184603153e-1};
const double B[4] = {2.36012909523441209e1, 2.44024637934444173e2, 1.28261652607737228e3,
                     2.84423683343917062e3};
const double C[9] = {5.64188496988670089e-1, 8.88314979438837594e0, 6.61191906371416295e1,
                     2.98635138197400131e2, 8.81952221241769090e2, 1.71204761263407058e3,
                     2.05107837782607147e3, 1.23033935479799725e3, 2.15311535474403846e-8};
const double D[8] = {1.57449261107098347e1, 1.17693950891312499e2, 5.37181101862009858e2,
                     1.62138957456669019e3, 3.29079923573345963e3, 4.36261909014324716e3,
                     3.43936767414372164e3, 1.23033935480374942e3};
const double P[6] = {3.05326634961232344e-1, 3.60344899949804439e-1, 1.25781726111229246e-1,
                     1.60837851487422766e-2, 6.58749161529837803e-4, 1.63153871373020978e-2};
const double Q[5] = {2.56852019228982242e0, 1.87295284992346047e0, 5.27905102951428412e-1,
                     6.05183413124413191e-2, 2.33520497626869185e-3};

double small_erf(double sq) {  // :92-101
    double num = A[4] * sq;
    double den = sq;
    for (int i = 0; i < 3; i++) {
        num = (num + A[i]) * sq;
        den = (den + B[i]) * sq;
    }
    return (num + A[3]) / (den + B[3]);
}

double cody_erfc(double x) {  // :55-89
    double y = std::fabs(x);
    double result;
    if (y <= THRESHOLD) {
        return 1.0 - x * small_erf(y * y);
    } else if (y <= 4.0) {
        double num = C[8] * y;
        double den = y;
        for (int i = 0; i < 7; i++) {
            num = (num + C[i]) * y;
            den = (den + D[i]) * y;
        }
        result = (num + C[7]) / (den + D[7]);
    } else {
        double inv = 1.0 / (y * y);
        double num = P[5] * inv;
        double den = inv;
        for (int i = 0; i < 4; i++) {
            num = (num + P[i]) * inv;
            den = (den + Q[i]) * inv;
        }
        result = inv * (num + P[4]) / (den + Q[4]);
        result = (INV_SQRT_PI - result) / y;
    }
    double trunc = jint(y * 16.0) / 16.0;
    double del = (y - trunc) * (y + trunc);
    result = result * (nfm::exp(-trunc * trunc) * nfm::exp(-del));
    return x < 0.0 ? 2.0 - result : result;
}

double cody_erf(double x) {  // :44-49
    double a = std::fabs(x);
    if (a <= THRESHOLD) return x * small_erf(a * a);
    double e = 1.0 - cody_erfc(a);
    return x < 0.0 ? -e : e;
}

double cdf(double x) { return 0.5 * cody_erfc(-x * SQRT_HALF); }      // :30-32
double survival(double x) { return 0.5 * cody_erfc(x * SQRT_HALF); }  // :35-37

// ---- the field, core/DefectField.java ----
const int CELLS_PER_LENGTH = 3;
const double KERNEL_RADII = 4.0;
const double CLAMP_SIGMA = 9.0;
const int64_t MAX_CACHE_NODES = int64_t(1) << 24;

// Failure codes of the Weibull record's checks, in the order the JVM would meet them.
enum { OK = 0, BAD_VOLUME = 1, BAD_COUNT = 2, BAD_SCALE = 3, BAD_PROBABILITY = 4 };

const char* message(int code) {
    switch (code) {
        case BAD_VOLUME: return "volume must be positive";
        case BAD_COUNT: return "sample count must be positive";
        case BAD_SCALE: return "scale must be positive";
        default: return "probability must lie in [0, 1]";
    }
}

uint64_t mix(uint64_t z) {  // :334-338
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double unit_open(uint64_t bits) { return ((double)(bits >> 11) + 0.5) * 0x1.0p-53; }  // :345-347

struct Field {
    double scale, modulus, v0, length;
    int64_t seed;
    double cell, sigmaCells;
    int radius;

    double white(int32_t i, int32_t j, int64_t s) const {  // :325-332
        uint64_t h = mix((uint64_t)s + 0x9E3779B97F4A7C15ULL);
        h = mix(h ^ ((uint64_t)(int64_t)i * 0xBF58476D1CE4E5B9ULL));
        h = mix(h ^ ((uint64_t)(int64_t)j * 0x94D049BB133111EBULL));
        double u1 = unit_open(h);
        double u2 = unit_open(mix(h ^ 0xD6E8FEB86659FD93ULL));
        return std::sqrt(-2.0 * nfm::log(u1)) * nfm::cos(2.0 * PI * u2);
    }

    double effective_volume(double rc, double area) const {  // :204-208
        double inPlane = jmax(1.0, area / (length * length));
        double hoop = jmax(1.0, 2.0 * PI * rc / length);
        return length * length * length * inPlane * hoop;
    }

    double slab_volume(double area, double depth) const {  // no counterpart (plane strain)
        double inPlane = jmax(1.0, area / (length * length));
        double through = jmax(1.0, depth / length);
        return length * length * length * inPlane * through;
    }

    // transform :261-271 with Weibull.forVolume / minimumOf / quantile /
    // quantileFromSurvival (core/Weibull.java:114-131, :141-169) and the record's check.
    int transform(double g, double volume, double& out) const {
        if (!(volume > 0.0)) return BAD_VOLUME;
        double n = volume / v0;
        if (!(n > 0.0)) return BAD_COUNT;
        double localScale = scale * nfm::pow(n, -1.0 / modulus);
        if (!(localScale > 0.0)) return BAD_SCALE;
        double clamped = jmax(-CLAMP_SIGMA, jmin(CLAMP_SIGMA, g));
        if (clamped <= 0.0) {
            double p = cdf(clamped);
            if (!(p >= 0.0 && p <= 1.0)) return BAD_PROBABILITY;
            out = localScale * nfm::pow(-nfm::log1p(-p), 1.0 / modulus);
        } else {
            double s = survival(clamped);
            if (!(s >= 0.0 && s <= 1.0)) return BAD_PROBABILITY;
            out = localScale * nfm::pow(-nfm::log(s), 1.0 / modulus);
        }
        return OK;
    }
};

// The memoised white noise of cacheFor :351-384 over a box of points; an empty grid is the
// uncached path.
struct Noise {
    const Field* f;
    int64_t seed;
    int32_t i0 = 0, j0 = 0, ni = 0, nj = 0;
    std::vector<double> v;

    Noise(const Field* field, int64_t s) : f(field), seed(s) {}

    double at(int32_t i, int32_t j) const {
        if (v.empty()) return f->white(i, j, seed);
        int32_t a = i - i0;
        int32_t b = j - j0;
        if (a < 0 || a >= ni || b < 0 || b >= nj) return f->white(i, j, seed);
        return v[(size_t)b * ni + a];
    }
};

Noise cache_for(const Field& f, int64_t seed, int32_t count, const double* r, const double* z,
                bool uncached) {
    Noise noise(&f, seed);
    if (uncached || count == 0) return noise;
    double rMin = 1.7976931348623157e308, rMax = -1.7976931348623157e308;
    double zMin = 1.7976931348623157e308, zMax = -1.7976931348623157e308;
    for (int n = 0; n < count; n++) {
        rMin = jmin(rMin, r[n]);
        rMax = jmax(rMax, r[n]);
        zMin = jmin(zMin, z[n]);
        zMax = jmax(zMax, z[n]);
    }
    int32_t i0 = jint(std::floor(rMin / f.cell)) - f.radius;
    int32_t j0 = jint(std::floor(zMin / f.cell)) - f.radius;
    int32_t ni = jint(std::floor(rMax / f.cell)) + f.radius + 2 - i0;
    int32_t nj = jint(std::floor(zMax / f.cell)) + f.radius + 2 - j0;
    int64_t nodes = (int64_t)ni * nj;
    if (nodes <= 0 || nodes > MAX_CACHE_NODES) return noise;
    noise.v.resize((size_t)nodes);
    for (int32_t b = 0; b < nj; b++)
        for (int32_t a = 0; a < ni; a++) noise.v[(size_t)b * ni + a] = f.white(i0 + a, j0 + b, seed);
    noise.i0 = i0;
    noise.j0 = j0;
    noise.ni = ni;
    noise.nj = nj;
    return noise;
}

double smooth(const Field& f, double r, double z, const Noise& noise) {  // :289-319
    const double gr = r / f.cell;
    const double gz = z / f.cell;
    const int32_t i0 = jint(std::floor(gr)) - f.radius;
    const int32_t j0 = jint(std::floor(gz)) - f.radius;
    const int taps = 2 * f.radius + 2;
    const double twoSigmaSq = 2.0 * f.sigmaCells * f.sigmaCells;
    std::vector<double> wr(taps), wz(taps);
    double sumRSq = 0.0;
    double sumZSq = 0.0;
    for (int k = 0; k < taps; k++) {
        double dr = gr - (double)(i0 + k);
        double dz = gz - (double)(j0 + k);
        wr[k] = nfm::exp(-dr * dr / twoSigmaSq);
        wz[k] = nfm::exp(-dz * dz / twoSigmaSq);
        sumRSq += wr[k] * wr[k];
        sumZSq += wz[k] * wz[k];
    }
    double sum = 0.0;
    for (int b = 0; b < taps; b++) {
        double row = 0.0;
        for (int a = 0; a < taps; a++) row += wr[a] * noise.at(i0 + a, j0 + b);
        sum += wz[b] * row;
    }
    return sum / std::sqrt(sumRSq * sumZSq);
}

// QuadMesh.signedArea :78-86 and the area centroid :96-118.
void element_geometry(const int32_t* conn, const double* r, const double* z, int e, double& rc,
                      double& zc, double& area) {
    // The three loops of the reference accumulate the same cross products in the same order,
    // so one loop gives all three.
    const int b = e * 4;
    double a = 0.0, mr = 0.0, mz = 0.0;
    for (int k = 0; k < 4; k++) {
        int i = conn[b + k];
        int j = conn[b + (k + 1) % 4];
        double cross = r[i] * z[j] - r[j] * z[i];
        a += cross;
        mr += (r[i] + r[j]) * cross;
        mz += (z[i] + z[j]) * cross;
    }
    rc = mr / (3.0 * a);
    zc = mz / (3.0 * a);
    area = 0.5 * a;
}

}  // namespace

struct nfo_defect {
    Field f;
};

extern "C" {

int nfo_defect_create(nfo_defect** out, const nf_defect_params* p) {
    if (!out || !p) return nfo_report(NF_E_ARGUMENT, "null argument");
    if (!(p->scale > 0.0)) return nfo_report(NF_E_ARGUMENT, "scale must be positive");
    if (!(p->modulus > 0.0)) return nfo_report(NF_E_ARGUMENT, "modulus must be positive");
    if (!(p->reference_volume > 0.0))
        return nfo_report(NF_E_ARGUMENT, "reference volume must be positive");
    if (!(p->correlation_length > 0.0))
        return nfo_report(NF_E_ARGUMENT, "correlation length must be positive");
    if (!std::isfinite(p->correlation_length))
        return nfo_report(NF_E_ARGUMENT, "correlation length must be finite");
    Field f;
    f.scale = p->scale;
    f.modulus = p->modulus;
    f.v0 = p->reference_volume;
    f.length = p->correlation_length;
    f.seed = p->seed;
    f.cell = f.length / CELLS_PER_LENGTH;  // :137-141
    f.sigmaCells = (f.length / std::sqrt(2.0)) / f.cell;
    f.radius = jint(std::ceil(KERNEL_RADII * f.sigmaCells));
    *out = new nfo_defect{f};
    return NF_OK;
}

int nfo_defect_destroy(nfo_defect* d) {
    delete d;
    return NF_OK;
}

int nfo_defect_sample_onto(nfo_defect* d, int32_t seed_count, const int64_t* seeds,
                           int32_t node_count, const double* r, const double* z,
                           int32_t element_count, const int32_t* conn, int32_t rule,
                           double depth, int32_t flags, double* out) {
    if (!d) return nfo_report(NF_E_ARGUMENT, "null handle");
    if (seeds == nullptr) seed_count = 1;
    if (seed_count < 0 || node_count < 0 || element_count < 0)
        return nfo_report(NF_E_ARGUMENT, "negative count");
    if (rule != NF_DEFECT_RING && rule != NF_DEFECT_SLAB)
        return nfo_report(NF_E_ARGUMENT, "unknown volume rule");
    if (rule == NF_DEFECT_SLAB && !(depth > 0.0))
        return nfo_report(NF_E_ARGUMENT, "depth must be positive");
    for (int64_t k = 0; k < 4 * (int64_t)element_count; k++)
        if (conn[k] < 0 || conn[k] >= node_count)
            return nfo_report(NF_E_ARGUMENT, "connectivity names a node outside the mesh");
    const Field& f = d->f;
    for (int32_t s = 0; s < seed_count; s++) {
        const int64_t seed = seeds ? seeds[s] : f.seed;
        Noise noise = cache_for(f, seed, node_count, r, z, (flags & NF_DEFECT_UNCACHED) != 0);
        for (int e = 0; e < element_count; e++) {  // :239-248
            double rc, zc, area;
            element_geometry(conn, r, z, e, rc, zc, area);
            double g = smooth(f, rc, zc, noise);
            double v = rule == NF_DEFECT_RING ? f.effective_volume(rc, area)
                                              : f.slab_volume(area, depth);
            int code = f.transform(g, v, out[(size_t)s * element_count + e]);
            if (code != OK) return nfo_report(NF_E_ARGUMENT, message(code));
        }
    }
    return NF_OK;
}

int nfo_defect_points(nfo_defect* d, int32_t function, int32_t n, const double* r,
                      const double* z, const double* w, double* out) {
    if (!d) return nfo_report(NF_E_ARGUMENT, "null handle");
    if (n < 0) return nfo_report(NF_E_ARGUMENT, "negative count");
    if (function < NF_DEFECT_GAUSSIAN || function > NF_DEFECT_WHITE)
        return nfo_report(NF_E_ARGUMENT, "unknown function");
    const Field& f = d->f;
    if (function == NF_DEFECT_WHITE) {
        for (int k = 0; k < n; k++) {
            if (!(r[k] == std::floor(r[k]) && z[k] == std::floor(z[k]) && r[k] >= -2147483648.0 &&
                  r[k] <= 2147483647.0 && z[k] >= -2147483648.0 && z[k] <= 2147483647.0))
                return nfo_report(NF_E_ARGUMENT, "white noise is indexed by int32 pairs");
            out[k] = f.white((int32_t)r[k], (int32_t)z[k], f.seed);
        }
        return NF_OK;
    }
    Noise noise(&f, f.seed);
    if (function <= NF_DEFECT_STRAIN_VOLUME) noise = cache_for(f, f.seed, n, r, z, false);
    for (int k = 0; k < n; k++) {
        int code = OK;
        switch (function) {
            case NF_DEFECT_GAUSSIAN: out[k] = smooth(f, r[k], z[k], noise); break;
            case NF_DEFECT_STRAIN_AREA:
                code = f.transform(smooth(f, r[k], z[k], noise), f.effective_volume(r[k], w[k]), out[k]);
                break;
            case NF_DEFECT_STRAIN_VOLUME:
                code = f.transform(smooth(f, r[k], z[k], noise), w[k], out[k]);
                break;
            case NF_DEFECT_EFFECTIVE_VOLUME: out[k] = f.effective_volume(r[k], w[k]); break;
            default: code = f.transform(r[k], w[k], out[k]); break;
        }
        if (code != OK) return nfo_report(NF_E_ARGUMENT, message(code));
    }
    return NF_OK;
}

int nfo_defect_timing(nfo_defect* d, double* ms) {
    if (!d || !ms) return nfo_report(NF_E_ARGUMENT, "null argument");
    ms[0] = ms[1] = ms[2] = 0.0;
    return NF_OK;
}

int nfo_normal_eval(int32_t function, int32_t n, const double* x, double* out) {
    if (n < 0) return nfo_report(NF_E_ARGUMENT, "negative count");
    for (int k = 0; k < n; k++) {
        switch (function) {
            case 0: out[k] = cody_erfc(x[k]); break;
            case 1: out[k] = cody_erf(x[k]); break;
            case 2: out[k] = cdf(x[k]); break;
            case 3: out[k] = survival(x[k]); break;
            default: return nfo_report(NF_E_ARGUMENT, "unknown function");
        }
    }
    return NF_OK;
}

}  // extern "C"

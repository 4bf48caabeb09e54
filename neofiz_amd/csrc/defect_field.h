// Defect-field seeding on the device: DefectField.sampleOnto (core/DefectField.java:239-332)
// with Weibull's volume scaling and quantiles (core/Weibull.java:114-169) and Cody's erfc
// (core/Normal.java:30-101). Included by neofiz_hip.hip after its helpers.
//
// The field value at (r, z) is a 20 x 20 Gaussian-weighted sum of white noise on the grid of
// spacing l/3 around it, normalised by the root sum of squares of the weights; the white
// noise at grid index (i, j) is Box-Muller on two SplitMix64 hashes of (seed, i, j). A
// failure strain is that value mapped through the normal cdf onto the Weibull quantile of the
// element's effective volume. sampleOnto is three kernels on the handle's stream:
//   k_df_bounds  bounding box of the nodes, two passes (Math.min / Math.max order-free: min
//                and max with Java's NaN rule are associative and commutative);
//   k_df_noise   the white noise on the box's grid, for a chunk of seeds (the reference's
//                memoisation, :351-384; the values are the hash's, so caching is invisible);
//   k_df_sample  one thread per (element, seed): area centroid and signed area
//                (mesh/QuadMesh.java:78-118), the 20 + 20 weights, the 400-tap sum in the
//                reference's order (a inner, b outer), the copula transform.
// Every operation is the reference's, in its order, with no fused multiply-add, and the
// transcendental functions are nf_math.h's, so the result is bitwise the oracle's.
//
// Layout in HBM: node coordinates r[N], z[N]; connectivity int4 conn[E]; the noise grid
// grid[seed][j][i] (i along r, contiguous, so a thread's 20 taps of one row are one 160-byte
// run and a wave of neighbouring elements shares rows in L2); out[seed][element].
#pragma once

namespace dfd {

constexpr int CELLS_PER_LENGTH = 3;  // :95
constexpr double KERNEL_RADII = 4.0; // :97
constexpr double CLAMP_SIGMA = 9.0;  // :104
// sigmaCells = (l / sqrt 2) / (l / 3) = 2.1213..., so radius = ceil(4 sigmaCells) = 9 for
// every finite positive l (the product stays far from an integer under rounding), and a
// window has 2 radius + 2 = 20 taps. The constructor on the host checks this.
constexpr int RADIUS = 9;
constexpr int TAPS = 2 * RADIUS + 2;
constexpr int BOUNDS_THREADS = 256;
constexpr int BOX_BLOCKS = 512;

enum { OK = 0, BAD_VOLUME = 1, BAD_COUNT = 2, BAD_SCALE = 3, BAD_PROBABILITY = 4 };
enum { SITES_MESH = 0, SITES_POINTS = 1 };
enum { OUT_GAUSSIAN = 0, OUT_RING = 1, OUT_VOLUME = 2, OUT_SLAB = 3 };

// The constructor's derived constants (:137-141) and the per-call ones Java recomputes
// identically each time (twoSigmaSq :295; -1/m and 1/m, core/Weibull.java:118, :130, :145).
struct Params {
    double scale, modulus, v0, length;
    double cell, sigmaCells, twoSigmaSq;
    double negInvM, invM;
    double depth;  // the SLAB rule's depth
};

__host__ __device__ __forceinline__ int32_t jint(double x) {  // Java (int) of a double
    if (x != x) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return -2147483647 - 1;
    return (int32_t)x;
}

__device__ __forceinline__ double dmin(double a, double b) {  // Math.min
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return __builtin_signbit(a) ? a : b;
    return a <= b ? a : b;
}
__device__ __forceinline__ double dmax(double a, double b) {  // Math.max
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return __builtin_signbit(a) ? b : a;
    return a >= b ? a : b;
}

// ---- SplitMix64 finaliser and the white noise, :325-347 ----
__host__ __device__ __forceinline__ uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// The seed's own hash, the first line of white(): the same for every grid index.
__host__ __device__ __forceinline__ uint64_t seed_hash(int64_t seed) {
    return mix((uint64_t)seed + 0x9E3779B97F4A7C15ULL);
}

__device__ __forceinline__ double unit_open(uint64_t bits) {
    return ((double)(bits >> 11) + 0.5) * 0x1.0p-53;
}

__device__ __forceinline__ double white(uint64_t h0, int32_t i, int32_t j) {
    uint64_t h = mix(h0 ^ ((uint64_t)(int64_t)i * 0xBF58476D1CE4E5B9ULL));
    h = mix(h ^ ((uint64_t)(int64_t)j * 0x94D049BB133111EBULL));
    const double u1 = unit_open(h);
    const double u2 = unit_open(mix(h ^ 0xD6E8FEB86659FD93ULL));
    return __builtin_sqrt(-2.0 * nfm::log(u1)) * nfm::cos(2.0 * 3.141592653589793 * u2);
}

// ---- Cody's erfc, core/Normal.java:55-101 (coefficients :104-129) ----
__device__ __forceinline__ double small_erf(double sq) {
    double num = 1.85777706184603153e-1 * sq;
    double den = sq;
    num = (num + 3.16112374387056560e0) * sq;
    den = (den + 2.36012909523441209e1) * sq;
    num = (num + 1.13864154151050156e2) * sq;
    den = (den + 2.44024637934444173e2) * sq;
    num = (num + 3.77485237685302021e2) * sq;
    den = (den + 1.28261652607737228e3) * sq;
    return (num + 3.20937758913846947e3) / (den + 2.84423683343917062e3);
}

__device__ __forceinline__ double erfc(double x) {
    const double y = __builtin_fabs(x);
    if (y <= 0.46875) return 1.0 - x * small_erf(y * y);
    double result;
    if (y <= 4.0) {
        constexpr double C[9] = {5.64188496988670089e-1, 8.88314979438837594e0, 6.61191906371416295e1,
                                 2.98635138197400131e2, 8.81952221241769090e2, 1.71204761263407058e3,
                                 2.05107837782607147e3, 1.23033935479799725e3, 2.15311535474403846e-8};
        constexpr double D[8] = {1.57449261107098347e1, 1.17693950891312499e2, 5.37181101862009858e2,
                                 1.62138957456669019e3, 3.29079923573345963e3, 4.36261909014324716e3,
                                 3.43936767414372164e3, 1.23033935480374942e3};
        double num = C[8] * y;
        double den = y;
#pragma unroll
        for (int i = 0; i < 7; i++) {
            num = (num + C[i]) * y;
            den = (den + D[i]) * y;
        }
        result = (num + C[7]) / (den + D[7]);
    } else {
        constexpr double P[6] = {3.05326634961232344e-1, 3.60344899949804439e-1, 1.25781726111229246e-1,
                                 1.60837851487422766e-2, 6.58749161529837803e-4, 1.63153871373020978e-2};
        constexpr double Q[5] = {2.56852019228982242e0, 1.87295284992346047e0, 5.27905102951428412e-1,
                                 6.05183413124413191e-2, 2.33520497626869185e-3};
        const double inv = 1.0 / (y * y);
        double num = P[5] * inv;
        double den = inv;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            num = (num + P[i]) * inv;
            den = (den + Q[i]) * inv;
        }
        result = inv * (num + P[4]) / (den + Q[4]);
        result = (5.6418958354775628695e-1 - result) / y;
    }
    const double trunc = jint(y * 16.0) / 16.0;
    const double del = (y - trunc) * (y + trunc);
    result = result * (nfm::exp(-trunc * trunc) * nfm::exp(-del));
    return x < 0.0 ? 2.0 - result : result;
}

__device__ __forceinline__ double erf(double x) {  // :47-52
    const double a = __builtin_fabs(x);
    if (a <= 0.46875) return x * small_erf(a * a);
    const double e = 1.0 - erfc(a);
    return x < 0.0 ? -e : e;
}

__device__ __forceinline__ double cdf(double x) { return 0.5 * erfc(-x * 0.70710678118654752440); }
__device__ __forceinline__ double survival(double x) { return 0.5 * erfc(x * 0.70710678118654752440); }

// ---- volumes and the copula ----
__device__ __forceinline__ double effective_volume(const Params& p, double rc, double area) {  // :204-208
    const double inPlane = dmax(1.0, area / (p.length * p.length));
    const double hoop = dmax(1.0, 2.0 * 3.141592653589793 * rc / p.length);
    return p.length * p.length * p.length * inPlane * hoop;
}

// The plane-strain rule (no counterpart in the reference): correlation lengths through the
// depth in place of around the circumference.
__device__ __forceinline__ double slab_volume(const Params& p, double area) {
    const double inPlane = dmax(1.0, area / (p.length * p.length));
    const double through = dmax(1.0, p.depth / p.length);
    return p.length * p.length * p.length * inPlane * through;
}

// transform :261-271 through Weibull.forVolume -> minimumOf -> the record's check ->
// quantile / quantileFromSurvival. A failed check returns its code (the Java exception).
__device__ __forceinline__ int transform(const Params& p, double g, double volume, double& out) {
    if (!(volume > 0.0)) return BAD_VOLUME;
    const double n = volume / p.v0;
    if (!(n > 0.0)) return BAD_COUNT;
    const double localScale = p.scale * nfm::pow(n, p.negInvM);
    if (!(localScale > 0.0)) return BAD_SCALE;
    const double clamped = dmax(-CLAMP_SIGMA, dmin(CLAMP_SIGMA, g));
    double prob, base;
    if (clamped <= 0.0) {
        prob = cdf(clamped);
        base = -nfm::log1p(-prob);
    } else {
        prob = survival(clamped);
        base = -nfm::log(prob);
    }
    if (!(prob >= 0.0 && prob <= 1.0)) return BAD_PROBABILITY;
    out = localScale * nfm::pow(base, p.invM);
    return OK;
}

// ---- the smoothing, :289-319 ----
// The memoised grid: noise of index (i, j) at grid[(j - j0) ni + (i - i0)], or the hash when
// (i, j) falls outside the box (or when there is no grid).
struct Grid {
    const double* v;  // this seed's grid, or nullptr
    int32_t i0, j0, ni, nj;
};

template <bool CACHED>
__device__ __forceinline__ double smooth(const Params& p, double r, double z, uint64_t h0, const Grid& g) {
    const double gr = r / p.cell;
    const double gz = z / p.cell;
    const int32_t i0 = jint(__builtin_floor(gr)) - RADIUS;
    const int32_t j0 = jint(__builtin_floor(gz)) - RADIUS;
    double wr[TAPS], wz[TAPS];
    double sumRSq = 0.0;
    double sumZSq = 0.0;
#pragma unroll
    for (int k = 0; k < TAPS; k++) {
        const double dr = gr - (double)(i0 + k);
        const double dz = gz - (double)(j0 + k);
        wr[k] = nfm::exp(-dr * dr / p.twoSigmaSq);
        wz[k] = nfm::exp(-dz * dz / p.twoSigmaSq);
        sumRSq += wr[k] * wr[k];
        sumZSq += wz[k] * wz[k];
    }
    double sum = 0.0;
    const int64_t a0 = (int64_t)i0 - g.i0;
    const int64_t b0 = (int64_t)j0 - g.j0;
    const bool inside = CACHED && a0 >= 0 && a0 + TAPS <= g.ni && b0 >= 0 && b0 + TAPS <= g.nj;
    // Both loops are unrolled where the weights are read, so wr and wz stay in registers.
    if (inside) {
        const double* base = g.v + b0 * g.ni + a0;
#pragma unroll
        for (int b = 0; b < TAPS; b++) {
            const double* row_v = base + (int64_t)b * g.ni;
            double row = 0.0;
#pragma unroll
            for (int a = 0; a < TAPS; a++) row += wr[a] * row_v[a];
            sum += wz[b] * row;
        }
    } else {
        // Taps outside the grid hash on demand. The row loop is not unrolled (each row inlines
        // twenty hashes), so its weight is recomputed: the same expression, the same bits.
        for (int b = 0; b < TAPS; b++) {
            const double dz = gz - (double)(j0 + b);
            const double wzb = nfm::exp(-dz * dz / p.twoSigmaSq);
            double row = 0.0;
#pragma unroll
            for (int a = 0; a < TAPS; a++) {
                const int64_t aa = a0 + a, bb = b0 + b;
                const bool hit = CACHED && aa >= 0 && aa < g.ni && bb >= 0 && bb < g.nj;
                const double n = hit ? g.v[bb * g.ni + aa] : white(h0, i0 + a, j0 + b);
                row += wr[a] * n;
            }
            sum += wzb * row;
        }
    }
    return sum / __builtin_sqrt(sumRSq * sumZSq);
}

// ---- kernels ----

// Math.min / Math.max over the coordinates in two passes: BOX_BLOCKS workgroups each reduce
// a grid-strided share to one partial [rMin, rMax, zMin, zMax], then one workgroup reduces the
// partials. dmin and dmax are associative and commutative (NaN wins either way round; -0 is
// below +0 either way round), so the result is the serial loop's.
__device__ __forceinline__ void box_reduce(double (*s)[BOUNDS_THREADS], double v0, double v1, double v2,
                                           double v3, double* __restrict__ dst) {
    s[0][threadIdx.x] = v0;
    s[1][threadIdx.x] = v1;
    s[2][threadIdx.x] = v2;
    s[3][threadIdx.x] = v3;
    __syncthreads();
    for (int h = BOUNDS_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            const int o = threadIdx.x + h;
            s[0][threadIdx.x] = dmin(s[0][threadIdx.x], s[0][o]);
            s[1][threadIdx.x] = dmax(s[1][threadIdx.x], s[1][o]);
            s[2][threadIdx.x] = dmin(s[2][threadIdx.x], s[2][o]);
            s[3][threadIdx.x] = dmax(s[3][threadIdx.x], s[3][o]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) dst[threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(BOUNDS_THREADS) k_df_bounds(const double* __restrict__ r,
                                                              const double* __restrict__ z, int n,
                                                              double* __restrict__ partial) {
    __shared__ double s[4][BOUNDS_THREADS];
    double rMin = 1.7976931348623157e308, rMax = -1.7976931348623157e308;
    double zMin = 1.7976931348623157e308, zMax = -1.7976931348623157e308;
    for (int k = blockIdx.x * BOUNDS_THREADS + threadIdx.x; k < n; k += gridDim.x * BOUNDS_THREADS) {
        rMin = dmin(rMin, r[k]);
        rMax = dmax(rMax, r[k]);
        zMin = dmin(zMin, z[k]);
        zMax = dmax(zMax, z[k]);
    }
    box_reduce(s, rMin, rMax, zMin, zMax, partial + 4 * blockIdx.x);
}

__global__ void __launch_bounds__(BOUNDS_THREADS) k_df_bounds_final(const double* __restrict__ partial,
                                                                    int blocks, double* __restrict__ box) {
    __shared__ double s[4][BOUNDS_THREADS];
    double rMin = 1.7976931348623157e308, rMax = -1.7976931348623157e308;
    double zMin = 1.7976931348623157e308, zMax = -1.7976931348623157e308;
    for (int k = threadIdx.x; k < blocks; k += BOUNDS_THREADS) {
        rMin = dmin(rMin, partial[4 * k]);
        rMax = dmax(rMax, partial[4 * k + 1]);
        zMin = dmin(zMin, partial[4 * k + 2]);
        zMax = dmax(zMax, partial[4 * k + 3]);
    }
    box_reduce(s, rMin, rMax, zMin, zMax, box);
}

// White noise of `seeds` seeds on an ni x nj grid from (i0, j0): grid[s][b][a].
__global__ void k_df_noise(const uint64_t* __restrict__ h0, int seeds, int32_t i0, int32_t j0,
                           int32_t ni, int32_t nj, double* __restrict__ grid) {
    const int64_t nodes = (int64_t)ni * nj;
    const int64_t total = nodes * seeds;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(k / nodes);
        const int64_t rem = k - (int64_t)s * nodes;
        const int32_t b = (int32_t)(rem / ni);
        const int32_t a = (int32_t)(rem - (int64_t)b * ni);
        grid[k] = white(h0[s], i0 + a, j0 + b);
    }
}

// One thread per (site, seed) of a chunk of seeds: out[(s0 + s) sites + e]. Sites are mesh
// elements (conn, r, z) or points (r, z, w). status receives the largest failure code.
template <int SITES, int OUT, bool CACHED>
__global__ void __launch_bounds__(256) k_df_sample(Params p, int sites, int seeds, int s0,
                                                    const uint64_t* __restrict__ h0,
                                                    const int4* __restrict__ conn,
                                                    const double* __restrict__ r,
                                                    const double* __restrict__ z,
                                                    const double* __restrict__ w, const double* grid,
                                                    int32_t gi0, int32_t gj0, int32_t gni,
                                                    int32_t gnj, double* __restrict__ out,
                                                    int* __restrict__ status) {
    const int64_t total = (int64_t)sites * seeds;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(k / sites);
        const int e = (int)(k - (int64_t)s * sites);
        double rc, zc, area;
        if (SITES == SITES_MESH) {
            // signedArea and the area centroid, mesh/QuadMesh.java:78-118: the three loops
            // accumulate the same cross products in the same order.
            const int4 c = conn[e];
            const int q[4] = {c.x, c.y, c.z, c.w};
            double a = 0.0, mr = 0.0, mz = 0.0;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int i = q[t], j = q[(t + 1) & 3];
                const double cross = r[i] * z[j] - r[j] * z[i];
                a += cross;
                mr += (r[i] + r[j]) * cross;
                mz += (z[i] + z[j]) * cross;
            }
            rc = mr / (3.0 * a);
            zc = mz / (3.0 * a);
            area = 0.5 * a;
        } else {
            rc = r[e];
            zc = z[e];
            area = w ? w[e] : 0.0;
        }
        const Grid g{CACHED ? grid + (int64_t)s * gni * gnj : nullptr, gi0, gj0, gni, gnj};
        const double gv = smooth<CACHED>(p, rc, zc, h0[s], g);
        double v = gv;
        int code = OK;
        if (OUT == OUT_RING) code = transform(p, gv, effective_volume(p, rc, area), v);
        if (OUT == OUT_VOLUME) code = transform(p, gv, area, v);
        if (OUT == OUT_SLAB) code = transform(p, gv, slab_volume(p, area), v);
        if (code != OK) {
            v = __builtin_nan("");
            atomicMax(status, code);
        }
        out[(int64_t)(s0 + s) * sites + e] = v;
    }
}

// The functions of single points that need no smoothing: effectiveVolume(r, w), transform(r,
// w), white(r, z) (integers checked on the host).
__global__ void k_df_pointwise(Params p, int fn, uint64_t h0, int n, const double* __restrict__ r,
                               const double* __restrict__ z, const double* __restrict__ w,
                               double* __restrict__ out, int* __restrict__ status) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double v = 0.0;
    if (fn == NF_DEFECT_EFFECTIVE_VOLUME) {
        v = effective_volume(p, r[k], w[k]);
    } else if (fn == NF_DEFECT_TRANSFORM) {
        const int code = transform(p, r[k], w[k], v);
        if (code != OK) {
            v = __builtin_nan("");
            atomicMax(status, code);
        }
    } else {
        v = white(h0, (int32_t)r[k], (int32_t)z[k]);
    }
    out[k] = v;
}

__global__ void k_normal_eval(int fn, int n, const double* __restrict__ x, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double v = x[k];
    out[k] = fn == 0 ? erfc(v) : (fn == 1 ? erf(v) : (fn == 2 ? cdf(v) : survival(v)));
}

inline const char* message(int code) {
    switch (code) {
        case BAD_VOLUME: return "volume must be positive";
        case BAD_COUNT: return "sample count must be positive";
        case BAD_SCALE: return "scale must be positive";
        default: return "probability must lie in [0, 1]";
    }
}

// Clears one int (a fill kernel rather than a memset node, so that the call records no
// memset if a caller captures its stream).
__global__ void k_df_zero_int(int* __restrict__ p) {
    if (threadIdx.x == 0) *p = 0;
}

// Largest grid, in doubles, that one chunk of seeds may hold (1 GiB); a box too large for
// even one seed hashes every tap instead.
constexpr int64_t GRID_BUDGET = int64_t(1) << 27;

}  // namespace dfd

// The device scratch one sampling call works in: the bounding box and its per-workgroup
// partials, the status word, the seeds' hashes and the white-noise grid (grid_cap doubles,
// h0_cap hashes). A handle's own call grows its buffers to fit; the no-allocation call hands
// in fixed slices of the library's static scratch array instead.
struct DfScratch {
    double* box;
    double* boxPartial;
    int* status;
    uint64_t* h0;
    size_t h0Cap;
    double* grid;
    size_t gridCap;
};

struct nf_defect {
    dfd::Params p{};
    int64_t seed = 0;
    hipStream_t stream = nullptr;
    // events: call start, bounding box done, chunk start, noise done, sampling done
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double lastMs[3] = {0.0, 0.0, 0.0};
    // device scratch, grown on demand
    double* box = nullptr;         // 4 doubles
    double* boxPartial = nullptr;  // 4 per bounding-box workgroup
    int* status = nullptr;
    uint64_t* h0 = nullptr;
    size_t h0Cap = 0;
    double* grid = nullptr;
    size_t gridCap = 0;
    double *r = nullptr, *z = nullptr, *w = nullptr, *out = nullptr;
    int4* conn = nullptr;
    size_t rCap = 0, zCap = 0, wCap = 0, outCap = 0, connCap = 0;

    ~nf_defect() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* q : {(void*)box, (void*)boxPartial, (void*)status, (void*)h0, (void*)grid, (void*)r, (void*)z,
                        (void*)w, (void*)out, (void*)conn})
            if (q) (void)hipFree(q);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <typename T>
    static int grow(T** p, size_t& cap, size_t need) {
        if (need <= cap) return NF_OK;
        if (*p) HIP_TRY(hipFree(*p));
        *p = nullptr;
        cap = 0;
        HIP_TRY(hipMalloc((void**)p, sizeof(T) * need));
        cap = need;
        return NF_OK;
    }

    // The grid box of cacheFor :351-366 from the device bounding box; false when the grid is
    // not used (flag, empty, or beyond the budget for a single seed).
    int box_for(const DfScratch& w, const double* r_d, const double* z_d, int n, bool uncached,
                dfd::Grid& g, bool& cached) {
        cached = false;
        g = dfd::Grid{nullptr, 0, 0, 0, 0};
        HIP_TRY(hipEventRecord(ev[0], stream));
        if (uncached || n == 0) {
            HIP_TRY(hipEventRecord(ev[1], stream));
            return NF_OK;
        }
        const int blocks = std::max(1, std::min(dfd::BOX_BLOCKS, (n + 4 * dfd::BOUNDS_THREADS - 1) /
                                                                     (4 * dfd::BOUNDS_THREADS)));
        hipLaunchKernelGGL(dfd::k_df_bounds, dim3(blocks), dim3(dfd::BOUNDS_THREADS), 0, stream, r_d, z_d, n,
                           w.boxPartial);
        hipLaunchKernelGGL(dfd::k_df_bounds_final, dim3(1), dim3(dfd::BOUNDS_THREADS), 0, stream,
                           (const double*)w.boxPartial, blocks, w.box);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ev[1], stream));
        double hb[4];
        HIP_TRY(hipMemcpyAsync(hb, w.box, sizeof hb, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        const int32_t i0 = dfd::jint(std::floor(hb[0] / p.cell)) - dfd::RADIUS;
        const int32_t j0 = dfd::jint(std::floor(hb[2] / p.cell)) - dfd::RADIUS;
        const int32_t ni = dfd::jint(std::floor(hb[1] / p.cell)) + dfd::RADIUS + 2 - i0;
        const int32_t nj = dfd::jint(std::floor(hb[3] / p.cell)) + dfd::RADIUS + 2 - j0;
        const int64_t nodes = (int64_t)ni * nj;
        if (ni <= 0 || nj <= 0 || nodes > dfd::GRID_BUDGET) return NF_OK;
        g = dfd::Grid{nullptr, i0, j0, ni, nj};
        cached = true;
        return NF_OK;
    }

    // The seeds' hashes into w.h0 (which holds at least `count`), through a fixed host buffer.
    int seeds_to_device(const DfScratch& w, int count, const int64_t* seeds) {
        uint64_t hh[256];
        for (int s0 = 0; s0 < count; s0 += 256) {
            const int n = std::min(256, count - s0);
            for (int s = 0; s < n; s++) hh[s] = dfd::seed_hash(seeds ? seeds[s0 + s] : seed);
            HIP_TRY(hipMemcpyAsync(w.h0 + s0, hh, sizeof(uint64_t) * n, hipMemcpyHostToDevice, stream));
            HIP_TRY(hipStreamSynchronize(stream));  // hh is reused
        }
        return NF_OK;
    }

    template <int SITES, int OUT>
    int launch_sample(const DfScratch& ws, bool cached, int sites, int seeds, int s0, const uint64_t* h0c,
                      const int4* c, const double* r_d, const double* z_d, const double* w_d,
                      const dfd::Grid& g, double* out_d) {
        const int64_t total = (int64_t)sites * seeds;
        const int blocks = (int)std::min<int64_t>((total + 255) / 256, 1 << 20);
        if (blocks == 0) return NF_OK;
        if (cached)
            hipLaunchKernelGGL((dfd::k_df_sample<SITES, OUT, true>), dim3(blocks), dim3(256), 0, stream, p,
                               sites, seeds, s0, h0c, c, r_d, z_d, w_d, ws.grid, g.i0, g.j0, g.ni, g.nj,
                               out_d, ws.status);
        else
            hipLaunchKernelGGL((dfd::k_df_sample<SITES, OUT, false>), dim3(blocks), dim3(256), 0, stream,
                               p, sites, seeds, s0, h0c, c, r_d, z_d, w_d, (const double*)nullptr, 0, 0,
                               0, 0, out_d, ws.status);
        HIP_TRY(hipGetLastError());
        return NF_OK;
    }

    // The shared body of sampleOnto and the smoothed point functions, on device arrays.
    // The handle's own scratch, grown to hold `seedCount` hashes (the grid grows in run).
    int own_scratch(int seedCount, DfScratch& w) {
        TRY(grow(&h0, h0Cap, (size_t)std::max(1, seedCount)));
        w = DfScratch{box, boxPartial, status, h0, h0Cap, grid, gridCap};
        return NF_OK;
    }

    // fixed: the scratch is fixed (the no-allocation call); the grid then holds as many
    // seeds as fit in w.gridCap, and a box too large for one seed hashes every tap, which
    // gives the same values. Otherwise the handle's own grid grows to the chunk.
    template <int SITES, int OUT>
    int run(DfScratch w, bool fixed, int seedCount, const int64_t* seeds, int nodes, const double* r_d,
            const double* z_d, const double* w_d, int sites, const int4* c, bool uncached, double* out_d) {
        TRY(seeds_to_device(w, seedCount, seeds));
        hipLaunchKernelGGL(dfd::k_df_zero_int, dim3(1), dim3(64), 0, stream, w.status);
        HIP_TRY(hipGetLastError());
        dfd::Grid g;
        bool cached;
        TRY(box_for(w, r_d, z_d, nodes, uncached, g, cached));
        int chunk = seedCount;
        if (cached) {
            const int64_t per = (int64_t)g.ni * g.nj;
            const int64_t room = fixed ? (int64_t)w.gridCap : dfd::GRID_BUDGET;
            if (per > room) {
                cached = false;  // only reachable with fixed scratch: GRID_BUDGET bounds per already
            } else {
                chunk = (int)std::max<int64_t>(1, std::min<int64_t>(seedCount, room / per));
                if (!fixed) {
                    TRY(grow(&grid, gridCap, (size_t)(per * chunk)));
                    w.grid = grid;
                    w.gridCap = gridCap;
                }
            }
        }
        float noiseMs = 0.0f, sampleMs = 0.0f;
        for (int s0 = 0; s0 < seedCount; s0 += chunk) {
            const int n = std::min(chunk, seedCount - s0);
            HIP_TRY(hipEventRecord(ev[2], stream));
            if (cached) {
                const int64_t total = (int64_t)g.ni * g.nj * n;
                const int blocks = (int)std::min<int64_t>((total + 255) / 256, 1 << 20);
                hipLaunchKernelGGL(dfd::k_df_noise, dim3(blocks), dim3(256), 0, stream, (const uint64_t*)(w.h0 + s0),
                                   n, g.i0, g.j0, g.ni, g.nj, w.grid);
                HIP_TRY(hipGetLastError());
            }
            HIP_TRY(hipEventRecord(ev[3], stream));
            TRY((launch_sample<SITES, OUT>(w, cached, sites, n, s0, w.h0 + s0, c, r_d, z_d, w_d, g, out_d)));
            HIP_TRY(hipEventRecord(ev[4], stream));
            HIP_TRY(hipEventSynchronize(ev[4]));
            float a = 0.0f, b = 0.0f;
            HIP_TRY(hipEventElapsedTime(&a, ev[2], ev[3]));
            HIP_TRY(hipEventElapsedTime(&b, ev[3], ev[4]));
            noiseMs += a;
            sampleMs += b;
        }
        float boxMs = 0.0f;
        HIP_TRY(hipEventElapsedTime(&boxMs, ev[0], ev[1]));
        int code = 0;
        HIP_TRY(hipMemcpyAsync(&code, w.status, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        lastMs[0] = boxMs;
        lastMs[1] = noiseMs;
        lastMs[2] = sampleMs;
        if (code != dfd::OK) return fail(NF_E_ARGUMENT, dfd::message(code));
        return NF_OK;
    }
};

namespace dfd {

// A connectivity entry outside [0, nodes) sets *bad: checked before any kernel reads
// coordinates through it.
__global__ void k_df_check_conn(const int4* __restrict__ conn, int elements, int nodes, int* __restrict__ bad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= elements) return;
    const int4 c = conn[e];
    if ((unsigned)c.x >= (unsigned)nodes || (unsigned)c.y >= (unsigned)nodes ||
        (unsigned)c.z >= (unsigned)nodes || (unsigned)c.w >= (unsigned)nodes)
        atomicMax(bad, 1);
}

}  // namespace dfd

namespace {

// sampleOnto on device arrays. fixed: null for the handle's own (grown) scratch, else the
// no-allocation call's fixed scratch.
int defect_sample(nf_defect* f, const DfScratch* fixed, int32_t seed_count, const int64_t* seeds,
                  int32_t node_count, const double* r_d, const double* z_d, int32_t element_count,
                  const int4* conn_d, int32_t rule, double depth, int32_t flags, double* out_d) {
    if (rule != NF_DEFECT_RING && rule != NF_DEFECT_SLAB) return fail(NF_E_ARGUMENT, "unknown volume rule");
    if (rule == NF_DEFECT_SLAB && !(depth > 0.0)) return fail(NF_E_ARGUMENT, "depth must be positive");
    if (element_count == 0 || seed_count == 0) return NF_OK;
    DfScratch w;
    if (fixed) {
        if ((size_t)seed_count > fixed->h0Cap)
            return fail(NF_E_ARGUMENT, "seed count " + std::to_string(seed_count) + " exceeds the " +
                                           std::to_string(fixed->h0Cap) + " this call's fixed scratch holds");
        w = *fixed;
    } else {
        TRY(f->own_scratch(seed_count, w));
    }
    // connectivity in range, before any kernel dereferences it
    hipLaunchKernelGGL(dfd::k_df_zero_int, dim3(1), dim3(64), 0, f->stream, w.status);
    hipLaunchKernelGGL(dfd::k_df_check_conn, dim3(blocks_for(element_count)), dim3(BLOCK), 0, f->stream,
                       conn_d, element_count, node_count, w.status);
    HIP_TRY(hipGetLastError());
    int bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, w.status, sizeof(int), hipMemcpyDeviceToHost, f->stream));
    HIP_TRY(hipStreamSynchronize(f->stream));
    if (bad) return fail(NF_E_ARGUMENT, "connectivity names a node outside the mesh");
    f->p.depth = rule == NF_DEFECT_SLAB ? depth : 0.0;
    const bool uncached = (flags & NF_DEFECT_UNCACHED) != 0;
    if (rule == NF_DEFECT_RING)
        return f->run<dfd::SITES_MESH, dfd::OUT_RING>(w, fixed != nullptr, seed_count, seeds, node_count, r_d, z_d,
                                                      nullptr, element_count, conn_d, uncached, out_d);
    return f->run<dfd::SITES_MESH, dfd::OUT_SLAB>(w, fixed != nullptr, seed_count, seeds, node_count, r_d, z_d,
                                                  nullptr, element_count, conn_d, uncached, out_d);
}

}  // namespace

extern "C" {

int nf_defect_create(nf_defect** out, const nf_defect_params* params) {
    g_error.clear();
    if (!out || !params) return fail(NF_E_ARGUMENT, "null argument");
    *out = nullptr;
    const nf_defect_params& q = *params;
    // new Weibull(scale, modulus) (core/Weibull.java:60-69), then the DefectField checks (:124-131)
    if (!(q.scale > 0.0)) return fail(NF_E_ARGUMENT, "scale must be positive, was " + std::to_string(q.scale));
    if (!(q.modulus > 0.0)) return fail(NF_E_ARGUMENT, "modulus must be positive, was " + std::to_string(q.modulus));
    if (!(q.reference_volume > 0.0))
        return fail(NF_E_ARGUMENT, "reference volume must be positive, was " + std::to_string(q.reference_volume));
    if (!(q.correlation_length > 0.0))
        return fail(NF_E_ARGUMENT, "correlation length must be positive, was " + std::to_string(q.correlation_length));
    if (!std::isfinite(q.correlation_length)) return fail(NF_E_ARGUMENT, "correlation length must be finite");
    dfd::Params p{};
    p.scale = q.scale;
    p.modulus = q.modulus;
    p.v0 = q.reference_volume;
    p.length = q.correlation_length;
    p.cell = p.length / dfd::CELLS_PER_LENGTH;  // :137-141
    p.sigmaCells = (p.length / std::sqrt(2.0)) / p.cell;
    const int radius = dfd::jint(std::ceil(dfd::KERNEL_RADII * p.sigmaCells));
    if (radius != dfd::RADIUS)
        return fail(NF_E_ARGUMENT, "kernel radius " + std::to_string(radius) + " cells; the kernels are built for 9");
    p.twoSigmaSq = 2.0 * p.sigmaCells * p.sigmaCells;  // :295
    p.negInvM = -1.0 / p.modulus;                      // core/Weibull.java:145
    p.invM = 1.0 / p.modulus;                          // core/Weibull.java:118, :130
    nf_defect* f = new nf_defect();
    f->p = p;
    f->seed = q.seed;
    auto setup = [&]() -> int {
        HIP_TRY(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
        for (hipEvent_t& e : f->ev) HIP_TRY(hipEventCreate(&e));
        TRY(dev_alloc(&f->box, 4));
        TRY(dev_alloc(&f->boxPartial, 4 * dfd::BOX_BLOCKS));
        TRY(dev_alloc(&f->status, 1));
        return NF_OK;
    };
    if (int rc = setup()) {
        delete f;
        return rc;
    }
    *out = f;
    return NF_OK;
}

int nf_defect_destroy(nf_defect* f) {
    g_error.clear();
    delete f;
    return NF_OK;
}

int nf_defect_sample_onto(nf_defect* f, int32_t seed_count, const int64_t* seeds, int32_t node_count,
                          const double* r, const double* z, int32_t element_count, const int32_t* conn,
                          int32_t volume_rule, double depth, int32_t flags, double* out) {
    g_error.clear();
    if (!f) return fail(NF_E_ARGUMENT, "null handle");
    if (!seeds) seed_count = 1;
    if (seed_count < 0 || node_count < 0 || element_count < 0) return fail(NF_E_ARGUMENT, "negative count");
    if ((element_count && (!conn || !r || !z)) || (seed_count && element_count && !out))
        return fail(NF_E_ARGUMENT, "null array");
    for (int64_t k = 0; k < 4 * (int64_t)element_count; k++)
        if (conn[k] < 0 || conn[k] >= node_count)
            return fail(NF_E_ARGUMENT, "connectivity names a node outside the mesh");
    const size_t total = (size_t)seed_count * element_count;
    if (total == 0) return NF_OK;
    TRY(nf_defect::grow(&f->r, f->rCap, (size_t)node_count));
    TRY(nf_defect::grow(&f->z, f->zCap, (size_t)node_count));
    TRY(nf_defect::grow(&f->conn, f->connCap, (size_t)element_count));
    TRY(nf_defect::grow(&f->out, f->outCap, total));
    HIP_TRY(hipMemcpyAsync(f->r, r, sizeof(double) * node_count, hipMemcpyHostToDevice, f->stream));
    HIP_TRY(hipMemcpyAsync(f->z, z, sizeof(double) * node_count, hipMemcpyHostToDevice, f->stream));
    HIP_TRY(hipMemcpyAsync(f->conn, conn, sizeof(int4) * element_count, hipMemcpyHostToDevice, f->stream));
    TRY(defect_sample(f, nullptr, seed_count, seeds, node_count, f->r, f->z, element_count, f->conn,
                      volume_rule, depth, flags, f->out));
    HIP_TRY(hipMemcpyAsync(out, f->out, sizeof(double) * total, hipMemcpyDeviceToHost, f->stream));
    HIP_TRY(hipStreamSynchronize(f->stream));
    return NF_OK;
}

int nf_defect_sample_onto_device(nf_defect* f, int32_t seed_count, const int64_t* seeds, int32_t node_count,
                                 const double* r, const double* z, int32_t element_count, const int32_t* conn,
                                 int32_t volume_rule, double depth, int32_t flags, double* out) {
    g_error.clear();
    if (!f) return fail(NF_E_ARGUMENT, "null handle");
    if (!seeds) seed_count = 1;
    if (seed_count < 0 || node_count < 0 || element_count < 0) return fail(NF_E_ARGUMENT, "negative count");
    if ((element_count && (!conn || !r || !z)) || (seed_count && element_count && !out))
        return fail(NF_E_ARGUMENT, "null array");
    if ((uintptr_t)conn % alignof(int4)) return fail(NF_E_ARGUMENT, "connectivity must be 16-byte aligned");
    TRY(defect_sample(f, nullptr, seed_count, seeds, node_count, r, z, element_count, (const int4*)conn,
                      volume_rule, depth, flags, out));
    HIP_TRY(hipStreamSynchronize(f->stream));
    return NF_OK;
}

}  // extern "C"

// ---- sampleOnto with no allocation: the library's static scratch ----------------------------
//
// Callers that may not allocate device memory (inside a captured stream, a serving loop, a
// harness that forbids it) sample through nf_defect_sample_onto_static, which works in the
// file-scope array below instead of the handle's grown buffers. Every GPU the library runs
// on holds its own copy, 16 MiB, until the process ends. Its layout, in doubles: the
// bounding box (4), the box partials (4 per bounding-box workgroup), the status word (one
// slot), the seeds' hashes (DF_STATIC_SEEDS), and the white-noise grid (the rest). A box
// whose grid does not fit even for one seed hashes every tap instead (the same values).
namespace dfd {
constexpr size_t STATIC_DOUBLES = size_t(1) << 21;  // 16 MiB
constexpr size_t STATIC_SEEDS = 16384;
constexpr size_t STATIC_GRID_AT = 4 + 4 * (size_t)BOX_BLOCKS + 4 + STATIC_SEEDS;  // a multiple of 4
}  // namespace dfd

static __device__ __align__(256) double scratch[dfd::STATIC_DOUBLES];

namespace {

// The device address of `scratch` on each GPU, looked up once per device with that device
// current (the host-side name `scratch` is not a device address and must never reach a
// kernel or a copy).
std::mutex g_scratchLock;
double* g_scratchOn[64] = {};

int static_scratch(hipStream_t stream, DfScratch& w) {
    int dev = 0;
    HIP_TRY(hipStreamGetDevice(stream, &dev));
    if (dev < 0 || dev >= 64) return fail(NF_E_DEVICE, "device index out of range for the static scratch table");
    double* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratchLock);
        if (!g_scratchOn[dev]) {
            int current = 0;
            HIP_TRY(hipGetDevice(&current));
            if (current != dev) HIP_TRY(hipSetDevice(dev));
            void* ptr = nullptr;
            const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (current != dev) HIP_TRY(hipSetDevice(current));
            if (err != hipSuccess || !ptr)
                return fail(NF_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
            g_scratchOn[dev] = (double*)ptr;
        }
        base = g_scratchOn[dev];
    }
    w.box = base;
    w.boxPartial = base + 4;
    w.status = (int*)(base + 4 + 4 * (size_t)dfd::BOX_BLOCKS);
    w.h0 = (uint64_t*)(base + 4 + 4 * (size_t)dfd::BOX_BLOCKS + 4);
    w.h0Cap = dfd::STATIC_SEEDS;
    w.grid = base + dfd::STATIC_GRID_AT;
    w.gridCap = dfd::STATIC_DOUBLES - dfd::STATIC_GRID_AT;
    return NF_OK;
}

}  // namespace

extern "C" {

int nf_defect_sample_onto_static(nf_defect* f, int32_t seed_count, const int64_t* seeds, int32_t node_count,
                                 const double* r, const double* z, int32_t element_count, const int32_t* conn,
                                 int32_t volume_rule, double depth, int32_t flags, double* out) {
    g_error.clear();
    if (!f) return fail(NF_E_ARGUMENT, "null handle");
    if (!seeds) seed_count = 1;
    if (seed_count < 0 || node_count < 0 || element_count < 0) return fail(NF_E_ARGUMENT, "negative count");
    if ((size_t)seed_count > dfd::STATIC_SEEDS)
        return fail(NF_E_ARGUMENT, "seed count " + std::to_string(seed_count) + " exceeds the static scratch's " +
                                       std::to_string(dfd::STATIC_SEEDS));
    if ((element_count && (!conn || !r || !z)) || (seed_count && element_count && !out))
        return fail(NF_E_ARGUMENT, "null array");
    if ((uintptr_t)conn % alignof(int4)) return fail(NF_E_ARGUMENT, "connectivity must be 16-byte aligned");
    DfScratch w;
    TRY(static_scratch(f->stream, w));
    TRY(defect_sample(f, &w, seed_count, seeds, node_count, r, z, element_count, (const int4*)conn, volume_rule,
                      depth, flags, out));
    HIP_TRY(hipStreamSynchronize(f->stream));
    return NF_OK;
}

int nf_defect_points(nf_defect* f, int32_t function, int32_t n, const double* r, const double* z,
                     const double* w, double* out) {
    g_error.clear();
    if (!f) return fail(NF_E_ARGUMENT, "null handle");
    if (n < 0) return fail(NF_E_ARGUMENT, "negative count");
    if (function < NF_DEFECT_GAUSSIAN || function > NF_DEFECT_WHITE) return fail(NF_E_ARGUMENT, "unknown function");
    if (n == 0) return NF_OK;
    const bool needZ = function <= NF_DEFECT_STRAIN_VOLUME || function == NF_DEFECT_WHITE;
    const bool needW = function != NF_DEFECT_GAUSSIAN && function != NF_DEFECT_WHITE;
    if (!r || !out || (needZ && !z) || (needW && !w)) return fail(NF_E_ARGUMENT, "null array");
    if (function == NF_DEFECT_WHITE)
        for (int k = 0; k < n; k++)
            if (!(r[k] == std::floor(r[k]) && z[k] == std::floor(z[k]) && r[k] >= -2147483648.0 &&
                  r[k] <= 2147483647.0 && z[k] >= -2147483648.0 && z[k] <= 2147483647.0))
                return fail(NF_E_ARGUMENT, "white noise is indexed by int32 pairs");
    TRY(nf_defect::grow(&f->r, f->rCap, (size_t)n));
    TRY(nf_defect::grow(&f->z, f->zCap, (size_t)n));
    TRY(nf_defect::grow(&f->w, f->wCap, (size_t)n));
    TRY(nf_defect::grow(&f->out, f->outCap, (size_t)n));
    HIP_TRY(hipMemcpyAsync(f->r, r, sizeof(double) * n, hipMemcpyHostToDevice, f->stream));
    if (needZ) HIP_TRY(hipMemcpyAsync(f->z, z, sizeof(double) * n, hipMemcpyHostToDevice, f->stream));
    if (needW) HIP_TRY(hipMemcpyAsync(f->w, w, sizeof(double) * n, hipMemcpyHostToDevice, f->stream));
    if (function <= NF_DEFECT_STRAIN_VOLUME) {
        // the smoothed functions, over a grid on the points' own bounding box
        DfScratch ws;
        TRY(f->own_scratch(1, ws));
        if (function == NF_DEFECT_GAUSSIAN)
            TRY((f->run<dfd::SITES_POINTS, dfd::OUT_GAUSSIAN>(ws, false, 1, nullptr, n, f->r, f->z, nullptr, n,
                                                               nullptr, false, f->out)));
        else if (function == NF_DEFECT_STRAIN_AREA)
            TRY((f->run<dfd::SITES_POINTS, dfd::OUT_RING>(ws, false, 1, nullptr, n, f->r, f->z, f->w, n, nullptr,
                                                           false, f->out)));
        else
            TRY((f->run<dfd::SITES_POINTS, dfd::OUT_VOLUME>(ws, false, 1, nullptr, n, f->r, f->z, f->w, n,
                                                             nullptr, false, f->out)));
    } else {
        HIP_TRY(hipMemsetAsync(f->status, 0, sizeof(int), f->stream));
        hipLaunchKernelGGL(dfd::k_df_pointwise, dim3(blocks_for(n)), dim3(BLOCK), 0, f->stream, f->p, function,
                           dfd::seed_hash(f->seed), n, (const double*)f->r, (const double*)f->z,
                           (const double*)f->w, f->out, f->status);
        HIP_TRY(hipGetLastError());
        int code = 0;
        HIP_TRY(hipMemcpyAsync(&code, f->status, sizeof(int), hipMemcpyDeviceToHost, f->stream));
        HIP_TRY(hipStreamSynchronize(f->stream));
        if (code != dfd::OK) return fail(NF_E_ARGUMENT, dfd::message(code));
    }
    HIP_TRY(hipMemcpyAsync(out, f->out, sizeof(double) * n, hipMemcpyDeviceToHost, f->stream));
    HIP_TRY(hipStreamSynchronize(f->stream));
    return NF_OK;
}

int nf_defect_timing(nf_defect* f, double* ms) {
    g_error.clear();
    if (!f || !ms) return fail(NF_E_ARGUMENT, "null argument");
    for (int k = 0; k < 3; k++) ms[k] = f->lastMs[k];
    return NF_OK;
}

int nf_normal_eval(int32_t function, int32_t n, const double* x, double* out) {
    g_error.clear();
    if (n < 0) return fail(NF_E_ARGUMENT, "negative count");
    if (function < 0 || function > 3) return fail(NF_E_ARGUMENT, "unknown function");
    if (n == 0) return NF_OK;
    if (!x || !out) return fail(NF_E_ARGUMENT, "null array");
    double *dx = nullptr, *dout = nullptr;
    TRY(dev_alloc(&dx, n));
    if (int rc = dev_alloc(&dout, n)) {
        (void)hipFree(dx);
        return rc;
    }
    hipError_t err = hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(dfd::k_normal_eval, dim3(blocks_for(n)), dim3(BLOCK), 0, 0, function, n,
                           (const double*)dx, dout);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    (void)hipFree(dout);
    if (err != hipSuccess) return fail(NF_E_DEVICE, hipGetErrorString(err));
    return NF_OK;
}

}  // extern "C"

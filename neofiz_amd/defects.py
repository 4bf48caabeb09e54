"""Defect-field seeding: the spatially correlated Weibull failure strain of each element.

Names follow the reference (src/main/java/org/neofiz/core/):

  Weibull      Weibull.java:53-236, the two-parameter weakest-link record. Its members are the
               closed forms of the distribution (forVolume, minimumOf, quantile, mean, ...),
               evaluated here as parameter algebra, as Material's wave speed is.
  Normal       Normal.java:17-130, Cody's rational erfc, evaluated on the device.
  DefectField  DefectField.java:92-385, the Gaussian-copula field, sampled on the device:
               gaussianAt, failureStrainAt, failureStrainForVolume, effectiveVolume and
               sampleOnto (for one seed or many at once).

The field is stateless: every value is a function of (seed, r, z), white noise drawn by a
SplitMix64 hash of the grid index and smoothed by a separable Gaussian, so two meshes of one
body see the same field bit for bit. sampleOnto runs three kernels (bounding box, white-noise
grid, smoothing + transform); there is no CPU path.
"""
import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _abi
from .solver import IllegalArgumentException, QuadMesh, check

RING, SLAB = 0, 1          # nf_defect_volume
UNCACHED = 1               # NF_DEFECT_UNCACHED
GAUSSIAN, STRAIN_AREA, STRAIN_VOLUME, EFFECTIVE_VOLUME, TRANSFORM, WHITE = range(6)

# Lanczos approximation with g = 7 and nine terms (Godfrey's published coefficients), the
# form Weibull.gamma uses (Weibull.java:211-235).
_LANCZOS = (0.99999999999980993, 676.5203681218851, -1259.1392167224028, 771.32342877765313,
            -176.61502916214059, 12.507343278686905, -0.13857109526572012,
            9.9843695780195716e-6, 1.5056327351493116e-7)


def _jpow(a, b):
    """Math.pow for the positive bases used here: overflow gives infinity, not an error."""
    try:
        return math.pow(a, b)
    except OverflowError:
        return math.inf


def _binding(binding):
    if binding is None:
        from . import _native
        binding = _native.binding()
    return binding


@dataclass(frozen=True)
class Weibull:
    """Weibull.java:53-236. scale is the characteristic value at the reference volume the
    caller uses; modulus m sets the scatter (small is scattered, large is tight)."""
    scale: float
    modulus: float

    MODULUS_DUCTILE_STEEL = 30.0
    MODULUS_BRITTLE = 8.0

    def __post_init__(self):  # :60-69
        if not self.scale > 0.0:
            raise IllegalArgumentException(f"scale must be positive, was {self.scale}")
        if not self.modulus > 0.0:
            raise IllegalArgumentException(f"modulus must be positive, was {self.modulus}")

    def cumulativeHazard(self, x):  # :88-91
        return 0.0 if x <= 0.0 else _jpow(x / self.scale, self.modulus)

    def survival(self, x):  # :94-97
        return 1.0 if x <= 0.0 else math.exp(-self.cumulativeHazard(x))

    def cdf(self, x):  # :100-103
        return 0.0 if x <= 0.0 else -math.expm1(-self.cumulativeHazard(x))

    def quantile(self, p):  # :114-119
        if not (0.0 <= p <= 1.0):
            raise IllegalArgumentException(f"probability must lie in [0, 1], was {p}")
        if p == 1.0:
            return math.inf
        return self.scale * _jpow(-math.log1p(-p), 1.0 / self.modulus)

    def quantileFromSurvival(self, s):  # :126-131
        if not (0.0 <= s <= 1.0):
            raise IllegalArgumentException(f"probability must lie in [0, 1], was {s}")
        if s == 0.0:
            return math.inf
        return self.scale * _jpow(-math.log(s), 1.0 / self.modulus)

    def minimumOf(self, n):  # :141-146
        if not n > 0.0:
            raise IllegalArgumentException(f"sample count must be positive, was {n}")
        return Weibull(self.scale * _jpow(n, -1.0 / self.modulus), self.modulus)

    def forVolume(self, volume, referenceVolume):  # :155-165
        if not volume > 0.0:
            raise IllegalArgumentException(f"volume must be positive, was {volume}")
        if not referenceVolume > 0.0:
            raise IllegalArgumentException(
                f"reference volume must be positive, was {referenceVolume}")
        return self.minimumOf(volume / referenceVolume)

    def sizeFactor(self, volume, referenceVolume):  # :173-175
        return self.forVolume(volume, referenceVolume).scale / self.scale

    def mean(self):  # :178-180
        return self.scale * Weibull.gamma(1.0 + 1.0 / self.modulus)

    def standardDeviation(self):  # :183-187
        g1 = Weibull.gamma(1.0 + 1.0 / self.modulus)
        g2 = Weibull.gamma(1.0 + 2.0 / self.modulus)
        return self.scale * math.sqrt(max(0.0, g2 - g1 * g1))

    def coefficientOfVariation(self):  # :194-198
        g1 = Weibull.gamma(1.0 + 1.0 / self.modulus)
        g2 = Weibull.gamma(1.0 + 2.0 / self.modulus)
        return math.sqrt(max(0.0, g2 - g1 * g1)) / g1

    @staticmethod
    def gamma(x):  # :211-222, reflection below 1/2
        if x < 0.5:
            return math.pi / (math.sin(math.pi * x) * Weibull.gamma(1.0 - x))
        z = x - 1.0
        a = _LANCZOS[0]
        for i in range(1, len(_LANCZOS)):
            a += _LANCZOS[i] / (z + i)
        t = z + 7.5
        return math.sqrt(2.0 * math.pi) * _jpow(t, z + 0.5) * math.exp(-t) * a


class Normal:
    """Normal.java:17-130 on the device: erfc, erf, cdf and survival of scalars or arrays."""

    @staticmethod
    def _eval(fn, x, binding=None):
        b = _binding(binding)
        arr = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))
        out = np.empty_like(arr)
        check(b, b.normal_eval(fn, arr.size, arr.ctypes.data_as(_abi.c_double_p),
                               out.ctypes.data_as(_abi.c_double_p)))
        return float(out[0]) if np.ndim(x) == 0 else out.reshape(np.shape(x))

    @staticmethod
    def erfc(x, binding=None):
        return Normal._eval(0, x, binding)

    @staticmethod
    def erf(x, binding=None):
        return Normal._eval(1, x, binding)

    @staticmethod
    def cdf(x, binding=None):
        return Normal._eval(2, x, binding)

    @staticmethod
    def survival(x, binding=None):
        return Normal._eval(3, x, binding)


class DefectField:
    """DefectField.java:92-385 over a native handle. `binding` selects the library; the
    default is the HIP library, which raises if it is not built."""

    def __init__(self, base: Weibull, referenceVolume, correlationLength, seed, binding=None):
        self._b = _binding(binding)
        self._base = base
        self._v0 = float(referenceVolume)
        self._length = float(correlationLength)
        self._seed = int(seed)
        params = _abi.DefectParamsC(base.scale, base.modulus, self._v0, self._length,
                                    C.c_int64(self._seed).value)
        handle = C.c_void_p()
        check(self._b, self._b.defect_create(C.byref(handle), C.byref(params)))
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            self._b.defect_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # accessors :144-158
    def base(self):
        return self._base

    def referenceVolume(self):
        return self._v0

    def correlationLength(self):
        return self._length

    def seed(self):
        return self._seed

    def correlation(self, d):  # :161-164, the closed form the field is built to have
        q = d / self._length
        return math.exp(-0.5 * q * q)

    def forBody(self, volume):  # :254-256
        return self._base.forVolume(volume, self._v0)

    # point evaluations, each on the device
    def _points(self, fn, r, z, w):
        shape = np.broadcast(np.asarray(r), np.asarray(z), np.asarray(w)).shape
        ra, za, wa = (np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), shape)).reshape(-1)
                      for v in (r, z, w))
        out = np.empty(ra.size)
        p = _abi.c_double_p
        check(self._b, self._b.defect_points(self._h, fn, ra.size, ra.ctypes.data_as(p),
                                             za.ctypes.data_as(p), wa.ctypes.data_as(p),
                                             out.ctypes.data_as(p)))
        return float(out[0]) if shape == () else out.reshape(shape)

    def gaussianAt(self, r, z):  # :172-174
        return self._points(GAUSSIAN, r, z, 0.0)

    def effectiveVolume(self, centroidRadius, crossSectionArea):  # :204-208
        return self._points(EFFECTIVE_VOLUME, centroidRadius, 0.0, crossSectionArea)

    def failureStrainAt(self, r, z, crossSectionArea):  # :217-219
        return self._points(STRAIN_AREA, r, z, crossSectionArea)

    def failureStrainForVolume(self, r, z, volume):  # :226-228
        return self._points(STRAIN_VOLUME, r, z, volume)

    def transform(self, g, volume):  # :261-271 (private in the reference)
        return self._points(TRANSFORM, g, 0.0, volume)

    def white(self, i, j):  # :325-332 (private in the reference)
        return self._points(WHITE, i, j, 0.0)

    def sampleOnto(self, mesh: QuadMesh, seeds=None, depth=None, cached=True):
        """sampleOnto :239-248. With `seeds`, samples the fields that differ from this one only
        in their seed, all in one launch, and returns an array [len(seeds), elements]. With
        `depth`, the plane-strain rule: each element holds l^3 max(1, A/l^2) max(1, depth/l)
        (no counterpart in the reference, whose sampleOnto is for rings)."""
        many = seeds is not None
        sd = np.ascontiguousarray(np.atleast_1d(seeds if many else [self._seed]), dtype=np.int64)
        out = np.empty((sd.size, mesh.elementCount))
        rule = RING if depth is None else SLAB
        check(self._b, self._b.defect_sample_onto(
            self._h, sd.size, sd.ctypes.data_as(C.POINTER(C.c_int64)), mesh.nodeCount,
            mesh.r.ctypes.data_as(_abi.c_double_p), mesh.z.ctypes.data_as(_abi.c_double_p),
            mesh.elementCount, mesh.conn.ctypes.data_as(_abi.c_int32_p), rule,
            float(depth or 0.0), 0 if cached else UNCACHED, out.ctypes.data_as(_abi.c_double_p)))
        return out if many else out[0]

    def sampleOntoDevice(self, count, r_ptr, z_ptr, elements, conn_ptr, out_ptr, seeds=None,
                         depth=None, cached=True):
        """sampleOnto with the mesh and the output in device memory (raw pointers, e.g. torch
        tensors' data_ptr()); returns when the kernels are done."""
        sd = np.ascontiguousarray(np.atleast_1d(seeds if seeds is not None else [self._seed]),
                                  dtype=np.int64)
        rule = RING if depth is None else SLAB
        check(self._b, self._b.defect_sample_onto_device(
            self._h, sd.size, sd.ctypes.data_as(C.POINTER(C.c_int64)), count, r_ptr, z_ptr,
            elements, conn_ptr, rule, float(depth or 0.0), 0 if cached else UNCACHED, out_ptr))

    def timing(self):
        """Device milliseconds of the last sampling call: (bounding box, noise grid,
        smoothing and transform)."""
        ms = np.zeros(3)
        check(self._b, self._b.defect_timing(self._h, ms.ctypes.data_as(_abi.c_double_p)))
        return tuple(float(v) for v in ms)

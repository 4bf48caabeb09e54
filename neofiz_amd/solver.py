"""Host-side mirror of Neofiz's solver API over the C ABI.

Class and method names follow the Java reference (src/main/java/org/neofiz/...), so a caller
of `org.neofiz.solver.ExplicitSolver` finds the same surface here:

  QuadMesh            mesh/QuadMesh.java:15-403 (solidCylinder, cylinderWall, axisNodes, ...)
  JohnsonCook         core/JohnsonCook.java:68-227 (STEEL_4340, COPPER_OFHC)
  Material            core/Material.java:17-136 (withJohnsonCook, yielding)
  Materials           core/Materials.java:23-129 (uniform, of, byElement)
  RigidWall           solver/RigidWall.java:95-316 (atZ)
  Contact             solver/Contact.java:132-364 (withFriction, withDamping, withStiffnessScale)
  ExplicitSolver      solver/ExplicitSolver.java:38-1966

Every compute call goes through the HIP library (`neofiz_amd._native`); there is no CPU path.
Errors map as the Java ones do: IllegalArgumentException and IllegalStateException below are
raised where the reference throws them.
"""
import ctypes as C
import math
from dataclasses import dataclass, replace
from typing import Optional, Sequence

import numpy as np

from . import _abi

AXISYMMETRIC, PLANE_STRAIN, PLANE_STRESS = 0, 1, 2   # core/Formulation.java:17-21
FULL, REDUCED = 0, 1                                 # solver/Integration.java:30-32
SMALL_STRAIN, FINITE_STRAIN = 0, 1                   # solver/Kinematics.java:20, :35


class IllegalArgumentException(ValueError):
    pass


class IllegalStateException(RuntimeError):
    pass


class DeviceError(RuntimeError):
    pass


def check(binding, rc):
    if rc == 0:
        return
    msg = binding.last_error().decode(errors="replace")
    if rc == -1:
        raise IllegalArgumentException(msg)
    if rc == -2:
        raise IllegalStateException(msg)
    raise DeviceError(f"status {rc}: {msg}")


# ---------------------------------------------------------------- materials

@dataclass(frozen=True)
class JohnsonCook:
    """core/JohnsonCook.java:68-71. Validation happens in the C ABI (nf_material_check)."""
    a: float
    b: float
    n: float
    c: float
    m: float
    referenceStrainRate: float
    roomTemperature: float
    meltTemperature: float
    specificHeat: float
    taylorQuinney: float

    def quasiStatic(self):  # :125-128
        return replace(self, c=0.0, taylorQuinney=0.0)


JohnsonCook.STEEL_4340 = JohnsonCook(792.0e6, 510.0e6, 0.26, 0.014, 1.03,
                                     1.0, 293.0, 1793.0, 477.0, 0.9)   # :95-97
JohnsonCook.COPPER_OFHC = JohnsonCook(90.0e6, 292.0e6, 0.31, 0.025, 1.09,
                                      1.0, 293.0, 1356.0, 383.0, 0.9)  # :100-102


@dataclass(frozen=True)
class Material:
    """core/Material.java:17-19."""
    name: str
    youngsModulus: float
    poissonRatio: float
    density: float
    yieldStress: float = math.inf
    isotropicHardening: float = 0.0
    kinematicHardening: float = 0.0
    johnsonCook: Optional[JohnsonCook] = None

    def isElastic(self):
        return math.isinf(self.yieldStress)

    def yielding(self, yieldStress, isotropic, kinematic=0.0):  # :71-80
        return replace(self, yieldStress=yieldStress, isotropicHardening=isotropic,
                       kinematicHardening=kinematic, johnsonCook=None)

    def withJohnsonCook(self, law):  # :87-90
        return replace(self, yieldStress=law.a, isotropicHardening=0.0, johnsonCook=law)

    def lambda_(self):
        e, nu = self.youngsModulus, self.poissonRatio
        return e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))

    def mu(self):
        return self.youngsModulus / (2.0 * (1.0 + self.poissonRatio))

    def dilatationalWaveSpeed(self):
        return math.sqrt((self.lambda_() + 2.0 * self.mu()) / self.density)

    def barWaveSpeed(self):  # :127-129
        return math.sqrt(self.youngsModulus / self.density)

    def to_c(self):
        out = _abi.MaterialC()
        out.youngs_modulus = self.youngsModulus
        out.poisson_ratio = self.poissonRatio
        out.density = self.density
        out.yield_stress = self.yieldStress
        out.isotropic_hardening = self.isotropicHardening
        out.kinematic_hardening = self.kinematicHardening
        if self.johnsonCook is not None:
            j = self.johnsonCook
            out.has_johnson_cook = 1
            out.johnson_cook = _abi.JohnsonCookC(j.a, j.b, j.n, j.c, j.m, j.referenceStrainRate,
                                                 j.roomTemperature, j.meltTemperature,
                                                 j.specificHeat, j.taylorQuinney)
        return out


Material.STEEL_4340 = Material("AISI 4340", 205.0e9, 0.29, 7850.0)      # :132
Material.COPPER_OFHC = Material("OFHC Cu", 117.0e9, 0.34, 8960.0)      # :135


class Materials:
    """core/Materials.java:23-129: one material, or one per element."""

    def __init__(self, distinct, index):
        self.distinct = list(distinct)
        self.index = index  # None when uniform

    @staticmethod
    def uniform(material):
        return Materials([material], None)

    @staticmethod
    def of(by_element: Sequence[Material]):
        distinct, index, seen = [], np.empty(len(by_element), np.int32), {}
        for e, m in enumerate(by_element):
            k = id(m)
            if k not in seen:
                seen[k] = len(distinct)
                distinct.append(m)
            index[e] = seen[k]
        return Materials(distinct, index)

    @staticmethod
    def byElement(count, pick):
        return Materials.of([pick(e) for e in range(count)])


# ---------------------------------------------------------------- mesh

class QuadMesh:
    """mesh/QuadMesh.java:15-403: SoA node coordinates, CCW connectivity, pressure edges."""

    def __init__(self, r, z, conn, pressureEdges=()):
        self.r = np.ascontiguousarray(r, dtype=np.float64)
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1)
        self.pressureEdges = np.ascontiguousarray(pressureEdges, dtype=np.int32).reshape(-1)
        if self.r.shape != self.z.shape:
            raise IllegalArgumentException("r and z must have the same length")
        if self.conn.size % 4:
            raise IllegalArgumentException("connectivity must be a multiple of 4")
        if self.pressureEdges.size % 2:
            raise IllegalArgumentException("pressure edges must be a multiple of 2")
        self.nodeCount = self.r.size
        self.elementCount = self.conn.size // 4
        if self.elementCount and self.conn.min() < 0 or self.conn.size and self.conn.max() >= self.nodeCount:
            raise IllegalArgumentException("connectivity refers to a node that does not exist")
        bad = np.nonzero(self.signedAreas() <= 0.0)[0]
        if bad.size:  # checkOrientation :228-236
            raise IllegalStateException(
                f"element {int(bad[0])} has non-positive area; node ordering must be "
                "counter-clockwise in (r, z)")

    @staticmethod
    def _lattice(r_of, z_of, nr, nz):
        nodes_r = nr + 1
        j, i = np.meshgrid(np.arange(nz + 1), np.arange(nodes_r), indexing="ij")
        r = r_of(i.reshape(-1))
        z = z_of(j.reshape(-1))
        jj, ii = np.meshgrid(np.arange(nz), np.arange(nr), indexing="ij")
        base = (jj * nodes_r + ii).reshape(-1)
        conn = np.stack([base, base + 1, base + nodes_r + 1, base + nodes_r], axis=1)
        return r, z, conn.astype(np.int32)

    @staticmethod
    def solidCylinder(radius, height, nr, nz):  # :324-356
        if radius <= 0.0:
            raise IllegalArgumentException("radius must be positive")
        if height <= 0.0:
            raise IllegalArgumentException("height must be positive")
        if nr < 1 or nz < 1:
            raise IllegalArgumentException("need at least one element in each direction")
        # radius * i / nr and height * j / nz, as the reference computes them (multiply, then
        # divide), element by element in double precision.
        r, z, conn = QuadMesh._lattice(lambda i: radius * i.astype(np.float64) / nr,
                                       lambda j: height * j.astype(np.float64) / nz, nr, nz)
        return QuadMesh(r, z, conn)

    @staticmethod
    def cylinderWall(inner, outer, height, nr, nz):  # :254-296
        if outer <= inner:
            raise IllegalArgumentException("outer radius must exceed inner radius")
        if nr < 1 or nz < 1:
            raise IllegalArgumentException("need at least one element in each direction")
        r, z, conn = QuadMesh._lattice(lambda i: inner + (outer - inner) * i.astype(np.float64) / nr,
                                       lambda j: height * j.astype(np.float64) / nz, nr, nz)
        nodes_r = nr + 1
        j = np.arange(nz)
        edges = np.stack([j * nodes_r, (j + 1) * nodes_r], axis=1)
        return QuadMesh(r, z, conn, edges)

    @staticmethod
    def axisNodes(nr, nz):  # :362-364
        return [j * (nr + 1) for j in range(nz + 1)]

    innerSurfaceNodes = axisNodes

    @staticmethod
    def outerSurfaceNodes(nr, nz):
        return [j * (nr + 1) + nr for j in range(nz + 1)]

    @staticmethod
    def nearFaceNodes(nr, nz):
        return list(range(nr + 1))

    @staticmethod
    def farFaceNodes(nr, nz):
        return [nz * (nr + 1) + i for i in range(nr + 1)]

    def signedAreas(self):
        """signedArea of every element, with the reference's summation order."""
        c = self.conn.reshape(-1, 4)
        a = np.zeros(c.shape[0])
        for k in range(4):
            i, j = c[:, k], c[:, (k + 1) % 4]
            a = a + (self.r[i] * self.z[j] - self.r[j] * self.z[i])
        return 0.5 * a

    def minimumEdgeLength(self):  # :61-75
        c = self.conn.reshape(-1, 4)
        best = 1.7976931348623157e308
        for k in range(4):
            i, j = c[:, k], c[:, (k + 1) % 4]
            dr, dz = self.r[j] - self.r[i], self.z[j] - self.z[i]
            ln = np.sqrt(dr * dr + dz * dz)
            if ln.size:
                best = min(best, float(ln.min()))
        return best

    def _centroid(self, e, coord):  # :106-118
        b = 4 * e
        a = moment = 0.0
        for k in range(4):
            i, j = self.conn[b + k], self.conn[b + (k + 1) % 4]
            cross = self.r[i] * self.z[j] - self.r[j] * self.z[i]
            a += cross
            moment += (coord[i] + coord[j]) * cross
        return float(moment / (3.0 * a))

    def centroidRadius(self, e):  # :96-98
        return self._centroid(e, self.r)

    def centroidZ(self, e):  # :101-103
        return self._centroid(e, self.z)

    def surface(self, skip=None):
        """The free surface by connectivity, :188-226: an edge used by exactly one live
        element is free; free edges are listed element by element, local edge by local
        edge, and `nodes` holds every node on them in ascending order."""
        c = self.conn.reshape(-1, 4)
        a = c.reshape(-1)
        b = c[:, [1, 2, 3, 0]].reshape(-1)
        live = np.ones(self.elementCount, dtype=bool) if skip is None else ~np.asarray(skip, dtype=bool)
        live4 = np.repeat(live, 4)
        key = (np.minimum(a, b).astype(np.int64) << 32) | np.maximum(a, b).astype(np.int64)
        uniq, inv, counts = np.unique(key[live4], return_inverse=True, return_counts=True)
        free_live = counts[inv] == 1
        idx = np.nonzero(live4)[0][free_live]
        edges = np.stack([a[idx], b[idx]], axis=1).reshape(-1).astype(np.int32)
        owner = (idx // 4).astype(np.int32)
        nodes = np.unique(edges).astype(np.int32)
        return Surface(edges, owner, nodes)

    def signedArea(self, e):
        c = self.conn[4 * e:4 * e + 4]
        a = 0.0
        for k in range(4):
            i, j = c[k], c[(k + 1) % 4]
            a += self.r[i] * self.z[j] - self.r[j] * self.z[i]
        return float(0.5 * a)

    def ringVolume(self, e):  # :132-134, Pappus: 2 pi rc A
        return 2.0 * math.pi * self.centroidRadius(e) * self.signedArea(e)

    def totalRingVolume(self):  # :137-141, summed in element order
        v = 0.0
        for e in range(self.elementCount):
            v += self.ringVolume(e)
        return v


@dataclass(frozen=True)
class Surface:
    """QuadMesh.Surface (mesh/QuadMesh.java:152-157): edges (a, b) pairs, owning element per
    edge, and the surface nodes ascending."""
    edges: np.ndarray
    owner: np.ndarray
    nodes: np.ndarray

    def edgeCount(self):
        return int(self.owner.size)


# ---------------------------------------------------------------- boundary objects

@dataclass(frozen=True)
class RigidWall:
    """solver/RigidWall.java:95-136: the plane z = plane, friction coefficient (inf = welded)."""
    plane: float
    friction: float = 0.0

    @staticmethod
    def atZ(z, friction=0.0):
        if friction < 0.0 or math.isnan(friction):
            raise IllegalArgumentException("friction coefficient must be >= 0")
        return RigidWall(z, friction)


class Contact:
    """solver/Contact.java:273-364: penalty contact settings (built into the solver's state)."""

    def __init__(self, mesh=None):
        self.friction = 0.0
        self.damping = 0.0
        self.scale = 0.1
        self._solver = None
        self._segments = mesh.surface().edgeCount() if mesh is not None else None

    # Audits (:371-404). The state lives in the solver this contact was handed to.
    def _state(self):
        if self._solver is None:
            return None
        return self._solver.state()

    def pairCount(self):
        st = self._state()
        return 0 if st is None else st.contact_pairs

    def energy(self):
        st = self._state()
        return 0.0 if st is None else st.contact_energy

    def dissipation(self):
        st = self._state()
        return 0.0 if st is None else st.contact_dissipation

    def maxPenetration(self):
        st = self._state()
        return 0.0 if st is None else st.contact_max_penetration

    def escapedNodes(self):
        st = self._state()
        return 0 if st is None else int(st.contact_escaped_nodes)

    def segmentCount(self):
        if self._solver is None:
            return self._segments
        return self._solver.contactSurface().shape[0]

    def withFriction(self, mu):
        if mu < 0.0:
            raise IllegalArgumentException("friction must not be negative")
        self.friction = mu
        return self

    def withDamping(self, zeta):
        if zeta < 0.0 or zeta > 1.0:
            raise IllegalArgumentException("contact damping is a fraction of critical, so it lies in [0, 1]")
        self.damping = zeta
        return self

    def withStiffnessScale(self, s):
        if not (s > 0.0) or s >= 2.0:
            raise IllegalArgumentException("the stiffness scale must lie in (0, 2)")
        self.scale = s
        return self


def _dptr(a):
    return a.ctypes.data_as(_abi.c_double_p)


def _iptr(a):
    return a.ctypes.data_as(_abi.c_int32_p)


class ExplicitSolver:
    """solver/ExplicitSolver.java:38-1966 over a native handle.

    `binding` selects the library; the default is the HIP library, loaded through
    `neofiz_amd._native.binding()`, which raises if it is not built.
    """

    def __init__(self, mesh: QuadMesh, materials, formulation=AXISYMMETRIC, integration=FULL,
                 kinematics=SMALL_STRAIN, thickness=1.0, cflSafety=0.5, binding=None):
        if binding is None:
            from . import _native
            binding = _native.binding()
        self._b = binding
        if isinstance(materials, Material):
            materials = Materials.uniform(materials)
        self.mesh = mesh
        self.materials = materials
        mats = (_abi.MaterialC * len(materials.distinct))(*[m.to_c() for m in materials.distinct])
        index = None
        if materials.index is not None:
            self._index = np.ascontiguousarray(materials.index, dtype=np.int32)
            index = _iptr(self._index)
        handle = C.c_void_p()
        edges = mesh.pressureEdges if mesh.pressureEdges.size else np.zeros(2, np.int32)
        check(self._b, self._b.solver_create(
            C.byref(handle), mesh.nodeCount, _dptr(mesh.r), _dptr(mesh.z), mesh.elementCount,
            _iptr(mesh.conn), mesh.pressureEdges.size // 2, _iptr(edges),
            len(materials.distinct), mats, index, formulation, integration, kinematics,
            float(thickness), float(cflSafety)))
        self._h = handle
        self.integration = integration
        self.formulation = formulation
        self.kinematics = kinematics
        self.wall = None
        self._contact = None

    # lifecycle
    def shutdown(self):
        if getattr(self, "_h", None):
            self._b.solver_destroy(self._h)
            self._h = None

    close = shutdown

    def __del__(self):
        try:
            self.shutdown()
        except Exception:
            pass

    def _call(self, name, *args):
        check(self._b, getattr(self._b, name)(self._h, *args))

    # setters, solver/ExplicitSolver.java:386-808
    def setThreads(self, n):
        self._call("set_threads", int(n))

    def fixRadial(self, node):
        self._call("fix_radial", int(node))

    def fixAxial(self, node):
        self._call("fix_axial", int(node))

    def fixAllAxial(self):
        self._call("fix_all_axial")

    def setPressureRamp(self, pascals, rampSeconds):
        self._call("set_pressure_ramp", float(pascals), float(rampSeconds))

    def setRelaxationDamping(self, fraction, omega):
        self._call("set_relaxation_damping", float(fraction), float(omega))

    def setHourglassCoefficient(self, q):
        self._call("set_hourglass_coefficient", float(q))

    def setRigidWall(self, wall: Optional[RigidWall]):
        if wall is None:
            self._call("set_rigid_wall", 0, 0.0, 0.0)
        else:
            self._call("set_rigid_wall", 1, float(wall.plane), float(wall.friction))
        self.wall = wall

    def setContact(self, contact: Optional[Contact]):
        if contact is None:
            self._call("set_contact", 0, 0.0, 0.0, 0.1)
        else:
            self._call("set_contact", 1, float(contact.friction), float(contact.damping),
                       float(contact.scale))
            contact._solver = self
        self._contact = contact

    def setContactLimits(self, captureCapacity=-1, maxTerms=-1):
        """nf_contact_set_limits (HIP library only; no counterpart in the reference): caps
        the capture grid's entries (0 forces the reference's nine-cell walk) and the pair
        terms a node may take in one step; -1 restores a default."""
        self._call("contact_set_limits", int(captureCapacity), int(maxTerms))

    def contact(self):
        """The contact model, or None if none was set (:492-494)."""
        return self._contact

    def setGravity(self, ar, az):
        self._call("set_gravity", float(ar), float(az))

    def setDamage(self, fractureEnergy, failureStrain):
        """setDamage(double, double[]) :713-717 for a scalar fracture energy, setDamage(double[],
        double[]) :729-791 for one per element."""
        E = self.mesh.elementCount
        eps = np.ascontiguousarray(failureStrain, dtype=np.float64)
        if np.ndim(fractureEnergy) == 0:
            if eps.size != E:
                raise IllegalArgumentException(
                    f"failure strain has {eps.size} entries for a mesh of {E} elements")
            self._call("set_damage_uniform", float(fractureEnergy), _dptr(eps))
        else:
            gf = np.ascontiguousarray(fractureEnergy, dtype=np.float64)
            if gf.size != E:
                raise IllegalArgumentException(
                    f"fracture energy has {gf.size} entries for a mesh of {E} elements")
            if eps.size != E:
                raise IllegalArgumentException(
                    f"failure strain has {eps.size} entries for a mesh of {E} elements")
            self._call("set_damage", _dptr(gf), _dptr(eps))
        self._failureStrain = eps.copy()

    def damage(self):
        """The damage state (:794-796), or None if none was set: a read-only view of
        solver/Damage.java over the solver's current arrays."""
        if getattr(self, "_failureStrain", None) is None:
            return None
        return Damage(self)

    def setErosion(self, threshold):
        self._call("set_erosion", float(threshold))

    def setMassScaling(self, targetStep):
        self._call("set_mass_scaling", float(targetStep))

    def setVelocity(self, node, vr, vz):
        self._call("set_velocity", int(node), float(vr), float(vz))

    def setUniformVelocity(self, vr, vz):
        self._call("set_uniform_velocity", float(vr), float(vz))

    # run
    def step(self):
        self._call("run", 1)

    def run(self, steps):
        self._call("run", int(steps))

    def runUntil(self, endTime, maxSteps=0):
        taken = C.c_int64()
        self._call("run_until", float(endTime), int(maxSteps), C.byref(taken))
        return taken.value

    def advance(self):
        self._call("advance")

    # state
    def nodeField(self, name):
        out = np.empty(self.mesh.nodeCount, np.float64)
        self._call("get_node_field", _abi.NODE_FIELDS[name], _dptr(out))
        return out

    def putNodeField(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._call("put_node_field", _abi.NODE_FIELDS[name], _dptr(v))

    def pointField(self, name):
        st = self.state()
        P = st.element_count * st.points_per_element
        width = 4 if name in ("stress", "back") else 1
        out = np.empty(P * width, np.float64)
        self._call("get_point_field", _abi.POINT_FIELDS[name], _dptr(out))
        return out

    def elementField(self, name):
        out = np.empty(self.mesh.elementCount, np.float64)
        self._call("get_element_field", _abi.ELEMENT_FIELDS[name], _dptr(out))
        return out

    def state(self):
        st = _abi.StateC()
        self._call("get_state", C.byref(st))
        return st

    def audits(self):
        a = _abi.AuditsC()
        self._call("get_audits", C.byref(a))
        return a

    # Java accessor names, solver/ExplicitSolver.java:830-932, :1610-1860
    def radialDisplacement(self):
        return self.nodeField("ur")

    def axialDisplacement(self):
        return self.nodeField("uz")

    def velocityR(self):
        return self.nodeField("vr")

    def velocityZ(self):
        return self.nodeField("vz")

    def forceR(self):
        return self.nodeField("fr")

    def forceZ(self):
        return self.nodeField("fz")

    def stress(self):
        return self.pointField("stress")

    def plasticStrain(self):
        return self.pointField("plastic_strain")

    def plasticWorkDensity(self):
        return self.pointField("plastic_work")

    def time(self):
        return self.state().time

    def timestep(self):
        return self.state().dt

    def steps(self):
        return self.state().steps

    def massScale(self):
        return self.state().mass_scale

    def erodedElements(self):
        return self.state().eroded_elements

    def invertedElements(self):
        return self.state().inverted_elements

    def erodedEnergy(self):
        return self.state().eroded_energy

    def isEroded(self, e):
        return bool(self.elementField("eroded")[e])

    def axialMomentum(self):
        return self.audits().axial_momentum

    def radialMomentum(self):
        return self.audits().radial_momentum

    def kineticEnergy(self):
        return self.audits().kinetic_energy

    def centredKineticEnergy(self):
        return self.audits().centred_kinetic_energy

    def strainEnergy(self):
        return self.audits().strain_energy

    def plasticDissipation(self):
        return self.audits().plastic_dissipation

    def hourglassEnergy(self):
        return self.audits().hourglass_energy

    def fractureDissipation(self):
        return self.audits().fracture_dissipation

    def maxPlasticStrain(self):
        return self.audits().max_plastic_strain

    def yieldedFraction(self):
        return self.audits().yielded_fraction

    def maxTemperature(self):
        t = self.audits().max_temperature
        if math.isnan(t):
            raise IllegalStateException("this material has no thermal model; temperature is not defined")
        return t

    def maxDamage(self):
        return self.audits().max_damage

    def damagedFraction(self):  # :1748-1751
        return self.audits().damaged_fraction

    def failedFraction(self):  # :1753-1756
        return self.audits().failed_fraction

    def temperature(self, point):
        """Adiabatic temperature at a Gauss point (:1832-1840); IllegalStateException when
        its material has no Johnson-Cook law."""
        out = C.c_double()
        self._call("temperature", int(point), C.byref(out))
        return out.value

    def plasticFrontRadius(self):  # :1867-1877
        out = C.c_double()
        self._call("plastic_front_radius", C.byref(out))
        return out.value

    def mostDamagedElement(self):  # :1778-1794
        out = C.c_int32()
        self._call("most_damaged_element", C.byref(out))
        return out.value

    def borePressure(self):  # :826-828
        out = C.c_double()
        self._call("bore_pressure", C.byref(out))
        return out.value

    def meshZ(self, node):  # :1928-1930
        out = C.c_double()
        self._call("mesh_z", int(node), C.byref(out))
        return out.value

    def centroidRadius(self, e):  # :1904-1908, reference coordinates
        c = self.mesh.conn[4 * e:4 * e + 4]
        r = self.mesh.r
        return 0.25 * (r[c[0]] + r[c[1]] + r[c[2]] + r[c[3]])

    def centroidZ(self, e):  # :1917-1925, current configuration
        c = self.mesh.conn[4 * e:4 * e + 4]
        uz = self.axialDisplacement()
        total = 0.0
        for n in c:
            total += self.mesh.z[n] + uz[n]
        return 0.25 * total

    # RigidWall audits live in the solver's state on this side of the boundary.
    def wallImpulse(self):
        return self.state().wall_impulse

    def wallEnergyLoss(self):
        return self.state().wall_energy_loss

    def wallFrictionDissipation(self):
        return self.state().wall_friction_dissipation

    def wallContactCount(self):
        return self.state().wall_contact_count

    # Parameter sensitivities (nf_enable_sensitivity / nf_get_node_sensitivity): the
    # derivative of the discrete step loop with respect to SENSITIVITY_PARAMETERS.
    SENSITIVITY_PARAMETERS = ("friction", "A", "B", "n", "C", "m")

    def enableSensitivity(self):
        self._call("enable_sensitivity")

    def nodeSensitivity(self, field, parameter):
        """d(node field)/d(parameter) for UR, UZ, VR, VZ, FR or FZ; parameter is an index
        into SENSITIVITY_PARAMETERS or its name."""
        if isinstance(parameter, str):
            parameter = self.SENSITIVITY_PARAMETERS.index(parameter)
        out = np.empty(self.mesh.nodeCount)
        self._call("get_node_sensitivity", _abi.NODE_FIELDS[field], parameter, _dptr(out))
        return out

    def enableAdjoint(self, checkpointInterval=64):
        """Records checkpoints every `checkpointInterval` steps of the runs that follow, for
        adjointGradient (nf_adjoint_enable)."""
        self._call("adjoint_enable", int(checkpointInterval))

    def adjointGradient(self, seeds):
        """Reverse-mode gradients of up to two objectives of the final state: `seeds` is
        [objectives][4][nodes] (d objective / d UR, UZ, VR, VZ); returns [objectives][6] over
        SENSITIVITY_PARAMETERS."""
        sd = np.ascontiguousarray(seeds, dtype=np.float64)
        if sd.ndim == 2:
            sd = sd[None]
        count = sd.shape[0]
        if sd.shape[1:] != (4, self.mesh.nodeCount):
            raise IllegalArgumentException("seeds are [objectives][4][nodes]")
        out = np.zeros((count, len(self.SENSITIVITY_PARAMETERS)))
        self._call("adjoint_gradient", count, _dptr(sd), _dptr(out))
        return out

    def adjointInfo(self):
        """(recorded steps, checkpoints, capped-solve element passes of the last sweep,
        checkpoint interval)."""
        out = (C.c_int64 * 4)()
        self._call("adjoint_info", out)
        return tuple(out)

    def centroidStress(self, e):
        """Mean stress over an element's Gauss points {rr, zz, tt, rz} (:1886-1894)."""
        ppe = self.state().points_per_element
        sig = self.pointField("stress").reshape(-1, 4)
        out = np.zeros(4)
        for g in range(ppe):
            out += sig[e * ppe + g]
        return out / ppe

    def elementDamage(self, e):  # :1764-1772, with Damage.at (solver/Damage.java:342-347)
        fail = getattr(self, "_failureStrain", None)
        if fail is None:
            return 0.0
        ppe = self.state().points_per_element
        strength = self.pointField("damage_strength")
        softening = self.pointField("damage_softening")
        ep = self.pointField("plastic_strain")
        total = 0.0
        for g in range(ppe):
            p = e * ppe + g
            sf = strength[p]
            if sf <= 0.0:
                d = 0.0
            else:
                d = -softening[p] * (ep[p] - fail[e]) / sf
                d = 0.0 if d <= 0.0 else (1.0 if d >= 1.0 else d)
            total += d
        return total / ppe

    def elementSpeed(self, e):  # :895-902
        vr, vz = self.velocityR(), self.velocityZ()
        total = 0.0
        for n in self.mesh.conn[4 * e:4 * e + 4]:
            total += math.sqrt(vr[n] * vr[n] + vz[n] * vz[n])
        return 0.25 * total

    def elementPlasticStrain(self, e):  # :1897-1901
        ppe = self.state().points_per_element
        ep = self.pointField("plastic_strain")
        total = 0.0
        for g in range(ppe):
            total += ep[e * ppe + g]
        return total / ppe

    def stepSensitivity(self):
        """(d timestep / d parameter, d previous timestep / d parameter), six each."""
        out = np.empty(2 * len(self.SENSITIVITY_PARAMETERS))
        self._call("get_step_sensitivity", _dptr(out))
        return out[:6], out[6:]

    def contactSurface(self):
        """The contact's current surface segments (a, b), in surface order (element by
        element, mesh/QuadMesh.java:188-226, over the elements contact has not forgotten)."""
        n = C.c_int32()
        self._call("contact_surface", 0, None, C.byref(n))
        out = np.empty(max(1, 2 * n.value), np.int32)
        self._call("contact_surface", n.value, _iptr(out), C.byref(n))
        return out[:2 * n.value].reshape(-1, 2)

    def contactPairs(self):
        """(node, segment end a, segment end b) for every pair of the last step, node order."""
        if not hasattr(self._b, "contact_pairs"):
            raise IllegalStateException("this library does not report contact pairs")
        n = C.c_int32()
        self._call("contact_pairs", 0, None, C.byref(n))
        out = np.empty(max(1, 3 * n.value), np.int32)
        self._call("contact_pairs", n.value, _iptr(out), C.byref(n))
        return out[:3 * n.value].reshape(-1, 3)


class Damage:
    """solver/Damage.java's queries over a solver's damage arrays (fracture energy, failure
    strain, and the onset strength, softening modulus and band width written by the return
    map), read from the solver when the view is made."""

    def __init__(self, solver):
        self._fe = solver.pointField("damage_fracture_energy")
        self._fs = solver.pointField("damage_failure_strain")
        self._st = solver.pointField("damage_strength")
        self._so = solver.pointField("damage_softening")
        self._w = solver.pointField("damage_width")
        youngs = np.array([m.youngsModulus for m in solver.materials.distinct])
        ppe = solver.state().points_per_element
        idx = (np.zeros(solver.mesh.elementCount, np.int64) if solver.materials.index is None
               else np.asarray(solver.materials.index))
        self._E = np.repeat(youngs[idx], ppe)

    def fractureEnergy(self, p):  # :266-268
        return float(self._fe[p])

    def points(self):  # :270-272
        return int(self._fs.size)

    def hasBegun(self, p):  # :275-277
        return bool(self._st[p] > 0.0)

    def bandWidth(self, p):  # :280-282
        return float(self._w[p])

    def strengthAtOnset(self, p):  # :285-287
        return float(self._st[p])

    def softeningModulus(self, p):  # :290-292
        return float(self._so[p])

    def at(self, p, plasticStrain):  # :342-347
        sf = self._st[p]
        if sf <= 0.0:
            return 0.0
        d = -self._so[p] * (plasticStrain - self._fs[p]) / sf
        return 0.0 if d <= 0.0 else (1.0 if d >= 1.0 else float(d))

    def dissipation(self, p, plasticStrain):  # :363-372
        sf = self._st[p]
        if sf <= 0.0:
            return 0.0
        slope = self._so[p]
        full = -sf / slope
        excess = plasticStrain - self._fs[p]
        if excess <= 0.0:
            return 0.0
        if excess > full:
            excess = full
        return float(excess * (sf + 0.5 * slope * excess))

    def maximum(self, plasticStrain):  # :375-382
        m = 0.0
        for p in range(self.points()):
            d = self.at(p, plasticStrain[p])
            if d > m:
                m = d
        return m

    def begunFraction(self):  # :385-389
        return int(np.count_nonzero(self._st > 0.0)) / self.points()

    def failedFraction(self, plasticStrain):  # :392-398
        n = sum(1 for p in range(self.points()) if self.at(p, plasticStrain[p]) >= 1.0)
        return n / self.points()

    def clampedPoints(self):  # :406-412
        return int(np.count_nonzero(self._so == -self._E))


class Batch:
    """Solvers stepped together on the GPU (nf_batch_* in include/neofiz.h).

    Every instance keeps its own time step and end time, and its state stays bit-identical
    to a solo run. After runUntil, scatter() copies the state back into the solvers, whose
    accessors then read it. The solvers must share one library (one binding)."""

    def __init__(self, solvers: Sequence[ExplicitSolver]):
        if not solvers:
            raise IllegalArgumentException("a batch needs at least one solver")
        self._b = solvers[0]._b
        if any(s._b is not self._b for s in solvers):
            raise IllegalArgumentException("every solver of a batch must use the same library")
        self.solvers = list(solvers)
        handles = (_abi._P * len(solvers))(*[s._h for s in solvers])
        h = _abi._P()
        check(self._b, self._b.batch_create(C.byref(h), len(solvers), handles))
        self._h = h

    def setTile(self, instance, threads):
        """Threads per tile (64, 128 or 256) for one instance, or every instance (-1)."""
        check(self._b, self._b.batch_set_tile(self._h, instance, threads))

    def tile(self, instance):
        t = C.c_int32()
        check(self._b, self._b.batch_get_tile(self._h, instance, C.byref(t)))
        return t.value

    def runUntil(self, endTime, maxSteps=0):
        check(self._b, self._b.batch_run_until(self._h, endTime, maxSteps))

    def scatter(self):
        check(self._b, self._b.batch_scatter(self._h))

    def close(self):
        if self._h:
            self._b.batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

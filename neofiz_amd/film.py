"""Film: frames of a run for the reference's scene viewer, captured from the device.

render/Film.java:40-402 over the C ABI. `Film(mesh).colourBy(...).aliveIf(...).readout(...)`
declares what a frame holds, `capture(solver)` records one, and `json(title)` / `write(path,
title)` produce the reference's format (Film.java:148-224) for viewer/scene-viewer.html.

A frame's deformed node positions and every built-in element quantity come from one device
launch (nf_film_capture): PLASTIC_STRAIN, DAMAGE, SPEED, MISES, STRESS_RR/ZZ/TT/RZ, CENTROID_Z
and ALIVE, each the solver accessor the reference's scenes colour by, in its summation order.
The reference's per-element lambdas are accepted too: a callable (solver, e) -> float is
evaluated element by element on the host, as in Java; a callable (solver) -> array, or a fixed
array, once per frame. `transform` post-processes a built-in (for instance to MPa).
"""
import inspect
import math
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _abi
from .report import java_g
from .solver import IllegalArgumentException, check

AUTO = math.nan
RAMP_CLIP = 0.005  # Film.java:247


class Builtin:
    """An element quantity computed on the device (enum nf_film_field)."""

    def __init__(self, code, name):
        self.code = code
        self.name = name

    def __repr__(self):
        return f"Film.{self.name}"


PLASTIC_STRAIN = Builtin(0, "PLASTIC_STRAIN")
DAMAGE = Builtin(1, "DAMAGE")
SPEED = Builtin(2, "SPEED")
MISES = Builtin(3, "MISES")
STRESS_RR = Builtin(4, "STRESS_RR")
STRESS_ZZ = Builtin(5, "STRESS_ZZ")
STRESS_TT = Builtin(6, "STRESS_TT")
STRESS_RZ = Builtin(7, "STRESS_RZ")
CENTROID_Z = Builtin(8, "CENTROID_Z")
ALIVE = Builtin(9, "ALIVE")


@dataclass(frozen=True)
class Field:
    """Film.Field (Film.java:60-66)."""
    name: str
    units: str
    low: float
    high: float

    def isAuto(self):
        return math.isnan(self.low) or math.isnan(self.high)


@dataclass
class Frame:
    time: float
    r: np.ndarray
    z: np.ndarray
    values: list
    scalars: list
    live: Optional[str]


def _arity(fn):
    try:
        return len(inspect.signature(fn).parameters)
    except (TypeError, ValueError):
        return 2


class Film:
    AUTO = AUTO
    PLASTIC_STRAIN, DAMAGE, SPEED, MISES = PLASTIC_STRAIN, DAMAGE, SPEED, MISES
    STRESS_RR, STRESS_ZZ, STRESS_TT, STRESS_RZ = STRESS_RR, STRESS_ZZ, STRESS_TT, STRESS_RZ
    CENTROID_Z, ALIVE = CENTROID_Z, ALIVE
    Field = Field

    def __init__(self, mesh):  # :109-111
        self.mesh = mesh
        self.axisymmetric = True
        self.fields = []
        self.sources = []
        self.readouts = []
        self.meters = []
        self.alive = None
        self.frames = []

    def colourBy(self, name, units, low, high, value, transform: Optional[Callable] = None):  # :121-126
        self.fields.append(Field(name, units, float(low), float(high)))
        self.sources.append((value, transform))
        return self

    def aliveIf(self, test):  # :132-135
        self.alive = test
        return self

    def readout(self, name, value):  # :141-145
        self.readouts.append(name)
        self.meters.append(value)
        return self

    def planar(self):  # :209-212
        self.axisymmetric = False
        return self

    def size(self):
        return len(self.frames)

    # ---------------------------------------------------------------- capture

    def _device(self, solver, builtins):
        E, N = self.mesh.elementCount, self.mesh.nodeCount
        ids = np.array([b.code for b in builtins], dtype=np.int32)
        r, z = np.empty(N), np.empty(N)
        vals = np.empty((len(builtins), E))
        b = solver._b
        check(b, b.film_capture(solver._h, ids.size, ids.ctypes.data_as(_abi.c_int32_p),
                                r.ctypes.data_as(_abi.c_double_p), z.ctypes.data_as(_abi.c_double_p),
                                vals.ctypes.data_as(_abi.c_double_p)))
        return r, z, {bi.code: vals[k] for k, bi in enumerate(builtins)}

    def _evaluate(self, solver, value, device):
        E = self.mesh.elementCount
        if isinstance(value, Builtin):
            return device[value.code]
        if callable(value):
            if _arity(value) == 1:
                out = np.asarray(value(solver), dtype=np.float64)
            else:
                out = np.array([value(solver, e) for e in range(E)], dtype=np.float64)
        else:
            out = np.asarray(value, dtype=np.float64)
        if out.shape != (E,):
            raise IllegalArgumentException(f"a field needs one value per element ({E}), got {out.shape}")
        return out

    def capture(self, solver):  # :159-189
        """Records one frame: deformed positions, the declared fields, readouts and the live
        flags, with every built-in quantity from a single device launch."""
        wanted = []
        for value, _ in self.sources:
            if isinstance(value, Builtin) and value not in wanted:
                wanted.append(value)
        if isinstance(self.alive, Builtin) and self.alive not in wanted:
            wanted.append(self.alive)
        r, z, device = self._device(solver, wanted)
        values = []
        for value, transform in self.sources:
            v = self._evaluate(solver, value, device)
            values.append(np.asarray(transform(v), dtype=np.float64) if transform else v)
        scalars = [float(m(solver)) for m in self.meters]
        live = None
        if self.alive is not None:
            if isinstance(self.alive, Builtin):
                flags = device[self.alive.code] != 0.0
            elif _arity(self.alive) == 1:
                flags = np.asarray(self.alive(solver), dtype=bool)
            else:
                flags = np.array([bool(self.alive(solver, e)) for e in range(self.mesh.elementCount)])
            live = "".join("1" if f else "0" for f in flags)
        self.frames.append(Frame(solver.time(), r, z, values, scalars, live))

    # ---------------------------------------------------------------- output

    def write(self, path, title):  # :195-207, to a path of the caller's choosing
        with open(path, "w", encoding="utf-8") as out:
            out.write("window.NEOFIZ_FILM = ")
            out.write(self.json(title))
            out.write(";\n")
            out.write("(window.NEOFIZ_FILMS = window.NEOFIZ_FILMS || []).push(window.NEOFIZ_FILM);\n")
        return path

    def json(self, title):  # :218-282
        parts = ['{"title":"', _escape(title), '"',
                 ',"axisymmetric":', "true" if self.axisymmetric else "false",
                 ',"nodes":', str(self.mesh.nodeCount),
                 ',"elements":', str(self.mesh.elementCount),
                 ',"conn":[', ",".join(str(int(c)) for c in self.mesh.conn), "]"]
        boundary = self.mesh.surface().edges
        parts += [',"boundary":[', ",".join(str(int(c)) for c in boundary), "]"]
        parts.append(',"fields":[')
        for f, field in enumerate(self.fields):
            rng = self._observed_range(f) if field.isAuto() else (field.low, field.high, field.high)
            if f > 0:
                parts.append(",")
            parts += ['{"name":"', _escape(field.name), '","units":"', _escape(field.units),
                      '","low":', number(rng[0]), ',"high":', number(rng[1]),
                      ',"peak":', number(rng[2]), ',"auto":', "true" if field.isAuto() else "false", "}"]
        parts.append("]")
        parts.append(',"readouts":[' + ",".join('"' + _escape(n) + '"' for n in self.readouts) + "]")
        parts.append(',"frames":[')
        for i, frame in enumerate(self.frames):
            if i > 0:
                parts.append(",")
            parts += ['{"t":', number(frame.time), _array(',"r":', frame.r), _array(',"z":', frame.z),
                      ',"v":[', ",".join(_array("", v) for v in frame.values), "]",
                      _array(',"s":', frame.scalars)]
            if frame.live is not None:
                parts += [',"live":"', frame.live, '"']
            parts.append("}")
        parts.append("]}")
        return "".join(parts)

    def _observed_range(self, field):  # :249-276
        finite = [v[np.isfinite(v)] for v in (fr.values[field] for fr in self.frames)]
        allv = np.sort(np.concatenate(finite)) if finite else np.empty(0)
        n = allv.size
        if n == 0:
            return (0.0, 1.0, 1.0)
        low, peak = float(allv[0]), float(allv[n - 1])
        above = int(math.floor(n * RAMP_CLIP))
        high = float(allv[max(0, n - 1 - above)])
        if high <= low:
            high = peak
        if high - low < 1e-12 * max(1.0, abs(high)):
            return (low, low + max(1e-9, abs(low) * 1e-6), peak)
        return (low, high, peak)


def _escape(text):  # :399-401
    return text.replace("\\", "\\\\").replace('"', '\\"')


def _array(key, values):  # :286-293
    return key + "[" + ",".join(number(float(v)) for v in values) + "]"


def number(value):
    """Film.number (Film.java:295-307): Java's String.format("%.6g") (report.java_g) with
    trailing zeros and a trailing point trimmed and the exponent written as a plain integer;
    non-finite values and zero print as 0. Java's digit generator is not the shortest for
    subnormal doubles (it prints Double.MIN_VALUE as 4.9e-324, this as 5e-324); no film
    quantity is subnormal."""
    if not math.isfinite(value) or value == 0.0:
        return "0"
    text = java_g(float(value), 6)
    if "e" in text:
        mant, exp = text.split("e")
        return _trim(mant) + "e" + str(int(exp))
    return _trim(text) if "." in text else text


def _trim(text):  # :309-314
    end = len(text)
    while end > 0 and text[end - 1] == "0":
        end -= 1
    if end > 0 and text[end - 1] == ".":
        end -= 1
    return text[:end]

"""Shapes meshed onto a square lattice: the host-side mirror of mesh/Outline.java.

An `Outline` is a closed polygon with optional holes; containment is the even-odd rule, so
neither winding nor which ring is a hole matters (mesh/Outline.java:160-185). Meshing keeps
every lattice cell whose centre lies inside a shape. The lattice is anchored at the origin,
so a point's cell depends on the point alone:

  mesh(h, *shapes)      one body, shared nodes, union of the shapes (:289-350)
  assemble(h, *shapes)  each shape a body of its own with its own nodes, on the same lattice
                        (:254-287), which is how stacked bodies begin exactly touching.

Node numbering is first appearance while the cells are visited row by row (j outer, i inner)
and each cell emits its corners counter-clockwise from the lower left (:321-345, :352-363).
The numbering is part of the mesh's identity: the gather order and the contact surface's
order follow from it.
"""
import math

import numpy as np

from .solver import IllegalArgumentException, QuadMesh

_MAX_CELLS = 20_000_000  # :311-315


def _crosses(ring, r, z):
    """Even-odd ray test of one closed ring (:174-185)."""
    inside = False
    n = len(ring) // 2
    j = n - 1
    for i in range(n):
        ri, zi = ring[2 * i], ring[2 * i + 1]
        rj, zj = ring[2 * j], ring[2 * j + 1]
        if (zi > z) != (zj > z) and r < (rj - ri) * (z - zi) / (zj - zi) + ri:
            inside = not inside
        j = i
    return inside


def _shoelace(ring):
    a = 0.0
    n = len(ring) // 2
    j = n - 1
    for i in range(n):
        a += ring[2 * j] * ring[2 * i + 1] - ring[2 * i] * ring[2 * j + 1]
        j = i
    return 0.5 * a


class Outline:
    """mesh/Outline.java:38-364."""

    def __init__(self, ring, holes=()):
        self.ring = [float(x) for x in ring]
        self.holes = [[float(x) for x in h] for h in holes]

    # ------------------------------------------------------------ construction (:68-95)

    @staticmethod
    def of(*rz):
        if len(rz) == 1 and not isinstance(rz[0], (int, float)):
            rz = tuple(rz[0])
        if len(rz) < 6 or len(rz) % 2:
            raise IllegalArgumentException(
                f"a polygon needs at least three (r, z) pairs; got {len(rz)} numbers")
        return Outline(rz)

    @staticmethod
    def rectangle(r0, z0, r1, z1):
        if r1 <= r0 or z1 <= z0:
            raise IllegalArgumentException("rectangle needs r1 > r0 and z1 > z0")
        return Outline.of(r0, z0, r1, z0, r1, z1, r0, z1)

    @staticmethod
    def circle(centreR, centreZ, radius, sides):
        if radius <= 0.0:
            raise IllegalArgumentException("radius must be positive")
        if sides < 3:
            raise IllegalArgumentException("need at least three sides")
        rz = []
        for k in range(sides):
            a = 2.0 * math.pi * k / sides
            rz += [centreR + radius * math.cos(a), centreZ + radius * math.sin(a)]
        return Outline(rz)

    # ------------------------------------------------------------ rigid moves (:97-157)

    def _mapped(self, move):
        def each(ring):
            out = []
            for i in range(0, len(ring), 2):
                out += list(move(ring[i], ring[i + 1]))
            return out
        return Outline(each(self.ring), [each(h) for h in self.holes])

    def translated(self, dr, dz):
        return self._mapped(lambda r, z: (r + dr, z + dz))

    def rotated(self, radians, aboutR, aboutZ):
        c, s = math.cos(radians), math.sin(radians)

        def move(r, z):
            dr, dz = r - aboutR, z - aboutZ
            return aboutR + c * dr - s * dz, aboutZ + s * dr + c * dz
        return self._mapped(move)

    def restingOn(self, z):
        lowest = min(self.ring[1::2])
        return self.translated(0.0, z - lowest)

    def centre(self):
        b = self.bounds()
        return [0.5 * (b[0] + b[2]), 0.5 * (b[1] + b[3])]

    def withHole(self, hole):
        return Outline(self.ring, self.holes + [list(hole.ring)] + [list(h) for h in hole.holes])

    # ------------------------------------------------------------ queries (:166-224)

    def contains(self, r, z):
        inside = _crosses(self.ring, r, z)
        for h in self.holes:
            if _crosses(h, r, z):
                inside = not inside
        return inside

    def area(self):
        a = abs(_shoelace(self.ring))
        for h in self.holes:
            a -= abs(_shoelace(h))
        return a

    def bounds(self):
        rs, zs = self.ring[0::2], self.ring[1::2]
        return [min(rs), min(zs), max(rs), max(zs)]

    # ------------------------------------------------------------ meshing

    def meshOne(self, cellSize):
        """outline.mesh(cellSize): this one shape on the lattice (:227-230)."""
        return Outline.meshAll(cellSize, self)

    @staticmethod
    def meshAll(cellSize, *shapes):
        """Outline.mesh(cellSize, shapes...) (:289-350): the union of the shapes, one body."""
        if not (cellSize > 0.0):
            raise IllegalArgumentException(f"cell size must be positive, was {cellSize}")
        if not shapes:
            raise IllegalArgumentException("nothing to mesh")
        rlo = min(s.bounds()[0] for s in shapes)
        zlo = min(s.bounds()[1] for s in shapes)
        rhi = max(s.bounds()[2] for s in shapes)
        zhi = max(s.bounds()[3] for s in shapes)
        i0, i1 = math.floor(rlo / cellSize), math.ceil(rhi / cellSize)
        j0, j1 = math.floor(zlo / cellSize), math.ceil(zhi / cellSize)
        if (i1 - i0) * (j1 - j0) > _MAX_CELLS:
            raise IllegalArgumentException(
                f"a cell size of {cellSize} m over this bounding box would need more than "
                "twenty million cells; the shape is large or the cell small")

        # Cell centres inside any shape, vectorised per shape with the same arithmetic as
        # the scalar even-odd test.
        ii = np.arange(i0, i1)
        jj = np.arange(j0, j1)
        cr = (ii + 0.5) * cellSize
        cz = (jj + 0.5) * cellSize
        CR, CZ = np.meshgrid(cr, cz)  # [rows j, cols i]
        keep = np.zeros(CR.shape, dtype=bool)
        for s in shapes:
            inside = _crosses_grid(s.ring, CR, CZ)
            for h in s.holes:
                inside ^= _crosses_grid(h, CR, CZ)
            keep |= inside
        if not keep.any():
            raise IllegalArgumentException(
                "no cell centre fell inside the outline; the shape is smaller than one "
                f"{cellSize} m cell, or the coordinates are not in metres")

        # First-appearance numbering of lattice points, cell by cell (:352-363).
        node_at = {}
        r, z, conn = [], [], []

        def node(i, j):
            k = (i, j)
            at = node_at.get(k)
            if at is None:
                at = len(r)
                node_at[k] = at
                r.append(i * cellSize)
                z.append(j * cellSize)
            return at

        rows, cols = np.nonzero(keep)
        for jr, ic in zip(rows.tolist(), cols.tolist()):
            i, j = i0 + ic, j0 + jr
            conn += [node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)]
        return QuadMesh(np.array(r), np.array(z), np.array(conn, dtype=np.int32))

    @staticmethod
    def assemble(cellSize, *shapes):
        """Outline.assemble (:254-287): each shape meshed alone, then concatenated with its
        node numbers offset, so touching shapes get coincident but separate nodes."""
        if not shapes:
            raise IllegalArgumentException("nothing to mesh")
        rs, zs, cs = [], [], []
        nodes = 0
        for s in shapes:
            body = Outline.meshAll(cellSize, s)
            rs.append(body.r)
            zs.append(body.z)
            cs.append(body.conn + nodes)
            nodes += body.nodeCount
        return QuadMesh(np.concatenate(rs), np.concatenate(zs),
                        np.concatenate(cs).astype(np.int32))


class _MeshMethod:
    """Outline.mesh is both the instance method mesh(cellSize) and the static
    mesh(cellSize, shapes...) in the reference; this dispatches on how it is reached."""

    def __get__(self, obj, cls):
        if obj is None:
            return Outline.meshAll
        return obj.meshOne


Outline.mesh = _MeshMethod()


def _crosses_grid(ring, R, Z):
    inside = np.zeros(R.shape, dtype=bool)
    n = len(ring) // 2
    j = n - 1
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(n):
            ri, zi = ring[2 * i], ring[2 * i + 1]
            rj, zj = ring[2 * j], ring[2 * j + 1]
            straddles = (zi > Z) != (zj > Z)
            if straddles.any():
                edge_r = (rj - ri) * (Z - zi) / (zj - zi) + ri
                inside ^= straddles & (R < edge_r)
            j = i
    return inside

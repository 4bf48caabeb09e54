"""The plane-strain multi-body perforation scene: a steel slug into a stacked copper wall.

The reference's template is Smash (Smash.java:82-211): rectangles meshed as separate bodies
on one lattice (Outline.assemble), steel slug elements picked by centroid, plane strain,
reduced integration, finite strain, gravity, a frictional floor, penalty contact between all
bodies, damage with erosion. The scene here has the workload's proportions: twelve copper
blocks 20 mm (r) x 10 mm (z) stacked on the floor, a 10 mm x 20 mm steel slug at 700 m/s
aimed at the third and fourth blocks, contact friction 0.2, floor friction 0.4, fracture
energy 10 kJ/m^2, erosion at full damage, no mass scaling.

The copper's failure strain is the reference's spatially correlated defect field
(core/DefectField.java:92-385), sampled on the device: Weibull modulus 10 and scale 1.2 at a
reference volume of 1 mm^3, correlation length 1 mm, seed 5, with the plane-strain volume rule
(each element holds l^3 max(1, A/l^2) max(1, depth/l), twenty correlation lengths through the
20 mm depth), which puts the elements' characteristic failure strain near 0.89. The steel slug
fails at 0.6.
"""
import numpy as np

from .mesh import Outline
from .solver import (Contact, ExplicitSolver, FINITE_STRAIN, JohnsonCook, Material, Materials,
                     PLANE_STRAIN, REDUCED, RigidWall)

MM = 1.0e-3
BLOCK_R, BLOCK_Z = 20.0 * MM, 10.0 * MM
BLOCKS = 12
SLUG_R, SLUG_Z = 10.0 * MM, 20.0 * MM
SLUG_CENTRE_Z = 3.0 * BLOCK_Z
GAP = 1.0 * MM
SPEED = 700.0
DEPTH = 20.0 * MM
STEEL = Material.STEEL_4340.withJohnsonCook(JohnsonCook.STEEL_4340)
COPPER = Material.COPPER_OFHC.withJohnsonCook(JohnsonCook.COPPER_OFHC)
FRACTURE_ENERGY = 10.0e3
WEIBULL_MODULUS = 10.0
COPPER_STRAIN_SCALE = 1.2        # Weibull scale at the reference volume
REFERENCE_VOLUME = 1.0e-9
CORRELATION_LENGTH = 1.0e-3
SLUG_FAILURE_STRAIN = 0.6


def outlines(blocks=BLOCKS):
    wall = [Outline.rectangle(0.0, j * BLOCK_Z, BLOCK_R, (j + 1) * BLOCK_Z) for j in range(blocks)]
    slug = Outline.rectangle(-GAP - SLUG_R, SLUG_CENTRE_Z - 0.5 * SLUG_Z, -GAP,
                             SLUG_CENTRE_Z + 0.5 * SLUG_Z)
    return wall, slug


def cell_for(elements, blocks=BLOCKS):
    """The lattice cell that gives about `elements` elements over the scene."""
    area = blocks * BLOCK_R * BLOCK_Z + SLUG_R * SLUG_Z
    return float(np.sqrt(area / elements))


def failure_strains(mesh, is_slug, seed=5, binding=None):
    """The copper's defect field on the mesh (the slug's elements get its fixed strain)."""
    from .defects import DefectField, Weibull
    field = DefectField(Weibull(COPPER_STRAIN_SCALE, WEIBULL_MODULUS), REFERENCE_VOLUME,
                        CORRELATION_LENGTH, seed, binding=binding)
    try:
        eps = field.sampleOnto(mesh, depth=DEPTH)
    finally:
        field.close()
    return np.where(is_slug, SLUG_FAILURE_STRAIN, eps)


def build(cell, solver_cls=ExplicitSolver, blocks=BLOCKS, seed=5, speed=SPEED, field_binding=None):
    """Returns (mesh, solver, is_slug). field_binding selects the library that samples the
    defect field (the HIP library by default)."""
    wall, slug = outlines(blocks)
    mesh = Outline.assemble(cell, *wall, slug)
    # The slug is the last body assembled; its elements are the ones whose centroid lies in
    # its rectangle (Smash.java:112-115), which for axis-aligned lattice cells is a box test.
    c = mesh.conn.reshape(-1, 4)
    rc = 0.25 * mesh.r[c].sum(axis=1)
    zc = 0.25 * mesh.z[c].sum(axis=1)
    r0, z0, r1, z1 = slug.bounds()
    is_slug = (rc > r0) & (rc < r1) & (zc > z0) & (zc < z1)
    mats = Materials.byElement(mesh.elementCount, lambda e: STEEL if is_slug[e] else COPPER)
    s = solver_cls(mesh, mats, PLANE_STRAIN, REDUCED, FINITE_STRAIN, DEPTH, 0.5)
    s.setGravity(0.0, -9.81)
    s.setRigidWall(RigidWall.atZ(0.0, 0.4))
    s.setContact(Contact(mesh).withFriction(0.2))
    eps = failure_strains(mesh, is_slug, seed, field_binding)
    s.setDamage(np.full(mesh.elementCount, FRACTURE_ENERGY), eps)
    s.setErosion(1.0)
    v = np.zeros(mesh.nodeCount)
    v[mesh.r < 0.0] = speed
    s.putNodeField("vr", v)
    return mesh, s, is_slug

"""neofiz_amd -- Neofiz's explicit-dynamics step loop on MI355X (gfx950, FP64 HIP kernels).

The compute lives in libneofiz_hip.so (C ABI in include/neofiz.h); this package is the
host-side mirror of the reference's Java solver API over that ABI.
"""
from .solver import (AXISYMMETRIC, FINITE_STRAIN, FULL, PLANE_STRAIN, PLANE_STRESS, REDUCED,
                     SMALL_STRAIN, Contact, ExplicitSolver, IllegalArgumentException,
                     IllegalStateException, JohnsonCook, Material, Materials, QuadMesh,
                     RigidWall)

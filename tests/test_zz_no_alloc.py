"""nf_defect_sample_onto_static, the sampling call that allocates nothing (its scratch is the
library's static device array), against nf_defect_sample_onto_device on the same inputs: bit
for bit, on two calls in a row (the second smaller), on a box too large for the static grid
(hashed taps), and its size limit."""
import ctypes as C

import numpy as np
import pytest

from conftest import same_bits
from test_gpu_defect_field import BASE, LENGTH, V0, DeviceArrays, MESHES
from neofiz_amd import _abi
from neofiz_amd.defects import DefectField
from neofiz_amd.solver import QuadMesh

pytestmark = pytest.mark.gpu

RING, SLAB = 0, 1


def call(field, name, mesh, seeds, mem, depth=None):
    """Runs entry point `name` on device copies of `mesh`; returns its output."""
    b = field._b
    r, z = mem.put(mesh.r), mem.put(mesh.z)
    conn = mem.put(mesh.conn.astype(np.int32))
    out = mem.put(np.full((seeds.size, mesh.elementCount), np.nan))
    rc = getattr(b, name)(field._h, seeds.size, seeds.ctypes.data_as(C.POINTER(C.c_int64)),
                          mesh.nodeCount, r, z, mesh.elementCount, conn,
                          RING if depth is None else SLAB, float(depth or 0.0), 0, out)
    assert rc == 0, b.last_error().decode()
    return mem.get(out, (seeds.size, mesh.elementCount))


def test_static_scratch_call_equals_the_allocating_call():
    field = DefectField(BASE, V0, LENGTH, 21)
    mem = DeviceArrays()
    try:
        print("[no-alloc] bodies mesh, 3 seeds, slab rule", flush=True)
        big = MESHES["bodies"]()
        seeds = np.array([21, 22, 23], dtype=np.int64)
        want = call(field, "defect_sample_onto_device", big, seeds, mem, depth=0.02)
        got = call(field, "defect_sample_onto_static", big, seeds, mem, depth=0.02)
        assert np.isfinite(want).all() and same_bits(got, want)
        print("[no-alloc] jittered mesh, 1 seed, ring rule (second, smaller call)", flush=True)
        small = MESHES["jittered"]()
        one = np.array([5], dtype=np.int64)
        want = call(field, "defect_sample_onto_device", small, one, mem)
        got = call(field, "defect_sample_onto_static", small, one, mem)
        assert same_bits(got, want)
        print("[no-alloc] a tall wall whose grid exceeds the static array: hashed taps", flush=True)
        tall = QuadMesh.cylinderWall(0.05, 0.10, 5.0, 4, 400)
        want = call(field, "defect_sample_onto_device", tall, one, mem)
        got = call(field, "defect_sample_onto_static", tall, one, mem)
        assert same_bits(got, want)
        too_many = np.arange(16385, dtype=np.int64)
        rc = field._b.defect_sample_onto_static(field._h, too_many.size,
                                                too_many.ctypes.data_as(C.POINTER(C.c_int64)),
                                                small.nodeCount, None, None, small.elementCount,
                                                None, RING, 0.0, 0, None)
        assert rc == -1 and b"exceeds" in field._b.last_error()
    finally:
        mem.free()
        field.close()

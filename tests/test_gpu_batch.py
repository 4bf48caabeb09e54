"""The batched ensemble against the oracle: every instance of a batch, stepped together on
the GPU with its own time step, ends bit-identical to the same rod run alone on the CPU."""
import numpy as np
import pytest

from conftest import first_difference
from neofiz_amd import ensemble
from neofiz_amd.solver import Batch

pytestmark = pytest.mark.gpu

FIELDS = ("ur", "uz", "vr", "vz", "fr", "fz", "on_wall")


def small_ensemble(count, seed):
    """Instances drawn as in the ensemble generator, meshes limited to 8x80 and 16x160."""
    out = []
    for inst in ensemble.draw(count * 4, seed):
        if inst.nr <= 16:
            out.append(inst)
        if len(out) == count:
            return out
    raise AssertionError("not enough small meshes drawn")


@pytest.mark.parametrize("tiles", [(256,), (64, 128, 256)])
def test_batch_matches_solo_runs_bitwise(tiles):
    from oracle.pyoracle import OracleSolver
    insts = small_ensemble(6, 11)
    built = [ensemble.build(i) for i in insts]
    batch = Batch([s for _, s in built])
    for k in range(len(insts)):
        batch.setTile(k, tiles[k % len(tiles)])
    end = 12e-6
    print(f"batch of {len(insts)} to {end}", flush=True)
    batch.runUntil(end, 300)          # a capped first call, then the rest
    batch.runUntil(end)
    batch.scatter()
    for k, inst in enumerate(insts):
        mesh, ref = ensemble.build(inst, OracleSolver)
        ref.runUntil(end, 300)
        ref.runUntil(end)
        gpu = built[k][1]
        assert gpu.steps() == ref.steps(), f"instance {k}: {gpu.steps()} vs {ref.steps()} steps"
        assert gpu.time() == ref.time() and gpu.timestep() == ref.timestep()
        for f in FIELDS:
            assert first_difference(gpu.nodeField(f), ref.nodeField(f)) is None, f"instance {k}: {f}"
        for f in ("stress", "plastic_strain", "plastic_work"):
            assert first_difference(gpu.pointField(f), ref.pointField(f)) is None, f"instance {k}: {f}"
        ref.shutdown()
    batch.close()


def test_batch_refuses_contact_and_damage():
    from neofiz_amd.solver import Contact, IllegalStateException
    inst = ensemble.draw(1, 3)[0]
    mesh, s = ensemble.build(inst)
    s.setDamage(np.full(mesh.elementCount, 1e4), np.full(mesh.elementCount, 0.5))
    with pytest.raises(IllegalStateException):
        Batch([s])


def test_batch_with_the_large_meshes_bitwise():
    """A batch holding a 32 x 320 and a 64 x 640 rod of the ensemble's draw (with 8 x 80 and
    16 x 160 beside them), 6000 steps each: every instance equals its solo oracle run."""
    import os
    from oracle.pyoracle import OracleSolver
    picks = {}
    for inst in ensemble.draw(64, 5):
        picks.setdefault((inst.nr, inst.nz), inst)
    insts = [picks[m] for m in ((8, 80), (16, 160), (32, 320), (64, 640))]
    built = [ensemble.build(i) for i in insts]
    batch = Batch([s for _, s in built])
    print("batch of 4 (8x80 .. 64x640), 6000 steps", flush=True)
    batch.runUntil(80e-6, 6000)
    batch.scatter()
    for k, inst in enumerate(insts):
        mesh, ref = ensemble.build(inst, OracleSolver)
        ref.setThreads(min(16, os.cpu_count() or 1))
        ref.runUntil(80e-6, 6000)
        gpu = built[k][1]
        assert gpu.steps() == ref.steps() == 6000, inst
        assert gpu.time() == ref.time() and gpu.timestep() == ref.timestep()
        for f in FIELDS:
            assert first_difference(gpu.nodeField(f), ref.nodeField(f)) is None, f"{inst.nr}x{inst.nz}: {f}"
        for f in ("stress", "plastic_strain", "plastic_work"):
            assert first_difference(gpu.pointField(f), ref.pointField(f)) is None, f"{inst.nr}x{inst.nz}: {f}"
        ref.shutdown()
    batch.close()


def test_capped_instance_collapses_in_the_oracle_too():
    """Many ensemble draws do not reach 80 us within 20,000 steps because the stable step
    collapses as the impact corner crushes. The fastest 64 x 640 rod of the ensemble's draw,
    run on the GPU and by the oracle (the reference's arithmetic) in chunks of 1000 steps:
    the step sizes agree bit for bit at every chunk, and both show the collapse (the step at
    20,000 steps is a small fraction of the first, and 80 us is not reached)."""
    import os
    from oracle.pyoracle import OracleSolver
    insts = [i for i in ensemble.draw(256) if i.nr == 64]
    inst = max(insts, key=lambda i: i.speed)
    traces = []
    for cls in (None, OracleSolver):
        mesh, s = ensemble.build(inst) if cls is None else ensemble.build(inst, cls)
        if cls is not None:
            s.setThreads(min(16, os.cpu_count() or 1))
        dt = [s.timestep()]
        for _ in range(20):
            s.runUntil(80e-6, 1000)
            dt.append(s.timestep())
        traces.append((np.array(dt), s.time(), s.steps()))
        print(f"{'oracle' if cls else 'gpu'}: dt {dt[0]:.3e} -> {dt[-1]:.3e} s over {s.steps()} steps, "
              f"t = {s.time() * 1e6:.2f} us", flush=True)
        s.shutdown()
    (dg, tg, ng), (dr, tr, nr_) = traces
    assert np.array_equal(dg.view(np.uint64), dr.view(np.uint64)) and tg == tr and ng == nr_ == 20000
    assert dr[-1] < 0.1 * dr[0] and tr < 80e-6


def test_persistent_batch_equals_graph_batch():
    """The persistent form of the batch (NEOFIZ_BATCH_PERSISTENT=1 at creation: 32 steps of
    every live instance in one cooperative launch, grid-wide barriers between the element and
    node walks) against the default graph-captured launches: every instance bit-identical,
    over mixed mesh sizes and a capped first call."""
    import os
    insts = [i for i in ensemble.draw(40, 21) if i.nr <= 32][:8]
    out = []
    for persistent in (False, True):
        old = os.environ.get("NEOFIZ_BATCH_PERSISTENT")
        os.environ["NEOFIZ_BATCH_PERSISTENT"] = "1" if persistent else "0"
        try:
            built = [ensemble.build(i) for i in insts]
            batch = Batch([s for _, s in built])
        finally:
            if old is None:
                del os.environ["NEOFIZ_BATCH_PERSISTENT"]
            else:
                os.environ["NEOFIZ_BATCH_PERSISTENT"] = old
        batch.runUntil(10e-6, 250)
        batch.runUntil(10e-6)
        batch.scatter()
        batch.close()
        out.append([(s.steps(), s.time(), {f: s.nodeField(f) for f in FIELDS},
                     {f: s.pointField(f) for f in ("stress", "plastic_strain", "plastic_work")})
                    for _, s in built])
        for _, s in built:
            s.shutdown()
    for k, (a, b) in enumerate(zip(*out)):
        assert a[0] == b[0] and a[1] == b[1], f"instance {k}: steps/time"
        for f in a[2]:
            assert first_difference(a[2][f], b[2][f]) is None, f"instance {k}: {f}"
        for f in a[3]:
            assert first_difference(a[3][f], b[3][f]) is None, f"instance {k}: {f}"

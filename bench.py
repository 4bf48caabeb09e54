#!/usr/bin/env python3
"""Benchmark: FP64 element-updates per second of the explicit step loop on one MI355X.

Workload (the HBM-roofline case): Taylor impact of a 25.4 mm x 7.62 mm copper rod at
190 m/s on a rigid anvil with Coulomb friction mu = 0.09, Johnson-Cook flow stress
(A 90 MPa, B 292 MPa, n 0.31, C 0.025, m 1.09, Tm 1356 K, beta 0.9), axisymmetric, reduced
integration with hourglass control, finite strain, CFL 0.5, meshed 128 x 1280
(163,840 elements, 165,249 nodes). One step is one explicit time step of the whole mesh:
nodal update, element pass, ordered gather, adaptive dt.

`python bench.py [--gpus 1] [--steps K] [--warmup W]` runs W untimed steps from t = 0, then
times exactly K steps with the device synchronised on both sides, and prints one JSON line.
Inputs are resident in HBM before the timed region starts.

Extra objects on the line:
  roofline      the dominant kernel's algorithmic bytes per launch over its mean launch
                duration on the timed steps (replayed on an identical twin of the rod, whose
                trajectory is bit-identical, with start/stop events recorded by each dispatch
                on the library's stream), against the 8 TB/s HBM peak; traffic (HBM bytes per
                launch) from the committed rocprofv3 PMC summary. The dominant kernel is the
                fused step kernel k_step (the node update and the element pass in one launch,
                every step of a run call after its first), so its bytes are the whole step's.
  full_run      the same rod from t = 0 to 80 us: steps, element-updates/s, HBM fraction,
                L/L0 and mushroom diameter, so the figure does not rest on one window.
  ladder        the rods of the refinement ladder to 80 us: 8x80 linear hardening without
                friction, 16x160 to 64x640 Johnson-Cook with mu = 0.09.
  cpu_baseline  the oracle (the C++ restatement, threaded like the reference's Parallel
                pool) on the host's cores: warm-up, then the best of three bursts.
  defect_field  device-side defect-field seeding of the ~200k-element perforation assembly
                (kernel time per call, inputs resident) and the size-effect population of
                3000 seeds, with the serial oracle on the same mesh beside it.
  perforation, sensitivity, batched
                the multi-body contact scene, the derivative's cost, the 256-instance
                ensemble against one-by-one runs (and its instances that reach 80 us alone).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "element_kernel_traffic.txt")
STEP_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_kernel_traffic.txt")


def element_kernel_bytes(E, N, kinematic=False):
    """Algorithmic bytes one launch of the reduced-integration element kernel must move.

    Per element: connectivity 16; point state read and written 2 x 64 (stress 32, plastic
    strain 8, plastic work 8, hourglass force 16), plus 2 x 32 of back stress when some
    material hardens kinematically (without it the back stress is +0 forever and the kernel
    neither loads nor stores it); hourglass stiffness written 8; four force slots written 64.
    Per node, read once: reference coordinates 16, displacement 16, velocity 16.
    """
    state = 96 if kinematic else 64
    return E * (16 + 2 * state + 8 + 64) + N * 48


def step_bytes(E, N):
    """Algorithmic bytes of one whole step (element + gather + nodal):
    B = 352 + 160 N/E per element-step for reduced integration with a rigid wall."""
    return 352.0 * E + 160.0 * N


def step_kernel_bytes(E, N, kinematic=False):
    """Algorithmic bytes one launch of the fused step kernel must move: the whole step's
    (step_bytes), less the back stress read and written (64 per element) when no material
    hardens kinematically and the kernel neither loads nor stores it."""
    return step_bytes(E, N) - (0.0 if kinematic else 64.0 * E)


def cpu_baseline(nr, nz, seconds, threads, bursts=3):
    """The oracle on the host's cores, measured as the reference measures (M0Cylinder.java:
    931-942, Bench.java:123-132): a warm-up, then `bursts` timed bursts of about
    `seconds` / bursts each, continuing the same run from t = 0; the best burst's rate."""
    from neofiz_amd import taylor
    from oracle.pyoracle import OracleSolver
    mesh, solver, _ = taylor.build(taylor.ROD_JC, taylor.ROD_LENGTH, taylor.ROD_DIAMETER,
                                   taylor.ROD_SPEED, nr, nz, friction=taylor.ROD_JC_FRICTION,
                                   solver_cls=OracleSolver)
    solver.setThreads(threads)
    solver.run(20)  # warm-up: the pool, the caches, the impact onset
    rates = []
    for _ in range(bursts):
        steps = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < seconds / bursts:
            solver.run(5)
            steps += 5
        rates.append(mesh.elementCount * steps / (time.perf_counter() - t0))
    last = solver.steps()
    solver.shutdown()
    return max(rates), rates, last


def host_cpu():
    """(CPU model, logical CPUs of the host, CPUs this process may run on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def stream_copy_gbps():
    """Device-to-device copy bandwidth of a 2 GiB buffer (read + write bytes / time), by the
    library's plain copy kernel."""
    import ctypes as C
    from neofiz_amd import _native
    b = _native.binding()
    gbps, ms = C.c_double(), C.c_double()
    rc = b.copy_bandwidth(1 << 31, 10, C.byref(gbps), C.byref(ms))
    if rc != 0:
        raise RuntimeError(b.last_error().decode())
    return gbps.value


def ensemble_leg(count, cap, per_size):
    """The batched ensemble against one-by-one on the same GPU.

    Batched: all `count` instances as one batch, each to 80 us or `cap` steps. One-by-one:
    `per_size` instances of each mesh size through the single-instance path with the same
    end test; the one-by-one time of the whole ensemble is then estimated per mesh size as
    (measured seconds per step) x (the steps every instance of that size took, known exactly
    from the batched run, whose states are bit-identical to solo runs).

    `reached`: the same comparison restricted to the instances that reach 80 us within the
    cap, as the ensemble is specified (each instance stops at 80 us): those instances run
    again as a batch of their own with no cap, and one by one, `per_size` of each mesh size
    measured and the rest estimated from them as above."""
    from neofiz_amd import ensemble, taylor
    insts = ensemble.draw(count)
    print(f"# ensemble: {count} instances, batched, cap {cap} steps", file=sys.stderr, flush=True)
    meshes, solvers, batched_s, tiles = ensemble.run_batched(insts, taylor.ROD_DURATION, cap)
    steps = [s.steps() for s in solvers]
    done = [s.time() >= taylor.ROD_DURATION for s in solvers]
    reached = sum(done)
    work = sum(m.elementCount * k for m, k in zip(meshes, steps))
    sample_checks = []
    print(f"# ensemble: one-by-one, {per_size} per mesh size", file=sys.stderr, flush=True)

    def one_by_one(indices):
        """Seconds per step per mesh size, measured on `per_size` instances of each."""
        by_size = {}
        for n in indices:
            by_size.setdefault(insts[n].elements, []).append(n)
        per_step = {}
        for e, ns in sorted(by_size.items()):
            picks = ns[:per_size]
            solo = ensemble.run_one_by_one([insts[n] for n in picks], taylor.ROD_DURATION, cap)
            per_step[e] = sum(sec for _, sec, _ in solo) / sum(k for k, _, _ in solo)
            for n, (k, _, res) in zip(picks, solo):
                sample_checks.append(k == steps[n] and res == ensemble.results(meshes[n], solvers[n]))
        estimate = sum(per_step[insts[n].elements] * steps[n] for n in indices)
        return per_step, estimate, {e: len(ns) for e, ns in by_size.items()}

    per_step, one_by_one_s, sizes = one_by_one(range(count))
    out = {
        "instances": count, "step_cap": cap, "reached_end_time": reached,
        "element_steps": work, "value": work / batched_s, "seconds": batched_s,
        "unit": "element-updates/s",
        "one_by_one_value": work / one_by_one_s, "one_by_one_seconds_estimated": one_by_one_s,
        "one_by_one_us_per_step": {str(e): v * 1e6 for e, v in per_step.items()},
        "speedup": one_by_one_s / batched_s,
        "tiles": {str(k): v for k, v in (tiles or {}).items()},
        "instances_per_size": {str(k): v for k, v in sizes.items()},
        "note": f"one-by-one measured on {per_size} instances per mesh size, scaled by the "
                "ensemble's exact step counts"}
    hit = [n for n in range(count) if done[n]]
    if hit:
        print(f"# ensemble: the {len(hit)} instances that reach 80 us, batched alone", file=sys.stderr, flush=True)
        sub = [insts[n] for n in hit]
        m2, s2, sec2, _ = ensemble.run_batched(sub, taylor.ROD_DURATION, 0)
        same = all(b.steps() == steps[n] and ensemble.results(mb, b) == ensemble.results(meshes[n], solvers[n])
                   for n, mb, b in zip(hit, m2, s2))
        for b in s2:
            b.shutdown()
        # The persistent form of the same batch: 32 steps per cooperative launch with
        # grid-wide barriers, against the graph-captured launches above.
        os.environ["NEOFIZ_BATCH_PERSISTENT"] = "1"
        try:
            m3, s3, sec3, _ = ensemble.run_batched(sub, taylor.ROD_DURATION, 0)
        finally:
            del os.environ["NEOFIZ_BATCH_PERSISTENT"]
        same3 = all(b.steps() == steps[n] and ensemble.results(mb, b) == ensemble.results(meshes[n], solvers[n])
                    for n, mb, b in zip(hit, m3, s3))
        for b in s3:
            b.shutdown()
        per_hit, est_hit, sizes_hit = one_by_one(hit)
        work_hit = sum(meshes[n].elementCount * steps[n] for n in hit)
        out["reached"] = {"instances": len(hit), "element_steps": work_hit,
                          "batched_seconds": sec2, "value": work_hit / sec2,
                          "one_by_one_seconds_estimated": est_hit,
                          "one_by_one_us_per_step": {str(e): v * 1e6 for e, v in per_hit.items()},
                          "instances_per_size": {str(k): v for k, v in sizes_hit.items()},
                          "speedup": est_hit / sec2, "same_results_as_full_batch": same,
                          "persistent_launch": {"batched_seconds": sec3, "same_results": same3,
                                                "graph_over_persistent": sec3 / sec2}}
    out["solo_equals_batched"] = all(sample_checks)
    for s_ in solvers:
        s_.shutdown()
    return out


def rod_run(nr, nz, johnson_cook, cap):
    """A whole rod from t = 0 to 80 us (or `cap` steps) through the single-instance path:
    steps, wall seconds, element-updates/s and the harness's L/L0 and mushroom diameter
    (validate/TaylorImpactCase.java:264-311)."""
    from neofiz_amd import taylor
    from neofiz_amd.solver import ExplicitSolver
    mat = taylor.ROD_JC if johnson_cook else taylor.ROD_LINEAR
    mu = taylor.ROD_JC_FRICTION if johnson_cook else 0.0
    mesh, solver, _ = taylor.build(mat, taylor.ROD_LENGTH, taylor.ROD_DIAMETER, taylor.ROD_SPEED,
                                   nr, nz, friction=mu, solver_cls=ExplicitSolver)
    a0 = solver.audits()
    t0 = time.perf_counter()
    steps = solver.runUntil(taylor.ROD_DURATION, cap)
    sec = time.perf_counter() - t0
    m = taylor.measure(mesh, solver, taylor.ROD_LENGTH, a0.axial_momentum, a0.kinetic_energy)
    st = solver.state()
    solver.shutdown()
    return {"mesh": f"{nr}x{nz}", "law": "johnson-cook, mu 0.09" if johnson_cook else "linear, mu 0",
            "elements": mesh.elementCount, "steps": steps, "time_us": st.time * 1e6,
            "reached_80us": st.time >= taylor.ROD_DURATION, "seconds": sec,
            "element_updates_per_s": mesh.elementCount * steps / sec,
            "L_over_L0": m["lengthRatio"], "mushroom_diameter_mm": m["mushroomDiameter"] * 1e3,
            "max_plastic_strain": m["maxPlasticStrain"],
            "momentum_balance_error": m["momentumBalanceError"],
            "energy_balance_error": m["energyBalanceError"]}


def sensitivity_leg(nr, nz):
    """Cost of the step loop's derivatives on the Johnson-Cook Taylor rod (nr x nz, to 80 us):
    the plain forward run; the discrete adjoint (the forward run recording a checkpoint every
    64 steps, then one reverse sweep carrying both objectives, recomputing each checkpoint's
    steps); the forward-mode tangent in six directions. The adjoint's cost ratio is
    (recording run + reverse sweep) / plain run; its twelve gradients are compared with the
    tangent's."""
    import numpy as np
    from neofiz_amd import sensitivity as S, taylor
    jc = taylor.ROD_JC.johnsonCook
    pt = S.Point(nr, nz, taylor.ROD_SPEED, taylor.ROD_JC_FRICTION, jc.a, jc.b, jc.n, jc.c, jc.m)
    print(f"# sensitivities: {nr}x{nz} rod to 80 us: plain, adjoint, tangent", file=sys.stderr, flush=True)
    obj, fwd_steps, fwd_s = S.forward(pt)
    obj_a, grad_a, steps_a, rec_s, rev_s, info = S.adjoint_gradients(pt, interval=64)
    obj_t, grad_t, steps_t, tan_s = S.gradients(pt)
    agree = float(np.max(np.abs(grad_a - grad_t) / np.abs(grad_t)))
    return {"workload": f"taylor-rod-jc-{nr}x{nz}, 80 us", "steps": fwd_steps,
            "forward_seconds": fwd_s,
            "adjoint": {"recording_forward_seconds": rec_s, "reverse_seconds": rev_s,
                        "cost_ratio": (rec_s + rev_s) / fwd_s, "checkpoint_interval": info[3],
                        "checkpoints": info[1], "capped_solve_element_passes": info[2]},
            "tangent": {"seconds": tan_s, "cost_ratio": tan_s / fwd_s},
            "cost_ratio": (rec_s + rev_s) / fwd_s,
            "adjoint_vs_tangent_max_rel": agree,
            "same_trajectory": fwd_steps == steps_a == steps_t and obj == obj_a == obj_t,
            "objectives": {"L/L0": obj[0], "mushroom_diameter": obj[1]},
            "gradient": {name: {"L/L0": grad_a[0, k], "mushroom_diameter": grad_a[1, k]}
                         for k, name in enumerate(S.PARAMETERS)},
            "mode": "reverse (discrete adjoint, checkpoints every 64 steps); tangent beside it"}


def perforation_leg(elements, warmup, steps):
    """The multi-body perforation scene at about `elements` elements: `steps` steps timed
    after `warmup` from t = 0 (the slug starts 1 mm from the wall and strikes it during the
    warmup), with the per-launch split of element pass, node pass and contact."""
    import ctypes as C
    from neofiz_amd import perforation
    cell = perforation.cell_for(elements)
    print(f"# perforation: building {elements} elements", file=sys.stderr, flush=True)
    mesh, s, _ = perforation.build(cell)
    s.run(warmup)
    t0 = time.perf_counter()
    s.run(steps)
    elapsed = time.perf_counter() - t0
    st = s.state()
    b = s._b
    ms = (C.c_double * 4)()
    counts = (C.c_int64 * 4)()
    b.profile(s._h, 1)
    b.profile_read(s._h, ms, counts)
    s.run(200)
    b.profile_read(s._h, ms, counts)
    b.profile(s._h, 0)
    s.shutdown()
    per = [ms[k] / max(1, counts[k]) for k in range(4)]
    return {"workload": f"perforation-{mesh.elementCount}", "elements": mesh.elementCount,
            "nodes": mesh.nodeCount, "cell_m": cell, "steps_from_t0": [warmup, warmup + steps],
            "value": mesh.elementCount * steps / elapsed, "unit": "element-updates/s",
            "ms_per_step": elapsed / steps * 1e3,
            "kernel_ms": {"node_pass": per[0], "element": per[1], "contact_launches": per[2]},
            "contact_pairs": st.contact_pairs, "eroded_elements": st.eroded_elements,
            "data": "synthetic scene (failure strains from the seeded defect field)"}


def defect_leg(reps, cpu):
    """Device-side defect-field seeding (DefectField.sampleOnto) for the perforation scene's
    ~200k-element assembly, plane-strain rule: device time of the three kernels per call,
    inputs already in device memory (the copies precede the first event), median of `reps`
    calls; then the weakest-link population of the reference's size-effect test (3000 seeds x
    128 elements on each of four tubes) as four multi-seed calls. With `cpu`, the oracle's
    serial sampleOnto on the same mesh beside it."""
    import numpy as np
    from neofiz_amd import perforation
    from neofiz_amd.defects import DefectField, Weibull
    from neofiz_amd.mesh import Outline
    from neofiz_amd.solver import QuadMesh
    cell = perforation.cell_for(200_000)
    wall, slug = perforation.outlines()
    mesh = Outline.assemble(cell, *wall, slug)
    E = mesh.elementCount
    base = Weibull(perforation.COPPER_STRAIN_SCALE, perforation.WEIBULL_MODULUS)
    print(f"# defect field: {E} elements, {reps} calls", file=sys.stderr, flush=True)
    field = DefectField(base, perforation.REFERENCE_VOLUME, perforation.CORRELATION_LENGTH, 5)
    first = field.sampleOnto(mesh, depth=perforation.DEPTH)
    calls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        field.sampleOnto(mesh, depth=perforation.DEPTH)
        calls.append((sum(field.timing()), field.timing(), time.perf_counter() - t0))
    calls.sort(key=lambda c: c[0])
    total_ms, parts, wall_s = calls[len(calls) // 2]
    # The smoothing's own arithmetic per element: 400 taps (multiply, add), 20 row weights
    # (multiply, add), 40 weights ((-d) d / 2s^2 and the squared sums): 1000 FP64 operations
    # beside 40 exponentials and the copula's 2 exponentials, a log and two powers.
    flops = 1000.0 * E
    size_seeds = 9000 + np.arange(3000)
    size_ms = 0.0
    size_elements = 0
    for ri in (0.02, 0.04, 0.08, 0.16):
        tube = QuadMesh.cylinderWall(ri, ri + 0.005, 0.02, 8, 16)
        f2 = DefectField(Weibull(0.42, 20.0), 1e-9, 1e-3, 9000)
        f2.sampleOnto(tube, seeds=size_seeds)
        size_ms += sum(f2.timing())
        size_elements += tube.elementCount * size_seeds.size
        f2.close()
    field.close()
    out = {"workload": f"defect-field-perforation-{E}", "elements": E, "cell_m": cell,
           "value": E / (total_ms * 1e-3), "unit": "elements/s",
           "kernel_ms": {"bounding_box": parts[0], "noise_grid": parts[1], "smooth_transform": parts[2]},
           "wall_ms_per_call_with_copies": wall_s * 1e3,
           "smoothing_fp64": {"flops_per_element": 1000, "achieved_tflops": flops / (parts[2] * 1e-3) / 1e12,
                              "peak_tflops": 78.6},
           "median_failure_strain": float(np.median(first)),
           "size_effect_population": {"samples": size_elements, "kernel_ms": size_ms,
                                      "value": size_elements / (size_ms * 1e-3), "unit": "element-samples/s"}}
    if cpu:
        from oracle.pyoracle import OracleDefectField
        ref = OracleDefectField(base, perforation.REFERENCE_VOLUME, perforation.CORRELATION_LENGTH, 5)
        t0 = time.perf_counter()
        want = ref.sampleOnto(mesh, depth=perforation.DEPTH)
        cpu_s = time.perf_counter() - t0
        ref.close()
        out["cpu_baseline"] = {"value": E / cpu_s, "unit": "elements/s", "cores": 1, "kind": "port",
                               "sample": f"the same {E}-element sampleOnto, oracle (serial C++ "
                                         "restatement with the reference's memoised noise grid)"}
        out["bitwise_equal_to_oracle"] = bool(np.array_equal(first.view(np.uint64), want.view(np.uint64)))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--nr", type=int, default=128)
    ap.add_argument("--nz", type=int, default=1280)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-copy", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-sensitivity", action="store_true")
    ap.add_argument("--no-perforation", action="store_true")
    ap.add_argument("--no-defect", action="store_true")
    ap.add_argument("--no-full-run", action="store_true")
    ap.add_argument("--no-ladder", action="store_true")
    ap.add_argument("--full-run-cap", type=int, default=400000,
                    help="step cap of the whole 80 us run of the benchmark rod")
    ap.add_argument("--defect-reps", type=int, default=20)
    ap.add_argument("--perforation-steps", type=int, default=2000)
    ap.add_argument("--perforation-warmup", type=int, default=1500)
    ap.add_argument("--ensemble", type=int, default=256, help="instances in the batched ensemble")
    ap.add_argument("--ensemble-cap", type=int, default=20000, help="step cap per instance")
    ap.add_argument("--ensemble-solo", type=int, default=8, help="one-by-one sample per mesh size")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; a multi-GPU figure is not "
                                   "produced (ensembles shard across GPUs with no "
                                   "communication)"}))
        sys.exit(2)

    import ctypes as C
    from neofiz_amd import taylor
    from neofiz_amd.solver import ExplicitSolver

    def rod():
        return taylor.build(taylor.ROD_JC, taylor.ROD_LENGTH, taylor.ROD_DIAMETER,
                            taylor.ROD_SPEED, args.nr, args.nz,
                            friction=taylor.ROD_JC_FRICTION, solver_cls=ExplicitSolver)

    print(f"# building the {args.nr}x{args.nz} rod", file=sys.stderr, flush=True)
    mesh, solver, _ = rod()
    E, N = mesh.elementCount, mesh.nodeCount
    solver.run(args.warmup)  # synchronous: returns with the device idle
    print(f"# warmup {args.warmup} steps done; timing {args.steps}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    solver.run(args.steps)   # synchronous
    elapsed = time.perf_counter() - t0
    st = solver.state()
    assert st.steps == args.warmup + args.steps, "step count mismatch"
    final_u = solver.axialDisplacement()
    solver.shutdown()

    # Kernel durations over the same steps: a twin of the rod, built and warmed identically
    # (the step loop is deterministic, so its trajectory is the timed one bit for bit), runs
    # the timed window again with start/stop events recorded by each dispatch
    # (hipExtLaunchKernelGGL), which the timed run itself does without.
    _, twin, _ = rod()
    twin.run(args.warmup)
    b = twin._b
    ms = (C.c_double * 4)()
    counts = (C.c_int64 * 4)()
    b.profile(twin._h, 1)
    b.profile_read(twin._h, ms, counts)
    twin.run(args.steps)
    b.profile_read(twin._h, ms, counts)
    b.profile(twin._h, 0)
    kernel_ms = [ms[k] / max(1, counts[k]) for k in range(4)]
    launches = [int(counts[k]) for k in range(4)]
    same_window = bool(np.array_equal(twin.axialDisplacement().view(np.uint64), final_u.view(np.uint64)))
    twin.shutdown()

    # The dominant kernel: the fused step kernel (node update + element pass, one launch per
    # step after a run call's first), or the element kernel where the step runs as two launches.
    fused = launches[3] > 0
    if fused:
        dom_name, dom_ms, dom_bytes, dom_file = "k_step", kernel_ms[3], step_kernel_bytes(E, N), STEP_TRAFFIC_FILE
    else:
        dom_name, dom_ms, dom_bytes, dom_file = "k_element", kernel_ms[1], element_kernel_bytes(E, N), TRAFFIC_FILE
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(dom_file):
        t = {}
        with open(dom_file) as f:
            for line in f:
                if ":" in line and not line.startswith("#"):
                    k, v = line.split(":", 1)
                    try:
                        t[k.strip()] = float(v)
                    except ValueError:
                        pass
        if t.get("nr") == args.nr and t.get("nz") == args.nz:
            traffic = int(t["bytes_per_launch"])

    value = E * args.steps / elapsed
    line = {
        "metric": "FP64 element-updates/s (Taylor rod, JC + Coulomb anvil)",
        "value": value,
        "unit": "element-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded-free: uniform initial velocity on a generated mesh)",
        "config": {"workload": f"taylor-rod-jc-{args.nr}x{args.nz}", "elements": E,
                   "nodes": N, "integration": "reduced+hourglass", "friction": 0.09,
                   "steps_from_t0": [args.warmup, args.warmup + args.steps],
                   "parallelism": "single-gpu"},
        "hbm_fraction_step": value * (step_bytes(E, N) / E) / (HBM_PEAK_GBPS * 1e9),
        "kernel_ms": {"step_kernel": kernel_ms[3] if fused else None, "element": kernel_ms[1],
                      "node_pass": kernel_ms[0],
                      "launches": {"step_kernel": launches[3], "element": launches[1], "node_pass": launches[0]},
                      "window": "the timed steps, replayed on an identical twin with events on",
                      "twin_trajectory_identical": same_window},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "bytes_per_launch": dom_bytes},
    }
    if not args.no_copy:
        try:
            line["stream_copy_GBps"] = stream_copy_gbps()
        except Exception as exc:  # no device: report, do not hide
            line["stream_copy_GBps"] = None
            line["stream_copy_error"] = str(exc)
    if not args.no_full_run:
        print(f"# full run: the {args.nr}x{args.nz} rod from t = 0 to 80 us", file=sys.stderr, flush=True)
        fr = rod_run(args.nr, args.nz, True, args.full_run_cap)
        fr["hbm_fraction_step"] = fr["element_updates_per_s"] * (step_bytes(E, N) / E) / (HBM_PEAK_GBPS * 1e9)
        line["full_run"] = fr
    if not args.no_ladder:
        print("# ladder: 8x80 linear, 16x160 .. 64x640 Johnson-Cook to 80 us", file=sys.stderr, flush=True)
        line["ladder"] = [rod_run(8, 80, False, 0)] + [rod_run(n, 10 * n, True, 0) for n in (16, 32, 64)]
    if not args.no_defect:
        line["defect_field"] = defect_leg(args.defect_reps, not args.no_cpu)
    if not args.no_perforation:
        line["perforation"] = perforation_leg(200000, args.perforation_warmup, args.perforation_steps)
    if not args.no_sensitivity:
        line["sensitivity"] = sensitivity_leg(32, 320)
    if not args.no_ensemble:
        line["batched"] = ensemble_leg(args.ensemble, args.ensemble_cap, args.ensemble_solo)
    if not args.no_cpu:
        model, host_cpus, allowed = host_cpu()
        # The pool is sized to the CPU share this process is given: OMP_NUM_THREADS where the
        # host sets it (the GPU host allots its GPU's share of cores that way), otherwise every
        # CPU this process may run on.
        share = os.environ.get("OMP_NUM_THREADS")
        threads = min(allowed, int(share)) if share and share.isdigit() and int(share) > 0 else allowed
        print(f"# CPU baseline: oracle on {threads} threads, 3 bursts over ~{args.cpu_seconds} s",
              file=sys.stderr, flush=True)
        rate, rates, last = cpu_baseline(args.nr, args.nz, args.cpu_seconds, threads)
        line["cpu_baseline"] = {
            "value": rate, "unit": "element-updates/s", "cores": threads, "kind": "port",
            "sample": f"the same {args.nr}x{args.nz} rod from t = 0: 20 warm-up steps, then the "
                      f"best of 3 bursts of ~{args.cpu_seconds / 3:.0f} s (to step {last}), oracle "
                      "(C++ restatement) threaded with the reference's cursor schedule",
            "bursts": rates, "cpu_model": model, "host_cpus": host_cpus,
            "cpus_allowed": allowed}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

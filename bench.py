#!/usr/bin/env python3
"""MPH-I time-step throughput on one MI355X.

Workload (default "3d_1m"): 1,000,000 fluid particles (100^3 lattice, seeded jitter
and Taylor-Green initial velocity, amplitude 0.1 m/s) resting in a closed 4-layer
wall box with a 140^3 interior (497,792 wall particles), mu = 1.0, dt = 5e-4, each
step solved to relative residual 1e-8 (rr/rr0 < 1e-16) with multigrid-preconditioned
conjugate residual in FP64.  A step is one pass of the whole time step (kinematics,
cell sort, neighbour search, explicit forces, assembly, multigrid build, solve).

Prints ONE JSON line (see the fields below).  ``value`` is fluid-particle-steps/s
with all state resident in HBM, with the particle matrix applied in the form that
--operator names (default: factored, the same matrix from one scalar per neighbour
entry; ``fp64_assembled`` reports the bit-exact assembled FP64 form beside it).
``roofline`` is measured live on the dominant kernel (the particle-matrix product
of the CR loop) with HIP events on the scene's stream.  ``cpu_baseline`` is the CPU
oracle (OpenMP) timed on a bounded sample of the same scene on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s, 1M 3D to 1e-8 res; mixed/batched vs plain FP64; p99 step"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = {"assembled": os.path.join(ROOT, "profiles", "spmv_traffic_3d_1m.json"),
                "factored": os.path.join(ROOT, "profiles", "fac_traffic_3d_1m.json")}
# the scene's counters of the FP64 CR loop's captured iterations
GRAPH_SCALARS = ("CrGraphLaunches", "CrCaptures", "CrInstantiations", "CrCaptureMs")
# a batch step's wall time per group of phases (mph_step_stats::ms_phase slots)
BATCH_GROUPS = {"cells_and_neighbours": 3, "explicit_to_multigrid_build": 8, "solve": 9, "advance": 11}


def spmv_bytes(blocks: float, nf: int, value_bytes: int = 8) -> tuple[float, float]:
    """Algorithmic bytes of one y = A s launch: (this format, CSR reference format).

    Sliced 3x3-block layout: 9 values (8 B in FP64, 4 B in FP32) + 4 B column per
    block, 4 B row length per particle, and the vectors once each (s read, y read
    for beta*y, y written: 3 x 8 B per row, n = 3*nf rows).  Reference CSR
    (main.cpp:4033-4050): 8 B value + 4 B column per nonzero (9 per block), 8 B row
    pointer per row, s read and y written once (16 B per row)."""
    n = 3 * nf
    mine = (9.0 * value_bytes + 4.0) * blocks + 4.0 * nf + 24.0 * n
    ref = 12.0 * 9.0 * blocks + 8.0 * (n + 1) + 16.0 * n
    return mine, ref


def roofline(kernel, ms, launches, bytes_sum, ref_sum, traffic):
    avg_s = ms / 1e3 / max(launches, 1)
    ok = launches > 0 and avg_s > 0
    achieved = bytes_sum / launches / avg_s / 1e9 if ok else None
    achieved_ref = ref_sum / launches / avg_s / 1e9 if ok else None
    return {
        "bound": "hbm",
        "kernel": kernel,
        "achieved": round(achieved, 1) if ok else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if ok else None,
        "traffic": traffic,
        "avg_launch_ms": round(1e3 * avg_s, 4),
        "launches": int(launches),
        "bytes_per_launch": round(bytes_sum / max(launches, 1)),
        "achieved_reference_csr_equivalent": round(achieved_ref, 1) if ok else None,
    }


def step_bytes(scene, factored, iters):
    """Algorithmic HBM bytes of one time step, each array counted once per logical
    pass, in the executed format and in the reference's CSR format (FP64 value, int32
    column, int64 row pointer), for `iters` CR iterations.  Terms:
      cells + neighbour search   36 Np (hash) + 16 Np (sort) + 104 Np (reorder)
                                 + 48 Np + 4 (E + E_P) (lists written)
      explicit forces            24 E + 100 Np
      assembly                   factored: lists, alpha (8 B per fluid pair), D, b,
                                 beta and C (8 + 24 B per P pair), C_ii, the records'
                                 positions, P; assembled: 3x3 blocks written (76 B) and
                                 read-modified-written by C Lambda C^T (144 B);
                                 reference: 12 nnzA + 12 nnzC + 4 E + 16 n + 8 Np
      multigrid build            factored: alpha and beta lists re-read, the G tables
                                 written and read, level 0 written; assembled and
                                 reference: A read once; coarser levels read and written
      per CR iteration           one product (factored: FactoredBytes; assembled: blocks
                                 + vectors; reference: 12 nnzA + 8 (n+1) + 16 n), one
                                 preconditioner (48 n + per sweep 8 B per stencil value +
                                 32 B per row: 4 sweeps on level 0, 6 on inner levels, 5
                                 on the coarsest; reference: 12 B per stencil value), and
                                 160 n of dot products and updates
      (I + 1) products and (I + 2) preconditioner applications per step.
    Returns (executed bytes, reference-format bytes, breakdown dict)."""
    Np, Nf = scene.scalar("ParticleCount"), scene.scalar("FluidCount")
    n = 3 * Nf
    E, EP = scene.scalar("NeighborIndCount"), scene.scalar("NeighborIndCountP")
    EA, EC = scene.scalar("ABlocks"), scene.scalar("CBlocks")
    depth = int(scene.scalar("MultiGridDepth"))
    levels = []
    for l in range(depth):
        rows = scene.field_bytes("MGInvD_L%d" % l) // 8
        vals = scene.field_bytes("MGCofA_L%d" % l) // 8
        levels.append((rows, vals))
    dim = 3 if levels and levels[0][0] and levels[0][1] // levels[0][0] > 27 else 2
    cells = 36 * Np + 16 * Np + 104 * Np + 48 * Np + 4 * (E + EP)
    explicit = 24 * E + 100 * Np
    c_common = 4 * EP + 24 * EP + 8 * Np + (4 + 24) * EC + 24 * Nf  # C, P; b -= C P
    if factored:
        assembly = (4 + 8) * EA + 72 * Nf + 24 * Nf + c_common + 8 * EP + 24 * Nf + 56 * Np
        gslots = 27 if dim == 3 else 9
        mg0 = 12 * EA + 12 * EP + 2 * 8 * 3 * gslots * Np + 24 * Nf
        product = scene.scalar("FactoredBytes")
    else:
        assembly = (4 + 76) * EA + 24 * Nf + c_common + 144 * EA
        mg0 = 76 * EA + 24 * Nf
        product = (9 * 8 + 4) * EA + 4 * Nf + 24 * n
    ref_assembly = 12 * 9 * EA + 12 * 3 * EC + 4 * E + 16 * n + 8 * Np
    ref_mg0 = 12 * 9 * EA + 24 * n
    mg_rest = sum(2 * 8 * vals for rows, vals in levels[1:])
    ref_mg_rest = sum(2 * 12 * vals for rows, vals in levels[1:])
    sweeps = [4] + [6] * max(0, depth - 2) + ([5] if depth >= 2 else [])
    precond = 48 * n + sum(s * (8 * vals + 32 * rows) for s, (rows, vals) in zip(sweeps, levels))
    ref_precond = 48 * n + sum(s * (12 * vals + 32 * rows) for s, (rows, vals) in zip(sweeps, levels))
    ref_product = 12 * 9 * EA + 8 * (n + 1) + 16 * n
    it = iters
    parts = {
        "cells_and_neighbours": cells,
        "explicit": explicit,
        "assembly": assembly,
        "multigrid_build": mg0 + mg_rest,
        "products": (it + 1) * product,
        "preconditioner": (it + 2) * precond,
        "vector_ops": it * 160 * n,
    }
    mine = float(sum(parts.values()))
    ref = float(cells + explicit + ref_assembly + ref_mg0 + ref_mg_rest + (it + 1) * ref_product
                + (it + 2) * ref_precond + it * 160 * n)
    return mine, ref, {k: round(v / 1e9, 3) for k, v in parts.items()}


PHASES = ["kinematics", "cells", "neighbor", "explicit", "matrix_a", "matrix_c", "multiply_c", "multigrid", "solve",
          "diagnostics", "advance"]


def form_for(args, precision):
    """Operator form of a leg (--operator; the same for every precision)."""
    return args.operator


def run_gpu(args, precision="fp64", form=None):
    """Create the scene, run warmup + timed steps; return timings, the SpMV roofline
    and the fields used for the precision comparison (Velocity and PressureP after
    the first step, whose output-step diagnostics run inside it, and after the last)."""
    from mphimplicit_amd import abi, scenes
    from mphimplicit_amd.scene import Scene

    params, grid = scenes.CONFIGS[args.config]()
    params.cr_tolerance = args.rel_residual ** 2
    params.end_time = 1.0e30
    params.precision = abi.PRECISION_MIXED if precision == "mixed" else abi.PRECISION_FP64
    form = form or form_for(args, precision)
    factored = form == "factored"
    params.operator_form = abi.OPERATOR_FACTORED if factored else abi.OPERATOR_ASSEMBLED
    nf = int((grid.type < 2).sum())
    t = time.perf_counter()
    scene = Scene(params, grid, device=0)
    if factored and args.factored_cells == "off":
        scene.set_option("FactoredCells", 0)  # the chunk kernel for the FP64 second pass
    if factored and args.factored_cells_xcd == "off":
        scene.set_option("FactoredCellsXcd", 0)
    scene.set_option("CrStop", args.cr_stop)
    setup_s = time.perf_counter() - t
    print("[bench] %s %s scene %s: %d particles (%d fluid), created in %.1f s"
          % (precision, form, args.config, grid.np, nf, setup_s), file=sys.stderr, flush=True)
    first = None
    for k in range(args.warmup):
        st = scene.step(1)
        if first is None:
            first = {"Velocity": scene.read("Velocity"), "PressureP": scene.read("PressureP")}
        print("[bench] warmup step %d: %d CR iterations" % (k, st.cr_iterations), file=sys.stderr, flush=True)
    lat, iters, refine = [], [], []
    g0 = {n: scene.scalar(n) for n in GRAPH_SCALARS}
    for k in range(args.steps):
        t0 = time.perf_counter()
        st = scene.step(1)
        lat.append(time.perf_counter() - t0)
        iters.append(st.cr_iterations)
        refine.append(int(scene.scalar("RefinementSteps")))
        if first is None:
            first = {"Velocity": scene.read("Velocity"), "PressureP": scene.read("PressureP")}
        print("[bench] step %d: %.1f ms, %d CR iterations" % (k, 1e3 * lat[-1], st.cr_iterations), file=sys.stderr,
              flush=True)
    graph = {n: scene.scalar(n) - g0[n] for n in GRAPH_SCALARS}
    graph["CrGraphError"] = scene.scalar("CrGraphError")
    # The products' launch durations: further steps with every product bracketed by HIP
    # events on the scene stream (the CR iterations are then launched one by one rather
    # than replayed from graphs; the kernels are the same)
    scene.set_timing(2)
    b64 = r64 = b32 = r32 = 0.0
    l64_prev = l32_prev = 0.0
    for k in range(args.spmv_timing_steps):
        st = scene.step(1)
        l64, l32 = scene.scalar("SpmvLaunches"), scene.scalar("SpmvLaunchesF32")
        blocks = scene.scalar("ABlocks")
        mine, ref = spmv_bytes(blocks, nf, 8)
        if factored:
            mine = scene.scalar("FactoredBytes")
        b64 += mine * (l64 - l64_prev)
        r64 += ref * (l64 - l64_prev)
        mine, ref = spmv_bytes(blocks, nf, 4)
        if factored and l32 > l32_prev:
            mine = scene.scalar("FactoredBytesF32")
        b32 += mine * (l32 - l32_prev)
        r32 += ref * (l32 - l32_prev)
        l64_prev, l32_prev = l64, l32
        print("[bench] SpMV-timed step %d: %d CR iterations" % (k, st.cr_iterations), file=sys.stderr, flush=True)
    print("[bench] SpMV timing: FP64 %.3f ms over %d launches, FP32 %.3f ms over %d launches (error code %d)"
          % (scene.scalar("SpmvMs"), l64_prev, scene.scalar("SpmvMsF32"), l32_prev,
             scene.scalar("SpmvTimingError")), file=sys.stderr, flush=True)
    traffic = None
    if os.path.exists(TRAFFIC_FILE[form]):
        with open(TRAFFIC_FILE[form]) as f:
            traffic = json.load(f).get("bytes_per_launch")
    if factored and scene.scalar("FactoredCellsUsed"):
        k64 = "k_fac_t + k_fac_y_cell (factored particle matrix, FP64 scalars, second pass staged per cell)"
        k32 = "k_fac_t + k_fac_y (factored particle matrix, FP32 scalars)"
    elif factored:
        k64 = "k_fac_t + k_fac_y (factored particle matrix, FP64 scalars)"
        k32 = "k_fac_t + k_fac_y (factored particle matrix, FP32 scalars)"
    else:
        k64 = "k_spmv (particle matrix, sliced 3x3 blocks, FP64 values)"
        k32 = "k_spmv (particle matrix, sliced 3x3 blocks, FP32 values)"
    rf64 = roofline(k64, scene.scalar("SpmvMs"), l64_prev, b64, r64, traffic if precision == "fp64" else None)
    if factored:
        rf64["note"] = ("the factored product moves ~2.8 GB of algorithmic HBM bytes where the assembled one moves "
                        "13.2 GB; its limit is the cache traffic of gathering neighbours' x, position and t "
                        "(three 16-byte pairs per entry) through L2 and the Infinity Cache, not HBM")
    rf32 = roofline(k32, scene.scalar("SpmvMsF32"), l32_prev, b32, r32, None)
    for rf in (rf64, rf32):
        rf["measured_over"] = ("%d steps after the timed ones, every product bracketed by HIP events on the scene "
                               "stream" % args.spmv_timing_steps)
    scene.set_timing(0)
    step = None
    if precision == "fp64":
        # the whole step against the roofline: algorithmic bytes at the mean iteration
        # count over the mean step time, and the device time per phase of one more
        # step (phase timing adds a synchronisation per phase: outside the timed steps)
        sb, sb_ref, parts = step_bytes(scene, factored, float(np.mean(iters)))
        ms = 1e3 * float(np.mean(lat))
        scene.set_timing(1)
        st = scene.step(1)
        step = {
            "bytes": round(sb),
            "achieved": round(sb / (ms / 1e3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(sb / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "ms_per_step": round(ms, 3),
            "bytes_by_part_gb": parts,
            "reference_csr_bytes": round(sb_ref),
            "achieved_reference_csr_equivalent": round(sb_ref / (ms / 1e3) / 1e9, 1),
            "phase_ms": {PHASES[k - 1]: round(st.ms_phase[k], 3) for k in range(1, 12) if st.ms_phase[k] > 0},
            "phase_ms_step_total": round(st.ms_total, 3),
        }
    else:
        # one more step, untimed, so that the fields read below are at the same
        # time step as the FP64 legs' (whose last step is the phase-timed one)
        scene.step(1)
    tuned = {"spmv_fp64_variant": int(scene.scalar("TunedSpmv64")), "spmv_fp32_variant": int(scene.scalar("TunedSpmv32")),
             "cac_variant": int(scene.scalar("TunedCac")), "fuse_variant": int(scene.scalar("TunedFuse")),
             "galerkin_variant": int(scene.scalar("TunedGalerkin"))}
    scene.run_phase(abi.PHASE_DIAGNOSTICS)
    last = {"Velocity": scene.read("Velocity"), "PressureP": scene.read("PressureP")}
    scene.close()
    return {
        "nf": nf,
        "np": grid.np,
        "lat": lat,
        "total": float(np.sum(lat)),
        "iters": iters,
        "refine": refine,
        "roofline": rf64,
        "roofline_f32": rf32,
        "setup_s": setup_s,
        "first": first,
        "last": last,
        "tuned": tuned,
        "graph": graph,
        "step": step,
    }


def field_error(a, b):
    """max |a - b| relative to max |b| for each field both carry."""
    out = {}
    for k in ("Velocity", "PressureP"):
        if a is None or b is None or k not in a or k not in b:
            continue
        scale = float(np.abs(b[k]).max())
        out[k] = float(np.abs(a[k] - b[k]).max() / scale) if scale > 0 else float(np.abs(a[k] - b[k]).max())
    return out


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or (
        len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count())


def calibrate_port(threads, steps=3):
    """Wall time of the port (the oracle, OpenMP) against the reference's own OpenMP
    executable (oracle/_ref/Mph_gcc3d, compiled from the reference's sources) on the
    same 3d_64k scene at the reference's hard-coded stop rule (rr/rr0 < 1e-12), over
    time steps 2..steps+1 (after the first step, which carries the output-step
    diagnostics).  The reference prints no per-step wall time, so its figure is the
    difference of two whole runs (1 step and steps+1 steps, identical reading,
    initialisation and file output).  None when the reference binary is absent."""
    import subprocess
    import tempfile
    import shutil

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from mphimplicit_amd import files, scenes

    exe = os.path.join(ROOT, "oracle", "_ref", "Mph_gcc3d")
    if not os.path.exists(exe):
        return None
    params, grid = scenes.config_3d_drop()
    params.cr_tolerance = 1.0e-12
    work = tempfile.mkdtemp(prefix="mph_calib_")
    try:
        files.write_grid_file(os.path.join(work, "box.grid"), grid)

        def ref_run(n):
            q = params.copy()
            q.end_time = (n - 1) * q.dt
            files.write_data_file(os.path.join(work, "box.data"), q)
            t = time.perf_counter()
            subprocess.run([exe, "box.data", "box.grid", "box%03d.prof", "box%03d.vtk", "box.log", str(threads)],
                           cwd=work, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            dt = time.perf_counter() - t
            its = [int(line.split("iter,=")[1].split(",")[0]) for line in open(os.path.join(work, "box.log"))
                   if "iter,=" in line]
            return dt, its

        t1, _ = ref_run(1)
        tn, ref_its = ref_run(steps + 1)
        q = params.copy()
        q.end_time = 1.0e30
        orc = pyoracle.OracleScene(q, grid, pyoracle.DOT_OMP)
        port_its = [orc.step(1).cr_iterations]
        t = time.perf_counter()
        for _ in range(steps):
            port_its.append(orc.step(1).cr_iterations)
        tp = time.perf_counter() - t
        orc.close()
    finally:
        shutil.rmtree(work, ignore_errors=True)
    tr = tn - t1
    print("[bench] cpu calibration on 3d_64k: reference %.2f s, port %.2f s for %d steps" % (tr, tp, steps),
          file=sys.stderr, flush=True)
    return {
        "port_vs_reference": round(tp / tr, 4),
        "reference_s": round(tr, 3),
        "port_s": round(tp, 3),
        "cr_iterations_reference": ref_its,
        "cr_iterations_port": port_its,
        "sample": "3d_64k (64,000 fluid + 118,016 wall particles), time steps 2-%d at rr/rr0 < 1e-12, %d threads; "
                  "reference = oracle/_ref/Mph_gcc3d (source/main.cpp built -O3 -fopenmp), its time the difference "
                  "of a %d-step and a 1-step run" % (steps + 1, threads, steps + 1),
    }


def run_cpu_baseline(args):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from mphimplicit_amd import scenes

    threads = cpu_threads()
    pyoracle.set_threads(threads)
    params, grid = scenes.CONFIGS[args.config]()
    params.cr_tolerance = args.rel_residual ** 2
    params.end_time = 1.0e30
    nf = int((grid.type < 2).sum())
    t = time.perf_counter()
    orc = pyoracle.OracleScene(params, grid, pyoracle.DOT_OMP)
    init_s = time.perf_counter() - t
    t = time.perf_counter()
    st = orc.step(1)  # warm-up: the first step, with the output-step diagnostics
    warm_s = time.perf_counter() - t
    warm_its = st.cr_iterations
    fields = {"Velocity": orc.read("Velocity"), "PressureP": orc.read("PressureP")}
    lat, its = [], []
    for _ in range(args.cpu_steps):
        t = time.perf_counter()
        st = orc.step(1)
        lat.append(time.perf_counter() - t)
        its.append(st.cr_iterations)
    orc.close()
    dt = float(np.sum(lat))
    print("[bench] cpu baseline: warm-up %.1f s, %d step(s) in %.1f s (CR iterations %s, %d threads, init %.1f s)"
          % (warm_s, args.cpu_steps, dt, its, threads, init_s), file=sys.stderr, flush=True)
    out = {
        "value": round(nf * args.cpu_steps / dt, 1),
        "unit": "fluid-particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": "time steps 2-%d of the same %s scene (step 1, with the output-step diagnostics, is the warm-up), "
                  "oracle/mph_oracle.cpp built -O3 -fopenmp (no -march), OpenMP dot products, %d threads, "
                  "stop rule rr/rr0 < %g; %.1f s wall"
                  % (args.cpu_steps + 1, args.config, threads, args.rel_residual ** 2, dt),
        "cr_iterations": its,
        "cr_iterations_warmup": warm_its,
        "p50_step_s": round(float(np.percentile(lat, 50)), 3),
        "p99_step_s": round(float(np.percentile(lat, 99)), 3),
    }
    if args.cpu_calibrate == "on":
        cal = calibrate_port(threads)
        out["calibration"] = cal
        if cal:
            out["port_vs_reference"] = cal["port_vs_reference"]
            # what the reference's own executable would take on the same sample
            out["reference_equivalent_value"] = round(out["value"] * cal["port_vs_reference"], 1)
    return out, fields


def run_batch(args):
    """Batched sweep: `count` seeded scenes (2D/3D mix, log-uniform sizes and
    viscosities, random gravity directions) advanced `steps` time steps, (a) one
    scene at a time in plain FP64 with mph_scene_step, (b) all together with
    mph_batch_step in mixed precision, (c) all together in plain FP64.  Scene
    creation (upload, initial sort) is outside the timed regions."""
    from mphimplicit_amd import abi, scenes
    from mphimplicit_amd.scene import Batch, Scene

    sweep = scenes.batch_sweep(args.batch_count, steps=args.batch_steps)
    nf = [int((g.type < 2).sum()) for _, g in sweep]
    npart = [int(g.np) for _, g in sweep]

    def form_of(precision):
        f = form_for(args, "mixed" if precision == abi.PRECISION_MIXED else "fp64")
        return abi.OPERATOR_FACTORED if f == "factored" else abi.OPERATOR_ASSEMBLED

    def mk(precision):
        out = []
        for p, g in sweep:
            q = p.copy()
            q.precision = precision
            q.cr_tolerance = args.rel_residual ** 2
            q.operator_form = form_of(precision)
            out.append(Scene(q, g))
        return out

    # (a) one scene at a time, plain FP64
    t_seq = 0.0
    for k, (p, g) in enumerate(sweep):
        q = p.copy()
        q.cr_tolerance = args.rel_residual ** 2
        q.operator_form = form_of(abi.PRECISION_FP64)
        s = Scene(q, g)
        t = time.perf_counter()
        for _ in range(args.batch_steps):
            s.step(1)
        t_seq += time.perf_counter() - t
        s.close()
    print("[bench] batch sweep: sequential FP64 %.2f s" % t_seq, file=sys.stderr, flush=True)
    res, splits = {}, {}
    for name, prec in (("batched_mixed", abi.PRECISION_MIXED), ("batched_fp64", abi.PRECISION_FP64)):
        members = mk(prec)
        with Batch() as b:
            for s in members:
                b.add(s)
            lat = []
            split = np.zeros(4)
            for _ in range(args.batch_steps):
                t = time.perf_counter()
                st = b.step()
                lat.append(time.perf_counter() - t)
                split += [st[0].ms_phase[k] for k in BATCH_GROUPS.values()]
        for s in members:
            s.close()
        res[name] = lat
        splits[name] = {g: round(float(v) / args.batch_steps, 2) for g, v in zip(BATCH_GROUPS, split)}
        print("[bench] batch sweep: %s %.2f s, mean ms per step %s" % (name, sum(lat), splits[name]),
              file=sys.stderr, flush=True)
    work = sum(nf) * args.batch_steps
    v_seq = work / t_seq
    out = {
        "workload": "%d seeded scenes (2D/3D, %d..%d fluid particles, %d fluid / %d total), %d steps each; "
                    "particle matrix %s in FP64, %s in mixed precision"
                    % (args.batch_count, min(nf), max(nf), sum(nf), sum(npart), args.batch_steps,
                       form_for(args, "fp64"), form_for(args, "mixed")),
        "unit": "fluid-particle-steps/s",
        "sequential_fp64": round(v_seq, 1),
    }
    for name, lat in res.items():
        v = work / sum(lat)
        out[name] = round(v, 1)
        out[name + "_speedup_vs_sequential_fp64"] = round(v / v_seq, 3)
        out[name + "_p50_step_ms"] = round(1e3 * float(np.percentile(lat, 50)), 2)
        out[name + "_p99_step_ms"] = round(1e3 * float(np.percentile(lat, 99)), 2)
        out[name + "_wall_ms_per_step"] = splits[name]
    return out


def run_serve(args, rate=None, seconds=None):
    """Serving stream: Poisson arrivals of scene-step requests against the wall
    clock for a bounded time (or until serve_requests requests), continuous batching
    over one resident batch.  Reports the offered load (fluid-particle-steps/s of the
    requests that arrived) beside the served throughput and latencies."""
    from mphimplicit_amd import abi, scenes
    from mphimplicit_amd.serve import Server, make_requests

    rate = args.serve_rate if rate is None else rate
    seconds = args.serve_seconds if seconds is None else seconds
    count = min(args.serve_requests, int(rate * seconds * 1.5) + 10)  # enough to outlast the run
    stream = scenes.serve_stream(count, rate=rate)
    mixed = args.serve_precision == "mixed"
    form = abi.OPERATOR_FACTORED if form_for(args, args.serve_precision) == "factored" else abi.OPERATOR_ASSEMBLED
    reqs = make_requests(stream, precision=abi.PRECISION_MIXED if mixed else abi.PRECISION_FP64,
                         tol=args.rel_residual ** 2, operator_form=form)
    srv = Server(lazy_mixed=mixed)
    try:
        r = srv.run(reqs, duration_s=seconds, log=lambda m: print(m, file=sys.stderr, flush=True))
    finally:
        srv.close()
    r["workload"] = ("Poisson arrivals at %g req/s, 1-20 steps, lognormal fluid counts (median 50k, sigma 1, "
                     "clipped 5k-500k), 2D/3D, %s, %s particle matrix, %g s run or %d requests"
                     % (rate, "mixed precision" if mixed else "FP64", form_for(args, args.serve_precision),
                        seconds, count))
    if r["fluid_particle_steps_per_s"] > 0:
        r["offered_over_served"] = r["offered_fluid_particle_steps_per_s"] / r["fluid_particle_steps_per_s"]
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}


def copy_probe(mib: int) -> None:
    """dst.copy_(src) over mib MiB: each launch reads and writes exactly mib MiB."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="3d_1m")
    ap.add_argument("--rel-residual", type=float, default=1.0e-8, help="CR stop: sqrt(rr/rr0) below this")
    ap.add_argument("--factored-cells", choices=["on", "off"], default="on",  # on: the scene's default rule
                    help="FP64 second pass of the factored product staged per cell in LDS (off: the chunk kernel)")
    ap.add_argument("--factored-cells-xcd", choices=["on", "off"], default="on",
                    help="cell-staged pass: each XCD takes a contiguous range of cells (off: round robin)")
    ap.add_argument("--cr-stop", type=int, choices=[0, 1, 2], default=1,
                    help="the FP64 CR stop test: 0 on the host per iteration, 1 on the device, 2 on the device with "
                         "the iterations replayed from captured HIP graphs")
    ap.add_argument("--spmv-timing-steps", type=int, default=2,
                    help="steps after the timed ones with every product bracketed by HIP events (its roofline)")
    ap.add_argument("--operator", choices=["assembled", "factored"], default="factored",
                    help="how the solve applies the particle matrix (mph_params.operator_form) in every leg")
    ap.add_argument("--assembled-leg", choices=["on", "off"], default="on",
                    help="when the FP64 leg is factored: also run plain FP64 with the assembled matrix (the "
                         "bit-exact form)")
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed CPU steps after one warm-up step")
    ap.add_argument("--cpu-calibrate", choices=["on", "off"], default="on",
                    help="time the CPU port against the reference's own OpenMP executable on 3d_64k")
    ap.add_argument("--mixed", choices=["on", "off"], default="on",
                    help="also run the mixed-precision form (FP32 matrices inside FP64 refinement)")
    ap.add_argument("--batch", choices=["on", "off"], default="on", help="batched-sweep leg")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--serve", choices=["on", "off"], default="on", help="serving leg")
    ap.add_argument("--serve-precision", choices=["fp64", "mixed"], default="fp64",
                    help="precision of the served scenes (FP64 steps the batch faster with the cell-staged product)")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-light-rate", type=float, default=5.0,
                    help="a second serving run at this arrival rate, below the GPU's capacity (0: skip)")
    ap.add_argument("--serve-light-seconds", type=float, default=20.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (the known-size kernel that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (scenes shard by scene, "
              "not by particle). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from mphimplicit_amd.scene import probe_bandwidth
    read_gbs, copy_gbs = probe_bandwidth(0, 8 << 30, 5)
    print("[bench] HBM ceilings: read %.0f GB/s, copy %.0f GB/s" % (read_gbs, copy_gbs), file=sys.stderr, flush=True)
    g = run_gpu(args)
    lat_ms = np.array(g["lat"]) * 1e3
    out = {
        "metric": METRIC,
        "value": round(g["nf"] * args.steps / g["total"], 1),
        "unit": "fluid-particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(lat_ms.mean()), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": args.config,
            "fluid_particles": g["nf"],
            "total_particles": g["np"],
            "rel_residual": args.rel_residual,
            "precision": "fp64",
            "operator": form_for(args, "fp64"),
        },
        "total_particle_steps_per_s": round(g["np"] * args.steps / g["total"], 1),
        "p50_step_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_step_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "cr_iterations": g["iters"],
        "autotuned": g["tuned"],
        "cr_stop": {"mode": args.cr_stop, "graph_launches": int(g["graph"]["CrGraphLaunches"]),
                    "captures": int(g["graph"]["CrCaptures"]),
                    "instantiations": int(g["graph"]["CrInstantiations"]),
                    "capture_ms_per_step": round(g["graph"]["CrCaptureMs"] / args.steps, 3),
                    "graph_error": int(g["graph"]["CrGraphError"])},
        "roofline": dict(g["roofline"], step=g["step"]),
        "hbm_ceiling": {"read_gbs": round(read_gbs, 1), "copy_gbs": round(copy_gbs, 1),
                        "spmv_frac_of_read": round(g["roofline"]["achieved"] / read_gbs, 4)
                        if g["roofline"]["achieved"] else None,
                        "probe": "mph_probe_bandwidth: best of 5 passes over 8 GiB; read with 16-byte nontemporal loads, copy with 16-byte loads and stores"},
        "cpu_baseline": None,
    }
    if form_for(args, "fp64") == "factored" and args.assembled_leg == "on":
        a = run_gpu(args, "fp64", "assembled")
        alat = np.array(a["lat"]) * 1e3
        av = a["nf"] * args.steps / a["total"]
        out["fp64_assembled"] = {
            "value": round(av, 1),
            "unit": "fluid-particle-steps/s",
            "ms_per_step": round(float(alat.mean()), 3),
            "p50_step_ms": round(float(np.percentile(alat, 50)), 3),
            "p99_step_ms": round(float(np.percentile(alat, 99)), 3),
            "cr_iterations": a["iters"],
            "arithmetic": "the reference's CSR product on the stored 3x3 blocks, bit-identical to the reference",
            "factored_speedup": round(out["value"] / av, 3),
            "roofline": a["roofline"],
            "rel_error_factored_vs_assembled_after_first_step": field_error(g["first"], a["first"]),
            "rel_error_factored_vs_assembled_after_last_step": field_error(g["last"], a["last"]),
        }
        if a["roofline"]["achieved"]:
            out["hbm_ceiling"]["assembled_spmv_frac_of_read"] = round(a["roofline"]["achieved"] / read_gbs, 4)
    m = None
    if args.mixed == "on":
        m = run_gpu(args, "mixed")
        mlat = np.array(m["lat"]) * 1e3
        mv = m["nf"] * args.steps / m["total"]
        out["mixed_precision"] = {
            "value": round(mv, 1),
            "unit": "fluid-particle-steps/s",
            "ms_per_step": round(float(mlat.mean()), 3),
            "p50_step_ms": round(float(np.percentile(mlat, 50)), 3),
            "p99_step_ms": round(float(np.percentile(mlat, 99)), 3),
            "cr_iterations_inner": m["iters"],
            "refinement_steps": m["refine"],
            "speedup_vs_plain_fp64": round(mv / out["value"], 3),
            "arithmetic": "FP32 particle-matrix values (%s form) and multigrid stencils (coarsest level FP64), FP64 "
                          "vectors, dots and outer residual; same stop rule ||b - A x||/||b|| < %g with the FP64 "
                          "matrix" % (form_for(args, "mixed"), args.rel_residual),
            "roofline_fp32_spmv": m["roofline_f32"],
            "roofline_fp64_residual_spmv": m["roofline"],
            "rel_error_vs_fp64_gpu_after_first_step": field_error(m["first"], g["first"]),
            "rel_error_vs_fp64_gpu_after_last_step": field_error(m["last"], g["last"]),
        }
    if args.batch == "on":
        out["batched"] = run_batch(args)
    if args.serve == "on":
        out["serving"] = run_serve(args)
        if args.serve_light_rate > 0:
            out["serving_below_capacity"] = run_serve(args, args.serve_light_rate, args.serve_light_seconds)
    if args.cpu_baseline == "on":
        cb, cfields = run_cpu_baseline(args)
        cb["rel_error_fp64_gpu_vs_cpu_after_first_step"] = field_error(g["first"], cfields)
        if m is not None:
            cb["rel_error_mixed_gpu_vs_cpu_after_first_step"] = field_error(m["first"], cfields)
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

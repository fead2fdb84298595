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
    setup_s = time.perf_counter() - t
    print("[bench] %s %s scene %s: %d particles (%d fluid), created in %.1f s"
          % (precision, form, args.config, grid.np, nf, setup_s), file=sys.stderr, flush=True)
    first = None
    for k in range(args.warmup):
        st = scene.step(1)
        if first is None:
            first = {"Velocity": scene.read("Velocity"), "PressureP": scene.read("PressureP")}
        print("[bench] warmup step %d: %d CR iterations" % (k, st.cr_iterations), file=sys.stderr, flush=True)
    scene.set_timing(2)  # bracket every SpMV launch with HIP events
    lat, iters, refine = [], [], []
    b64 = r64 = b32 = r32 = 0.0
    l64_prev = l32_prev = 0.0
    for k in range(args.steps):
        t0 = time.perf_counter()
        st = scene.step(1)
        lat.append(time.perf_counter() - t0)
        iters.append(st.cr_iterations)
        refine.append(int(scene.scalar("RefinementSteps")))
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
        if first is None:
            first = {"Velocity": scene.read("Velocity"), "PressureP": scene.read("PressureP")}
        print("[bench] step %d: %.1f ms, %d CR iterations" % (k, 1e3 * lat[-1], st.cr_iterations), file=sys.stderr,
              flush=True)
    print("[bench] SpMV timing: FP64 %.3f ms over %d launches, FP32 %.3f ms over %d launches (error code %d)"
          % (scene.scalar("SpmvMs"), l64_prev, scene.scalar("SpmvMsF32"), l32_prev,
             scene.scalar("SpmvTimingError")), file=sys.stderr, flush=True)
    traffic = None
    if os.path.exists(TRAFFIC_FILE[form]):
        with open(TRAFFIC_FILE[form]) as f:
            traffic = json.load(f).get("bytes_per_launch")
    if factored:
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


def run_cpu_baseline(args):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from mphimplicit_amd import scenes

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count())
    pyoracle.set_threads(threads)
    params, grid = scenes.CONFIGS[args.config]()
    params.cr_tolerance = args.rel_residual ** 2
    params.end_time = 1.0e30
    nf = int((grid.type < 2).sum())
    t = time.perf_counter()
    orc = pyoracle.OracleScene(params, grid, pyoracle.DOT_OMP)
    init_s = time.perf_counter() - t
    t = time.perf_counter()
    st = orc.step(args.cpu_steps)
    dt = time.perf_counter() - t
    fields = {"Velocity": orc.read("Velocity"), "PressureP": orc.read("PressureP")}
    orc.close()
    print("[bench] cpu baseline: %d step(s) in %.1f s (%d CR iterations, %d threads, init %.1f s)"
          % (args.cpu_steps, dt, st.cr_iterations, threads, init_s), file=sys.stderr, flush=True)
    return {
        "value": round(nf * args.cpu_steps / dt, 1),
        "unit": "fluid-particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": "first %d time step(s) of the same %s scene after its initial sort and neighbour search, "
                  "oracle/mph_oracle.cpp built -O3 -fopenmp (no -march), OpenMP dot products, %d threads; "
                  "%d CR iterations, %.1f s wall" % (args.cpu_steps, args.config, threads, st.cr_iterations, dt),
    }, fields


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
    res = {}
    for name, prec in (("batched_mixed", abi.PRECISION_MIXED), ("batched_fp64", abi.PRECISION_FP64)):
        members = mk(prec)
        with Batch() as b:
            for s in members:
                b.add(s)
            lat = []
            for _ in range(args.batch_steps):
                t = time.perf_counter()
                b.step()
                lat.append(time.perf_counter() - t)
        for s in members:
            s.close()
        res[name] = lat
        print("[bench] batch sweep: %s %.2f s" % (name, sum(lat)), file=sys.stderr, flush=True)
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
    return out


def run_serve(args, rate=None):
    """Serving stream: Poisson arrivals of scene-step requests against the wall
    clock for a bounded time, continuous batching over one resident batch."""
    from mphimplicit_amd import abi, scenes
    from mphimplicit_amd.serve import Server, make_requests

    rate = args.serve_rate if rate is None else rate
    count = min(args.serve_requests, int(rate * args.serve_seconds * 1.5) + 10)  # enough to outlast the run
    stream = scenes.serve_stream(count, rate=rate)
    form = abi.OPERATOR_FACTORED if form_for(args, "mixed") == "factored" else abi.OPERATOR_ASSEMBLED
    reqs = make_requests(stream, tol=args.rel_residual ** 2, operator_form=form)
    srv = Server(lazy_mixed=True)
    try:
        r = srv.run(reqs, duration_s=args.serve_seconds,
                    log=lambda m: print(m, file=sys.stderr, flush=True))
    finally:
        srv.close()
    r["workload"] = ("Poisson arrivals at %g req/s, 1-20 steps, lognormal fluid counts (median 50k, sigma 1, "
                     "clipped 5k-500k), 2D/3D, mixed precision, %s particle matrix, %g s run"
                     % (rate, form_for(args, "mixed"), args.serve_seconds))
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
    ap.add_argument("--operator", choices=["assembled", "factored"], default="factored",
                    help="how the solve applies the particle matrix (mph_params.operator_form) in every leg")
    ap.add_argument("--assembled-leg", choices=["on", "off"], default="on",
                    help="when the FP64 leg is factored: also run plain FP64 with the assembled matrix (the "
                         "bit-exact form)")
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--mixed", choices=["on", "off"], default="on",
                    help="also run the mixed-precision form (FP32 matrices inside FP64 refinement)")
    ap.add_argument("--batch", choices=["on", "off"], default="on", help="batched-sweep leg")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--serve", choices=["on", "off"], default="on", help="serving leg")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-light-rate", type=float, default=5.0,
                    help="a second serving run at this arrival rate, below the GPU's capacity (0: skip)")
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
        "roofline": g["roofline"],
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
            out["serving_below_capacity"] = run_serve(args, args.serve_light_rate)
    if args.cpu_baseline == "on":
        cb, cfields = run_cpu_baseline(args)
        if args.cpu_steps == 1:
            cb["rel_error_fp64_gpu_vs_cpu_after_first_step"] = field_error(g["first"], cfields)
            if m is not None:
                cb["rel_error_mixed_gpu_vs_cpu_after_first_step"] = field_error(m["first"], cfields)
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

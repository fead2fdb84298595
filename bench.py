#!/usr/bin/env python3
"""MPH-I time-step throughput on one MI355X.

Workload (default "3d_1m"): 1,000,000 fluid particles (100^3 lattice, seeded jitter
and Taylor-Green initial velocity, amplitude 0.1 m/s) resting in a closed 4-layer
wall box with a 140^3 interior (497,792 wall particles), mu = 1.0, dt = 5e-4, each
step solved to relative residual 1e-8 (rr/rr0 < 1e-16) with multigrid-preconditioned
conjugate residual in FP64.  A step is one pass of the whole time step (kinematics,
cell sort, neighbour search, explicit forces, assembly, multigrid build, solve).

Prints ONE JSON line (see the fields below).  ``value`` is fluid-particle-steps/s
with all state resident in HBM.  ``roofline`` is measured live on the dominant
kernel (the particle-matrix SpMV of the CR loop) with HIP events on the scene's
stream.  ``cpu_baseline`` is the CPU oracle (OpenMP) timed on a bounded sample of
the same scene on this host.
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "spmv_traffic_3d_1m.json")


def spmv_bytes(blocks: float, nf: int) -> tuple[float, float]:
    """Algorithmic bytes of one y = A s launch: (this format, CSR reference format).

    Sliced 3x3-block layout: 72 B of values + 4 B column per block, 4 B row length
    per particle, and the vectors once each (s read, y read for beta*y, y written:
    3 x 8 B per row, n = 3*nf rows).  Reference CSR (main.cpp:4033-4050): 8 B value +
    4 B column per nonzero (9 per block), 8 B row pointer per row, s read and y
    written once (16 B per row)."""
    n = 3 * nf
    mine = 76.0 * blocks + 4.0 * nf + 24.0 * n
    ref = 12.0 * 9.0 * blocks + 8.0 * (n + 1) + 16.0 * n
    return mine, ref


def run_gpu(args):
    from mphimplicit_amd import scenes
    from mphimplicit_amd.scene import Scene

    params, grid = scenes.CONFIGS[args.config]()
    params.cr_tolerance = args.rel_residual ** 2
    params.end_time = 1.0e30
    nf = int((grid.type < 2).sum())
    t = time.perf_counter()
    scene = Scene(params, grid, device=0)
    setup_s = time.perf_counter() - t
    print("[bench] scene %s: %d particles (%d fluid), created in %.1f s" % (args.config, grid.np, nf, setup_s),
          file=sys.stderr, flush=True)
    for k in range(args.warmup):
        st = scene.step(1)
        print("[bench] warmup step %d: %d CR iterations" % (k, st.cr_iterations), file=sys.stderr, flush=True)
    scene.set_timing(2)  # bracket every SpMV launch with HIP events
    lat, iters, spmv_bytes_sum, spmv_ref_sum, launches_prev = [], [], 0.0, 0.0, 0.0
    for k in range(args.steps):
        t0 = time.perf_counter()
        st = scene.step(1)
        lat.append(time.perf_counter() - t0)
        iters.append(st.cr_iterations)
        launches = scene.scalar("SpmvLaunches")
        blocks = scene.scalar("ABlocks")
        mine, ref = spmv_bytes(blocks, nf)
        spmv_bytes_sum += mine * (launches - launches_prev)
        spmv_ref_sum += ref * (launches - launches_prev)
        launches_prev = launches
        print("[bench] step %d: %.1f ms, %d CR iterations" % (k, 1e3 * lat[-1], st.cr_iterations), file=sys.stderr,
              flush=True)
    spmv_ms = scene.scalar("SpmvMs")
    launches = scene.scalar("SpmvLaunches")
    print("[bench] SpMV timing: %.3f ms over %d launches (error code %d)"
          % (spmv_ms, launches, scene.scalar("SpmvTimingError")), file=sys.stderr, flush=True)
    total = float(np.sum(lat))
    avg_launch_s = spmv_ms / 1e3 / max(launches, 1)
    ok = launches > 0 and avg_launch_s > 0
    achieved = spmv_bytes_sum / launches / avg_launch_s / 1e9 if ok else None
    achieved_ref = spmv_ref_sum / launches / avg_launch_s / 1e9 if ok else None
    scene.close()
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_launch")
    return {
        "nf": nf,
        "np": grid.np,
        "lat": lat,
        "total": total,
        "iters": iters,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_spmv (particle matrix, sliced 3x3 blocks, FP64)",
            "achieved": round(achieved, 1) if ok else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if ok else None,
            "traffic": traffic,
            "avg_launch_ms": round(1e3 * avg_launch_s, 4),
            "launches": int(launches),
            "bytes_per_launch": round(spmv_bytes_sum / max(launches, 1)),
            "achieved_reference_csr_equivalent": round(achieved_ref, 1) if ok else None,
        },
        "setup_s": setup_s,
    }


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
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="3d_1m")
    ap.add_argument("--rel-residual", type=float, default=1.0e-8, help="CR stop: sqrt(rr/rr0) below this")
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--cpu-steps", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (scenes shard by scene, "
              "not by particle). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
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
        },
        "total_particle_steps_per_s": round(g["np"] * args.steps / g["total"], 1),
        "p50_step_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_step_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "cr_iterations": g["iters"],
        "roofline": g["roofline"],
        "cpu_baseline": None,
    }
    if args.cpu_baseline == "on":
        out["cpu_baseline"] = run_cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

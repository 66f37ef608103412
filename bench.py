#!/usr/bin/env python3
"""Throughput of the chicken FP64 explicit step (libchkn, HIP/gfx950) on one MI355X.

One "step" is one reference time step: three TVD-RK3 stages over every active cell, with
the CFL time-step reduction every cfl_count steps.  Default workload: the 256x128x128
Mach 5 compression-ramp block (FP64, 1% seeded density perturbation), inputs resident in
HBM before timing.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload scale|plain|batched]
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gdtk-rmhd_amd"))

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_CELL_STEP = 864  # algorithmic FP64 bytes per cell-update (3 stages)
BYTES_PER_CELL_DT = 144    # extra bytes per cell per CFL reduction


def build_workload(name, steps_total):
    from chkn import cases
    if name == "scale":
        job = cases.ramp(256, 128, 128, max_step=steps_total)
        desc = {"workload": "plain-at-scale: one 256x128x128 block, Mach 5 flow over a 15-degree "
                            "ramp, p=2e3 Pa, T=220 K, 1% seeded density perturbation, FP64",
                "cells": 256 * 128 * 128, "blocks": 1}
    elif name == "plain":
        job = cases.sod_box(64, 32, 32, max_step=steps_total)
        desc = {"workload": "plain: one 64x32x32 block, 3D Sod shock tube, slip walls, FP64",
                "cells": 64 * 32 * 32, "blocks": 1}
    elif name == "batched":
        job = cases.ramp_array(seed=7, max_step=steps_total)
        desc = {"workload": "batched: 5x4x2 array of tensor-product uneven blocks (seed 7), "
                            "Mach 5 ramp, seeded exchange/wall connection graph, FP64",
                "cells": job.active_cells, "blocks": len(job.blocks)}
    else:
        raise SystemExit("unknown workload " + name)
    desc.update({"scheme": "MUSCL/van Albada + AUSMDV, TVD-RK3 (3 stages)", "cfl": 0.5,
                 "cfl_count": job.cfl_count})
    return job, desc


def cpu_baseline(job, steps):
    """Time the reference's own sources (oracle/_ref) on the first `steps` steps of the same
    job; fall back to the C oracle when the reference build is absent."""
    ref = os.path.join(ROOT, "oracle", "_ref", "chkn_ref")
    cells = job.active_cells
    nblocks = sum(1 for b in job.blocks if b.active)
    if os.path.exists(ref):
        from chkn import prep
        d = tempfile.mkdtemp(prefix="chkn_cpu_")
        try:
            prep.write_initial_files(True, d)
            threads = min(nblocks, os.cpu_count() or 1)
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            out = subprocess.run([ref, "run", d, "0", str(steps), "-"], env=env, check=True,
                                 capture_output=True, text=True).stdout
            secs = float(out.split("march_seconds=")[1].split()[0])
        finally:
            shutil.rmtree(d, ignore_errors=True)
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bind
        threads = 1
        o = oracle_bind.Oracle(job)
        t0 = time.perf_counter()
        o.march(steps)
        secs = time.perf_counter() - t0
        kind = "port"
    return {"value": cells * steps / secs, "unit": "cell-updates/s", "cores": threads, "kind": kind,
            "sample": "first %d steps of the same job (%d active cells in %d block(s)); the "
                      "reference parallelises over blocks only, so %d thread(s)"
                      % (steps, cells, nblocks, threads)}


def copy_calibration(torch):
    """Three copies of a 2 GiB FP64 buffer: each dispatch reads and writes exactly 2 GiB, so
    the HBM byte counters of these dispatches give the counters' scale factor."""
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="scale", choices=["scale", "plain", "batched"])
    ap.add_argument("--tile", default="", help="TIxTJxTK launch tile, e.g. 32x8x8")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--per-block", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 2 GiB device copy (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one block array "
              "per device); run with --gpus 1", file=sys.stderr)
        raise SystemExit(2)

    import torch
    if args.copy_calibration:
        copy_calibration(torch)
        return
    from chkn.lib import Sim

    job, desc = build_workload(args.workload, args.warmup + args.steps)
    sim = Sim(job)
    opts = {"use_graph": bool(args.graph), "per_block_launch": bool(args.per_block)}
    if args.tile:
        ti, tj, tk = (int(v) for v in args.tile.split("x"))
        opts.update(tile_i=ti, tile_j=tj, tile_k=tk)
    sim.set_launch(**opts)
    if args.warmup:
        sim.step(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep = sim.step(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms, launches = sim.last_timing()
    cells = sim.active_cells
    first = args.warmup
    reductions = sum(1 for s in range(first + 1, first + args.steps + 1)
                     if s % job.cfl_count == 0)
    alg_bytes = cells * (BYTES_PER_CELL_STEP * args.steps + BYTES_PER_CELL_DT * reductions)
    value = cells * args.steps / wall
    achieved = alg_bytes / (kernel_ms * 1e-3)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "stage_kernel_traffic_%s.txt" % args.workload)
    if os.path.exists(tfile):
        with open(tfile) as f:
            for ln in f:
                if ln.startswith("hbm_bytes_per_launch:"):
                    traffic = float(ln.split(":")[1])
    line = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": dict(desc, launch=opts, steps_completed=rep.step),
        "hbm_pct_of_peak": 100.0 * achieved / HBM_PEAK,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "stage_kernel (3 launches per step)",
                     "avg_launch_us": kernel_ms * 1e3 / max(launches, 1),
                     "alg_bytes_per_launch": alg_bytes / max(launches, 1)},
    }
    sim.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(job, args.cpu_steps)
    print(json.dumps(line))


if __name__ == "__main__":
    main()

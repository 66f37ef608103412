#!/usr/bin/env python3
"""Throughput of the chicken FP64 explicit step (libchkn, HIP/gfx950) on one MI355X.

One "step" is one reference time step: three TVD-RK3 stages over every active cell, with
the CFL time-step reduction every cfl_count steps.  Default workload: the 256x128x128
Mach 5 compression-ramp block (FP64, 1% seeded density perturbation), inputs resident in
HBM before timing.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload scale|plain|batched|ensemble|...]

The launch shape is autotuned first (chkn.tune) unless --tile is given.
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
# Two-stage predictor-corrector: stage 1 reads U0 and writes U1 and L0 (18 words), stage 2
# reads U1, U0 and L0 and writes U (24 words), plus 10 geometry words per stage.
BYTES_PER_CELL_STEP_PC = 496
# Viscous step: also the cell centre and the three face centres a cell's clouds need, per stage.
BYTES_PER_CELL_STEP_VISCOUS = BYTES_PER_CELL_STEP + 3 * 96


WORKLOAD_STEPS = {"scale": 500, "plain": 200, "batched": 300, "ensemble": 300, "diff": 500, "diff-pc": 500,
                  "ensemble-diff": 300, "viscous": 300, "shard": 500, "pc": 500, "diff-scale": 500,
                  "xl": 100}
DIFF_WORKLOADS = ("diff", "diff-pc", "ensemble-diff", "diff-scale")
# Launch shape of workloads too large to sweep in a bench run: the tuner's pick for the
# 256x128x128 block of the same flow.
FIXED_LAUNCH = {"xl": dict(tile_i=32, tile_j=8, tile_k=0)}
# Steps of the reference CPU build timed for cpu_baseline: a sample of about 10-30 s on one
# core (the viscous step runs about 6x slower there than the inviscid one).
CPU_STEPS = {"viscous": 1}


def build_workload(name, steps_total):
    """(list of JobData, description) of a named workload."""
    from chkn import cases
    if name == "scale":
        jobs = [cases.ramp(256, 128, 128, max_step=steps_total)]
        desc = {"workload": "plain-at-scale: one 256x128x128 block, Mach 5 flow over a 15-degree "
                            "ramp, p=2e3 Pa, T=220 K, 1% seeded density perturbation, FP64"}
    elif name == "pc":
        jobs = [cases.ramp(256, 128, 128, max_step=steps_total)]
        jobs[0].time_scheme = 1
        desc = {"workload": "predictor-corrector at scale: the 256x128x128 Mach 5 ramp block of "
                            "plain-at-scale stepped by the two-stage predictor-corrector "
                            "(time_scheme predictor_corrector), FP64"}
    elif name == "viscous":
        jobs = [cases.ramp(256, 128, 128, max_step=steps_total, viscous=True)]
        desc = {"workload": "viscous: one 256x128x128 block, Mach 5 flow over a 15-degree ramp "
                            "with viscous fluxes (least-squares face gradients, Sutherland air), "
                            "no-slip ramp wall at 300 K, 1% seeded density perturbation, FP64"}
    elif name == "shard":
        jobs = [cases.ramp_slab(64, 32, max_step=steps_total)]
        desc = {"workload": "shard: the 32x128x128 i-slab (cells 64-95, the ramp corner) of the "
                            "256x128x128 Mach 5 ramp block, one GPU's share of an 8-way split "
                            "along i, run standalone (inflow and outflow in place of the "
                            "neighbouring slabs), FP64"}
    elif name == "plain":
        jobs = [cases.sod_box(64, 32, 32, max_step=steps_total)]
        desc = {"workload": "plain: one 64x32x32 block, 3D Sod shock tube, slip walls, FP64"}
    elif name == "batched":
        jobs = [cases.ramp_array(seed=7, max_step=steps_total)]
        desc = {"workload": "batched: 5x4x2 array of 40 uneven blocks (slab sizes seeded in "
                            "[16,96], seed 7), Mach 5 ramp, seeded exchange/wall connection graph, FP64"}
    elif name == "ensemble":
        jobs = cases.ramp_ensemble(seed=11, n=64, max_step=steps_total)
        desc = {"workload": "ensemble: 64 independent single-block Mach 3-8 ramp jobs (seed 11), "
                            "grids 32-96 x 32-64 x 32-64, per-instance dt, FP64"}
    elif name == "diff":
        jobs = [cases.ramp(128, 64, 64, max_step=1 << 30)]
        desc = {"workload": "differentiation: one 128x64x64 block, Mach 5 ramp; forward + adjoint "
                            "of the wall-pressure loss w.r.t. U0 and inflow (Mach, p, T), "
                            "checkpoint every 50 steps (segments taped by the forward pass while "
                            "they fit in the device's free memory, recomputed otherwise), dt "
                            "sequence frozen, FP64"}
    elif name == "diff-scale":
        jobs = [cases.ramp(256, 128, 128, max_step=1 << 30)]
        desc = {"workload": "differentiation at scale: the 256x128x128 block of plain-at-scale; "
                            "forward + adjoint of the wall-pressure loss w.r.t. U0 and inflow "
                            "(Mach, p, T) over 500 steps, checkpoint every 50 steps (30 GB per "
                            "segment tape: the segments that fit beside one recompute tape are "
                            "taped by the forward pass, the rest recomputed), dt sequence frozen, "
                            "FP64"}
    elif name == "xl":
        jobs = [cases.ramp(768, 384, 384, max_step=steps_total)]
        jobs[0].dt_init = 2.5e-7  # cells 3x smaller than the 256x128x128 block's: CFL-stable start
        desc = {"workload": "xl: one 768x384x384 block (113M cells) of the Mach 5 ramp flow of "
                            "plain-at-scale, dt_init 2.5e-7 s, FP64"}
    elif name == "diff-pc":
        jobs = [cases.ramp(128, 64, 64, max_step=1 << 30)]
        jobs[0].time_scheme = 1
        desc = {"workload": "differentiation, predictor-corrector: one 128x64x64 block, Mach 5 ramp "
                            "stepped by the two-stage predictor-corrector; forward + adjoint of the "
                            "wall-pressure loss w.r.t. U0 and inflow (Mach, p, T), checkpoint every "
                            "50 steps (segments taped while they fit in the device's free memory, "
                            "recomputed otherwise), dt sequence frozen, FP64"}
    elif name == "ensemble-diff":
        jobs = cases.ramp_ensemble(seed=11, n=64, max_step=1 << 30)
        desc = {"workload": "batched differentiation: 64 independent Mach 3-8 ramp jobs (seed 11), "
                            "forward + adjoint of each wall-pressure loss w.r.t. its inflow "
                            "(Mach, p, T), checkpoint every 50 steps (segments taped while they fit "
                            "beside one recompute tape in the device's free memory, recomputed "
                            "otherwise), FP64"}
    else:
        raise SystemExit("unknown workload " + name)
    stepper = ("predictor-corrector (2 stages)" if jobs[0].time_scheme == 1 else "TVD-RK3 (3 stages)")
    desc.update({"cells": sum(j.active_cells for j in jobs),
                 "blocks": sum(len(j.blocks) for j in jobs), "instances": len(jobs),
                 "scheme": "MUSCL/van Albada + AUSMDV, " + stepper
                           + (", viscous fluxes" if jobs[0].viscous else ""), "cfl": 0.5,
                 "cfl_count": jobs[0].cfl_count})
    return jobs, desc


def new_sim(jobs):
    from chkn.lib import Sim
    return Sim(jobs=jobs) if len(jobs) > 1 else Sim(jobs[0])


def heat_sim(jobs, opts):
    """A handle of the same jobs that steps without a step limit, for heat_device."""
    import copy
    unbounded = []
    for j in jobs:
        c = copy.copy(j)
        c.max_step = 1 << 30
        c.max_time = 1.0e30
        unbounded.append(c)
    h = new_sim(unbounded)
    h.set_launch(**opts)
    return h


def heat_device(h, seconds=0.25):
    """Untimed steps on the handle h (not the measured one) for `seconds`, right before a timed
    run: after even a short idle gap (building or freeing a handle) the device takes some 50 ms
    of work to return to its full clocks -- the first 25 steps of a 256x128x128 block then run
    up to 24% slower -- which a few warmup steps of the measured handle do not cover.  The
    measured handle still starts from the initial state and runs exactly its warmup and timed
    steps."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        h.step(10)


def run_timed(jobs, opts, steps, warmup):
    """A fresh handle at the initial state, the device heated on a second handle
    (heat_device), warmup steps, then `steps` timed steps.  Returns (wall_s, kernel_ms,
    launches, report)."""
    import torch
    sim = new_sim(jobs)
    sim.set_launch(**opts)
    h = heat_sim(jobs, opts)
    heat_device(h)
    if warmup:
        sim.step(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep = sim.step(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms, launches = sim.last_timing()
    cells = sim.active_cells
    sim.close()
    h.close()
    return wall, kernel_ms, launches, rep, cells


def dedicated_tape_budget(sim, jobs, steps, checkpoint_every):
    """Resident-tape budget of a device that runs this unroll alone: every checkpoint segment
    if they fit in the free memory less 8 GB, else as many as fit beside one segment's
    recompute tape (the library's default, half of the free memory, leaves room for a
    caller's other work)."""
    import torch
    nst = 2 if jobs[0].time_scheme == 1 else 3
    seg = 8 * 6 * sim.active_cells * (nst * checkpoint_every + 1)
    nseg = -(-steps // checkpoint_every)
    free, _ = torch.cuda.mem_get_info()
    avail = free - 8e9
    if nseg * seg <= avail:
        return nseg * seg
    return max(0, int((avail - seg) // seg)) * seg


def run_diff(jobs, opts, steps, warmup, checkpoint_every=50):
    """Forward (checkpointed) + adjoint of the wall-pressure loss against the plain forward
    of the same steps, with the tapes of a device that runs the unroll alone (the forward
    pass tapes the checkpoint segments that fit, the backward pass recomputes the rest) and
    with every segment recomputed.  Returns a dict of timings."""
    import torch
    from chkn import adjoint
    from chkn.lib import Sim
    sim = Sim(jobs=jobs) if len(jobs) > 1 else Sim(jobs[0])
    sim.set_launch(**opts)
    h = heat_sim(jobs, opts)  # built before the tapes are sized, which counts its memory
    budget = dedicated_tape_budget(sim, jobs, steps, checkpoint_every)
    sim.set_tape_budget(budget)
    loss_fn = adjoint.WallPressure(sim, jobs)
    u0 = adjoint.initial_state(sim)
    mpts = []
    for j in jobs:
        p, T, _, vx = (float(v) for v in j.flow_states[0][:4])
        mpts.append([vx / (1.4 * 287.1 * T) ** 0.5, p, T])
    state_inst = [q for q, j in enumerate(jobs) for _ in j.flow_states]

    def gradient(n):
        mp = [torch.tensor(m, dtype=torch.float64, requires_grad=True) for m in mpts]
        inflow = torch.stack([adjoint.inflow_from_mpt(mp[q]) for q in state_inst])
        u = u0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        uN = adjoint.unroll(sim, u, inflow, n, checkpoint_every)
        L = loss_fn(uN)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, float(L.detach()), [m.grad for m in mp], u.grad

    gradient(steps)  # allocates tapes and checkpoints
    heat_device(h)
    t_fwd, t_adj, L, gm, gu = gradient(steps)
    resident = sim.grad_resident_segments
    sim.set_tape_budget(0)
    gradient(steps)
    rc_fwd, rc_adj, rc_L, rc_gm, rc_gu = gradient(steps)
    same = rc_L == L and all(bool(torch.equal(a, b)) for a, b in zip(gm, rc_gm)) and bool(torch.equal(gu, rc_gu))
    sim.set_tape_budget(budget)
    # Plain forward of the same steps from the same state (stepping ring, graph replay).
    sim.write_conserved(u0)
    sim.step(1)
    sim.write_conserved(u0)
    heat_device(h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    cells = sim.active_cells
    sim.close()
    h.close()
    nseg = -(-steps // checkpoint_every)
    return {"seconds_forward_taped": t_fwd, "seconds_adjoint": t_adj, "seconds_plain_forward": t_plain,
            "fwd_adj_over_fwd": (t_fwd + t_adj) / t_plain,
            "resident_segments": "%d of %d" % (resident, nseg), "tape_budget_bytes": budget,
            "recompute_all": {"seconds_forward": rc_fwd, "seconds_adjoint": rc_adj,
                              "fwd_adj_over_fwd": (rc_fwd + rc_adj) / t_plain,
                              "gradient_identical": same},
            "loss": L,
            "dL_dMpT_first": [float(v) for v in gm[0]], "grad_u0_norm": float(gu.norm()),
            "cells": cells}


def cpu_baseline(job, steps):
    """Time the reference's own sources (oracle/_ref) on the first `steps` steps of `job`,
    written out through chkn.prep (which holds the job built last); fall back to the C
    oracle when the reference build is absent.  OpenMP threads: one per block, at most 16
    (the host share of one GPU)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "chkn_ref")
    cells = job.active_cells
    nblocks = sum(1 for b in job.blocks if b.active)
    if os.path.exists(ref):
        from chkn import prep
        held = sum(b.nic * b.njc * b.nkc for b in prep.fluidBlocksList if b.active)
        if held != cells:
            raise RuntimeError("chkn.prep holds another job (%d active cells, not %d)" % (held, cells))
        d = tempfile.mkdtemp(prefix="chkn_cpu_")
        try:
            prep.write_initial_files(True, d)
            threads = min(nblocks, 16, os.cpu_count() or 1)
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
    the HBM byte counters of these dispatches give the counters' scale factor.  Then the
    device's streaming rates for reference: a timed 2 GiB copy (read + write) and a timed
    2 GiB sum (read only)."""
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    copy_ms = e0.elapsed_time(e1) / 10
    for _ in range(3):
        s = a.sum()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(10):
        s = a.sum()
    e1.record()
    torch.cuda.synchronize()
    sum_ms = e0.elapsed_time(e1) / 10
    del s
    print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8,
                      "copy_TBps_read_plus_write": 2 * 8 * n / copy_ms / 1e9,
                      "sum_TBps_read": 8 * n / sum_ms / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default: the workload's)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="scale", choices=sorted(WORKLOAD_STEPS))
    ap.add_argument("--tile", default="", help="TIxTJxTK launch tile, e.g. 32x8x8 (skips tuning)")
    ap.add_argument("--order", type=int, default=0, help="tile order with --tile")
    ap.add_argument("--persistent", type=int, default=0, help="persistent work-groups with --tile")
    ap.add_argument("--vertex-frames", type=int, default=0,
                    help="with --tile: the stage kernel forms face frames from the vertices")
    ap.add_argument("--tune", type=int, default=1, help="autotune the launch shape first")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--per-block", type=int, default=0)
    ap.add_argument("--compare-per-block", type=int, default=-1,
                    help="also time one launch per block per stage (default: batched only)")
    ap.add_argument("--cpu-steps", type=int, default=0, help="reference CPU steps timed (default 4; viscous 1)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sub-records", type=int, default=1,
                    help="scale workload: also report the batched and diff forms (default 1)")
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
    from chkn import tune

    steps = args.steps or WORKLOAD_STEPS[args.workload]
    t_start = time.perf_counter()

    def progress(msg):
        print("[bench %6.1f s] %s" % (time.perf_counter() - t_start, msg), file=sys.stderr, flush=True)

    jobs, desc = build_workload(args.workload, args.warmup + steps)
    progress("workload %s built: %d cells" % (args.workload, sum(j.active_cells for j in jobs)))
    tuned = None
    if args.tile:
        ti, tj, tk = (int(v) for v in args.tile.split("x"))
        opts = dict(tile_i=ti, tile_j=tj, tile_k=tk, tile_order=args.order,
                    persistent=bool(args.persistent), vertex_frames=bool(args.vertex_frames))
    elif args.workload in FIXED_LAUNCH:
        opts = dict(FIXED_LAUNCH[args.workload])
    elif args.tune and not jobs[0].viscous:  # the face-list stage has no tile shape
        t0 = time.perf_counter()
        opts, table = tune.autotune(jobs=jobs, per_block_launch=bool(args.per_block))
        tuned = {"candidates": len(table), "seconds": round(time.perf_counter() - t0, 1),
                 "best_ms_per_step": table[0][0], "worst_ms_per_step": table[-1][0],
                 "chosen_sweep_ms_per_step": next(ms for ms, o in table if o == opts)}
    else:
        opts = {}
    opts = dict(opts, use_graph=bool(args.graph), per_block_launch=bool(args.per_block))
    progress("launch %s%s" % ({k: opts.get(k) for k in ("tile_i", "tile_j", "tile_k")},
                              " (tuned)" if tuned else ""))
    if args.workload in DIFF_WORKLOADS:
        r = run_diff(jobs, opts, steps, args.warmup)
        wall = r["seconds_forward_taped"] + r["seconds_adjoint"]
        line = {"metric": "cell-updates/s (forward + adjoint)",
                "value": r["cells"] * steps / wall, "unit": "cell-updates/s", "n_gpus": 1,
                "steps": steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / steps,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                "data": "synthetic",
                "config": dict(desc, launch=opts, autotune=tuned, checkpoint_every=50,
                               plain_forward_value=r["cells"] * steps / r["seconds_plain_forward"],
                               timings=r)}
        print(json.dumps(line))
        return
    wall, kernel_ms, launches, rep, cells = run_timed(jobs, opts, steps, args.warmup)
    progress("timed %d steps: %.3f ms per step" % (steps, wall * 1e3 / steps))
    rec = roofline_record(jobs, opts, steps, args.warmup, args.workload, kernel_ms, launches, cells)
    config = dict(desc, launch=opts, steps_completed=rep.step)
    if tuned:
        config["autotune"] = tuned
    compare = args.compare_per_block if args.compare_per_block >= 0 else int(args.workload == "batched")
    if compare and not args.per_block:
        pb_opts, _ = tune.autotune(jobs=jobs, per_block_launch=True)
        pb_opts = dict(pb_opts, use_graph=bool(args.graph), per_block_launch=True)
        pw, _, pl, _, _ = run_timed(jobs, pb_opts, steps, args.warmup)
        config["per_block_comparator"] = {
            "value": cells * steps / pw, "launches": pl, "launch": pb_opts,
            "fused_speedup": pw / wall}
    line = {
        "metric": "cell-updates/s",
        "value": cells * steps / wall,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": config,
        "hbm_pct_of_peak": 100.0 * rec["frac"],
        "roofline": rec,
    }
    if not args.no_cpu:
        # chkn.prep holds the job built last (before the sub-records build theirs): for an
        # ensemble, its last member (the reference runs one job per process).
        line["cpu_baseline"] = cpu_baseline(jobs[-1], args.cpu_steps or CPU_STEPS.get(args.workload, 4))
    if args.workload == "scale" and args.sub_records:
        # The rest of the metric: the batched form against the reference's per-block launch
        # structure, and forward + adjoint against the plain forward.
        line["batched"] = batched_record(args.graph, cpu=not args.no_cpu)
        line["diff"] = diff_record(args.graph)
    print(json.dumps(line))


def roofline_record(jobs, opts, steps, warmup, workload, kernel_ms, launches, cells):
    """The roofline object of a timed run: algorithmic bytes of the timed steps over the stage
    kernels' device time (HIP events on the library's stream), and the measured HBM bytes per
    launch of this launch shape from the workload's profile summary (null if not profiled)."""
    reductions = sum(1 for s in range(warmup + 1, warmup + steps + 1) if s % jobs[0].cfl_count == 0)
    viscous = bool(jobs[0].viscous)
    pc = jobs[0].time_scheme == 1
    per_step = (BYTES_PER_CELL_STEP_VISCOUS if viscous else
                BYTES_PER_CELL_STEP_PC if pc else BYTES_PER_CELL_STEP)
    alg_bytes = cells * (per_step * steps + BYTES_PER_CELL_DT * reductions)
    achieved = alg_bytes / (kernel_ms * 1e-3)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "stage_kernel_traffic_%s.txt" % workload)
    if os.path.exists(tfile) and opts:
        shape = "%dx%dx%d%s" % (opts.get("tile_i", 0), opts.get("tile_j", 0), opts.get("tile_k", 0),
                                "-vf" if opts.get("vertex_frames") else "")
        with open(tfile) as f:
            for ln in f:
                if ln.startswith("hbm_bytes_per_launch %s:" % shape):
                    traffic = float(ln.split(":")[1])
    nst = 2 if pc else 3
    kernel = ("split_face_kernel + split_viscous_kernel + split_update_kernel (%d launches per step)"
              % (3 * nst) if viscous else "stage_kernel (%d launches per step)" % nst)
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": kernel,
            "avg_launch_us": kernel_ms * 1e3 / max(launches, 1),
            "alg_bytes_per_launch": alg_bytes / max(launches, 1)}


# Launch shapes of the sub-records, from the tuner's choices on these workloads (fused: the
# stage kernel's default k-chunking; per block: one plane per work-group, without which a
# small block's launch cannot fill the device).
SUB_FUSED = [dict(tile_i=32, tile_j=8, tile_k=0), dict(tile_i=16, tile_j=16, tile_k=0),
             dict(tile_i=32, tile_j=8, tile_k=4)]
SUB_PER_BLOCK = [dict(tile_i=32, tile_j=8, tile_k=1), dict(tile_i=16, tile_j=16, tile_k=1),
                 dict(tile_i=64, tile_j=4, tile_k=1)]


def _best_of(jobs, candidates, steps, warmup, graph, per_block):
    """Time each candidate launch shape on a fresh handle; the fastest run's numbers."""
    best = None
    for c in candidates:
        opts = dict(c, use_graph=bool(graph), per_block_launch=per_block)
        wall, kernel_ms, launches, rep, cells = run_timed(jobs, opts, steps, warmup)
        if best is None or wall < best[0]:
            best = (wall, kernel_ms, launches, rep, cells, opts)
    return best


def batched_record(graph, steps=WORKLOAD_STEPS["batched"], warmup=10, cpu=True):
    """The 40-block uneven array (5x4x2, slab sizes seeded in [16, 96]) in one launch per
    stage for all blocks, against one launch per block per stage (the reference's launch
    structure, simulate.cu:605-625), each with its best launch shape of a short list; and the
    reference's CPU path on the same array (OpenMP over its 40 blocks, at most 16 threads)."""
    jobs, desc = build_workload("batched", warmup + steps)
    wall, kms, launches, rep, cells, opts = _best_of(jobs, SUB_FUSED, steps, warmup, graph, False)
    rec = roofline_record(jobs, opts, steps, warmup, "batched", kms, launches, cells)
    pw, _, pl, _, _, popts = _best_of(jobs, SUB_PER_BLOCK, steps, warmup, graph, True)
    out = {"workload": desc["workload"], "cells": cells, "steps": steps, "warmup": warmup,
           "value": cells * steps / wall, "unit": "cell-updates/s", "ms_per_step": wall * 1e3 / steps,
           "launch": opts, "roofline": rec,
           "per_block_value": cells * steps / pw, "per_block_launches": pl, "per_block_launch": popts,
           "fused_over_per_block": pw / wall}
    if cpu:
        out["cpu_baseline"] = cpu_baseline(jobs[0], 2)
    return out


def diff_record(graph, steps=WORKLOAD_STEPS["diff"]):
    """Forward (checkpoint every 50 steps) + adjoint of the wall-pressure loss of the
    128x64x64 ramp w.r.t. U0 and the inflow (Mach, p, T), against the plain forward of the
    same steps: the differentiation workload's 500-step unroll (shorter unrolls reverse a
    larger share of steps that the adjoint has not reached yet, which are cheaper)."""
    jobs, desc = build_workload("diff", steps)
    opts = dict(tile_i=32, tile_j=8, tile_k=0, use_graph=bool(graph), per_block_launch=False)
    r = run_diff(jobs, opts, steps, 5)
    wall = r["seconds_forward_taped"] + r["seconds_adjoint"]
    return {"workload": desc["workload"], "cells": r["cells"], "steps": steps, "checkpoint_every": 50,
            "value": r["cells"] * steps / wall, "unit": "cell-updates/s (forward + adjoint)",
            "plain_forward_value": r["cells"] * steps / r["seconds_plain_forward"],
            "fwd_adj_over_fwd": r["fwd_adj_over_fwd"], "resident_segments": r["resident_segments"],
            "recompute_all_fwd_adj_over_fwd": r["recompute_all"]["fwd_adj_over_fwd"], "launch": opts,
            "seconds": {k: r[k] for k in ("seconds_forward_taped", "seconds_adjoint",
                                          "seconds_plain_forward")}}


if __name__ == "__main__":
    main()

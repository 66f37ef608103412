"""Launch-shape autotuner for libchkn.

The stage kernel's speed depends on the tile shape (TI x TJ cells per work-group), the
number of k-planes a work-group marches (tile_k), the order of the tile table and whether
work-groups are persistent.  None of these changes the results (the parity tests check
that), so the tuner times each candidate on a scratch handle built from the same job(s),
on the real block sizes, and returns the fastest.  Results are cached per problem shape.

    from chkn import tune
    best, table = tune.autotune(job)            # or tune.autotune(jobs=[...])
    sim.set_launch(**best)
"""
import copy
import hashlib
import json
import os
import time

from .lib import BadCellError, Sim

TILES = [(32, 8), (64, 4), (16, 16), (64, 8), (16, 4), (8, 8)]
TILE_K = [0, 1, 2, 4, 8, 16]
ORDERS = [0, 1]
PERSISTENT = [0, 1]
# Shapes with the vertex-frame form of the stage kernel (face frames formed from the vertices,
# x_order 2): tried with the default k-chunking and plane-major order.
VERTEX_FRAME_TILES = [(32, 8), (16, 16)]


def candidates(tiles=TILES, tile_k=TILE_K, orders=ORDERS, persistent=PERSISTENT, x_order=2):
    """Every launch option combination to try (graph replay is always on)."""
    out = []
    for ti, tj in tiles:
        for tk in tile_k:
            for o in orders:
                for p in persistent:
                    out.append(dict(tile_i=ti, tile_j=tj, tile_k=tk, tile_order=o,
                                    persistent=bool(p)))
    if x_order > 1:
        for ti, tj in tiles:
            if (ti, tj) in VERTEX_FRAME_TILES:
                out.append(dict(tile_i=ti, tile_j=tj, tile_k=0, tile_order=0, persistent=False,
                                vertex_frames=True))
    return out


def shape_key(jobs):
    """Cache key: block sizes, activity, x_order and the GPU's name."""
    h = hashlib.sha1()
    for j in jobs:
        h.update(repr((j.x_order, [(b.nic, b.njc, b.nkc, b.active) for b in j.blocks])).encode())
    try:
        import torch
        h.update(torch.cuda.get_device_name(0).encode())
    except Exception:
        pass
    return h.hexdigest()


def _scratch(jobs):
    out = []
    for j in jobs:
        c = copy.copy(j)
        c.max_step = 1 << 30
        c.max_time = 1.0e30
        out.append(c)
    return out


def autotune(job=None, jobs=None, steps=10, warmup=2, cands=None, cache_path=None,
             per_block_launch=False, verbose=False, refine=8, refine_steps=50, heat_seconds=0.1):
    """Time every candidate for `steps` steps (after `warmup` untimed steps) on a scratch
    handle, then time the `refine` fastest twice more for `refine_steps` steps each (each run
    on a fresh handle from the initial state), and
    return (best_options, [(ms_per_step, options), ...] of the sweep, fastest first).  The
    short sweep cannot separate shapes a few percent apart; the longer runs pick among them.
    The first steps of a fresh handle run slower than later ones (the flow's first steps cost
    more, and on some devices the clocks are still low after the idle time of building the
    handle: up to 2.4% on a 256x128x128 block, more than the shapes being ranked differ), so
    the sweep's first candidate and every refine run are timed only after `heat_seconds` of
    untimed steps.

    Timing is the device time of the step kernels (HIP events around the launches).  The
    scratch handle's state advances during the sweep; a candidate that meets a bad cell
    rebuilds the handle from the initial state and is timed again."""
    jobs = list(jobs) if jobs is not None else [job]
    cands = cands if cands is not None else candidates(x_order=jobs[0].x_order)
    key = shape_key(jobs) + ("-pb" if per_block_launch else "")
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            return cache[key]["best"], [tuple(r) for r in cache[key]["table"]]
    scratch = _scratch(jobs)

    def fresh():
        return Sim(jobs=scratch) if len(scratch) > 1 else Sim(scratch[0])

    sim = [fresh()]

    def time_steps(opts, n, warm=warmup):
        for attempt in range(2):
            try:
                sim[0].set_launch(**opts)
                if warm:
                    sim[0].step(warm)
                sim[0].step(n)
                break
            except BadCellError:
                sim[0].close()
                sim[0] = fresh()
                if attempt == 1:
                    raise
        ms, _ = sim[0].last_timing()
        return ms / n

    def heat(opts):
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < heat_seconds:
            time_steps(opts, 10, 0)

    table = []
    if cands:
        heat(dict(cands[0], use_graph=True, per_block_launch=per_block_launch))
    for c in cands:
        opts = dict(c, use_graph=True, per_block_launch=per_block_launch)
        table.append((time_steps(opts, steps), opts))
        if verbose:
            print("%8.3f ms/step  %s" % (table[-1][0], opts), flush=True)
    table.sort(key=lambda r: r[0])
    best = table[0][1]
    sim[0].close()
    top = [o for _, o in table[:max(1, refine)]]
    if len(top) > 1:
        # Each from a fresh handle: the sweep's handle has stepped thousands of times, and the
        # ranking of shapes a few percent apart can differ on that developed flow.
        best_ms = [float("inf")] * len(top)
        for _ in range(2):
            for q, opts in enumerate(top):
                sim[0] = fresh()
                heat(opts)
                best_ms[q] = min(best_ms[q], time_steps(opts, refine_steps))
                sim[0].close()
        best = top[min(range(len(top)), key=lambda q: best_ms[q])]
        if verbose:
            for q, opts in enumerate(top):
                print("refined %8.3f ms/step  %s" % (best_ms[q], opts), flush=True)
    if cache_path:
        cache[key] = {"best": best, "table": table}
        d = os.path.dirname(os.path.abspath(cache_path))
        os.makedirs(d, exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return best, table

"""Autotuner for the launch parameters of the particle step and the tracer.

Each candidate is timed with HIP events on the library stream over a few
repetitions on the caller's real workload; the fastest is applied and
returned.  Results can be cached in a JSON file keyed by a workload
signature so that later runs skip the sweep.
"""
import json
import os
import time

import numpy as np

from . import _lib as L

PUSH_TILES = (1024, 2048, 4096, 8192)
# register budgets of the push: 1 = the compiler's allocation, 4 = held to
# 128 VGPRs for four wavefronts per SIMD
PUSH_WAVES = (1, 4)
# nodes per wavefront of the deposit node gather
NODE_BATCHES = (4, 8, 16)
# (cell-pass kernel, chunks): 0 = per-lane particle loop, 1 = shared-memory
# record batches, 2 = FP64 matrix cores, 3 = eight lanes per cell (chunks =
# particles per cell per record batch)
DEPOSIT_VARIANTS = ((3, 8), (3, 4), (2, 4), (1, 4), (1, 6), (1, 12), (0, 6))


def set_deposit_variant(ctx, kernel, nch):
    """Select a deposit cell-pass variant (a chunk count of 4 suits every kernel)."""
    ctx.set_param("deposit_chunks", 4)
    ctx.set_param("deposit_kernel", kernel)
    ctx.set_param("deposit_chunks", nch)


def _load(path):
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}
    return {}


def _save(path, d):
    if not path:
        return
    try:
        with open(path, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
    except OSError:
        pass


def weigh_steps(ctx, state, opts_kw, deposit_dt, variants, reps=2):
    """The deposit variant whose whole steps (push, migration, both deposit
    passes, by HIP events over reps + 1 eager steps) are fastest, and each
    variant's time per step; leaves the winner set."""
    ctx.set_timing(True)
    times = {}
    for var in variants:
        set_deposit_variant(ctx, *var)
        ctx.reset_timing()
        state.run_steps(reps + 1, deposit_dt, use_graph=False, **opts_kw)
        ms, n = ctx.timing()
        times[var] = sum(ms[k] for k in (L.K_PUSH, L.K_COMPACT, L.K_DEPOSIT_CELL, L.K_DEPOSIT_NODE)) / \
            max(n[L.K_PUSH], 1)
    best = min(variants, key=lambda v: times[v])
    set_deposit_variant(ctx, *best)
    return best, times


def tune_pic(ctx, state, opts_kw, deposit_dt, signature, cache=None, reps=2):
    """Pick the push tile and register budget (occupancy), the deposit
    cell-pass variant and the node gather's batch for a PicState."""
    db = _load(cache)
    if signature in db and "node_batch" in db[signature]:
        best = db[signature]
        ctx.set_param("tile", best["tile"])
        ctx.set_param("push_waves", best["push_waves"])
        set_deposit_variant(ctx, best["deposit_kernel"], best["deposit_chunks"])
        ctx.set_param("node_batch", best["node_batch"])
        return best, True
    ctx.set_timing(True)
    res = {}

    def push_ms():
        ctx.reset_timing()
        for _ in range(reps):
            state.MoveParticles(compact=True, **opts_kw)
        ms, n = ctx.timing()
        return ms[L.K_PUSH] / max(n[L.K_PUSH], 1)

    ctx.set_param("push_waves", PUSH_WAVES[0])
    for tile in PUSH_TILES:
        ctx.set_param("tile", tile)
        res[("tile", tile)] = push_ms()
    best_tile = min(PUSH_TILES, key=lambda t: res[("tile", t)])
    ctx.set_param("tile", best_tile)
    res[("waves", PUSH_WAVES[0])] = res[("tile", best_tile)]
    for w in PUSH_WAVES[1:]:
        ctx.set_param("push_waves", w)
        res[("waves", w)] = push_ms()
    best_waves = min(PUSH_WAVES, key=lambda w: res[("waves", w)])
    ctx.set_param("push_waves", best_waves)
    for var in DEPOSIT_VARIANTS:
        set_deposit_variant(ctx, *var)
        ctx.reset_timing()
        for _ in range(reps):
            state.UpdateJMassMatrix(deposit_dt, copy_out=False)
        ms, n = ctx.timing()
        res[("dep", var)] = (ms[L.K_DEPOSIT_CELL] + ms[L.K_DEPOSIT_NODE]) / max(n[L.K_DEPOSIT_CELL], 1)
    best_var = min(DEPOSIT_VARIANTS, key=lambda v: res[("dep", v)])
    # the matrix-core and the lane passes (kernels 2 and 3) read a step's
    # particles where the migration found them, so that the step skips the
    # record scatter: when another pass wins on the deposit alone, the whole
    # steps decide against the faster of those two
    fusable = [v for v in DEPOSIT_VARIANTS if v[0] in (2, 3)]
    if fusable and best_var not in fusable:
        alt = min(fusable, key=lambda v: res[("dep", v)])
        best_var, steps = weigh_steps(ctx, state, opts_kw, deposit_dt, (best_var, alt), reps)
        res.update({("step", v): t for v, t in steps.items()})
    set_deposit_variant(ctx, *best_var)
    for nb in NODE_BATCHES:
        ctx.set_param("node_batch", nb)
        ctx.reset_timing()
        for _ in range(reps):
            state.UpdateJMassMatrix(deposit_dt, copy_out=False)
        ms, n = ctx.timing()
        res[("node", nb)] = ms[L.K_DEPOSIT_NODE] / max(n[L.K_DEPOSIT_NODE], 1)
    best_nb = min(NODE_BATCHES, key=lambda b: res[("node", b)])
    ctx.set_param("node_batch", best_nb)
    ctx.set_timing(False)
    best = {"tile": int(best_tile), "push_waves": int(best_waves), "deposit_kernel": int(best_var[0]),
            "deposit_chunks": int(best_var[1]), "node_batch": int(best_nb),
            "push_ms": {str(t): round(res[("tile", t)], 4) for t in PUSH_TILES},
            "push_waves_ms": {str(w): round(res[("waves", w)], 4) for w in PUSH_WAVES},
            "deposit_ms": {f"k{k}c{c}": round(res[("dep", (k, c))], 4) for k, c in DEPOSIT_VARIANTS},
            "step_ms": {f"k{key[1][0]}c{key[1][1]}": round(v, 4) for key, v in res.items() if key[0] == "step"},
            "node_batch_ms": {str(b): round(res[("node", b)], 4) for b in NODE_BATCHES}}
    db[signature] = best
    _save(cache, db)
    return best, False


def tune_tracer(ctx, run, n_cu, signature, cache=None):
    """Pick the persistent grid size and refill chunk for the tracer; run()
    executes one representative tracing call."""
    from . import cutoff as CU

    db = _load(cache)
    if signature in db:
        best = db[signature]
        CU.set_trace_param(ctx, "workgroups", best["workgroups"])
        CU.set_trace_param(ctx, "chunk", best["chunk"])
        return best, True
    res = {}
    for wg in (2 * n_cu, 3 * n_cu, 4 * n_cu, 8 * n_cu):
        for ch in (16, 64):
            CU.set_trace_param(ctx, "workgroups", wg)
            CU.set_trace_param(ctx, "chunk", ch)
            t0 = time.perf_counter()
            run()
            res[(wg, ch)] = time.perf_counter() - t0
    (wg, ch) = min(res, key=res.get)
    CU.set_trace_param(ctx, "workgroups", wg)
    CU.set_trace_param(ctx, "chunk", ch)
    best = {"workgroups": int(wg), "chunk": int(ch),
            "seconds": {f"{k[0]}x{k[1]}": round(v, 4) for k, v in res.items()}}
    db[signature] = best
    _save(cache, db)
    return best, False

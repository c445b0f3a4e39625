"""Small autotuner for the list path's launch and list parameters.

Results of the list paths do not depend on the skin, the cell budget or the
capacities (lists are sorted and pairs outside r_cut contribute nothing;
tests/test_gpu_parity.py checks it), and block size and tiles per block only
regroup the FP64 reductions, so the tuner measures speed alone.  It runs
short trajectories from the caller's state and keeps the fastest setting, one
parameter at a time (skin, block size, tiles per block, cell-grid budget).

The caller should pass a state typical of the run to be tuned for: under a
barostat the density, and with it the list length and the rebuild interval,
can change by several times over a run.
"""
from __future__ import annotations

import time

from .engine import Engine

SKINS = (0.1, 0.2, 0.35, 0.5, 0.7, 1.0)
BLOCKS = (128, 256, 512)
TILES = (0, 1, 2, 4, 8)
CELLS = (0.5, 1.0, 2.0, 4.0)


def _trial(counts, masses, box, x, v, dt, controls, steps, warmup, force_mode, opts, repeats=1):
    e = Engine(counts, masses, box, force_mode=force_mode, **opts)
    try:
        e.upload(x, v, box)
        e.update_force()
        th, ba = controls()
        e.step(dt, warmup, th, ba)
        # the fastest of `repeats` consecutive windows: a one-off stall of the
        # host or the GPU should not decide between settings a few per cent
        # apart
        t = float("inf")
        for _ in range(max(1, repeats)):
            t0 = time.perf_counter()
            e.step(dt, steps, th, ba)
            t = min(t, (time.perf_counter() - t0) / steps)
        return t, e.stats()
    finally:
        e.close()


def autotune(counts, masses, box, x, v, dt, controls, steps=300, warmup=20,
             force_mode="fp32_list", skins=SKINS, blocks=BLOCKS, tiles=TILES, cells=CELLS,
             log=None, repeats=1):
    """controls() -> (thermostat, barostat) as passed to Engine.step (fresh objects).
    Each setting is timed over `repeats` windows of `steps` steps (the fastest
    counts).

    Returns (best_options, table) where table lists (options, seconds_per_step, stats)."""
    best = {"skin": 0.3 * 0.3418, "block_size": 256, "tiles_per_block": 0, "cells_per_atom": 2.0}
    table = []

    def run(opts):
        t, st = _trial(counts, masses, box, x, v, dt, controls, steps, warmup, force_mode, opts, repeats)
        table.append((dict(opts), t, st))
        if log:
            log(f"[autotune] {opts} -> {t * 1e6:.2f} us/step (rebuilds {st['rebuilds']}, "
                f"<n_l> {st['mean_list_len']:.3f})")
        return t

    for key, values in (("skin", skins), ("block_size", blocks), ("tiles_per_block", tiles),
                        ("cells_per_atom", cells)):
        scores = {}
        for val in values:
            opts = dict(best)
            opts[key] = val
            scores[val] = run(opts)
        best[key] = min(scores, key=scores.get)
    return best, table

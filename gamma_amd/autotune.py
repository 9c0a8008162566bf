"""Autotuner for the launch options of the HIP backend.

Candidates: threads per workgroup of the per-cell kernels (64, 128, 256),
XCD-aware tile order (off/on), two-step hipGraph replay (off/on) and, when a
trained network is given, the learned cons2prim guess (off, or on with
trigger exponents 2 and 4).  Each candidate is timed on the live grid (the
production state, not a toy): a short untimed run to settle (and capture the
graph), then `steps` timed steps between stream synchronisations.  The grid
keeps evolving during tuning, so candidates are interleaved over two rounds
and each scored by its best round.  The winner is applied to the grid and
returned with the full table.
"""
from __future__ import annotations

import itertools
import time


def candidates(graph_options=(0, 1), mlp_options=(None,)):
    for block, remap, graph, mlp in itertools.product((64, 128, 256), (0, 1), graph_options, mlp_options):
        yield {"block": block, "remap": remap, "graph": graph, "mlp_thr": mlp}


def apply(grid, cand: dict, surrogate=None):
    grid.set_options(block=cand["block"], remap=cand["remap"], graph=cand["graph"])
    if cand.get("mlp_thr") is None:
        if surrogate is not None:
            grid.load_surrogate(None)
    else:
        grid.load_surrogate(surrogate)
        grid.set_options(mlp_thr=cand["mlp_thr"])


def tune(grid, steps: int = 6, settle: int = 2, rounds: int = 2, graph_options=(0, 1),
         surrogate=None, mlp_options=(None, 2, 4)) -> dict:
    cands = list(candidates(graph_options, mlp_options if surrogate is not None else (None,)))
    best_ms = {i: float("inf") for i in range(len(cands))}
    for _ in range(rounds):
        for i, c in enumerate(cands):
            apply(grid, c, surrogate)
            grid.step(settle)
            grid.sync()
            t0 = time.perf_counter()
            grid.step(steps)
            grid.sync()
            ms = 1e3 * (time.perf_counter() - t0) / steps
            best_ms[i] = min(best_ms[i], ms)
    win = min(best_ms, key=best_ms.get)
    apply(grid, cands[win], surrogate)
    table = [dict(cands[i], ms_per_step=round(best_ms[i], 4)) for i in range(len(cands))]
    return {"best": dict(cands[win], ms_per_step=round(best_ms[win], 4)), "table": table}

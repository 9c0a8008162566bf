"""Autotuner for the launch options of the HIP backend.

Candidates: threads per workgroup of the per-cell kernels (64, 128, 256),
XCD-aware tile order (off/on) and two-step hipGraph replay (off/on).  Each
candidate is timed on the live grid (the production state, not a toy): a
short untimed run to settle (and capture the graph), then `steps` timed steps
between stream synchronisations.  The grid keeps evolving during tuning, so
candidates are interleaved over two rounds and each scored by its best round.
The winner is applied to the grid and returned with the full table.
"""
from __future__ import annotations

import itertools
import time


def candidates(graph_options=(0, 1)):
    for block, remap, graph in itertools.product((64, 128, 256), (0, 1), graph_options):
        yield {"block": block, "remap": remap, "graph": graph}


def tune(grid, steps: int = 6, settle: int = 2, rounds: int = 2, graph_options=(0, 1)) -> dict:
    cands = list(candidates(graph_options))
    best_ms = {i: float("inf") for i in range(len(cands))}
    for _ in range(rounds):
        for i, c in enumerate(cands):
            grid.set_options(**c)
            grid.step(settle)
            grid.sync()
            t0 = time.perf_counter()
            grid.step(steps)
            grid.sync()
            ms = 1e3 * (time.perf_counter() - t0) / steps
            best_ms[i] = min(best_ms[i], ms)
    win = min(best_ms, key=best_ms.get)
    grid.set_options(**cands[win])
    table = [dict(cands[i], ms_per_step=round(best_ms[i], 4)) for i in range(len(cands))]
    return {"best": dict(cands[win], ms_per_step=round(best_ms[win], 4)), "table": table}

"""Starts of the cons2prim solve compared on the solver's own states.

The CPU build of the kernel bodies (tests/libgammahip_emu.so) records the
update's cons2prim solves of a run -- the conserved state (D, |S|, tau), the
previous pressure the reference starts from (rel_cart.cpp:170-172,
rel_sph.cpp:245), the root, the cell's shock flag -- and the FP64 Newton solve
of the backend (gammahip_c2p_batch on the GPU) then runs each state from three
starts: the previous pressure, the learned network (alone, and judged against
the previous pressure by |f| as the update kernel does) and a start without
history, (gamma - 1) tau.  Newton returns the double below the exact root
from any start, so the three runs solve the same problems and their
residual-evaluation counts compare directly.

Run as a script on a GPU host for the report on the benchmark configurations
(16 M states: the 4000-cell blast wave's 2000 steps and a 32-track strip of
the 2048 x 2048 Riemann problem for its 500 steps):

    python tests/c2p_starts.py [--states N] > report.json
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

from gamma_amd import hip, setups, surrogate as SG  # noqa: E402

EMU_LIB = os.path.join(HERE, "libgammahip_emu.so")
COLS = ("D", "S", "tau", "p_prev", "p", "shocked", "stage")


def capture(setup, nsteps: int, stride: int = 1, cap: int = 1 << 20, warmup: int = 0,
            solver=hip.C2P_BRENT) -> np.ndarray:
    """Rows [n][7] (COLS) of the update's cons2prim solves over `nsteps`
    steps after `warmup`, one in every `stride`, at most `cap`."""
    L = hip.load_library(EMU_LIB)
    L.gammahip_emu_capture.restype = C.c_longlong
    L.gammahip_emu_capture.argtypes = [C.c_void_p, C.c_longlong, C.c_longlong]
    g = hip.Grid(setup, solver=solver, lib_path=EMU_LIB)
    if warmup:
        g.step(warmup)
    buf = np.zeros((cap, 7))
    L.gammahip_emu_capture(buf.ctypes.data, cap, stride)
    try:
        done = 0
        while done < nsteps:              # in chunks, reporting progress on long runs
            k = min(100, nsteps - done)
            g.step(k)
            done += k
            if nsteps > 100:
                print(f"[c2p starts] captured {done}/{nsteps} steps", file=sys.stderr, flush=True)
    finally:
        n = L.gammahip_emu_capture(None, 0, 0)
    g.close()
    return buf[:n].copy()


def starts(rows: np.ndarray, gamma: float, w=None, maxit: int = 100) -> dict:
    """Residual evaluations per solve and failures of the GPU Newton solve
    from each start, overall and on the rows of shocked cells."""
    st = np.ascontiguousarray(rows[:, :3])
    prev = rows[:, 3].copy()
    runs = {"previous_p": SG.c2p_batch(st, prev, gamma, maxit=maxit),
            "cold": SG.c2p_batch(st, SG.cold_guess(st, gamma), gamma, maxit=maxit)}
    if w is not None:
        net = SG.infer(w, st, mfma=True)
        runs["network"] = SG.c2p_batch(st, net, gamma, maxit=maxit)
        runs["network_judged"] = SG.c2p_batch(st, net, gamma, guess2=prev, maxit=maxit)
    shocked = rows[:, 5] != 0
    ok = np.ones(len(rows), dtype=bool)
    for r in runs.values():
        ok &= r["fail"] == 0
    # the update kernel's trigger: a wavefront asks the network only where the
    # previous pressure leaves |f| above 10^-n of the magnitude of f's terms
    from mlp_ref import c2p_residual
    f, _, fs = c2p_residual(prev, st[:, 0], st[:, 1], st[:, 2], gamma)
    rel = np.abs(f) / fs
    out = {"states": int(len(rows)), "shocked_states": int(shocked.sum())}
    mean = lambda a, m: float(a[m].mean()) if m.any() else None
    for k, r in runs.items():
        e = {"evals_per_solve": mean(r["iters"], ok), "failures": int((r["fail"] != 0).sum()),
             "evals_per_solve_shocked": mean(r["iters"], ok & shocked)}
        for n in (2, 3, 4):
            e[f"evals_per_solve_where_trigger_{n}"] = mean(r["iters"], ok & (rel > 10.0 ** -n))
        if k == "network_judged":
            e["network_won"] = mean(r["used"].astype(float), ok)
        out[k] = e
    for n in (2, 3, 4):
        t = rel > 10.0 ** -n
        out[f"trigger_{n}_fraction"] = float(t.mean())
        if "network_judged" in runs:
            # the kernel's scheme: the judged start where the trigger fires,
            # the previous pressure elsewhere
            ev = np.where(t, runs["network_judged"]["iters"], runs["previous_p"]["iters"])
            out[f"triggered_{n}"] = {"evals_per_solve": mean(ev, ok),
                                     "evals_per_solve_shocked": mean(ev, ok & shocked)}
    # every start must land on the same root (Newton's answer does not depend on it)
    ref = runs["previous_p"]["p"]
    out["same_root"] = bool(all(np.array_equal(r["p"][ok], ref[ok]) for r in runs.values()))
    return out


def train_network(rows: np.ndarray, gamma: float, n_train: int = 1 << 20, steps: int = 3000, seed: int = 5):
    """The network trained on the GPU on solver states (D, |S|, tau, root)
    plus synthetic states over the stated ranges."""
    rng = np.random.default_rng(seed)
    take = rng.choice(len(rows), min(n_train, len(rows)), replace=False)
    sol = np.stack([rows[take, 0], rows[take, 1], rows[take, 2], rows[take, 4]], axis=1)
    sol = sol[np.isfinite(sol).all(axis=1) & (sol[:, 3] > 0)]
    train = np.concatenate([sol, SG.synthetic_states(len(sol), gamma, seed=seed)])
    w, loss = SG.train(train, steps=steps, batch=4096, lr=2e-3, seed=seed)
    return w, float(loss[-100:].mean())


def report(n_states: int = 16 << 20) -> dict:
    """The comparison on the benchmark configurations' own solves."""
    out = {}
    half = n_states // 2
    cases = [("blandford_mckee_1d 4000 cells, 2000 steps", setups.blandford_mckee_1d(4000), 2000, 0),
             ("riemann_2d 2048 cells x 32 tracks, 500 steps", setups.riemann_2d(2048, 32, ly=32 / 2048), 500, 0)]
    for name, s, n, warm in cases:
        t0 = time.perf_counter()
        # solves per run: ~5 per active cell per step
        total = 5 * n * (s.x.size if s.cfg.nd == 1 else s.x.size)
        stride = max(1, total // half)
        rows = capture(s, n, stride=stride, cap=half, warmup=warm)
        t1 = time.perf_counter()
        w, loss = train_network(rows, s.cfg.gamma)
        res = starts(rows, s.cfg.gamma, w)
        res.update(capture_s=round(t1 - t0, 1), train_loss=loss, stride=int(stride))
        out[name] = res
        print(f"[c2p starts] {name}: {res}", file=sys.stderr, flush=True)
    return out


if __name__ == "__main__":
    n = int(sys.argv[sys.argv.index("--states") + 1]) if "--states" in sys.argv else 16 << 20
    print(json.dumps(report(n)))

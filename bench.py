"""Benchmark of the MI355X backend on the 2D relativistic Riemann problem.

Workload (the configuration the headline metric is quoted on): 2D cartesian
relativistic Riemann problem on x-directed Lagrangian tracks, 2048 tracks x
2048 cells, FP64, left rho=10 p=13.33, right rho=1 p=1e-6, gamma=5/3, 1e-3
seeded density perturbation, CFL 0.2, host-rule AMR (dx/dy split 3 / merge
0.2), plain per-stage kernels.  A step is one full RK3 step (regrid, ghosts,
three flux/update stages).  Metric: cell-updates/s = sum over timed steps of
the active cells of every track, divided by wall time; inputs resident in HBM.

Prints one JSON line.  Only one GPU: --gpus N>1 exits.
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

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s

# HIP-event timed kernels (ids of gammahip_kernel_times)
KNAMES = {0: "k_update", 1: "k_faces", 2: "k_f1faces", 3: "k_f1grad", 4: "k_f1sum", 5: "k_f1cell"}


def algorithmic_bytes(kid: int, cells: float, f1faces: float, nq: int, stage_mix: float) -> float:
    """Bytes one launch must move at minimum: every array element the kernel
    reads or writes, once, for `cells` active cells (neighbour stencil reads
    shared between cells are counted once).  stage_mix: fraction of launches
    that are RK stages 2/3 (which also read the RK base)."""
    d, i4 = 8.0, 4.0
    if kid == 0:   # k_update: prim, x, dx, face x, v, flux in; prim, x, dx, face x out;
        # RK base (U0, dV0, x0) written in stage 1 / read in stages 2-3; net F1 flux in (2D);
        # in stages 2-3 also the dt candidates of the stage's faces: the faces' two
        # signal speeds in, the cell's dt_loc read and written
        per = (nq + 2 + 2 + nq) * d + (nq + 3) * d + (nq + 2) * d + stage_mix * 4 * d
        if f1faces > 0:
            per += nq * d
        return cells * per
    if kid == 1:   # k_faces: prim, centroid inputs (x, dx), face x in; v, lL, lR, flux, grad_mv out
        return cells * ((nq + 2 + 1) * d + (3 + nq + nq) * d)
    if kid == 2:   # k_f1faces (one thread per F1 face): owner cell prim, grad_mv, new grad_f1,
        # x, dx once per cell; per face the upper cell's prim, grad_mv, old grad_f1,
        # x, dx, six slot/list ints, and the flux out
        return cells * (3 * nq + 2) * d + f1faces * ((3 * nq + 2) * d + nq * d + 6 * i4)
    if kid == 3:   # k_f1grad: prim, grad_mv, x, dx and the neighbour lists of each cell
        # (neighbours' values shared between cells), grad_f1 out
        return cells * ((3 * nq + 2) * d + 6 * i4)
    if kid == 4:   # k_f1sum: each F1 face flux read twice (by the lower cell, whose
        # slot holds it, and by the upper cell), slot counts/offsets in, net flux out
        return cells * (nq * d + 4 * i4) + 2 * f1faces * nq * d
    if kid == 5:   # k_f1cell (k_f1grad + k_f1faces fused): per cell prim, grad_mv, x, dx and
        # the neighbour lists once, grad_f1 out; per face the upper cell's old grad_f1,
        # slot ints and the flux out (its prim, grad_mv, x, dx are neighbour reads
        # already counted once per cell)
        return cells * ((3 * nq + 2) * d + 6 * i4) + f1faces * (nq * d + nq * d + 6 * i4)
    return 0.0


def cpu_baseline(sample_tracks: int = 256, nx: int = 2048, steps: int = 20) -> dict:
    """Oracle (the CPU restatement: track loops on all host cores, as the
    reference's OpenMP loops run) on a bounded sample of the same workload:
    the bottom `sample_tracks` tracks of the 2048 x 2048 grid (same cells,
    same seeded perturbation), `steps` RK3 steps after 2 warm-up steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as oc
    from gamma_amd import setups
    L = oc.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    L.og_set_threads(threads)
    s = setups.riemann_2d(nx, sample_tracks, ly=sample_tracks / nx)
    g = oc.OracleGrid(s)
    g.run(2)
    n0 = sum(len(a) for a in g.active(0))
    t0 = time.perf_counter()
    g.run(steps)
    el = time.perf_counter() - t0
    n1 = sum(len(a) for a in g.active(0))
    cu = 0.5 * (n0 + n1) * steps
    return {"value": cu / el, "unit": "cell-updates/s", "cores": int(L.og_threads()), "kind": "port",
            "sample": f"oracle, riemann_2d {nx} cells x {sample_tracks} tracks, {steps} RK3 steps "
                      f"after 2 warm-up steps, {el:.1f} s on {int(L.og_threads())} threads"}


def b_cu(rc) -> float:
    """Algorithmic bytes per cell-update of a fused RK3 step: 8 * [9 (Q_live + 1)
    + (H + 1)], Q_live = NUM_D momenta + D + tau + 2 tracers (+ 4 with
    synchrotron), H = NUM_D + 2 hydro primitives."""
    q_live = rc.nd + 2 + 2 + (4 if rc.sync else 0)
    h = rc.nd + 2
    return 8.0 * (9 * (q_live + 1) + (h + 1))


COPY_PROBE = r"""
import json, torch
n = 1 << 31
a = torch.empty(n // 8, dtype=torch.float64, device="cuda"); b = torch.empty_like(a); a.fill_(1.0)
for _ in range(3):
    b.copy_(a)
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
torch.cuda.synchronize(); e0.record()
for _ in range(10):
    b.copy_(a)
e1.record(); torch.cuda.synchronize()
print(json.dumps({"gbs": 2 * n / (e0.elapsed_time(e1) / 10 * 1e-3) / 1e9}))
"""


def copy_peak_gbs():
    """Measured streaming-copy bandwidth (bytes read + written per second) of a
    2 GiB device-to-device copy (torch's copy kernel, timed with events).  It
    runs in a process of its own: torch carries its own HIP runtime, which
    cannot open the device in a process where the library's runtime already
    has."""
    import subprocess
    r = subprocess.run([sys.executable, "-c", COPY_PROBE], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-300:])
    return json.loads(r.stdout.strip().splitlines()[-1])["gbs"]


def run_form(setup, solver, options, steps, warmup, timing_steps, surrogate=None):
    """Time `steps` RK3 steps after `warmup` on a fresh grid; then per-kernel
    HIP-event times over `timing_steps` extra steps."""
    from gamma_amd import hip
    g = hip.Grid(setup, solver=solver)
    if surrogate is not None:
        g.load_surrogate(surrogate)
    if options:
        g.set_options(**options)
    g.step(warmup)
    g.sync()
    n0 = track_cells(g)
    g.sync()
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    el = time.perf_counter() - t0
    n1 = track_cells(g)
    st = g.stats()
    nf1 = float(sum(a.sum() for a in g.active(hip.F_NF1))) if setup.cfg.nd == 2 else 0.
    g.kernel_times()                      # switches event timing on and clears
    g.step(timing_steps)
    g.sync()
    ms, nl = g.kernel_times()
    g.close()
    return {"value": 0.5 * (n0 + n1) * steps / el, "ms_per_step": 1e3 * el / steps,
            "stats": st, "ms": ms, "launches": nl, "cells": n1, "f1faces": nf1}


def track_cells(g):
    """Active cells summed over the active tracks (the metric's unit count)."""
    tr = g.tracks()
    rc = g.rc
    nact = tr["nact"]
    return int(nact[rc.ngst:-rc.ngst].sum()) if rc.nd == 2 else int(nact.sum())


def run_other(setup, steps, warmup, options=None, solver=None, surrogate=None):
    """cell-updates/s of a secondary configuration (no per-kernel timing)."""
    from gamma_amd import hip
    g = hip.Grid(setup, solver=hip.C2P_BRENT if solver is None else solver)
    if surrogate is not None:
        g.load_surrogate(surrogate)
    if options:
        g.set_options(**options)
    g.step(warmup)
    g.sync()
    n0 = track_cells(g)
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    el = time.perf_counter() - t0
    n1 = track_cells(g)
    st = g.stats()
    lens = g.tracks()["ntrack"] - 2 * g.rc.ngst
    g.close()
    return {"value": 0.5 * (n0 + n1) * steps / el, "ms_per_step": 1e3 * el / steps, "cells": n1,
            "track_cells_min": int(lens.min()), "track_cells_max": int(lens.max()),
            "c2p_iters_per_call": st["c2p_iters"] / max(st["c2p_calls"], 1),
            "network_starts": st["mlp_used"],
            "failures": {k: st[k] for k in ("c2p_nobracket", "c2p_maxiter", "nan_flux", "negative_dt")},
            "splits": st["splits"], "merges": st["merges"]}


def other_configs(steps, warmup):
    """The uneven-track configurations: the 2D spherical jet blast wave (1024
    theta tracks, seeded lengths 500-4000) and the ensemble of 256 1D runs
    (512-16384 cells), each plain and batched (cons2prim solver and launch
    options -- packed persistent grids, XCD order, graphs, block size --
    picked by the autotuner on a scratch copy)."""
    from gamma_amd import autotune, hip, setups
    out = {}
    for name, mk in (("bm_jet_2d", lambda: setups.bm_jet_2d()),
                     ("ensemble_1d", lambda: setups.ensemble_1d()[0])):
        s = mk()
        print(f"[bench] {name}: plain", file=sys.stderr, flush=True)
        p = run_other(s, steps, warmup)
        print(f"[bench] {name}: autotuning and batched", file=sys.stderr, flush=True)
        scratch = hip.Grid(s, solver=hip.C2P_NEWTON)
        scratch.step(warmup)
        from gamma_amd import surrogate as SG
        w, _ = SG.train(np.concatenate([SG.states_from_grid(scratch, s.cfg.gamma, max_states=1 << 20),
                                        SG.synthetic_states(1 << 20, s.cfg.gamma, seed=3)]),
                        steps=3000, batch=4096, lr=2e-3, seed=5)
        tuned = autotune.tune(scratch, steps=4, settle=2, sweeps=1, surrogate=w)
        scratch.close()
        batched = {k: tuned["best"][k] for k in autotune.OPTION_KEYS}
        b = run_other(s, steps, warmup, batched, batched["solver"],
                      surrogate=w if batched["mlp_thr"] is not None else None)
        out[name] = {"plain": p, "batched": dict(b, options=batched),
                     "speedup": b["value"] / p["value"]}
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "riemann_2d_2048_hbm_traffic.json")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by what
    the counter reads for a copy kernel of known size), or None."""
    t = _traffic()
    for name, v in t.get("kernels", {}).items():
        if name.split("<")[0] == kernel:
            return v["bytes_per_launch"]
    return None


def _traffic() -> dict:
    try:
        return json.load(open(TRAFFIC_FILE))
    except (OSError, ValueError):
        return {}


def measured_step_traffic():
    """Counter bytes of one whole step (every solver kernel's bytes per launch
    times its launches per step) from the same summary, or None.  The
    runtime's fill and copy kernels in the summary run at set-up (each plane
    zeroed as it is allocated, the initial state uploaded) and are left out."""
    t = _traffic()
    steps = t.get("steps")
    if not steps:
        return None
    return sum(v["bytes_per_launch"] * v["dispatches"] for k, v in t.get("kernels", {}).items()
               if "__amd_rocclr" not in k) / steps


def kernel_table(res, nq):
    f1faces = res["f1faces"]
    kern = {}
    for kid, name in KNAMES.items():
        if res["launches"][kid]:
            by = algorithmic_bytes(kid, res["cells"], f1faces, nq, 2. / 3.)
            ms = res["ms"][kid]
            kern[name] = {"ms": ms, "launches": int(res["launches"][kid]), "bytes": by,
                          "GBs": by / (ms * 1e-3) / 1e9}
    return kern


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=2)
    ap.add_argument("--no-other", action="store_true",
                    help="skip the uneven-track configurations")
    ap.add_argument("--other-steps", type=int, default=100)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the 2 GiB device copy whose counters calibrate the traffic figures")
    ap.add_argument("--profile-config", choices=("bm_jet_2d", "ensemble_1d"), default=None,
                    help="run only that configuration's plain form for --steps steps (for profilers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU only; sharding the tracks across GPUs "
                                   "(ghost-track exchange per stage, one dt reduction per step) is "
                                   "designed, not built", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.copy_probe:
        exec(COPY_PROBE, {})
        return
    from gamma_amd import autotune, hip, setups
    if args.profile_config:
        s = setups.bm_jet_2d() if args.profile_config == "bm_jet_2d" else setups.ensemble_1d()[0]
        print(json.dumps({"profile_config": args.profile_config,
                          "plain": run_other(s, args.steps, args.warmup)}))
        return
    s = setups.riemann_2d(args.nx, args.ny)
    nq = s.cfg.nq
    print(f"[bench] riemann_2d {args.nx}x{args.ny}: plain form", file=sys.stderr, flush=True)
    plain = run_form(s, hip.C2P_BRENT, None, args.steps, args.warmup, args.timing_steps)
    print(f"[bench] plain {plain['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    kern = kernel_table(plain, nq)
    dom = max(kern, key=lambda k: kern[k]["ms"] * kern[k]["launches"])
    bcu = b_cu(s.cfg)
    # the whole-step byte model: a fused step moves B_cu per cell-update, a third
    # of it per stage; the dominant (per-stage) kernel is charged one stage's
    # share of it for the cells it advances
    alg = bcu / 3.0 * plain["cells"]
    ach = alg / (kern[dom]["ms"] * 1e-3) / 1e9
    step_bytes = measured_step_traffic()
    roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom), "kernel": dom,
            "kernel_ms": round(kern[dom]["ms"], 4), "algorithmic_bytes": alg,
            "bytes_per_cell_update": bcu,
            "step_frac": round(plain["value"] * bcu / (HBM_PEAK_GBS * 1e9), 4),
            "step_counter_bytes": step_bytes,
            "kernel_own_model": {"bytes": kern[dom]["bytes"], "achieved": round(kern[dom]["GBs"], 1),
                                 "frac": round(kern[dom]["GBs"] / HBM_PEAK_GBS, 4)}}
    try:
        roof["copy_peak"] = round(copy_peak_gbs(), 1)
    except Exception as e:                # torch missing or no device memory left: say so
        roof["copy_peak"] = None
        roof["copy_peak_error"] = str(e)[:200]

    batched = None
    if not args.no_batched:
        from gamma_amd import surrogate as SG
        print("[bench] training the cons2prim network on solver states", file=sys.stderr, flush=True)
        scratch = hip.Grid(s, solver=hip.C2P_NEWTON)
        scratch.step(args.warmup)
        t0 = time.perf_counter()
        train_states = np.concatenate([
            SG.states_from_grid(scratch, s.cfg.gamma, max_states=1 << 20),
            SG.synthetic_states(1 << 20, s.cfg.gamma, seed=3)])
        w, loss = SG.train(train_states, steps=3000, batch=4096, lr=2e-3, seed=5)
        train_s = time.perf_counter() - t0
        print("[bench] autotuning the batched form", file=sys.stderr, flush=True)
        tuned = autotune.tune(scratch, surrogate=w)
        scratch.close()
        opts = {k: tuned["best"][k] for k in autotune.OPTION_KEYS}
        b = run_form(s, opts["solver"], opts, args.steps, args.warmup, args.timing_steps,
                     surrogate=w if opts["mlp_thr"] is not None else None)
        print(f"[bench] batched {b['ms_per_step']:.3f} ms/step with {opts}", file=sys.stderr, flush=True)
        bst = b["stats"]
        batched = {"value": b["value"], "ms_per_step": b["ms_per_step"], "options": opts,
                   "cons2prim": ("Brent (the reference's algorithm)" if opts["solver"] == hip.C2P_BRENT else
                                 "Newton to the rounded exact root, previous-pressure start"
                                 + ("" if opts["mlp_thr"] is None else
                                    ", learned start judged against it by |f|")),
                   "speedup_vs_plain": b["value"] / plain["value"],
                   "c2p_iters_per_call": bst["c2p_iters"] / max(bst["c2p_calls"], 1),
                   "network_starts": bst["mlp_used"], "c2p_calls": bst["c2p_calls"],
                   "failures": {k: bst[k] for k in ("c2p_nobracket", "c2p_maxiter", "c2p_phicap")},
                   "network": {"train_states": int(train_states.shape[0]), "train_steps": 3000,
                               "final_loss": float(loss[-100:].mean()), "train_s": round(train_s, 2)},
                   "kernels": kernel_table(b, nq), "autotune": tuned["table"]}

    other = None if args.no_other else other_configs(args.other_steps, 3)
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    st = plain["stats"]
    out = {
        "metric": "cell-updates/s (full RK3 step)", "value": plain["value"], "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded SplitMix64 perturbation)",
        "config": {"workload": f"riemann_2d {args.ny} tracks x {args.nx} cells, 2D cartesian, "
                               f"plain per-stage kernels, Brent cons2prim",
                   "tracks": args.ny, "cells_per_track": args.nx, "parallelism": "single GPU",
                   "timed_steps": f"{args.warmup + 1}..{args.warmup + args.steps} of the run"},
        "roofline": roof, "cpu_baseline": cpu,
        "kernels": kern,
        "c2p_iters_per_call": st["c2p_iters"] / max(st["c2p_calls"], 1),
        "failures": {k: st[k] for k in ("c2p_nobracket", "c2p_maxiter", "negative_dt", "nan_flux",
                                        "f1_overflow")},
        "regrid": {"splits": st["splits"], "merges": st["merges"]},
        "batched": batched,
        "other_configs": other,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""The learned cons2prim start against the start the reference uses (the
previous pressure, rel_cart.cpp:170-172, rel_sph.cpp:245) and a start without
history, on the solver's own states (tests/c2p_starts.py)."""
from __future__ import annotations

import numpy as np
import pytest

from gamma_amd import hip, setups
import c2p_starts as CS


def test_emu_capture_records_every_solve(emu_lib):
    """One row per cons2prim call of the update (stage 1 one per cell, stages
    2-3 two), each a state with its previous pressure and root."""
    s = setups.blandford_mckee_1d(300)
    g = hip.Grid(s, lib_path=emu_lib)
    g.step(30)
    before = g.stats()["c2p_calls"]
    g.step(20)
    calls = g.stats()["c2p_calls"] - before
    g.close()
    rows = CS.capture(s, 20, cap=1 << 18, warmup=30)
    assert len(rows) == calls
    D, S, tau, pprev, p, shk, stage = rows.T
    assert np.all(D > 0) and np.all(S >= 0) and np.all(np.isfinite(tau))
    assert np.all(p > 0) and np.all(np.isfinite(pprev))
    assert set(np.unique(stage)) == {1., 2., 3.}
    assert (shk != 0).any() and (shk == 0).any()           # the blast wave's shocked cells
    # stages 2-3 start their second solve from the first one's root
    assert np.mean(pprev == p) > 0.1


@pytest.mark.gpu
def test_gpu_starts_on_the_blast_wave_solves():
    """On the blast wave's own solves (1D spherical, Lorentz factor 100 shock):
    every start lands on the same root, neither the reference's start nor the
    judged one fails, and the network alone needs fewer residual evaluations
    than the history-free start.  Where the previous pressure is far from the
    root -- |f| above 1 % of the magnitude of f's terms, the update kernel's
    trigger at n = 2 -- the network alone also needs fewer evaluations than
    the previous pressure.  Elsewhere the previous pressure is within the
    rounding of the root and the network cannot gain (counts printed)."""
    s = setups.blandford_mckee_1d(1000)
    print("[gpu c2p starts] capturing the blast wave's solves on the CPU build", flush=True)
    rows = CS.capture(s, 300, cap=1 << 21, warmup=20)
    w, loss = CS.train_network(rows, s.cfg.gamma, n_train=1 << 18, steps=2000)
    res = CS.starts(rows, s.cfg.gamma, w)
    print(f"[gpu c2p starts] {res}", flush=True)
    assert res["same_root"]
    assert res["shocked_states"] > 1000
    assert res["previous_p"]["failures"] == 0 and res["network_judged"]["failures"] == 0
    assert res["network"]["evals_per_solve"] < res["cold"]["evals_per_solve"]
    assert res["trigger_2_fraction"] * res["states"] > 200
    assert (res["network"]["evals_per_solve_where_trigger_2"]
            < res["previous_p"]["evals_per_solve_where_trigger_2"])

"""The Newton cons2prim of the kernels (gamma_amd/csrc/hydro.hpp: newton,
c2p_fdd): its answer is the double just below the exact root of the
reference's residual, whatever the start.

The reference solves f(p) = E + p - D W - G/(G-1) p W^2 = 0 with
W = 1/sqrt(1 - S^2/(E+p)^2) by Brent's method to a 1e-13 relative interval
(src/Hydro/rel_cart.cpp:108-123 and :125-229).  The kernels' Newton solver
evaluates an algebraically equal residual in double-double arithmetic and
returns the largest double p with f(p) > 0.  These tests evaluate the
reference's own formula in 60-digit decimal arithmetic (E = D + tau formed
exactly) at the returned p and at the next double up, and require the sign
change between them; then they require the same bits from different starts,
on the CPU build of the kernel bodies and (marked gpu) on the GPU.
"""
from __future__ import annotations

import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from gamma_amd import hip as H
from gamma_amd import surrogate as SG

G53 = 5.0 / 3.0


def f_exact(p: float, D: float, S: float, tau: float, G: float) -> Decimal:
    """rel_cart.cpp:108-123 in 60-digit decimal arithmetic, E = D + tau exact."""
    getcontext().prec = 60
    p, D, S, tau, G = (Decimal(float(v)) for v in (p, D, S, tau, G))
    E = D + tau
    Ep = E + p
    W = 1 / (1 - (S * S) / (Ep * Ep)).sqrt()
    return E + p - D * W - (G * p * W * W) / (G - 1)


def _states(n: int, seed: int) -> np.ndarray:
    """Slow, fast, cold and hot states, with cold gas in motion (kinetic energy
    many orders above the thermal one) well represented."""
    parts = [SG.synthetic_states(n, G53, seed=seed, lf=(1.0, 30.0), p_over_rho=(1e-6, 1e2)),
             SG.synthetic_states(n, G53, seed=seed + 1, lf=(1.0, 1e3), p_over_rho=(1e-8, 1e-2)),
             SG.synthetic_states(n, G53, seed=seed + 2, lf=(1.0, 1.01), p_over_rho=(1e-6, 1e-3))]
    return np.concatenate(parts)


def _starts(st: np.ndarray):
    p = st[:, 3]
    return [p, 1.7 * p, 0.3 * p, SG.cold_guess(st, G53)]


def test_newton_returns_the_double_below_the_exact_root(emu_lib):
    st = _states(200, seed=31)
    r = SG.c2p_batch(st, st[:, 3] * 1.3, G53, lib_path=emu_lib)
    checked = 0
    for k in range(len(st)):
        if r["fail"][k]:
            # only the bracket test may refuse (no root above the lower bound)
            assert r["fail"][k] == 1, (k, r["fail"][k])
            continue
        p = float(r["p"][k])
        D, S, tau = st[k, :3]
        lo, hi = f_exact(p, D, S, tau, G53), f_exact(math.nextafter(p, math.inf), D, S, tau, G53)
        assert lo > 0 >= hi, (k, p, float(lo), float(hi))
        checked += 1
    assert checked >= 0.99 * len(st)


def test_newton_same_bits_from_every_start(emu_lib):
    st = _states(2000, seed=7)
    res = [SG.c2p_batch(st, g, G53, lib_path=emu_lib) for g in _starts(st)]
    ok = np.all([r["fail"] == 0 for r in res], axis=0)
    assert ok.mean() > 0.99
    for r in res[1:]:
        assert np.array_equal(r["p"][ok], res[0]["p"][ok])


def test_newton_second_start_does_not_change_the_answer(emu_lib):
    """The judged second start (the learned guess's slot) changes the path,
    not the answer."""
    st = _states(1000, seed=9)
    base = SG.c2p_batch(st, st[:, 3], G53, lib_path=emu_lib)
    two = SG.c2p_batch(st, 0.5 * st[:, 3], G53, guess2=st[:, 3] * 3.0, lib_path=emu_lib)
    ok = (base["fail"] == 0) & (two["fail"] == 0)
    assert np.array_equal(base["p"][ok], two["p"][ok])


def test_newton_close_to_brent(emu_lib):
    """Brent stops on a 1e-13 interval around the same root."""
    st = _states(1000, seed=11)
    nw = SG.c2p_batch(st, st[:, 3], G53, lib_path=emu_lib)
    br = SG.c2p_batch(st, st[:, 3], G53, solver=H.C2P_BRENT, maxit=1000, lib_path=emu_lib)
    ok = (nw["fail"] == 0) & (br["fail"] == 0) & (st[:, 1] / st[:, 0] < 1.0)
    rel = np.abs(nw["p"][ok] - br["p"][ok]) / nw["p"][ok]
    # Brent's residual rounds at eps E: its root is known to eps E / (p |f'|)
    assert np.median(rel) < 1e-12, np.median(rel)


@pytest.mark.gpu
def test_gpu_newton_equals_cpu_build_bitwise(emu_lib):
    st = _states(4000, seed=13)
    for g in _starts(st)[:2]:
        a = SG.c2p_batch(st, g, G53)
        b = SG.c2p_batch(st, g, G53, lib_path=emu_lib)
        assert np.array_equal(a["p"], b["p"]) and np.array_equal(a["fail"], b["fail"])

"""Surrogate starting pressure for cons2prim: the network's CPU reference,
the GPU training/inference kernels against it, and the gain of the trained
network over starts without it.

The network has no counterpart in the reference; what it feeds is the
reference's cons2prim root search (src/Hydro/rel_cart.cpp:125-229), whose
result must not depend on the start beyond the solver tolerance.
"""
from __future__ import annotations

import numpy as np
import pytest

from gamma_amd import surrogate as SG
from tests import mlp_ref as R

G53 = 5.0 / 3.0


def _params(seed=3, n=4096):
    st = SG.synthetic_states(n, G53, seed=seed)
    w = SG.init_params(st, seed=seed).astype(np.float64)
    rng = np.random.default_rng(seed)
    w[SG.OB1:SG.OW2] = rng.uniform(-0.1, 0.1, SG.NH)
    w[SG.OB2:SG.OW3] = rng.uniform(-0.1, 0.1, SG.NH)
    w[SG.OB3] = 0.05
    return w.astype(np.float32), st


# ---------------------------------------------------------------- CPU pins
def test_reference_gradient_finite_differences():
    w, st = _params(n=96)
    X, T = R.features(w, st), R.targets(w, st)
    w64 = w.astype(np.float64)
    _, g = R.loss_grad(w64, X, T)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b in (
        (SG.OW1, SG.OB1), (SG.OB1, SG.OW2), (SG.OW2, SG.OB2), (SG.OB2, SG.OW3), (SG.OW3, SG.OB3))]
        + [np.array([SG.OB3])])
    h = 1e-6
    for k in idx:
        wp, wm = w64.copy(), w64.copy()
        wp[k] += h
        wm[k] -= h
        fd = (R.loss_grad(wp, X, T)[0] - R.loss_grad(wm, X, T)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_reference_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.140625, 65504.0], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even mantissa (1); 1 + 3*2^-9 rounds up
    np.testing.assert_array_equal(R.bf16(x), [1.0, 1.0, 1.0 + 2 ** -7, -3.140625, 65536.0])


def test_synthetic_states_match_oracle_prim2cons():
    from gamma_amd import setups
    from tests import oracle_ctypes as O
    rc = setups.RunConfig(nd=1, geom=setups.GEOM_CART, gamma=G53)
    st = SG.synthetic_states(64, G53, seed=2, lf=(1.0, 30.0), p_over_rho=(1e-4, 1e2))
    for D, S, tau, p in st:
        # recover the primitive state from (D, S, tau, p) and map it back through the oracle
        E = D + tau
        v = S / (E + p)
        lf = 1 / np.sqrt(1 - v * v)
        pr = np.zeros(rc.nq)
        pr[0], pr[1], pr[2] = D / lf, p, lf * v
        cons = O.prim2cons(rc, pr)
        np.testing.assert_allclose([cons[0], cons[2], cons[1]], [D, S, tau], rtol=1e-9)


def test_feature_layout_constants():
    assert SG.NPARAM == 3 * 64 + 64 + 64 * 64 + 64 + 64 + 1 + 8


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_mfma_lane_layout():
    """v_mfma_f32_32x32x16_bf16 operand/result layout assumed by the inference path."""
    import ctypes as C
    from gamma_amd import hip as H
    L = H.load_library()
    rng = np.random.default_rng(0)
    A = (rng.integers(-8, 9, (32, 16)) / 4).astype(np.float32)
    B = (rng.integers(-8, 9, (16, 32)) / 4).astype(np.float32)
    Dm = np.zeros((32, 32), dtype=np.float32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    assert L.gammahip_mfma_probe(p(A), p(B), p(Dm)) == 0
    np.testing.assert_array_equal(Dm, A.astype(np.float64) @ B.astype(np.float64))


@pytest.mark.gpu
def test_gpu_gradient_matches_cpu_reference():
    w, st = _params(n=1000)
    g, loss = SG.gradient(w, st)
    X, T = R.features(w, st), R.targets(w, st)
    lref, gref = R.loss_grad(w.astype(np.float64), X, T)
    assert abs(loss - lref) <= 1e-5 * lref
    for a, b in ((SG.OW1, SG.OB1), (SG.OB1, SG.OW2), (SG.OW2, SG.OB2), (SG.OB2, SG.OW3),
                 (SG.OW3, SG.OB3), (SG.OB3, SG.ONORM)):
        err = np.linalg.norm(g[a:b] - gref[a:b]) / max(np.linalg.norm(gref[a:b]), 1e-30)
        assert err < 1e-4, (a, b, err)


@pytest.mark.gpu
def test_gpu_inference_paths():
    w, st = _params(n=5000)
    X = R.features(w, st)
    y32 = R.y_of_pressure(w, st, SG.infer(w, st, mfma=False))
    ymf = R.y_of_pressure(w, st, SG.infer(w, st, mfma=True))
    yref, _ = R.forward(w.astype(np.float64), X)
    yemu = R.forward_mfma(w.astype(np.float64), X)
    spread = np.std(yref)
    assert np.max(np.abs(y32 - yref)) < 1e-4 * max(spread, 1.0)
    assert np.max(np.abs(ymf - yemu)) < 1e-3 * max(spread, 1.0)
    # BF16 weights cost little next to the spread of the output
    assert np.sqrt(np.mean((ymf - yref) ** 2)) < 0.05 * spread


def _train_set(n_syn=1 << 18):
    syn = SG.synthetic_states(n_syn, G53, seed=21)
    return syn


@pytest.mark.gpu
def test_training_converges_and_gains_over_cold_start():
    """Train on the GPU, then count FP64 Newton work per state from the
    network's start and from the history-free start (gamma - 1) tau."""
    train = _train_set()
    w, loss = SG.train(train, steps=3000, batch=4096, lr=2e-3, seed=5)
    assert np.all(np.isfinite(loss))
    assert loss[-200:].mean() < 0.05 * loss[:20].mean(), (loss[:5], loss[-5:])
    test = SG.synthetic_states(1 << 16, G53, seed=99)
    guess = SG.infer(w, test, mfma=True)
    cold = SG.cold_guess(test, G53)
    rm = SG.c2p_batch(test, guess, G53)
    rc = SG.c2p_batch(test, cold, G53)
    # Failures come from states whose rounded conserved variables admit no
    # root above the solver's lower bracket (cold, ultra-relativistic corner);
    # the bracket test does not depend on the start, so both starts fail there.
    fm, fc = rm["fail"] != 0, rc["fail"] != 0
    print(f"\nfailures: network start {fm.sum()} (bits {np.unique(rm['fail'])}), "
          f"cold start {fc.sum()} (bits {np.unique(rc['fail'])}) of {len(test)}")
    assert fm.sum() <= fc.sum() and fm.sum() <= 1e-3 * len(test)
    ok = ~(fm | fc)
    # Newton returns the double below the root of the conserved state whatever
    # its start: both starts give the same bits, a root to working precision
    assert np.array_equal(rm["p"][ok], rc["p"][ok])
    D, S, tau = test[ok, 0], test[ok, 1], test[ok, 2]
    eps = np.finfo(np.float64).eps
    f, _, fs = R.c2p_residual(rc["p"][ok], D, S, tau, G53)
    assert np.all(np.abs(f) <= 16 * eps * fs)
    it_m, it_c = rm["iters"][ok].mean(), rc["iters"][ok].mean()
    print(f"residual evaluations per state: network start {it_m:.2f}, cold start {it_c:.2f}, "
          f"final loss {loss[-100:].mean():.4g}")
    assert it_m < 0.8 * it_c

"""Learned starting pressure for cons2prim (no counterpart in the reference).

The reference starts its cons2prim root search from the cell's previous
pressure (src/Hydro/rel_cart.cpp:125-229, src/Hydro/rel_sph.cpp:213-320).
This module trains and runs a 3 -> 64 -> 64 -> 1 MLP that predicts log p of a
conserved state from (log D, |S|/D, tau/D) (mapped to log D, log1p(|S|/D) and
the log of the non-kinetic part of tau/D); the FP64 root finder still
converges to its own tolerance from whatever start it is given, so the network
only changes how many iterations that takes.

Everything numeric runs in the HIP library: training is FP32 forward/backward
kernels with Adam (``gammahip_mlp_train``), inference runs the 64x64 hidden
layer as BF16 MFMA with FP32 accumulation (``gammahip_mlp_infer``), and the
FP64 solve from a given start is ``gammahip_c2p_batch``.  This module holds the
host side: parameter layout, initialisation, the feature normalisation (fixed
from the training data before training) and the state generators.

The network has no adiabatic-index input: a parameter set belongs to the
adiabatic index of the states it was trained on.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip as H

NH = 64
OW1 = 0
OB1 = OW1 + NH * 3
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
ONORM = OB3 + 1
NPARAM = ONORM + 8


def _lib():
    L = H.load_library()
    if L.gammahip_mlp_nparams() != NPARAM:
        raise RuntimeError("parameter layout of the library and gamma_amd.surrogate differ")
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def raw_features(states: np.ndarray) -> np.ndarray:
    """(log D, log1p(q), log e) of states [n][>=3] = (D, |S|, tau, ...), with
    q = |S|/D, t = tau/D and e = t + 1 - sqrt(1 + q^2) (the non-kinetic part
    of t, floored at 1e-16 (1 + |t|)): for a cold gas e ~ (k-1)(p/rho)/W."""
    D = np.maximum(states[:, 0], 1e-300)
    q = states[:, 1] / D
    t = states[:, 2] / D
    e = t + 1.0 - np.sqrt(1.0 + q * q)
    fl = 1e-16 * (1.0 + np.abs(t))
    e = np.where(e > fl, e, fl)
    return np.stack([np.log(D), np.log1p(q), np.log(e)], axis=1)


def init_params(states: np.ndarray, seed: int = 1) -> np.ndarray:
    """He-uniform weights, zero biases, and the fixed normalisation of the
    features and of the target log(p/D) taken from ``states`` [n][4]."""
    rng = np.random.default_rng(seed)
    w = np.zeros(NPARAM, dtype=np.float64)
    w[OW1:OB1] = rng.uniform(-1, 1, NH * 3) * np.sqrt(6.0 / 3)
    w[OW2:OB2] = rng.uniform(-1, 1, NH * NH) * np.sqrt(6.0 / NH)
    w[OW3:OB3] = rng.uniform(-1, 1, NH) * np.sqrt(3.0 / NH)
    f = raw_features(states)
    lpd = np.log(states[:, 3] / np.maximum(states[:, 0], 1e-300))
    w[ONORM:ONORM + 3] = f.mean(axis=0)
    w[ONORM + 3:ONORM + 6] = 1.0 / np.maximum(f.std(axis=0), 1e-12)
    w[ONORM + 6] = lpd.mean()
    w[ONORM + 7] = 1.0 / max(lpd.std(), 1e-12)
    return w.astype(np.float32)


def prim_to_states(rho, p, u, gamma: float) -> np.ndarray:
    """Cartesian conserved states [n][4] = (D, |S|, tau, p) of primitive
    (rho, p, |u|) with u the spatial part of the four-velocity
    (FluidState::prim2cons, src/Hydro/rel_cart.cpp:22-48)."""
    rho, p, u = (np.asarray(a, dtype=np.float64).ravel() for a in (rho, p, u))
    lf = np.sqrt(1.0 + u * u)
    h = 1.0 + p * gamma / (gamma - 1.0) / rho
    D = rho * lf
    S = rho * h * lf * lf * (u / lf)
    tau = rho * h * lf * lf - p - D
    return np.stack([D, np.abs(S), tau, p], axis=1)


def states_from_grid(grid, gamma: float, max_states: int = 1 << 20, seed: int = 0) -> np.ndarray:
    """Conserved states [n][4] of the active cells of a running grid (a random
    subset of at most ``max_states``): the states the solver itself produces."""
    rho = np.concatenate(grid.active(0))
    p = np.concatenate(grid.active(1))
    u = np.abs(np.concatenate(grid.active(2)))
    if grid.rc.nd == 2:
        u = np.hypot(u, np.concatenate(grid.active(3)))
    st = prim_to_states(rho, p, u, gamma)
    if st.shape[0] > max_states:
        st = st[np.random.default_rng(seed).choice(st.shape[0], max_states, replace=False)]
    return st


def synthetic_states(n: int, gamma: float, seed: int = 7, lf=(1.0, 1e3), p_over_rho=(1e-10, 1e2),
                     rho=(1e-3, 1e3)) -> np.ndarray:
    """States with Lorentz factor, p/rho and rho log-uniform in the given ranges."""
    rng = np.random.default_rng(seed)
    lu = lambda a, b: np.exp(rng.uniform(np.log(a), np.log(b), n))
    L = lu(*lf)
    r = lu(*rho)
    pr = lu(*p_over_rho) * r
    u = np.sqrt(np.maximum(L * L - 1.0, 0.0))
    return prim_to_states(r, pr, u, gamma)


def train(states: np.ndarray, steps: int = 2000, batch: int = 4096, lr: float = 2e-3, seed: int = 11,
          params: np.ndarray | None = None):
    """Adam on the mean squared error of the normalised log(p/D); returns
    (params, per-step loss).  Runs on the GPU."""
    st = np.ascontiguousarray(states[:, :4], dtype=np.float64)
    w = np.ascontiguousarray(init_params(st, seed) if params is None else params, dtype=np.float32).copy()
    loss = np.zeros(steps, dtype=np.float32)
    r = _lib().gammahip_mlp_train(_p(st), st.shape[0], _p(w), steps, batch, lr, seed, _p(loss))
    if r != 0:
        raise RuntimeError("gammahip_mlp_train failed")
    return w, loss


def gradient(params: np.ndarray, states: np.ndarray):
    """Gradient of the mean squared error over all ``states`` (one batch) and the loss."""
    st = np.ascontiguousarray(states[:, :4], dtype=np.float64)
    w = np.ascontiguousarray(params, dtype=np.float32)
    g = np.zeros(ONORM, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    if _lib().gammahip_mlp_grad(_p(w), _p(st), st.shape[0], _p(g), _p(loss)) != 0:
        raise RuntimeError("gammahip_mlp_grad failed")
    return g, float(loss[0])


def infer(params: np.ndarray, states: np.ndarray, mfma: bool = True) -> np.ndarray:
    """Pressure guesses of states [n][>=3]; BF16 MFMA hidden layer or FP32."""
    st = np.ascontiguousarray(states[:, :3], dtype=np.float64)
    w = np.ascontiguousarray(params, dtype=np.float32)
    out = np.zeros(st.shape[0])
    if _lib().gammahip_mlp_infer(_p(w), _p(st), st.shape[0], _p(out), int(mfma)) != 0:
        raise RuntimeError("gammahip_mlp_infer failed")
    return out


def c2p_batch(states: np.ndarray, guess: np.ndarray, gamma: float, guess2: np.ndarray | None = None,
              solver: int = H.C2P_NEWTON, maxit: int = 100, lib_path: str = H.LIB_PATH) -> dict:
    """FP64 cons2prim of states [n][>=3] from the given starting pressures
    (the kernels' cons2prim; ``lib_path`` may name the CPU build of the same
    bodies)."""
    st = np.ascontiguousarray(states[:, :3], dtype=np.float64)
    n = st.shape[0]
    g = np.ascontiguousarray(guess, dtype=np.float64)
    g2 = None if guess2 is None else np.ascontiguousarray(guess2, dtype=np.float64)
    p = np.zeros(n)
    it = np.zeros(n, dtype=np.int32)
    fl = np.zeros(n, dtype=np.int32)
    us = np.zeros(n, dtype=np.int32)
    r = H.load_library(lib_path).gammahip_c2p_batch(_p(st), _p(g), None if g2 is None else _p(g2), n, gamma,
                                                    solver, maxit, _p(p), _p(it), _p(fl), _p(us))
    if r != 0:
        raise RuntimeError("gammahip_c2p_batch failed")
    return {"p": p, "iters": it, "fail": fl, "used": us}


def cold_guess(states: np.ndarray, gamma: float) -> np.ndarray:
    """The guess a solver without history starts from: (gamma - 1) tau, the
    root for a static state (upper end for moving ones)."""
    return np.maximum((gamma - 1.0) * states[:, 2], 1e-300)

"""Float64 NumPy restatement of the surrogate network (test infrastructure).

Forward, mean-squared-error loss and its gradient for the parameter layout of
gamma_amd/csrc/mlp.hpp, plus a BF16-rounding emulation of the MFMA inference
path.  Pinned by central finite differences (tests/test_mlp.py).  The network
has no counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

from gamma_amd import surrogate as SG


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return dict(
        W1=w[SG.OW1:SG.OB1].reshape(SG.NH, 3), b1=w[SG.OB1:SG.OW2],
        W2=w[SG.OW2:SG.OB2].reshape(SG.NH, SG.NH), b2=w[SG.OB2:SG.OW3],
        w3=w[SG.OW3:SG.OB3], b3=w[SG.OB3], mean=w[SG.ONORM:SG.ONORM + 3],
        scale=w[SG.ONORM + 3:SG.ONORM + 6], out_mean=w[SG.ONORM + 6], out_scale=w[SG.ONORM + 7])


def features(w, states):
    P = unpack(w)
    return ((SG.raw_features(states) - P["mean"]) * P["scale"]).astype(np.float32).astype(np.float64)


def targets(w, states):
    P = unpack(w)
    lpd = np.log(states[:, 3] / np.maximum(states[:, 0], 1e-300))
    return ((lpd - P["out_mean"]) * P["out_scale"]).astype(np.float32).astype(np.float64)


def forward(w, X):
    P = unpack(w)
    z1 = X @ P["W1"].T + P["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.maximum(z2, 0)
    return h2 @ P["w3"] + P["b3"], (z1, h1, z2, h2)


def loss_grad(w, X, T):
    """Loss mean((y - T)^2) and its gradient in the trained entries [0, ONORM)."""
    P = unpack(w)
    y, (z1, h1, z2, h2) = forward(w, X)
    e = y - T
    n = X.shape[0]
    dy = 2.0 * e / n
    g = np.zeros(SG.ONORM)
    g[SG.OW3:SG.OB3] = h2.T @ dy
    g[SG.OB3] = dy.sum()
    dz2 = dy[:, None] * P["w3"][None, :] * (z2 > 0)
    g[SG.OW2:SG.OB2] = (dz2.T @ h1).ravel()
    g[SG.OB2:SG.OW3] = dz2.sum(axis=0)
    dz1 = (dz2 @ P["W2"]) * (z1 > 0)
    g[SG.OW1:SG.OB1] = (dz1.T @ X).ravel()
    g[SG.OB1:SG.OW2] = dz1.sum(axis=0)
    return float(np.mean(e * e)), g


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), as float64."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_mfma(w, X):
    """Inference with the first layer in FP32, its output and W2 rounded to
    BF16, products accumulated exactly (the MFMA accumulates in FP32)."""
    P = unpack(w)
    X32 = X.astype(np.float32)
    W1 = P["W1"].astype(np.float32)
    b1 = P["b1"].astype(np.float32)
    z1 = b1[None, :] + W1[None, :, 0] * X32[:, 0:1]
    z1 = z1 + W1[None, :, 1] * X32[:, 1:2]
    z1 = z1 + W1[None, :, 2] * X32[:, 2:3]
    h1 = bf16(np.maximum(z1, 0))
    z2 = h1 @ bf16(P["W2"]).T + P["b2"]
    return np.maximum(z2, 0) @ P["w3"] + P["b3"]


def y_of_pressure(w, states, p):
    """Network output recovered from a returned pressure guess."""
    P = unpack(w)
    return (np.log(p / states[:, 0]) - P["out_mean"]) * P["out_scale"]


def c2p_residual(p, D, S, tau, G):
    """Cartesian cons2prim residual f(p) (src/Hydro/rel_cart.cpp:108-123), its
    derivative and its rounding scale: the magnitudes of its terms, those
    holding the Lorentz factor W weighted by 1 + W^2 (W is formed from
    1 - S^2/(E+p)^2, a cancellation that costs a factor W^2 in precision)."""
    E = D + tau
    Ep = E + p
    lf = 1.0 / np.sqrt(1.0 - S * S / (Ep * Ep))
    gg = G / (G - 1.0)
    f = E + p - D * lf - gg * p * lf * lf
    tw = np.abs(D * lf) + np.abs(gg * p * lf * lf)
    scale = np.abs(E) + np.abs(p) + (1.0 + lf * lf) * tw
    dlf = -lf ** 3 * S * S / Ep ** 3
    df = 1.0 - D * dlf - gg * (lf * lf + 2.0 * p * lf * dlf)
    return f, df, scale

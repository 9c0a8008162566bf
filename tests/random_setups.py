"""Randomised setups for parity runs (test infrastructure).

Each seed draws a small run the benchmark setups do not cover: dimension
and geometry, track count (2D) and lengths, irregular cell widths, two arbitrary
states (density, pressure, velocity in any direction, tracer) either side of
a discontinuity at a random position, the adiabatic index, and the
compile-time and hook choices (RK3 or Euler, PLM or piecewise constant,
shock detection, synchrotron particles on spherical grids, static or free
boundary faces, the aspect-ratio rule or none, cells from flat to tall).  States are kept moderate (density and pressure jumps within two
decades, speeds below 0.6) so that every draw runs without solver failures;
the runs are for comparing the backend with the oracle, not for physics.
"""
from __future__ import annotations

import numpy as np

from gamma_amd import setups as S


def _widths(rng, n, length):
    w = rng.uniform(0.5, 1.5, n)
    return w * (length / w.sum())


def _states(rng, nd, strong=False):
    """Two primitive states (rho, p, v1, v2, tracer)."""
    out = []
    for side in range(2):
        rho = 10. ** rng.uniform(-2., 2.) if strong else 10. ** rng.uniform(-1., 1.)
        p = 10. ** rng.uniform(-6., 1.) if strong else 10. ** rng.uniform(-2., 1.)
        # strong: Lorentz factors up to 20
        speed = np.sqrt(1. - 1. / rng.uniform(1., 20.) ** 2) if strong else rng.uniform(0., 0.6)
        ang = rng.uniform(0., 2. * np.pi)
        v1 = speed * np.cos(ang)
        v2 = speed * np.sin(ang) if nd == 2 else 0.
        out.append((rho, p, v1, v2, 1. + side))
    return out


def random_setup(seed: int, strong: bool = False) -> S.Setup:
    rng = np.random.default_rng(seed)
    nd = 2 if rng.uniform() < 0.75 else 1
    geom = S.GEOM_SPH if rng.uniform() < 0.4 else S.GEOM_CART
    ntr = int(rng.integers(3, 9)) if nd == 2 else 1
    lens = rng.integers(12, 49, ntr).astype(np.int64)
    nmv = int(lens.max())
    gamma = 4. / 3. if rng.uniform() < 0.5 else 5. / 3.
    x0 = 1. if geom == S.GEOM_SPH else 0.
    length = 1.
    regrid = rng.uniform() < 0.6
    kin = S.KIN_FIXED if rng.uniform() < 0.5 else S.KIN_NONE
    sync = int(geom == S.GEOM_SPH and rng.uniform() < 0.4)
    na, nmp, nme = S.normalized_constants(S.MP, S.C_LIGHT, S.C_LIGHT)
    cfg = S.RunConfig(nd=nd, geom=geom, shock=int(sync or rng.uniform() < 0.5), sync=sync,
                      n_alpha=na, n_mp=nmp, n_me=nme, nmax=(nmv * 3) // 2,
                      ncell_mv=nmv, ncell_f1=ntr, gamma=gamma, kin_mode=kin, v_in=0., v_out=0.,
                      rg_mode=S.RG_AR2D if regrid else S.RG_NONE, rg=(3., 0.2, 0., 0., 0., 0.),
                      timestep=S.TS_EULER if rng.uniform() < 0.2 else S.TS_RK3,
                      recon=S.RECON_PC if rng.uniform() < 0.2 else S.RECON_PLM)
    if nd == 2:
        # cells from flat to tall: aspect ratios either side of the rule's bounds
        aspect = 10. ** rng.uniform(-0.7, 0.7)
        dyc = aspect * length / lens.mean() / (1.5 if geom == S.GEOM_SPH else 1.)
        y0 = 0.6 if geom == S.GEOM_SPH else 0.
        y = y0 + (np.arange(ntr) + 0.5) * dyc
        dy = np.full(ntr, dyc)
    else:
        y, dy = np.zeros(ntr), np.ones(ntr)
    x = np.zeros((ntr, nmv))
    dx = np.ones((ntr, nmv))
    prim = np.zeros((cfg.nq, ntr, nmv))
    (rl, pl, ul, wl, tl), (rr, pr, ur, wr, tr) = _states(rng, nd, strong)
    xd = x0 + rng.uniform(0.3, 0.7) * length
    for a in range(ntr):
        m = int(lens[a])
        w = _widths(rng, m, length)
        faces = x0 + np.concatenate([[0.], np.cumsum(w)])
        xa = 0.5 * (faces[:-1] + faces[1:])
        x[a, :m], dx[a, :m] = xa, w
        x[a, m:] = xa[-1]
        left = xa <= xd
        prim[S.RHO, a, :m] = np.where(left, rl, rr)
        prim[S.PPP, a, :m] = np.where(left, pl, pr)
        prim[S.UU1, a, :m] = np.where(left, ul, ur)
        if nd == 2:
            prim[S.UU2, a, :m] = np.where(left, wl, wr)
        prim[S.TR1, a, :m] = np.where(left, tl, tr)
        prim[S.RHO, a, m:] = 1.
        prim[S.PPP, a, m:] = 1.
    return S.Setup(f"random_{seed}", cfg, 0., x, dx, y, dy, prim,
                   meta=dict(seed=seed), ncells=lens if nd == 2 else None)

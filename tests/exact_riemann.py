"""Exact solution of the 1D special-relativistic Riemann problem (ideal gas, no
tangential velocity), used to pin the oracle's whole scheme.

Method: Marti & Mueller, J. Fluid Mech. 258 (1994) 317, as summarised in
Living Rev. Relativ. 6 (2003) 7, section 4.  The star pressure p* solves
v*_L(p*) = v*_R(p*) where each side is a rarefaction (p* < p_a: Riemann
invariant atanh(v) +- F(c_s) with F(c) = 2/sqrt(g-1) atanh(c/sqrt(g-1)) along the
isentrope) or a shock (p* > p_a: Taub adiabat, mass flux j, shock speed V_s).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.optimize import brentq


def _h(rho, p, g):
    return 1. + g / (g - 1.) * p / rho


def _cs(rho, p, g):
    return math.sqrt(g * p / (rho * _h(rho, p, g)))


def _F(c, g):
    a = math.sqrt(g - 1.)
    return 2. / a * math.atanh(c / a)


def _rarefaction_v(pa, rhoa, va, p, g, sign):
    """Velocity behind a rarefaction from state a to pressure p.
    sign = -1: left-facing wave (left state), +1: right-facing."""
    rho = rhoa * (p / pa) ** (1. / g)
    ca, c = _cs(rhoa, pa, g), _cs(rho, p, g)
    # left wave keeps atanh(v) + F(c) fixed, right wave atanh(v) - F(c)
    return math.tanh(math.atanh(va) - sign * (_F(ca, g) - _F(c, g))), rho


def _shock(pa, rhoa, va, p, g, sign):
    """State behind a shock from state a to pressure p > pa: (v, rho, V_shock)."""
    ha = _h(rhoa, pa, g)
    A = (g - 1.) * (p - pa) / (g * p)
    a2, a1, a0 = 1. - A, A, -(ha * ha + ha * (p - pa) / rhoa)
    hb = (-a1 + math.sqrt(a1 * a1 - 4. * a2 * a0)) / (2. * a2)
    rhob = g * p / ((g - 1.) * (hb - 1.))
    j2 = (p - pa) / (ha / rhoa - hb / rhob)
    j = math.sqrt(j2)
    Wa = 1. / math.sqrt(1. - va * va)
    Vs = (rhoa ** 2 * Wa ** 2 * va + sign * j * math.sqrt(j2 + rhoa ** 2 * Wa ** 2 * (1. - va * va))) \
        / (rhoa ** 2 * Wa ** 2 + j2)
    Ws = 1. / math.sqrt(1. - Vs * Vs)
    js = sign * j
    vb = (ha * Wa * va + Ws * (p - pa) / js) / (ha * Wa + (p - pa) * (Ws * va / js + 1. / (rhoa * Wa)))
    return vb, rhob, Vs


def _side(pa, rhoa, va, p, g, sign):
    if p <= pa:
        v, rho = _rarefaction_v(pa, rhoa, va, p, g, sign)
        return v, rho, None
    return _shock(pa, rhoa, va, p, g, sign)


def star(rhoL, pL, vL, rhoR, pR, vR, g):
    """(p*, v*, rho*_L, rho*_R, V_shock_R or None, V_shock_L or None)."""
    def f(p):
        return _side(pL, rhoL, vL, p, g, -1)[0] - _side(pR, rhoR, vR, p, g, +1)[0]
    lo, hi = 1e-12 * min(pL, pR), max(pL, pR)
    while f(hi) > 0:
        hi *= 1.5
    ps = brentq(f, lo, hi, xtol=1e-15, rtol=1e-15, maxiter=500)
    vL_, rL, VL = _side(pL, rhoL, vL, ps, g, -1)
    vR_, rR, VR = _side(pR, rhoR, vR, ps, g, +1)
    return ps, 0.5 * (vL_ + vR_), rL, rR, VR, VL


def sample(x, t, x0, rhoL, pL, vL, rhoR, pR, vR, g):
    """(rho, p, v) at positions x and time t for left state rarefaction /
    right state shock configurations (the standard blast-wave problems)."""
    ps, vs, rLs, rRs, VR, _ = star(rhoL, pL, vL, rhoR, pR, vR, g)
    assert ps < pL and ps > pR, "sampler covers rarefaction-left / shock-right only"
    x = np.asarray(x, dtype=float)
    xi = (x - x0) / t
    rho = np.empty_like(x)
    p = np.empty_like(x)
    v = np.empty_like(x)
    cL = _cs(rhoL, pL, g)
    head = (vL - cL) / (1. - vL * cL)
    cst = _cs(rLs, ps, g)
    tail = (vs - cst) / (1. - vs * cst)
    # inside the fan: self-similar, xi = (v - c)/(1 - v c); solve for p on the isentrope
    for k, z in enumerate(xi):
        if z <= head:
            rho[k], p[k], v[k] = rhoL, pL, vL
        elif z <= tail:
            def fan(pp):
                vv, rr = _rarefaction_v(pL, rhoL, vL, pp, g, -1)
                cc = _cs(rr, pp, g)
                return (vv - cc) / (1. - vv * cc) - z
            pp = brentq(fan, ps * (1 - 1e-12), pL * (1 + 1e-12), xtol=1e-15, rtol=1e-14)
            vv, rr = _rarefaction_v(pL, rhoL, vL, pp, g, -1)
            rho[k], p[k], v[k] = rr, pp, vv
        elif z <= vs:
            rho[k], p[k], v[k] = rLs, ps, vs
        elif z <= VR:
            rho[k], p[k], v[k] = rRs, ps, vs
        else:
            rho[k], p[k], v[k] = rhoR, pR, vR
    return rho, p, v

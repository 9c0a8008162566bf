"""Pins of the oracle's grid-level semantics, each against a fact found
without the oracle: brute-force interval overlap, closed forms on hand-built
tracks, conservation, the order of the time integrator, and the carry-over
rule recomputed from the faces' recorded signal speeds.

Every check takes the path of an oracle build, so that the same check can be
run on builds with a planted mistake (gamma_oracle.c, -DOG_MUT=n): each check
must pass on the oracle and fail on the mistake it is meant to catch
(test_planted_mistakes_are_caught).

Reference semantics pinned here:
  (i)   F1 neighbour lists, Grid::computeNeighbors (2d.cpp:534-606)
  (ii)  transverse gradients, Grid::computeTransGradients (2d.cpp:366-430)
  (iii) Grid::split / Grid::merge (1d.cpp:560-728)
  (iv)  checkCellForRegrid of the setups (2DST.cpp:304-322, BM1D.cpp:277-319)
  (v)   RK3 stage weights (rk3.cpp:27-92)
  (vi)  Cell::update_dt and collect_dt's carry-over (cell.cpp:47-99,
        1d.cpp:779-795, 2d.cpp:954-971)
  (vii) Grid::updateGhosts whole-object copies (1d.cpp:217-246, 2d.cpp:241-321)
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from gamma_amd import setups as S
import oracle_ctypes as oc

RHO, PPP, UU1, UU2, TR1 = 0, 1, 2, 3, 4


# ---------------------------------------------------------------------------
# toy grids
def _track_cells(widths, x0=0.):
    f = x0 + np.concatenate([[0.], np.cumsum(widths)])
    return (f[:-1] + f[1:]) / 2., np.diff(f)


def toy_2d(widths, dy=0.1, field=None, geom=S.GEOM_CART, y0=0.5, rg_mode=S.RG_NONE, rg=(3., 0.2, 0, 0, 0, 0)):
    """Active tracks j = 0.. with the given cell widths (uneven counts), static
    end faces; field(x, y) -> dict of primitive arrays (3-velocities)."""
    nt = len(widths)
    nm = max(len(w) for w in widths)
    cfg = S.RunConfig(nd=2, geom=geom, nmax=nm + 8, ncell_mv=nm, ncell_f1=nt, kin_mode=S.KIN_FIXED,
                      rg_mode=rg_mode, rg=rg)
    x = np.zeros((nt, nm))
    dx = np.ones((nt, nm))
    y = y0 + (np.arange(nt) + 0.5) * dy
    prim = np.zeros((cfg.nq, nt, nm))
    prim[RHO] = 1.
    prim[PPP] = 1.
    for j, w in enumerate(widths):
        n = len(w)
        x[j, :n], dx[j, :n] = _track_cells(w)
        x[j, n:] = x[j, n - 1]
        if field is not None:
            for q, v in field(x[j, :n], np.full(n, y[j])).items():
                prim[q, j, :n] = v
    return S.Setup("toy_2d", cfg, 0., x, dx, y, np.full(nt, dy), prim,
                   ncells=np.array([len(w) for w in widths]))


def toy_1d(widths, field=None, geom=S.GEOM_CART, x0=0., **cfgkw):
    n = len(widths)
    cfg = S.RunConfig(nd=1, geom=geom, nmax=n + 8, ncell_mv=n, kin_mode=S.KIN_FIXED, **cfgkw)
    x, dx = _track_cells(widths, x0)
    prim = np.zeros((cfg.nq, 1, n))
    prim[RHO] = 1.
    prim[PPP] = 1.
    if field is not None:
        for q, v in field(x).items():
            prim[q, 0] = v
    return S.Setup("toy_1d", cfg, 0., x[None], dx[None], np.zeros(1), np.ones(1), prim)


def _random_widths(rng, counts):
    out = []
    for n in counts:
        w = rng.uniform(0.3, 1., n)
        out.append(w / w.sum())
    return out


def _minmod(a, b):
    if abs(a) < abs(b) and a * b > 0:
        return a
    if abs(b) < abs(a) and a * b > 0:
        return b
    return 0.


# ---------------------------------------------------------------------------
# (i) neighbour lists = cells of the adjacent track overlapping the cell
def check_neighbour_lists(lib_path=None, seed=1):
    rng = np.random.default_rng(seed)
    s = toy_2d(_random_widths(rng, [17, 23, 11, 30, 19]))
    o = oc.OracleGrid(s, lib_path=lib_path)
    o.prep()
    fx = o.faces(oc.I_X)
    cx, cdx = o.cells(oc.F_X), o.cells(oc.F_DX)
    nt = o.tracks()["ntrack"]
    checked = multi = 0
    # cells away from the track ends: there the right ghost face sits on the
    # last active face (1d.cpp:241) and rounding at the shared end decides
    # whether zero-width overlaps enter the sweep
    for j in range(1, o.nj - 1):
        for i in range(3, nt[j] - 3):
            xL, xR = cx[j, i] - cdx[j, i] / 2., cx[j, i] + cdx[j, i] / 2.
            lists = o.neighbors(j, i)
            for side, ja in ((0, j - 1), (1, j + 1)):
                f = fx[ja, :nt[ja] - 1]
                # cell k of track ja lies between faces k-1 and k
                ks = [k for k in range(1, nt[ja] - 1) if min(xR, f[k]) - max(xL, f[k - 1]) > 0.]
                tj, lo, cnt = lists[side]
                assert tj == ja, (j, i, side, tj)
                assert list(range(lo, lo + cnt)) == ks, (j, i, side, (lo, cnt), ks)
                checked += 1
                multi += len(ks) > 1
    assert checked > 150 and multi > 30, (checked, multi)


# ---------------------------------------------------------------------------
# (ii) transverse gradients
def check_trans_gradients_closed_form(lib_path=None, seed=2):
    """A field f(y) = a + c y + d y^2 on misaligned tracks: every MV gradient
    is exactly 0 (f is constant along a track), the extrapolated states are
    the cell values, and the slope to a neighbour is the secant
    c + d (y0 + yn), the same for every neighbour on one side.  The minmod
    chain of 2d.cpp:417-426 starts from the area-weighted average, which lies
    strictly between the lower and the upper secants, takes the smaller lower
    secant at the first lower neighbour, and drops to zero at a second lower
    neighbour, whose equal slope is a tie of minmod (1d.cpp:250-258); the
    larger upper secants never change it.  So the gradient is the lower
    secant for a cell with one lower neighbour and zero for a cell with two
    or more, whatever the area weights."""
    rng = np.random.default_rng(seed)
    a, c, d = 1.0, 0.8, 0.5
    fld = lambda x, y: {RHO: a + c * y + d * y * y, PPP: 2. + c * y + d * y * y}
    s = toy_2d(_random_widths(rng, [14, 21, 9, 26, 15, 18]), field=fld)
    o = oc.OracleGrid(s, lib_path=lib_path)
    o.prep()
    gf = [o.cells(oc.F_GRADF10 + q) for q in (RHO, PPP)]
    gm = o.cells(oc.F_GRADMV0 + RHO)
    y = o.cells(oc.F_Y)
    nt = o.tracks()["ntrack"]
    ng = s.cfg.ngst
    single = zero = 0
    for j in range(ng + 1, o.nj - ng - 1):           # tracks whose neighbours are active
        for i in range(3, nt[j] - 3):
            dn, up = o.neighbors(j, i)
            assert gm[j, i] == 0.
            y0, yl = y[j, i], y[j - 1, 0]
            want = c + d * (y0 + yl)
            for g in gf:
                if dn[2] == 1:
                    assert abs(g[j, i] - want) <= 1e-9 * want, (j, i, g[j, i], want)
                else:
                    assert g[j, i] == 0., (j, i, dn, up, g[j, i])
            single += dn[2] == 1
            zero += dn[2] > 1
    assert single > 10 and zero > 10, (single, zero)


def check_trans_gradients_brute_force(lib_path=None, seed=3):
    """A field curved in x and y on misaligned tracks: each cell's gradient
    equals the one formed from its geometric overlaps (not the oracle's
    lists) with the stored states and MV gradients: the states extrapolated
    along MV to the overlap's midpoint, divided by the distance of the
    centroids in y, the minmod chain over the lower then the upper
    neighbours."""
    rng = np.random.default_rng(seed)
    fld = lambda x, y: {RHO: 1. + 0.3 * x + 0.4 * x * x + 0.7 * y + 0.2 * y * y,
                        PPP: 1. + 0.5 * np.sin(3. * x) + y}
    s = toy_2d(_random_widths(rng, [15, 19, 12, 24, 16, 20]), field=fld)
    o = oc.OracleGrid(s, lib_path=lib_path)
    o.prep()
    fx = o.faces(oc.I_X)
    cx, cdx, y = o.cells(oc.F_X), o.cells(oc.F_DX), o.cells(oc.F_Y)
    nt = o.tracks()["ntrack"]
    ng = s.cfg.ngst
    nonzero = 0
    for q in (RHO, PPP):
        p, gm, gf = o.cells(q), o.cells(oc.F_GRADMV0 + q), o.cells(oc.F_GRADF10 + q)
        for j in range(ng + 1, o.nj - ng - 1):
            for i in range(3, nt[j] - 3):
                xL0, xR0 = cx[j, i] - cdx[j, i] / 2., cx[j, i] + cdx[j, i] / 2.
                slopes = []
                for ja in (j - 1, j + 1):
                    f = fx[ja]
                    for k in range(1, nt[ja] - 1):
                        if min(xR0, f[k]) - max(xL0, f[k - 1]) <= 0.:
                            continue
                        xLn, xRn = cx[ja, k] - cdx[ja, k] / 2., cx[ja, k] + cdx[ja, k] / 2.
                        xm = (max(xL0, xLn) + min(xR0, xRn)) / 2.
                        qL = p[j, i] + gm[j, i] * (xm - cx[j, i])
                        qR = p[ja, k] + gm[ja, k] * (xm - cx[ja, k])
                        slopes.append((qR - qL) / (y[ja, k] - y[j, i]))
                g = float(np.mean(slopes))       # any value between the slopes
                for sl in slopes:
                    g = _minmod(g, sl)
                if g == 0.:
                    assert gf[j, i] == 0., (q, j, i, gf[j, i])
                else:
                    assert abs(gf[j, i] - g) <= 1e-12 * abs(g), (q, j, i, gf[j, i], g)
                    nonzero += 1
    assert nonzero > 20, nonzero


# ---------------------------------------------------------------------------
# (iii) split and merge
def _conserved_total(o):
    dV = o.cells(oc.F_DV)
    tr = o.tracks()
    a, b = tr["iLbnd"][0] + 1, tr["iRbnd"][0]
    return np.array([(o.cells(100 + q)[0, a:b] * dV[0, a:b]).sum() for q in range(4)])


def check_merge(lib_path=None, seed=4):
    """Merge conserves the total of U dV (cartesian: the volumes add), and
    cell i absorbs its smaller neighbour, except the first active cell (always
    the right one) and the last (always the left one)."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 1.5, 24)
    w[5], w[7] = 0.4, 1.6           # cell 6: left neighbour smaller
    w[10], w[12] = 1.6, 0.4         # cell 11: right neighbour smaller
    w[0], w[2] = 0.3, 1.7           # cell 1 (first active): left smaller, still right
    w[23], w[21] = 1.7, 0.3         # cell 22 (last active... index 22): right smaller, still left
    w /= w.sum()
    fld = lambda x: {RHO: 1. + 0.5 * np.sin(5. * x), PPP: 1. + 0.3 * x, UU1: 0.2 * np.cos(4. * x)}
    ng = 2
    for ii, left in ((6, True), (11, False), (0, False), (23, True)):
        s = toy_1d(w, field=fld)
        o = oc.OracleGrid(s, lib_path=lib_path)
        tot0 = _conserved_total(o)
        x0, dx0 = o.cells(oc.F_X)[0].copy(), o.cells(oc.F_DX)[0].copy()
        i = ii + ng
        o.merge(0, i)
        tot1 = _conserved_total(o)
        assert np.allclose(tot1, tot0, rtol=1e-12, atol=0.), (ii, tot0, tot1)
        x1, dx1 = o.cells(oc.F_X)[0], o.cells(oc.F_DX)[0]
        v = i - 1 if left else i + 1
        lo = min(x0[i] - dx0[i] / 2., x0[v] - dx0[v] / 2.)
        hi = max(x0[i] + dx0[i] / 2., x0[v] + dx0[v] / 2.)
        k = i - 1 if left else i
        assert abs((x1[k] - dx1[k] / 2.) - lo) < 1e-14 and abs((x1[k] + dx1[k] / 2.) - hi) < 1e-14, \
            (ii, left, x1[k], dx1[k], lo, hi)
        assert o.tracks()["nact"][0] == len(w) - 1


def check_split(lib_path=None):
    """A cartesian split halves the cell around its centre, puts the new face
    at the old centre, and keeps each primitive's mean; the halves follow the
    limited slope: the smaller one-sided secant on a convex monotone profile,
    none at a local maximum."""
    n = 20
    w = np.full(n, 1. / n)
    ng = 2
    for prof, kind in ((lambda x: 1. + x * x, "convex"), (lambda x: 2. - (x - 0.425) ** 2 * 40., "peak")):
        s = toy_1d(w, field=lambda x: {RHO: prof(x), PPP: 1. + 0. * x})
        o = oc.OracleGrid(s, lib_path=lib_path)
        ii = 8
        i = ii + ng
        x0, dx0 = o.cells(oc.F_X)[0].copy(), o.cells(oc.F_DX)[0].copy()
        r0 = o.cells(RHO)[0].copy()
        o.split(0, i)
        x1, dx1, r1 = o.cells(oc.F_X)[0], o.cells(oc.F_DX)[0], o.cells(RHO)[0]
        assert x1[i] == x0[i] - dx0[i] / 4. and x1[i + 1] == x0[i] + dx0[i] / 4.
        assert dx1[i] == dx0[i] / 2. and dx1[i + 1] == dx0[i] / 2.
        assert o.faces(oc.I_X)[0, i] == x0[i]
        assert abs((r1[i] + r1[i + 1]) / 2. - r0[i]) <= 2e-16 * r0[i]
        gl = (r0[i] - r0[i - 1]) / (x0[i] - x0[i - 1])
        gr = (r0[i + 1] - r0[i]) / (x0[i + 1] - x0[i])
        if kind == "convex":
            g = min(gl, gr)
            assert 0 < gl < gr
            assert abs(r1[i + 1] - r1[i] - g * dx0[i] / 2.) <= 1e-13 * r0[i], (r1[i], r1[i + 1], g)
        else:
            assert gl > 0 > gr
            assert r1[i] == r0[i] and r1[i + 1] == r0[i]
        # the cells after the split moved up one slot, unchanged
        assert np.array_equal(r1[i + 2:n + 2 * ng + 1], r0[i + 1:n + 2 * ng])


# ---------------------------------------------------------------------------
# (iv) the setups' regrid rules on hand-built tracks
def check_regrid_rules(lib_path=None):
    # 2DST aspect ratio dx/dy: split above 3, merge below 0.2
    dy = 0.1
    dxs = [0.35, 0.31, 0.29, 0.1, 0.25, 0.019, 0.021, 0.1, 0.1, 0.1]
    s = toy_2d([np.array(dxs)], dy=dy, rg_mode=S.RG_AR2D, rg=(3., 0.2, 0, 0, 0, 0))
    o = oc.OracleGrid(s, lib_path=lib_path)
    j = s.cfg.ngst
    got = [o.regrid_decision(j, k + 2) for k in range(len(dxs))]
    assert got == [2, 2, 0, 0, 0, 1, 0, 0, 0, 0], got
    # BM1D: ar = dr / (r_out dth), times 10 for the nine cells ahead of the
    # shock (the last cell with p > 1.5 eta); split above split*target/W^1.5,
    # merge below merge*target/W^1.5
    n, eta, w0 = 60, 1e-6, 0.01
    wid = np.full(n, w0)
    wid[40] *= 0.1                       # ar 0.1: merge at W = 1 ...
    wid[45] *= 0.1                       # ... but not at W^1.5 = 8 (threshold 0.025)
    wid[50] *= 0.5                       # ar 0.5 at W^1.5 = 8: threshold 0.375 -> split
    r_out = 1. + wid[:-1].sum() + wid[-1] / 2.
    dth = w0 / r_out                     # ar = 1 for the unmodified cells
    rg = (3., 0.2, dth, eta, 1., 0.)
    W = np.ones(n)
    W[45] = W[50] = 4.
    v = np.sqrt(1. - 1. / W ** 2)
    ishock = 20
    pr = np.where(np.arange(n) <= ishock, 1., eta)
    s = toy_1d(wid, field=lambda x: {PPP: pr, UU1: v}, geom=S.GEOM_SPH, x0=1.,
               rg_mode=S.RG_BM1D, rg=rg)
    o = oc.OracleGrid(s, lib_path=lib_path)
    rout = o.cells(oc.F_X)[0, o.tracks()["iRbnd"][0] - 1]
    assert abs(rout * dth - w0) < 1e-12       # the hand-built ar is 1 (r_out from the loaded cells)
    got = [o.regrid_decision(0, k + 2) for k in range(n)]
    want = [0] * n
    for k in range(ishock + 1, ishock + 10):
        want[k] = 2
    want[40] = 1
    want[50] = 2
    assert got == want, [(k, g, w_) for k, (g, w_) in enumerate(zip(got, want)) if g != w_]


# ---------------------------------------------------------------------------
# (v) the time integrator's order at a fixed mesh
def _smooth_wave():
    """A smooth pulse riding a uniform flow v = 0.2.  The flow keeps |v| away
    from 0: cons2prim rebuilds the velocity as W sqrt(1 - 1/W^2)
    (rel_cart.cpp:199-212), which keeps few digits where |v| ~ 1e-6 and would
    add a step-size independent error of ~1e-10 per step there."""
    n = 100
    bump = lambda x: np.exp(-((x - 0.5) / 0.08) ** 2)
    fld = lambda x: {RHO: 1. + 0.3 * bump(x), PPP: 1. + 0.4 * bump(x), UU1: 0.2 + 0.1 * bump(x)}
    return toy_1d(np.full(n, 1. / n), field=fld, recon=S.RECON_PC)


def rk3_errors(lib_path=None, T=0.02, ns=(8, 16, 32), nref=512):
    s = _smooth_wave()

    def run(nsteps):
        o = oc.OracleGrid(s, dt_fixed=T / nsteps, lib_path=lib_path)
        o.run(nsteps)
        assert abs(o.t - T) < 1e-12
        return np.concatenate([o.active(RHO)[0], o.active(oc.F_X)[0]])

    ref = run(nref)
    return [np.abs(run(n) - ref).sum() / np.abs(ref).sum() for n in ns]


def check_rk3_order(lib_path=None):
    """Piecewise-constant states on a fixed smooth problem make the
    semi-discrete system smooth (no limiter switches), so halving the step
    must cut the error of the fully discrete RK3 solution eight times."""
    e = rk3_errors(lib_path)
    orders = [np.log2(e[k] / e[k + 1]) for k in range(len(e) - 1)]
    assert all(2.7 < p < 3.3 for p in orders), (e, orders)


# ---------------------------------------------------------------------------
# (vi) dt from the recorded signal speeds under the carry-over rule
def _candidates(v, lL, lR, dl, nt, lo_off):
    """Cell::update_dt of cells lo_off .. nt-1-lo_off of one track."""
    out = []
    for i in range(lo_off, nt - lo_off):
        a = dl[i]
        dtL = a / (lR[i - 1] - v[i]) if lR[i - 1] > v[i] else 1e15
        dtR = a / (v[i - 1] - lL[i]) if lL[i] < v[i - 1] else 1e15
        out.append(min(dtL, dtR))
    return out


def check_dt_carry_over(lib_path=None, nd=1):
    """collect_dt's minimum after a step's first prepForUpdate sees that
    stage's candidates on cells 1 .. ntrack-2 and, carried over, the
    previous step's stage-2 and stage-3 candidates of the cells updateGhosts
    did not overwrite since (2 .. ntrack-3); dt = CFL min."""
    if nd == 1:
        s = S.shock_tube_1d(160)
        tracks = [0]
    else:
        s = S.riemann_2d(96, 8, regrid=False)
        tracks = list(range(s.cfg.ngst, 8 + s.cfg.ngst))
    o = oc.OracleGrid(s, lib_path=lib_path)
    o.record_stages()
    o.run(3)
    differs = 0
    for step in range(12):
        prev = [o.stage(k, f) for k in (2, 3) for f in range(4)]
        o.run(1)
        cur = [o.stage(1, f) for f in range(4)]
        nt = o.tracks()["ntrack"]
        carried, fresh = [], []
        for j in tracks:
            for k in range(2):
                v, lL, lR, dl = (a[j] for a in prev[4 * k:4 * k + 4])
                carried += _candidates(v, lL, lR, dl, nt[j], 2)
            v, lL, lR, dl = (a[j] for a in cur)
            fresh += _candidates(v, lL, lR, dl, nt[j], 1)
        want = s.cfg.cfl * min(min(carried), min(fresh))
        assert o.dt == want, (step, o.dt, want)
        differs += want != s.cfg.cfl * min(fresh)
    assert differs > 0, "the carried-over candidates never set dt: the check has no power here"


# ---------------------------------------------------------------------------
# (vii) ghosts
GHOST_FIELDS = list(range(6)) + [oc.F_DX, oc.F_DTLOC, oc.F_SD, oc.F_PSPEC, oc.F_SHOCKED] + \
    [oc.F_GRADMV0 + q for q in range(6)] + [oc.F_GRADF10 + q for q in range(6)]


def check_ghosts(lib_path=None):
    """After updateGhosts every ghost cell is its track's edge cell, field by
    field (primitives, widths, dt_loc, shock values, gradients, neighbour
    lists), shifted by whole cell widths; in 2D the ghost tracks are the edge
    tracks shifted by whole track widths."""
    for s in (S.shock_tube_1d(100, shock=1), S.riemann_2d(48, 6)):
        o = oc.OracleGrid(s, lib_path=lib_path)
        o.run(3)
        o.update_ghosts()
        tr = o.tracks()
        cells = {f: o.cells(f) for f in GHOST_FIELDS + [oc.F_X, oc.F_Y]}
        ng = s.cfg.ngst
        js = range(o.nj)
        for j in js:
            iL, iR, nt = tr["iLbnd"][j], tr["iRbnd"][j], tr["ntrack"][j]
            for i, e, k in [(i, iL + 1, iL - i + 1) for i in range(iL + 1)] + \
                           [(i, iR - 1, -(i - iR + 1)) for i in range(iR, nt)]:
                for f in GHOST_FIELDS:
                    assert cells[f][j, i] == cells[f][j, e], (s.name, j, i, f)
                assert cells[oc.F_X][j, i] == cells[oc.F_X][j, e] - k * cells[oc.F_DX][j, e], (j, i)
                assert o.neighbors(j, i) == o.neighbors(j, e), (j, i)
        if s.cfg.nd == 2:
            dyc = s.dy[0]
            for j, e in [(j, ng) for j in range(ng)] + [(j, o.nj - ng - 1) for j in range(o.nj - ng, o.nj)]:
                assert tr["ntrack"][j] == tr["ntrack"][e]
                n = tr["ntrack"][e]
                for f in GHOST_FIELDS + [oc.F_X]:
                    assert np.array_equal(cells[f][j, :n], cells[f][e, :n]), (j, f)
                assert np.allclose(cells[oc.F_Y][j, :n], cells[oc.F_Y][e, :n] + (j - e) * dyc, rtol=0, atol=1e-15)


# ---------------------------------------------------------------------------
def test_neighbour_lists_are_the_overlapping_cells():
    check_neighbour_lists()


def test_transverse_gradient_closed_form():
    check_trans_gradients_closed_form()


def test_transverse_gradient_from_geometric_overlaps():
    check_trans_gradients_brute_force()


def test_merge_conserves_and_takes_the_smaller_neighbour():
    check_merge()


def test_split_halves_keep_the_mean():
    check_split()


def test_regrid_rules_on_hand_built_tracks():
    check_regrid_rules()


def test_rk3_is_third_order_in_time():
    check_rk3_order()


@pytest.mark.parametrize("nd", [1, 2])
def test_dt_from_recorded_signal_speeds_with_carry_over(nd):
    check_dt_carry_over(nd=nd)


def test_ghosts_are_edge_copies():
    check_ghosts()


# ---------------------------------------------------------------------------
# planted mistakes (gamma_oracle.c, OG_MUT)
MUTATIONS = {
    1: ("RK3 stage-2 weights swapped", check_rk3_order),
    10: ("RK3 stage-3 weights swapped", check_rk3_order),
    2: ("dt_loc reset before every stage", lambda p: check_dt_carry_over(p, nd=1)),
    3: ("neighbour sweep without its step back", check_neighbour_lists),
    4: ("ghost cells copied from the second active cell", check_ghosts),
    5: ("merge ignores which neighbour is smaller", check_merge),
    7: ("BM1D refinement window includes the shock cell", check_regrid_rules),
    8: ("split with the unlimited central gradient", check_split),
    9: ("AR2D rule reads dy/dx", check_regrid_rules),
}


@pytest.fixture(scope="session")
def mutant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("oracle_mutants")


def _mutant(d, n):
    out = os.path.join(str(d), f"libgamma_oracle_mut{n}.so")
    if not os.path.exists(out):
        src = os.path.join(oc.ROOT, "oracle", "gamma_oracle.c")
        subprocess.check_call(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off",
                               "-fno-fast-math", f"-DOG_MUT={n}", src, "-o", out, "-lm"])
    return out


@pytest.mark.parametrize("n", sorted(MUTATIONS))
def test_planted_mistakes_are_caught(n, mutant_dir):
    what, check = MUTATIONS[n]
    path = _mutant(mutant_dir, n)
    with pytest.raises(AssertionError):
        check(path)


# ---------------------------------------------------------------- threads
def test_oracle_threads_do_not_change_bits():
    """The oracle's track loops run in parallel (OpenMP) where the reference's
    do (2d.cpp omp loops over tracks); its F1 loop keeps the reference's
    upward walk by computing gradients and faces in passes and summing in the
    walk's order.  One thread and all threads must give the same bits: 2D
    Riemann problem with regrid, 2D spherical jet with shock detection and
    synchrotron, and the 1D blast wave."""
    L = oc.lib()
    nmax = L.og_threads()
    cases = [(S.riemann_2d(96, 12), 25), (S.bm_jet_2d(10, 0.3, 40, 120), 25),
             (S.blandford_mckee_1d(300), 40)]
    L.og_set_par_min(0)                  # parallel loops on these small grids too
    try:
        for s, n in cases:
            runs = []
            for th in (1, max(nmax, 4)):
                L.og_set_threads(th)
                o = oc.OracleGrid(s)
                o.run(n)
                runs.append([o.cells(f) for f in range(s.cfg.nq)] + [o.faces(0)] + [o.tracks()["ntrack"]])
            for a, b in zip(*runs):
                assert np.array_equal(a, b, equal_nan=True)
    finally:
        L.og_set_threads(nmax)
        L.og_set_par_min(65536)

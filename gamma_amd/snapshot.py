"""Snapshot files in the reference's text format (Grid::printCols).

``write_phys`` downloads a grid's state and writes ``phys%010d.out`` with the
header and columns of src/IO/text1d.cpp:135-243 (1D) or
src/IO/text2d.cpp:146-283 (2D, single rank): one row per active cell, in
track order, with the same printf conversions (``%1.15le`` / ``%le`` /
``%d``; Python's ``%.15e`` / ``%e`` / ``%d`` produce the same text).  The
columns follow the reference's conventions:

* ``vx`` (``vy``) is u / Lorentz factor, with the factor formed as
  FluidState::lfac does (src/fluid.cpp:10-23: sqrt(1 + u^2), u = |u|);
* ``D``, ``sx`` (``sy``), ``tau`` are the conserved variables of the stored
  state (cons2prim ends with prim2cons, so the reference's stored cons are
  these);
* ``Sd``: the second tracer slot, which printCols overwrites with isShocked
  (1D) or Sd (2D) when shock detection is on;
* with synchrotron tracers, ``gmin``/``gmax`` pass through
  Cell::radiation_apply_trac2gammae (src/Radiation/radiation_sph.cpp:206-244)
  with the C library's pow, as the reference's host code does; 2D also prints
  ``gmid`` and ``psyn`` as stored.

``read_phys`` loads such a file into a dict of numpy columns (the layout
plotting_scripts.py reads); ``last_snapshot`` and ``setup_from_phys`` resume a
run from one as Simu::reinitialise does.  ``save_exact``/``load_exact`` keep
the exact device state (gammahip_state_save/restore) for a bit-for-bit resume,
which the text format, rounded to the printed digits, cannot give.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import hip as H
from . import setups as S

RHO, PPP, UU1, UU2, TR1 = 0, 1, 2, 3, 4
GMN, GMX, GMD, PSN = 6, 7, 8, 9


def _lfac(u1, u2=None):
    """FluidState::lfac elementwise (IEEE sqrt, the same rounding as libm's)."""
    s = u1 * u1 if u2 is None else u1 * u1 + u2 * u2
    u = np.sqrt(s)
    return np.sqrt(1 + u * u)


def _trac2gammae(g_trac, lf, rho):
    """radiation_trac2gammae applied to gmin/gmax (radiation_sph.cpp:206-244)."""
    out = np.zeros_like(g_trac)
    for k in range(g_trac.size):
        g = g_trac.flat[k] * (lf.flat[k] * rho.flat[k])
        out.flat[k] = lf.flat[k] * math.pow(rho.flat[k], 4. / 3.) / g if g != 0. else 0.
    return out


def columns(grid, t=None, track: int = 0) -> tuple:
    """(header, rows) of printCols for the grid's current state.

    ``grid`` offers cells(field), tracks() and time() as gamma_amd.hip.Grid
    does; ``t`` overrides the time column.  A 1D grid holding an ensemble of
    runs prints the run on ``track``."""
    rc = grid.rc
    nd, nq = rc.nd, rc.nq
    sync = bool(rc.sync)
    tr = grid.tracks()
    tt = grid.time()[0] if t is None else np.full(grid.nj, float(t))
    prim = {q: grid.cells(q) for q in range(nq) if not (nd == 1 and q == UU2)}
    cons = {q: grid.cells(100 + q) for q in (0, 1, 2, 3) if not (nd == 1 and q == 3)}
    x, dx = grid.cells(H.F_X), grid.cells(H.F_DX)
    second = prim[TR1 + 1]
    if rc.shock:
        second = grid.cells(H.F_SHOCKED if nd == 1 else H.F_SD)
    ng = rc.ngst
    js = [track] if nd == 1 else range(ng, grid.nj - ng)
    rows = []
    if nd == 2:
        y, dy = grid.cells(H.F_Y), grid.cells(H.F_DY)
    for j in js:
        nt = int(tr["ntrack"][j])
        sl = slice(ng, nt - ng)
        rho, p = prim[RHO][j, sl], prim[PPP][j, sl]
        u1 = prim[UU1][j, sl]
        u2 = prim[UU2][j, sl] if nd == 2 else None
        lf = _lfac(u1, u2)
        nact = nt - 2 * ng
        xs, dxs = x[j, sl], dx[j, sl]
        if nd == 2:
            # Cell::computedl: r dtheta in spherical geometry, dy in cartesian
            dly = xs * dy[j, sl] if rc.geom == S.GEOM_SPH else dy[j, sl]
        if sync:
            gmin = _trac2gammae(prim[GMN][j, sl], lf, rho)
            gmax = _trac2gammae(prim[GMX][j, sl], lf, rho)
        for m in range(nact):
            i = ng + m
            if nd == 1:
                vals = [xs[m], dxs[m], dxs[m], rho[m], u1[m] / lf[m], p[m], cons[0][j, i], cons[2][j, i],
                        cons[1][j, i], prim[TR1][j, i], second[j, i]]
                if sync:
                    vals += [gmin[m], gmax[m]]
                rows.append("%.15e %d %d " % (tt[j], nact, m) + " ".join("%.15e" % v for v in vals))
            else:
                head = [xs[m], y[j, i], dxs[m], dy[j, i], dxs[m], dly[m]]
                tail = [rho[m], u1[m] / lf[m], u2[m] / lf[m], p[m], cons[0][j, i], cons[2][j, i],
                        cons[3][j, i], cons[1][j, i], prim[TR1][j, i], second[j, i]]
                if sync:
                    tail += [gmin[m], gmax[m], prim[GMD][j, i], prim[PSN][j, i]]
                    fmt_head = ["%.15e"] * 5 + ["%e"]
                else:
                    fmt_head = ["%.15e"] * 6
                s = "%.15e %d %d %d " % (tt[j], nact, j - ng, m)
                s += " ".join(f % v for f, v in zip(fmt_head, head)) + " "
                s += " ".join("%e" % v for v in tail)
                rows.append(s)
    return header_of(nd, sync), rows


def header_of(nd: int, sync: bool) -> str:
    """First line of a phys file (text1d.cpp:168-172, text2d.cpp:195-199)."""
    if nd == 1:
        return "t nact i x dx dlx rho vx p D sx tau trac Sd" + (" gmin gmax" if sync else "")
    return "t nact j i x y dx dy dlx dly rho vx vy p D sx sy tau trac Sd" + (" gmin gmax gmid psyn" if sync else "")


def write_phys(grid, it: int, directory: str, t=None, track: int = 0) -> str:
    """Write phys%010d.out of the grid's current state into directory."""
    header, rows = columns(grid, t, track)
    path = os.path.join(directory, "phys%010d.out" % it)
    with open(path, "w") as f:
        f.write(header + "\n")
        for r in rows:
            f.write(r + "\n")
    return path


def read_phys(path: str) -> dict:
    """Columns of a phys file, by header name."""
    with open(path) as f:
        names = f.readline().split()
    data = np.loadtxt(path, skiprows=1, ndmin=2)
    return {n: data[:, k] for k, n in enumerate(names)}


def last_snapshot(directory: str) -> tuple:
    """(path, it) of the lexicographically last phys file in directory, the one
    Simu::reinitialise resumes from (text1d.cpp:67-99)."""
    files = sorted(f for f in os.listdir(directory) if "phys" in f)
    if not files:
        raise FileNotFoundError(f"no phys files in {directory}")
    name = files[-1]
    digits = "".join(ch for ch in name.split("phys", 1)[1] if ch.isdigit())
    return os.path.join(directory, name), int(digits or 0)


def setup_from_phys(path: str, base: "S.Setup") -> "S.Setup":
    """Setup that resumes a run from a phys file, as Simu::reinitialise does
    (text1d.cpp:103-134, text2d.cpp:108-145): the run's parameters and per-track
    hooks come from ``base`` (loadParams, grid.initialise), each row gives a
    cell's x, dx (and the track's y, dy), rho, v, p and the first tracer
    (Data::toCell); with synchrotron tracers gmin, gmax (and gmid) go back
    through g -> rho^(1/3) / g and psyn is kept; every other primitive starts at
    zero, the track lengths are the rows' nact, and the time is the first
    row's.  The text carries 16 significant digits in 1D and 7 for the 2D
    fluid columns, so the resumed state is the snapshot's, not the run's."""
    rc = base.cfg
    cols = read_phys(path)
    nd = rc.nd
    n = cols["nact"].size
    jcol = cols["j"].astype(int) if nd == 2 else np.zeros(n, dtype=int)
    icol = cols["i"].astype(int)
    ntr = int(jcol.max()) + 1
    nact = np.zeros(ntr, dtype=np.int32)
    for jj, na in zip(jcol, cols["nact"].astype(int)):
        nact[jj] = na
    nm = int(nact.max())
    x = np.zeros((ntr, nm))
    dx = np.ones((ntr, nm))
    prim = np.zeros((rc.nq, ntr, nm))
    x[jcol, icol] = cols["x"]
    dx[jcol, icol] = cols["dx"]
    rho = cols["rho"]
    prim[RHO, jcol, icol] = rho
    prim[PPP, jcol, icol] = cols["p"]
    prim[UU1, jcol, icol] = cols["vx"]
    if nd == 2:
        prim[UU2, jcol, icol] = cols["vy"]
    prim[TR1, jcol, icol] = cols["trac"]
    if rc.sync:
        r13 = np.array([math.pow(r, 1. / 3.) for r in rho])
        with np.errstate(divide="ignore"):
            prim[GMN, jcol, icol] = r13 / cols["gmin"]
            prim[GMX, jcol, icol] = r13 / cols["gmax"]
            if nd == 2:
                prim[GMD, jcol, icol] = r13 / cols["gmid"]
        if nd == 2:
            prim[PSN, jcol, icol] = cols["psyn"]
    y = np.zeros(ntr)
    dy = np.ones(ntr)
    if nd == 2:
        first = np.unique(jcol, return_index=True)[1]
        y[jcol[first]] = cols["y"][first]
        dy[jcol[first]] = cols["dy"][first]
    cfg = rc.replace(ncell_mv=nm, nmax=max(rc.nmax, nm), ncell_f1=ntr if nd == 2 else rc.ncell_f1)
    return S.Setup(name=base.name + "@resume", cfg=cfg, t0=float(cols["t"][0]), x=x, dx=dx, y=y, dy=dy,
                   prim=prim, meta=dict(base.meta, resumed_from=path), ncells=nact,
                   hooks=base.hooks)


def save_exact(grid, path: str) -> None:
    """Write the grid's exact state (gammahip_state_save) to a .npy file."""
    np.save(path, grid.save_state(), allow_pickle=False)


def load_exact(grid, path: str) -> None:
    """Continue ``grid`` (same setup) from a state file of save_exact."""
    grid.restore_state(np.load(path, allow_pickle=False))

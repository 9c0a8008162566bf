"""Snapshots: the text files of Grid::printCols (src/IO/text1d.cpp:135-243,
src/IO/text2d.cpp:146-283), resuming from them as Simu::reinitialise does
(text1d.cpp:67-134, text2d.cpp:72-145), and the exact binary state.

The text must be what the reference would print for the same state: the
header line, one row per active cell, the reference's printf conversions.
The oracle holds the same state as the kernels (bit for bit on these
cases), so the file written from the kernels' grid must equal, byte for
byte, the file written from the oracle's; the conversions themselves are
checked against the C library's printf.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import math
import os

import numpy as np
import pytest

from gamma_amd import hip, setups
from gamma_amd import snapshot as SN
import oracle_ctypes as oc


class _OracleView:
    """The oracle's grid with the time accessor of gamma_amd.hip.Grid."""

    def __init__(self, o):
        self.o, self.rc, self.nj, self.nax = o, o.rc, o.nj, o.nax

    def cells(self, f):
        return self.o.cells(f)

    def tracks(self):
        return self.o.tracks()

    def time(self):
        return np.full(self.nj, self.o.t), None


CASES = {
    "shock_tube_cart": (lambda: setups.shock_tube_1d(200), 40),
    "shock_tube_sph_shock": (lambda: setups.shock_tube_1d(200, xmin=0.2, xmax=1.2, xmid=0.7,
                                                          geom=setups.GEOM_SPH, shock=1), 30),
    "blandford_mckee_sync": (lambda: setups.blandford_mckee_1d(400), 30),
    "riemann_2d": (lambda: setups.riemann_2d(48, 40), 20),
    "bm_jet_2d_sync": (lambda: setups.bm_jet_2d(12, 0.3, 40, 160), 20),
}


def _pair(name, emu_lib):
    mk, n = CASES[name]
    s = mk()
    o = oc.OracleGrid(s)
    o.run(n)
    g = hip.Grid(s, lib_path=emu_lib)
    g.step(n)
    return s, o, g


def test_printf_conversions_match_the_c_library():
    """%.15e / %e / %d of Python produce the C library's %1.15le / %le / %d."""
    libc = C.CDLL(ctypes.util.find_library("c"))
    buf = C.create_string_buffer(128)
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.standard_normal(300) * 10.0 ** rng.integers(-300, 300, 300),
                           [0.0, -0.0, 1.0, 0.1, 1e-310, 5e-324, 1.7976931348623157e308, 2.5e-7, 9.9999995e-5]])
    for v in vals:
        for cf, pf in ((b"%1.15le", "%.15e"), (b"%le", "%e")):
            libc.snprintf(buf, 128, cf, C.c_double(float(v)))
            assert buf.value.decode() == pf % v, (v, cf)


def test_headers_are_the_reference_headers():
    assert SN.header_of(1, False) == "t nact i x dx dlx rho vx p D sx tau trac Sd"
    assert SN.header_of(1, True) == "t nact i x dx dlx rho vx p D sx tau trac Sd gmin gmax"
    assert SN.header_of(2, False) == "t nact j i x y dx dy dlx dly rho vx vy p D sx sy tau trac Sd"
    assert SN.header_of(2, True) == ("t nact j i x y dx dy dlx dly rho vx vy p D sx sy tau trac Sd"
                                     " gmin gmax gmid psyn")


@pytest.mark.parametrize("name", list(CASES))
def test_snapshot_text_equals_the_oracle_snapshot(name, emu_lib, tmp_path):
    s, o, g = _pair(name, emu_lib)
    os.makedirs(tmp_path / "k")
    os.makedirs(tmp_path / "o")
    pg = SN.write_phys(g, 7, str(tmp_path / "k"))
    po = SN.write_phys(_OracleView(o), 7, str(tmp_path / "o"))
    assert os.path.basename(pg) == "phys0000000007.out"
    a, b = open(pg).read(), open(po).read()
    assert a == b
    lines = a.splitlines()
    assert lines[0] == SN.header_of(s.cfg.nd, bool(s.cfg.sync))
    nact = g.tracks()["nact"]
    rows = int(nact.sum()) if s.cfg.nd == 1 else int(nact[s.cfg.ngst:g.nj - s.cfg.ngst].sum())
    assert len(lines) == 1 + rows
    ncol = len(lines[0].split())
    assert all(len(r.split()) == ncol for r in lines[1:])


def test_snapshot_columns_are_the_state(emu_lib, tmp_path):
    """Each column is the quantity its header names, to the printed digits:
    positions, 3-velocity u/W, conserved variables of the state (oracle
    prim2cons), and the reference's overwrite of the second tracer with
    isShocked (1D) or Sd (2D)."""
    for name in ("shock_tube_sph_shock", "riemann_2d"):
        s, o, g = _pair(name, emu_lib)
        d = SN.read_phys(SN.write_phys(g, 1, str(tmp_path)))
        rc = s.cfg
        js = range(g.nj) if rc.nd == 1 else range(rc.ngst, g.nj - rc.ngst)
        cat = lambda f: np.concatenate([o.active(f)[k] for k in range(len(js))])
        rho, p, u1 = cat(0), cat(1), cat(2)
        u2 = cat(3) if rc.nd == 2 else 0.0
        lf = np.sqrt(1 + (u1 * u1 + u2 * u2))
        rtol = 1e-15 if rc.nd == 1 else 5e-7
        np.testing.assert_allclose(d["x"], cat(oc.F_X), rtol=1e-15)
        np.testing.assert_allclose(d["dx"], cat(oc.F_DX), rtol=1e-15)
        np.testing.assert_allclose(d["rho"], rho, rtol=rtol)
        np.testing.assert_allclose(d["p"], p, rtol=rtol)
        np.testing.assert_allclose(d["vx"], u1 / lf, rtol=rtol, atol=1e-300)
        np.testing.assert_allclose(d["D"], cat(100), rtol=rtol)
        np.testing.assert_allclose(d["tau"], cat(101), rtol=rtol)
        np.testing.assert_allclose(d["sx"], cat(102), rtol=rtol, atol=1e-300)
        flag = cat(oc.F_SHOCKED) if rc.nd == 1 else cat(4 + 1)
        if rc.shock:
            np.testing.assert_array_equal(d["Sd"], cat(oc.F_SHOCKED if rc.nd == 1 else oc.F_SD))
        else:
            np.testing.assert_allclose(d["Sd"], flag, rtol=rtol)
        if rc.nd == 2:
            np.testing.assert_allclose(d["dlx"], d["dx"], rtol=0)
            ys = np.concatenate([np.full(int(g.tracks()["nact"][j]), o.cells(oc.F_Y)[j, 0]) for j in js])
            np.testing.assert_allclose(d["y"], ys, rtol=1e-15)
        assert np.all(d["i"] == np.concatenate([np.arange(n) for n in g.tracks()["nact"][list(js)]]))


def test_sync_gamma_columns(emu_lib, tmp_path):
    """gmin/gmax go through radiation_apply_trac2gammae
    (radiation_sph.cpp:206-244): W rho^(4/3) / (g W rho)."""
    s, o, g = _pair("blandford_mckee_sync", emu_lib)
    d = SN.read_phys(SN.write_phys(g, 2, str(tmp_path)))
    rho, u = o.active(0)[0], o.active(2)[0]
    lf = np.sqrt(1 + u * u)
    for col, q in (("gmin", 6), ("gmax", 7)):
        tr = o.active(q)[0]
        want = np.array([lf[k] * math.pow(rho[k], 4. / 3.) / (tr[k] * (lf[k] * rho[k])) if tr[k] != 0 else 0.
                         for k in range(len(tr))])
        np.testing.assert_allclose(d[col], want, rtol=1e-15)


@pytest.mark.parametrize("name", ["shock_tube_cart", "blandford_mckee_sync", "riemann_2d", "bm_jet_2d_sync"])
def test_resume_from_text_snapshot(name, emu_lib, tmp_path):
    """Simu::reinitialise: the last phys file of the directory gives the
    cells, track lengths and time; the kernels and the oracle, resumed from
    the same file, continue bit for bit alike."""
    s, o, g = _pair(name, emu_lib)
    SN.write_phys(g, 3, str(tmp_path))
    SN.write_phys(g, 12, str(tmp_path))
    path, it = SN.last_snapshot(str(tmp_path))
    assert it == 12 and path.endswith("phys0000000012.out")
    r = SN.setup_from_phys(path, s)
    assert r.t0 == float("%.15e" % g.time()[0][0 if s.cfg.nd == 1 else s.cfg.ngst])
    act = g.tracks()["nact"]
    want = act if s.cfg.nd == 1 else act[s.cfg.ngst:g.nj - s.cfg.ngst]
    np.testing.assert_array_equal(r.ncells, want)
    g2 = hip.Grid(r, lib_path=emu_lib)
    o2 = oc.OracleGrid(r)
    # the resumed state is the printed one
    d = SN.read_phys(path)
    rho = np.concatenate(g2.active(0))
    rtol = 1e-15 if s.cfg.nd == 1 else 0
    np.testing.assert_allclose(rho, d["rho"], rtol=rtol)
    g2.step(5)
    o2.run(5)
    for f in (0, 1, 2):
        for a, b in zip(g2.active(f), o2.active(f)):
            assert np.array_equal(a, b), f


@pytest.mark.parametrize("name", ["shock_tube_sph_shock", "blandford_mckee_sync", "riemann_2d", "bm_jet_2d_sync"])
def test_exact_state_resume(name, emu_lib, tmp_path):
    """gammahip_state_save / restore: a run split by a save and a restore into
    a fresh context is the unsplit run, bit for bit."""
    mk, n = CASES[name]
    s = mk()
    a = hip.Grid(s, lib_path=emu_lib)
    a.step(n)
    path = str(tmp_path / "state.npy")
    SN.save_exact(a, path)
    a.step(n)
    b = hip.Grid(s, lib_path=emu_lib)
    SN.load_exact(b, path)
    b.step(n)
    np.testing.assert_array_equal(a.tracks()["ntrack"], b.tracks()["ntrack"])
    np.testing.assert_array_equal(a.time()[0], b.time()[0])
    for f in list(range(s.cfg.nq)) + [hip.F_X, hip.F_DX, 1000]:
        np.testing.assert_array_equal(a.cells(f), b.cells(f))


def test_exact_state_checks_dimensions(emu_lib):
    a = hip.Grid(setups.shock_tube_1d(200), lib_path=emu_lib)
    b = hip.Grid(setups.shock_tube_1d(300), lib_path=emu_lib)
    buf = a.save_state()
    with pytest.raises(RuntimeError, match="size|dimensions"):
        b.restore_state(buf)
    bad = buf.copy()
    bad[:4] = 0
    with pytest.raises(RuntimeError, match="magic"):
        a.restore_state(bad)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["blandford_mckee_sync", "bm_jet_2d_sync"])
def test_gpu_snapshot_and_exact_resume(name, tmp_path):
    """The HIP build writes the oracle's text, and a saved and restored run
    continues as the unsplit one."""
    mk, n = CASES[name]
    s = mk()
    o = oc.OracleGrid(s)
    o.run(n)
    g = hip.Grid(s)
    g.step(n)
    os.makedirs(tmp_path / "k")
    os.makedirs(tmp_path / "o")
    assert open(SN.write_phys(g, 1, str(tmp_path / "k"))).read() == \
        open(SN.write_phys(_OracleView(o), 1, str(tmp_path / "o"))).read()
    buf = g.save_state()
    g.step(n)
    h = hip.Grid(s)
    h.restore_state(buf)
    h.step(n)
    for f in list(range(s.cfg.nq)) + [hip.F_X, hip.F_DX, 1000]:
        np.testing.assert_array_equal(g.cells(f), h.cells(f))

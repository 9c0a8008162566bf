"""The user setup's hooks (gamma_amd/csrc/user_hooks.h as shipped: a
reflecting box) are called where the reference calls userKinematics,
userBoundaries and checkCellForRegrid, on the CPU build of the kernels and
in the oracle alike.  What the shipped functions define is checked on each
side on its own -- not by comparing the two -- so that a hook both sides
skipped, or called in the wrong place, fails here; the bitwise comparison of
the two runs is in test_parity (reflecting_box_1d, reflecting_box_2d).

* a fixed mesh (par[0] = 0, no regrid) keeps every face where it started
  (to rounding: the RK3 stages average a face's positions, 3/4 x + 1/4 x);
* with the mesh moving, the walls (the faces between the ghost and the active
  cells) stay put while the faces between them move;
* after a ghost update the ghost cells hold the edge cell's state with the
  normal four-velocity reversed;
* the split/merge rule acts only when its widths are set;
* the parameters are checked at the boundary;
* a hooks file named at build time (a piston with a body force and a
  density floor, tests/user_hooks_piston.h) replaces the shipped one on both
  sides, the per-cell hooks (userSourceTerms, cons2prim_user) included, on
  the CPU build and on the GPU.
"""
from __future__ import annotations

import dataclasses
import os
import subprocess

import numpy as np
import pytest

from gamma_amd import hip, setups
import oracle_ctypes as oc


def _grids(emu_lib, s):
    return hip.Grid(s, lib_path=emu_lib), oc.OracleGrid(s)


def _face_x(g):
    return g.faces(1000) if isinstance(g, hip.Grid) else g.faces(oc.I_X)


@pytest.mark.parametrize("ny", [0, 8])
def test_fixed_mesh_keeps_every_face(emu_lib, ny):
    s = setups.reflecting_box(64, ny, mesh=0., regrid=False)
    for g in _grids(emu_lib, s):
        nt = g.tracks()["ntrack"]
        x0 = _face_x(g).copy()
        if isinstance(g, hip.Grid):
            g.step(40)
        else:
            g.run(40)
        x1 = _face_x(g)
        assert np.array_equal(g.tracks()["ntrack"], nt)
        for j in range(len(nt)):
            np.testing.assert_allclose(x1[j, :nt[j] - 1], x0[j, :nt[j] - 1], rtol=0, atol=1e-15,
                                       err_msg=f"{type(g).__name__} track {j}")


@pytest.mark.parametrize("ny", [0, 8])
def test_walls_stay_and_the_mesh_moves(emu_lib, ny):
    s = setups.reflecting_box(64, ny, mesh=1., regrid=False)
    ng = s.cfg.ngst
    for g in _grids(emu_lib, s):
        nt = g.tracks()["ntrack"]
        x0 = _face_x(g).copy()
        if isinstance(g, hip.Grid):
            g.step(40)
        else:
            g.run(40)
        x1 = _face_x(g)
        js = range(ng, len(nt) - ng) if ny else range(len(nt))
        for j in js:
            n = nt[j]
            walls = [ng - 1, n - 1 - ng]        # between the ghost and the active cells
            np.testing.assert_allclose(x1[j, walls], x0[j, walls], rtol=0, atol=1e-15,
                                       err_msg=f"{type(g).__name__} track {j}")
            inner = np.arange(ng, n - 1 - ng)
            assert np.max(np.abs(x1[j, inner] - x0[j, inner])) > 1e-3, (type(g).__name__, j)


@pytest.mark.parametrize("ny", [0, 8])
def test_ghosts_reflect_the_edge_cells(emu_lib, ny):
    s = setups.reflecting_box(64, ny, mesh=0.5)
    ng = s.cfg.ngst
    for g in _grids(emu_lib, s):
        if isinstance(g, hip.Grid):
            g.step(30)
            g.updateGhosts()
        else:
            g.run(30)
            g.update_ghosts()
        tr = g.tracks()
        rho, p, u = g.cells(0), g.cells(1), g.cells(2)
        js = range(ng, len(tr["ntrack"]) - ng) if ny else range(len(tr["ntrack"]))
        for j in js:
            iL, iR = tr["iLbnd"][j], tr["iRbnd"][j]
            assert np.any(u[j, iL + 1:iR] != 0.)
            for i, e in [(i, iL + 1) for i in range(iL + 1)] + [(i, iR - 1) for i in range(iR, tr["ntrack"][j])]:
                assert rho[j, i] == rho[j, e] and p[j, i] == p[j, e], (type(g).__name__, j, i)
                assert u[j, i] == -u[j, e], (type(g).__name__, j, i)


def test_regrid_rule_acts_only_with_its_widths(emu_lib):
    on = setups.reflecting_box(200, mesh=1.)
    off = setups.reflecting_box(200, mesh=1., regrid=False)
    for s, acts in ((on, True), (off, False)):
        g, o = _grids(emu_lib, s)
        g.step(150)
        o.run(150)
        for x in (g, o):
            sp, mg = x.regrid_counts()
            assert (sp.sum() > 0 and mg.sum() > 0) == acts and (sp.sum() + mg.sum() > 0) == acts


def test_user_params_checked(emu_lib):
    g = hip.Grid(setups.reflecting_box(32), lib_path=emu_lib)
    with pytest.raises(RuntimeError, match="user parameters"):
        g.set_user_params(np.zeros(17))
    g.set_user_params(np.zeros(16))
    g.set_user_params([])


# ---- a hooks file given at build time (tests/user_hooks_piston.h) ----
PISTON = os.path.join(os.path.dirname(os.path.abspath(__file__)), "user_hooks_piston.h")
PISTON_PAR = (0.05, 0.5, 0.9)          # piston speed, force, density floor


@pytest.fixture(scope="module")
def piston_libs(tmp_path_factory):
    """The CPU build of the kernels and the oracle, both built with the piston
    hooks file, as GAMMAHIP_USER_HOOKS builds them."""
    from gamma_amd import build
    d = str(tmp_path_factory.mktemp("user_hooks"))
    emu = os.path.join(d, "libgammahip_emu_piston.so")
    orc = os.path.join(d, "libgamma_oracle_piston.so")
    flag = f'-D{{}}_USER_HOOKS="{PISTON}"'
    subprocess.check_call(["g++", *build.EMU_FLAGS, flag.format("GH"), build.SRC, "-o", emu])
    subprocess.check_call(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                           flag.format("OG"), os.path.join(oc.ROOT, "oracle", "gamma_oracle.c"), "-o", orc,
                           "-lm"])
    return emu, orc


def _piston(lib, orc, ny, user_cell, nsteps, opts=None):
    """The piston setup run by the library at `lib` (under launch options
    `opts`) and the oracle at `orc`: bit for bit equal; returns (grid, oracle)."""
    s = setups.reflecting_box(64 if ny else 128, ny, mesh=1., regrid=False)
    s = dataclasses.replace(s, cfg=s.cfg.replace(upar=PISTON_PAR, user_cell=user_cell))
    g = hip.Grid(s, lib_path=lib)
    if opts:
        g.set_options(**opts)
    o = oc.OracleGrid(s, lib_path=orc)
    g.step(nsteps)
    o.run(nsteps)
    assert np.array_equal(o.tracks()["ntrack"], g.tracks()["ntrack"])
    for f in range(s.cfg.nq):
        for j, (a, b) in enumerate(zip(g.active(f), o.active(f))):
            assert np.array_equal(a, b), (f, j)
    nt = g.tracks()["ntrack"]
    xg, xo = g.faces(1000), o.faces(oc.I_X)
    ng = s.cfg.ngst
    for j in (range(ng, len(nt) - ng) if ny else range(len(nt))):
        assert np.array_equal(xg[j, :nt[j] - 1], xo[j, :nt[j] - 1]), j
    assert g.time()[0][0] == o.t
    return g, o


def _piston_checks(lib, orc):
    for ny in (0, 8):
        g0, o0 = _piston(lib, orc, ny, 0, 60)
        # the inner walls (faces between ghost and active cells) move with the piston
        w = g0.rc.ngst - 1
        js = range(g0.rc.ngst, g0.nj - g0.rc.ngst) if ny else range(g0.nj)
        x = g0.faces(1000)
        x_init = hip.Grid(g0.setup, lib_path=lib).faces(1000)
        for j in js:
            np.testing.assert_allclose(x[j, w] - x_init[j, w], PISTON_PAR[0] * o0.t, rtol=1e-12)
        # the per-cell hooks act: the density floor holds where it did not
        g3, o3 = _piston(lib, orc, ny, 3, 60)
        rho0 = np.concatenate(g0.active(0))
        rho3 = np.concatenate(g3.active(0))
        assert rho0.min() < PISTON_PAR[2] and rho3.min() >= PISTON_PAR[2]


def test_hooks_file_given_at_build_time(piston_libs):
    """Both sides call the file they were built with, all five hooks (the
    per-cell ones under GAMMAHIP_OPT_USER_CELL): bit for bit equal runs, the
    inner walls moving as the piston file says (0.05 t), the density floor
    holding only with the per-cell hooks on (CPU build)."""
    _piston_checks(*piston_libs)


@pytest.mark.gpu
def test_gpu_hooks_file_given_at_build_time():
    """The same on the GPU: the HIP library built with the piston file
    (tests/libgammahip_piston.so, made by the build) against the oracle built
    with it."""
    here = os.path.dirname(os.path.abspath(__file__))
    lib, orc = (os.path.join(here, n) for n in ("libgammahip_piston.so", "libgamma_oracle_piston.so"))
    for f in (lib, orc):
        assert os.path.exists(f), f"{f} is missing: build with __graft_entry__.build()"
    print("[gpu user hooks] piston hooks file, 1D and 2D, per-cell hooks off and on", flush=True)
    _piston_checks(lib, orc)


@pytest.mark.gpu
@pytest.mark.parametrize("opts", [dict(persist=4, graph=1), dict(block=64, remap=1, fuse=3, f1fuse=1, occ=3),
                                  dict(block=256, remap=0, f1cell=1, fuse=4)],
                         ids=["packed-graph", "fused-capped", "f1cell-per-owner"])
def test_gpu_hooks_file_under_launch_options(opts):
    """The piston file's five hooks inside the packed, persistent, fused and
    graph-replayed kernels: bit for bit the oracle built with it (1D and 2D,
    per-cell hooks on)."""
    here = os.path.dirname(os.path.abspath(__file__))
    lib, orc = (os.path.join(here, n) for n in ("libgammahip_piston.so", "libgamma_oracle_piston.so"))
    for f in (lib, orc):
        assert os.path.exists(f), f"{f} is missing: build with __graft_entry__.build()"
    print(f"[gpu user hooks] piston hooks file under {opts}", flush=True)
    for ny in (0, 8):
        g, _ = _piston(lib, orc, ny, 3, 60, opts)
        assert np.concatenate(g.active(0)).min() >= PISTON_PAR[2]

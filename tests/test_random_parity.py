"""Parity on randomised setups: the backend equals the oracle bit for bit on
runs drawn by tests/random_setups.py (1D and 2D, cartesian and spherical,
uneven tracks of irregular cells, arbitrary states either side of a
discontinuity, RK3 or Euler, PLM or piecewise constant, shock detection,
synchrotron particles, static or free boundary faces, the aspect-ratio rule
splitting or merging), 30 steps each.  Every field of every active cell,
the cells' positions and widths, shock flags, track lengths, regrid counts
and time are compared.  The CPU build runs every seed; the GPU runs the
same seeds with the default launch and a share of them under other launch
options.  A second set draws strong jumps and Lorentz factors up to 20.
"""
from __future__ import annotations

import numpy as np
import pytest

from gamma_amd import hip
import oracle_ctypes as oc
from random_setups import random_setup

SEEDS = list(range(48))
NSTEPS = 30
# Strong draws (density over four decades, pressure over seven, Lorentz factors
# up to 20): the seeds among 0-40 whose runs have no cons2prim failure.  On the
# others the solver finds no bracket within a few steps; the reference then
# stops in GSL's error handler (no sign change at gsl_root_fsolver_set), so
# those runs are outside what it defines.
STRONG_SEEDS = [3, 4, 5, 7, 9, 10, 13, 14, 15, 17, 18, 20, 21, 23, 24, 25, 27, 28, 29, 30, 31, 33,
                34, 36, 37, 40]


def _run_and_compare(seed, lib_path=None, opts=None, strong=False):
    s = random_setup(seed, strong)
    c = s.cfg
    o = oc.OracleGrid(s)
    o.run(NSTEPS)
    g = hip.Grid(s, lib_path=lib_path) if lib_path else hip.Grid(s)
    if opts:
        g.set_options(**opts)
    g.step(NSTEPS)
    assert np.array_equal(o.tracks()["ntrack"], g.tracks()["ntrack"]), "track lengths differ"
    so, mo = o.regrid_counts()
    sg, mg = g.regrid_counts()
    assert np.array_equal(so, sg) and np.array_equal(mo, mg), "regrid counts differ"
    fields = [(q, q) for q in range(c.nq)] + [(hip.F_X, oc.F_X), (hip.F_DX, oc.F_DX)]
    if c.shock:
        fields.append((hip.F_SHOCKED, oc.F_SHOCKED))
    for fg, fo in fields:
        for j, (a, b) in enumerate(zip(g.active(fg), o.active(fo))):
            assert np.array_equal(a, b, equal_nan=True), f"field {fg} differs on active track {j}"
    assert g.time()[0][0] == o.t
    st = g.stats()
    for k in ("c2p_nobracket", "c2p_maxiter", "nan_flux", "negative_dt", "f1_overflow"):
        assert st[k] == 0, (k, st)
    return s


@pytest.mark.parametrize("seed", SEEDS)
def test_emu_random_setup_bitwise(emu_lib, seed):
    _run_and_compare(seed, lib_path=emu_lib)


@pytest.mark.parametrize("seed", STRONG_SEEDS)
def test_emu_strong_random_setup_bitwise(emu_lib, seed):
    _run_and_compare(seed, lib_path=emu_lib, strong=True)


def test_random_setups_cover_the_options():
    """The seeds reach every option the draws offer, and AMR acts."""
    seen = set()
    for seed in SEEDS:
        c = random_setup(seed).cfg
        seen |= {("nd", c.nd), ("geom", c.geom), ("shock", c.shock), ("sync", c.sync), ("ts", c.timestep),
                 ("recon", c.recon), ("kin", c.kin_mode), ("rg", c.rg_mode)}
    want = {("nd", 1), ("nd", 2), ("geom", 0), ("geom", 1), ("shock", 0), ("shock", 1), ("sync", 1),
            ("ts", 0), ("ts", 1), ("recon", 0), ("recon", 1), ("kin", 0), ("kin", 1), ("rg", 0), ("rg", 1)}
    assert want <= seen, want - seen


@pytest.mark.gpu
@pytest.mark.parametrize("seed", SEEDS)
def test_gpu_random_setup_bitwise(seed):
    print(f"[gpu random] seed={seed}", flush=True)
    _run_and_compare(seed)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", STRONG_SEEDS)
def test_gpu_strong_random_setup_bitwise(seed):
    print(f"[gpu random] strong seed={seed}", flush=True)
    _run_and_compare(seed, strong=True)


GPU_OPTS = [dict(persist=4, graph=1), dict(block=64, remap=1, fuse=1, f1fuse=1),
            dict(block=256, remap=0, f1cell=1, occ=3), dict(fuse=6, graph=1)]


@pytest.mark.gpu
@pytest.mark.parametrize("seed", SEEDS[:16])
def test_gpu_random_setup_bitwise_under_launch_options(seed):
    opts = GPU_OPTS[seed % len(GPU_OPTS)]
    print(f"[gpu random] seed={seed} opts={opts}", flush=True)
    _run_and_compare(seed, opts=opts)

"""CIRC_REGRID_ (Grid::applyRegrid, 1d.cpp:526-557, 2d.cpp:650-683; victims
from Grid::targetRegridVictims, 1d.cpp:498-520, 2d.cpp:625-646).

Each split of a cell first merges the track's smallest cell (the first cell
of the smallest Cell::regridVal among cells iLbnd+3 .. iRbnd-3), and each
merge is followed by a split of the track's biggest one, so the track keeps
its length.  A victim equal to the acting cell ends the reference with
exit(10): the oracle returns -10 there and the kernels count the clash
(gammahip_stats.regrid_clash).

The pair semantics are pinned by a model written from the reference's
description alone: under the 2DST rule (aspect ratio dx/dy, 2DST.cpp:304-322)
on Cartesian tracks both the rule and regridVal (cartesian.cpp:58-63: dx)
read nothing but the cell widths, and splits halve a width while merges add
two, so one regrid pass is a function of the widths that numpy can follow
exactly.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from gamma_amd import hip, setups
import oracle_ctypes as oc


def _circ(setup):
    return dataclasses.replace(setup, cfg=setup.cfg.replace(circ_regrid=1))


def model_pass(dx, dy, ngst, split_ar, merge_ar, nax):
    """One Grid::regrid pass with CIRC_REGRID_ over the widths of a Cartesian
    track under the aspect-ratio rule; returns the widths and (splits, merges)."""
    dx = [float(v) for v in dx]

    def victims():
        iL, iR = ngst - 1, len(dx) - ngst
        mn, mx, vs, vb = 1.e20, 0., -1, -1
        for i in range(iL + 3, iR - 2):
            v = dx[i]
            if v < 0:
                continue
            if v < mn:
                mn, vs = v, i
            if v > mx:
                mx, vb = v, i
        return vs, vb

    def split(i):
        dx[i:i + 1] = [dx[i] / 2., dx[i] / 2.]

    def merge(i):
        nact = len(dx) - 2 * ngst
        left = (dx[i - 1] < dx[i + 1] and i != ngst) or i == nact + ngst - 1
        if left:
            dx[i - 1:i + 1] = [dx[i] + dx[i - 1]]
        else:
            dx[i:i + 2] = [dx[i] + dx[i + 1]]

    vs, vb = victims()
    ns = nm = 0
    i = ngst
    while i <= len(dx) - ngst - 1:
        ar = dx[i] / dy
        act = 2 if ar > split_ar else 1 if ar < merge_ar else 0
        acted = False
        if act == 2 and len(dx) < nax - 1:
            assert vs >= 0 and vs != i, "victim clash"
            merge(vs)
            split(i - 1 if i > vs else i)
            acted = True
        if act == 1 and len(dx) > 2 * ngst + 2:
            assert vb >= 0 and vb != i, "victim clash"
            merge(i)
            split(vb - 1 if vb > i else vb)
            acted = True
        if acted:
            ns += 1
            nm += 1
            vs, vb = victims()
        i += 1
    return np.array(dx), ns, nm


def _check_model(g, s, passes):
    """Alternate steps and single regrid passes; each pass must be the model's."""
    rc = s.cfg
    dy = float(s.dy[0])
    total = 0
    for _ in range(passes):
        g.step(3)
        nt = g.tracks()["ntrack"]
        dx0 = g.cells(hip.F_DX)
        s0, m0 = g.regrid_counts()
        g.regrid()
        dx1 = g.cells(hip.F_DX)
        s1, m1 = g.regrid_counts()
        assert np.array_equal(g.tracks()["ntrack"], nt)
        for j in range(rc.ngst, g.nj - rc.ngst):
            want, ns, nm = model_pass(dx0[j, :nt[j]], dy, rc.ngst, rc.rg[0], rc.rg[1], g.nax)
            assert np.array_equal(dx1[j, :nt[j]], want), j
            assert (s1[j] - s0[j], m1[j] - m0[j]) == (ns, nm), j
            total += ns
    assert g.stats()["regrid_clash"] == 0
    return total


def test_model_pass_by_hand():
    """The model itself on a hand-made track (ngst 2, dy 1): cell 4 (width 4)
    asks for a split, the smallest cell among 4..7 is cell 6 (0.5, first of the
    two), which merges with its smaller neighbour 7; cell 4 splits in place."""
    dx = [1., 1., 1., 1., 4., 1., .5, .5, 1., 1., 1., 1.]
    out, ns, nm = model_pass(dx, 1., 2, 3., 0.2, 100)
    assert (ns, nm) == (1, 1)
    assert out.tolist() == [1., 1., 1., 1., 2., 2., 1., 1., 1., 1., 1., 1.]
    # a merge request (width 0.1 at cell 3): merged into its smaller neighbour,
    # then the biggest cell (cell 6, width 3, index shifted by the merge) splits
    dx = [1., 1., 1., .1, 1., 2., 3., 1., 1., 1., 1., 1.]
    out, ns, nm = model_pass(dx, 1., 2, 3.5, 0.2, 100)
    assert (ns, nm) == (1, 1)
    assert out.tolist() == [1., 1., 1., 1.1, 2., 1.5, 1.5, 1., 1., 1., 1., 1.]


def test_emu_circ_pass_is_the_model(emu_lib):
    s = _circ(setups.riemann_2d(64, 16))
    g = hip.Grid(s, lib_path=emu_lib)
    assert _check_model(g, s, 6) > 0


def test_emu_circ_keeps_track_lengths(emu_lib):
    """Every action is a pair: lengths stay, splits equal merges per track,
    on the oracle and on the kernels alike."""
    s = _circ(setups.riemann_2d(64, 16))
    o = oc.OracleGrid(s)
    g = hip.Grid(s, lib_path=emu_lib)
    nt0 = g.tracks()["ntrack"].copy()
    o.run(40)
    g.step(40)
    for gr in (o, g):
        sp, me = gr.regrid_counts()
        assert np.array_equal(sp, me) and sp.sum() > 0
        assert np.array_equal(gr.tracks()["ntrack"], nt0)


def _clash_step(s, emu_lib, nmax_steps):
    """Step oracle and kernels together until the oracle stops with -10."""
    o = oc.OracleGrid(s)
    g = hip.Grid(s, lib_path=emu_lib)
    for it in range(nmax_steps):
        try:
            o.run(1)
        except RuntimeError as e:
            assert "-10" in str(e)
            g.step(1)
            return it, o, g
        g.step(1)
        assert g.stats()["regrid_clash"] == 0, it
        for f in (0, 1, 2):
            for a, b in zip(g.active(f), o.active(f)):
                assert np.array_equal(a, b), (it, f)
    return None, o, g


def test_emu_circ_clash_is_the_reference_exit(emu_lib):
    """Under the BM1D hooks the cell refined next to the shock is also the
    track's smallest: the reference's exit(10).  The kernels count the clash
    on the step where the oracle stops, and agree bit for bit before it."""
    it, o, g = _clash_step(_circ(setups.blandford_mckee_1d(400)), emu_lib, 80)
    assert it is not None
    assert g.stats()["regrid_clash"] > 0


@pytest.mark.gpu
def test_gpu_circ_pass_is_the_model():
    s = _circ(setups.riemann_2d(64, 16))
    g = hip.Grid(s)
    assert _check_model(g, s, 6) > 0


@pytest.mark.gpu
def test_gpu_circ_clash_is_the_reference_exit():
    it, o, g = _clash_step(_circ(setups.blandford_mckee_1d(400)), hip.LIB_PATH, 80)
    assert it is not None
    assert g.stats()["regrid_clash"] > 0

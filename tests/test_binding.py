"""The reference-side binding: Grid methods called one by one, the way
Simu::run (src/simu.cpp:41-78) calls them, with the caller's dt.

* collect_dt returns the reference's value (the minimum before CFL_), so the
  caller's dt = CFL_ * collect_dt() has the device's own bits, and a run of
  Grid method calls equals gammahip_step bit for bit;
* evolve(it, t, dt) follows the caller's dt: a driver that fixes dt is
  followed, and matches the oracle run with the same fixed dt;
* the binding sources (gamma_amd/host/binding) compile against the
  reference's own headers, for NUM_D 1 and 2, and define every Grid method
  the reference's Dimensions and Timestep modules define.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from gamma_amd import hip, setups
import oracle_ctypes as oc

REF_SRC = "/root/reference/src"
BINDING = os.path.join(ROOT, "gamma_amd", "host", "binding")


def _methods_step(g: hip.Grid, dt_of):
    """One iteration of the Simu::run loop body through the Grid methods."""
    g.regrid()
    g.apply(hip.RESET_SHOCK)
    g.updateGhosts()
    g.prepForUpdate()
    g.apply(hip.DETECT_SHOCK)
    m = g.collect_dt_min()
    g.evolve(dt_of(m))
    g.apply(hip.ADVANCE_TIME)
    g.apply(hip.INJECT)


def _same(a, b, fields):
    assert np.array_equal(a.tracks()["ntrack"], b.tracks()["ntrack"])
    for f in fields:
        assert np.array_equal(a.cells(f), b.cells(f), equal_nan=True), f
    np.testing.assert_array_equal(a.time()[0], b.time()[0])


CASES = [("shock_tube_1d", lambda: setups.shock_tube_1d(128), 30),
         ("blandford_mckee_1d", lambda: setups.blandford_mckee_1d(300), 30),
         ("riemann_2d", lambda: setups.riemann_2d(96, 8), 12)]


def _check_methods_equal_step(lib_path, name, mk, n):
    s = mk()
    cfl = s.cfg.cfl
    a = hip.Grid(s, lib_path=lib_path)
    a.step(n)
    b = hip.Grid(s, lib_path=lib_path)
    for _ in range(n):
        _methods_step(b, lambda m: cfl * m)
    _same(a, b, list(range(s.cfg.nq)) + [hip.F_X, hip.F_DX])


def _check_fixed_dt_follows_oracle(lib_path, n=25):
    s = setups.shock_tube_1d(128)
    o = oc.OracleGrid(s)
    o.run(3)
    dt = 0.5 * o.dt                      # below every step's CFL dt
    o = oc.OracleGrid(s, dt_fixed=dt)
    o.run(n)
    g = hip.Grid(s, lib_path=lib_path)
    for _ in range(n):
        _methods_step(g, lambda m: dt)
    for f in (0, 1, 2):
        assert np.array_equal(np.concatenate(g.active(f)), np.concatenate(o.active(f))), f
    assert g.time()[0][0] == o.t
    free = hip.Grid(s, lib_path=lib_path)
    free.step(n)
    assert free.time()[0][0] != g.time()[0][0]      # the caller's dt was used


@pytest.mark.parametrize("name,mk,n", CASES, ids=[c[0] for c in CASES])
def test_emu_methods_with_caller_dt_equal_step(emu_lib, name, mk, n):
    _check_methods_equal_step(emu_lib, name, mk, n)


def test_emu_fixed_dt_follows_oracle(emu_lib):
    _check_fixed_dt_follows_oracle(emu_lib)


def test_emu_caller_dt_rejected_for_ensembles(emu_lib):
    ens, _ = setups.ensemble_1d(4, sizes=(64, 128))
    g = hip.Grid(ens, lib_path=emu_lib)
    with pytest.raises(RuntimeError, match="several runs"):
        g.collect_dt_min()
    with pytest.raises(RuntimeError, match="several runs"):
        g.evolve(1e-3)
    g2 = hip.Grid(setups.shock_tube_1d(64), lib_path=emu_lib)
    with pytest.raises(RuntimeError, match="positive and finite"):
        g2.evolve(-1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("name,mk,n", CASES, ids=[c[0] for c in CASES])
def test_gpu_methods_with_caller_dt_equal_step(name, mk, n):
    print(f"[gpu binding] {name}: Grid methods with the caller's dt vs gammahip_step", flush=True)
    _check_methods_equal_step(hip.LIB_PATH, name, mk, n)


@pytest.mark.gpu
def test_gpu_fixed_dt_follows_oracle():
    print("[gpu binding] fixed dt vs the oracle", flush=True)
    _check_fixed_dt_follows_oracle(hip.LIB_PATH)


# ---------------------------------------------------------------- compile check
def _reference_headers(tmp, num_d):
    """A copy of the reference's headers (outside the repository) with
    environment.h's NUM_D set, as a maintainer's build would configure it."""
    src = os.path.join(tmp, "src")
    os.makedirs(os.path.join(src, "Dimensions"), exist_ok=True)
    os.makedirs(os.path.join(src, "Timestep"), exist_ok=True)
    os.makedirs(os.path.join(src, "Initial", "Tests"), exist_ok=True)
    for f in os.listdir(REF_SRC):
        if f.endswith(".h"):
            shutil.copy(os.path.join(REF_SRC, f), src)
    env = os.path.join(src, "environment.h")
    text = open(env).read()
    text, n = re.subn(r"#define NUM_D\s+\d+", f"#define NUM_D  {num_d}", text)
    assert n == 1
    open(env, "w").write(text)
    return src


def _defined_grid_methods(obj):
    out = subprocess.run(["nm", "-C", "--defined-only", obj], capture_output=True, text=True,
                         check=True).stdout
    return set(re.findall(r"\bGrid::([A-Za-z_0-9]+)\(", out))


def _reference_module_methods(files):
    names = set()
    for f in files:
        names |= set(re.findall(r"^\s*[a-z][a-z ]*\bGrid::([A-Za-z_0-9]+)\(",
                                open(os.path.join(REF_SRC, f)).read(), re.M))
    return names


@pytest.mark.skipif(not os.path.isdir(REF_SRC) or not os.path.exists("/opt/conda/include/mpi.h"),
                    reason="needs the reference's headers and mpi.h")
@pytest.mark.parametrize("num_d", [1, 2])
def test_binding_compiles_against_reference_headers(tmp_path, num_d):
    src = _reference_headers(str(tmp_path), num_d)
    dim = os.path.join(BINDING, "Dimensions", f"{num_d}d_hip.cpp")
    ts = os.path.join(BINDING, "Timestep", "hip.cpp")
    files = [dim, ts] + ([os.path.join(BINDING, "Initial", "2DST_hip.cpp")] if num_d == 2 else [])
    objs = []
    for f in files:
        o = os.path.join(str(tmp_path), os.path.basename(f) + ".o")
        r = subprocess.run(["g++", "-std=c++11", "-fopenmp", "-Wall", "-Wextra", "-Werror=return-type",
                            "-c", f, "-o", o, "-I/opt/conda/include",
                            "-I" + os.path.join(src, "Dimensions"), "-I" + src,
                            "-I" + os.path.join(src, "Initial", "Tests"),
                            "-I" + os.path.join(ROOT, "include"), "-I" + BINDING],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]
        objs.append(o)
    defined = set()
    for o in objs:
        defined |= _defined_grid_methods(o)
    want = _reference_module_methods([f"Dimensions/{num_d}d.cpp", "Timestep/rk3.cpp"])
    assert want, "no Grid methods found in the reference modules"
    missing = sorted(want - defined)
    assert not missing, missing

"""The C++ host side (gamma_amd/host/gammahip_grid.hpp): a C++ program drives
the hot path through the reference's Grid method sequence (Simu::run,
src/simu.cpp:41-78) or through whole-step calls, and its result is the
oracle's, bit for bit.  The CPU test links the CPU build of the kernels; the
GPU test links the HIP library."""
import os
import subprocess

import numpy as np
import pytest

import oracle_ctypes as oc
from gamma_amd import setups

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_cpp(exe, n, steps, mode):
    if not os.path.exists(exe):
        pytest.fail(f"{exe} missing: build with __graft_entry__.build()")
    r = subprocess.run([exe, str(n), str(steps), mode], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    t = float.fromhex(lines[0].split()[1])
    rows = np.array([[float.fromhex(v) for v in l.split()] for l in lines[1:] if l.strip()])
    return t, rows


def check(exe, mode, n=200, steps=100):
    t, rows = run_cpp(exe, n, steps, mode)
    o = oc.OracleGrid(setups.shock_tube_1d(n))
    o.run(steps)
    assert t == o.t
    for col, f in ((1, 0), (2, 1), (3, 2)):
        assert np.array_equal(rows[:, col], o.active(f)[0]), (mode, f)


@pytest.mark.parametrize("mode", ["methods", "step"])
def test_cpp_host_cpu_build(mode):
    check(os.path.join(ROOT, "tests", "shock_tube_emu"), mode)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["methods", "step"])
def test_cpp_host_gpu(mode):
    print(f"[gpu cpp host] mode={mode}", flush=True)
    check(os.path.join(ROOT, "gamma_amd", "host", "shock_tube"), mode)

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


def run_cpp(exe, n, steps, mode, outdir=None):
    if not os.path.exists(exe):
        pytest.fail(f"{exe} missing: build with __graft_entry__.build()")
    args = [exe, str(n), str(steps), mode] + ([outdir] if outdir else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
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


def check_phys(exe, tmp_path, n=200, steps=60):
    from gamma_amd import snapshot as SN

    class View:
        def __init__(self, o):
            self.o, self.rc, self.nj, self.nax = o, o.rc, o.nj, o.nax
            self.cells, self.tracks = o.cells, o.tracks

        def time(self):
            return np.full(self.nj, self.o.t), None

    os.makedirs(tmp_path / "cpp")
    os.makedirs(tmp_path / "py")
    run_cpp(exe, n, steps, "step", str(tmp_path / "cpp"))
    o = oc.OracleGrid(setups.shock_tube_1d(n))
    o.run(steps)
    SN.write_phys(View(o), steps, str(tmp_path / "py"))
    name = "phys%010d.out" % steps
    assert open(tmp_path / "cpp" / name).read() == open(tmp_path / "py" / name).read()


@pytest.mark.parametrize("mode", ["methods", "step", "split"])
def test_cpp_host_cpu_build(mode):
    check(os.path.join(ROOT, "tests", "shock_tube_emu"), mode)


def test_cpp_print_cols_cpu_build(tmp_path):
    check_phys(os.path.join(ROOT, "tests", "shock_tube_emu"), tmp_path)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["methods", "step", "split"])
def test_cpp_host_gpu(mode):
    print(f"[gpu cpp host] mode={mode}", flush=True)
    check(os.path.join(ROOT, "gamma_amd", "host", "shock_tube"), mode)


@pytest.mark.gpu
def test_cpp_print_cols_gpu(tmp_path):
    print("[gpu cpp host] printCols", flush=True)
    check_phys(os.path.join(ROOT, "gamma_amd", "host", "shock_tube"), tmp_path)

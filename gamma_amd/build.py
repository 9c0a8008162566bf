"""Build recipes for the native pieces (run in-tree; outputs are .so files
next to their sources, git-ignored but shipped with the tree).

* ``gamma_amd/libgammahip.so`` -- the product: HIP for gfx950.
* ``tests/libgammahip_emu.so`` -- the same kernel bodies compiled for the CPU
  (GH_EMU); a test tool that checks indexing and arithmetic against the oracle
  without a GPU.  Never loaded by the package.
* ``oracle/libgamma_oracle.so`` -- the CPU oracle (test infrastructure).
* ``tests/libgammahip_piston.so``, ``tests/libgamma_oracle_piston.so`` -- test
  tools: the HIP library and the oracle built with the test hooks file
  ``tests/user_hooks_piston.h``, so that the GPU tests can check a hooks file
  given at build time on both sides.
* ``gamma_amd/host/shock_tube`` (and ``tests/shock_tube_emu``) -- a C++
  caller of the hot path through ``gamma_amd/host/gammahip_grid.hpp``, linked
  against the HIP library (and against the CPU build, for tests).

The user setup's hooks come from ``gamma_amd/csrc/user_hooks.h``, or from the
file the environment variable GAMMAHIP_USER_HOOKS names; the kernels (HIP and
CPU builds) and the oracle are built with the same file.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "gammahip.cpp")
DEPS = [SRC, os.path.join(HERE, "csrc", "hydro.hpp"), os.path.join(HERE, "csrc", "mlp.hpp"),
        os.path.join(HERE, "csrc", "ddpow.hpp"), os.path.join(HERE, "csrc", "ddtables.h"),
        os.path.join(HERE, "csrc", "user_hooks.h"), os.path.join(ROOT, "include", "gammahip.h")]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# -ffp-contract=off: no fused multiply-add, every expression rounds as written
# (the oracle is built the same way); packed FP32 ops disabled as required on gfx950.
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fPIC", "-shared", "-x", "hip"]
EMU_FLAGS = ["-O2", "-std=c++17", "-DGH_EMU", "-x", "c++", "-ffp-contract=off", "-fPIC",
             "-shared"]


def user_hooks() -> str | None:
    """The user hooks file named by GAMMAHIP_USER_HOOKS (absolute), or None."""
    p = os.environ.get("GAMMAHIP_USER_HOOKS")
    if not p:
        return None
    p = os.path.abspath(p)
    if not os.path.isfile(p):
        raise RuntimeError(f"GAMMAHIP_USER_HOOKS={p}: no such file")
    return p


def _hooks_flags() -> list:
    p = user_hooks()
    return [f'-DGH_USER_HOOKS="{p}"'] if p else []


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)


def build_hip(out: str | None = None) -> str:
    out = out or os.path.join(HERE, "libgammahip.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, *_hooks_flags(), SRC, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_emu(out: str | None = None) -> str:
    out = out or os.path.join(ROOT, "tests", "libgammahip_emu.so")
    tmp = out + ".tmp"
    _run(["g++", *EMU_FLAGS, *_hooks_flags(), SRC, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    d = os.path.join(ROOT, "oracle")
    p = user_hooks()
    _run(["make", "-s", "-B", "-C", d] + ([f"USER_HOOKS={p}"] if p else []))
    return os.path.join(d, "libgamma_oracle.so")


PISTON_HOOKS = os.path.join(ROOT, "tests", "user_hooks_piston.h")


def build_hooks_variant() -> list:
    """The HIP library and the oracle built with tests/user_hooks_piston.h."""
    hip_out = os.path.join(ROOT, "tests", "libgammahip_piston.so")
    _run([HIPCC, *HIP_FLAGS, f'-DGH_USER_HOOKS="{PISTON_HOOKS}"', SRC, "-o", hip_out + ".tmp"])
    os.replace(hip_out + ".tmp", hip_out)
    orc_out = os.path.join(ROOT, "tests", "libgamma_oracle_piston.so")
    _run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          f'-DOG_USER_HOOKS="{PISTON_HOOKS}"', os.path.join(ROOT, "oracle", "gamma_oracle.c"),
          "-o", orc_out, "-lm"])
    return [hip_out, orc_out]


def build_host() -> list:
    src = os.path.join(HERE, "host", "shock_tube.cpp")
    outs = []
    for libdir, lib, out in ((HERE, "gammahip", os.path.join(HERE, "host", "shock_tube")),
                             (os.path.join(ROOT, "tests"), "gammahip_emu", os.path.join(ROOT, "tests", "shock_tube_emu"))):
        rpath = os.path.relpath(libdir, os.path.dirname(out))
        _run(["g++", "-O2", "-std=c++17", src, "-L" + libdir, "-l" + lib,
              "-Wl,-rpath,$ORIGIN/" + rpath, "-o", out])
        outs.append(out)
    return outs


def build_all() -> None:
    build_hip()
    build_emu()
    build_oracle()
    build_host()
    build_hooks_variant()


if __name__ == "__main__":
    build_all()

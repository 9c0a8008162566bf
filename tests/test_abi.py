"""The C-ABI library: it loads, exports every entry point include/gammahip.h
declares, rejects bad configurations with an error code and a message, and
(without a GPU) fails cleanly instead of falling back to anything."""
import ctypes as C
import os
import subprocess

import pytest

from gamma_amd import hip, setups


def test_header_symbols_exported():
    names = hip.exported_symbols()
    assert len(names) >= 19
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_library_loads_and_validates_config():
    L = hip.load_library()
    cfg = hip.make_config(setups.shock_tube_1d(64).cfg, 1)
    cfg.nq = 7                                   # not a valid NUM_Q
    h = C.c_void_p()
    assert L.gammahip_create(C.byref(cfg), C.byref(h)) != 0
    assert b"nq" in L.gammahip_last_error(h)
    L.gammahip_destroy(h)


def test_emulation_is_not_the_product():
    """The package only ever loads the HIP build."""
    assert os.path.basename(hip.LIB_PATH) == "libgammahip.so"
    assert "emu" not in open(hip.__file__).read().lower()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_fails_loudly():
    with pytest.raises(RuntimeError, match="no HIP device"):
        hip.Grid(setups.shock_tube_1d(64))

"""The allocation-free inference entry point (gammahip_mlp_infer_noalloc)
against the allocating one (gammahip_mlp_infer), through the C-ABI: same
bits, on two calls in a row, the second with fewer states and other
parameters (the library's scratch array keeps the first call's packed
operands until the second rewrites them).  Device memory for the inputs and
outputs is the caller's, allocated here with the HIP runtime the library
itself is linked against."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from gamma_amd import hip, surrogate as SG

G53 = 5.0 / 3.0
H2D, D2H = 1, 2


def _hip():
    """The HIP runtime of the library's process (the one libgammahip.so loads)."""
    hip.load_library()
    R = C.CDLL("libamdhip64.so.7")
    R.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    R.hipFree.argtypes = [C.c_void_p]
    R.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    R.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    R.hipStreamSynchronize.argtypes = [C.c_void_p]
    R.hipStreamDestroy.argtypes = [C.c_void_p]
    for f in (R.hipMalloc, R.hipFree, R.hipMemcpy, R.hipStreamCreate, R.hipStreamSynchronize,
              R.hipStreamDestroy):
        f.restype = C.c_int
    return R


class _Dev:
    """A device buffer holding a copy of a host array (or room for one)."""

    def __init__(self, R, a: np.ndarray, upload=True):
        self.R, self.nbytes = R, a.nbytes
        self.p = C.c_void_p()
        assert R.hipMalloc(C.byref(self.p), a.nbytes) == 0
        if upload:
            assert R.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, out: np.ndarray):
        assert self.R.hipMemcpy(out.ctypes.data, self.p, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.R.hipFree(self.p)


def _params(seed):
    st = SG.synthetic_states(4096, G53, seed=seed)
    w = SG.init_params(st, seed=seed).astype(np.float64)
    rng = np.random.default_rng(seed)
    w[SG.OB1:SG.OW2] = rng.uniform(-0.1, 0.1, SG.NH)
    w[SG.OB2:SG.OW3] = rng.uniform(-0.1, 0.1, SG.NH)
    return w.astype(np.float32)


@pytest.mark.gpu
def test_mlp_infer_noalloc_equals_allocating_entry():
    L = hip.load_library()
    R = _hip()
    stream = C.c_void_p()
    assert R.hipStreamCreate(C.byref(stream)) == 0
    try:
        for call, (n, seed) in enumerate(((70001, 3), (4097, 8))):
            w = _params(seed)
            st = np.ascontiguousarray(SG.synthetic_states(n, G53, seed=100 + seed)[:, :3])
            for mfma in (1, 0):
                print(f"[no-alloc] call {call + 1}: n = {n}, mfma = {mfma}", flush=True)
                ref = np.zeros(n)
                assert L.gammahip_mlp_infer(w.ctypes.data, st.ctypes.data, n, ref.ctypes.data, mfma) == 0
                dw, ds = _Dev(R, w), _Dev(R, st)
                out = np.full(n, np.nan)
                do = _Dev(R, out)                    # NaN-filled: every value must be written
                rc = L.gammahip_mlp_infer_noalloc(dw.p, ds.p, n, do.p, mfma, stream)
                assert rc == 0
                assert R.hipStreamSynchronize(stream) == 0
                got = do.get(np.empty(n))
                for b in (dw, ds, do):
                    b.free()
                assert np.all(np.isfinite(ref))
                assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), \
                    (call, mfma, int(np.sum(got != ref)))
    finally:
        R.hipStreamDestroy(stream)


@pytest.mark.gpu
def test_mlp_infer_noalloc_rejects_bad_arguments():
    L = hip.load_library()
    R = _hip()
    w, st, out = _Dev(R, _params(3)), _Dev(R, np.zeros((64, 3))), _Dev(R, np.zeros(64))
    try:
        assert L.gammahip_mlp_infer_noalloc(w.p, st.p, 0, out.p, 1, None) != 0
        assert L.gammahip_mlp_infer_noalloc(None, st.p, 64, out.p, 1, None) != 0
        assert L.gammahip_mlp_infer_noalloc(w.p, st.p, 64, out.p, 2, None) != 0
    finally:
        for b in (w, st, out):
            b.free()

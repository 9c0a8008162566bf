"""ctypes binding of the CPU oracle (oracle/libgamma_oracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the reference of the parity
checks / CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "libgamma_oracle.so")

QMAX = 10


class OgConfig(C.Structure):
    _fields_ = [
        ("nd", C.c_int), ("geom", C.c_int), ("nq", C.c_int), ("shock", C.c_int),
        ("sync", C.c_int), ("ngst", C.c_int), ("nmax", C.c_int), ("ncell_mv", C.c_int),
        ("ncell_f1", C.c_int),
        ("gamma", C.c_double), ("cfl", C.c_double), ("vi", C.c_double),
        ("shock_thr", C.c_double), ("gamma_max_init", C.c_double),
        ("Nalpha", C.c_double), ("Nmp", C.c_double), ("Nme", C.c_double),
        ("kin_mode", C.c_int), ("v_in", C.c_double), ("v_out", C.c_double),
        ("bnd_mode", C.c_int), ("bm", C.c_double * 6),
        ("rg_mode", C.c_int), ("rg", C.c_double * 6),
        ("c2p_maxiter", C.c_int), ("timestep", C.c_int), ("recon", C.c_int), ("vpspec", C.c_int),
        ("eos", C.c_int), ("circ_regrid", C.c_int),
        ("c2p_tol", C.c_double), ("c2p_form", C.c_int), ("dt_fixed", C.c_double),
        ("upar", C.c_double * 16), ("user_cell", C.c_int),
    ]


def make_config(rc, c2p_tol=1e-13, c2p_form=0, dt_fixed=0.) -> OgConfig:
    c = OgConfig()
    c.nd, c.geom, c.nq, c.shock, c.sync = rc.nd, rc.geom, rc.nq, rc.shock, rc.sync
    c.ngst, c.nmax, c.ncell_mv, c.ncell_f1 = rc.ngst, rc.nmax, rc.ncell_mv, rc.ncell_f1
    c.gamma, c.cfl, c.vi = rc.gamma, rc.cfl, rc.vi
    c.shock_thr, c.gamma_max_init = rc.shock_thr, rc.gamma_max_init
    c.Nalpha, c.Nmp, c.Nme = rc.n_alpha, rc.n_mp, rc.n_me
    c.kin_mode, c.v_in, c.v_out = rc.kin_mode, rc.v_in, rc.v_out
    c.bnd_mode = rc.bnd_mode
    for k in range(6):
        c.bm[k] = rc.bm[k]
        c.rg[k] = rc.rg[k]
    c.rg_mode = rc.rg_mode
    c.c2p_maxiter = rc.c2p_maxiter
    c.timestep, c.recon, c.vpspec, c.eos = rc.timestep, rc.recon, rc.vpspec, rc.eos
    c.circ_regrid = rc.circ_regrid
    c.c2p_tol = c2p_tol
    c.c2p_form = c2p_form
    c.dt_fixed = dt_fixed
    for k, v in enumerate(rc.upar):
        c.upar[k] = v
    c.user_cell = rc.user_cell
    return c


_libs = {}


def lib(path: str | None = None):
    """The oracle library (default build), or another build of it at `path`
    (the pin tests load builds with planted mistakes)."""
    path = path or LIB
    if path not in _libs:
        if path == LIB and not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(path)
        P, D, I, V = C.c_void_p, C.c_double, C.c_int, None
        cfgp = C.POINTER(OgConfig)
        sig = {
            "og_create": (P, [cfgp]), "og_destroy": (V, [P]),
            "og_load": (I, [P, P, P, P, P, P]), "og_init": (I, [P, D]),
            "og_set_track_cells": (I, [P, P]), "og_set_track_bm": (I, [P, P]),
            "og_pow_batch": (V, [P, P, P, C.c_long]), "og_fn_batch": (V, [I, P, P, C.c_long]),
            "og_run": (I, [P, I]), "og_ntracks": (I, [P]), "og_nax": (I, [P]),
            "og_track_info": (V, [P, P, P, P, P]), "og_time": (D, [P]),
            "og_iter": (C.c_long, [P]), "og_last_dt": (D, [P]),
            "og_get_cells": (V, [P, I, P]), "og_get_faces": (V, [P, I, P]),
            "og_counters": (V, [P, P]), "og_regrid_counts": (V, [P, P, P]),
            "og_prim2cons": (V, [cfgp, P, D, P]), "og_state2flux": (V, [cfgp, P, P, D, P]),
            "og_cons2prim": (I, [cfgp, P, P, D]), "og_c2p_residual": (D, [cfgp, D, D, D, D]),
            "og_face_flux": (V, [cfgp, I, P, P, D, D, D, P, P]),
            "og_compute_sd": (D, [cfgp, P, P, I, I, P]),
            "og_cell_geom": (V, [cfgp, D, D, D, D, P]), "og_face_area": (D, [cfgp, I, D, D, D]),
            "og_sync_source": (D, [cfgp, P, D, D]), "og_inject": (V, [cfgp, P, I, D]),
            "og_prep": (I, [P]), "og_update_ghosts": (I, [P]), "og_get_neighbors": (V, [P, I, I, P]),
            "og_split": (V, [P, I, I]), "og_merge": (V, [P, I, I]), "og_regrid_decision": (I, [P, I, I]),
            "og_record_stages": (I, [P, I]), "og_get_stage": (I, [P, I, I, P]),
            "og_threads": (I, []), "og_set_threads": (V, [I]), "og_set_par_min": (V, [C.c_long]),
        }
        for name, (res, args) in sig.items():
            if not hasattr(L, name):
                continue
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _libs[path] = L
    return _libs[path]


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# field ids (gamma_oracle.h)
F_X, F_DX, F_Y, F_DY, F_DV, F_CEN, F_DTLOC, F_SD, F_SHOCKED, F_PSPEC = range(200, 210)
F_GRADMV0, F_GRADF10 = 300, 400
I_X, I_V, I_LL, I_LR, I_LS, I_DA, I_FLUX0 = 0, 1, 2, 3, 4, 5, 10


class OracleGrid:
    """Grid + Simu::run loop of the oracle for one setup."""

    def __init__(self, setup, c2p_tol=1e-13, c2p_form=0, dt_fixed=0., lib_path=None):
        self.setup = setup
        self.rc = setup.cfg
        self.cfg = make_config(self.rc, c2p_tol, c2p_form, dt_fixed)
        self.L = L = lib(lib_path)
        self.h = L.og_create(C.byref(self.cfg))
        if not self.h:
            raise MemoryError("oracle allocation failed")
        x = np.ascontiguousarray(setup.x, dtype=np.float64)
        dx = np.ascontiguousarray(setup.dx, dtype=np.float64)
        y = np.ascontiguousarray(setup.y, dtype=np.float64)
        dy = np.ascontiguousarray(setup.dy, dtype=np.float64)
        prim = np.ascontiguousarray(setup.prim, dtype=np.float64)
        if setup.ncells is not None:
            nc = np.ascontiguousarray(setup.ncells, dtype=np.int32)
            if self.L.og_set_track_cells(self.h, _p(nc)) != 0:
                raise ValueError("bad per-track cell counts")
        if setup.hooks is not None:
            bm = np.array([list(h.get("bm", self.rc.bm))[:4] for h in setup.hooks], dtype=np.float64)
            self.L.og_set_track_bm(self.h, _p(bm))
        self.L.og_load(self.h, _p(x), _p(dx), _p(y), _p(dy), _p(prim))
        self.L.og_init(self.h, setup.t0)
        self.nj = self.L.og_ntracks(self.h)
        self.nax = self.L.og_nax(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.og_destroy(self.h)
            self.h = None

    def run(self, nsteps: int) -> None:
        e = self.L.og_run(self.h, nsteps)
        if e:
            raise RuntimeError(f"oracle run failed with code {e}")

    @property
    def t(self) -> float:
        return self.L.og_time(self.h)

    @property
    def it(self) -> int:
        return self.L.og_iter(self.h)

    @property
    def dt(self) -> float:
        return self.L.og_last_dt(self.h)

    def tracks(self):
        n = [np.zeros(self.nj, dtype=np.int32) for _ in range(4)]
        self.L.og_track_info(self.h, *[_p(a) for a in n])
        return dict(ntrack=n[0], nact=n[1], iLbnd=n[2], iRbnd=n[3])

    def cells(self, field: int) -> np.ndarray:
        out = np.zeros((self.nj, self.nax))
        self.L.og_get_cells(self.h, field, _p(out))
        return out

    def faces(self, field: int) -> np.ndarray:
        out = np.zeros((self.nj, self.nax - 1))
        self.L.og_get_faces(self.h, field, _p(out))
        return out

    def counters(self) -> dict:
        out = np.zeros(8, dtype=np.int64)
        self.L.og_counters(self.h, _p(out))
        keys = ["c2p_calls", "brent_iters", "no_bracket", "max_iter", "splits", "merges",
                "phi_cap", "tangled"]
        return dict(zip(keys, out.tolist()))

    def regrid_counts(self):
        s = np.zeros(self.nj, dtype=np.int64)
        m = np.zeros(self.nj, dtype=np.int64)
        self.L.og_regrid_counts(self.h, _p(s), _p(m))
        return s, m

    # ---- pieces of a step (pin tests) ----
    def prep(self) -> None:
        """updateGhosts + prepForUpdate, as the step runs them before collect_dt."""
        e = self.L.og_prep(self.h)
        if e:
            raise RuntimeError(f"oracle prep failed with code {e}")

    def update_ghosts(self) -> None:
        self.L.og_update_ghosts(self.h)

    def neighbors(self, j: int, i: int):
        """((j_down, first, count), (j_up, first, count)) of cell (j, i)."""
        out = np.zeros(6, dtype=np.int32)
        self.L.og_get_neighbors(self.h, j, i, _p(out))
        return tuple(out[:3].tolist()), tuple(out[3:].tolist())

    def split(self, j: int, i: int) -> None:
        self.L.og_split(self.h, j, i)

    def merge(self, j: int, i: int) -> None:
        self.L.og_merge(self.h, j, i)

    def regrid_decision(self, j: int, i: int) -> int:
        return self.L.og_regrid_decision(self.h, j, i)

    def record_stages(self, on: bool = True) -> None:
        if self.L.og_record_stages(self.h, int(on)):
            raise MemoryError("oracle record allocation failed")

    def stage(self, s: int, field: int) -> np.ndarray:
        """Faces' v (0), lL (1), lR (2) [nj][nax-1] or cells' dl (3) [nj][nax]
        as the prepForUpdate of stage s of the last step left them."""
        shape = (self.nj, self.nax) if field == 3 else (self.nj, self.nax - 1)
        out = np.zeros(shape)
        if self.L.og_get_stage(self.h, s, field, _p(out)):
            raise RuntimeError("stage recording is off")
        return out

    def active(self, field: int):
        """List over active tracks of the active-cell values of a cell field."""
        tr = self.tracks()
        a = self.cells(field)
        rc = self.rc
        js = range(rc.ngst, self.nj - rc.ngst) if rc.nd == 2 else range(1)
        return [a[j, tr["iLbnd"][j] + 1: tr["iRbnd"][j]] for j in js]


# ---- pointwise helpers ----
def prim2cons(rc, prim, r=1.0):
    cfg = make_config(rc)
    p = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.zeros(rc.nq)
    lib().og_prim2cons(C.byref(cfg), _p(p), r, _p(out))
    return out


def state2flux(rc, prim, cons, r=1.0):
    cfg = make_config(rc)
    p = np.ascontiguousarray(prim, dtype=np.float64)
    c = np.ascontiguousarray(cons, dtype=np.float64)
    out = np.zeros(rc.nd * rc.nq)
    lib().og_state2flux(C.byref(cfg), _p(p), _p(c), r, _p(out))
    return out.reshape(rc.nd, rc.nq)


def cons2prim(rc, cons, prim_guess, r=1.0):
    cfg = make_config(rc)
    p = np.array(prim_guess, dtype=np.float64)
    c = np.array(cons, dtype=np.float64)
    it = lib().og_cons2prim(C.byref(cfg), _p(p), _p(c), r)
    return p, c, it


def c2p_residual(rc, p, D, S, E):
    return lib().og_c2p_residual(C.byref(make_config(rc)), p, D, S, E)


def face_flux(rc, dim, primL, primR, xf=1.0, v=float("nan"), dA=1.0):
    cfg = make_config(rc)
    a = np.ascontiguousarray(primL, dtype=np.float64)
    b = np.ascontiguousarray(primR, dtype=np.float64)
    lam = np.zeros(3)
    fl = np.zeros(rc.nq)
    lib().og_face_flux(C.byref(cfg), dim, _p(a), _p(b), xf, v, dA, _p(lam), _p(fl))
    return lam, fl


def compute_sd(rc, prim1, prim2, dim=0, reverse=0):
    cfg = make_config(rc)
    a = np.ascontiguousarray(prim1, dtype=np.float64)
    b = np.ascontiguousarray(prim2, dtype=np.float64)
    ps = np.zeros(1)
    sd = lib().og_compute_sd(C.byref(cfg), _p(a), _p(b), dim, reverse, _p(ps))
    return sd, ps[0]


def cell_geom(rc, x, dx, y=0.0, dy=0.0):
    out = np.zeros(5)
    lib().og_cell_geom(C.byref(make_config(rc)), x, dx, y, dy, _p(out))
    return out


def face_area(rc, dim, x, y=0.0, dxf=0.0):
    return lib().og_face_area(C.byref(make_config(rc)), dim, x, y, dxf)


def sync_source(rc, prim, dV, dt):
    p = np.ascontiguousarray(prim, dtype=np.float64)
    return lib().og_sync_source(C.byref(make_config(rc)), _p(p), dV, dt)


def inject(rc, prim, shocked, pspec):
    p = np.array(prim, dtype=np.float64)
    lib().og_inject(C.byref(make_config(rc)), _p(p), int(shocked), pspec)
    return p

"""Host side of the HIP backend: ctypes binding of include/gammahip.h and a
Grid / Simu pair mirroring the reference's interface for the hot path.

Reference interface (eliotayache-nv/GAMMA): ``Grid`` methods of src/grid.h:61-112
(``updateGhosts``, ``prepForUpdate``, ``collect_dt``, ``evolve``, ``regrid``,
``apply``, ``prepForRun``) called by ``Simu::run`` (src/simu.cpp:41-78).
The native library is required: importing without a built
``gamma_amd/libgammahip.so`` raises, there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import setups as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgammahip.so")

F_X, F_DX, F_Y, F_DY, F_DV, F_CEN, F_DTLOC, F_SD, F_SHOCKED, F_PSPEC, F_NF1 = range(200, 211)
F_GRADMV0, F_GRADF10 = 300, 400
FACE_X, FACE_V, FACE_LL, FACE_LR, FACE_FLUX0 = 1000, 1001, 1002, 1003, 1010
RESET_SHOCK, DETECT_SHOCK, INJECT, INIT_GAMMAS, ADVANCE_TIME = range(5)
C2P_BRENT, C2P_NEWTON = 0, 1
OPT_BLOCK, OPT_REMAP, OPT_GRAPH, OPT_MLP_THR, OPT_PERSIST, OPT_OCC, OPT_F1FUSE, OPT_SOLVER, OPT_F1CELL, OPT_FUSE, OPT_USER_CELL = range(11)


class Config(C.Structure):
    _fields_ = [
        ("nd", C.c_int), ("nq", C.c_int), ("geom", C.c_int), ("shock", C.c_int),
        ("sync", C.c_int), ("ngst", C.c_int), ("nmax", C.c_int), ("ncell_mv", C.c_int),
        ("ntracks_active", C.c_int),
        ("gamma", C.c_double), ("cfl", C.c_double), ("vi", C.c_double),
        ("shock_thr", C.c_double), ("gamma_max_init", C.c_double),
        ("Nalpha", C.c_double), ("Nmp", C.c_double), ("Nme", C.c_double),
        ("rg_mode", C.c_int), ("rg", C.c_double * 6),
        ("c2p_solver", C.c_int), ("c2p_guess", C.c_int), ("c2p_maxiter", C.c_int),
        ("timestep", C.c_int), ("recon", C.c_int), ("vpspec", C.c_int), ("eos", C.c_int),
        ("circ_regrid", C.c_int),
    ]


class TrackHooks(C.Structure):
    _fields_ = [
        ("geom", C.c_int), ("kin_mode", C.c_int), ("v_in", C.c_double), ("v_out", C.c_double),
        ("bnd_mode", C.c_int), ("bm", C.c_double * 4), ("rg_mode", C.c_int), ("ncell", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [(n, C.c_ulonglong) for n in (
        "c2p_calls", "c2p_iters", "c2p_nobracket", "c2p_maxiter", "c2p_phicap",
        "negative_dt", "nan_flux", "f1_overflow", "splits", "merges", "mlp_used", "regrid_clash")] + [
        ("reserved", C.c_ulonglong * 4)]


_libs: dict = {}


def load_library(path: str = LIB_PATH):
    """Load and type the C-ABI library (the HIP build by default)."""
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (the HIP backend has no fallback)")
    L = C.CDLL(path)
    P, I, D, V = C.c_void_p, C.c_int, C.c_double, None
    sig = {
        "gammahip_create": (I, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
        "gammahip_destroy": (I, [P]), "gammahip_last_error": (C.c_char_p, [P]),
        "gammahip_load": (I, [P, P, P, P, P, P, C.POINTER(TrackHooks), D]),
        "gammahip_step": (I, [P, I]), "gammahip_regrid": (I, [P]),
        "gammahip_update_ghosts": (I, [P]), "gammahip_prep_for_update": (I, [P]),
        "gammahip_collect_dt": (I, [P, P]), "gammahip_evolve": (I, [P]),
        "gammahip_collect_dt_min": (I, [P, P]), "gammahip_evolve_dt": (I, [P, D]),
        "gammahip_apply": (I, [P, I]), "gammahip_download": (I, [P, I, P]),
        "gammahip_track_info": (I, [P, P, P, P, P]), "gammahip_dims": (I, [P, P, P]),
        "gammahip_time": (I, [P, P, P]), "gammahip_stats_get": (I, [P, C.POINTER(Stats)]),
        "gammahip_regrid_counts": (I, [P, P, P]), "gammahip_kernel_times": (I, [P, P, P]),
        "gammahip_sync": (I, [P]), "gammahip_set_option": (I, [P, I, C.c_longlong]),
        "gammahip_set_user_params": (I, [P, P, I]),
        "gammahip_mlp_nparams": (I, []),
        "gammahip_mlp_train": (I, [P, C.c_longlong, P, I, I, C.c_float, C.c_ulonglong, P]),
        "gammahip_mlp_grad": (I, [P, P, I, P, P]),
        "gammahip_mlp_infer": (I, [P, P, C.c_longlong, P, I]),
        "gammahip_mlp_infer_noalloc": (I, [P, P, C.c_longlong, P, I, P]),
        "gammahip_mfma_probe": (I, [P, P, P]),
        "gammahip_c2p_batch": (I, [P, P, P, C.c_longlong, D, I, I, P, P, P, P]),
        "gammahip_mlp_load": (I, [P, P]),
        "gammahip_pow_batch": (I, [P, P, P, C.c_longlong]),
        "gammahip_fn_batch": (I, [I, P, P, C.c_longlong]),
        "gammahip_state_size": (I, [P, C.POINTER(C.c_ulonglong)]),
        "gammahip_state_save": (I, [P, P, C.c_ulonglong]),
        "gammahip_state_restore": (I, [P, P, C.c_ulonglong]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = L
    return L


def exported_symbols():
    """Names the C-ABI header declares (checked against the built library)."""
    hdr = os.path.join(os.path.dirname(HERE), "include", "gammahip.h")
    import re
    return sorted(set(re.findall(r"\b(gammahip_[a-z_]+)\s*\(", open(hdr).read())))


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def make_config(rc: S.RunConfig, ntracks_active: int, solver=C2P_BRENT, guess=0) -> Config:
    c = Config()
    c.nd, c.nq, c.geom, c.shock, c.sync = rc.nd, rc.nq, rc.geom, rc.shock, rc.sync
    c.ngst, c.nmax, c.ncell_mv, c.ntracks_active = rc.ngst, rc.nmax, rc.ncell_mv, ntracks_active
    c.gamma, c.cfl, c.vi, c.shock_thr, c.gamma_max_init = (
        rc.gamma, rc.cfl, rc.vi, rc.shock_thr, rc.gamma_max_init)
    c.Nalpha, c.Nmp, c.Nme = rc.n_alpha, rc.n_mp, rc.n_me
    c.rg_mode = rc.rg_mode
    for k in range(6):
        c.rg[k] = rc.rg[k]
    c.c2p_solver, c.c2p_guess, c.c2p_maxiter = solver, guess, rc.c2p_maxiter
    c.timestep, c.recon, c.vpspec, c.eos = rc.timestep, rc.recon, rc.vpspec, rc.eos
    c.circ_regrid = rc.circ_regrid
    return c


def make_hooks(rc: S.RunConfig, ncell: int = 0, over: dict | None = None) -> TrackHooks:
    """Hooks of one track: the run config's, with per-track overrides."""
    o = dict(geom=rc.geom, kin_mode=rc.kin_mode, v_in=rc.v_in, v_out=rc.v_out,
             bnd_mode=rc.bnd_mode, bm=rc.bm, rg_mode=-1)
    o.update(over or {})
    h = TrackHooks()
    h.geom, h.kin_mode, h.v_in, h.v_out = o["geom"], o["kin_mode"], o["v_in"], o["v_out"]
    h.bnd_mode, h.rg_mode, h.ncell = o["bnd_mode"], o["rg_mode"], int(ncell)
    for k in range(4):
        h.bm[k] = o["bm"][k]
    return h


def setup_hooks(setup: S.Setup) -> list:
    """One hooks entry per active track of a setup."""
    rc = setup.cfg
    na = rc.ncell_f1 if rc.nd == 2 else setup.x.shape[0]
    nc = setup.ncells if setup.ncells is not None else [0] * na
    ov = setup.hooks if setup.hooks is not None else [None] * na
    return [make_hooks(rc, nc[a], ov[a]) for a in range(na)]


class Grid:
    """Device-resident grid; methods named after the reference's Grid."""

    def __init__(self, setup: S.Setup, solver=C2P_BRENT, lib_path: str = LIB_PATH,
                 hooks_per_track=None, geom=None):
        self.L = load_library(lib_path)
        self.setup = setup
        rc = setup.cfg
        self.rc = rc
        na = rc.ncell_f1 if rc.nd == 2 else setup.x.shape[0]
        self.cfg = make_config(rc, na, solver)
        self.solver = solver
        if geom is not None:
            self.cfg.geom = geom
        h = C.c_void_p()
        r = self.L.gammahip_create(C.byref(self.cfg), C.byref(h))
        self.h = h
        if r != 0:
            msg = self.L.gammahip_last_error(h).decode()
            self.L.gammahip_destroy(h)
            self.h = None
            raise RuntimeError("gammahip_create: " + msg)
        if rc.upar:
            self.set_user_params(rc.upar)       # before load: its first ghost update calls the hooks
        if rc.user_cell:
            self._check(self.L.gammahip_set_option(self.h, OPT_USER_CELL, int(rc.user_cell)))
        if hooks_per_track is None:
            hooks_per_track = setup_hooks(setup)
        arr = (TrackHooks * len(hooks_per_track))(*hooks_per_track)
        self._keep = [np.ascontiguousarray(a, dtype=np.float64)
                      for a in (setup.x, setup.dx, setup.y, setup.dy, setup.prim)]
        self._check(self.L.gammahip_load(self.h, *[_p(a) for a in self._keep], arr, setup.t0))
        nj, nax = C.c_int(), C.c_int()
        self.L.gammahip_dims(self.h, C.byref(nj), C.byref(nax))
        self.nj, self.nax = nj.value, nax.value

    def _check(self, r):
        if r != 0:
            raise RuntimeError(self.L.gammahip_last_error(self.h).decode())

    def set_user_params(self, par):
        """Parameters of the user setup's hooks (gammahip_set_user_params)."""
        a = np.ascontiguousarray(par, dtype=np.float64)
        self._check(self.L.gammahip_set_user_params(self.h, _p(a), int(a.size)))

    def close(self):
        if getattr(self, "h", None):
            self.L.gammahip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- the reference's Grid methods ----
    def regrid(self):
        self._check(self.L.gammahip_regrid(self.h))

    def updateGhosts(self):
        self._check(self.L.gammahip_update_ghosts(self.h))

    def prepForUpdate(self):
        self._check(self.L.gammahip_prep_for_update(self.h))

    def collect_dt(self) -> float:
        dt = np.zeros(1)
        self._check(self.L.gammahip_collect_dt(self.h, _p(dt)))
        return float(dt[0])

    def evolve(self, dt: float | None = None):
        """Grid::evolve(it, t, dt): with dt given, the device advances by it
        (gammahip_evolve_dt); without, by its own collected dt."""
        if dt is None:
            self._check(self.L.gammahip_evolve(self.h))
        else:
            self._check(self.L.gammahip_evolve_dt(self.h, float(dt)))

    def collect_dt_min(self) -> float:
        """Grid::collect_dt as the reference returns it: the minimum before CFL_."""
        m = np.zeros(1)
        self._check(self.L.gammahip_collect_dt_min(self.h, _p(m)))
        return float(m[0])

    def apply(self, what: int):
        self._check(self.L.gammahip_apply(self.h, what))

    # ---- Simu::run loop ----
    def step(self, n: int = 1):
        self._check(self.L.gammahip_step(self.h, n))

    def sync(self):
        self._check(self.L.gammahip_sync(self.h))

    def set_options(self, block=None, remap=None, graph=None, mlp_thr=None, persist=None, occ=None,
                    f1fuse=None, solver=None, f1cell=None, fuse=None):
        """Launch options: threads per workgroup; tile order (0 hardware, 1 XCD
        bands, 2 track groups); hipGraph replay; trigger exponent n of the
        learned guess (threshold 10^-n); persistent packed launch with n
        workgroups per CU or 0; register-capped builds (bit 0 the F1 face
        kernel, bit 1 the update kernel); F1 flux sums inside the update
        kernel; cons2prim solver (C2P_BRENT or C2P_NEWTON); fused
        transverse-gradient and own-face F1 pass; fused passes (bit 0: the F1
        neighbour search inside the transverse-gradient pass; bit 1: MV
        gradients recomputed in the F1 passes; bit 2: F1 faces one thread per
        owner cell)."""
        for key, val in ((OPT_SOLVER, solver), (OPT_BLOCK, block), (OPT_REMAP, remap), (OPT_GRAPH, graph),
                         (OPT_MLP_THR, mlp_thr), (OPT_PERSIST, persist), (OPT_OCC, occ),
                         (OPT_F1FUSE, f1fuse), (OPT_F1CELL, f1cell), (OPT_FUSE, fuse)):
            if val is not None:
                self._check(self.L.gammahip_set_option(self.h, key, int(val)))
        if solver is not None:
            self.solver = int(solver)

    def load_surrogate(self, params):
        """Seed the Newton cons2prim with the learned guess (None switches it off)."""
        if params is None:
            self._check(self.L.gammahip_mlp_load(self.h, None))
            return
        self._mlp = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.L.gammahip_mlp_load(self.h, _p(self._mlp)))

    # ---- exact state (binary resume) ----
    def save_state(self) -> np.ndarray:
        """The grid's exact state between steps as bytes (gammahip_state_save)."""
        n = C.c_ulonglong()
        self._check(self.L.gammahip_state_size(self.h, C.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint8)
        self._check(self.L.gammahip_state_save(self.h, _p(buf), n.value))
        return buf

    def restore_state(self, buf: np.ndarray):
        """Continue from a state saved by a grid of the same setup."""
        b = np.ascontiguousarray(buf, dtype=np.uint8)
        self._check(self.L.gammahip_state_restore(self.h, _p(b), b.size))

    # ---- state access ----
    def cells(self, field: int) -> np.ndarray:
        out = np.zeros((self.nj, self.nax))
        self._check(self.L.gammahip_download(self.h, field, _p(out)))
        return out

    faces = cells

    def tracks(self) -> dict:
        a = [np.zeros(self.nj, dtype=np.int32) for _ in range(4)]
        self._check(self.L.gammahip_track_info(self.h, *[_p(x) for x in a]))
        return dict(ntrack=a[0], nact=a[1], iLbnd=a[2], iRbnd=a[3])

    def time(self):
        t = np.zeros(self.nj)
        dt = np.zeros(self.nj)
        self._check(self.L.gammahip_time(self.h, _p(t), _p(dt)))
        return t, dt

    def stats(self) -> dict:
        st = Stats()
        self._check(self.L.gammahip_stats_get(self.h, C.byref(st)))
        return {n: getattr(st, n) for n, _ in Stats._fields_ if n != "reserved"}

    def regrid_counts(self):
        s = np.zeros(self.nj, dtype=np.int64)
        m = np.zeros(self.nj, dtype=np.int64)
        self._check(self.L.gammahip_regrid_counts(self.h, _p(s), _p(m)))
        return s, m

    def kernel_times(self):
        ms = np.zeros(16)
        n = np.zeros(16, dtype=np.int64)
        self._check(self.L.gammahip_kernel_times(self.h, _p(ms), _p(n)))
        return ms, n

    def active(self, field: int):
        """Active-cell values of a cell field, one array per active track."""
        tr = self.tracks()
        a = self.cells(field)
        rc = self.rc
        js = range(rc.ngst, self.nj - rc.ngst) if rc.nd == 2 else range(self.nj)
        return [a[j, tr["iLbnd"][j] + 1: tr["iRbnd"][j]] for j in js]

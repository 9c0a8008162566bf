"""Problem setups written against GAMMA's hook API.

A setup plays the role of one ``src/Initial/Tests/*.cpp`` file of the
reference: ``loadParams`` (sizes), ``Grid::initialGeometry`` /
``Grid::initialValues`` (the initial active cells), ``userKinematics`` /
``userBoundaries`` (boundary faces and ghost cells) and
``checkCellForRegrid`` (AMR rule).  Initial data is produced here as
structure-of-arrays numpy buffers; the run-time hooks are selected by kind and
parameters in :class:`RunConfig` and evaluated by the solver itself (device
code in the HIP backend).

Templates followed (reference files):
  * 1D spherical Blandford-McKee blast wave -- Initial/Tests/BM1D.cpp:62-332
    (code units c = 1, ambient rho = 1, R_shock = 1 at t0)
  * planar relativistic shock tube -- Initial/Tests/1DST.cpp:19-78
  * 2D cartesian Riemann problem on x tracks -- Initial/Tests/2DST.cpp:166-322
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np

# variable slots (reference environment.h:33-35, 70-82)
RHO, PPP, UU1, UU2 = 0, 1, 2, 3
TR1, GMN, GMX, GMD, PSN = 4, 6, 7, 8, 9

GEOM_CART, GEOM_SPH = 0, 1
TS_RK3, TS_EULER = 0, 1
RECON_PLM, RECON_PC = 0, 1
EOS_IDEAL, EOS_SYNGE = 0, 1
# hook kinds; the USER kinds call the user setup's functions (gamma_amd/csrc/user_hooks.h)
KIN_NONE, KIN_FIXED, KIN_USER = 0, 1, 2
BND_NONE, BND_BM, BND_USER = 0, 1, 2
RG_NONE, RG_AR2D, RG_BM1D, RG_USER = 0, 1, 2, 3

# physical constants (reference constants.h:8-28)
MP = 1.6726219e-24
ME = 9.1093835e-28
C_LIGHT = 2.99792458e10
ALPHA_SYN = 1.29251816e-09


def normalized_constants(rho_norm: float, v_norm: float, l_norm: float):
    """normalizeConstants (reference constants.cpp:12-20): (Nalpha, Nmp, Nme)."""
    n_alpha = ALPHA_SYN / (1. / (rho_norm * v_norm * l_norm))
    n_mp = MP / (rho_norm * math.pow(l_norm, 3))
    n_me = ME / (rho_norm * math.pow(l_norm, 3))
    return n_alpha, n_mp, n_me


@dataclass
class RunConfig:
    """Compile-time switches of environment.h plus the setup's hook choices."""
    nd: int = 1                   # NUM_D
    geom: int = GEOM_CART         # Hydro + Geometry module pair
    shock: int = 0                # SHOCK_DETECTION_
    sync: int = 0                 # LOCAL_SYNCHROTRON_
    ngst: int = 2
    nmax: int = 0                 # max active cells per track
    ncell_mv: int = 0             # initial active cells per track
    ncell_f1: int = 1             # active tracks (2D)
    gamma: float = 5. / 3.        # GAMMA_
    cfl: float = 0.2              # CFL_
    vi: float = 1.                # VI
    shock_thr: float = 0.01       # DETECT_SHOCK_THRESHOLD_
    gamma_max_init: float = 1.e8  # GAMMA_MAX_INIT_
    n_alpha: float = 0.
    n_mp: float = 1.
    n_me: float = 1.
    kin_mode: int = KIN_NONE
    v_in: float = 0.
    v_out: float = 0.
    bnd_mode: int = BND_NONE
    bm: tuple = (0., 0., 0., 0., 0., 0.)
    rg_mode: int = RG_NONE
    rg: tuple = (0., 0., 0., 0., 0., 0.)
    c2p_maxiter: int = 1000
    timestep: int = TS_RK3        # TIMESTEP_
    recon: int = RECON_PLM        # SPATIAL_RECONSTRUCTION_
    vpspec: int = 0               # VARIABLE_PSPEC_
    eos: int = EOS_IDEAL          # EOS_
    circ_regrid: int = 0          # CIRC_REGRID_
    upar: tuple = ()              # parameters of the user setup's hooks (up to 16)
    user_cell: int = 0            # the user setup's per-cell hooks: 1 userSourceTerms, 2 cons2prim_user

    @property
    def nq(self) -> int:          # NUM_Q = NUM_C(4) + NUM_TR(2) + NUM_S
        return 6 + (4 if self.sync else 0)

    def replace(self, **kw) -> "RunConfig":
        return dataclasses.replace(self, **kw)


@dataclass
class Setup:
    """Initial active cells: x,dx [ncell_f1][ncell_mv], y,dy [ncell_f1],
    prim [nq][ncell_f1][ncell_mv] with 3-velocities in the UU slots
    (prepForRun converts them, reference 1d.cpp:934-951)."""
    name: str
    cfg: RunConfig
    t0: float
    x: np.ndarray
    dx: np.ndarray
    y: np.ndarray
    dy: np.ndarray
    prim: np.ndarray
    meta: dict = field(default_factory=dict)
    # initial active cells per track (None: ncell_mv for every track); the
    # arrays above stay [track][ncell_mv], read up to each count
    ncells: np.ndarray | None = None
    # per-track hook overrides for ensembles of 1D runs: dicts with keys of
    # RunConfig (geom, kin_mode, v_in, v_out, bnd_mode, bm, rg_mode)
    hooks: list | None = None


# ---------------------------------------------------------------------------
# counter-based PRNG (SplitMix64, Steele/Lea/Flood 2014), identical on host
# and device: u(k) = mix(seed + (k+1) * golden) -> uniform [-1, 1)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix_uniform(seed: int, idx: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * _GOLD
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


DEFAULT_SEED = 20261202


# ---------------------------------------------------------------------------
def shock_tube_1d(n: int = 400, rho_l=10., p_l=13.33, rho_r=1., p_r=1.e-6,
                  gamma=5. / 3., xmin=0., xmax=1., xmid=0.5, geom=GEOM_CART,
                  nmax=None, shock=0) -> Setup:
    """Planar relativistic shock tube (Marti & Mueller problem 1 by default).
    Static boundary faces as in 2DST.cpp:261-278, no AMR."""
    cfg = RunConfig(nd=1, geom=geom, shock=shock, nmax=nmax or n + 8, ncell_mv=n,
                    gamma=gamma, kin_mode=KIN_FIXED, v_in=0., v_out=0.)
    dxc = (xmax - xmin) / n
    x = xmin + (np.arange(n) + 0.5) * dxc
    prim = np.zeros((cfg.nq, 1, n))
    left = x <= xmid
    prim[RHO, 0] = np.where(left, rho_l, rho_r)
    prim[PPP, 0] = np.where(left, p_l, p_r)
    prim[TR1, 0] = np.where(left, 1., 2.)
    return Setup("shock_tube_1d", cfg, 0., x[None, :], np.full((1, n), dxc),
                 np.zeros(1), np.ones(1), prim)


def expanding_tube_1d(n: int = 200, v_edge=0.5) -> Setup:
    """Shock tube whose boundary faces move outward (userKinematics v_in = -v,
    v_out = +v), with the aspect-ratio rule of BM1D.cpp:277-319 on a flat
    geometry: the edge cells grow and split repeatedly next to the ghosts."""
    s = shock_tube_1d(n, p_r=1.e-6, xmid=0.1)
    s.prim[RHO, 0] *= 1. + 0.5 * s.x[0]          # density ramp at uniform pressure
    s.cfg = s.cfg.replace(nmax=2 * n, kin_mode=KIN_FIXED, v_in=-v_edge, v_out=v_edge,
                          rg_mode=RG_BM1D, rg=(3., 0.1, 1. / n, 1.e-6, 1., 0.))
    s.name = "expanding_tube_1d"
    return s


def bm_profile(r, t, lfac_shock2_0, t0, rho_a, eta, gamma):
    """Blandford-McKee self-similar profile in code units (c = 1, k = 0);
    BM1D.cpp:62-86 with lfacShock2(t) = lfacShock2(t0) (t0/t)^3."""
    tr = t0 / t
    L = lfac_shock2_0 * tr * tr * tr
    pa = rho_a * eta
    ha = 1. + pa * gamma / (gamma - 1.) / rho_a
    pf = 2. / 3. * L * ha * rho_a
    lfacf = math.sqrt(max(1., .5 * L))
    Df = 2. * L * rho_a
    chi = (1. + 8. * L) * (1. - r / t)
    p = pf * np.power(chi, -17. / 12.)
    lfac = np.sqrt(lfacf * lfacf / chi + 1)
    D = Df * np.power(chi, -7. / 4.)
    rho = D / lfac
    v = np.sqrt(1. - 1. / (lfac * lfac))
    return rho, v, p


def blandford_mckee_1d(n: int = 4000, lfac_shock=100., eta=1.e-10, gamma=4. / 3.,
                       nmax=None, sync=1, shock=1, dth=5.e-4, t0=None) -> Setup:
    """Config: 1D spherical BM blast wave, shock Lorentz factor 100, ambient
    rho = 1, p = eta rho c^2, cells geometrically spaced on [0.5, 1.5] R_sh
    with r_{i+1/2} = 0.5 R_sh 3^(i/n).  Boundary faces as BM1D.cpp:239-244
    (v_in = 0, v_out = 1.5), inner ghosts from the analytic solution
    (BM1D.cpp:254-274), AMR rule of BM1D.cpp:277-319.  The start time puts
    the shock at R_sh = 1 (BM1D.cpp:42) unless t0 is given."""
    L0 = lfac_shock * lfac_shock
    if t0 is None:
        t0 = 1. / (1. - 1. / (1. + 8. * L0))      # R_shock(t0) = 1 (BM1D.cpp:42)
    rshock = t0 * (1. - 1. / (1. + 8. * L0))
    rho_norm, v_norm, l_norm = MP, C_LIGHT, C_LIGHT  # BM1D.cpp:20-23
    na, nmp, nme = normalized_constants(rho_norm, v_norm, l_norm)
    cfg = RunConfig(nd=1, geom=GEOM_SPH, shock=shock, sync=sync, nmax=nmax or int(n * 1.25),
                    ncell_mv=n, gamma=gamma, n_alpha=na, n_mp=nmp, n_me=nme,
                    kin_mode=KIN_FIXED, v_in=0., v_out=1.5,
                    bnd_mode=BND_BM, bm=(L0, t0, 1., eta, 0., 0.),
                    rg_mode=RG_BM1D, rg=(3., 0.1, dth, eta, 1., 0.))
    i = np.arange(n + 1)
    edges = rshock * 0.5 * np.power(3., i / n)
    x = 0.5 * (edges[1:] + edges[:-1])
    dx = edges[1:] - edges[:-1]
    prim = np.zeros((cfg.nq, 1, n))
    inside = x <= rshock
    rho_b, v_b, p_b = bm_profile(np.minimum(x, rshock * 0.999999), t0, L0, t0, 1., eta, gamma)
    prim[RHO, 0] = np.where(inside, rho_b, 1.)
    prim[UU1, 0] = np.where(inside, v_b, 0.)
    prim[PPP, 0] = np.where(inside, p_b, eta)
    prim[TR1, 0] = np.where(inside, 2., 1.)
    return Setup("blandford_mckee_1d", cfg, t0, x[None, :], dx[None, :],
                 np.zeros(1), np.ones(1), prim, meta=dict(L0=L0, eta=eta))


def jet_lfac(theta):
    """Angular Lorentz-factor profile of the 2D blast wave: 100 exp(-(theta/0.1)^2) + 2."""
    return 100. * np.exp(-(np.asarray(theta) / 0.1) ** 2) + 2.


def bm_jet_2d(ntheta: int = 1024, theta_max: float = 0.5, n_lo: int = 500, n_hi: int = 4000,
              eta=1.e-10, gamma=4. / 3., nmax=None, sync=1, shock=1, t0=1.) -> Setup:
    """Config: 2D spherical (r, theta) relativistic blast wave whose shock
    Lorentz factor follows jet_lfac(theta), on ntheta theta tracks over
    [0, theta_max].  Each track starts as the BM solution of its own Lorentz
    factor at the common time t0 (BM1D.cpp:62-86 per track), on cells spaced
    geometrically over [0.5, 1.5] R_shock(theta) as in the 1D case; its length
    is seeded between n_lo (Lorentz factor 2) and n_hi (102) in proportion to
    the Lorentz factor, so tracks are uneven from the start.  Inner boundary:
    each track's own BM inflow (BM1D.cpp:254-274); boundary faces v_in = 0,
    v_out = 1.5; the per-track AMR rule of BM1D.cpp:277-319 with dtheta the
    track width; shock detection (threshold 0.01) and local synchrotron
    particles (gamma_max 1e8) on."""
    dth = theta_max / ntheta
    th = (np.arange(ntheta) + 0.5) * dth
    lf = jet_lfac(th)
    n = np.rint(n_lo + (n_hi - n_lo) * (lf - 2.) / 100.).astype(np.int64)
    nmv = int(n.max())
    rho_norm, v_norm, l_norm = MP, C_LIGHT, C_LIGHT
    na, nmp, nme = normalized_constants(rho_norm, v_norm, l_norm)
    cfg = RunConfig(nd=2, geom=GEOM_SPH, shock=shock, sync=sync, nmax=nmax or int(nmv * 1.25),
                    ncell_mv=nmv, ncell_f1=ntheta, gamma=gamma, n_alpha=na, n_mp=nmp, n_me=nme,
                    kin_mode=KIN_FIXED, v_in=0., v_out=1.5, bnd_mode=BND_BM,
                    bm=(lf[0] ** 2, t0, 1., eta, 0., 0.),
                    rg_mode=RG_BM1D, rg=(3., 0.1, dth, eta, 1., 0.))
    x = np.zeros((ntheta, nmv))
    dx = np.ones((ntheta, nmv))
    prim = np.zeros((cfg.nq, ntheta, nmv))
    prim[RHO] = 1.
    prim[PPP] = eta
    prim[TR1] = 1.
    hooks = []
    for a in range(ntheta):
        L0 = lf[a] ** 2
        rsh = t0 * (1. - 1. / (1. + 8. * L0))
        m = int(n[a])
        edges = rsh * 0.5 * np.power(3., np.arange(m + 1) / m)
        xa = 0.5 * (edges[1:] + edges[:-1])
        x[a, :m] = xa
        dx[a, :m] = edges[1:] - edges[:-1]
        x[a, m:] = xa[-1]
        inside = xa <= rsh
        rb, vb, pb = bm_profile(np.minimum(xa, rsh * 0.999999), t0, L0, t0, 1., eta, gamma)
        prim[RHO, a, :m] = np.where(inside, rb, 1.)
        prim[UU1, a, :m] = np.where(inside, vb, 0.)
        prim[PPP, a, :m] = np.where(inside, pb, eta)
        prim[TR1, a, :m] = np.where(inside, 2., 1.)
        hooks.append(dict(bm=(L0, t0, 1., eta)))
    return Setup("bm_jet_2d", cfg, t0, x, dx, th, np.full(ntheta, dth), prim,
                 meta=dict(theta_max=theta_max, lengths="seeded"), ncells=n, hooks=hooks)


def lorentz_tube_1d(n: int, lfac: float, gamma=4. / 3., nmax=None, shock=1, rho=1., p=1.e-4,
                    t0=1.) -> Setup:
    """Planar relativistic shock tube: gas with Lorentz factor lfac flowing
    right (x < 0.5) into gas at rest (x > 0.5), both rho = 1, p = 1e-4, on
    [0, 1] with n cells; no kinematics, boundary or regrid hooks (1DST.cpp)."""
    cfg = RunConfig(nd=1, geom=GEOM_CART, shock=shock, nmax=nmax or int(n * 1.25), ncell_mv=n,
                    gamma=gamma)
    dxc = 1. / n
    x = (np.arange(n) + 0.5) * dxc
    prim = np.zeros((cfg.nq, 1, n))
    left = x <= 0.5
    prim[RHO, 0] = rho
    prim[PPP, 0] = p
    prim[UU1, 0] = np.where(left, math.sqrt(1. - 1. / (lfac * lfac)), 0.)
    prim[TR1, 0] = np.where(left, 1., 2.)
    return Setup("lorentz_tube_1d", cfg, t0, x[None, :], np.full((1, n), dxc), np.zeros(1),
                 np.ones(1), prim, meta=dict(lfac=lfac))


def ensemble_1d(n_runs: int = 256, sizes=(512, 16384), lfacs=(2., 300.), seed=DEFAULT_SEED,
                eta=1.e-10, dth=5.e-4):
    """Config: n_runs independent 1D relativistic runs advanced together, one
    track each: even members are spherical Blandford-McKee blast waves, odd
    members planar Lorentz shock tubes.  Sizes and Lorentz factors are drawn
    log-uniform from the given ranges with SplitMix64 (counters 2m and 2m+1
    of member m).  All members share gamma = 4/3, CFL 0.2, shock detection
    and the start time t0 = 1, and no synchrotron; each keeps its own
    geometry, kinematics, boundary, regrid rule and time step.

    Returns (ensemble, members): the ensemble Setup (one track per member,
    per-track hooks and lengths) and the member Setups, each runnable alone
    with the same capacity (nmax), so that member m of the ensemble must
    match member m run by itself."""
    u = splitmix_uniform(seed, np.arange(2 * n_runs, dtype=np.uint64))
    u = 0.5 * (u + 1.)                                        # [0, 1)
    lo, hi = np.log(sizes[0]), np.log(sizes[1])
    ns = np.rint(np.exp(lo + (hi - lo) * u[0::2])).astype(np.int64)
    lf = np.exp(np.log(lfacs[0]) + (np.log(lfacs[1]) - np.log(lfacs[0])) * u[1::2])
    nmax = int(ns.max() * 1.25)
    members = []
    for m in range(n_runs):
        if m % 2 == 0:
            members.append(blandford_mckee_1d(int(ns[m]), lfac_shock=float(lf[m]), eta=eta, nmax=nmax,
                                              sync=0, shock=1, dth=dth, t0=1.))
        else:
            members.append(lorentz_tube_1d(int(ns[m]), float(lf[m]), nmax=nmax, shock=1))
    nmv = int(ns.max())
    bm_cfg = members[0].cfg
    cfg = RunConfig(nd=1, geom=-1, shock=1, sync=0, nmax=nmax, ncell_mv=nmv, gamma=4. / 3.,
                    n_alpha=bm_cfg.n_alpha, n_mp=bm_cfg.n_mp, n_me=bm_cfg.n_me,
                    rg_mode=RG_NONE, rg=bm_cfg.rg)
    x = np.zeros((n_runs, nmv))
    dx = np.ones((n_runs, nmv))
    prim = np.zeros((cfg.nq, n_runs, nmv))
    hooks = []
    for m, ms in enumerate(members):
        k = int(ns[m])
        x[m, :k], dx[m, :k] = ms.x[0], ms.dx[0]
        x[m, k:] = ms.x[0, -1]
        prim[:, m, :k] = ms.prim[:, 0, :]
        c = ms.cfg
        hooks.append(dict(geom=c.geom, kin_mode=c.kin_mode, v_in=c.v_in, v_out=c.v_out,
                          bnd_mode=c.bnd_mode, bm=c.bm, rg_mode=c.rg_mode))
    ens = Setup("ensemble_1d", cfg, 1., x, dx, np.zeros(n_runs), np.ones(n_runs), prim,
                meta=dict(seed=seed, sizes=ns.tolist(), lfacs=lf.tolist()), ncells=ns, hooks=hooks)
    return ens, members


def riemann_2d_uneven(nx: int = 96, ny: int = 16, gamma=5. / 3., seed=DEFAULT_SEED) -> Setup:
    """Test configuration (not a benchmark): the 2D Riemann problem with
    tracks of different lengths, track a holding n_a = nx - (37 a mod nx/2)
    uniform cells over [0, 1] (between nx/2 and nx), so that neighbouring
    tracks' faces are misaligned from the start and a cell's F1 neighbours sit
    at indices far from its own (up to half the track away): lists of one to
    three cells per side, neighbour searches that leave the cell's
    neighbourhood, incoming faces from several lower cells.  Same states,
    perturbation, boundaries and AR2D rule as riemann_2d; dy = 1/nx."""
    n = np.array([nx - (37 * a) % (nx // 2) for a in range(ny)], dtype=np.int64)
    cfg = RunConfig(nd=2, geom=GEOM_CART, nmax=(nx * 5) // 4, ncell_mv=nx,
                    ncell_f1=ny, gamma=gamma, kin_mode=KIN_FIXED, v_in=0., v_out=0.,
                    rg_mode=RG_AR2D, rg=(3., 0.2, 0., 0., 0., 0.))
    dyc = 1. / nx
    y = (np.arange(ny) + 0.5) * dyc
    x = np.zeros((ny, nx))
    dx = np.ones((ny, nx))
    prim = np.zeros((cfg.nq, ny, nx))
    xi = splitmix_uniform(seed, np.arange(nx * ny, dtype=np.uint64)).reshape(ny, nx)
    for a in range(ny):
        m = int(n[a])
        xa = (np.arange(m) + 0.5) / m
        x[a, :m] = xa
        dx[a, :m] = 1. / m
        x[a, m:] = xa[-1]
        left = xa <= 0.5
        prim[RHO, a, :m] = np.where(left, 10., 1.) * (1. + 1.e-3 * xi[a, :m])
        prim[PPP, a, :m] = np.where(left, 13.33, 1.e-6)
        prim[TR1, a, :m] = np.where(left, 1., 2.)
        prim[RHO, a, m:] = 1.
        prim[PPP, a, m:] = 1.e-6
    return Setup("riemann_2d_uneven", cfg, 0., x, dx, y, np.full(ny, dyc), prim,
                 meta=dict(seed=seed), ncells=n)


def riemann_2d(nx: int = 2048, ny: int = 2048, rho_l=10., p_l=13.33, rho_r=1.,
               p_r=1.e-6, gamma=5. / 3., amp=1.e-3, seed=DEFAULT_SEED,
               nmax=None, regrid=True, ly: float = 1.) -> Setup:
    """Config: 2D cartesian relativistic Riemann problem on x-directed tracks
    over [0,1] x [0,ly], diaphragm at x = 0.5, density seeded with a 1e-3
    SplitMix64 perturbation per cell (counter j*nx + i); static boundary faces
    and the aspect-ratio AMR rule of 2DST.cpp:261-322 (split above 3, merge
    below 0.2).  With ly = ny/2048 and nx = 2048 the grid is the bottom strip
    of the 2048 x 2048 configuration, cell for cell."""
    cfg = RunConfig(nd=2, geom=GEOM_CART, nmax=nmax or (nx * 5) // 4, ncell_mv=nx,
                    ncell_f1=ny, gamma=gamma, kin_mode=KIN_FIXED, v_in=0., v_out=0.,
                    rg_mode=RG_AR2D if regrid else RG_NONE, rg=(3., 0.2, 0., 0., 0., 0.))
    dxc, dyc = 1. / nx, ly / ny
    x1 = (np.arange(nx) + 0.5) * dxc
    y = (np.arange(ny) + 0.5) * dyc
    x = np.broadcast_to(x1, (ny, nx)).copy()
    prim = np.zeros((cfg.nq, ny, nx))
    left = x <= 0.5
    xi = splitmix_uniform(seed, np.arange(nx * ny, dtype=np.uint64)).reshape(ny, nx)
    prim[RHO] = np.where(left, rho_l, rho_r) * (1. + amp * xi)
    prim[PPP] = np.where(left, p_l, p_r)
    prim[TR1] = np.where(left, 1., 2.)
    return Setup("riemann_2d", cfg, 0., x, np.full((ny, nx), dxc), y, np.full(ny, dyc), prim,
                 meta=dict(seed=seed, amp=amp))


def reflecting_box(nx: int = 200, ny: int = 0, mesh: float = 0.5, rho0: float = 1., p0: float = 0.01,
                   p_hot: float = 10., r_hot: float = 0.1, gamma=4. / 3., shock: int = 1,
                   regrid: bool = True, seed=DEFAULT_SEED) -> Setup:
    """A setup written against the user hooks as shipped (gamma_amd/csrc/
    user_hooks.h): a hot region of radius r_hot at the centre of the box
    [0, 1] (1D, ny = 0) or [0, 1] x [0, ny/nx] (2D, ny x-tracks) expands
    against reflecting walls at the track ends (gh_user_ghost: the edge
    cell's state with the normal velocity reversed; gh_user_face_speed: the
    end faces still); the other faces move at `mesh` times the solver's face
    speeds; with `regrid`, cells wider than 1.5/nx split (where the density is
    above 0.1 rho0) and cells narrower than 0.3/nx merge (gh_user_regrid).  The 2D density
    carries a 1e-3 SplitMix64 perturbation per cell (counter j*nx + i)."""
    nd = 2 if ny else 1
    ntr = ny if ny else 1
    upar = (mesh, 1.5 / nx if regrid else 0., 0.3 / nx if regrid else 0., 0.1 * rho0)
    cfg = RunConfig(nd=nd, geom=GEOM_CART, shock=shock, nmax=(nx * 3) // 2, ncell_mv=nx, ncell_f1=ntr,
                    gamma=gamma, kin_mode=KIN_USER, bnd_mode=BND_USER, rg_mode=RG_USER, upar=upar)
    dxc = 1. / nx
    x1 = (np.arange(nx) + 0.5) * dxc
    y = (np.arange(ntr) + 0.5) * dxc if nd == 2 else np.zeros(1)
    dy = np.full(ntr, dxc) if nd == 2 else np.ones(1)
    x = np.broadcast_to(x1, (ntr, nx)).copy()
    yc = 0.5 * ntr * dxc if nd == 2 else 0.
    r = np.sqrt((x - 0.5) ** 2 + ((y[:, None] - yc) ** 2 if nd == 2 else 0.))
    hot = r < r_hot
    prim = np.zeros((cfg.nq, ntr, nx))
    prim[RHO] = rho0
    if nd == 2:
        prim[RHO] *= 1. + 1.e-3 * splitmix_uniform(seed, np.arange(nx * ntr, dtype=np.uint64)).reshape(ntr, nx)
    prim[RHO] = np.where(hot, 2.5 * prim[RHO], prim[RHO])
    prim[PPP] = np.where(hot, p_hot, p0)
    prim[TR1] = np.where(hot, 2., 1.)
    return Setup("reflecting_box", cfg, 0., x, np.full((ntr, nx), dxc), y, dy, prim,
                 meta=dict(seed=seed, mesh=mesh))

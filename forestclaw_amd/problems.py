"""Initial data, refinement tagging and forests for the workloads the batched
patch update is measured on.  This is harness code (the reference's
qinit/setprob and initial regrid), evaluated on the host with numpy; it feeds
identical forests and initial data to the GPU path and to the CPU oracle.

Initial values are sampled at cell centres.

  advection  : 2D scalar advection, velocity (1.0, 0.5), periodic unit
               square, q = 1 inside the disk r < 0.2 about (0.5, 0.5)
  dam break  : shallow water, g = 1, h = h_in inside r < r0 about
               (0.5, 0.5), 1 outside, zero momentum, outflow boundaries
  shock-bubble: Euler, gamma = 1.4, Mach-1.5 shock at x = 0.2 moving into
               rho = 1, p = 1; bubble rho = 0.1, r = 0.1 at (0.5, 0.25);
               4 x 1 brick over [0, 2] x [0, 0.5], walls top and bottom
"""
from dataclasses import dataclass, field

import numpy as np

from . import RP_ADVECTION, RP_EULER, RP_SHALLOW, Forest, solver_shape


@dataclass
class Problem:
    name: str
    solver: int
    extent: tuple          # (ax, bx, ay, by)
    nbx: int
    nby: int
    periodic: tuple
    mthbc: tuple
    minlevel: int
    maxlevel: int
    mx: int
    cfl: float
    qfun: object           # (x, y) -> array (meqn, ...)
    auxfun: object = None  # (x, y) -> array (maux, ...)
    tag_field: int = 0
    tag_threshold: float = 0.01
    grav: float = 1.0
    gamma: float = 1.4
    params: dict = field(default_factory=dict)
    rp_params: tuple = ()     # a registered solver's constants (hipclaw_config_t.rp_params)
    mthlim: tuple = None      # limiter per wave family (None: MC for every family)
    max_speed: float = None   # largest wave speed, for solvers stable_dt does not know
    oracle_solver: int = None # the oracle's number for this solver (None: the same)
    my: int = None            # cells per patch in y (None: mx)

    def __post_init__(self):
        if self.my is None:
            self.my = self.mx

    @property
    def meqn(self):
        return solver_shape(self.solver)[0]

    @property
    def maux(self):
        return solver_shape(self.solver)[2]


def advection_problem(minlevel=4, maxlevel=4, mx=16):
    def q(x, y):
        return ((x - 0.5) ** 2 + (y - 0.5) ** 2 < 0.2 ** 2).astype(np.float64)[None]

    def aux(x, y):
        return np.stack([np.full_like(x, 1.0), np.full_like(x, 0.5)])

    return Problem("advection", RP_ADVECTION, (0.0, 1.0, 0.0, 1.0), 1, 1, (True, True),
                   (2, 2, 2, 2), minlevel, maxlevel, mx, 0.9, q, aux, tag_threshold=0.5)


def dambreak_problem(minlevel=3, maxlevel=6, mx=32, r0=0.2, hin=2.0, hout=1.0, my=None):
    def q(x, y):
        h = np.where((x - 0.5) ** 2 + (y - 0.5) ** 2 < r0 ** 2, hin, hout)
        return np.stack([h, np.zeros_like(h), np.zeros_like(h)])

    return Problem("dambreak", RP_SHALLOW, (0.0, 1.0, 0.0, 1.0), 1, 1, (False, False),
                   (1, 1, 1, 1), minlevel, maxlevel, mx, 0.9, q, grav=1.0,
                   params=dict(r0=r0, hin=hin, hout=hout), my=my)


def shockbubble_problem(minlevel=2, maxlevel=7, mx=16, gamma=1.4, mach=1.5):
    g = gamma
    # Rankine-Hugoniot post-shock state for a shock moving into rho = 1, p = 1, u = 0
    p2 = 1.0 + 2.0 * g / (g + 1.0) * (mach ** 2 - 1.0)
    rho2 = (g + 1.0) * mach ** 2 / ((g - 1.0) * mach ** 2 + 2.0)
    c1 = np.sqrt(g * 1.0 / 1.0)
    u2 = mach * c1 * (1.0 - 1.0 / rho2)

    def q(x, y):
        behind = x < 0.2
        bubble = (x - 0.5) ** 2 + (y - 0.25) ** 2 < 0.1 ** 2
        rho = np.where(behind, rho2, np.where(bubble, 0.1, 1.0))
        u = np.where(behind, u2, 0.0)
        p = np.where(behind, p2, 1.0)
        E = p / (g - 1.0) + 0.5 * rho * u * u
        return np.stack([rho, rho * u, np.zeros_like(rho), E])

    return Problem("shockbubble", RP_EULER, (0.0, 2.0, 0.0, 0.5), 4, 1, (False, False),
                   (1, 1, 3, 3), minlevel, maxlevel, mx, 0.8, q, gamma=gamma,
                   params=dict(p2=p2, rho2=rho2, u2=u2))


def acoustics_problem(minlevel=2, maxlevel=4, mx=16, rho=1.0, bulk=4.0):
    """The reference's radial acoustics example
    (applications/cudaclaw/acoustics/2d/radial: fclaw_options.ini, cuda.ini,
    user_cuda/qinit.f90): a ring of pressure 1 + cos(pi (r - 0.5) / 0.2) about
    r = 0.5 on [-1, 1]^2, at rest, rho = 1, bulk = 4, extrapolation
    boundaries; run with the application's own Riemann solver (the example
    solver library user_solvers/libacoustics.so)."""
    from . import example_solver
    cc = float(np.sqrt(bulk / rho))

    def q(x, y):
        r = np.sqrt(x * x + y * y)
        p = np.where(np.abs(r - 0.5) <= 0.2, 1.0 + np.cos(np.pi * (r - 0.5) / 0.2), 0.0)
        return np.stack([p, np.zeros_like(p), np.zeros_like(p)])

    return Problem("acoustics", example_solver("acoustics"), (-1.0, 1.0, -1.0, 1.0), 1, 1,
                   (False, False), (1, 1, 1, 1), minlevel, maxlevel, mx, 0.9, q, tag_threshold=1.0,
                   rp_params=(rho, bulk, cc, rho * cc), max_speed=cc, oracle_solver=3,
                   params=dict(rho=rho, bulk=bulk))


def tracers_problem(minlevel=2, maxlevel=4, mx=16):
    """Ten colour equations on the periodic unit square (the example solver
    library user_solvers/libtracers.so): tracer k is a disk of radius 0.15
    about a centre that turns with k, moved by c_k (u, v) with
    u = 1 + 0.5 sin(2 pi y), v = 0.5 + 0.25 cos(2 pi x), c_k = (10 - k) / 10;
    each wave family takes its own limiter (1..6, then 1..4)."""
    from . import example_solver
    n = 10

    def q(x, y):
        out = []
        for k in range(n):
            a = 2 * np.pi * k / n
            cx, cy = 0.5 + 0.25 * np.cos(a), 0.5 + 0.25 * np.sin(a)
            out.append(((x - cx) ** 2 + (y - cy) ** 2 < 0.15 ** 2).astype(np.float64))
        return np.stack(out)

    def aux(x, y):
        return np.stack([1.0 + 0.5 * np.sin(2 * np.pi * y), 0.5 + 0.25 * np.cos(2 * np.pi * x)])

    return Problem("tracers", example_solver("tracers"), (0.0, 1.0, 0.0, 1.0), 1, 1, (True, True),
                   (2, 2, 2, 2), minlevel, maxlevel, mx, 0.9, q, aux, tag_threshold=0.5,
                   mthlim=tuple(1 + k % 6 for k in range(n)), max_speed=1.5, oracle_solver=4)


def ensemble_problem(nruns=64, mx=16, minlevel=2, maxlevels=(4, 7), seed=2026):
    """An ensemble of radial dam breaks, one per block of an nruns x 1 brick of
    isolated unit squares (block b is [b, b+1] x [0, 1]): r0 ~ U[0.1, 0.3],
    h_in ~ U[1.5, 3.0], maxlevel ~ U{maxlevels}, drawn from numpy's PCG64
    with the given seed.  Each run keeps its own stable dt (a per-slot dt
    scale on the device)."""
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(0.1, 0.3, nruns)
    hin = rng.uniform(1.5, 3.0, nruns)
    mlev = rng.integers(maxlevels[0], maxlevels[1] + 1, nruns)

    def q(x, y):
        b = np.clip(np.floor(x).astype(np.int64), 0, nruns - 1)
        xr = x - b
        h = np.where((xr - 0.5) ** 2 + (y - 0.5) ** 2 < r0[b] ** 2, hin[b], 1.0)
        return np.stack([h, np.zeros_like(h), np.zeros_like(h)])

    return Problem("ensemble", RP_SHALLOW, (0.0, float(nruns), 0.0, 1.0), nruns, 1, (False, False),
                   (1, 1, 1, 1), minlevel, int(mlev.max()), mx, 0.9, q, grav=1.0,
                   params=dict(r0=r0, hin=hin, maxlevel=mlev, seed=seed))


def member_problem(ens, b):
    """Run b of an ensemble as a problem of its own on the unit square."""
    pr = ens.params
    return dambreak_problem(ens.minlevel, int(pr["maxlevel"][b]), mx=ens.mx, r0=float(pr["r0"][b]),
                            hin=float(pr["hin"][b]))


def build_ensemble_forest(ens, mbc=2):
    """Initial refinement of every run of an ensemble up to its own maxlevel,
    in one forest of isolated blocks."""
    forest = Forest(ens.nbx, ens.nby, ens.periodic, ens.minlevel, isolated=True)
    mlev = ens.params["maxlevel"]
    for _ in range(ens.maxlevel - ens.minlevel):
        q, _ = initial_data(ens, forest, mbc)
        lv, gx, _ = forest.leaves()
        flags = tag_gradient(ens, forest, q, mbc) & (lv < mlev[gx >> lv]).astype(np.int32)
        if not flags.any():
            break
        forest.refine(flags, ens.maxlevel)
    return forest


def member_leaves(forest, b):
    """Leaves of block b of a 1-row brick, moved to block 0."""
    lv, gx, gy = forest.leaves()
    sel = (gx >> lv) == b
    return lv[sel], gx[sel] - (b << lv[sel]), gy[sel], np.nonzero(sel)[0]


def ensemble_dt(ens, forest, q, mbc=2):
    """Per-slot dt: each run's stable dt (Courant number 0.9 on its finest
    patches for the initial data)."""
    _, _, dx, _ = forest.geometry(ens.extent, ens.mx, ens.mx)
    inner = q[:, :, mbc:-mbc, mbc:-mbc]
    h = inner[:, 0]
    s = np.maximum((np.abs(inner[:, 1] / h) + np.sqrt(ens.grav * h)).max(axis=(1, 2)),
                   (np.abs(inner[:, 2] / h) + np.sqrt(ens.grav * h)).max(axis=(1, 2)))
    lv, gx, _ = forest.leaves()
    blk = gx >> lv
    per_run = np.full(ens.nbx, np.inf)
    np.minimum.at(per_run, blk, dx / s)
    return ens.cfl * per_run[blk], ens.cfl * per_run


def cell_centres(xlower, ylower, dx, dy, mx, my, mbc):
    """Cell-centre coordinates including ghost cells: arrays (npatch, ny, nx)."""
    i = np.arange(1 - mbc, mx + mbc + 1)
    j = np.arange(1 - mbc, my + mbc + 1)
    x = xlower[:, None, None] + (i[None, None, :] - 0.5) * dx[:, None, None]
    y = ylower[:, None, None] + (j[None, :, None] - 0.5) * dy[:, None, None]
    return np.broadcast_arrays(x, y)


def initial_data(prob, forest, mbc=2):
    xl, yl, dx, dy = forest.geometry(prob.extent, prob.mx, prob.my)
    x, y = cell_centres(xl, yl, dx, dy, prob.mx, prob.my, mbc)
    q = np.ascontiguousarray(np.moveaxis(prob.qfun(x, y), 0, 1))  # (npatch, meqn, ny, nx)
    aux = None
    if prob.auxfun is not None:
        aux = np.ascontiguousarray(np.moveaxis(prob.auxfun(x, y), 0, 1))
    return q, aux


def tag_gradient(prob, forest, q, mbc=2):
    """Patches where the centred gradient of the tag field exceeds the
    threshold in some interior cell (fclaw2d_clawpatch_gradient_exceeds_th)."""
    _, _, dx, dy = forest.geometry(prob.extent, prob.mx, prob.my)
    f = q[:, prob.tag_field]
    dqx = (f[:, mbc:-mbc, mbc + 1:f.shape[2] - mbc + 1] - f[:, mbc:-mbc, mbc - 1:f.shape[2] - mbc - 1]) \
        / (2 * dx[:, None, None])
    dqy = (f[:, mbc + 1:f.shape[1] - mbc + 1, mbc:-mbc] - f[:, mbc - 1:f.shape[1] - mbc - 1, mbc:-mbc]) \
        / (2 * dy[:, None, None])
    g = np.sqrt(dqx ** 2 + dqy ** 2)
    return (g.reshape(len(g), -1).max(axis=1) > prob.tag_threshold).astype(np.int32)


def build_forest(prob, mbc=2):
    """Initial refinement: tag on the initial data and refine from minlevel to
    maxlevel, keeping the forest 2:1 balanced (initial regrid of
    fclaw2d_initialize)."""
    forest = Forest(prob.nbx, prob.nby, prob.periodic, prob.minlevel)
    for _ in range(prob.maxlevel - prob.minlevel):
        q, _ = initial_data(prob, forest, mbc)
        flags = tag_gradient(prob, forest, q, mbc)
        if not flags.any():
            break
        forest.refine(flags, prob.maxlevel)
    return forest


def stable_dt(prob, forest, q, mbc=2):
    """dt giving Courant number prob.cfl on the finest patches for the initial
    data (the reference sets initial_dt; here it is derived from the data)."""
    _, _, dx, dy = forest.geometry(prob.extent, prob.mx, prob.my)
    dx = np.minimum(dx, dy)
    inner = q[:, :, mbc:-mbc, mbc:-mbc]
    if prob.solver == RP_SHALLOW:
        h = inner[:, 0]
        s = (np.abs(inner[:, 1] / h) + np.sqrt(prob.grav * h)).max(axis=(1, 2))
        s = np.maximum(s, (np.abs(inner[:, 2] / h) + np.sqrt(prob.grav * h)).max(axis=(1, 2)))
    elif prob.solver == RP_EULER:
        rho = inner[:, 0]
        u, v = inner[:, 1] / rho, inner[:, 2] / rho
        p = (prob.gamma - 1) * (inner[:, 3] - 0.5 * rho * (u * u + v * v))
        c = np.sqrt(prob.gamma * p / rho)
        s = np.maximum(np.abs(u) + c, np.abs(v) + c).max(axis=(1, 2))
    else:
        s = np.full(len(dx), prob.max_speed if prob.max_speed is not None else 1.0)
    return prob.cfl * float(np.min(dx / s))

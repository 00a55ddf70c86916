"""Shared test helpers: random admissible patch states, forest setups and
oracle runs on the same inputs as the GPU path."""
import ctypes as C

import numpy as np

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from oracle_lib import domain_struct, make_claw, oracle


def random_state(solver, npatch, mx, my, rng, mbc=2, gamma=1.4, rough=True):
    """Admissible states with jumps, shocks and transonic regions."""
    shape = (npatch, my + 2 * mbc, mx + 2 * mbc)
    if solver == fa.RP_SHALLOW:
        h = 0.5 + 1.5 * rng.random(shape)
        u = rng.normal(0, 0.8, shape)
        v = rng.normal(0, 0.8, shape)
        q = np.stack([h, h * u, h * v], axis=1)
    elif solver == fa.RP_EULER:
        r = 0.5 + 1.5 * rng.random(shape)
        u = rng.normal(0, 1.0, shape)
        v = rng.normal(0, 1.0, shape)
        p = 0.5 + 1.5 * rng.random(shape)
        q = np.stack([r, r * u, r * v, p / (gamma - 1) + 0.5 * r * (u * u + v * v)], axis=1)
    else:
        q = rng.random(shape)[:, None]
    if not rough:
        q = q.mean(axis=(2, 3), keepdims=True) + 0.05 * (q - q.mean(axis=(2, 3), keepdims=True))
    return np.ascontiguousarray(q)


def random_aux(npatch, mx, my, rng, mbc=2, constant=False):
    shape = (npatch, 2, my + 2 * mbc, mx + 2 * mbc)
    if constant:
        aux = np.zeros(shape)
        aux[:, 0] = 1.0
        aux[:, 1] = 0.5
        return aux
    return rng.normal(0, 1, shape)


def claw_for(solver, mthlim=None, order=(2, 2), grav=1.0, gamma=1.4):
    meqn, mwaves, maux = fa.SOLVER_SHAPES[solver]
    return make_claw(solver, meqn, maux, mwaves, order=order, mthlim=mthlim, grav=grav,
                     gamma=gamma)


def prob_claw(prob, order=(2, 2)):
    """The oracle's method parameters for a problem (its solver, constants and
    limiters)."""
    meqn, mwaves, maux = fa.solver_shape(prob.solver)
    solver = prob.oracle_solver if prob.oracle_solver is not None else prob.solver
    return make_claw(solver, meqn, maux, mwaves, order=order, mthlim=prob.mthlim,
                     grav=prob.grav, gamma=prob.gamma, params=prob.rp_params)


def oracle_domain(prob, forest, mx):
    lv, gx, gy = forest.leaves()
    my = prob.my if mx == prob.mx else mx
    return domain_struct(lv, gx, gy, prob.nbx, prob.nby, prob.periodic, mx, my, 2, prob.meqn,
                         prob.maux, prob.mthbc, prob.extent)


def oracle_ghost_update(prob, forest, q, mx):
    d = oracle_domain(prob, forest, mx)
    out = np.ascontiguousarray(q, dtype=np.float64).copy()
    oracle().oracle_ghost_update(C.byref(d), out.reshape(-1))
    return out


def oracle_advance(prob, forest, q, aux, mx, dt, nsteps):
    d = oracle_domain(prob, forest, mx)
    claw = prob_claw(prob)
    out = np.ascontiguousarray(q, dtype=np.float64).copy()
    auxp = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    cfl = oracle().oracle_advance(C.byref(d), C.byref(claw), out.reshape(-1),
                                  None if auxp is None else auxp.ctypes.data, dt, nsteps)
    return out, cfl


def oracle_advance_ts(prob, forest, q, aux, mx, dt, nsteps, time_sync=True, order=(2, 2)):
    """nsteps of (update all levels, ghost update, time sync) on the oracle."""
    d = oracle_domain(prob, forest, mx)
    claw = prob_claw(prob, order=order)
    out = np.ascontiguousarray(q, dtype=np.float64).copy()
    auxp = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    cfl = oracle().oracle_advance_ts(C.byref(d), C.byref(claw), out.reshape(-1),
                                     None if auxp is None else auxp.ctypes.data, dt, nsteps,
                                     int(time_sync))
    return out, cfl


def oracle_advance_subcycle(prob, forest, q, aux, mx, dt, ncoarse=1, time_sync=True):
    """ncoarse calls of fclaw2d_advance_all_levels(dt) with subcycling on the oracle."""
    d = oracle_domain(prob, forest, mx)
    claw = prob_claw(prob)
    out = np.ascontiguousarray(q, dtype=np.float64).copy()
    auxp = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    cfl = 0.0
    for _ in range(ncoarse):
        cfl = max(cfl, oracle().oracle_advance_subcycle(
            C.byref(d), C.byref(claw), out.reshape(-1),
            None if auxp is None else auxp.ctypes.data, dt, int(time_sync)))
    return out, cfl


def oracle_mass(prob, forest, q, mx):
    d = oracle_domain(prob, forest, mx)
    s = np.zeros(prob.meqn)
    oracle().oracle_mass(C.byref(d), np.ascontiguousarray(q, dtype=np.float64).reshape(-1), s)
    return s


def interior(q, mbc=2):
    return q[..., mbc:-mbc, mbc:-mbc]


def setup_problem(prob, precision=fa.FP64, tile=0, time_sync=False):
    """Forest, initial data (ghost cells filled by the oracle) and a GPU solver."""
    forest = P.build_forest(prob)
    q0, aux = P.initial_data(prob, forest)
    q0 = oracle_ghost_update(prob, forest, q0, prob.mx)
    sv = fa.Solver(prob.solver, prob.mx, prob.my, precision=precision, mthbc=prob.mthbc,
                   grav=prob.grav, gamma=prob.gamma, tile=tile, max_patches=forest.npatch,
                   time_sync=time_sync, rp_params=prob.rp_params, mthlim=prob.mthlim)
    sv.set_forest(forest, prob.extent)
    sv.set_q(q0)
    if aux is not None:
        sv.set_aux(aux)
    return forest, q0, aux, sv


def rel_max_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

"""The workloads at their stated sizes, on the GPU against the oracle.

* Shallow-water dam break on the AMR forest at levels 3-8 with 8x8 and 16x16
  patches (the patch sizes the autotuner sweeps besides 32x32): FP64 steps
  with time sync bit-for-bit against the oracle (tolerance 1e-12 relative,
  max norm), and the FP32-bulk form within 1e-5 relative L1 per field of the
  FP64 form, with the total of every field unchanged to 1e-12 relative.
* The ensemble of 64 independent dam-break runs (radius U[0.1, 0.3], inner
  height U[1.5, 3.0], maxlevel U{4..7}, 16x16 patches, one run per isolated
  block, each at its own dt): every run's FP64 steps against the oracle
  advancing that run alone, and every run's FP32-bulk state within 1e-5
  relative L1 of its FP64 state after 20 steps, with each run's mass kept.
* The method orders of Clawpack's method(2:3) that the CPU path allows besides
  (2, 2): (2, 1), (1, 2), (2, 0), (1, 0) -- the order2 / transverse branches of
  the update kernel -- for all three Riemann solvers, with a tile that covers
  the patch (in place) and with one that does not (two slabs).
"""
import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import (claw_for, interior, oracle_advance_ts, oracle_ghost_update, random_aux,
                     random_state, rel_max_err, setup_problem)
from oracle_lib import step2_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _l1_per_field(a, b, area):
    """Relative L1 difference of interior cells per field, area-weighted."""
    a, b = interior(a), interior(b)
    w = area[:, None, None, None]
    return (np.abs(a - b) * w).sum(axis=(0, 2, 3)) / (np.abs(b) * w).sum(axis=(0, 2, 3))


def _totals(q, area):
    return (interior(q) * area[:, None, None, None]).sum(axis=(0, 2, 3))


@pytest.mark.parametrize("mx", [8, 16])
def test_dambreak_l3_8_small_patches(gpu, mx):
    prob = P.dambreak_problem(3, 8, mx=mx)
    forest, q0, aux, s64 = setup_problem(prob, precision=fa.FP64, time_sync=True)
    _, _, dx, dy = forest.geometry(prob.extent, mx, mx)
    area = dx * dy
    dt = P.stable_dt(prob, forest, q0)
    print("L3-8 mx=%d: %d patches, levels %d-%d, dt %.3e" % (
        mx, forest.npatch, forest.minlevel, forest.maxlevel, dt), flush=True)
    s64.advance(dt, 3)
    q64 = s64.get_q()
    qo, _ = oracle_advance_ts(prob, forest, q0, aux, mx, dt, 3, time_sync=True)
    err = rel_max_err(q64, qo)
    assert err <= TOL, "FP64 vs oracle: %.3e (bit-exact %d of %d)" % (
        err, int(np.sum(q64 == qo)), qo.size)
    s64.advance(dt, 17)
    q64 = s64.get_q()
    _, _, _, s32 = setup_problem(prob, precision=fa.FP32_BULK, time_sync=True)
    s32.advance(dt, 20)
    q32 = s32.get_q()
    l1 = _l1_per_field(q32, q64, area)
    t0, t32 = _totals(q0, area), _totals(q32, area)
    print("  FP32-bulk L1 vs FP64 per field %s, mass change %.2e" % (
        np.array2string(l1, precision=2), abs(t32[0] - t0[0]) / abs(t0[0])), flush=True)
    assert np.all(l1 <= 1e-5), l1
    assert abs(t32[0] - t0[0]) <= 1e-12 * abs(t0[0])


def _ensemble(nruns, precision, maxlevels=(4, 7), seed=2026):
    ens = P.ensemble_problem(nruns, mx=16, minlevel=2, maxlevels=maxlevels, seed=seed)
    forest = P.build_ensemble_forest(ens)
    q0, _ = P.initial_data(ens, forest)
    members = []
    for b in range(nruns):
        lv, gx, gy, idx = P.member_leaves(forest, b)
        fb = fa.Forest(1, 1, (False, False), leaves=(lv, gx, gy))
        mp = P.member_problem(ens, b)
        q0[idx] = oracle_ghost_update(mp, fb, q0[idx], ens.mx)
        members.append((mp, fb, idx))
    dt_slot, dt_run = P.ensemble_dt(ens, forest, q0)
    sv = fa.Solver(ens.solver, ens.mx, ens.mx, precision=precision, mthbc=ens.mthbc,
                   grav=ens.grav, max_patches=forest.npatch, time_sync=True)
    sv.set_forest(forest, ens.extent)
    sv.set_q(q0)
    sv.set_dt_scale(dt_slot)
    return ens, forest, q0, sv, members, dt_run


def test_ensemble_64_runs_at_size(gpu):
    ens, forest, q0, s64, members, dt_run = _ensemble(64, fa.FP64)
    mlev = ens.params["maxlevel"]
    print("ensemble: %d runs, %d patches (%d..%d per run), maxlevel %d..%d" % (
        len(members), forest.npatch, min(len(m[2]) for m in members),
        max(len(m[2]) for m in members), mlev.min(), mlev.max()), flush=True)
    assert set(mlev.tolist()) == {4, 5, 6, 7}
    s64.advance(1.0, 2)
    q64 = s64.get_q()
    worst = 0.0
    for b, (mp, fb, idx) in enumerate(members):
        qo, _ = oracle_advance_ts(mp, fb, q0[idx], None, ens.mx, float(dt_run[b]), 2,
                                  time_sync=True)
        err = rel_max_err(q64[idx], qo)
        worst = max(worst, err)
        assert err <= TOL, (b, err)
    print("  FP64, 2 steps: worst run error vs the oracle %.2e" % worst, flush=True)
    s64.advance(1.0, 18)
    q64 = s64.get_q()
    s32 = _ensemble(64, fa.FP32_BULK)[3]
    s32.advance(1.0, 20)
    q32 = s32.get_q()
    _, _, dx, dy = forest.geometry(ens.extent, ens.mx, ens.mx)
    area = dx * dy
    l1max = np.zeros(3)
    for b, (_, _, idx) in enumerate(members):
        l1 = _l1_per_field(q32[idx], q64[idx], area[idx])
        l1max = np.maximum(l1max, l1)
        assert np.all(l1 <= 1e-5), (b, l1)
        t0, t32 = _totals(q0[idx], area[idx]), _totals(q32[idx], area[idx])
        assert abs(t32[0] - t0[0]) <= 1e-12 * abs(t0[0]), b
    print("  FP32-bulk, 20 steps: largest per-run L1 vs FP64 %s" % np.array2string(
        l1max, precision=2), flush=True)


ORDERS = [(2, 1), (1, 2), (2, 0), (1, 0)]


@pytest.mark.parametrize("solver", [fa.RP_ADVECTION, fa.RP_SHALLOW, fa.RP_EULER])
@pytest.mark.parametrize("order", ORDERS)
def test_method_orders_fp64(gpu, solver, order):
    """One batched step of random patches (random ghost cells) with method
    order `order` against oracle_step2_patch (itself bit-equal to the compiled
    clawpack46_step2 for these orders); tile 16 covers a 16x16 patch (in
    place), tile 8 does not (two slabs)."""
    meqn, mwaves, _ = fa.SOLVER_SHAPES[solver]
    for tile in (16, 8):
        rng = np.random.default_rng(500 + 10 * solver + order[0] * 3 + order[1] + tile)
        mx, npatch = 16, 4
        q = random_state(solver, npatch, mx, mx, rng)
        aux = random_aux(npatch, mx, mx, rng) if solver == fa.RP_ADVECTION else None
        sv = fa.Solver(solver, mx, mx, order=order, max_patches=npatch, tile=tile)
        assert sv.tile == tile
        dx = np.full(npatch, 1.0 / mx)
        sv.set_geometry(dx, dx)
        sv.set_ghost_program(np.zeros((0, 8), np.int32), np.zeros(7, np.int32))
        sv.set_q(q)
        if aux is not None:
            sv.set_aux(aux)
        dt = 0.25 / mx
        cfl = sv.step_batch(np.arange(npatch, dtype=np.int32), dt)
        qg = sv.get_q()
        claw = claw_for(solver, order=order)
        cfl_o = 0.0
        for p in range(npatch):
            qo, c = step2_oracle(claw, q[p], None if aux is None else aux[p], dx[p], dx[p], dt)
            cfl_o = max(cfl_o, c)
            err = rel_max_err(interior(qg[p]), interior(qo))
            assert err <= TOL, "order %s tile %d patch %d: %.3e (bit-exact %d of %d)" % (
                order, tile, p, err, int(np.sum(interior(qg[p]) == interior(qo))),
                interior(qo).size)
        assert abs(cfl - cfl_o) <= TOL * cfl_o


@pytest.mark.parametrize("order", ORDERS)
def test_method_orders_amr_with_time_sync(gpu, order):
    """Method orders through the whole step on an AMR forest (ghost fill, time
    sync): FP64 against the oracle, FP32-bulk within 1e-5 relative L1."""
    prob = P.dambreak_problem(2, 5, mx=16)
    forest = P.build_forest(prob)
    q0, aux = P.initial_data(prob, forest)
    q0 = oracle_ghost_update(prob, forest, q0, prob.mx)
    # without transverse terms the unsplit method is stable only to Courant
    # number 1/2 (at 0.9 the oracle itself reaches negative depths by step 6)
    dt = P.stable_dt(prob, forest, q0) * (1.0 if order[1] > 0 else 0.5)
    out = {}
    for precision in (fa.FP64, fa.FP32_BULK):
        sv = fa.Solver(prob.solver, 16, 16, precision=precision, order=order, mthbc=prob.mthbc,
                       max_patches=forest.npatch, time_sync=True, tile=8)
        sv.set_forest(forest, prob.extent)
        sv.set_q(q0)
        sv.advance(dt, 6)
        out[precision] = sv.get_q()
    qo, _ = oracle_advance_ts(prob, forest, q0, aux, 16, dt, 6, time_sync=True, order=order)
    assert np.isfinite(qo).all()
    assert rel_max_err(out[fa.FP64], qo) <= TOL
    _, _, dx, dy = forest.geometry(prob.extent, 16, 16)
    l1 = _l1_per_field(out[fa.FP32_BULK], qo, dx * dy)
    assert np.all(l1 <= 1e-5), l1


def test_ghost_program_refuses_unsupported_physical_bc(gpu):
    """A physical face with mthbc 2 (periodic) or 0 (user) is refused when the
    program is installed: the reference stops there
    (clawpack46_bc2_default.f:22-25, 41-43)."""
    prob = P.dambreak_problem(1, 2, mx=8)
    forest = P.build_forest(prob)
    for bad in ((2, 1, 1, 1), (1, 1, 0, 1)):
        sv = fa.Solver(prob.solver, 8, 8, mthbc=bad, max_patches=forest.npatch)
        with pytest.raises(fa.HipclawError, match="boundary condition"):
            sv.set_forest(forest, prob.extent)

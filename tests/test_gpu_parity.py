"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same inputs.

FP64 form: every case is checked to 1e-12 relative in max norm (the stated
tolerance) and the CFL numbers must agree to the same tolerance; the
arithmetic is written to be bit-identical, and the bit-exact count is
reported in the assertion messages.
"""
import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import (claw_for, interior, oracle_advance, oracle_ghost_update, random_aux,
                     random_state, rel_max_err, setup_problem)
from oracle_lib import step2_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12
SOLVERS = [fa.RP_ADVECTION, fa.RP_SHALLOW, fa.RP_EULER]


def _tiles(solver):
    out = [(8, 8), (16, 16), (16, 8), (32, 16)]
    if solver == fa.RP_SHALLOW:
        out.append((32, 32))
    return out


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("limiter", [1, 2, 3, 4, 5, 6])
def test_patch_step_fp64_all_limiters(gpu, solver, limiter):
    """One batched step of random patches (random ghost cells) vs oracle_step2_patch."""
    rng = np.random.default_rng(100 + 10 * solver + limiter)
    mx = 16
    npatch = 5
    q = random_state(solver, npatch, mx, mx, rng)
    aux = random_aux(npatch, mx, mx, rng) if solver == fa.RP_ADVECTION else None
    meqn, mwaves, _ = fa.SOLVER_SHAPES[solver]
    sv = fa.Solver(solver, mx, mx, mthlim=[limiter] * mwaves, max_patches=npatch, tile=16)
    dx = np.full(npatch, 1.0 / mx)
    sv.set_geometry(dx, dx)
    sv.set_ghost_program(np.zeros((0, 8), np.int32), np.zeros(7, np.int32))
    sv.set_q(q)
    if aux is not None:
        sv.set_aux(aux)
    dt = 0.2 / mx
    slots = np.arange(npatch, dtype=np.int32)
    cfl = sv.step_batch(slots, dt)
    qg = sv.get_q()
    claw = claw_for(solver, mthlim=[limiter] * mwaves)
    cfl_o = 0.0
    for p in range(npatch):
        qo, c = step2_oracle(claw, q[p], None if aux is None else aux[p], dx[p], dx[p], dt)
        cfl_o = max(cfl_o, c)
        err = rel_max_err(interior(qg[p]), interior(qo))
        exact = int(np.sum(interior(qg[p]) == interior(qo)))
        assert err <= TOL, "patch %d: rel err %.3e (bit-exact %d of %d)" % (
            p, err, exact, interior(qo).size)
    assert abs(cfl - cfl_o) <= TOL * cfl_o, (cfl, cfl_o)


@pytest.mark.parametrize("solver", SOLVERS)
def test_patch_step_fp64_tiles(gpu, solver):
    """Every tile size gives the oracle's result (tiles recompute their halo)."""
    for mx, tile in _tiles(solver):
        rng = np.random.default_rng(7 + mx + tile + solver)
        npatch = 3
        q = random_state(solver, npatch, mx, mx, rng)
        aux = random_aux(npatch, mx, mx, rng) if solver == fa.RP_ADVECTION else None
        sv = fa.Solver(solver, mx, mx, max_patches=npatch, tile=tile)
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
        claw = claw_for(solver)
        cfl_o = 0.0
        for p in range(npatch):
            qo, c = step2_oracle(claw, q[p], None if aux is None else aux[p], dx[p], dx[p], dt)
            cfl_o = max(cfl_o, c)
            err = rel_max_err(interior(qg[p]), interior(qo))
            assert err <= TOL, "mx=%d tile=%d patch %d: %.3e" % (mx, tile, p, err)
            # ghost cells are not touched by the device step
            np.testing.assert_array_equal(qg[p][:, :2, :], q[p][:, :2, :])
        assert abs(cfl - cfl_o) <= TOL * cfl_o


def _small_problems():
    return [P.dambreak_problem(2, 4, mx=8), P.advection_problem(2, 4, mx=8),
            P.shockbubble_problem(1, 3, mx=8)]


@pytest.mark.parametrize("layout", [0, 1, 2])
@pytest.mark.parametrize("k", [0, 1, 2])
def test_ghost_fill_fp64(gpu, k, layout):
    """Device ghost fill (copy/average/interpolate/physical BC) vs the oracle,
    in both launch layouts."""
    prob = _small_problems()[k]
    forest, q0, aux, sv = setup_problem(prob)
    sv.set_ghost_layout(layout)
    rng = np.random.default_rng(3 + k)
    q = random_state(prob.solver, forest.npatch, prob.mx, prob.mx, rng)
    sv.set_q(q)
    sv.ghost_fill()
    qg = sv.get_q()
    qo = oracle_ghost_update(prob, forest, q, prob.mx)
    np.testing.assert_array_equal(qg, qo)


@pytest.mark.parametrize("layout", [0, 1, 2])
@pytest.mark.parametrize("k", [0, 1, 2])
def test_advance_fp64_amr(gpu, k, layout):
    """Ten steps of (update all levels, ghost fill) on an AMR forest, in both
    ghost-fill layouts."""
    prob = _small_problems()[k]
    forest, q0, aux, sv = setup_problem(prob)
    sv.set_ghost_layout(layout)
    dt = P.stable_dt(prob, forest, q0) * 0.9
    nsteps = 10
    cfl = sv.advance(dt, nsteps)
    qg = sv.get_q()
    qo, cfl_o = oracle_advance(prob, forest, q0, aux, prob.mx, dt, nsteps)
    err = rel_max_err(qg, qo)
    exact = int(np.sum(qg == qo))
    assert err <= TOL, "%s: rel err %.3e (bit-exact %d of %d)" % (prob.name, err, exact, qo.size)
    assert abs(cfl - cfl_o) <= TOL * cfl_o, (cfl, cfl_o)


def test_step_batch_per_level(gpu):
    """The drop-in path: one batch per level (fclaw2d_update_single_step) then a
    ghost fill equals the oracle's all-level step."""
    prob = P.dambreak_problem(2, 4, mx=16)
    forest, q0, aux, sv = setup_problem(prob, tile=8)
    lv = forest.leaves()[0]
    dt = P.stable_dt(prob, forest, q0) * 0.9
    cfl = 0.0
    for level in range(lv.max(), lv.min() - 1, -1):
        slots = np.nonzero(lv == level)[0].astype(np.int32)
        cfl = max(cfl, sv.step_batch(slots, dt))
    sv.ghost_fill()
    qg = sv.get_q()
    qo, cfl_o = oracle_advance(prob, forest, q0, aux, prob.mx, dt, 1)
    assert rel_max_err(qg, qo) <= TOL
    assert abs(cfl - cfl_o) <= TOL * cfl_o


@pytest.mark.parametrize("layout", [0, 1, 2])
def test_advance_two_slabs(gpu, layout):
    """Two slabs (16x16 patches, tile 8: the update writes the other slab)
    with either ghost-fill layout, bit-identical to the oracle with and
    without time sync."""
    from helpers import oracle_advance_ts
    prob = P.dambreak_problem(2, 5, mx=16)
    for time_sync in (False, True):
        forest, q0, aux, sv = setup_problem(prob, tile=8, time_sync=time_sync)
        sv.set_ghost_layout(layout)
        dt = P.stable_dt(prob, forest, q0) * 0.9
        cfl = sv.advance(dt, 6)
        qg = sv.get_q()
        qo, cfl_o = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 6, time_sync=time_sync)
        err = rel_max_err(qg, qo)
        assert err <= TOL, "time_sync %d: %.3e (bit-exact %d of %d)" % (time_sync, err,
                                                                        int(np.sum(qg == qo)), qo.size)
        assert abs(cfl - cfl_o) <= TOL * cfl_o


@pytest.mark.parametrize("solver_case", ["shallow", "euler"])
def test_waves_per_workgroup(gpu, solver_case):
    """16 x 16 tiles with 2, 3 and 4 waves per workgroup: FP64 bit-identical to the
    oracle with time sync, and the FP32-bulk form the same bits either way
    (the waves only change which lanes do the same arithmetic)."""
    from helpers import oracle_advance_ts
    prob = P.dambreak_problem(2, 5, mx=16) if solver_case == "shallow" else P.shockbubble_problem(1, 4, mx=16)
    results = {}
    for precision in (fa.FP64, fa.FP32_BULK):
        for waves in (2, 3, 4):
            forest, q0, aux, sv = setup_problem(prob, precision=precision, tile=16, time_sync=True)
            sv.set_waves(waves)
            assert sv.waves == waves
            dt = P.stable_dt(prob, forest, q0) * 0.9
            sv.advance(dt, 6)
            results[(precision, waves)] = sv.get_q()
        if precision == fa.FP64:
            qo, _ = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 6, time_sync=True)
            for waves in (2, 3, 4):
                assert rel_max_err(results[(precision, waves)], qo) <= TOL, waves
        else:
            np.testing.assert_array_equal(results[(precision, 2)], results[(precision, 4)])
            np.testing.assert_array_equal(results[(precision, 2)], results[(precision, 3)])

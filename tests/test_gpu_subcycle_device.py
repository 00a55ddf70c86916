"""The subcycled advance enqueued on the device as one sequence
(hipclaw_advance_subcycle: per-level updates with the state saved before each,
time-interpolated coarse copies, partial ghost fills and time syncs between
the finest steps, src/fclaw2d_advance.c:54-246), eagerly and as a captured
graph, against the CPU oracle's subcycled advance.  FP64: 1e-12 relative in
max norm (bit-exact arithmetic).  FP32-bulk: 1e-4 relative, mass conserved to
FP64 rounding."""
import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import (interior, oracle_advance_subcycle, oracle_ghost_update, oracle_mass,
                     rel_max_err)

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _setup(prob, time_sync, precision=fa.FP64, tile=0):
    forest = P.build_forest(prob)
    q0, aux = P.initial_data(prob, forest)
    q0 = oracle_ghost_update(prob, forest, q0, prob.mx)
    sv = fa.Solver(prob.solver, prob.mx, prob.my, precision=precision, mthbc=prob.mthbc,
                   grav=prob.grav, gamma=prob.gamma, tile=tile, max_patches=forest.npatch,
                   time_sync=time_sync, subcycle=True)
    sv.set_forest(forest, prob.extent)
    sv.set_q(q0)
    if aux is not None:
        sv.set_aux(aux)
    lv = forest.leaves()[0]
    dt = P.stable_dt(prob, forest, q0) * 2 ** (int(lv.max()) - int(lv.min()))
    return forest, q0, aux, sv, dt


@pytest.mark.parametrize("time_sync", [False, True])
@pytest.mark.parametrize("name", ["dambreak", "shockbubble"])
def test_device_subcycle_matches_oracle(gpu, time_sync, name):
    prob = P.dambreak_problem(2, 5, mx=16) if name == "dambreak" else P.shockbubble_problem(1, 3, mx=16)
    print("case", name, "time_sync", time_sync, flush=True)
    forest, q0, aux, sv, dt = _setup(prob, time_sync)
    assert forest.maxlevel - forest.minlevel >= 2
    cfl = sv.advance_subcycle(dt, 2)
    qg = sv.get_q()
    qo, cfl_o = oracle_advance_subcycle(prob, forest, q0, aux, prob.mx, dt, 2, time_sync=time_sync)
    err = rel_max_err(interior(qg), interior(qo))
    exact = int(np.sum(interior(qg) == interior(qo)))
    assert err <= TOL, "rel err %.3e (bit-exact %d of %d)" % (err, exact, interior(qo).size)
    assert abs(cfl - cfl_o) <= TOL * cfl_o


@pytest.mark.parametrize("time_sync", [False, True])
def test_subcycle_graph_replays_match_oracle(gpu, time_sync):
    """One captured coarse step replayed three times, with eager host work
    and synchronisations between the replays, each replay against the oracle."""
    prob = P.dambreak_problem(2, 5, mx=16)
    forest, q0, aux, sv, dt = _setup(prob, time_sync)
    sv.capture_subcycle_graph(dt, 1)
    qo = q0
    for r in range(3):
        sv.replay_graph()
        sv.sync()
        qg = sv.get_q()
        qo, _ = oracle_advance_subcycle(prob, forest, qo, aux, prob.mx, dt, 1, time_sync=time_sync)
        err = rel_max_err(interior(qg), interior(qo))
        assert err <= TOL, "replay %d: rel err %.3e" % (r, err)
        _ = sv.mass()   # eager work on the same stream between replays


def test_autotune_on_a_subcycling_state(gpu):
    """The tuner times subcycled coarse steps; the state it leaves stays
    bit-faithful and its registers clear."""
    prob = P.dambreak_problem(2, 5, mx=16)
    forest, q0, aux, sv, dt = _setup(prob, True)
    trials = sv.autotune(dt, 1)
    assert len(trials) >= 3
    cfl = sv.advance_subcycle(dt, 1)
    qo, cfl_o = oracle_advance_subcycle(prob, forest, q0, aux, prob.mx, dt, 1, time_sync=True)
    assert rel_max_err(interior(sv.get_q()), interior(qo)) <= TOL
    assert abs(cfl - cfl_o) <= TOL * cfl_o


def test_fp32_subcycle_conserves(gpu):
    prob = P.dambreak_problem(2, 5, mx=16)
    forest, q0, aux, sv, dt = _setup(prob, True, precision=fa.FP32_BULK)
    m0 = oracle_mass(prob, forest, q0, prob.mx)
    sv.advance_subcycle(dt, 3)
    qg = sv.get_q()
    m1 = oracle_mass(prob, forest, qg, prob.mx)
    qo, _ = oracle_advance_subcycle(prob, forest, q0, aux, prob.mx, dt, 3, time_sync=True)
    assert rel_max_err(interior(qg), interior(qo)) <= 1e-4
    assert abs(m1[0] - m0[0]) <= 1e-13 * abs(m0[0])


def test_subcycle_needs_levels_and_programs(gpu):
    prob = P.dambreak_problem(2, 4, mx=16)
    forest = P.build_forest(prob)
    sv = fa.Solver(prob.solver, 16, 16, max_patches=forest.npatch, subcycle=True)
    _, _, dx, dy = forest.geometry(prob.extent, 16, 16)
    sv.set_geometry(dx, dy)
    ops, ps = forest.ghost_program(16, 16, 2, prob.mthbc)
    sv.set_ghost_program(ops, ps)
    with pytest.raises(fa.HipclawError, match="levels"):
        sv.advance_subcycle(1e-3, 1)
    plain = fa.Solver(prob.solver, 16, 16, max_patches=forest.npatch)
    plain.set_forest(forest, prob.extent)
    with pytest.raises(fa.HipclawError, match="subcycle"):
        plain.advance_subcycle(1e-3, 1)

"""Patch shapes other than square multiples of the tile, against the CPU
oracle.  The reference takes any mx x my in its batched update
(src/solvers/fc2d_cudaclaw/cuda_source/cudaclaw_flux2.cu:278-281); here a
tile that does not divide the patch overlaps the patch's last tile with the
one before it.

FP64: 1e-12 relative in max norm (bit-exact arithmetic; the exact count is
reported).  FP32-bulk: 1e-4 relative in max norm against the FP64 oracle.
"""
import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import (claw_for, interior, oracle_advance, oracle_advance_ts, random_aux,
                     random_state, rel_max_err, setup_problem)
from oracle_lib import step2_oracle

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-12, 1e-4
SHAPES = [(20, 20), (24, 24), (12, 10), (20, 12), (40, 24), (9, 17)]


@pytest.mark.parametrize("solver", [fa.RP_ADVECTION, fa.RP_SHALLOW, fa.RP_EULER])
@pytest.mark.parametrize("shape", SHAPES, ids=["%dx%d" % s for s in SHAPES])
@pytest.mark.parametrize("precision", [fa.FP64, fa.FP32_BULK])
def test_patch_step_any_shape(gpu, solver, shape, precision):
    """One batched step of random patches of mx x my cells, every tile that
    fits, vs oracle_step2_patch."""
    mx, my = shape
    rng = np.random.default_rng(7 * mx + my + 100 * solver + precision)
    npatch = 4
    q = random_state(solver, npatch, mx, my, rng)
    aux = random_aux(npatch, mx, my, rng) if solver == fa.RP_ADVECTION else None
    dx, dy = np.full(npatch, 1.0 / mx), np.full(npatch, 1.0 / my)
    dt = 0.2 / max(mx, my)
    claw = claw_for(solver)
    ref = [step2_oracle(claw, q[p], None if aux is None else aux[p], dx[p], dy[p], dt)
           for p in range(npatch)]
    cfl_o = max(c for _, c in ref)
    for tile in (8, 16):
        if tile > min(mx, my):
            continue
        print("case solver", solver, "shape", shape, "tile", tile, "precision", precision, flush=True)
        sv = fa.Solver(solver, mx, my, precision=precision, max_patches=npatch, tile=tile)
        assert sv.tile == tile
        sv.set_geometry(dx, dy)
        sv.set_ghost_program(np.zeros((0, 8), np.int32), np.zeros(7, np.int32))
        sv.set_q(q)
        if aux is not None:
            sv.set_aux(aux)
        cfl = sv.step_batch(np.arange(npatch, dtype=np.int32), dt)
        qg = sv.get_q()
        tol = TOL64 if precision == fa.FP64 else TOL32
        for p in range(npatch):
            qo = ref[p][0]
            err = rel_max_err(interior(qg[p]), interior(qo))
            exact = int(np.sum(interior(qg[p]) == interior(qo)))
            assert err <= tol, "tile %d patch %d: rel err %.3e (bit-exact %d of %d)" % (
                tile, p, err, exact, interior(qo).size)
        assert abs(cfl - cfl_o) <= (TOL64 if precision == fa.FP64 else 1e-5) * cfl_o


def test_default_tile_for_other_sizes(gpu):
    """16 x 16 where it divides the patch, else the tile with fewer Riemann
    solves per patch."""
    for (mx, my), want in (((32, 32), 16), ((16, 16), 16), ((20, 20), 8), ((24, 24), 8),
                           ((48, 48), 16), ((12, 10), 8), ((40, 24), 8)):
        assert fa.Solver(fa.RP_SHALLOW, mx, my).tile == want, (mx, my)


@pytest.mark.parametrize("mx", [18, 20])
@pytest.mark.parametrize("layout", [0, 1, 2])
@pytest.mark.parametrize("tile", [8, 16])
def test_amr_time_sync_mx20(gpu, layout, tile, mx):
    """AMR dam break with 20 x 20 and 18 x 18 patches (neither tile divides
    them), time sync on, every ghost layout, against the oracle's step / ghost
    fill / sync sequence.  With mx = 18 the overlapped last tile leaves the
    tile before it two cells from the patch's edge, so that tile stages the
    edge cells' deferred time-sync increments too (layout 2)."""
    prob = P.dambreak_problem(2, 4, mx=mx)
    print("case mx", mx, "layout", layout, "tile", tile, flush=True)
    forest, q0, aux, sv = setup_problem(prob, tile=tile, time_sync=True)
    assert forest.maxlevel > forest.minlevel
    sv.set_ghost_layout(layout)
    dt = P.stable_dt(prob, forest, q0)
    nsteps = 5
    cfl = sv.advance(dt, nsteps)
    qg = sv.get_q()
    qo, cfl_o = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, nsteps)
    err = rel_max_err(interior(qg), interior(qo))
    assert err <= TOL64, "rel err %.3e" % err
    assert abs(cfl - cfl_o) <= TOL64 * cfl_o


@pytest.mark.parametrize("precision", [fa.FP64, fa.FP32_BULK])
def test_uniform_forest_mx_ne_my(gpu, precision):
    """A uniform forest of 24 x 12 patches (ghost fill: copies and physical
    BC only), several steps against the oracle."""
    prob = P.dambreak_problem(3, 3, mx=24, my=12)
    print("case 24 x 12 precision", precision, flush=True)
    forest, q0, aux, sv = setup_problem(prob, precision=precision)
    assert forest.npatch == 64
    dt = P.stable_dt(prob, forest, q0)
    nsteps = 6
    cfl = sv.advance(dt, nsteps)
    qg = sv.get_q()
    qo, cfl_o = oracle_advance(prob, forest, q0, aux, prob.mx, dt, nsteps)
    err = rel_max_err(interior(qg), interior(qo))
    assert err <= (TOL64 if precision == fa.FP64 else TOL32), "rel err %.3e" % err
    assert abs(cfl - cfl_o) <= (TOL64 if precision == fa.FP64 else 1e-5) * cfl_o


@pytest.mark.parametrize("layout", [0, 2])
@pytest.mark.parametrize("tile", [8, 16])
def test_uniform_forest_odd_mx(gpu, tile, layout):
    """A uniform forest of 17 x 17 patches: the tile before the overlapped
    last one reads the patch's ghost column and row, which the fused layout
    (2) stages straight from the same-size neighbours."""
    prob = P.dambreak_problem(3, 3, mx=17)
    print("case 17 x 17 tile", tile, "layout", layout, flush=True)
    forest, q0, aux, sv = setup_problem(prob, tile=tile)
    sv.set_ghost_layout(layout)
    assert sv.tile == tile
    dt = P.stable_dt(prob, forest, q0)
    nsteps = 6
    cfl = sv.advance(dt, nsteps)
    qg = sv.get_q()
    qo, cfl_o = oracle_advance(prob, forest, q0, aux, prob.mx, dt, nsteps)
    err = rel_max_err(interior(qg), interior(qo))
    assert err <= TOL64, "rel err %.3e" % err
    assert abs(cfl - cfl_o) <= TOL64 * cfl_o


def test_coarse_fine_work_refused_when_mx_ne_my(gpu):
    """The reference's coarse/fine exchanges assume square patches; an AMR
    program, a refining regrid and time sync are refused for mx != my."""
    prob = P.dambreak_problem(2, 3, mx=24, my=16)
    forest = P.build_forest(prob)
    assert forest.maxlevel > forest.minlevel
    sv = fa.Solver(fa.RP_SHALLOW, 24, 16, max_patches=forest.npatch)
    ops, ps = forest.ghost_program(24, 16, 2, prob.mthbc)
    with pytest.raises(fa.HipclawError, match="mx = my"):
        sv.set_ghost_program(ops, ps)
    with pytest.raises(fa.HipclawError, match="mx = my"):
        fa.Solver(fa.RP_SHALLOW, 24, 16, time_sync=True)

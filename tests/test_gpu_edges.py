"""Edges of the batched step that the configurations above leave out.

  * a patch width that is a multiple of 8 but not of 16 (mx = 24: three 8x8
    tiles per row) on an AMR forest, FP64 with time sync, against the oracle;
  * a patch narrower than the smallest tile (mx = 6) is refused with a
    message (widths the tiles do not divide are covered in test_gpu_shapes);
  * a forest of one patch (level 0: every ghost cell comes from the physical
    boundary conditions), FP64 against the oracle and FP32-bulk within 1e-5;
  * an advance of zero steps leaves q unchanged and reports a Courant
    number of 0;
  * a ghost program with a bad op is refused and leaves the installed
    program (and the fused step's tables built from it) as it was, and a
    graph captured before stays valid.
"""
import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import oracle_advance, oracle_advance_ts, rel_max_err, setup_problem

pytestmark = pytest.mark.gpu


def test_mx24_amr_fp64_time_sync(gpu):
    prob = P.dambreak_problem(2, 4, mx=24)
    forest, q0, aux, sv = setup_problem(prob, time_sync=True)
    assert sv.tile == 8 and len(set(forest.leaves()[0].tolist())) > 1
    dt = P.stable_dt(prob, forest, q0)
    cfl = sv.advance(dt, 5)
    qo, cfl_o = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 5, time_sync=True)
    err = rel_max_err(sv.get_q(), qo)
    print("mx=24, %d patches: err %.3e, cfl %.6f" % (forest.npatch, err, cfl), flush=True)
    assert err <= 1e-12
    assert abs(cfl - cfl_o) <= 1e-12 * cfl_o


def test_patch_narrower_than_a_tile_is_refused(gpu):
    with pytest.raises(fa.HipclawError, match="at least the tile size"):
        fa.Solver(fa.RP_SHALLOW, 6, 6, precision=fa.FP64, mthbc=(1, 1, 1, 1), max_patches=4)
    with pytest.raises(fa.HipclawError, match="at least the tile size"):
        fa.Solver(fa.RP_SHALLOW, 32, 12, precision=fa.FP64, tile=16, max_patches=4)


@pytest.mark.parametrize("precision", [fa.FP64, fa.FP32_BULK])
def test_one_patch_forest(gpu, precision):
    prob = P.dambreak_problem(0, 0, mx=32)
    forest, q0, aux, sv = setup_problem(prob, precision=precision)
    assert forest.npatch == 1
    dt = P.stable_dt(prob, forest, q0)
    sv.advance(dt, 10)
    qo, _ = oracle_advance(prob, forest, q0, aux, prob.mx, dt, 10)
    a, b = sv.get_q()[..., 2:-2, 2:-2], qo[..., 2:-2, 2:-2]
    if precision == fa.FP64:
        err = rel_max_err(a, b)
        print("one patch FP64: err %.3e" % err, flush=True)
        assert err <= 1e-12
    else:
        l1 = np.abs(a - b).sum(axis=(0, 2, 3)) / np.maximum(np.abs(b).sum(axis=(0, 2, 3)),
                                                               np.abs(b[:, 0]).sum())
        print("one patch FP32-bulk: L1 per field", l1, flush=True)
        assert np.all(l1 <= 1e-5)


def test_zero_steps(gpu):
    prob = P.advection_problem(2, 2, mx=16)
    forest, q0, aux, sv = setup_problem(prob)
    before = sv.get_q()
    assert sv.advance(0.01, 0) == 0.0
    assert np.array_equal(sv.get_q(), before)


@pytest.mark.parametrize("precision", [fa.FP64, fa.FP32_BULK])
@pytest.mark.parametrize("layout", [1, 2])
def test_bad_ghost_program_leaves_the_installed_one(gpu, layout, precision):
    """Ops are checked on the host before anything of the state changes: a
    copy from a slot out of range, or one that reads outside its neighbour's
    interior, in the middle of the list, raises and the next steps run the
    program installed before, bit for bit as a state that never saw it."""
    prob = P.dambreak_problem(2, 4, mx=16)
    print("case layout", layout, "precision", precision, flush=True)
    forest, q0, aux, sv = setup_problem(prob, precision=precision, time_sync=True)
    _, _, _, ref = setup_problem(prob, precision=precision, time_sync=True)
    for s in (sv, ref):
        s.set_ghost_layout(layout)
    ops, ps = forest.ghost_program(prob.mx, prob.my, 2, prob.mthbc)
    assert ps[2] - ps[1] >= 2   # copies (phase 1)
    k = (ps[1] + ps[2]) // 2
    bad = ops.copy()
    bad[k, 2] = forest.npatch + 3
    with pytest.raises(fa.HipclawError, match="bad source"):
        sv.set_ghost_program(bad, ps)
    bad = ops.copy()
    bad[k, 4] += 4 * prob.mx
    with pytest.raises(fa.HipclawError, match="outside"):
        sv.set_ghost_program(bad, ps)
    dt = P.stable_dt(prob, forest, q0)
    c1 = sv.advance(dt, 4)
    c2 = ref.advance(dt, 4)
    np.testing.assert_array_equal(sv.get_q(), ref.get_q())
    assert c1 == c2
    # a refused program leaves a captured graph in place
    for s in (sv, ref):
        s.capture_graph(dt, 2)
    with pytest.raises(fa.HipclawError, match="bad source"):
        bad = ops.copy()
        bad[k, 2] = forest.npatch + 3
        sv.set_ghost_program(bad, ps)
    for s in (sv, ref):
        s.replay_graph()
        s.sync()
    np.testing.assert_array_equal(sv.get_q(), ref.get_q())

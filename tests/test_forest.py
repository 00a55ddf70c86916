"""Host forest and ghost-fill program (CPU only).

The product's op list, interpreted on the host with the device kernel's
per-op semantics (tests/ghost_emulator.py), must reproduce the oracle's
independent ghost fill bit for bit and overwrite every ghost cell.
"""
import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
import ghost_emulator as G
from helpers import oracle_ghost_update, random_state


def _neighbour_levels(forest):
    """Level pairs of all leaves that touch across a face or a corner
    (brute force: closed-interval overlap on the finest lattice, with the
    periodic images of the brick)."""
    lv, gx, gy = (a.astype(np.int64) for a in forest.leaves())
    maxl = int(lv.max())
    s = maxl - lv
    x0, x1 = gx << s, (gx + 1) << s
    y0, y1 = gy << s, (gy + 1) << s
    W, H = forest.nbx << maxl, forest.nby << maxl
    pairs = []
    shifts_x = (-W, 0, W) if forest.periodic[0] else (0,)
    shifts_y = (-H, 0, H) if forest.periodic[1] else (0,)
    for sx in shifts_x:
        for sy in shifts_y:
            touch = ((x0[:, None] <= x1[None, :] + sx) & (x0[None, :] + sx <= x1[:, None]) &
                     (y0[:, None] <= y1[None, :] + sy) & (y0[None, :] + sy <= y1[:, None]))
            if sx == 0 and sy == 0:
                np.fill_diagonal(touch, False)
            a, b = np.nonzero(touch)
            pairs.extend(zip(lv[a].tolist(), lv[b].tolist()))
    return pairs


@pytest.mark.parametrize("prob", [P.dambreak_problem(2, 6, mx=8), P.advection_problem(2, 5, mx=8),
                                  P.shockbubble_problem(1, 4, mx=8)], ids=lambda p: p.name)
def test_forest_is_balanced_and_tiles_the_domain(prob):
    forest = P.build_forest(prob)
    lv, gx, gy = forest.leaves()
    area = np.sum(np.ldexp(1.0, -2 * lv.astype(np.int64)))
    assert area == prob.nbx * prob.nby
    assert len(set(zip(lv.tolist(), gx.tolist(), gy.tolist()))) == len(lv)
    for l, m in _neighbour_levels(forest):
        assert abs(l - m) <= 1


def test_random_refinement_stays_balanced():
    rng = np.random.default_rng(0)
    f = fa.Forest(2, 1, (True, False), 1)
    for _ in range(5):
        flags = (rng.random(f.npatch) < 0.3).astype(np.int32)
        f.refine(flags, 6)
    for l, m in _neighbour_levels(f):
        assert abs(l - m) <= 1


def test_forest_from_leaves_rejects_gaps_and_overlaps():
    with pytest.raises(ValueError):
        fa.Forest(1, 1, (False, False), leaves=([1, 1, 1], [0, 1, 0], [0, 0, 1]))
    with pytest.raises(ValueError):
        fa.Forest(1, 1, (False, False), leaves=([0, 1, 1, 1, 1], [0, 0, 1, 0, 1], [0, 0, 0, 1, 1]))


@pytest.mark.parametrize("prob", [P.dambreak_problem(2, 4, mx=8), P.advection_problem(2, 4, mx=8),
                                  P.shockbubble_problem(1, 3, mx=8)], ids=lambda p: p.name)
def test_ghost_program_matches_oracle(prob):
    forest = P.build_forest(prob)
    rng = np.random.default_rng(1)
    q = random_state(prob.solver, forest.npatch, prob.mx, prob.mx, rng)
    ops, ps = forest.ghost_program(prob.mx, prob.mx, 2, prob.mthbc)
    qe = G.run(q.copy(), ops, ps, prob.mx, prob.mx, 2)
    qo = oracle_ghost_update(prob, forest, q, prob.mx)
    np.testing.assert_array_equal(qe, qo)
    # every ghost cell of every patch is written by the fill
    marked = np.full_like(q, np.nan)
    marked[:, :, 2:-2, 2:-2] = q[:, :, 2:-2, 2:-2]
    filled = G.run(marked, ops, ps, prob.mx, prob.mx, 2)
    assert np.isfinite(filled).all()


def test_ghost_program_phases_and_op_bounds():
    prob = P.dambreak_problem(2, 5, mx=16)
    forest = P.build_forest(prob)
    ops, ps = forest.ghost_program(16, 16, 2, prob.mthbc)
    assert ps[0] == 0 and ps[6] == len(ops) and np.all(np.diff(ps) >= 0)
    n = forest.npatch
    assert ops[:, 1].min() >= 0 and ops[:, 1].max() < n
    nonbc = ops[:, 0] != 6
    assert ops[nonbc, 2].min() >= 0 and ops[nonbc, 2].max() < n
    # op types live in their phase
    for ph, types in enumerate(({2, 3}, {0, 1}, {6}, {4, 5}, {6}, {6})):
        assert set(ops[ps[ph]:ps[ph + 1], 0].tolist()) <= types

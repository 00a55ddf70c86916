"""Saved state and retake (fclaw2d_save_time_step / fclaw2d_restore_time_step,
src/fclaw2d_run.c:47-80, clawpatch_save_step / clawpatch_restore_step).
With two slabs hipclaw_save_step keeps the saved state in its slab without
copying it, and copies it only before something may overwrite that slab:
every path must put back exactly the state that was saved (bit for bit), and
a one-step advance retaken from it must equal the same step taken once."""
import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import setup_problem

pytestmark = pytest.mark.gpu


def _solver(tile=16, time_sync=True):
    # (FP64 tiles of 32 keep no time-sync registers)
    time_sync = time_sync and tile != 32
    prob = P.dambreak_problem(2, 4, mx=32)
    forest, q0, aux, sv = setup_problem(prob, tile=tile, time_sync=time_sync)
    return prob, forest, q0, sv


@pytest.mark.parametrize("tile", [16, 32])
def test_restore_after_one_step_and_retake(gpu, tile):
    """save, one step, restore -> the saved state; retaking the step gives the
    bits of the same step taken without a retake (tile 32: one slab, copied)."""
    prob, forest, q0, sv = _solver(tile)
    dt = P.stable_dt(prob, forest, q0)
    sv.advance(dt, 1)
    base = sv.get_q()
    sv.save_step()
    sv.advance(2.0 * dt, 1)
    assert not np.array_equal(sv.get_q(), base)
    sv.restore_step()
    np.testing.assert_array_equal(sv.get_q(), base)
    sv.advance(dt, 1)
    retaken = sv.get_q()
    prob2, forest2, q02, ref = _solver(tile)
    ref.advance(dt, 1)
    ref.advance(dt, 1)
    np.testing.assert_array_equal(retaken, ref.get_q())


@pytest.mark.parametrize("what", ["two_steps", "two_advances", "ghost_fill", "set_q", "time_sync",
                                  "step_batch"])
def test_restore_after_writes_to_the_saved_slab(gpu, what):
    """Whatever runs between save and restore, the saved state comes back."""
    prob, forest, q0, sv = _solver(16)
    dt = P.stable_dt(prob, forest, q0)
    sv.advance(dt, 1)
    base = sv.get_q()
    print("case", what, flush=True)
    sv.save_step()
    if what == "two_steps":
        sv.advance(dt, 2)
    elif what == "two_advances":
        sv.advance(dt, 1)
        sv.advance(dt, 1)
    elif what == "ghost_fill":
        sv.ghost_fill()
        sv.advance(dt, 1)
    elif what == "set_q":
        sv.set_q(np.ascontiguousarray(base[::-1]))
    elif what == "time_sync":
        sv.advance(dt, 1)
        sv.time_sync_now()
    else:
        sv.step_batch(np.arange(sv.npatch, dtype=np.int32), dt)
    sv.restore_step()
    np.testing.assert_array_equal(sv.get_q(), base)


def test_graph_refuses_to_start_from_the_other_slab(gpu):
    """A restore that makes the other slab current again invalidates the
    captured graph's starting slab; the replay says so instead of running."""
    prob, forest, q0, sv = _solver(16)
    dt = P.stable_dt(prob, forest, q0)
    sv.capture_graph(dt, 2)
    sv.save_step()
    sv.advance(dt, 1)
    with pytest.raises(fa.HipclawError, match="other slab"):
        sv.replay_graph()
    sv.restore_step()
    sv.replay_graph()
    sv.sync()


def test_graph_dropped_when_the_programs_change(gpu):
    """A graph's launches hold the ghost program's tables and counts: after a
    program is installed again, replaying the old graph is refused."""
    prob, forest, q0, sv = _solver(16)
    dt = P.stable_dt(prob, forest, q0)
    sv.capture_graph(dt, 2)
    sv.replay_graph()
    sv.sync()
    sv.set_forest(forest, prob.extent)
    with pytest.raises(fa.HipclawError, match="no captured graph"):
        sv.replay_graph()
    sv.capture_graph(dt, 2)
    sv.replay_graph()
    sv.sync()

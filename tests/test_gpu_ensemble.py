"""An ensemble of independent dam-break runs advanced as one batch: one run
per block of a brick of isolated blocks (no ghost exchange across block
faces, physical boundary conditions on every block face), each run with its
own stable dt through the per-slot dt scale.

Every run must match the oracle advancing that run alone on the unit square
with its own dt (FP64: 1e-12 relative, max norm, time sync on); the FP32-bulk
form must stay within 1e-5 (relative L1) of it per run.
"""
import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import oracle_advance_ts, oracle_ghost_update, rel_max_err

pytestmark = pytest.mark.gpu


def _setup(nruns, precision):
    ens = P.ensemble_problem(nruns, mx=16, minlevel=2, maxlevels=(3, 5), seed=11)
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
    sv = fa.Solver(ens.solver, ens.mx, ens.mx, precision=precision, mthbc=ens.mthbc, grav=ens.grav,
                   max_patches=forest.npatch, time_sync=True)
    sv.set_forest(forest, ens.extent)
    sv.set_q(q0)
    sv.set_dt_scale(dt_slot)
    return ens, forest, q0, sv, members, dt_run


def test_fp64_ensemble_members_match_oracle_runs_alone(gpu):
    ens, forest, q0, sv, members, dt_run = _setup(8, fa.FP64)
    assert len(set(dt_run.tolist())) > 1          # the runs really step with different dt
    nsteps = 4
    cfl = sv.advance(1.0, nsteps)                 # dt = 1.0 x each slot's scale
    qg = sv.get_q()
    cfl_o = 0.0
    for b, (mp, fb, idx) in enumerate(members):
        qo, c = oracle_advance_ts(mp, fb, q0[idx], None, ens.mx, float(dt_run[b]), nsteps, time_sync=True)
        err = rel_max_err(qg[idx], qo)
        print("run %d: %d patches, maxlevel %d, dt %.4e, err %.2e" %
              (b, len(idx), int(fb.leaves()[0].max()), dt_run[b], err), flush=True)
        assert err <= 1e-12, b
        cfl_o = max(cfl_o, c)
    assert abs(cfl - cfl_o) <= 1e-12 * cfl_o


def test_fp32_ensemble_members_close_to_fp64(gpu):
    ens, forest, q0, s32, members, dt_run = _setup(8, fa.FP32_BULK)
    s64 = _setup(8, fa.FP64)[3]
    s32.advance(1.0, 20)
    s64.advance(1.0, 20)
    a, b = s32.get_q()[:, :, 2:-2, 2:-2], s64.get_q()[:, :, 2:-2, 2:-2]
    _, _, dx, dy = forest.geometry(ens.extent, ens.mx, ens.mx)
    area = (dx * dy)[:, None, None, None]
    for r, (_, _, idx) in enumerate(members):
        l1 = (np.abs(a[idx] - b[idx]) * area[idx]).sum(axis=(0, 2, 3)) / \
            (np.abs(b[idx]) * area[idx]).sum(axis=(0, 2, 3))
        assert np.all(l1 <= 1e-5), (r, l1)


def test_dt_scale_validation_and_reset(gpu):
    ens, forest, q0, sv, members, dt_run = _setup(3, fa.FP64)
    bad = np.ones(forest.npatch)
    bad[1] = -1.0
    with pytest.raises(RuntimeError):
        sv.set_dt_scale(bad)
    bad[1] = np.nan
    with pytest.raises(RuntimeError):
        sv.set_dt_scale(bad)
    # back to one common dt: every run steps with the smallest dt of the ensemble
    sv.set_dt_scale(None)
    dt = float(dt_run.min())
    sv.advance(dt, 2)
    qg = sv.get_q()
    for mp, fb, idx in members:
        qo, _ = oracle_advance_ts(mp, fb, q0[idx], None, ens.mx, dt, 2, time_sync=True)
        assert rel_max_err(qg[idx], qo) <= 1e-12

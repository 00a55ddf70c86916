"""The benchmark workload at its full size (shallow-water dam break, AMR
levels 3-8, 2032 patches of 32x32, time sync on):

  * FP64 form against the oracle over a few fused steps (1e-12 relative,
    max norm);
  * size-independent properties over many steps, for both forms: the radial
    dam break on its transpose-symmetric forest stays symmetric under
    x <-> y (h(x,y) = h(y,x), hu(x,y) = hv(y,x)) -- a transposed operand or a
    wrong index in one sweep breaks this at the 1e-2 level;
  * conservation on the periodic version of the same forest (mass change
    relative to the total, both forms).
"""
import dataclasses

import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import oracle_advance_ts, rel_max_err, setup_problem

pytestmark = pytest.mark.gpu


def _full():
    return P.dambreak_problem(3, 8, mx=32)


def test_fp64_full_workload_matches_oracle(gpu):
    prob = _full()
    forest, q0, aux, sv = setup_problem(prob, time_sync=True)
    assert forest.npatch == 2032
    dt = P.stable_dt(prob, forest, q0)
    nsteps = 3
    cfl = sv.advance(dt, nsteps)
    qg = sv.get_q()
    qo, cfl_o = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, nsteps, time_sync=True)
    err = rel_max_err(qg, qo)
    print("full size FP64: %d patches, err %.3e, bit-exact %d of %d" %
          (forest.npatch, err, int(np.sum(qg == qo)), qo.size), flush=True)
    assert err <= 1e-12
    assert abs(cfl - cfl_o) <= 1e-12 * cfl_o


def _transpose_map(forest):
    lv, gx, gy = forest.leaves()
    index = {(int(l), int(x), int(y)): p for p, (l, x, y) in enumerate(zip(lv, gx, gy))}
    return np.array([index[(int(l), int(y), int(x))] for l, x, y in zip(lv, gx, gy)])


@pytest.mark.parametrize("precision,tol", [(fa.FP64, 1e-10), (fa.FP32_BULK, 1e-5)])
def test_full_workload_stays_transpose_symmetric(gpu, precision, tol):
    prob = _full()
    forest, q0, aux, sv = setup_problem(prob, precision=precision, time_sync=True)
    tmap = _transpose_map(forest)
    dt = P.stable_dt(prob, forest, q0)
    sv.advance(dt, 40)
    q = sv.get_q()[:, :, 2:-2, 2:-2]
    qt = q[tmap]
    scale = np.abs(q).max(axis=(0, 2, 3))
    err_h = np.abs(q[:, 0] - np.swapaxes(qt[:, 0], 1, 2)).max() / scale[0]
    err_u = np.abs(q[:, 1] - np.swapaxes(qt[:, 2], 1, 2)).max() / scale[1]
    err_v = np.abs(q[:, 2] - np.swapaxes(qt[:, 1], 1, 2)).max() / scale[2]
    print("form %d: symmetry errors h %.2e hu %.2e hv %.2e" % (precision, err_h, err_u, err_v), flush=True)
    assert max(err_h, err_u, err_v) <= tol
    assert scale[1] > 1e-3   # the flow has developed


@pytest.mark.parametrize("precision", [fa.FP64, fa.FP32_BULK])
def test_full_workload_conserves_on_periodic_forest(gpu, precision):
    prob = dataclasses.replace(_full(), periodic=(True, True), mthbc=(2, 2, 2, 2))
    forest, q0, aux, sv = setup_problem(prob, precision=precision, time_sync=True)
    dt = P.stable_dt(prob, forest, q0)
    m0, _ = sv.mass()
    sv.advance(dt, 40)
    m1, _ = sv.mass()
    rel = abs(m1[0] - m0[0]) / abs(m0[0])
    mom = max(abs(m1[1] - m0[1]), abs(m1[2] - m0[2])) / abs(m0[0])
    print("form %d: mass change %.2e, momentum change %.2e (relative to the mass)" % (precision, rel, mom),
          flush=True)
    assert rel <= 1e-13 and mom <= 1e-13

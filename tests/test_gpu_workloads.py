"""The other workloads at their stated sizes.

  * periodic advection of a disk (velocity (1, 0.5), radius 0.2) on the
    uniform level-4 forest of 16x16 patches of 16x16 cells, Courant number
    0.9, 200 steps, FP64: bit parity with the oracle over the whole run, and
    the total of q unchanged (constant velocity: the fluctuation form is
    conservative);
  * the Euler shock-bubble on the 4x1 brick at levels 2-7 with 16x16
    patches: FP64 steps with time sync against the oracle (1e-12 relative),
    and the FP32-bulk form within 1e-5 of the FP64 form (relative L1 per
    field; the y-momentum, whose total is ~0 by symmetry, against the
    density total).
"""
import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import oracle_advance, oracle_advance_ts, oracle_mass, rel_max_err, setup_problem

pytestmark = pytest.mark.gpu


def test_advection_uniform_l4_200_steps(gpu):
    prob = P.advection_problem(4, 4, mx=16)
    forest, q0, aux, sv = setup_problem(prob)
    assert forest.npatch == 256
    dt = P.stable_dt(prob, forest, q0)
    cfl = sv.advance(dt, 200)
    qg = sv.get_q()
    qo, cfl_o = oracle_advance(prob, forest, q0, aux, prob.mx, dt, 200)
    err = rel_max_err(qg, qo)
    m0, m1 = oracle_mass(prob, forest, q0, prob.mx), oracle_mass(prob, forest, qg, prob.mx)
    print("advection L4, 200 steps: err %.3e, bit-exact %d of %d, mass change %.2e"
          % (err, int(np.sum(qg == qo)), qo.size, abs(m1[0] - m0[0]) / m0[0]), flush=True)
    assert err <= 1e-12
    assert abs(cfl - cfl_o) <= 1e-12 * cfl_o and abs(cfl - 0.9) <= 1e-12
    assert abs(m1[0] - m0[0]) <= 1e-13 * m0[0]


def test_shockbubble_l2_7_fp64_and_fp32(gpu):
    prob = P.shockbubble_problem(2, 7, mx=16)
    forest, q0, aux, s64 = setup_problem(prob, time_sync=True)
    assert forest.npatch > 1000 and int(forest.leaves()[0].max()) == 7
    dt = P.stable_dt(prob, forest, q0)
    s64.advance(dt, 3)
    qo, _ = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 3, time_sync=True)
    err = rel_max_err(s64.get_q(), qo)
    print("shock-bubble L2-7: %d patches, FP64 err %.3e" % (forest.npatch, err), flush=True)
    assert err <= 1e-12

    s64.set_q(q0)
    s32 = fa.Solver(prob.solver, prob.mx, prob.mx, precision=fa.FP32_BULK, mthbc=prob.mthbc,
                    gamma=prob.gamma, max_patches=forest.npatch, time_sync=True)
    s32.set_forest(forest, prob.extent)
    s32.set_q(q0)
    s64.advance(dt, 20)
    s32.advance(dt, 20)
    a, b = s32.get_q()[:, :, 2:-2, 2:-2], s64.get_q()[:, :, 2:-2, 2:-2]
    _, _, dx, dy = forest.geometry(prob.extent, prob.mx, prob.mx)
    area = (dx * dy)[:, None, None, None]
    tot = (np.abs(b) * area).sum(axis=(0, 2, 3))
    scale = np.maximum(tot, tot[0])
    l1 = (np.abs(a - b) * area).sum(axis=(0, 2, 3)) / scale
    print("FP32-bulk vs FP64 after 20 steps, L1 per field:", l1, flush=True)
    assert np.all(l1 <= 1e-5), l1

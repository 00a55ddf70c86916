"""CPU pins of the oracle's two extra Riemann solvers, the ones the
registered-solver path is tested with (no GPU).

acoustics : the oracle's restatement against the reference's own
            clawpack46_rpn2_acoustics.f / clawpack46_rpt2_acoustics.f driven by
            its Clawpack-4.6 patch step (oracle/_ref/libref46_acoustics.so),
            bit for bit, every limiter and order; and, without the reference,
            the closed form of a right-going simple wave translated by exactly
            one cell at Courant number 1.
tracers   : meqn = mwaves = n independent colour equations, tracer k moved by
            c_k (u, v).  Decoupled by construction, so each tracer must equal,
            bit for bit, the oracle's scalar advection (pinned against the
            reference's clawpack46_rpn2adv.f / _rpt2adv.f in test_oracle.py)
            run with velocity c_k (u, v) and that family's limiter.
"""
import math

import numpy as np
import pytest

from oracle_lib import REF_ACOUSTICS_SO, make_claw, step2_oracle, step2_ref_acoustics

ORACLE_RP_ADVECTION, ORACLE_RP_ACOUSTICS, ORACLE_RP_TRACERS = 0, 3, 4


def acoustics_params(rho=1.0, bulk=4.0):
    cc = math.sqrt(bulk / rho)
    return (rho, bulk, cc, rho * cc)


def acoustics_claw(order=(2, 2), lim=1, rho=1.0, bulk=4.0):
    return make_claw(ORACLE_RP_ACOUSTICS, 3, 0, 2, order=order, mthlim=[lim] * 2,
                     params=acoustics_params(rho, bulk))


def random_acoustics(mx, my, rng):
    return rng.normal(0.0, 1.0, (3, my + 4, mx + 4))


@pytest.mark.skipif(not __import__("os").path.exists(REF_ACOUSTICS_SO),
                    reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("order", [(2, 2), (2, 1), (1, 2), (2, 0), (1, 0)])
def test_acoustics_oracle_equals_reference_fortran(order):
    for lim in range(1, 7):
        for (mx, my), (rho, bulk) in (((8, 8), (1.0, 4.0)), ((10, 7), (1.3, 2.7))):
            rng = np.random.default_rng(1000 * lim + mx + 7 * order[0] + order[1])
            q = random_acoustics(mx, my, rng)
            claw = acoustics_claw(order, lim, rho, bulk)
            dt = 0.4 / (math.sqrt(bulk / rho) * max(mx, my))
            qo, co, fo = step2_oracle(claw, q, None, 1.0 / mx, 1.0 / my, dt, want_fluxes=True)
            qr, cr, fr = step2_ref_acoustics(claw, q, 1.0 / mx, 1.0 / my, dt)
            np.testing.assert_array_equal(qo, qr)
            assert co == cr
            for a, b in zip(fo, fr):
                np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("ixy", [1, 2])
def test_acoustics_simple_wave_moves_one_cell_at_courant_one(ixy):
    # p = zz * u is the right-going characteristic (speed +cc): with first-order
    # upwinding at dt = dx / cc the profile moves exactly one cell.  rho = 1,
    # bulk = 4 give cc = zz = 2, so every product below is exact.
    mx = 12
    rho, bulk, cc, zz = acoustics_params(1.0, 4.0)
    x = np.arange(mx + 4)
    prof = np.where((x > 4) & (x < 8), 1.0, 0.0) + 0.25 * (x == 8)
    q = np.zeros((3, mx + 4, mx + 4))
    if ixy == 1:
        q[0] = zz * prof[None, :]
        q[1] = prof[None, :]
    else:
        q[0] = zz * prof[:, None]
        q[2] = prof[:, None]
    claw = acoustics_claw(order=(1, 0))
    dx = 1.0 / mx
    qo, cfl = step2_oracle(claw, q, None, dx, dx, dx / cc)
    assert cfl == 1.0
    want = np.roll(q, 1, axis=2 if ixy == 1 else 1)
    sl = np.s_[:, 3:-2, 3:-2]
    np.testing.assert_array_equal(qo[sl], want[sl])


@pytest.mark.parametrize("n", [1, 5, 10])
@pytest.mark.parametrize("order", [(2, 2), (2, 1), (1, 2)])
def test_tracers_equal_scalar_advection_per_component(n, order):
    rng = np.random.default_rng(n * 10 + order[0] * 3 + order[1])
    mx, my = 10, 7
    q = rng.random((n, my + 4, mx + 4))
    aux = rng.normal(0.0, 1.0, (2, my + 4, mx + 4))
    lims = [1 + (k % 6) for k in range(n)]
    claw = make_claw(ORACLE_RP_TRACERS, n, 2, n, order=order, mthlim=lims)
    dt = 0.2 / max(mx, my)
    qt, ct = step2_oracle(claw, q, aux, 1.0 / mx, 1.0 / my, dt)
    cmax = 0.0
    for k in range(n):
        ck = float(n - k) / float(n)
        ca = make_claw(ORACLE_RP_ADVECTION, 1, 2, 1, order=order, mthlim=[lims[k]])
        qa, c1 = step2_oracle(ca, q[k:k + 1], ck * aux, 1.0 / mx, 1.0 / my, dt)
        np.testing.assert_array_equal(qt[k:k + 1], qa)
        cmax = max(cmax, c1)
    assert ct == cmax

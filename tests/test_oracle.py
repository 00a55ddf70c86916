"""CPU checks that pin the oracle (no GPU).

1. Golden vectors written from the reference's own Fortran
   (tests/golden/make_golden.py) must be reproduced bit for bit.
2. Where oracle/_ref/libref46.so is present, the restatement must equal the
   reference Fortran bit for bit on fresh random cases (all solvers, all six
   limiters, first/second order, with and without transverse terms).
3. Closed forms and invariants that a dropped term, wrong sign, wrong index or
   transposed operand would break: exact preservation of constant states,
   CFL = max|lambda| dt/dx, exact translation for first-order advection at
   Courant number 1, x/y symmetry, ghost copy as index translation, averaging
   of constants, exact interpolation of linear data, Kahan mass sums.
"""
import ctypes as C
import glob
import math
import os

import numpy as np
import pytest

import forestclaw_amd as fa
from helpers import claw_for, interior, random_aux, random_state
from oracle_lib import (OracleClaw, make_claw, oracle, ref, ref_available, step2_oracle,
                        step2_ref)

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "step2_*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_reproduces_golden_step(path):
    g = np.load(path)
    solver, meqn, mwaves, maux = (int(g[k]) for k in ("solver", "meqn", "mwaves", "maux"))
    claw = make_claw(solver, meqn, maux, mwaves, mthlim=[int(g["limiter"])] * mwaves)
    aux = g["aux"] if maux else None
    q, cfl, (fm, fp, gm, gp) = step2_oracle(claw, g["q_in"], aux, float(g["dx"]),
                                            float(g["dy"]), float(g["dt"]), want_fluxes=True)
    np.testing.assert_array_equal(q, g["q_out"])
    assert cfl == float(g["cfl"])
    for a, b in ((fm, g["fm"]), (fp, g["fp"]), (gm, g["gm"]), (gp, g["gp"])):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "bc2_*.npz"))))
def test_oracle_reproduces_golden_bc2(path):
    g = np.load(path)
    q = np.ascontiguousarray(g["q_in"]).copy()
    oracle().oracle_bc2(6, 5, 2, 3, q.reshape(-1), np.asarray(g["mthbc"], dtype=np.int32))
    np.testing.assert_array_equal(q, g["q_out"])


requires_ref = pytest.mark.skipif(not ref_available(),
                                  reason="oracle/_ref not built (needs the reference tree)")


@requires_ref
@pytest.mark.parametrize("solver", [fa.RP_ADVECTION, fa.RP_SHALLOW, fa.RP_EULER])
@pytest.mark.parametrize("order", [(2, 2), (2, 1), (1, 2), (2, 0), (1, 0)])
def test_oracle_equals_reference_fortran(solver, order):
    meqn, mwaves, maux = fa.SOLVER_SHAPES[solver]
    for lim in range(1, 7):
        for mx, my in ((8, 8), (10, 7)):
            rng = np.random.default_rng(hash((solver, order, lim, mx)) % 2 ** 32)
            q = random_state(solver, 1, mx, my, rng)[0]
            aux = random_aux(1, mx, my, rng)[0] if maux else None
            claw = make_claw(solver, meqn, maux, mwaves, order=order, mthlim=[lim] * mwaves)
            dt = 0.25 * min(1.0 / mx, 1.0 / my)
            qo, co, fo = step2_oracle(claw, q, aux, 1.0 / mx, 1.0 / my, dt, want_fluxes=True)
            qr, cr, fr = step2_ref(claw, q, aux, 1.0 / mx, 1.0 / my, dt, want_fluxes=True)
            np.testing.assert_array_equal(qo, qr)
            assert co == cr
            for a, b in zip(fo, fr):
                np.testing.assert_array_equal(a, b)


@requires_ref
def test_oracle_bc2_equals_reference_fortran():
    rng = np.random.default_rng(5)
    for trial in range(20):
        mth = rng.choice([-1, 1, 3], size=4).astype(np.int32)
        q = random_state(fa.RP_SHALLOW, 1, 7, 9, rng)[0]
        a, b = q.copy(), q.copy()
        oracle().oracle_bc2(7, 9, 2, 3, a.reshape(-1), mth)
        ref().ref46_bc2(7, 9, 2, 3, b.reshape(-1), mth)
        np.testing.assert_array_equal(a, b)


@requires_ref
def test_oracle_limiter_equals_reference_fortran():
    rng = np.random.default_rng(9)
    maxm, mbc, mx, meqn, mwaves = 12, 2, 12, 3, 3
    n = maxm + 2 * mbc
    for lim in range(0, 8):
        wave = rng.normal(size=(mwaves, meqn, n))
        wave[:, :, 5] = 0.0  # a zero-norm wave is skipped
        s = rng.normal(size=(mwaves, n))
        mthlim = np.array([lim] * mwaves + [0] * (10 - mwaves), dtype=np.int32)
        a, b = wave.copy(), wave.copy()
        oracle().oracle_inlinelimiter(maxm, meqn, mwaves, mbc, mx, a.reshape(-1),
                                      np.ascontiguousarray(s).reshape(-1), mthlim)
        ref().ref46_inlinelimiter(maxm, meqn, mwaves, mbc, mx, b.reshape(-1),
                                  np.ascontiguousarray(s).reshape(-1).copy(), mthlim)
        np.testing.assert_array_equal(a, b)


# ----------------------------------------------------------------- closed forms

@pytest.mark.parametrize("solver", [fa.RP_SHALLOW, fa.RP_EULER])
def test_constant_state_preserved_and_cfl(solver):
    mx = my = 8
    shape = (my + 4, mx + 4)
    if solver == fa.RP_SHALLOW:
        h, u, v, g = 1.7, 0.3, -0.2, 1.0
        q = np.stack([np.full(shape, h), np.full(shape, h * u), np.full(shape, h * v)])
        smax = max(abs(u) + math.sqrt(g * h), abs(v) + math.sqrt(g * h))
    else:
        rho, u, v, p, gam = 1.3, 0.4, 0.1, 0.9, 1.4
        E = p / (gam - 1) + 0.5 * rho * (u * u + v * v)
        q = np.stack([np.full(shape, rho), np.full(shape, rho * u), np.full(shape, rho * v),
                      np.full(shape, E)])
        c = math.sqrt(gam * p / rho)
        smax = max(abs(u) + c, abs(v) + c)
    dx = 1.0 / mx
    dt = 0.5 * dx
    qo, cfl = step2_oracle(claw_for(solver), q, None, dx, dx, dt)
    np.testing.assert_array_equal(interior(qo), interior(q))
    assert abs(cfl - smax * dt / dx) <= 1e-14 * cfl


def test_first_order_advection_translates_exactly():
    mx = my = 8
    rng = np.random.default_rng(3)
    q = rng.random((1, my + 4, mx + 4))
    aux = np.zeros((2, my + 4, mx + 4))
    aux[0] = 1.0
    dx = 1.0 / mx
    claw = claw_for(fa.RP_ADVECTION, order=(1, 0))
    qo, cfl = step2_oracle(claw, q, aux, dx, dx, dx)  # Courant number 1
    np.testing.assert_array_equal(qo[0, 2:-2, 2:-2], q[0, 2:-2, 1:-3])
    assert cfl == 1.0


def test_xy_symmetry_of_step():
    """Transposing the data (and swapping momenta) transposes the result."""
    mx = 8
    rng = np.random.default_rng(4)
    q = random_state(fa.RP_SHALLOW, 1, mx, mx, rng)[0]
    qt = np.stack([q[0].T, q[2].T, q[1].T])
    claw = claw_for(fa.RP_SHALLOW)
    a, ca = step2_oracle(claw, q, None, 1 / mx, 1 / mx, 0.2 / mx)
    b, cb = step2_oracle(claw, np.ascontiguousarray(qt), None, 1 / mx, 1 / mx, 0.2 / mx)
    bt = np.stack([b[0].T, b[2].T, b[1].T])
    np.testing.assert_allclose(interior(a), interior(bt), rtol=0, atol=1e-14)
    assert ca == cb


def test_ghost_copy_is_translation():
    mx, my, mbc, meqn = 6, 6, 2, 2
    rng = np.random.default_rng(6)
    nb = rng.random((meqn, my + 4, mx + 4))
    for iface in range(4):
        this = np.zeros_like(nb)
        ti = {0: mx, 1: -mx, 2: 0, 3: 0}[iface]
        tj = {0: 0, 1: 0, 2: my, 3: -my}[iface]
        oracle().oracle_copy_face(mx, my, mbc, meqn, this.reshape(-1), nb.reshape(-1), iface, ti, tj)
        if iface == 0:
            np.testing.assert_array_equal(this[:, 2:-2, 0:2], nb[:, 2:-2, mx:mx + 2])
        if iface == 3:
            np.testing.assert_array_equal(this[:, my + 2:, 2:-2], nb[:, 2:4, 2:-2])


def test_average_and_interpolate_closed_forms():
    mx = my = 8
    mbc, meqn = 2, 1
    jj, ii = np.mgrid[1 - mbc:my + mbc + 1, 1 - mbc:mx + mbc + 1]
    # linear data on the coarse patch, in coarse-cell units
    coarse = (0.5 + 0.25 * ii + 0.125 * jj)[None].astype(np.float64)
    fine = np.zeros_like(coarse)
    # fine patch to the right of the coarse patch's face 1, lower half:
    # first fine child of coarse cell (ic, jc) is 2*ic - 1 + ti with ti = -2*mx
    oracle().oracle_interpolate_face(mx, my, mbc, meqn, np.ascontiguousarray(coarse).reshape(-1),
                                     fine.reshape(-1), 0, 1, -2 * mx, 0)
    # fine ghost cells i = -1, 0 (array 0, 1) hold the linear function at fine centres
    for j in range(1, my + 1):
        for i in (-1, 0):
            xc = (i + 2 * mx - 0.5) / 2 + 0.5  # coarse-cell coordinate of the fine centre
            yc = (j - 0.5) / 2 + 0.5
            assert fine[0, j + 1, i + 1] == pytest.approx(0.5 + 0.25 * xc + 0.125 * yc, abs=1e-15)
    # averaging four equal fine values returns that value exactly
    fineconst = np.full((1, my + 4, mx + 4), 0.3)
    c2 = np.zeros((1, my + 4, mx + 4))
    oracle().oracle_average_face(mx, my, mbc, meqn, c2.reshape(-1), fineconst.reshape(-1), 0, 1,
                                 -2 * mx, 0)
    assert np.all(c2[0, 2:2 + my // 2, mx + 2:] == 0.3)


def test_mass_is_kahan_sum():
    from forestclaw_amd import problems as P
    from helpers import oracle_mass
    prob = P.dambreak_problem(2, 3, mx=8)
    forest = P.build_forest(prob)
    q, _ = P.initial_data(prob, forest)
    s = oracle_mass(prob, forest, q, 8)
    _, _, dx, dy = forest.geometry(prob.extent, 8, 8)
    exact = math.fsum((interior(q)[:, 0] * (dx * dy)[:, None, None]).ravel())
    assert abs(s[0] - exact) <= 2e-16 * exact

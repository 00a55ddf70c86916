"""CPU checks that pin the oracle's time synchronisation (the coarse/fine and
same-size flux-register corrections of the reference's fclaw2d_time_sync).

The reference's clawpack46_time_sync.f does not compile unmodified here
(`xor` on logicals), so it cannot be run; the restatement is pinned by

1. the shallow-water conservative flux against the reference's own
   rpn2cons_update.f (compiled in place into oracle/_ref), bit for bit;
2. conservation: on a periodic AMR forest the total of every conserved
   field is unchanged to rounding after steps with time sync, while without
   it the coarse/fine flux mismatch changes it by orders of magnitude more --
   a dropped register, a wrong sign or a wrong fine index breaks this;
3. identities: on a uniform forest every same-size correction is exactly
   zero, and for a constant state on an AMR forest every correction is
   exactly zero, so q is unchanged bit for bit.
"""
import ctypes as C
import dataclasses
import math

import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import claw_for, oracle_advance, oracle_advance_ts, oracle_ghost_update, \
    oracle_mass, random_state, interior
from oracle_lib import oracle, ref, ref_available


def _periodic(prob):
    return dataclasses.replace(prob, periodic=(True, True), mthbc=(2, 2, 2, 2))


def _setup(prob):
    forest = P.build_forest(prob)
    q0, aux = P.initial_data(prob, forest)
    q0 = oracle_ghost_update(prob, forest, q0, prob.mx)
    dt = P.stable_dt(prob, forest, q0)
    return forest, q0, aux, dt


@pytest.mark.skipif(not ref_available(), reason="reference Fortran not built (make -C oracle ref)")
def test_swe_cons_flux_equals_reference_fortran():
    rng = np.random.default_rng(5)
    L = ref()
    C.c_double.in_dll(L, "cparam_").value = 9.81
    claw = claw_for(fa.RP_SHALLOW, grav=9.81)
    for _ in range(200):
        q = np.array([rng.uniform(0.1, 3.0), rng.normal(), rng.normal()])
        for idir in (0, 1):
            got = np.zeros(3)
            oracle().oracle_cons_flux(C.byref(claw), idir, q, np.zeros(1), got)
            want = np.zeros(3)
            i = C.c_int
            L.rpn2_cons_update_(C.byref(i(3)), C.byref(i(0)), C.byref(i(idir)),
                                C.byref(i(2 * idir)), q.copy(), np.zeros(1), np.zeros(1), want)
            assert np.array_equal(got, want), (q, idir, got, want)


@pytest.mark.parametrize("which", ["shallow", "euler"])
def test_time_sync_conserves_on_periodic_amr_forest(which):
    if which == "shallow":
        prob = _periodic(P.dambreak_problem(2, 4, mx=8))
        fields = [0, 1, 2]
    else:
        prob = dataclasses.replace(P.shockbubble_problem(1, 3, mx=8), periodic=(True, True),
                                   mthbc=(2, 2, 2, 2))
        fields = [0, 1, 2, 3]
    forest, q0, aux, dt = _setup(prob)
    assert len(set(forest.leaves()[0].tolist())) >= 2   # a real coarse/fine forest
    m0 = oracle_mass(prob, forest, q0, prob.mx)
    nsteps = 6
    qs, _ = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, nsteps, time_sync=True)
    qn, _ = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, nsteps, time_sync=False)
    ms = oracle_mass(prob, forest, qs, prob.mx)
    mn = oracle_mass(prob, forest, qn, prob.mx)
    # momentum starts at zero: measure it against the total |momentum| reached
    scale = np.maximum(oracle_mass(prob, forest, np.abs(q0), prob.mx),
                       oracle_mass(prob, forest, np.abs(qs), prob.mx))
    drift_sync = np.abs(ms - m0) / scale
    drift_none = np.abs(mn - m0) / scale
    # the density/height field always has a coarse/fine mismatch without sync
    assert drift_none[0] > 1e-9, drift_none
    for m in fields:
        assert drift_sync[m] <= 2e-15, (m, drift_sync, drift_none)


def test_time_sync_without_flag_matches_plain_advance():
    prob = P.dambreak_problem(2, 4, mx=8)
    forest, q0, aux, dt = _setup(prob)
    a, ca = oracle_advance(prob, forest, q0, aux, prob.mx, dt, 3)
    b, cb = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 3, time_sync=False)
    assert np.array_equal(a, b) and ca == cb


def test_time_sync_is_identity_on_uniform_forest():
    prob = _periodic(P.dambreak_problem(3, 3, mx=8))
    forest, q0, aux, dt = _setup(prob)
    assert len(set(forest.leaves()[0].tolist())) == 1
    a, _ = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 3, time_sync=True)
    b, _ = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 3, time_sync=False)
    assert np.array_equal(a, b)


def test_time_sync_is_identity_for_constant_state_on_amr_forest():
    prob = P.dambreak_problem(2, 4, mx=8)
    forest = P.build_forest(prob)           # forest from the dam-break data ...
    q0, aux = P.initial_data(prob, forest)
    q0[:, 0] = 1.5                          # ... but a constant state on it
    q0[:, 1] = 0.25
    q0[:, 2] = -0.125
    q0 = oracle_ghost_update(prob, forest, q0, prob.mx)
    dt = P.stable_dt(prob, forest, q0)
    assert len(set(forest.leaves()[0].tolist())) >= 2
    a, _ = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 2, time_sync=True)
    b, _ = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 2, time_sync=False)
    assert np.array_equal(a, b)
    assert np.array_equal(a, q0)


# ---------------------------------------------------------------- register routines
from oracle_lib import ref_sync, ref_sync_available


def _ref_rpn2_cons(grav):
    L = ref()
    C.c_double.in_dll(L, "cparam_").value = grav
    return C.cast(L.rpn2_cons_update_, C.c_void_p).value


@pytest.mark.skipif(not ref_sync_available(), reason="reference Fortran not built (make -C oracle ref)")
@pytest.mark.parametrize("mx,my", [(8, 8), (16, 16)])
def test_store_flux_equals_reference_fortran(mx, my):
    """oracle_store_flux (shallow water) adds the same bits to the same
    registers as clawpack46_time_sync_store_flux with rpn2cons_update."""
    mbc, meqn, grav = 2, 3, 0.75
    rng = np.random.default_rng(mx * my)
    q = np.ascontiguousarray(random_state(fa.RP_SHALLOW, 1, mx, my, rng)[0])
    dt, dx, dy = 0.013, 1.0 / mx, 0.9 / my
    claw = claw_for(fa.RP_SHALLOW, grav=grav)
    reg = rng.normal(size=(4, 4, meqn, max(mx, my)))   # registers accumulate onto old values
    want = [np.ascontiguousarray(reg[f, 2:4, :, :n]) for f, n in ((0, my), (1, my), (2, mx), (3, mx))]
    got = reg.copy()
    oracle().oracle_store_flux(C.byref(claw), mx, my, mbc, q, None, dt, dx, dy, got.reshape(-1))
    el = [np.full(my, dy), np.full(my, dy), np.full(mx, dx), np.full(mx, dx)]
    i = lambda v: C.byref(C.c_int(v))
    ref_sync().clawpack46_time_sync_store_flux_(
        i(mx), i(my), i(mbc), i(meqn), i(0), i(0), i(0), C.byref(C.c_double(dt)), *el, q, None,
        *want, _ref_rpn2_cons(grav), np.zeros(meqn), np.zeros(1), np.zeros(1), np.zeros(meqn))
    for f in range(4):
        np.testing.assert_array_equal(got[f, 2:4], want[f])
    np.testing.assert_array_equal(got[:, 0:2], reg[:, 0:2])   # fm/fp registers untouched


@pytest.mark.skipif(not ref_sync_available(), reason="reference Fortran not built (make -C oracle ref)")
def test_accumulate_waves_equals_reference_fortran():
    """oracle_accumulate_waves: fp with its sign flipped, fm as is, scaled by
    dt * dy (x-faces) and dt * dx (y-faces), added to the old registers, as
    clawpack46_time_sync_accumulate_waves does."""
    mx = my = 8
    mbc, meqn = 2, 3
    rng = np.random.default_rng(11)
    shape = (meqn, my + 2 * mbc, mx + 2 * mbc)
    fp, fm, gp, gm = (rng.normal(size=shape) for _ in range(4))
    dt, dx, dy = 0.02, 0.125, 0.1
    reg = rng.normal(size=(4, 4, meqn, mx))
    got = reg.copy()
    oracle().oracle_accumulate_waves(mx, my, mbc, meqn, dt, dx, dy, fp.reshape(-1), fm.reshape(-1),
                                     gp.reshape(-1), gm.reshape(-1), got.reshape(-1))
    # Fortran arrays fp_left(my, meqn) etc. are (meqn, my) in C order
    R = {k: np.ascontiguousarray(reg[f, kind]) for k, (f, kind) in {
        "fp_left": (0, 1), "fp_right": (1, 1), "fm_left": (0, 0), "fm_right": (1, 0),
        "gp_bottom": (2, 1), "gp_top": (3, 1), "gm_bottom": (2, 0), "gm_top": (3, 0)}.items()}
    el = [np.full(my, dy), np.full(my, dy), np.full(mx, dx), np.full(mx, dx)]
    i = lambda v: C.byref(C.c_int(v))
    d = lambda v: C.byref(C.c_double(v))
    ref_sync().clawpack46_time_sync_accumulate_waves_(
        i(mx), i(my), i(mbc), i(meqn), d(dt), d(dx), d(dy), i(0), *el, fp, fm, gp, gm,
        R["fp_left"], R["fp_right"], R["fm_left"], R["fm_right"],
        R["gp_bottom"], R["gp_top"], R["gm_bottom"], R["gm_top"])
    for k, (f, kind) in {"fp_left": (0, 1), "fp_right": (1, 1), "fm_left": (0, 0),
                         "fm_right": (1, 0), "gp_bottom": (2, 1), "gp_top": (3, 1),
                         "gm_bottom": (2, 0), "gm_top": (3, 0)}.items():
        np.testing.assert_array_equal(got[f, kind], R[k], err_msg=k)
    np.testing.assert_array_equal(got[:, 2:], reg[:, 2:])


# ---------------------------------------------------------------- where the fix-up lands
def _face_kinds(forest, extent, mx):
    """Per patch and face: 'phys', 'same', 'finer' or 'coarser' neighbour, by
    point location just outside the face (independent of the neighbour
    searches of the oracle and the product)."""
    lv, gx, gy = forest.leaves()
    ax, bx, ay, by = extent
    wx, wy = (bx - ax) / forest.nbx, (by - ay) / forest.nby
    s = np.ldexp(1.0, -lv.astype(np.int64))
    x0, y0 = ax + gx * wx * s, ay + gy * wy * s
    x1, y1 = x0 + wx * s, y0 + wy * s
    W, H = bx - ax, by - ay

    def owner(x, y):
        if forest.periodic[0]:
            x = ax + (x - ax) % W
        if forest.periodic[1]:
            y = ay + (y - ay) % H
        hit = np.nonzero((x >= x0) & (x < x1) & (y >= y0) & (y < y1))[0]
        return int(hit[0]) if len(hit) else -1

    out = []
    for p in range(len(lv)):
        h = (x1[p] - x0[p]) / mx
        xm, ym = 0.5 * (x0[p] + x1[p]), 0.5 * (y0[p] + y1[p])
        # a quarter of the way along the face, half a fine cell outside
        pts = [(x0[p] - h / 4, y0[p] + h / 4 + 0.25 * (y1[p] - y0[p])),
               (x1[p] + h / 4, y0[p] + h / 4 + 0.25 * (y1[p] - y0[p])),
               (x0[p] + h / 4 + 0.25 * (x1[p] - x0[p]), y0[p] - h / 4),
               (x0[p] + h / 4 + 0.25 * (x1[p] - x0[p]), y1[p] + h / 4)]
        kinds = []
        for (x, y) in pts:
            o = owner(x, y)
            kinds.append("phys" if o < 0 else "same" if lv[o] == lv[p] else
                         "finer" if lv[o] > lv[p] else "coarser")
        out.append(kinds)
    return out


@pytest.mark.parametrize("which", ["shallow", "euler"])
def test_time_sync_lands_on_coarse_and_same_size_edges(which):
    """sync - no sync after one step: zero in every ghost cell (the ghost
    update precedes the fix-up, src/fclaw2d_advance.c:342-349), zero in every
    cell off the first interior layer, zero on edges facing a coarser
    neighbour or a physical boundary (only the coarse side of a coarse/fine
    face and both sides of a same-size face are corrected,
    src/fclaw2d_time_sync.c:118-158); non-zero somewhere on coarse/fine
    edges; and its total is minus the mass change without the fix-up."""
    if which == "shallow":
        prob = _periodic(P.dambreak_problem(2, 5, mx=8, r0=0.23))
        fields = [0, 1, 2]
    else:
        prob = dataclasses.replace(P.shockbubble_problem(1, 4, mx=8), periodic=(True, True),
                                   mthbc=(2, 2, 2, 2))
        fields = [0, 1, 2, 3]
    forest, _, aux, _ = _setup(prob)
    mx = prob.mx
    # rough random states: every coarse/fine face carries a flux mismatch
    rng = np.random.default_rng(31)
    q0 = oracle_ghost_update(prob, forest, random_state(prob.solver, forest.npatch, mx, mx, rng), mx)
    dt = 0.5 * P.stable_dt(prob, forest, q0)
    qs, _ = oracle_advance_ts(prob, forest, q0, aux, mx, dt, 1, time_sync=True)
    qn, _ = oracle_advance_ts(prob, forest, q0, aux, mx, dt, 1, time_sync=False)
    D = qs - qn
    mbc = 2
    inner = D[:, :, mbc:-mbc, mbc:-mbc]
    ghost = D.copy()
    ghost[:, :, mbc:-mbc, mbc:-mbc] = 0
    assert not ghost.any(), "the fix-up touched ghost cells"
    assert not inner[:, :, 1:-1, 1:-1].any(), "the fix-up touched cells off the patch edges"
    kinds = _face_kinds(forest, prob.extent, mx)
    n_cf = 0
    for p, k in enumerate(kinds):
        # edge strips without the corner cells; a corner cell belongs to two faces
        strips = [inner[p, :, 1:-1, 0], inner[p, :, 1:-1, -1], inner[p, :, 0, 1:-1], inner[p, :, -1, 1:-1]]
        for face in range(4):
            if k[face] in ("phys", "coarser"):
                assert not strips[face].any(), (p, face, k[face])
            if k[face] == "finer" and strips[face].any():
                n_cf += 1
        for (jc, ic, fa_, fb_) in ((0, 0, 0, 2), (0, -1, 1, 2), (-1, 0, 0, 3), (-1, -1, 1, 3)):
            if k[fa_] in ("phys", "coarser") and k[fb_] in ("phys", "coarser"):
                assert not inner[p, :, jc, ic].any(), (p, "corner", jc, ic)
    assert n_cf > 0, "no coarse/fine edge was corrected: the check has no teeth"
    _, _, dx, dy = forest.geometry(prob.extent, mx, mx)
    area = (dx * dy)[:, None, None]
    m0 = oracle_mass(prob, forest, q0, mx)
    mn = oracle_mass(prob, forest, qn, mx)
    ms = oracle_mass(prob, forest, qs, mx)
    for m in fields:
        total = math.fsum((inner[:, m] * area).ravel())
        scale = math.fsum(np.abs(interior(q0)[:, m] * area).ravel())
        assert abs(total + (mn[m] - m0[m])) <= 1e-13 * scale, (m, total, mn[m] - m0[m])
        assert abs(ms[m] - m0[m]) <= 1e-13 * scale


def test_time_sync_keeps_mirror_symmetry():
    """A dam break centred in the unit square on its mirror-symmetric forest
    stays symmetric under x -> 1-x, y -> 1-y and x <-> y with the fix-up in
    every step: a correction placed one cell off along an edge, or applied to
    the cells of the other axis, breaks it by the size of the correction."""
    prob = _periodic(P.dambreak_problem(2, 5, mx=8))
    forest, q0, aux, dt = _setup(prob)
    lv, gx, gy = forest.leaves()
    n = 1 << lv
    keys = set(zip(lv.tolist(), gx.tolist(), gy.tolist()))
    assert keys == set(zip(lv.tolist(), (n - 1 - gx).tolist(), gy.tolist()))
    assert keys == set(zip(lv.tolist(), gy.tolist(), gx.tolist()))
    qs, _ = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 6, time_sync=True)
    qn, _ = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, 6, time_sync=False)
    assert np.abs(qs - qn).max() > 1e-6      # the fix-up is large against the tolerance
    index = {k: p for p, k in enumerate(zip(lv.tolist(), gx.tolist(), gy.tolist()))}
    h, hu, hv = (interior(qs)[:, m] for m in range(3))
    for p, (L_, x_, y_) in enumerate(zip(lv.tolist(), gx.tolist(), gy.tolist())):
        mp = index[(L_, (1 << L_) - 1 - x_, y_)]      # x mirror: hu changes sign
        np.testing.assert_allclose(h[p], h[mp][:, ::-1], rtol=0, atol=1e-13)
        np.testing.assert_allclose(hu[p], -hu[mp][:, ::-1], rtol=0, atol=1e-13)
        np.testing.assert_allclose(hv[p], hv[mp][:, ::-1], rtol=0, atol=1e-13)
        tp = index[(L_, y_, x_)]                      # transpose: hu <-> hv
        np.testing.assert_allclose(h[p], h[tp].T, rtol=0, atol=1e-13)
        np.testing.assert_allclose(hu[p], hv[tp].T, rtol=0, atol=1e-13)

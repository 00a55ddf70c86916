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

import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import claw_for, oracle_advance, oracle_advance_ts, oracle_ghost_update, \
    oracle_mass
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

"""GPU time synchronisation (the conservation fix-up of fclaw2d_time_sync):
FP64 form against the oracle (stated tolerance 1e-12 relative, max norm) and
conservation of both forms on periodic AMR forests."""
import dataclasses

import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
import dropin_lib as D
from helpers import oracle_advance_ts, oracle_ghost_update, rel_max_err, setup_problem

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _periodic(prob):
    return dataclasses.replace(prob, periodic=(True, True), mthbc=(2, 2, 2, 2))


CASES = {
    "dambreak": lambda: P.dambreak_problem(2, 4, mx=16),
    "dambreak_periodic": lambda: _periodic(P.dambreak_problem(2, 4, mx=16)),
    "shockbubble": lambda: P.shockbubble_problem(1, 3, mx=16),
    "advection_amr": lambda: P.advection_problem(2, 4, mx=16),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("tile,layout", [(0, 0), (8, 1), (0, 2), (8, 2)])
def test_fp64_time_sync_matches_oracle(gpu, case, tile, layout):
    prob = CASES[case]()
    forest, q0, aux, sv = setup_problem(prob, tile=tile, time_sync=True)
    sv.set_ghost_layout(layout)
    assert len(set(forest.leaves()[0].tolist())) >= 2
    dt = P.stable_dt(prob, forest, q0) * (0.5 if prob.solver == fa.RP_ADVECTION else 1.0)
    nsteps = 6
    cfl = sv.advance(dt, nsteps)
    qg = sv.get_q()
    qo, cfl_o = oracle_advance_ts(prob, forest, q0, aux, prob.mx, dt, nsteps, time_sync=True)
    err = rel_max_err(qg, qo)
    assert err <= TOL, "%s tile %d: %.3e (bit-exact %d of %d)" % (case, sv.tile, err,
                                                                   int(np.sum(qg == qo)), qo.size)
    assert abs(cfl - cfl_o) <= TOL * cfl_o


@pytest.mark.parametrize("precision", [fa.FP64, fa.FP32_BULK])
@pytest.mark.parametrize("which", ["shallow", "euler"])
def test_time_sync_conserves_on_periodic_amr(gpu, precision, which):
    if which == "shallow":
        prob = _periodic(P.dambreak_problem(2, 5, mx=16))
    else:
        prob = _periodic(P.shockbubble_problem(1, 4, mx=16))
    forest, q0, aux, sv = setup_problem(prob, precision=precision, time_sync=True)
    assert len(set(forest.leaves()[0].tolist())) >= 3
    dt = P.stable_dt(prob, forest, q0)
    m0, _ = sv.mass()
    sv.advance(dt, 40)
    m1, _ = sv.mass()
    q = sv.get_q()[:, :, 2:-2, 2:-2]
    _, _, dx, dy = forest.geometry(prob.extent, prob.mx, prob.mx)
    absmass = (np.abs(q) * (dx * dy)[:, None, None, None]).sum(axis=(0, 2, 3))
    for m in range(len(m0)):
        # a field whose total is ~0 (the y-momentum of a y-symmetric flow) is
        # measured against the density total, not against its own rounding noise
        scale = max(abs(m0[m]), absmass[m], abs(m0[0]))
        assert abs(m1[m] - m0[m]) / scale <= 1e-13, (m, m0, m1)


def test_fp32_with_time_sync_close_to_fp64(gpu):
    prob = P.dambreak_problem(2, 5, mx=16)
    forest, q0, aux, s64 = setup_problem(prob, time_sync=True)
    s32 = fa.Solver(prob.solver, prob.mx, prob.mx, precision=fa.FP32_BULK, mthbc=prob.mthbc,
                    grav=prob.grav, max_patches=forest.npatch, time_sync=True)
    s32.set_forest(forest, prob.extent)
    s32.set_q(q0)
    dt = P.stable_dt(prob, forest, q0)
    s64.advance(dt, 30)
    s32.advance(dt, 30)
    a, b = s32.get_q()[:, :, 2:-2, 2:-2], s64.get_q()[:, :, 2:-2, 2:-2]
    _, _, dx, dy = forest.geometry(prob.extent, prob.mx, prob.mx)
    area = (dx * dy)[:, None, None, None]
    l1 = (np.abs(a - b) * area).sum(axis=(0, 2, 3)) / (np.abs(b) * area).sum(axis=(0, 2, 3))
    assert np.all(l1 <= 1e-5), l1


def test_dropin_time_sync_per_level(gpu):
    """fclaw2d_update_single_step per level, fclaw2d_ghost_update, then
    fclaw2d_time_sync, against the oracle."""
    import ctypes as C
    L = D.bind()
    prob = P.dambreak_problem(2, 4, mx=16)
    forest = P.build_forest(prob)
    lv, gx, gy = forest.leaves()
    npatch = len(lv)
    fh = L.fclaw2d_hip_forest_from_leaves(1, 1, 0, 0, npatch, lv, gx, gy, 0)
    dom = L.fclaw2d_domain_new_from_forest(fh, 1, 1, 0, 0, 0.0, 1.0, 0.0, 1.0)
    glob = L.fclaw2d_global_new()
    L.fclaw2d_global_store_domain(glob, dom)
    glob.contents.advance_one_step = 1
    glob.contents.time_sync = 1
    cp = glob.contents.clawpatch_opt
    cp.mx, cp.my, cp.mbc, cp.meqn, cp.maux = 16, 16, 2, 3, 0
    opt = D.Options()
    L.fc2d_hipclaw_options_default(C.byref(opt), 3)
    for k in range(3):
        opt.mthlim[k] = 4
    opt.buffer_len = 13
    L.fc2d_hipclaw_options_store(glob, C.byref(opt))
    L.fc2d_hipclaw_solver_initialize(glob)
    L.fc2d_hipclaw_vt(glob).contents.riemann_solver = fa.RP_SHALLOW
    L.fc2d_hipclaw_initialize_GPUs(glob)
    L.fc2d_hipclaw_allocate_buffers(glob)
    q0, _ = P.initial_data(prob, forest)
    q0 = oracle_ghost_update(prob, forest, q0, 16)
    D.host_q(glob, npatch, (3, 20, 20))[:] = q0
    L.fc2d_hipclaw_upload(glob)
    dt = P.stable_dt(prob, forest, q0)
    t = 0.0
    for n in range(3):
        for level in range(int(lv.max()), int(lv.min()) - 1, -1):
            L.fclaw2d_update_single_step(glob, level, t, dt)
        L.fclaw2d_ghost_update(glob, int(lv.min()), int(lv.max()), t + dt, 0, 0)
        L.fclaw2d_time_sync(glob, int(lv.min()), int(lv.max()))
        t += dt
    L.fc2d_hipclaw_download(glob)
    qg = D.host_q(glob, npatch, (3, 20, 20)).copy()
    qo, _ = oracle_advance_ts(prob, forest, q0, None, 16, dt, 3, time_sync=True)
    assert rel_max_err(qg, qo) <= TOL
    # the fused all-level step with time sync gives the same state
    L.fclaw2d_advance_all_levels(glob, t, dt)
    L.fc2d_hipclaw_download(glob)
    qg2 = D.host_q(glob, npatch, (3, 20, 20)).copy()
    qo2, _ = oracle_advance_ts(prob, forest, qo, None, 16, dt, 1, time_sync=True)
    assert rel_max_err(qg2, qo2) <= TOL
    L.fclaw2d_global_destroy(glob)


@pytest.mark.parametrize("case", ["dambreak", "dambreak_periodic", "shockbubble"])
def test_fp32_time_sync_same_bits_in_every_layout(gpu, case):
    """The FP32 flux form's fused step computes one increment per same-size
    face pair and writes it to both patches; the two-launch layouts compute
    each side's increment on its own.  The states after several steps must
    agree bit for bit."""
    prob = CASES[case]()
    out = []
    for layout in (0, 1, 2):
        print("case", case, "layout", layout, flush=True)
        forest, q0, aux, sv = setup_problem(prob, precision=fa.FP32_BULK, time_sync=True)
        sv.set_ghost_layout(layout)
        dt = P.stable_dt(prob, forest, q0)
        sv.advance(dt, 5)
        out.append(sv.get_q())
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])

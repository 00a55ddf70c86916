"""GPU regrid: the device gradient tags and the move of the solution onto
adapted leaves (copy / interpolate2fine / average2coarse), bit for bit
against the oracle; then the device ghost fill and further steps against the
oracle's, and the drop-in cycle of fclaw2d_run (advance, regrid every few
steps, src/fclaw2d_run.c:281-287) against the same cycle driven on the CPU.
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
import dropin_lib as D
from helpers import oracle_advance, oracle_ghost_update, rel_max_err, setup_problem
from oracle_lib import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12
MBC = 2


def _periodic(prob):
    return dataclasses.replace(prob, periodic=(True, True), mthbc=(2, 2, 2, 2))


CASES = {
    "dambreak": lambda: P.dambreak_problem(2, 5, mx=16),
    "shockbubble": lambda: P.shockbubble_problem(1, 4, mx=16),
    "advection_periodic": lambda: dataclasses.replace(P.advection_problem(2, 4, mx=16)),
    "dambreak_periodic": lambda: _periodic(P.dambreak_problem(2, 4, mx=16)),
}


def _oracle_tags(prob, forest, q, ivar, thr):
    _, _, dx, dy = forest.geometry(prob.extent, prob.mx, prob.mx)
    o = oracle()
    return np.array([o.oracle_patch_exceeds(prob.mx, prob.mx, MBC, np.ascontiguousarray(q[p]), ivar,
                                            float(dx[p]), float(dy[p]), thr)
                     for p in range(len(q))], np.int32)


def _oracle_remap(prob, old, qold, new):
    qnew = np.zeros((len(new[0]),) + qold.shape[1:])
    rc = oracle().oracle_regrid_remap(prob.mx, prob.mx, MBC, prob.meqn, len(old[0]),
                                      *(np.ascontiguousarray(a, dtype=np.int32) for a in old),
                                      np.ascontiguousarray(qold), len(new[0]),
                                      *(np.ascontiguousarray(a, dtype=np.int32) for a in new), qnew)
    assert rc == 0
    return qnew


def _patch_max_gradients(prob, forest, q):
    _, _, dx, dy = forest.geometry(prob.extent, prob.mx, prob.mx)
    f = q[:, prob.tag_field]
    I = slice(MBC, MBC + prob.mx)
    dqx = (f[:, I, MBC + 1:MBC + prob.mx + 1] - f[:, I, MBC - 1:MBC + prob.mx - 1]) / (2 * dx[:, None, None])
    dqy = (f[:, MBC + 1:MBC + prob.mx + 1, I] - f[:, MBC - 1:MBC + prob.mx - 1, I]) / (2 * dy[:, None, None])
    return np.sqrt(dqx * dqx + dqy * dqy).reshape(len(q), -1).max(axis=1)


@pytest.mark.parametrize("case", sorted(CASES))
def test_device_tags_match_oracle(gpu, case):
    prob = CASES[case]()
    forest, q0, aux, sv = setup_problem(prob)
    sv.advance(P.stable_dt(prob, forest, q0) * 0.5, 3)
    q = sv.get_q()
    g = _patch_max_gradients(prob, forest, q)
    u = np.unique(g[g > 0])
    assert len(u) >= 2
    # refine: halfway below the largest patch maximum; coarsen: below the
    # smallest nonzero one, so that both answers occur for both thresholds
    thr_r, thr_c = 0.5 * (u[-2] + u[-1]), 0.5 * u[0]
    er, ec = sv.tag(prob.tag_field, thr_r, thr_c)
    np.testing.assert_array_equal(er, _oracle_tags(prob, forest, q, prob.tag_field, thr_r))
    np.testing.assert_array_equal(ec, _oracle_tags(prob, forest, q, prob.tag_field, thr_c))
    assert 0 < er.sum() < len(er) and 0 < ec.sum()
    # negative refine threshold: always refine; coarsen threshold <= 0: never coarsen
    er, ec = sv.tag(prob.tag_field, -1.0, 0.0)
    assert er.all() and ec.all()


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("seed", [0, 1])
def test_regrid_matches_oracle(gpu, case, seed):
    prob = CASES[case]()
    forest, q0, aux, sv = setup_problem(prob)
    dt = P.stable_dt(prob, forest, q0) * 0.5
    sv.advance(dt, 2)
    qold = sv.get_q()
    old = [a.copy() for a in forest.leaves()]
    rng = np.random.default_rng(seed)
    refine = (rng.random(forest.npatch) < 0.15).astype(np.int32)
    coarsen = (rng.random(forest.npatch) < 0.7).astype(np.int32)
    kind, src = forest.adapt(refine, coarsen, prob.minlevel, prob.maxlevel)
    assert (kind == fa.REGRID_CHILD).any() and (kind == fa.REGRID_PARENT).any()
    new = forest.leaves()
    print("case %s seed %d: %d -> %d patches" % (case, seed, len(old[0]), len(new[0])), flush=True)

    sv.regrid(forest, kind, src, prob.extent)
    qg = sv.get_q()
    qo = _oracle_remap(prob, old, qold, new)
    np.testing.assert_array_equal(qg, qo)

    aux_new = None
    if aux is not None:
        _, aux_new = P.initial_data(prob, forest)
        sv.set_aux(aux_new)
    sv.ghost_fill()
    qg = sv.get_q()
    qo = oracle_ghost_update(prob, forest, qo, prob.mx)
    np.testing.assert_array_equal(qg, qo)

    sv.advance(dt, 2)
    qg = sv.get_q()
    qo, _ = oracle_advance(prob, forest, qo, aux_new, prob.mx, dt, 2)
    assert rel_max_err(qg, qo) <= TOL


def _dambreak_qinit():
    """fort_qinit of the radial dam break (reference qinit signature)."""
    def qinit(maxmx, maxmy, meqn, mbc, mx, my, xlower, ylower, dx, dy, q, maux, aux):
        mx_, my_, mbc_, meqn_ = mx[0], my[0], mbc[0], meqn[0]
        nx, ny = mx_ + 2 * mbc_, my_ + 2 * mbc_
        arr = np.ctypeslib.as_array(q, shape=(meqn_ * ny * nx,)).reshape(meqn_, ny, nx)
        i = np.arange(1 - mbc_, mx_ + mbc_ + 1)
        j = np.arange(1 - mbc_, my_ + mbc_ + 1)
        X, Y = np.meshgrid(xlower[0] + (i - 0.5) * dx[0], ylower[0] + (j - 0.5) * dy[0])
        arr[0] = np.where((X - 0.5) ** 2 + (Y - 0.5) ** 2 < 0.2 ** 2, 2.0, 1.0)
        arr[1] = 0.0
        arr[2] = 0.0
    return D.FORT_QINIT(qinit)


def _domain_leaves(glob):
    dom = glob.contents.domain.contents
    lv, gx, gy = [], [], []
    for b in range(dom.num_blocks):
        blk = dom.blocks[b]
        for p in range(blk.num_patches):
            pt = blk.patches[p]
            lv.append(pt.level)
            w = 2.0 ** -pt.level
            gx.append(int(round(pt.xlower / w)))
            gy.append(int(round(pt.ylower / w)))
    return [np.array(a, np.int32) for a in (lv, gx, gy)]


def test_dropin_initialize_and_regrid_cycle(gpu):
    """fclaw2d_initialize, then advance / fclaw2d_regrid every 2 steps, on the
    dam break; the CPU side repeats every decision with the oracle's tags,
    the same forest adapt and the oracle's transfer, ghost fill and step."""
    L = D.bind()
    prob = P.dambreak_problem(2, 5, mx=16)
    # a coarsen threshold above every gradient: each family not tagged for
    # refinement coarsens where the balance allows, so regrids both refine
    # and coarsen
    thr_r, thr_c = 0.05, 1000.0
    fh = L.fclaw2d_hip_forest_new(1, 1, 0, 0, prob.minlevel)
    dom = L.fclaw2d_domain_new_from_forest(fh, 1, 1, 0, 0, 0.0, 1.0, 0.0, 1.0)
    glob = L.fclaw2d_global_new()
    L.fclaw2d_global_store_domain(glob, dom)
    g = glob.contents
    g.advance_one_step = 1
    g.minlevel, g.maxlevel = prob.minlevel, prob.maxlevel
    g.refine_threshold, g.coarsen_threshold, g.threshold_variable = thr_r, thr_c, 1
    g.refinement_criteria = 3   # FCLAW_REFINE_CRITERIA_GRADIENT
    cp = g.clawpatch_opt
    cp.mx, cp.my, cp.mbc, cp.meqn, cp.maux = 16, 16, 2, 3, 0
    opt = D.Options()
    L.fc2d_hipclaw_options_default(C.byref(opt), 3)
    for k in range(3):
        opt.mthlim[k] = 4
    L.fc2d_hipclaw_options_store(glob, C.byref(opt))
    L.fc2d_hipclaw_solver_initialize(glob)
    vt = L.fc2d_hipclaw_vt(glob).contents
    vt.riemann_solver = fa.RP_SHALLOW
    qinit_cb = _dambreak_qinit()
    vt.fort_qinit = C.cast(qinit_cb, C.c_void_p).value
    L.fc2d_hipclaw_initialize_GPUs(glob)
    L.fclaw2d_initialize(glob)

    # CPU replica of the initial refinement
    forest = fa.Forest(1, 1, (False, False), prob.minlevel)
    zeros = lambda n: np.zeros(n, np.int32)
    for _ in range(prob.minlevel, prob.maxlevel):
        q, _ = P.initial_data(prob, forest)
        q = oracle_ghost_update(prob, forest, q, 16)
        kind, src = forest.adapt(_oracle_tags(prob, forest, q, 0, thr_r), zeros(forest.npatch),
                                 prob.minlevel, prob.maxlevel)
        if np.all(kind == fa.REGRID_SAME) and np.array_equal(src[:, 0], np.arange(len(kind))):
            break
    qo, _ = P.initial_data(prob, forest)
    qo = oracle_ghost_update(prob, forest, qo, 16)
    for a, b in zip(_domain_leaves(glob), forest.leaves()):
        np.testing.assert_array_equal(a, b)
    npatch = forest.npatch
    shape = (3, 20, 20)
    L.fc2d_hipclaw_download(glob)
    np.testing.assert_array_equal(D.host_q(glob, npatch, shape), qo)
    assert len(set(forest.leaves()[0].tolist())) >= 3

    dt = P.stable_dt(prob, forest, qo) * 0.8
    t, new_domains = 0.0, 0
    for cycle in range(4):
        for n in range(2):
            L.fclaw2d_advance_all_levels(glob, t, dt)
            t += dt
        qo, _ = oracle_advance(prob, forest, qo, None, 16, dt, 2)
        L.fclaw2d_regrid(glob)
        old = [a.copy() for a in forest.leaves()]
        er = _oracle_tags(prob, forest, qo, 0, thr_r)
        ec = _oracle_tags(prob, forest, qo, 0, thr_c)
        kind, src = forest.adapt(er, (1 - ec).astype(np.int32), prob.minlevel, prob.maxlevel)
        changed = not (np.all(kind == fa.REGRID_SAME) and np.array_equal(src[:, 0], np.arange(len(kind))))
        if changed:
            qo = _oracle_remap(prob, old, qo, forest.leaves())
            qo = oracle_ghost_update(prob, forest, qo, 16)
            new_domains += 1
        for a, b in zip(_domain_leaves(glob), forest.leaves()):
            np.testing.assert_array_equal(a, b)
        npatch = forest.npatch
        L.fc2d_hipclaw_download(glob)
        qg = D.host_q(glob, npatch, shape).copy()
        err = rel_max_err(qg, qo)
        print("cycle %d: %d patches, changed %d, err %.3e" % (cycle, npatch, changed, err), flush=True)
        assert err <= TOL
    assert g.count_amr_regrid == 4 and g.count_amr_new_domain == new_domains
    assert new_domains >= 1
    L.fclaw2d_global_destroy(glob)

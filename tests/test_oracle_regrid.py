"""Regrid oracle (CPU only): the gradient tag, interpolate2fine,
average2coarse and the repopulation of new leaves.

Pins, each chosen so that a plausible slip fails one of them:
  * the reference's own interpolate2fine / average2coarse / compute_slopes
    Fortran (oracle/_ref/libref46_amr.so), bit for bit, on random patches
    with every ghost cell random;
  * interpolation reproduces a linear field exactly at the fine cell centres
    (a wrong shift sign, a transposed gradient or a wrong child offset fails);
  * averaging the interpolation of a patch gives the patch back (the four
    shifts cancel), and averaging conserves the sum;
  * the gradient tag against closed forms: a linear field has |grad q| =
    sqrt(b^2 + c^2) in every interior cell; a one-cell spike is seen by its
    interior neighbours at s/(2 dx) or s/(2 dy), by one cell from the first
    ghost layer and by none from the second (dx != dy: an axis swap or a
    missing factor 2 moves the switch);
  * repopulation on a small brick equals the per-patch routines.
"""
import numpy as np
import pytest

from oracle_lib import (average2coarse_ref, compute_slopes_ref, interpolate2fine_ref, oracle,
                        ref_amr_available)

requires_ref = pytest.mark.skipif(not ref_amr_available(),
                                  reason="oracle/_ref/libref46_amr.so not built (no reference tree)")

MX, MBC, MEQN = 8, 2, 3
N = MX + 2 * MBC


def patch(rng, meqn=MEQN, mx=MX, mbc=MBC):
    n = mx + 2 * mbc
    return np.ascontiguousarray(rng.standard_normal((meqn, n, n)))


def centres(mx, mbc, x0, h):
    """Cell-centre coordinates of a patch row, ghost cells included."""
    return x0 + (np.arange(-mbc, mx + mbc) + 0.5) * h


def written_fine_range(igrid, mx=MX):
    """Fine index ranges (0-based, into the padded array) interpolate2fine
    writes: its loops cover one ghost layer toward the sibling patches."""
    ig, jg = igrid % 2, igrid // 2
    i1, i2 = 1 - ig, mx // 2 + (1 - ig)
    j1, j2 = 1 - jg, mx // 2 + (1 - jg)
    fi = np.arange((i1 - 1) * 2 + 1, i2 * 2 + 1) + MBC - 1
    fj = np.arange((j1 - 1) * 2 + 1, j2 * 2 + 1) + MBC - 1
    return fi, fj


@requires_ref
def test_compute_slopes_matches_reference():
    o = oracle()
    rng = np.random.default_rng(3)
    vals = list(rng.standard_normal((200, 2))) + [(0.0, 1.0), (1.0, 0.0), (0.0, 0.0), (1.0, -1.0),
                                                   (-2.0, -0.5), (3.0, 3.0), (1e-300, -1e-300)]
    for sl, sr in vals:
        assert o.oracle_compute_slopes(float(sl), float(sr)) == compute_slopes_ref(float(sl), float(sr))


@requires_ref
@pytest.mark.parametrize("igrid", range(4))
def test_interpolate2fine_matches_reference(igrid):
    o = oracle()
    rng = np.random.default_rng(10 + igrid)
    qc = patch(rng)
    qf_o = np.full((MEQN, N, N), 7.25)
    qf_r = qf_o.copy()
    o.oracle_interpolate2fine(MX, MX, MBC, MEQN, qc, qf_o, igrid)
    interpolate2fine_ref(MX, MX, MBC, MEQN, qc, qf_r, igrid)
    np.testing.assert_array_equal(qf_o, qf_r)
    assert np.count_nonzero(qf_o != 7.25) > 0


@requires_ref
@pytest.mark.parametrize("igrid", range(4))
def test_average2coarse_matches_reference(igrid):
    o = oracle()
    rng = np.random.default_rng(20 + igrid)
    qf = patch(rng)
    qc_o = np.full((MEQN, N, N), -3.5)
    qc_r = qc_o.copy()
    o.oracle_average2coarse(MX, MX, MBC, MEQN, qc_o, qf, igrid)
    average2coarse_ref(MX, MX, MBC, MEQN, qc_r, qf, igrid)
    np.testing.assert_array_equal(qc_o, qc_r)


@pytest.mark.parametrize("igrid", range(4))
def test_interpolation_is_exact_for_linear_fields(igrid):
    o = oracle()
    a, b, c = 0.7, -1.3, 2.1
    h = 1.0 / MX
    xc = centres(MX, MBC, 0.0, h)
    qc = np.ascontiguousarray(np.broadcast_to(a + b * xc[None, None, :] + c * xc[None, :, None],
                                              (MEQN, N, N)))
    qf = np.full((MEQN, N, N), np.nan)
    o.oracle_interpolate2fine(MX, MX, MBC, MEQN, qc, qf, igrid)
    ig, jg = igrid % 2, igrid // 2
    xf = centres(MX, MBC, ig * 0.5, h / 2)
    yf = centres(MX, MBC, jg * 0.5, h / 2)
    exact = a + b * xf[None, :] + c * yf[:, None]
    fi, fj = written_fine_range(igrid)
    got = qf[:, fj][:, :, fi]
    np.testing.assert_allclose(got, np.broadcast_to(exact[np.ix_(fj, fi)], got.shape),
                               rtol=0, atol=1e-13)
    # nothing outside the reference's loop ranges is written
    mask = np.ones((N, N), bool)
    mask[np.ix_(fj, fi)] = False
    assert np.all(np.isnan(qf[:, mask]))


@pytest.mark.parametrize("igrid", range(4))
def test_average_of_interpolation_is_identity_and_sums_are_kept(igrid):
    o = oracle()
    rng = np.random.default_rng(30 + igrid)
    qc = patch(rng)
    qf = np.zeros((MEQN, N, N))
    o.oracle_interpolate2fine(MX, MX, MBC, MEQN, qc, qf, igrid)
    back = np.zeros((MEQN, N, N))
    o.oracle_average2coarse(MX, MX, MBC, MEQN, back, qf, igrid)
    ig, jg = igrid % 2, igrid // 2
    sl = (slice(None), slice(MBC + jg * MX // 2, MBC + (jg + 1) * MX // 2),
          slice(MBC + ig * MX // 2, MBC + (ig + 1) * MX // 2))
    np.testing.assert_allclose(back[sl], qc[sl], rtol=0, atol=1e-14)
    # conservation: coarse cell = mean of its four fine cells
    fine_int = qf[:, MBC:MBC + MX, MBC:MBC + MX]
    np.testing.assert_allclose(4 * back[sl].sum(axis=(1, 2)), fine_int.sum(axis=(1, 2)),
                               rtol=1e-13, atol=1e-12)


def test_gradient_tag_linear_field_closed_form():
    """q = a + b x + c y: every centred difference is exact, so every interior
    cell has |grad q| = sqrt(b^2 + c^2) in physical units
    (fclaw2d_clawpatch_gradient_exceeds_th.f90: (q(i+1)-q(i-1))/(2 dx));
    the tag flips between thresholds just below and just above it.  dx != dy
    and b != c, so a swapped axis or a missing factor 2 moves the switch."""
    o = oracle()
    idx = np.arange(1 - MBC, MX + MBC + 1)
    for (b, c, dx, dy) in ((3.0, -1.5, 0.1, 0.05), (0.0, 2.0, 0.02, 0.2), (-4.0, 0.0, 0.3, 0.1)):
        x = (idx - 0.5) * dx
        y = (idx - 0.5) * dy
        q = np.zeros((MEQN, N, N))
        q[1] = 0.7 + b * x[None, :] + c * y[:, None]
        g = float(np.hypot(b, c))
        assert o.oracle_patch_exceeds(MX, MX, MBC, q, 1, dx, dy, g * (1 - 1e-9)) == 1
        assert o.oracle_patch_exceeds(MX, MX, MBC, q, 1, dx, dy, g * (1 + 1e-9)) == 0
        # a constant field in the other variables never tags
        assert o.oracle_patch_exceeds(MX, MX, MBC, q, 0, dx, dy, 1e-300) == 0
        assert o.oracle_patch_exceeds(MX, MX, MBC, q, 1, dx, dy, -1.0) == 1   # always refine


def test_gradient_tag_spike_closed_form():
    """One cell of height s in a zero field: its four neighbours see
    |grad q| = s / (2 dx) (left/right) or s / (2 dy) (below/above); the tag is
    the largest of those that are interior cells.  A spike in the first ghost
    layer is seen by one interior cell only, one in the second ghost layer or
    a ghost corner by none."""
    o = oracle()
    dx, dy, s = 0.1, 0.04, 1.0
    gx, gy = s / (2 * dx), s / (2 * dy)      # 5 and 12.5

    def tag_threshold_switch(i, j):
        q = np.zeros((MEQN, N, N))
        q[2, j + MBC - 1, i + MBC - 1] = s   # cell (i, j), 1-based interior indices
        lo = o.oracle_patch_exceeds(MX, MX, MBC, q, 2, dx, dy, 0.0)
        cases = {}
        for name, g in (("x", gx), ("y", gy)):
            cases[name] = (o.oracle_patch_exceeds(MX, MX, MBC, q, 2, dx, dy, g * (1 - 1e-12)),
                           o.oracle_patch_exceeds(MX, MX, MBC, q, 2, dx, dy, g * (1 + 1e-12)))
        return lo, cases

    # interior spike: both neighbour pairs are interior, the y one is larger
    lo, c = tag_threshold_switch(4, 5)
    assert lo == 1 and c["y"] == (1, 0) and c["x"] == (1, 1)
    # first ghost column i = 0: only cell (1, j) sees it, in x
    lo, c = tag_threshold_switch(0, 5)
    assert lo == 1 and c["x"] == (1, 0) and c["y"] == (0, 0)
    # first ghost row j = my + 1: only cell (i, my) sees it, in y
    lo, c = tag_threshold_switch(3, MX + 1)
    assert lo == 1 and c["y"] == (1, 0)
    # second ghost layer and a ghost corner: no interior cell is adjacent
    for (i, j) in ((-1, 4), (4, MX + 2), (0, 0), (MX + 1, MX + 1)):
        lo, _ = tag_threshold_switch(i, j)
        assert lo == 0, (i, j)


def test_gradient_tag_ignores_ghost_only_gradients():
    o = oracle()
    q = np.zeros((MEQN, N, N))
    q[0, 0, :] = 100.0          # outermost ghost row: not adjacent to any interior cell
    assert o.oracle_patch_exceeds(MX, MX, MBC, q, 0, 0.1, 0.1, 0.01) == 0
    q[0, MBC - 1, MBC + 3] = 1.0  # first ghost row: centred difference of cell (4, 1)
    assert o.oracle_patch_exceeds(MX, MX, MBC, q, 0, 0.1, 0.1, 0.01) == 1
    assert o.oracle_patch_exceeds(MX, MX, MBC, q, 1, 0.1, 0.1, 0.01) == 0


def _leaves(*triples):
    a = np.array(triples, dtype=np.int32).reshape(-1, 3)
    return [np.ascontiguousarray(a[:, k]) for k in range(3)]


def test_repopulation_matches_per_patch_routines():
    o = oracle()
    rng = np.random.default_rng(7)
    # old: level-1 leaves of a 2x1 brick, with block 1's lower-left leaf split
    old = _leaves((1, 0, 0), (1, 1, 0), (1, 0, 1), (1, 1, 1),
                  (2, 4, 0), (2, 5, 0), (2, 4, 1), (2, 5, 1), (1, 3, 0), (1, 2, 1), (1, 3, 1))
    # new: block 0's leaf (0,0) refined, block 1's split leaf coarsened back
    new = _leaves((2, 0, 0), (2, 1, 0), (2, 0, 1), (2, 1, 1), (1, 1, 0), (1, 0, 1), (1, 1, 1),
                  (1, 2, 0), (1, 3, 0), (1, 2, 1), (1, 3, 1))
    qold = np.ascontiguousarray(rng.standard_normal((len(old[0]), MEQN, N, N)))
    qnew = np.zeros((len(new[0]), MEQN, N, N))
    assert o.oracle_regrid_remap(MX, MX, MBC, MEQN, len(old[0]), *old, qold, len(new[0]), *new,
                                 qnew) == 0
    for g in range(4):
        want = np.zeros((MEQN, N, N))
        o.oracle_interpolate2fine(MX, MX, MBC, MEQN, np.ascontiguousarray(qold[0]), want, g)
        np.testing.assert_array_equal(qnew[g], want)
    for p_new, p_old in ((4, 1), (5, 2), (6, 3), (8, 8), (9, 9), (10, 10)):
        np.testing.assert_array_equal(qnew[p_new], qold[p_old])
    want = np.zeros((MEQN, N, N))
    for g in range(4):
        o.oracle_average2coarse(MX, MX, MBC, MEQN, want, np.ascontiguousarray(qold[4 + g]), g)
    np.testing.assert_array_equal(qnew[7], want)
    # a leaf that is none of same / child / parent of four is an error
    bad = _leaves((3, 0, 0))
    assert o.oracle_regrid_remap(MX, MX, MBC, MEQN, len(old[0]), *old, qold, 1, *bad,
                                 np.zeros((1, MEQN, N, N))) == -1

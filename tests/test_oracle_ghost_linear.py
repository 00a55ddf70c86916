"""Pins of the oracle's whole ghost update (copy, average and interpolation
across faces and corners, physical BCs, and the order of fclaw2d_ghost_update)
by what the mathematics fixes, independently of the forest codes:

* A globally linear field q = a + b x + c y, written into the interior cells of
  every leaf of a random 2:1-balanced forest, with every ghost cell set to NaN,
  must come back from one ghost update as the same linear function in every
  ghost cell that lies inside the domain:
    - a same-size copy moves values unchanged,
    - the average of the four fine cells of a coarse ghost cell is the value at
      its centre (fclaw2d_clawpatch46_average.f),
    - the limited interpolation (fclaw2d_clawpatch46_interpolate.f, AMRClaw
      slopes of fclaw2d_clawpatch_utils.f) is exact on linear data wherever
      its coarse 3 x 3 stencil holds the linear field,
  and a NaN left anywhere means a ghost cell read before it was written.  A
  wrong child offset, a mirrored interpolation shift or x and y swapped move
  a value by a multiple of b dx or c dy and fail.
* Across a periodic seam the field is continued periodically, so a ghost cell
  equals the field at its wrapped position; interpolation stencils that cross
  a seam or a physical boundary see the discontinuity (or the extrapolated
  values), and only those cells are left to the closed-form BC check below
  and to the GPU's bit-for-bit comparison.
* Ghost cells outside the domain satisfy the BC relations of
  clawpack46_bc2_default.f on the final data (extrapolation
  q(1-k) = q(1), wall q(1-k) = q(k) with the normal momentum negated), the
  bottom and top faces last, so that the corner cells follow them.

The covering leaf of every ghost cell (same level: copy, finer: average,
coarser: interpolation) is found here by point location in the leaves'
squares, not by either neighbour search.
"""
import numpy as np
import pytest

import forestclaw_amd as fa
from helpers import oracle_ghost_update
from forestclaw_amd.problems import Problem

MBC = 2


def random_forest(nbx, nby, periodic, minlevel, maxlevel, seed, frac=0.3):
    rng = np.random.default_rng(seed)
    forest = fa.Forest(nbx, nby, periodic, minlevel)
    for _ in range(maxlevel - minlevel):
        flags = (rng.random(forest.npatch) < frac).astype(np.int32)
        if flags.any():
            forest.refine(flags, maxlevel)
    return forest


def leaf_squares(forest, extent):
    """Per leaf: level, [x0, x1) x [y0, y1) in physical coordinates, dx, dy."""
    lv, gx, gy = forest.leaves()
    ax, bx, ay, by = extent
    wx, wy = (bx - ax) / forest.nbx, (by - ay) / forest.nby
    s = np.ldexp(1.0, -lv.astype(np.int64))
    x0, y0 = ax + gx * wx * s, ay + gy * wy * s
    return lv, x0, x0 + wx * s, y0, y0 + wy * s


def locate(points_x, points_y, sq):
    """Index of the leaf whose square contains each point (-1: none)."""
    lv, x0, x1, y0, y1 = sq
    out = np.full(points_x.shape, -1, np.int64)
    for p in range(len(lv)):
        inside = (points_x >= x0[p]) & (points_x < x1[p]) & (points_y >= y0[p]) & (points_y < y1[p])
        out[inside] = p
    return out


def field(coef, x, y):
    return np.stack([a + b * x + c * y for a, b, c in coef])


COEF = [(0.75, 0.375, -0.2), (-1.25, -0.15, 0.65), (2.0, 0.05, 0.011)]


def _problem(nbx, nby, periodic, mthbc, extent, mx):
    return Problem("linear", fa.RP_SHALLOW, extent, nbx, nby, periodic, mthbc, 0, 0, mx, 0.9,
                   None)


def _check_linear(nbx, nby, periodic, mthbc, minlevel, maxlevel, seed, mx=8,
                  extent=None):
    extent = extent or (0.0, float(nbx), 0.0, float(nby))
    forest = random_forest(nbx, nby, periodic, minlevel, maxlevel, seed)
    prob = _problem(nbx, nby, periodic, mthbc, extent, mx)
    ax, bx, ay, by = extent
    W, H = bx - ax, by - ay
    sq = leaf_squares(forest, extent)
    lv, x0, x1, y0, y1 = sq
    n = len(lv)
    dx = (x1 - x0) / mx
    dy = (y1 - y0) / mx
    idx = np.arange(1 - MBC, mx + MBC + 1)
    xc = x0[:, None, None] + (idx[None, None, :] - 0.5) * dx[:, None, None]
    yc = y0[:, None, None] + (idx[None, :, None] - 0.5) * dy[:, None, None]
    xc, yc = np.broadcast_arrays(xc, yc)
    q = np.full((n, 3, mx + 2 * MBC, mx + 2 * MBC), np.nan)
    inner = (slice(None), slice(None), slice(MBC, -MBC), slice(MBC, -MBC))
    q[inner] = np.moveaxis(field(COEF, xc, yc), 0, 1)[inner]
    out = oracle_ghost_update(prob, forest, q, mx)
    np.testing.assert_array_equal(out[inner], q[inner])   # interiors untouched

    ghost = np.ones((mx + 2 * MBC, mx + 2 * MBC), bool)
    ghost[MBC:-MBC, MBC:-MBC] = False
    ghost = np.broadcast_to(ghost, xc.shape)
    # wrapped position of every ghost centre; inside the domain or not
    wx_, wy_ = xc.copy(), yc.copy()
    if periodic[0]:
        wx_ = ax + np.mod(wx_ - ax, W)
    if periodic[1]:
        wy_ = ay + np.mod(wy_ - ay, H)
    inside = (wx_ > ax) & (wx_ < bx) & (wy_ > ay) & (wy_ < by)
    g_in = ghost & inside
    owner = locate(wx_[g_in], wy_[g_in], sq)
    assert (owner >= 0).all()
    plev = np.broadcast_to(lv[:, None, None], xc.shape)[g_in]
    olev = lv[owner]
    assert (np.abs(olev - plev) <= 1).all(), "2:1 balance across faces and corners"
    # interpolated cells: the coarse stencil (3 x 3 coarse cells around the
    # coarse cell holding the ghost centre) must not cross a seam or boundary
    dxc = dx[owner]
    dyc = dy[owner]
    interp = olev < plev
    margin_x = np.minimum(wx_[g_in] - ax, bx - wx_[g_in])
    margin_y = np.minimum(wy_[g_in] - ay, by - wy_[g_in])
    # the coarse cell's centre is within half a coarse cell of the point; its
    # neighbours one more cell out: they stay inside when the point is more
    # than two coarse cells from every edge of the domain
    clear = (margin_x > 2 * dxc) & (margin_y > 2 * dyc)
    checked = ~interp | clear
    expect = np.moveaxis(field(COEF, wx_[g_in], wy_[g_in]), 0, 1)   # (cells, meqn)
    got = np.moveaxis(out, 1, -1)[g_in]                              # (cells, meqn)
    assert not np.isnan(got).any(), "%d ghost cells inside the domain never written" % int(
        np.isnan(got).any(axis=1).sum())
    scale = max(abs(a) + abs(b) * max(abs(ax), abs(bx)) + abs(c) * max(abs(ay), abs(by))
                for a, b, c in COEF)
    err = np.abs(got - expect)[checked]
    kinds = np.where(olev == plev, "copy", np.where(olev > plev, "average", "interpolate"))
    bad = np.nonzero(err.max(axis=1) > 1e-13 * scale)[0]
    assert len(bad) == 0, "linear field not reproduced in %d ghost cells (%s), max err %.3e" % (
        len(bad), sorted(set(kinds[checked][bad])), err.max())
    counts = {k: int(np.sum((kinds == k) & checked)) for k in ("copy", "average", "interpolate")}
    # ghost cells outside the domain: the BC relations on the final data
    out_c = ghost & ~inside
    assert not np.isnan(out[:, :, out_c[0]]).any(), "physical ghost cells never written"
    _check_bc(out, forest, mthbc, mx, periodic)
    return counts, int(np.sum(~checked))


def _check_bc(out, forest, mthbc, mx, periodic):
    """clawpack46_bc2_default.f on the final state: left/right faces for the
    rows bottom/top do not rewrite, bottom/top faces for every column."""
    lv, gx, gy = forest.leaves()
    ncx, ncy = forest.nbx << lv, forest.nby << lv
    phys = [(gx == 0) & (not periodic[0]), (gx == ncx - 1) & (not periodic[0]),
            (gy == 0) & (not periodic[1]), (gy == ncy - 1) & (not periodic[1])]
    n = mx + 2 * MBC
    for p in range(len(lv)):
        qp = out[p]
        for face in range(4):
            if not phys[face][p]:
                continue
            bc = mthbc[face]
            for k in range(1, MBC + 1):
                if face < 2:
                    g = MBC - k if face == 0 else MBC + mx - 1 + k
                    src = (MBC if face == 0 else MBC + mx - 1) if bc == 1 else \
                        (MBC - 1 + k if face == 0 else MBC + mx - k)
                    rows = np.arange(n)
                    if phys[2][p]:
                        rows = rows[rows >= MBC]
                    if phys[3][p]:
                        rows = rows[rows < MBC + mx]
                    sign = np.ones(3)
                    if bc == 3:
                        sign[1] = -1
                    np.testing.assert_array_equal(qp[:, rows, g], sign[:, None] * qp[:, rows, src])
                else:
                    g = MBC - k if face == 2 else MBC + mx - 1 + k
                    src = (MBC if face == 2 else MBC + mx - 1) if bc == 1 else \
                        (MBC - 1 + k if face == 2 else MBC + mx - k)
                    sign = np.ones(3)
                    if bc == 3:
                        sign[2] = -1
                    np.testing.assert_array_equal(qp[:, g, :], sign[:, None] * qp[:, src, :])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_linear_field_single_block(seed):
    counts, unchecked = _check_linear(1, 1, (False, False), (1, 1, 1, 1), 2, 5, seed)
    assert all(v > 0 for v in counts.values()), counts
    assert unchecked < sum(counts.values())


@pytest.mark.parametrize("seed", [3, 4])
def test_linear_field_brick_walls(seed):
    """3 x 2 brick of blocks (faces and corners across block boundaries), walls
    left/right and extrapolation top/bottom, on a non-unit extent."""
    counts, _ = _check_linear(3, 2, (False, False), (3, 3, 1, 1), 1, 4, seed,
                              extent=(-1.5, 1.5, 0.0, 1.0))
    assert all(v > 0 for v in counts.values()), counts


@pytest.mark.parametrize("periodic", [(True, False), (False, True), (True, True)])
def test_linear_field_periodic(periodic):
    """Periodic seams: every copy and average across the seam lands on the
    field's periodic continuation."""
    mthbc = tuple(2 if periodic[k // 2] else 1 for k in range(4))
    counts, _ = _check_linear(2, 2, periodic, mthbc, 1, 4, 7 + sum(periodic))
    assert all(v > 0 for v in counts.values()), counts


def test_linear_check_catches_planted_mistakes():
    """The check is sharp: a ghost value moved by one fine cell (a wrong child
    offset) or with x and y swapped is outside its tolerance."""
    mx, L = 8, 3
    dxf = 1.0 / (mx << L)
    x, y = 0.3, 0.6
    exact = field(COEF, np.array(x), np.array(y))
    shifted = field(COEF, np.array(x + dxf), np.array(y))
    swapped = field(COEF, np.array(y), np.array(x))
    scale = 3.0
    assert np.abs(shifted - exact).max() > 1e-6 * scale
    assert np.abs(swapped - exact).max() > 1e-6 * scale

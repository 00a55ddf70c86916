"""Ghost-patch message oracle (CPU only).

Pins:
  * the q/area band against the reference's own
    fclaw2d_clawpatch46_fort_local_ghost_pack (oracle/_ref/libref46_amr.so),
    value for value, for packing and unpacking; the reference's own size
    check (ierror = 0 only when the buffer length equals the band it walks)
    pins the message size;
  * closed form: the band is exactly the frame of cells within mbc ghost and
    refratio*mbc interior layers of the patch boundary, each cell once;
  * the register part: round trip, and every register value appears exactly
    once, followed per face by the non-manifold edge length and area.
"""
import numpy as np
import pytest

from oracle_lib import local_ghost_pack_ref, oracle, ref_amr_available

requires_ref = pytest.mark.skipif(not ref_amr_available(),
                                  reason="oracle/_ref/libref46_amr.so not built (no reference tree)")

MBC = 2
MINT = 4   # refratio 2 * mbc


def _band_size(mx):
    return (mx + 2 * MBC) ** 2 - (mx - 2 * MINT) ** 2


@requires_ref
@pytest.mark.parametrize("mx,meqn", [(8, 1), (16, 3), (32, 3), (16, 4)])
@pytest.mark.parametrize("packarea", [0, 1])
def test_local_ghost_pack_matches_reference(mx, meqn, packarea):
    o = oracle()
    rng = np.random.default_rng(mx * 10 + meqn)
    n = mx + 2 * MBC
    q = np.ascontiguousarray(rng.standard_normal((meqn, n, n)))
    area = np.ascontiguousarray(rng.standard_normal((n + 2, n + 2)))
    size = _band_size(mx) * (meqn + packarea)
    assert o.oracle_ghost_pack_elems(mx, mx, MBC, meqn, 2, packarea, 0) == size
    p_o = np.full(size, np.nan)
    p_r = np.full(size, np.nan)
    assert o.oracle_local_ghost_pack(mx, mx, MBC, meqn, MINT, q, area.ctypes.data, p_o, 2 * packarea) == size
    assert local_ghost_pack_ref(mx, mx, MBC, meqn, MINT, q, area, p_r, 2 * packarea) == 0
    np.testing.assert_array_equal(p_o, p_r)
    # unpacking into fresh arrays restores exactly the band
    qa, qb = np.zeros_like(q), np.zeros_like(q)
    aa, ab = np.zeros_like(area), np.zeros_like(area)
    assert o.oracle_local_ghost_pack(mx, mx, MBC, meqn, MINT, qa, aa.ctypes.data, p_o, 2 * packarea + 1) == size
    assert local_ghost_pack_ref(mx, mx, MBC, meqn, MINT, qb, ab, p_r, 2 * packarea + 1) == 0
    np.testing.assert_array_equal(qa, qb)
    np.testing.assert_array_equal(aa, ab)


@requires_ref
def test_reference_size_check_pins_the_message_length():
    mx, meqn = 16, 3
    n = mx + 2 * MBC
    q = np.zeros((meqn, n, n))
    area = np.zeros((n + 2, n + 2))
    size = _band_size(mx) * meqn
    assert local_ghost_pack_ref(mx, mx, MBC, meqn, MINT, q, area, np.zeros(size), 0) == 0
    assert local_ghost_pack_ref(mx, mx, MBC, meqn, MINT, q, area, np.zeros(size + 1), 0) == 2


@pytest.mark.parametrize("mx", [8, 16, 32])
def test_band_is_the_boundary_frame_once(mx):
    o = oracle()
    n = mx + 2 * MBC
    q = np.ascontiguousarray(np.arange(n * n, dtype=np.float64).reshape(1, n, n))
    pack = np.zeros(_band_size(mx))
    o.oracle_local_ghost_pack(mx, mx, MBC, 1, MINT, q, None, pack, 0)
    got = sorted(pack.astype(int).tolist())
    i = np.arange(n) - MBC + 1          # 1-based cell index incl. ghosts
    I, J = np.meshgrid(i, i)
    frame = (np.minimum(np.minimum(I, J), np.minimum(mx + 1 - I, mx + 1 - J)) <= MINT)
    want = sorted(q[0][frame].astype(int).tolist())
    assert got == want and len(set(got)) == len(got)


@pytest.mark.parametrize("mx,meqn", [(8, 1), (16, 3)])
def test_register_message_layout_and_round_trip(mx, meqn):
    o = oracle()
    nreg = 16 * mx * meqn
    reg = np.ascontiguousarray(np.arange(1, nreg + 1, dtype=np.float64))
    frsize = 2 * (4 * meqn + 2) * (2 * mx)
    pack = np.zeros(frsize)
    dx, dy = 0.25, 0.125
    assert o.oracle_register_pack(mx, mx, meqn, dx, dy, reg, pack, 0) == frsize
    r = reg.reshape(4, 4, meqn, mx)
    k = 0
    for face in range(4):
        for j in range(meqn * mx):
            m, pos = divmod(j, mx)
            np.testing.assert_array_equal(pack[k:k + 4], r[face, :, m, pos])
            k += 4
        for pos in range(mx):
            assert pack[k] == (dy if face < 2 else dx) and pack[k + 1] == dx * dy
            k += 2
    assert k == frsize
    back = np.zeros_like(reg)
    o.oracle_register_pack(mx, mx, meqn, dx, dy, back, pack, 1)
    np.testing.assert_array_equal(back, reg)

"""ASCII frame output (fort.tXXXX / fort.qXXXX) against the reference's own
Fortran writer (src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_output_ascii.f,
compiled in place into oracle/_ref): byte-identical files for patches with
positive, negative, zero, tiny (< 1e-99, written as 0) and huge (three-digit
exponent) values.  No GPU is used."""
import ctypes as C
import os

import numpy as np
import pytest

import forestclaw_amd as fa
from oracle_lib import ref, ref_available

pytestmark = pytest.mark.skipif(not ref_available(), reason="reference Fortran not built")


def _lib():
    L = fa.lib()
    L.fclaw2d_clawpatch_time_header_ascii_files.argtypes = [C.c_char_p, C.c_char_p, C.c_double,
                                                           C.c_int, C.c_int, C.c_int]
    L.fclaw2d_clawpatch_output_ascii_patch.argtypes = [C.c_char_p] + [C.c_int] * 4 + \
        [C.c_double] * 4 + [np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")] + [C.c_int] * 4
    return L


def _ref_write(dirname, time, patches, meqn, maux):
    R = ref()
    ip = lambda v: C.byref(C.c_int(v))
    dp = lambda v: C.byref(C.c_double(v))
    n1 = C.create_string_buffer(b"fort.q0007 ", 11)
    n2 = C.create_string_buffer(b"fort.t0007 ", 11)
    cwd = os.getcwd()
    os.chdir(dirname)
    try:
        R.fclaw2d_clawpatch46_fort_header_ascii_(n1, n2, dp(time), ip(meqn), ip(maux),
                                                 ip(len(patches)), C.c_size_t(11), C.c_size_t(11))
        for (mx, mbc, xl, yl, dx, dy, q, num, level, blk) in patches:
            qq = np.ascontiguousarray(q).copy()
            R.fclaw2d_clawpatch46_fort_output_ascii_(
                n1, ip(mx), ip(mx), ip(meqn), ip(mbc), dp(xl), dp(yl), dp(dx), dp(dy),
                qq.ctypes.data_as(C.c_void_p), ip(num), ip(level), ip(blk), ip(0), C.c_size_t(11))
    finally:
        os.chdir(cwd)


def test_ascii_frame_matches_reference_writer(tmp_path):
    L = _lib()
    rng = np.random.default_rng(3)
    meqn, mbc = 3, 2
    patches = []
    for k, mx in enumerate((4, 8)):
        q = rng.normal(0, 1, (meqn, mx + 2 * mbc, mx + 2 * mbc)) * 10.0 ** rng.integers(-30, 30, (meqn, 1, 1))
        q[0, mbc, mbc] = 1e-120
        q[1, mbc, mbc + 1] = -2.5e150
        q[2, mbc + 1, mbc] = 0.0
        q[0, mbc + 1, mbc + 1] = 9.99999999999999999e99
        q[1, mbc + 2, mbc] = -1e-99
        patches.append((mx, mbc, 0.125 * k, -0.375, 1.0 / mx, 1.0 / mx, q, k, 3 + k, k))
    mine, theirs = tmp_path / "mine", tmp_path / "ref"
    mine.mkdir()
    theirs.mkdir()
    time = 0.123456789012345678
    assert L.fclaw2d_clawpatch_time_header_ascii_files(str(mine / "fort.q0007").encode(),
                                                       str(mine / "fort.t0007").encode(),
                                                       time, meqn, 0, len(patches)) == 0
    for (mx, mbc_, xl, yl, dx, dy, q, num, level, blk) in patches:
        assert L.fclaw2d_clawpatch_output_ascii_patch(str(mine / "fort.q0007").encode(), mx, mx, meqn,
                                                      mbc_, xl, yl, dx, dy,
                                                      np.ascontiguousarray(q), num, level, blk, 0) == 0
    _ref_write(str(theirs), time, patches, meqn, 0)
    for name in ("fort.t0007", "fort.q0007"):
        a = (mine / name).read_bytes()
        b = (theirs / name).read_bytes()
        assert a == b, (name, a[:400], b[:400])

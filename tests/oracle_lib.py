"""ctypes bindings to the CPU oracle (oracle/liboracle46.so) and to the
reference's own Fortran built in place (oracle/_ref/libref46.so, and
oracle/_ref/libref46_amr.so for the regrid transfer routines).

Test infrastructure only: the product never imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle46.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref46.so")
REF_ACOUSTICS_SO = os.path.join(ROOT, "oracle", "_ref", "libref46_acoustics.so")
REF_AMR_SO = os.path.join(ROOT, "oracle", "_ref", "libref46_amr.so")
REF_SYNC_SO = os.path.join(ROOT, "oracle", "_ref", "libref46_sync.so")

RP_ADVECTION, RP_SHALLOW, RP_EULER = 0, 1, 2

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OracleClaw(C.Structure):
    _fields_ = [("solver", C.c_int), ("meqn", C.c_int), ("maux", C.c_int),
                ("mwaves", C.c_int), ("method", C.c_int * 7), ("mthlim", C.c_int * 10),
                ("grav", C.c_double), ("gamma", C.c_double), ("p", C.c_double * 8)]


class OracleDomain(C.Structure):
    _fields_ = [("npatch", C.c_int), ("level", C.POINTER(C.c_int)),
                ("gx", C.POINTER(C.c_int)), ("gy", C.POINTER(C.c_int)),
                ("nbx", C.c_int), ("nby", C.c_int), ("periodic_x", C.c_int),
                ("periodic_y", C.c_int), ("mx", C.c_int), ("my", C.c_int),
                ("mbc", C.c_int), ("meqn", C.c_int), ("maux", C.c_int),
                ("mthbc", C.c_int * 4), ("ax", C.c_double), ("bx", C.c_double),
                ("ay", C.c_double), ("by", C.c_double)]


_oracle = None
_ref = None
_ref_amr = None
_ref_sync = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        lib = C.CDLL(ORACLE_SO)
        lib.oracle_step2_patch.argtypes = [C.POINTER(OracleClaw), C.c_int, C.c_int, C.c_int,
                                           _dp, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                           C.POINTER(C.c_double), C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p]
        lib.oracle_step2_patch.restype = C.c_int
        lib.oracle_ghost_update.argtypes = [C.POINTER(OracleDomain), _dp]
        lib.oracle_update_all.argtypes = [C.POINTER(OracleDomain), C.POINTER(OracleClaw), _dp,
                                          C.c_void_p, C.c_double]
        lib.oracle_update_all.restype = C.c_double
        lib.oracle_advance.argtypes = [C.POINTER(OracleDomain), C.POINTER(OracleClaw), _dp,
                                       C.c_void_p, C.c_double, C.c_int]
        lib.oracle_advance.restype = C.c_double
        lib.oracle_advance_ts.argtypes = [C.POINTER(OracleDomain), C.POINTER(OracleClaw), _dp,
                                          C.c_void_p, C.c_double, C.c_int, C.c_int]
        lib.oracle_advance_ts.restype = C.c_double
        lib.oracle_advance_subcycle.argtypes = [C.POINTER(OracleDomain), C.POINTER(OracleClaw),
                                                _dp, C.c_void_p, C.c_double, C.c_int]
        lib.oracle_advance_subcycle.restype = C.c_double
        lib.oracle_patch_exceeds.argtypes = [C.c_int] * 3 + [_dp, C.c_int] + [C.c_double] * 3
        lib.oracle_regrid_remap.argtypes = [C.c_int] * 5 + [_ip] * 3 + [_dp, C.c_int] + [_ip] * 3 + [_dp]
        lib.oracle_ghost_pack_elems.argtypes = [C.c_int] * 7
        lib.oracle_ghost_pack_elems.restype = C.c_size_t
        lib.oracle_local_ghost_pack.argtypes = [C.c_int] * 5 + [_dp, C.c_void_p, _dp, C.c_int]
        lib.oracle_local_ghost_pack.restype = C.c_long
        lib.oracle_register_pack.argtypes = [C.c_int] * 3 + [C.c_double] * 2 + [_dp, _dp, C.c_int]
        lib.oracle_register_pack.restype = C.c_long
        lib.oracle_compute_slopes.argtypes = [C.c_double, C.c_double]
        lib.oracle_compute_slopes.restype = C.c_double
        lib.oracle_interpolate2fine.argtypes = [C.c_int] * 4 + [_dp, _dp, C.c_int]
        lib.oracle_average2coarse.argtypes = [C.c_int] * 4 + [_dp, _dp, C.c_int]
        lib.oracle_cons_flux.argtypes = [C.POINTER(OracleClaw), C.c_int, _dp, _dp, _dp]
        lib.oracle_mass.argtypes = [C.POINTER(OracleDomain), _dp, _dp]
        lib.oracle_bc2.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _ip]
        lib.oracle_inlinelimiter.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                             _dp, _dp, _ip]
        for name, nint in (("oracle_copy_face", 3), ("oracle_copy_corner", 3),
                           ("oracle_average_face", 4), ("oracle_average_corner", 3)):
            getattr(lib, name).argtypes = [C.c_int] * 4 + [_dp, _dp] + [C.c_int] * nint
        lib.oracle_interpolate_face.argtypes = [C.c_int] * 4 + [_dp, _dp] + [C.c_int] * 4
        lib.oracle_interpolate_corner.argtypes = [C.c_int] * 4 + [_dp, _dp] + [C.c_int] * 3
        lib.oracle_timeinterp_patch.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, C.c_double]
        lib.oracle_store_flux.argtypes = [C.POINTER(OracleClaw), C.c_int, C.c_int, C.c_int, _dp,
                                          C.c_void_p, C.c_double, C.c_double, C.c_double, _dp]
        lib.oracle_accumulate_waves.argtypes = [C.c_int] * 4 + [C.c_double] * 3 + [_dp] * 5
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        lib.ref46_step2.argtypes = [C.c_int] * 7 + [_ip, _ip, _dp, C.c_void_p] + \
            [C.c_double] * 5 + [C.POINTER(C.c_double), _dp, _dp, _dp, _dp]
        lib.ref46_step2.restype = C.c_int
        lib.ref46_bc2.argtypes = [C.c_int] * 4 + [_dp, _ip]
        lib.ref46_inlinelimiter.argtypes = [C.c_int] * 5 + [_dp, _dp, _ip]
        ip = C.POINTER(C.c_int)
        lib.rpn2_cons_update_.argtypes = [ip, ip, ip, ip, _dp, _dp, _dp, _dp]
        _ref = lib
    return _ref


def make_claw(solver, meqn, maux, mwaves, order=(2, 2), mthlim=None, grav=1.0, gamma=1.4,
              params=None):
    c = OracleClaw()
    c.solver, c.meqn, c.maux, c.mwaves = solver, meqn, maux, mwaves
    method = [0, order[0], order[1], 0, 0, 0, maux]
    for k in range(7):
        c.method[k] = method[k]
    lim = list(mthlim) if mthlim is not None else [4] * mwaves
    for k in range(10):
        c.mthlim[k] = lim[k] if k < len(lim) else 0
    c.grav, c.gamma = grav, gamma
    for k, v in enumerate(params or ()):
        c.p[k] = v
    return c


def step2_oracle(claw, q, aux, dx, dy, dt, want_fluxes=False):
    """q: (meqn, my+2mbc, mx+2mbc) float64, updated copy returned with cfl."""
    meqn, ny, nx = q.shape
    mbc = 2
    mx, my = nx - 2 * mbc, ny - 2 * mbc
    qq = np.ascontiguousarray(q, dtype=np.float64).copy()
    cfl = C.c_double(0.0)
    fl = [np.zeros_like(qq) for _ in range(4)] if want_fluxes else [None] * 4
    auxp = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    rc = oracle().oracle_step2_patch(C.byref(claw), mx, my, mbc, qq,
                                     None if auxp is None else auxp.ctypes.data,
                                     dx, dy, dt, C.byref(cfl),
                                     *[None if f is None else f.ctypes.data for f in fl])
    assert rc == 0
    if want_fluxes:
        return qq, cfl.value, fl
    return qq, cfl.value


def step2_ref(claw, q, aux, dx, dy, dt, want_fluxes=False):
    meqn, ny, nx = q.shape
    mbc = 2
    mx, my = nx - 2 * mbc, ny - 2 * mbc
    qq = np.ascontiguousarray(q, dtype=np.float64).copy()
    cfl = C.c_double(0.0)
    fm, fp, gm, gp = (np.zeros_like(qq) for _ in range(4))
    method = np.array(list(claw.method), dtype=np.int32)
    mthlim = np.array(list(claw.mthlim), dtype=np.int32)
    auxp = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    rc = ref().ref46_step2(claw.solver, mx, my, mbc, meqn, claw.maux, claw.mwaves, method,
                           mthlim, qq, None if auxp is None else auxp.ctypes.data,
                           dx, dy, dt, claw.grav, claw.gamma, C.byref(cfl), fm, fp, gm, gp)
    assert rc == 0
    if want_fluxes:
        return qq, cfl.value, [fm, fp, gm, gp]
    return qq, cfl.value


_ref_ac = None


def step2_ref_acoustics(claw, q, dx, dy, dt):
    """The reference's own patch step with its acoustics solvers
    (oracle/_ref/libref46_acoustics.so); claw.p = (rho, bulk, ...)."""
    global _ref_ac
    if _ref_ac is None:
        lib = C.CDLL(REF_ACOUSTICS_SO)
        lib.ref46_acoustics_step2.argtypes = [C.c_int] * 3 + [_ip, _ip, _dp] + [C.c_double] * 5 + \
            [C.POINTER(C.c_double), _dp, _dp, _dp, _dp]
        lib.ref46_acoustics_step2.restype = C.c_int
        _ref_ac = lib
    meqn, ny, nx = q.shape
    mbc = 2
    qq = np.ascontiguousarray(q, dtype=np.float64).copy()
    cfl = C.c_double(0.0)
    fl = [np.zeros_like(qq) for _ in range(4)]
    rc = _ref_ac.ref46_acoustics_step2(nx - 2 * mbc, ny - 2 * mbc, mbc,
                                       np.array(list(claw.method), dtype=np.int32),
                                       np.array(list(claw.mthlim), dtype=np.int32), qq,
                                       dx, dy, dt, claw.p[0], claw.p[1], C.byref(cfl), *fl)
    assert rc == 0
    return qq, cfl.value, fl


def domain_struct(level, gx, gy, nbx, nby, periodic, mx, my, mbc, meqn, maux, mthbc, extent):
    """Build an OracleDomain; keeps the int arrays alive on the returned object."""
    d = OracleDomain()
    arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (level, gx, gy)]
    d.npatch = len(arrs[0])
    d.level, d.gx, d.gy = (a.ctypes.data_as(C.POINTER(C.c_int)) for a in arrs)
    d.nbx, d.nby = nbx, nby
    d.periodic_x, d.periodic_y = int(periodic[0]), int(periodic[1])
    d.mx, d.my, d.mbc, d.meqn, d.maux = mx, my, mbc, meqn, maux
    for k in range(4):
        d.mthbc[k] = mthbc[k]
    d.ax, d.bx, d.ay, d.by = extent
    d._keep = arrs
    return d


def ref_amr_available():
    return os.path.exists(REF_AMR_SO)


def ref_amr():
    """The reference's fclaw2d_clawpatch46_fort_interpolate2fine /
    fclaw2d_clawpatch46_fort_average2coarse / fclaw2d_clawpatch_compute_slopes
    (Fortran calling convention: every argument by reference)."""
    global _ref_amr
    if _ref_amr is None:
        lib = C.CDLL(REF_AMR_SO)
        ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
        for name in ("fclaw2d_clawpatch46_fort_interpolate2fine_",
                     "fclaw2d_clawpatch46_fort_average2coarse_"):
            getattr(lib, name).argtypes = [ip] * 4 + [_dp, _dp, _dp, _dp, ip, ip]
        lib.fclaw2d_clawpatch46_fort_local_ghost_pack_.argtypes = [ip] * 5 + [_dp, _dp, _dp] + [ip] * 3
        lib.fclaw2d_clawpatch46_fort_timeinterp_.argtypes = [ip] * 5 + [_dp, _dp, _dp, dp, ip]
        lib.fclaw2d_clawpatch_compute_slopes_.argtypes = [dp, dp, ip]
        lib.fclaw2d_clawpatch_compute_slopes_.restype = C.c_double
        _ref_amr = lib
    return _ref_amr


def _transfer_ref(name, mx, my, mbc, meqn, qcoarse, qfine, igrid):
    area = np.zeros((my + 2 * mbc + 2) * (mx + 2 * mbc + 2))
    i = lambda v: C.byref(C.c_int(v))
    getattr(ref_amr(), name)(i(mx), i(my), i(mbc), i(meqn), qcoarse, qfine, area, area.copy(),
                             i(igrid), i(0))


def interpolate2fine_ref(mx, my, mbc, meqn, qcoarse, qfine, igrid):
    """Reference interpolate2fine (manifold = 0) into qfine, in place."""
    _transfer_ref("fclaw2d_clawpatch46_fort_interpolate2fine_", mx, my, mbc, meqn, qcoarse, qfine,
                  igrid)


def average2coarse_ref(mx, my, mbc, meqn, qcoarse, qfine, igrid):
    """Reference average2coarse (manifold = 0) into qcoarse, in place."""
    _transfer_ref("fclaw2d_clawpatch46_fort_average2coarse_", mx, my, mbc, meqn, qcoarse, qfine,
                  igrid)


def compute_slopes_ref(sl, sr, mth=5):
    return ref_amr().fclaw2d_clawpatch_compute_slopes_(C.byref(C.c_double(sl)),
                                                       C.byref(C.c_double(sr)),
                                                       C.byref(C.c_int(mth)))


def local_ghost_pack_ref(mx, my, mbc, meqn, mint, q, area, qpack, packmode):
    """Reference fclaw2d_clawpatch46_fort_local_ghost_pack; returns ierror."""
    i = lambda v: C.byref(C.c_int(v))
    ierr = C.c_int(-1)
    ref_amr().fclaw2d_clawpatch46_fort_local_ghost_pack_(i(mx), i(my), i(mbc), i(meqn), i(mint),
                                                         q, area, qpack, i(len(qpack)),
                                                         i(packmode), C.byref(ierr))
    return ierr.value


def timeinterp_ref(mx, my, mbc, meqn, qcurr, qlast, qinterp, alpha, psize):
    """Reference fclaw2d_clawpatch46_fort_timeinterp into qinterp, in place;
    returns its ierror (2 when psize is not the band's size)."""
    i = lambda v: C.byref(C.c_int(v))
    ierr = C.c_int(-1)
    ref_amr().fclaw2d_clawpatch46_fort_timeinterp_(i(mx), i(my), i(mbc), i(meqn), i(psize), qcurr,
                                                   qlast, qinterp, C.byref(C.c_double(alpha)),
                                                   C.byref(ierr))
    return ierr.value


def ref_sync_available():
    return os.path.exists(REF_SYNC_SO) and os.path.exists(REF_SO)


def ref_sync():
    """The reference's clawpack46_time_sync_store_flux / _accumulate_waves."""
    global _ref_sync
    if _ref_sync is None:
        lib = C.CDLL(REF_SYNC_SO)
        ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
        lib.clawpack46_time_sync_store_flux_.argtypes = [ip] * 7 + [dp] + [_dp] * 4 + \
            [_dp, C.c_void_p] + [_dp] * 4 + [C.c_void_p] + [_dp] * 4
        lib.clawpack46_time_sync_accumulate_waves_.argtypes = [ip] * 4 + [dp] * 3 + [ip] + \
            [_dp] * 4 + [_dp] * 4 + [_dp] * 8
        _ref_sync = lib
    return _ref_sync

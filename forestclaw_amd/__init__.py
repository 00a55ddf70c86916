"""forestclaw_amd: MI355X-native batched patch update for ForestClaw.

Python access to the in-tree library libforestclaw_amd.so (HIP kernels for
gfx950 + the C host layer).  The library is required: importing the package
works without it, but every object that computes raises if it is missing --
there is no CPU fallback.

Main objects
  Forest  -- leaves of a 2:1-balanced forest of quadtrees over a brick of
             blocks, and the device ghost-fill program for it
  Solver  -- one resident solution on the GPU (include/hipclaw.h)
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libforestclaw_amd.so")

RP_ADVECTION, RP_SHALLOW, RP_EULER = 0, 1, 2
FP64, FP32_BULK = 0, 1
AUTOTUNE_MAX_TRIALS = 16   # HIPCLAW_AUTOTUNE_MAX_TRIALS
# provenance of a leaf after a forest adapt (fclaw2d_hip_forest_provenance)
REGRID_SAME, REGRID_CHILD, REGRID_PARENT = 0, 1, 2
SOLVER_SHAPES = {RP_ADVECTION: (1, 1, 2), RP_SHALLOW: (3, 3, 0), RP_EULER: (4, 4, 0)}

_lib = None
_user_libs = []      # loaded solver libraries (kept alive)
_user_shapes = {}    # registered solver number -> (meqn, mwaves, maux)


class HipclawSolverDesc(C.Structure):
    """hipclaw_solver_t (include/hipclaw_solver.h)."""
    _fields_ = [("abi", C.c_int), ("step_args_size", C.c_int), ("name", C.c_char_p),
                ("meqn", C.c_int), ("mwaves", C.c_int), ("maux", C.c_int),
                ("conservative", C.c_int), ("time_sync", C.c_int), ("tiles", C.c_int),
                ("launch_fp64", C.c_void_p), ("launch_fp32", C.c_void_p)]


class HipclawConfig(C.Structure):
    _fields_ = [("solver", C.c_int), ("precision", C.c_int), ("mx", C.c_int),
                ("my", C.c_int), ("mbc", C.c_int), ("meqn", C.c_int), ("maux", C.c_int),
                ("mwaves", C.c_int), ("order", C.c_int * 2), ("mthlim", C.c_int * 10),
                ("mthbc", C.c_int * 4), ("grav", C.c_double), ("gamma", C.c_double),
                ("tile", C.c_int), ("time_sync", C.c_int), ("subcycle", C.c_int),
                ("rp_params", C.c_double * 8)]


def _dp():
    return np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _ip():
    return np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    """Load libforestclaw_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("forestclaw_amd: %s is missing; build it with "
                           "`python -m forestclaw_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    sig = {
        "hipclaw_last_error": ([], C.c_char_p),
        "hipclaw_initialize_device": ([i, C.c_char_p, C.c_size_t], i),
        "hipclaw_create": ([C.POINTER(HipclawConfig), i, C.POINTER(vp)], i),
        "hipclaw_destroy": ([vp], None),
        "hipclaw_tile": ([vp], i),
        "hipclaw_set_geometry": ([vp, i, _dp(), _dp()], i),
        "hipclaw_set_q": ([vp, _dp(), i, i], i),
        "hipclaw_get_q": ([vp, _dp(), i, i], i),
        "hipclaw_set_aux": ([vp, _dp(), i, i], i),
        "hipclaw_set_ghost_program": ([vp, _ip(), i, _ip()], i),
        "hipclaw_ghost_fill": ([vp], i),
        "hipclaw_set_sync_program": ([vp, _ip(), i], i),
        "hipclaw_time_sync": ([vp], i),
        "hipclaw_step_batch": ([vp, _ip(), i, d, d, C.POINTER(d)], i),
        "hipclaw_advance": ([vp, d, i, C.POINTER(d)], i),
        "hipclaw_advance_async": ([vp, d, i], i),
        "hipclaw_sync": ([vp], i),
        "hipclaw_capture_graph": ([vp, d, i], i),
        "hipclaw_replay_graph": ([vp], i),
        "hipclaw_mass": ([vp, _dp(), _dp()], i),
        "hipclaw_timing": ([vp, i, C.POINTER(d), C.POINTER(d)], i),
        "fclaw2d_hip_forest_new": ([i, i, i, i, i], vp),
        "fclaw2d_hip_forest_from_leaves": ([i, i, i, i, i, _ip(), _ip(), _ip(), i], vp),
        "fclaw2d_hip_forest_destroy": ([vp], None),
        "fclaw2d_hip_forest_num_patches": ([vp], i),
        "fclaw2d_hip_forest_min_level": ([vp], i),
        "fclaw2d_hip_forest_max_level": ([vp], i),
        "fclaw2d_hip_forest_leaves": ([vp, _ip(), _ip(), _ip()], None),
        "fclaw2d_hip_forest_refine": ([vp, _ip(), i], i),
        "fclaw2d_hip_forest_ghost_program": ([vp, i, i, i, _ip(), vp, i, _ip()], i),
        "fclaw2d_hip_forest_face_relation": ([vp, i, i, _ip()], i),
        "fclaw2d_hip_forest_sync_program": ([vp, i, i, vp, i], i),
        "fclaw2d_hip_forest_adapt": ([vp, _ip(), _ip(), i, i], i),
        "fclaw2d_hip_forest_set_isolated": ([vp, i], i),
        "hipclaw_set_dt_scale": ([vp, vp], i),
        "hipclaw_save_step": ([vp], i),
        "hipclaw_restore_step": ([vp], i),
        "hipclaw_autotune": ([vp, C.c_double, i, vp, vp, vp, vp, vp], i),
        "hipclaw_set_waves": ([vp, i], i),
        "hipclaw_waves": ([vp], i),
        "hipclaw_set_ghost_layout": ([vp, i], i),
        "hipclaw_ghost_layout": ([vp], i),
        "fclaw2d_hip_forest_provenance": ([vp, _ip(), _ip()], i),
        "hipclaw_tag_patches": ([vp, i, d, d, _ip(), _ip()], i),
        "hipclaw_regrid": ([vp, i, _ip(), _ip(), _dp(), _dp()], i),
        "hipclaw_ghost_pack_elems": ([vp, C.POINTER(C.c_longlong)], i),
        "hipclaw_ghost_pack": ([vp, _ip(), i, i, vp], i),
        "hipclaw_ghost_unpack": ([vp, _ip(), i, i, vp], i),
        "hipclaw_get_registers": ([vp, i, i, _dp()], i),
        "hipclaw_register_solver": ([vp, C.POINTER(i)], i),
        "fclaw2d_hip_forest_ghost_program_range": ([vp, i, i, i, _ip(), i, i, i, vp, i, _ip()], i),
        "fclaw2d_hip_forest_sync_program_range": ([vp, i, i, i, vp, i, vp, i, C.POINTER(i)], i),
        "hipclaw_set_ghost_program_k": ([vp, i, _ip(), i, _ip()], i),
        "hipclaw_set_sync_program_k": ([vp, i, _ip(), i, _ip(), i], i),
        "hipclaw_set_levels": ([vp, _ip(), i], i),
        "hipclaw_advance_subcycle": ([vp, d, i, C.POINTER(d)], i),
        "hipclaw_advance_subcycle_async": ([vp, d, i], i),
        "hipclaw_capture_subcycle_graph": ([vp, d, i], i),
        "hipclaw_step_batch_noalloc": ([C.POINTER(HipclawConfig), i, vp, vp, vp, vp, vp, _ip(), i,
                                        d, vp, C.POINTER(d)], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


class HipclawError(RuntimeError):
    pass


def solver_shape(solver):
    """(meqn, mwaves, maux) of a built-in or registered Riemann solver."""
    if solver in SOLVER_SHAPES:
        return SOLVER_SHAPES[solver]
    if solver in _user_shapes:
        return _user_shapes[solver]
    raise ValueError("unknown Riemann solver %r" % (solver,))


def register_solver(path, name):
    """Load the shared library of an application's own Riemann solver (built
    with csrc/hipclaw_user_solver.h, exporting `<name>_solver`) and register
    it; returns the solver number for Solver(solver=...)."""
    L = C.CDLL(os.path.abspath(path))
    fn = getattr(L, name + "_solver")
    fn.restype = C.POINTER(HipclawSolverDesc)
    fn.argtypes = []
    d = fn()
    sid = C.c_int(-1)
    _check(lib().hipclaw_register_solver(C.cast(d, C.c_void_p), C.byref(sid)),
           "hipclaw_register_solver")
    _user_libs.append(L)
    _user_shapes[sid.value] = (d.contents.meqn, d.contents.mwaves, d.contents.maux)
    return sid.value


def example_solver(name):
    """Register one of the example solvers built in user_solvers/
    ("acoustics", "tracers")."""
    from .build import user_solver_lib
    return register_solver(user_solver_lib(name), name)


def _check(rc, what):
    if rc != 0:
        msg = lib().hipclaw_last_error()
        raise HipclawError("%s failed (code %d): %s" % (what, rc, msg.decode() if msg else ""))


def initialize_device(device=0):
    """Select the GPU (fc2d_cudaclaw_initialize_GPUs); returns its name."""
    buf = C.create_string_buffer(256)
    _check(lib().hipclaw_initialize_device(device, buf, 256), "hipclaw_initialize_device")
    return buf.value.decode()


class Forest:
    """Leaves of a balanced forest of quadtrees over an nbx x nby brick."""

    def __init__(self, nbx=1, nby=1, periodic=(False, False), minlevel=0, leaves=None,
                 balance=True, isolated=False):
        L = lib()
        self.nbx, self.nby = nbx, nby
        self.periodic = (bool(periodic[0]), bool(periodic[1]))
        self.isolated = bool(isolated)
        if leaves is None:
            self._h = L.fclaw2d_hip_forest_new(nbx, nby, int(periodic[0]), int(periodic[1]),
                                               minlevel)
        else:
            lv, gx, gy = (np.ascontiguousarray(a, dtype=np.int32) for a in leaves)
            self._h = L.fclaw2d_hip_forest_from_leaves(nbx, nby, int(periodic[0]),
                                                       int(periodic[1]), len(lv), lv, gx, gy,
                                                       int(balance))
        if not self._h:
            raise ValueError("invalid forest")
        if isolated:
            L.fclaw2d_hip_forest_set_isolated(self._h, 1)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.fclaw2d_hip_forest_destroy(self._h)
            self._h = None

    @property
    def npatch(self):
        return lib().fclaw2d_hip_forest_num_patches(self._h)

    @property
    def minlevel(self):
        return lib().fclaw2d_hip_forest_min_level(self._h)

    @property
    def maxlevel(self):
        return lib().fclaw2d_hip_forest_max_level(self._h)

    def leaves(self):
        n = self.npatch
        lv, gx, gy = (np.zeros(n, np.int32) for _ in range(3))
        lib().fclaw2d_hip_forest_leaves(self._h, lv, gx, gy)
        return lv, gx, gy

    def refine(self, flags, maxlevel):
        flags = np.ascontiguousarray(flags, dtype=np.int32)
        assert len(flags) == self.npatch
        n = lib().fclaw2d_hip_forest_refine(self._h, flags, maxlevel)
        if n < 0:
            raise RuntimeError("forest refinement failed")
        return n

    def ghost_program(self, mx, my, mbc, mthbc):
        mth = np.ascontiguousarray(mthbc, dtype=np.int32)
        ps = np.zeros(7, np.int32)
        ops = np.empty((16 * self.npatch + 16, 8), np.int32)   # room for the usual forest
        n = lib().fclaw2d_hip_forest_ghost_program(self._h, mx, my, mbc, mth, ops.ctypes.data,
                                                   len(ops), ps)
        if n < 0:
            raise RuntimeError("ghost program failed")
        if n > len(ops):
            ops = np.zeros((n, 8), np.int32)
            n2 = lib().fclaw2d_hip_forest_ghost_program(self._h, mx, my, mbc, mth,
                                                        ops.ctypes.data, n, ps)
            assert n2 == n
        return ops[:n], ps

    def ghost_program_range(self, mx, my, mbc, mthbc, minlevel, maxlevel, time_interp):
        """Ghost fill of levels minlevel..maxlevel, level minlevel - 1 taken
        from its time-interpolated copies when time_interp (the partial
        fclaw2d_ghost_update of a subcycled step)."""
        mth = np.ascontiguousarray(mthbc, dtype=np.int32)
        ps = np.zeros(7, np.int32)
        n = lib().fclaw2d_hip_forest_ghost_program_range(self._h, mx, my, mbc, mth, minlevel,
                                                         maxlevel, int(time_interp), None, 0, ps)
        if n < 0:
            raise RuntimeError("ghost program failed")
        ops = np.zeros((max(n, 1), 8), np.int32)
        lib().fclaw2d_hip_forest_ghost_program_range(self._h, mx, my, mbc, mth, minlevel, maxlevel,
                                                     int(time_interp), ops.ctypes.data, n, ps)
        return ops[:n], ps

    def sync_program_range(self, mx, my, minlevel):
        """Time-sync ops of levels minlevel.. and their register resets."""
        nres = C.c_int(0)
        n = lib().fclaw2d_hip_forest_sync_program_range(self._h, mx, my, minlevel, None, 0, None, 0,
                                                        C.byref(nres))
        if n < 0:
            raise RuntimeError("sync program failed")
        ops = np.zeros((max(n, 1), 8), np.int32)
        res = np.zeros((max(nres.value, 1), 2), np.int32)
        lib().fclaw2d_hip_forest_sync_program_range(self._h, mx, my, minlevel, ops.ctypes.data, n,
                                                    res.ctypes.data, nres.value, C.byref(nres))
        return ops[:n], res[:nres.value]

    def adapt(self, refine, coarsen, minlevel, maxlevel):
        """One regrid of the leaves; returns (kind, src) provenance of the new
        leaves (see fclaw2d_hip_forest_adapt)."""
        r = np.ascontiguousarray(refine, dtype=np.int32)
        c = np.ascontiguousarray(coarsen, dtype=np.int32)
        assert len(r) == self.npatch and len(c) == self.npatch
        n = lib().fclaw2d_hip_forest_adapt(self._h, r, c, minlevel, maxlevel)
        if n < 0:
            raise RuntimeError("forest adapt failed")
        kind = np.zeros(n, np.int32)
        src = np.zeros(4 * n, np.int32)
        if lib().fclaw2d_hip_forest_provenance(self._h, kind, src) != n:
            raise RuntimeError("forest provenance failed")
        return kind, src.reshape(n, 4)

    def sync_program(self, mx, my):
        """Time-synchronisation ops (hipclaw_set_sync_program)."""
        ops = np.empty((8 * self.npatch + 8, 8), np.int32)   # at most two ops per face
        n = lib().fclaw2d_hip_forest_sync_program(self._h, mx, my, ops.ctypes.data, len(ops))
        if n < 0:
            raise RuntimeError("sync program failed")
        if n > len(ops):
            ops = np.zeros((n, 8), np.int32)
            n2 = lib().fclaw2d_hip_forest_sync_program(self._h, mx, my, ops.ctypes.data, n)
            assert n2 == n
        return ops[:n]

    def geometry(self, extent, mx, my):
        """Per-patch xlower, ylower, dx, dy for physical extent (ax, bx, ay, by)."""
        ax, bx, ay, by = extent
        lv, gx, gy = self.leaves()
        wx, wy = (bx - ax) / self.nbx, (by - ay) / self.nby
        scale = np.ldexp(1.0, -lv.astype(np.int64))
        xlower = ax + gx * wx * scale
        ylower = ay + gy * wy * scale
        return xlower, ylower, wx * scale / mx, wy * scale / my


class Solver:
    """A solution resident on the GPU (include/hipclaw.h)."""

    def __init__(self, solver, mx, my, mbc=2, precision=FP64, order=(2, 2), mthlim=None,
                 mthbc=(1, 1, 1, 1), grav=1.0, gamma=1.4, tile=0, max_patches=1,
                 time_sync=False, subcycle=False, rp_params=None):
        meqn, mwaves, maux = solver_shape(solver)
        cfg = HipclawConfig()
        cfg.solver, cfg.precision = solver, precision
        cfg.mx, cfg.my, cfg.mbc = mx, my, mbc
        cfg.meqn, cfg.maux, cfg.mwaves = meqn, maux, mwaves
        cfg.order[0], cfg.order[1] = order
        lim = list(mthlim) if mthlim is not None else [4] * mwaves
        for k in range(10):
            cfg.mthlim[k] = lim[k] if k < len(lim) else 0
        for k in range(4):
            cfg.mthbc[k] = mthbc[k]
        cfg.grav, cfg.gamma, cfg.tile = grav, gamma, tile
        for k, v in enumerate(rp_params or ()):
            cfg.rp_params[k] = v
        cfg.time_sync = int(bool(time_sync))
        cfg.subcycle = int(bool(subcycle))
        self.time_sync = bool(time_sync)
        self.subcycle = bool(subcycle)
        self.cfg = cfg
        self.meqn, self.maux, self.mwaves = meqn, maux, mwaves
        self.mx, self.my, self.mbc = mx, my, mbc
        self.mthbc = tuple(mthbc)
        self.max_patches = max_patches
        self.npatch = 0
        h = C.c_void_p()
        _check(lib().hipclaw_create(C.byref(cfg), max_patches, C.byref(h)), "hipclaw_create")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hipclaw_destroy(self._h)
            self._h = None

    @property
    def tile(self):
        return lib().hipclaw_tile(self._h)

    @property
    def patch_shape(self):
        return (self.meqn, self.my + 2 * self.mbc, self.mx + 2 * self.mbc)

    def set_forest(self, forest, extent):
        """Geometry and ghost-fill program of `forest` over physical `extent`;
        for a subcycling state also the levels and the programs of every level
        range (as fc2d_hipclaw_allocate_buffers installs them)."""
        _, _, dx, dy = forest.geometry(extent, self.mx, self.my)
        self.set_geometry(dx, dy)
        ops, ps = forest.ghost_program(self.mx, self.my, self.mbc, self.mthbc)
        self.set_ghost_program(ops, ps)
        if not self.subcycle:
            if self.time_sync:
                self.set_sync_program(forest.sync_program(self.mx, self.my))
            return
        lv = forest.leaves()[0]
        lo, hi = int(lv.min()), int(lv.max())
        _check(lib().hipclaw_set_levels(self._h, np.ascontiguousarray(lv, dtype=np.int32), len(lv)),
               "hipclaw_set_levels")
        for ti in range(lo, hi):
            ops, ps = forest.ghost_program_range(self.mx, self.my, self.mbc, self.mthbc, ti + 1, hi, True)
            _check(lib().hipclaw_set_ghost_program_k(self._h, 1 + ti - lo, ops.reshape(-1), len(ops), ps),
                   "hipclaw_set_ghost_program_k")
        if self.time_sync:
            for minl in range(lo, hi + 1):
                ops, res = forest.sync_program_range(self.mx, self.my, minl)
                _check(lib().hipclaw_set_sync_program_k(self._h, minl - lo, ops.reshape(-1), len(ops),
                                                        res.reshape(-1), len(res)),
                       "hipclaw_set_sync_program_k")

    def advance_subcycle(self, dt, ncoarse=1):
        """ncoarse coarse steps of the subcycled advance, enqueued on the device
        (hipclaw_advance_subcycle); returns the largest Courant number."""
        cfl = C.c_double(0.0)
        _check(lib().hipclaw_advance_subcycle(self._h, dt, ncoarse, C.byref(cfl)),
               "hipclaw_advance_subcycle")
        return cfl.value

    def advance_subcycle_async(self, dt, ncoarse=1):
        _check(lib().hipclaw_advance_subcycle_async(self._h, dt, ncoarse), "hipclaw_advance_subcycle")

    def capture_subcycle_graph(self, dt, ncoarse=1):
        _check(lib().hipclaw_capture_subcycle_graph(self._h, dt, ncoarse),
               "hipclaw_capture_subcycle_graph")

    def set_sync_program(self, ops):
        ops = np.ascontiguousarray(ops, dtype=np.int32).reshape(-1)
        _check(lib().hipclaw_set_sync_program(self._h, ops, len(ops) // 8),
               "hipclaw_set_sync_program")

    def time_sync_now(self):
        """Apply the conservation fix-up from the registers of the last update."""
        _check(lib().hipclaw_time_sync(self._h), "hipclaw_time_sync")

    def tag(self, ivar, refine_threshold, coarsen_threshold):
        """Per patch: (exceeds refine threshold, exceeds coarsen threshold)."""
        er = np.zeros(self.npatch, np.int32)
        ec = np.zeros(self.npatch, np.int32)
        _check(lib().hipclaw_tag_patches(self._h, ivar, refine_threshold, coarsen_threshold, er, ec),
               "hipclaw_tag_patches")
        return er, ec

    def regrid(self, forest, kind, src, extent, times=None):
        """Move the solution onto the adapted `forest` (provenance kind, src),
        then install its geometry and programs (aux must be set again).
        `times`, a dict, receives the seconds of the transfer and of the
        programs."""
        import time
        t0 = time.perf_counter()
        _, _, dx, dy = forest.geometry(extent, self.mx, self.my)
        kind = np.ascontiguousarray(kind, dtype=np.int32)
        src = np.ascontiguousarray(src, dtype=np.int32).reshape(-1)
        _check(lib().hipclaw_regrid(self._h, len(kind), kind, src,
                                    np.ascontiguousarray(dx, dtype=np.float64),
                                    np.ascontiguousarray(dy, dtype=np.float64)), "hipclaw_regrid")
        self.npatch = len(kind)
        self.max_patches = max(self.max_patches, self.npatch)
        t1 = time.perf_counter()
        self.set_forest(forest, extent)
        if times is not None:
            times["transfer"] = t1 - t0
            times["programs"] = time.perf_counter() - t1

    def save_step(self):
        """Keep a device copy of every patch's state (to retake a step)."""
        _check(lib().hipclaw_save_step(self._h), "hipclaw_save_step")

    def restore_step(self):
        """Put the state kept by save_step back."""
        _check(lib().hipclaw_restore_step(self._h), "hipclaw_restore_step")

    def autotune(self, dt, nsteps=8):
        """Time every allowed tile (16 x 16 with 2 and 4 waves) in every
        ghost-fill layout over nsteps steps of dt (the state is put back
        after each) and keep the fastest; returns
        {(tile, waves, layout): ms per step}."""
        tiles = np.zeros(AUTOTUNE_MAX_TRIALS, np.int32)
        waves = np.zeros(AUTOTUNE_MAX_TRIALS, np.int32)
        layouts = np.zeros(AUTOTUNE_MAX_TRIALS, np.int32)
        ms = np.zeros(AUTOTUNE_MAX_TRIALS, np.float64)
        n = C.c_int()
        _check(lib().hipclaw_autotune(self._h, dt, nsteps, tiles.ctypes.data, waves.ctypes.data,
                                      layouts.ctypes.data, ms.ctypes.data, C.byref(n)),
               "hipclaw_autotune")
        k = n.value
        return {(int(t), int(w), int(l)): float(m)
                for t, w, l, m in zip(tiles[:k], waves[:k], layouts[:k], ms[:k])}

    @property
    def waves(self):
        return lib().hipclaw_waves(self._h)

    def set_waves(self, waves):
        """Waves per workgroup of 16 x 16 tiles: 2, 4 or 0 (the default)."""
        _check(lib().hipclaw_set_waves(self._h, waves), "hipclaw_set_waves")

    @property
    def ghost_layout(self):
        return lib().hipclaw_ghost_layout(self._h)

    def set_ghost_layout(self, layout):
        """0: one launch per ghost-fill phase; 1: per-destination groups."""
        _check(lib().hipclaw_set_ghost_layout(self._h, layout), "hipclaw_set_ghost_layout")

    def set_dt_scale(self, scale):
        """Per-slot factor on dt (None: a common dt)."""
        if scale is None:
            _check(lib().hipclaw_set_dt_scale(self._h, None), "hipclaw_set_dt_scale")
            return
        sc = np.ascontiguousarray(scale, dtype=np.float64)
        assert len(sc) == self.npatch
        _check(lib().hipclaw_set_dt_scale(self._h, sc.ctypes.data), "hipclaw_set_dt_scale")

    def ghost_pack_elems(self):
        """Doubles in one patch's ghost-patch message."""
        n = C.c_longlong()
        _check(lib().hipclaw_ghost_pack_elems(self._h, C.byref(n)), "hipclaw_ghost_pack_elems")
        return n.value

    def ghost_pack(self, slots, device_ptr, time_interp=False):
        """Pack the messages of `slots` into device memory at device_ptr."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        _check(lib().hipclaw_ghost_pack(self._h, sl, len(sl), int(time_interp), device_ptr),
               "hipclaw_ghost_pack")

    def ghost_unpack(self, slots, device_ptr, time_interp=False):
        """Unpack messages from device memory at device_ptr into `slots`."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        _check(lib().hipclaw_ghost_unpack(self._h, sl, len(sl), int(time_interp), device_ptr),
               "hipclaw_ghost_unpack")

    def get_registers(self, first=0, count=None):
        """Time-sync registers, shape (count, 4 faces, 4 kinds, meqn, mx)."""
        count = self.npatch - first if count is None else count
        out = np.zeros((count, 4, 4, self.meqn, self.mx))
        _check(lib().hipclaw_get_registers(self._h, first, count, out.reshape(-1)),
               "hipclaw_get_registers")
        return out

    def set_geometry(self, dx, dy):
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        dy = np.ascontiguousarray(dy, dtype=np.float64)
        _check(lib().hipclaw_set_geometry(self._h, len(dx), dx, dy), "hipclaw_set_geometry")
        self.npatch = len(dx)

    def set_ghost_program(self, ops, phase_start):
        ops = np.ascontiguousarray(ops, dtype=np.int32).reshape(-1)
        ps = np.ascontiguousarray(phase_start, dtype=np.int32)
        _check(lib().hipclaw_set_ghost_program(self._h, ops, len(ops) // 8, ps),
               "hipclaw_set_ghost_program")

    def set_q(self, q, first=0):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.ndim != 4 or q.shape[1:] != self.patch_shape:
            raise ValueError("q must have shape (npatch,) + %s (ghost cells included), got %s"
                             % (self.patch_shape, q.shape))
        _check(lib().hipclaw_set_q(self._h, q.reshape(-1), first, q.shape[0]), "hipclaw_set_q")

    def get_q(self, first=0, count=None):
        count = self.npatch - first if count is None else count
        out = np.zeros((count,) + self.patch_shape, np.float64)
        _check(lib().hipclaw_get_q(self._h, out.reshape(-1), first, count), "hipclaw_get_q")
        return out

    def set_aux(self, aux, first=0):
        aux = np.ascontiguousarray(aux, dtype=np.float64)
        ashape = (self.maux, self.my + 2 * self.mbc, self.mx + 2 * self.mbc)
        if aux.ndim != 4 or aux.shape[1:] != ashape:
            raise ValueError("aux must have shape (npatch,) + %s (ghost cells included), got %s"
                             % (ashape, aux.shape))
        _check(lib().hipclaw_set_aux(self._h, aux.reshape(-1), first, aux.shape[0]),
               "hipclaw_set_aux")

    def ghost_fill(self):
        _check(lib().hipclaw_ghost_fill(self._h), "hipclaw_ghost_fill")

    def step_batch(self, slots, dt, t=0.0):
        slots = np.ascontiguousarray(slots, dtype=np.int32)
        cfl = C.c_double(0.0)
        _check(lib().hipclaw_step_batch(self._h, slots, len(slots), t, dt, C.byref(cfl)),
               "hipclaw_step_batch")
        return cfl.value

    def advance(self, dt, nsteps=1):
        cfl = C.c_double(0.0)
        _check(lib().hipclaw_advance(self._h, dt, nsteps, C.byref(cfl)), "hipclaw_advance")
        return cfl.value

    def advance_async(self, dt, nsteps=1):
        _check(lib().hipclaw_advance_async(self._h, dt, nsteps), "hipclaw_advance_async")

    def sync(self):
        _check(lib().hipclaw_sync(self._h), "hipclaw_sync")

    def capture_graph(self, dt, nsteps):
        _check(lib().hipclaw_capture_graph(self._h, dt, nsteps), "hipclaw_capture_graph")

    def replay_graph(self):
        _check(lib().hipclaw_replay_graph(self._h), "hipclaw_replay_graph")

    def timing(self, nsteps):
        tot, ker = C.c_double(0.0), C.c_double(0.0)
        _check(lib().hipclaw_timing(self._h, nsteps, C.byref(tot), C.byref(ker)),
               "hipclaw_timing")
        return tot.value, ker.value

    def mass(self):
        hi = np.zeros(self.meqn)
        lo = np.zeros(self.meqn)
        _check(lib().hipclaw_mass(self._h, hi, lo), "hipclaw_mass")
        return hi, lo

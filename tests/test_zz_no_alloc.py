"""hipclaw_step_batch_noalloc: the batched update on caller-owned device
memory, with its slot list and CFL word in the library's fixed device scratch
array.  Checked against hipclaw_step_batch (the entry point it is a variant
of) on the same patches: FP64 bit for bit (tolerance 1e-12 relative, max
norm), both forms, on two calls in a row, the second with a smaller batch and
fewer patches, so whatever the first call left in the scratch must not leak
into the second."""
import ctypes as C

import numpy as np
import pytest

import forestclaw_amd as fa
from helpers import interior, random_state, rel_max_err

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers for the test itself (the entry point allocates nothing;
    its caller does)."""

    def __init__(self):
        fa.lib()  # loads the HIP runtime the library links
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, shape):
        out = np.empty(shape, np.float64)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc(sv, hip, q, slots, dx, dt, aux=None):
    npatch = q.shape[0]
    qi, qo = hip.put(q), hip.put(np.zeros_like(q))
    dxp, dyp = hip.put(dx), hip.put(dx)
    auxp = hip.put(aux) if aux is not None else None
    cfl = C.c_double(-1.0)
    slots = np.ascontiguousarray(slots, dtype=np.int32)
    rc = fa.lib().hipclaw_step_batch_noalloc(C.byref(sv.cfg), npatch, qi, qo, auxp, dxp, dyp,
                                             slots, len(slots), dt, None, C.byref(cfl))
    if rc != 0:
        raise fa.HipclawError(fa.lib().hipclaw_last_error().decode())
    return hip.get(qo, q.shape), cfl.value


@pytest.mark.parametrize("precision", [fa.FP64, fa.FP32_BULK])
@pytest.mark.parametrize("solver", [fa.RP_SHALLOW, fa.RP_ADVECTION])
def test_noalloc_matches_step_batch_twice(gpu, solver, precision):
    hip = _Hip()
    try:
        mx = 32
        tol = 1e-12 if precision == fa.FP64 else 0.0
        for call, (npatch, nslots) in enumerate([(24, 24), (9, 5)]):
            rng = np.random.default_rng(77 + 10 * call + solver + 3 * precision)
            q = random_state(solver, npatch, mx, mx, rng)
            aux = None
            if solver == fa.RP_ADVECTION:
                aux = rng.uniform(-1.0, 1.0, (npatch, 2, mx + 4, mx + 4))
            dx = np.full(npatch, 1.0 / mx)
            slots = rng.permutation(npatch)[:nslots].astype(np.int32)
            dt = 0.3 / mx
            sv = fa.Solver(solver, mx, mx, precision=precision, max_patches=npatch)
            sv.set_geometry(dx, dx)
            sv.set_q(q)
            if aux is not None:
                sv.set_aux(aux)
            cfl_ref = sv.step_batch(slots, dt)
            q_ref = sv.get_q()
            q_new, cfl = _noalloc(sv, hip, q, slots, dx, dt, aux)
            print("call %d: %d patches, %d slots, tile %d, cfl %.6f (step_batch %.6f)" % (
                call, npatch, nslots, sv.tile, cfl, cfl_ref), flush=True)
            for p in slots:
                err = rel_max_err(interior(q_new[p]), interior(q_ref[p]))
                assert err <= tol, (call, int(p), err)
            assert cfl == cfl_ref
            untouched = np.setdiff1d(np.arange(npatch), slots)
            assert np.all(q_new[untouched] == 0.0)   # only the listed slots are written
    finally:
        hip.free()


def test_noalloc_refuses_what_it_cannot_hold(gpu):
    sv = fa.Solver(fa.RP_SHALLOW, 32, 32, max_patches=1)
    big = np.zeros(70000, np.int32)
    rc = fa.lib().hipclaw_step_batch_noalloc(C.byref(sv.cfg), 1, C.c_void_p(8), C.c_void_p(8), None,
                                             C.c_void_p(8), C.c_void_p(8), big, len(big), 0.01,
                                             None, None)
    assert rc == 1 and b"HIPCLAW_NOALLOC_MAX_SLOTS" in fa.lib().hipclaw_last_error()
    one = np.zeros(1, np.int32)
    q = C.c_void_p(8)
    rc = fa.lib().hipclaw_step_batch_noalloc(C.byref(sv.cfg), 1, q, q, None, q, q, one, 1, 0.01,
                                             None, None)
    assert rc == 1 and b"must differ" in fa.lib().hipclaw_last_error()

"""GPU ghost-patch messages (the multi-rank wire format of
clawpatch_local_ghost_pack / clawpatch_remote_ghost_unpack): the device pack
of the q band and the time-sync registers equals the oracle's pack of the
same host data value for value; unpacking into another state and packing
again gives the same messages; bad buffers are refused on the host."""
import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
from helpers import setup_problem
from oracle_lib import oracle

pytestmark = pytest.mark.gpu

MBC, MINT = 2, 4


class DeviceBuffer:
    """n doubles of device memory from the HIP runtime the library uses."""

    def __init__(self, n):
        import ctypes as C
        self._C = C
        self.hip = C.CDLL("libamdhip64.so")
        self.n = n
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * n)) == 0
        assert self.hip.hipMemset(self.ptr, 0, C.c_size_t(8 * n)) == 0

    def data_ptr(self):
        return self.ptr.value

    def numpy(self):
        out = np.zeros(self.n)
        assert self.hip.hipMemcpy(self._C.c_void_p(out.ctypes.data), self.ptr,
                                  self._C.c_size_t(8 * self.n), 2) == 0  # device to host
        return out

    def __del__(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = self._C.c_void_p()




def _oracle_messages(sv, forest, prob, slots, q, reg):
    o = oracle()
    _, _, dx, dy = forest.geometry(prob.extent, prob.mx, prob.mx)
    elems = sv.ghost_pack_elems()
    out = np.zeros((len(slots), elems))
    nq = ((prob.mx + 2 * MBC) ** 2 - (prob.mx - 2 * MINT) ** 2) * prob.meqn
    for k, s in enumerate(slots):
        msg = np.zeros(elems)
        assert o.oracle_local_ghost_pack(prob.mx, prob.mx, MBC, prob.meqn, MINT,
                                         np.ascontiguousarray(q[s]), None, msg, 0) == nq
        if reg is not None:
            part = np.zeros(elems - nq)
            assert o.oracle_register_pack(prob.mx, prob.mx, prob.meqn, float(dx[s]), float(dy[s]),
                                          np.ascontiguousarray(reg[s]).reshape(-1), part, 0) == elems - nq
            msg[nq:] = part
        out[k] = msg
    return out


@pytest.mark.parametrize("time_sync", [False, True])
def test_pack_matches_oracle_and_round_trips(gpu, time_sync):
    prob = P.dambreak_problem(2, 4, mx=16)
    forest, q0, aux, sv = setup_problem(prob, time_sync=time_sync)
    sv.advance(P.stable_dt(prob, forest, q0) * 0.5, 2)
    q = sv.get_q()
    reg = sv.get_registers() if time_sync else None
    elems = sv.ghost_pack_elems()
    nq = ((16 + 2 * MBC) ** 2 - (16 - 2 * MINT) ** 2) * prob.meqn
    assert elems == nq + (2 * (4 * prob.meqn + 2) * 32 if time_sync else 0)

    rng = np.random.default_rng(3)
    slots = rng.permutation(forest.npatch)[: min(40, forest.npatch)].astype(np.int32)
    print("slots %d, elems %d, time_sync %d" % (len(slots), elems, time_sync), flush=True)
    buf = DeviceBuffer(len(slots) * elems)
    sv.ghost_pack(slots, buf.data_ptr())
    got = buf.numpy().reshape(len(slots), elems)
    np.testing.assert_array_equal(got, _oracle_messages(sv, forest, prob, slots, q, reg))

    # unpack into other slots of a second state, then pack those again
    _, _, _, sv2 = setup_problem(prob, time_sync=time_sync)
    dest = rng.permutation(forest.npatch)[: len(slots)].astype(np.int32)
    sv2.ghost_unpack(dest, buf.data_ptr())
    buf2 = DeviceBuffer(len(slots) * elems)
    sv2.ghost_pack(dest, buf2.data_ptr())
    m1 = buf.numpy().reshape(len(slots), elems)
    m2 = buf2.numpy().reshape(len(slots), elems)
    # everything but the edge lengths and areas, which are the destination's
    keep = np.ones(elems, bool)
    if time_sync:
        k = nq
        for face in range(4):
            k += 4 * prob.meqn * prob.mx
            keep[k:k + 2 * prob.mx] = False
            k += 2 * prob.mx
    np.testing.assert_array_equal(m2[:, keep], m1[:, keep])
    np.testing.assert_array_equal(m2, _oracle_messages(sv2, forest, prob, dest, sv2.get_q(),
                                                       sv2.get_registers() if time_sync else None))
    # the band of each destination holds the source's band; its centre is untouched
    q2 = sv2.get_q()
    frame = np.zeros((20, 20), bool)
    frame[:, :MBC + MINT] = frame[:, -(MBC + MINT):] = True
    frame[:MBC + MINT, :] = frame[-(MBC + MINT):, :] = True
    q2_init = setup_problem(prob)[1]
    for s, d in zip(slots, dest):
        np.testing.assert_array_equal(q2[d][:, frame], q[s][:, frame])
        np.testing.assert_array_equal(q2[d][:, ~frame], q2_init[d][:, ~frame])
    if time_sync:
        np.testing.assert_array_equal(sv2.get_registers()[dest], reg[slots])


def test_bad_message_buffers_are_refused(gpu):
    prob = P.dambreak_problem(2, 3, mx=16)
    forest, q0, aux, sv = setup_problem(prob)
    elems = sv.ghost_pack_elems()
    host = np.zeros(4 * elems)
    with pytest.raises(fa.HipclawError):
        sv.ghost_pack([0, 1, 2, 3], host.ctypes.data)       # host memory
    # an allocation of its own (a caching allocator may place a small buffer
    # inside a larger block, which the size check can only see as a whole)
    small = DeviceBuffer(2 * elems)
    with pytest.raises(fa.HipclawError):
        sv.ghost_pack([0, 1, 2, 3], small.data_ptr())          # room for two messages
    with pytest.raises(fa.HipclawError):
        sv.ghost_pack([0, forest.npatch], small.data_ptr())    # bad slot
    with pytest.raises(fa.HipclawError):
        sv.ghost_pack([0], small.data_ptr(), time_interp=True)  # no subcycle copies
    sv.ghost_pack([0, 1], small.data_ptr())                    # fits
    with pytest.raises(fa.HipclawError):
        sv.get_registers()                                    # no time sync

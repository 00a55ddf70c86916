"""CPU baseline timing for bench.py (test infrastructure; never part of the
product path).

Times the reference's own Clawpack-4.6 Fortran patch step
(oracle/_ref/libref46.so, kind "reference") -- or, if that library was not
built, the oracle's C restatement (oracle/liboracle46.so, kind "port") -- on a
bounded sample of the benchmark's patches, with P worker processes (the
Fortran's COMMON blocks rule out threads).  Each worker steps its share of
the sample patches repeatedly; the rate is sample cell updates / wall time.
"""
import ctypes as C
import multiprocessing as mp
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "_ref", "libref46.so")
PORT_SO = os.path.join(HERE, "liboracle46.so")


def _load():
    if os.path.exists(REF_SO):
        lib = C.CDLL(REF_SO)
        return lib, "reference"
    return C.CDLL(PORT_SO), "port"


def _worker(args):
    solver, q, aux, dx, dt, mwaves, maux, reps, grav, gamma, barrier = args
    lib, kind = _load()
    meqn = q.shape[1]
    mx = q.shape[3] - 4
    method = np.array([0, 2, 2, 0, 0, 0, maux], dtype=np.int32)
    mthlim = np.array([4] * mwaves + [0] * (10 - mwaves), dtype=np.int32)
    cfl = C.c_double(0.0)
    flux = [np.zeros(q.shape[1:]) for _ in range(4)]
    work = q.copy()
    barrier.wait()
    t0 = time.perf_counter()
    for _ in range(reps):
        for p in range(work.shape[0]):
            work[p] = q[p]
            if kind == "reference":
                lib.ref46_step2(C.c_int(solver), C.c_int(mx), C.c_int(mx), C.c_int(2),
                                C.c_int(meqn), C.c_int(maux), C.c_int(mwaves),
                                method.ctypes.data_as(C.POINTER(C.c_int)),
                                mthlim.ctypes.data_as(C.POINTER(C.c_int)),
                                work[p].ctypes.data_as(C.POINTER(C.c_double)),
                                None if aux is None else aux[p].ctypes.data_as(C.c_void_p),
                                C.c_double(dx[p]), C.c_double(dx[p]), C.c_double(dt),
                                C.c_double(grav), C.c_double(gamma), C.byref(cfl),
                                *[f.ctypes.data_as(C.POINTER(C.c_double)) for f in flux])
    return time.perf_counter() - t0


def _port_step(lib, solver, q, aux, dx, dt, mwaves, maux, grav, gamma):
    class Claw(C.Structure):
        _fields_ = [("solver", C.c_int), ("meqn", C.c_int), ("maux", C.c_int),
                    ("mwaves", C.c_int), ("method", C.c_int * 7), ("mthlim", C.c_int * 10),
                    ("grav", C.c_double), ("gamma", C.c_double), ("p", C.c_double * 8)]
    c = Claw()
    c.solver, c.meqn, c.maux, c.mwaves = solver, q.shape[0], maux, mwaves
    for k, v in enumerate([0, 2, 2, 0, 0, 0, maux]):
        c.method[k] = v
    for k in range(mwaves):
        c.mthlim[k] = 4
    c.grav, c.gamma = grav, gamma
    cfl = C.c_double(0.0)
    mx = q.shape[2] - 4
    lib.oracle_step2_patch(C.byref(c), mx, mx, 2, q.ctypes.data_as(C.c_void_p),
                           None if aux is None else aux.ctypes.data_as(C.c_void_p),
                           C.c_double(dx), C.c_double(dx), C.c_double(dt), C.byref(cfl),
                           None, None, None, None)


def _worker_port(args):
    solver, q, aux, dx, dt, mwaves, maux, reps, grav, gamma, barrier = args
    lib = C.CDLL(PORT_SO)
    work = q.copy()
    barrier.wait()
    t0 = time.perf_counter()
    for _ in range(reps):
        for p in range(work.shape[0]):
            work[p] = q[p]
            _port_step(lib, solver, work[p], None if aux is None else aux[p], dx[p], dt,
                       mwaves, maux, grav, gamma)
    return time.perf_counter() - t0


# Host cores the machine allots to one GPU: a host shared by several GPUs
# gives each GPU's processes this share, whatever os.cpu_count() reports.
HOST_CORES_PER_GPU = 16


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measure(solver, q, aux, dx, dt, mwaves, maux, grav=1.0, gamma=1.4, nproc=None,
            target_seconds=10.0):
    """Returns dict(value=cell updates/s, cores=P, kind, sample=str): the patch
    step only (clawpack46_step2 with its Riemann solvers), no ghost fill and no
    time synchronisation."""
    _, kind = _load()
    nproc = nproc or max(1, min(HOST_CORES_PER_GPU, os.cpu_count() or 1))
    npatch = q.shape[0]
    mx = q.shape[3] - 4
    # calibrate: one pass over a few patches in this process
    probe = min(npatch, 8)
    wfn = _worker if kind == "reference" else _worker_port

    class _NoBarrier:
        def wait(self):
            pass
    t = wfn((solver, q[:probe].copy(), None if aux is None else aux[:probe].copy(), dx[:probe],
             dt, mwaves, maux, 1, grav, gamma, _NoBarrier()))
    per_patch = max(t / probe, 1e-6)
    # each process steps its share `reps` times, about target_seconds of CPU work overall
    share = np.array_split(np.arange(npatch), nproc)
    reps = max(1, int(target_seconds / (per_patch * max(len(s) for s in share))))
    ctx = mp.get_context("fork")
    barrier = ctx.Manager().Barrier(nproc)
    jobs = [(solver, q[s].copy(), None if aux is None else aux[s].copy(), dx[s], dt, mwaves,
             maux, reps, grav, gamma, barrier) for s in share]
    with ctx.Pool(nproc) as pool:
        times = pool.map(wfn, jobs)
    wall = max(times)
    cells = npatch * mx * mx * reps
    # the same sample on one process, about 2 s of it
    reps1 = max(1, int(2.0 / (per_patch * npatch)))
    t1 = wfn((solver, q.copy(), None if aux is None else aux.copy(), dx, dt, mwaves, maux, reps1,
              grav, gamma, _NoBarrier()))
    single = npatch * mx * mx * reps1 / t1
    return dict(value=cells / wall, unit="cell-updates/s", cores=nproc, kind=kind,
                single_process_value=single, cpu_model=cpu_model(),
                host_cpus_visible=os.cpu_count(),
                covers="the patch update only (clawpack46_step2_wrap with the Roe solvers); "
                       "no ghost fill, no time synchronisation",
                sample="%d patches of %dx%d x %d repetitions of one Clawpack-4.6 step "
                       "(%s), %d processes, %.1f s wall; one process: %d repetitions, %.1f s"
                       % (npatch, mx, mx, reps,
                          "reference Fortran via oracle/_ref" if kind == "reference" else
                          "oracle C port", nproc, wall, reps1, t1))

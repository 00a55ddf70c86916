"""Registration of an application's own Riemann solver (include/hipclaw_solver.h),
on the CPU: the descriptors the example libraries export, the checks of
hipclaw_register_solver, and the configuration checks that refuse what a
registered solver does not provide (all before any device call)."""
import ctypes as C

import pytest

import forestclaw_amd as fa
from forestclaw_amd.build import user_solver_lib


def _desc(name):
    L = C.CDLL(user_solver_lib(name))
    fn = getattr(L, name + "_solver")
    fn.restype = C.POINTER(fa.HipclawSolverDesc)
    return L, fn()


def test_example_descriptors():
    _, a = _desc("acoustics")
    d = a.contents
    assert (d.meqn, d.mwaves, d.maux, d.conservative, d.time_sync) == (3, 2, 0, 1, 0)
    assert d.tiles == 8 | 16 and d.launch_fp64 and d.launch_fp32
    assert d.name == b"acoustics"
    _, t = _desc("tracers")
    d = t.contents
    assert (d.meqn, d.mwaves, d.maux, d.conservative, d.time_sync) == (10, 10, 2, 0, 0)
    assert d.tiles & 8 and d.launch_fp64 and d.launch_fp32


def test_registration_is_idempotent_and_numbers_from_user_base():
    a = fa.example_solver("acoustics")
    assert a >= 16                      # HIPCLAW_RP_USER
    assert fa.example_solver("acoustics") == a
    t = fa.example_solver("tracers")
    assert t != a and t >= 16
    assert fa.solver_shape(a) == (3, 2, 0) and fa.solver_shape(t) == (10, 10, 2)


def test_registration_refuses_a_foreign_build():
    L, a = _desc("acoustics")
    bad = fa.HipclawSolverDesc.from_buffer_copy(a.contents)
    sid = C.c_int(-1)
    bad.abi = 99
    assert fa.lib().hipclaw_register_solver(C.cast(C.pointer(bad), C.c_void_p), C.byref(sid)) != 0
    bad = fa.HipclawSolverDesc.from_buffer_copy(a.contents)
    bad.step_args_size += 8
    assert fa.lib().hipclaw_register_solver(C.cast(C.pointer(bad), C.c_void_p), C.byref(sid)) != 0
    bad = fa.HipclawSolverDesc.from_buffer_copy(a.contents)
    bad.mwaves = 11
    assert fa.lib().hipclaw_register_solver(C.cast(C.pointer(bad), C.c_void_p), C.byref(sid)) != 0


def test_configuration_checks_against_the_registered_solver():
    a = fa.example_solver("acoustics")
    # no conservative flux for the time-sync registers: time sync is refused
    with pytest.raises(fa.HipclawError, match="time sync"):
        fa.Solver(a, 16, 16, time_sync=True)
    # meqn / mwaves must be the solver's
    cfg = fa.HipclawConfig()
    cfg.solver, cfg.mx, cfg.my, cfg.mbc = a, 16, 16, 2
    cfg.meqn, cfg.mwaves, cfg.maux = 3, 3, 0
    cfg.order[0], cfg.order[1] = 2, 2
    h = C.c_void_p()
    assert fa.lib().hipclaw_create(C.byref(cfg), 1, C.byref(h)) != 0
    assert b"mwaves" in fa.lib().hipclaw_last_error()
    # an unregistered number
    cfg.solver = 31
    assert fa.lib().hipclaw_create(C.byref(cfg), 1, C.byref(h)) != 0
    assert b"unknown Riemann solver" in fa.lib().hipclaw_last_error()


def test_ten_wave_families_take_their_own_limiters():
    t = fa.example_solver("tracers")
    with pytest.raises(fa.HipclawError, match="mthlim"):
        fa.Solver(t, 16, 16, mthlim=[4] * 9 + [7])

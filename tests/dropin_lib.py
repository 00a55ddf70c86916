"""ctypes mirrors of the drop-in C surface (include/fclaw2d_hip_core.h,
include/fc2d_hipclaw.h) so tests can drive the library the way ForestClaw
drives fc2d_cudaclaw."""
import ctypes as C

import numpy as np

import forestclaw_amd as fa


class Patch(C.Structure):
    _fields_ = [("level", C.c_int), ("xlower", C.c_double), ("xupper", C.c_double),
                ("ylower", C.c_double), ("yupper", C.c_double), ("blockno", C.c_int),
                ("patchno", C.c_int), ("slot", C.c_int), ("user", C.c_void_p)]


class Block(C.Structure):
    _fields_ = [("num_patches", C.c_int), ("patchno_offset", C.c_int),
                ("patches", C.POINTER(Patch))]


class Domain(C.Structure):
    _fields_ = [("num_blocks", C.c_int), ("blocks", C.POINTER(Block)),
                ("local_num_patches", C.c_int), ("global_minlevel", C.c_int),
                ("global_maxlevel", C.c_int), ("nbx", C.c_int), ("nby", C.c_int),
                ("periodic_x", C.c_int), ("periodic_y", C.c_int), ("ax", C.c_double),
                ("bx", C.c_double), ("ay", C.c_double), ("by", C.c_double),
                ("forest", C.c_void_p)]


PATCH_INIT = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(Patch), C.c_int, C.c_int)
SINGLE_STEP = C.CFUNCTYPE(C.c_double, C.c_void_p, C.POINTER(Patch), C.c_int, C.c_int, C.c_double,
                          C.c_double, C.c_void_p)


class PatchVtable(C.Structure):
    _fields_ = [("initialize", C.c_void_p), ("setup", C.c_void_p), ("physical_bc", C.c_void_p),
                ("single_step_update", C.c_void_p), ("is_set", C.c_int)]


class ClawpatchOptions(C.Structure):
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mbc", C.c_int), ("meqn", C.c_int),
                ("maux", C.c_int)]


class Global(C.Structure):
    _fields_ = [("domain", C.POINTER(Domain)), ("count_single_step", C.c_int),
                ("count_ghost_update", C.c_int), ("curr_time", C.c_double),
                ("curr_dt", C.c_double), ("patch_vt", PatchVtable),
                ("clawpatch_opt", ClawpatchOptions), ("hipclaw_vt", C.c_void_p),
                ("hipclaw_opt", C.c_void_p), ("state", C.c_void_p),
                ("qhost", C.POINTER(C.c_double)), ("auxhost", C.POINTER(C.c_double)),
                ("time_sync", C.c_int), ("subcycle", C.c_int), ("advance_one_step", C.c_int),
                ("minlevel", C.c_int), ("maxlevel", C.c_int), ("refine_threshold", C.c_double),
                ("coarsen_threshold", C.c_double), ("threshold_variable", C.c_int),
                ("refinement_criteria", C.c_int),
                ("count_amr_regrid", C.c_int), ("count_amr_new_domain", C.c_int),
                ("prefix", C.c_char_p), ("initial_dt", C.c_double), ("max_cfl", C.c_double),
                ("desired_cfl", C.c_double), ("tfinal", C.c_double), ("reduce_cfl", C.c_int),
                ("use_fixed_dt", C.c_int), ("outstyle", C.c_int), ("nout", C.c_int),
                ("nstep", C.c_int), ("outstyle_uses_maxlevel", C.c_int),
                ("regrid_interval", C.c_int), ("output", C.c_int), ("count_retake", C.c_int),
                ("conservation_check", C.c_int), ("cons_mass", C.c_double * 4),
                ("cons_mass0", C.c_double * 4)]


class Options(C.Structure):
    _fields_ = [("mwaves", C.c_int), ("order", C.c_int * 2), ("mthlim", C.c_int * 10),
                ("mthbc", C.c_int * 4), ("method", C.c_int * 7), ("mcapa", C.c_int),
                ("src_term", C.c_int), ("use_fwaves", C.c_int), ("ascii_out", C.c_int),
                ("vtk_out", C.c_int), ("buffer_len", C.c_int), ("precision", C.c_int),
                ("tile", C.c_int), ("grav", C.c_double), ("gamma", C.c_double),
                ("rp_params", C.c_double * 8), ("is_registered", C.c_int)]


FORT_QINIT = C.CFUNCTYPE(None, *([C.POINTER(C.c_int)] * 6 + [C.POINTER(C.c_double)] * 4 +
                                 [C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double)]))


class Vtable(C.Structure):
    _fields_ = [("b4step2", C.c_void_p), ("src2", C.c_void_p), ("fort_setprob", C.c_void_p),
                ("fort_bc2", C.c_void_p), ("fort_qinit", C.c_void_p), ("fort_setaux", C.c_void_p),
                ("fort_src2", C.c_void_p), ("riemann_solver", C.c_int), ("rp_solver", C.c_void_p),
                ("is_set", C.c_int)]


def bind():
    L = fa.lib()
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    sig = {
        "fclaw2d_global_new": ([], C.POINTER(Global)),
        "fclaw2d_global_destroy": ([C.POINTER(Global)], None),
        "fclaw2d_domain_new_from_forest": ([vp, i, i, i, i, d, d, d, d], C.POINTER(Domain)),
        "fclaw2d_global_store_domain": ([C.POINTER(Global), C.POINTER(Domain)], None),
        "fclaw2d_update_single_step": ([C.POINTER(Global), i, d, d], d),
        "fclaw2d_ghost_update": ([C.POINTER(Global), i, i, d, i, i], None),
        "fclaw2d_advance_all_levels": ([C.POINTER(Global), d, d], d),
        "fclaw2d_time_sync": ([C.POINTER(Global), i, i], None),
        "fclaw2d_timeinterp": ([C.POINTER(Global), i, d], None),
        "fclaw2d_initialize": ([C.POINTER(Global)], None),
        "fclaw2d_regrid": ([C.POINTER(Global)], None),
        "fclaw2d_run": ([C.POINTER(Global)], None),
        "fclaw2d_diagnostics_gather": ([C.POINTER(Global), i], None),
        "fclaw2d_save_time_step": ([C.POINTER(Global)], None),
        "fclaw2d_restore_time_step": ([C.POINTER(Global)], None),
        "fc2d_hipclaw_options_default": ([C.POINTER(Options), i], None),
        "fc2d_hipclaw_options_store": ([C.POINTER(Global), C.POINTER(Options)], None),
        "fc2d_hipclaw_check": ([C.POINTER(Options)], i),
        "hipclaw_check_parameters": ([i], i),
        "fc2d_hipclaw_solver_initialize": ([C.POINTER(Global)], None),
        "fc2d_hipclaw_vt": ([C.POINTER(Global)], C.POINTER(Vtable)),
        "fc2d_hipclaw_initialize_GPUs": ([C.POINTER(Global)], None),
        "fc2d_hipclaw_allocate_buffers": ([C.POINTER(Global)], None),
        "fc2d_hipclaw_upload": ([C.POINTER(Global)], None),
        "fc2d_hipclaw_download": ([C.POINTER(Global)], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def patches(glob):
    """(block index, patch index within block, Patch pointer) in slot order."""
    dom = glob.contents.domain.contents
    out = []
    for b in range(dom.num_blocks):
        blk = dom.blocks[b]
        for p in range(blk.num_patches):
            out.append((b, p, C.pointer(blk.patches[p])))
    return out


def host_q(glob, npatch, shape):
    n = npatch * int(np.prod(shape))
    return np.ctypeslib.as_array(glob.contents.qhost, shape=(n,)).reshape((npatch,) + shape)

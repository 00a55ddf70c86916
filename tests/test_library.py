"""The C-ABI library loads and exports every function declared in include/*.h
(no GPU calls)."""
import ctypes as C
import glob
import os
import re

import forestclaw_amd as fa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"typedef[^;]*;", "", text)
        text = re.sub(r"(struct|enum)\s+\w*\s*\{[^}]*\}\s*\w*\s*;", "", text, flags=re.S)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", text):
            name = m.group(1)
            start = text.rfind("\n", 0, m.start())
            line = text[start + 1:m.start()]
            if line.strip().startswith("#") or name in ("sizeof", "defined"):
                continue
            if re.search(r"(^|[\s\*])(int|void|double|char|const|fclaw2d_\w+|fc2d_\w+|hipclaw_\w+)[\s\*]*$",
                         line.rstrip() + " ") or line.strip() == "":
                names.add(name)
    return names


def test_library_exports_all_declared_symbols():
    lib = fa.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_missing_library_fails_loudly(tmp_path):
    import importlib
    # the reload below restores the module (its path, a fresh library handle
    # and the struct classes the handle's signatures name)
    fa.LIB_PATH = str(tmp_path / "nope.so")
    fa._lib = None
    try:
        fa.lib()
    except RuntimeError as e:
        assert "missing" in str(e)
    else:
        raise AssertionError("expected a RuntimeError")
    finally:
        importlib.reload(fa)


def test_options_checks_follow_cudaclaw():
    """fc2d_cudaclaw_check rejects mwaves > FC2D_CUDACLAW_MWAVES
    (fc2d_cudaclaw_options.c:110-115); defaults follow fc2d_cudaclaw_options.c:36-71."""
    import ctypes as C
    import dropin_lib as D
    L = D.bind()
    assert L.hipclaw_check_parameters(10) == 0
    assert L.hipclaw_check_parameters(11) == 1
    opt = D.Options()
    L.fc2d_hipclaw_options_default(C.byref(opt), 3)
    assert (opt.order[0], opt.order[1], opt.buffer_len, opt.mcapa, opt.src_term) == (2, 2, 1, 0, 0)
    assert list(opt.mthbc) == [1, 1, 1, 1]
    assert L.fc2d_hipclaw_check(C.byref(opt)) == 0
    opt.buffer_len = 0
    assert L.fc2d_hipclaw_check(C.byref(opt)) == 1


def test_struct_layouts_match_the_build():
    """The ctypes copies of the C structs have the sizes the library was built with."""
    import dropin_lib as D
    L = fa.lib()
    L.hipclaw_config_size.restype = C.c_size_t
    assert C.sizeof(fa.HipclawConfig) == L.hipclaw_config_size()
    sizes = (C.c_size_t * 6)()
    L.fclaw2d_hip_struct_sizes(sizes)
    assert list(sizes) == [C.sizeof(D.Global), C.sizeof(D.Domain), C.sizeof(D.Block),
                           C.sizeof(D.Patch), C.sizeof(D.Options), C.sizeof(D.Vtable)]


def test_options_check_rejects_unknown_boundary_codes():
    """mthbc 1 (extrapolation), 2 (periodic, left to the topology) and 3 (wall)
    pass the options check; 0 (user BC: the reference stops,
    clawpack46_bc2_default.f:22-25) and codes the BC routine would skip are
    refused."""
    import dropin_lib as D
    L = D.bind()
    opt = D.Options()
    L.fc2d_hipclaw_options_default(C.byref(opt), 3)
    for code, ok in ((1, True), (2, True), (3, True), (0, False), (4, False), (-1, False)):
        for face in range(4):
            opt.mthbc[face] = code
            assert (L.fc2d_hipclaw_check(C.byref(opt)) == 0) == ok, (face, code)
            opt.mthbc[face] = 1


def test_global_defaults_follow_the_reference():
    """fclaw2d_global_new: time-stepping defaults of src/fclaw_options.c and the
    clawpatch tagging defaults (threshold-variable 1, refinement-criteria
    minmax: fclaw2d_clawpatch_options.c:107-114)."""
    import dropin_lib as D
    L = D.bind()
    g = L.fclaw2d_global_new()
    try:
        c = g.contents
        assert (c.initial_dt, c.max_cfl, c.desired_cfl, c.regrid_interval) == (0.1, 1.0, 0.9, 1)
        assert (c.threshold_variable, c.refinement_criteria) == (1, 2)
    finally:
        L.fclaw2d_global_destroy(g)


def test_set_q_and_aux_check_the_patch_shape():
    """Arrays without ghost cells (or of another solver) are refused on the host
    before any copy: the C side copies count * patch-size doubles."""
    import numpy as np
    import pytest
    sv = fa.Solver.__new__(fa.Solver)   # no device state: the check runs first
    sv.meqn, sv.maux, sv.mx, sv.my, sv.mbc, sv._h = 3, 0, 8, 8, 2, None
    with pytest.raises(ValueError):
        sv.set_q(np.zeros((2, 3, 8, 8)))          # interior only
    with pytest.raises(ValueError):
        sv.set_q(np.zeros((2, 4, 12, 12)))        # meqn of another solver
    with pytest.raises(ValueError):
        sv.set_q(np.zeros((3, 12, 12)))           # one patch without the patch axis
    sv.maux = 2
    with pytest.raises(ValueError):
        sv.set_aux(np.zeros((2, 2, 8, 8)))

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


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    import importlib
    monkeypatch.setattr(fa, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(fa, "_lib", None)
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

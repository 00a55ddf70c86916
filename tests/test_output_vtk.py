"""VTK frames (<prefix>_frame_XXXX.vtu) as the reference's fclaw2d_vtk_write_file
writes them.

The reference writer needs p4est/libsc and is not built here.  The oracle's
numpy restatement (oracle/fclaw46_vtk.py) is pinned by decoding its file
against closed forms (points on the patch grids, quad connectivity, offsets,
cell types, block and patch numbers, float values of q); the product's writer
must then match it byte for byte: on the CPU with the product's coordinate
callback and a value callback from the test, and on the GPU through
fclaw2d_clawpatch_output_vtk with the values converted on the device.
"""
import ctypes as C
import os
import re
import sys

import numpy as np
import pytest

import forestclaw_amd as fa
from forestclaw_amd import problems as P
import dropin_lib as D

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import fclaw46_vtk  # noqa: E402

VTK_CB = C.CFUNCTYPE(None, C.POINTER(D.Global), C.POINTER(D.Patch), C.c_int, C.c_int, C.c_void_p)


def _decode(data):
    """Header attributes and the appended arrays of a .vtu file."""
    head, rest = data.split(b'<AppendedData encoding="raw">\n  _', 1)
    text = head.decode()
    assert rest.endswith(b'\n </AppendedData>\n</VTKFile>\n')
    blob = rest[:-len(b'\n </AppendedData>\n</VTKFile>\n')]
    npts, ncells = (int(v) for v in re.search(r'NumberOfPoints="(\d+)" NumberOfCells="(\d+)"', text).groups())
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)" Name="(\w+)"[^>]*offset="(\d+)"', text):
        t, name, off = m.group(1), m.group(2), int(m.group(3))
        nbytes = int(np.frombuffer(blob[off:off + 8], np.uint64)[0])
        dt = {"Float64": "<f8", "Float32": "<f4", "Int32": "<i4", "Int64": "<i8", "UInt8": "u1"}[t]
        arrays[name] = np.frombuffer(blob[off + 8:off + 8 + nbytes], dt)
    return npts, ncells, arrays


def _patches(prob, forest, q):
    xl, yl, dx, dy = forest.geometry(prob.extent, prob.mx, prob.mx)
    lv, gx, gy = forest.leaves()
    blk = (gx >> lv) + prob.nbx * (gy >> lv)
    return [(int(blk[p]), xl[p], yl[p], dx[p], dy[p], q[p]) for p in range(forest.npatch)]


@pytest.mark.parametrize("meqn", [1, 3])
def test_oracle_file_decodes_to_closed_forms(meqn):
    prob = P.shockbubble_problem(1, 2, mx=8)
    forest = P.build_forest(prob)
    rng = np.random.default_rng(meqn)
    q = rng.standard_normal((forest.npatch, meqn, 12, 12))
    pats = _patches(prob, forest, q)
    npts, ncells, a = _decode(fclaw46_vtk.vtk_bytes(pats, 8, 8, meqn))
    n = forest.npatch
    assert npts == 81 * n and ncells == 64 * n
    pos = a["Position"].reshape(n, 9, 9, 3)
    for p, (b, xl, yl, dx, dy, _) in enumerate(pats):
        np.testing.assert_array_equal(pos[p, :, :, 0], np.broadcast_to(xl + np.arange(9)[None, :] * dx, (9, 9)))
        np.testing.assert_array_equal(pos[p, :, :, 1], np.broadcast_to(yl + np.arange(9)[:, None] * dy, (9, 9)))
        assert not pos[p, :, :, 2].any()
    conn = a["connectivity"].reshape(n, 8, 8, 4)
    # each quad's corners are the patch-grid points (i,j), (i+1,j), (i+1,j+1), (i,j+1)
    p_, j_, i_ = np.meshgrid(np.arange(n), np.arange(8), np.arange(8), indexing="ij")
    base = 81 * p_ + 9 * j_ + i_
    np.testing.assert_array_equal(conn, np.stack([base, base + 1, base + 10, base + 9], -1))
    np.testing.assert_array_equal(a["offsets"], 4 * (np.arange(ncells) + 1))
    assert np.all(a["types"] == 9) and not a["mpirank"].any()
    np.testing.assert_array_equal(a["blockno"], np.repeat([b for b, *_ in pats], 64))
    np.testing.assert_array_equal(a["patchno"], np.repeat(np.arange(n), 64))
    vals = a["meqn"].reshape(n, 8, 8, meqn)
    np.testing.assert_array_equal(vals, np.moveaxis(q[:, :, 2:10, 2:10], 1, -1).astype(np.float32))


def _cpu_glob(L, prob, forest, meqn):
    lv, gx, gy = forest.leaves()
    fh = L.fclaw2d_hip_forest_from_leaves(prob.nbx, prob.nby, int(prob.periodic[0]),
                                          int(prob.periodic[1]), forest.npatch, lv, gx, gy, 0)
    ax, bx, ay, by = prob.extent
    dom = L.fclaw2d_domain_new_from_forest(fh, prob.nbx, prob.nby, int(prob.periodic[0]),
                                           int(prob.periodic[1]), ax, bx, ay, by)
    glob = L.fclaw2d_global_new()
    L.fclaw2d_global_store_domain(glob, dom)
    cp = glob.contents.clawpatch_opt
    cp.mx, cp.my, cp.mbc, cp.meqn, cp.maux = prob.mx, prob.mx, 2, meqn, 0
    return glob


def _bind_vtk(L):
    L.fclaw2d_vtk_write_file.argtypes = [C.POINTER(D.Global), C.c_char_p] + [C.c_int] * 3 + \
        [C.c_double, C.c_int, C.c_void_p, C.c_void_p]
    L.fclaw2d_vtk_write_file.restype = C.c_int
    L.fclaw2d_clawpatch_output_vtk.argtypes = [C.POINTER(D.Global), C.c_int]
    return L


@pytest.mark.parametrize("meqn", [1, 4])
def test_writer_matches_oracle_bytes(tmp_path, meqn):
    L = _bind_vtk(D.bind())
    prob = P.shockbubble_problem(1, 3, mx=8)
    forest = P.build_forest(prob)
    rng = np.random.default_rng(7)
    q = rng.standard_normal((forest.npatch, meqn, 12, 12))
    glob = _cpu_glob(L, prob, forest, meqn)

    def value_cb(g, patch, blockno, patchno, a):
        slot = patch.contents.slot
        vals = np.moveaxis(q[slot][:, 2:10, 2:10], 0, -1).astype(np.float32).reshape(-1)
        C.memmove(a, vals.ctypes.data, vals.nbytes)
    vcb = VTK_CB(value_cb)
    base = str(tmp_path / "frame")
    assert L.fclaw2d_vtk_write_file(glob, base.encode(), 8, 8, meqn, 0.0, 0,
                                    C.cast(L.fclaw2d_output_vtk_coordinate_cb, C.c_void_p),
                                    C.cast(vcb, C.c_void_p)) == 0
    got = open(base + ".vtu", "rb").read()
    assert got == fclaw46_vtk.vtk_bytes(_patches(prob, forest, q), 8, 8, meqn)
    L.fclaw2d_global_destroy(glob)


@pytest.mark.gpu
def test_device_frame_matches_oracle_bytes(gpu, tmp_path):
    """fclaw2d_clawpatch_output_vtk after a few steps of the dam break."""
    from test_gpu_regrid import _dambreak_qinit
    L = _bind_vtk(D.bind())
    prob = P.dambreak_problem(2, 4, mx=16)
    forest = P.build_forest(prob)
    glob = _cpu_glob(L, prob, forest, 3)
    glob.contents.advance_one_step = 1
    opt = D.Options()
    L.fc2d_hipclaw_options_default(C.byref(opt), 3)
    for k in range(3):
        opt.mthlim[k] = 4
    L.fc2d_hipclaw_options_store(glob, C.byref(opt))
    L.fc2d_hipclaw_solver_initialize(glob)
    vt = L.fc2d_hipclaw_vt(glob).contents
    vt.riemann_solver = fa.RP_SHALLOW
    qinit_cb = _dambreak_qinit()
    vt.fort_qinit = C.cast(qinit_cb, C.c_void_p).value
    L.fc2d_hipclaw_initialize_GPUs(glob)
    glob.contents.minlevel, glob.contents.maxlevel = prob.minlevel, prob.minlevel
    L.fclaw2d_initialize(glob)
    q0, _ = P.initial_data(prob, forest)
    dt = P.stable_dt(prob, forest, q0) * 0.8
    for n in range(3):
        L.fclaw2d_advance_all_levels(glob, n * dt, dt)
    prefix = str(tmp_path / "dam").encode()
    glob.contents.prefix = prefix
    L.fclaw2d_clawpatch_output_vtk(glob, 3)
    got = open(str(tmp_path / "dam_frame_0003.vtu"), "rb").read()
    L.fc2d_hipclaw_download(glob)
    q = D.host_q(glob, forest.npatch, (3, 20, 20)).copy()
    assert got == fclaw46_vtk.vtk_bytes(_patches(prob, forest, q), 16, 16, 3)
    # and the host-view callback path writes the same file
    base = str(tmp_path / "host")
    assert L.fclaw2d_vtk_write_file(glob, base.encode(), 16, 16, 3, 0.0, 0,
                                    C.cast(L.fclaw2d_output_vtk_coordinate_cb, C.c_void_p),
                                    C.cast(L.fclaw2d_output_vtk_value_cb, C.c_void_p)) == 0
    assert open(base + ".vtu", "rb").read() == got
    L.fclaw2d_global_destroy(glob)

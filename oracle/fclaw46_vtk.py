"""Oracle (test infrastructure only): numpy restatement of the reference's VTK
frame writer, fclaw2d_vtk_write_file with the non-manifold coordinate and
value callbacks (src/patches/clawpatch/fclaw2d_clawpatch_output_vtk.c:97-900):

  header   : fclaw2d_vtk_write_header (:97-201), one rank (mpirank 0)
  arrays   : sizes and offsets of fclaw2d_vtk_write_file (:597-663); per
             array a UInt64 byte count, then one block per patch in patch
             order (write_position_cb .. write_meqn_cb, :221-452)
  points   : fclaw2d_output_vtk_coordinate_cb (:698-743), manifold = F
  values   : fclaw2d_output_vtk_value_cb (:781-808), float of q(i, j, m)
  footer   : fclaw2d_vtk_write_footer (:536-567)

The reference writer needs p4est/libsc, so it is not built here; this
restatement is pinned by decoding its output against closed forms (tests).
"""
import numpy as np


def vtk_bytes(patches, mx, my, meqn):
    """patches: list of (blockno, xlower, ylower, dx, dy, q) in global patch
    order, q of shape (meqn, my + 2 mbc, mx + 2 mbc).  Returns the file bytes."""
    n = len(patches)
    ppp, cpp = (mx + 1) * (my + 1), mx * my
    npoints, ncells = ppp * n, cpp * n
    fits32 = npoints <= 2 ** 31 - 1 and 4 * (ncells + 1) <= 2 ** 31 - 1
    itype, idt, isz = ("Int32", np.int32, 4) if fits32 else ("Int64", np.int64, 8)
    psize = [ppp * 3 * 8, cpp * 4 * isz, cpp * isz, cpp, cpp * 4, cpp * 4, cpp * isz, cpp * meqn * 4]
    off = [0]
    for p in psize:
        off.append(off[-1] + 8 + p * n)
    h = []
    h.append('<?xml version="1.0"?>\n')
    h.append('<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" '
             'header_type="UInt64" >\n')
    h.append(' <UnstructuredGrid>\n')
    h.append('  <Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (npoints, ncells))
    h.append('   <Points>\n')
    h.append('    <DataArray type="Float64" Name="Position" NumberOfComponents="3" '
             'format="appended" offset="%d">\n' % off[0])
    h.append('    </DataArray>\n')
    h.append('   </Points>\n')
    h.append('   <Cells>\n')
    for name, k in (("connectivity", 1), ("offsets", 2)):
        h.append('    <DataArray type="%s" Name="%s" format="appended" offset="%d">\n'
                 % (itype, name, off[k]))
        h.append('    </DataArray>\n')
    h.append('    <DataArray type="UInt8" Name="types" format="appended" offset="%d">\n' % off[3])
    h.append('    </DataArray>\n')
    h.append('   </Cells>\n')
    h.append('   <CellData Scalars="mpirank,blockno,patchno" Fields="meqn">\n')
    h.append('    <DataArray type="Int32" Name="mpirank" format="appended" offset="%d">\n' % off[4])
    h.append('    </DataArray>\n')
    h.append('    <DataArray type="Int32" Name="blockno" format="appended" offset="%d">\n' % off[5])
    h.append('    </DataArray>\n')
    h.append('    <DataArray type="%s" Name="patchno" format="appended" offset="%d">\n'
             % (itype, off[6]))
    h.append('    </DataArray>\n')
    if meqn == 1:
        h.append('    <DataArray type="Float32" Name="meqn" format="appended" offset="%d">\n' % off[7])
    else:
        h.append('    <DataArray type="Float32" Name="meqn" NumberOfComponents="%d" '
                 'format="appended" offset="%d">\n' % (meqn, off[7]))
    h.append('    </DataArray>\n')
    h.append('   </CellData>\n')
    h.append('   <PointData>\n')
    h.append('   </PointData>\n')
    h.append('  </Piece>\n')
    h.append(' </UnstructuredGrid>\n')
    h.append(' <AppendedData encoding="raw">\n  _')
    out = [''.join(h).encode()]

    def count(k):
        out.append(np.uint64(psize[k] * n).tobytes())

    i = np.arange(mx + 1)
    j = np.arange(my + 1)
    count(0)
    for blockno, xl, yl, dx, dy, q in patches:
        pts = np.zeros((my + 1, mx + 1, 3))
        pts[:, :, 0] = xl + i[None, :] * dx
        pts[:, :, 1] = yl + j[:, None] * dy
        out.append(pts.astype('<f8').tobytes())
    count(1)
    for gp in range(n):
        l = ppp * gp + np.arange(mx)[None, :] + np.arange(my)[:, None] * (mx + 1)
        quad = np.stack([l, l + 1, l + mx + 2, l + mx + 1], axis=-1)
        out.append(quad.astype(idt).tobytes())
    count(2)
    for gp in range(n):
        out.append((4 * (cpp * gp + 1) + 4 * np.arange(cpp)).astype(idt).tobytes())
    count(3)
    for gp in range(n):
        out.append(np.full(cpp, 9, np.uint8).tobytes())
    count(4)
    for gp in range(n):
        out.append(np.zeros(cpp, np.int32).tobytes())
    count(5)
    for blockno, *_ in patches:
        out.append(np.full(cpp, blockno, np.int32).tobytes())
    count(6)
    for gp in range(n):
        out.append(np.full(cpp, gp, idt).tobytes())
    count(7)
    for blockno, xl, yl, dx, dy, q in patches:
        mbc = (q.shape[2] - mx) // 2
        interior = q[:, mbc:mbc + my, mbc:mbc + mx]             # (meqn, my, mx)
        out.append(np.ascontiguousarray(np.moveaxis(interior, 0, -1)).astype('<f4').tobytes())
    out.append(b'\n </AppendedData>\n</VTKFile>\n')
    return b''.join(out)

"""Solution export: the modal displacement evaluated at the four vertices of
every tetrahedron (discontinuous nodal values, as EXPORT_SOLUTION computes
them, src/solution_processing.f90:100-180) and written as a VTK unstructured
grid (.vtu, binary appended data), readable by ParaView.

The vertex values are computed on the GPU by the receiver kernel (one
receiver per tet vertex, located in its own tet); this module only writes
the file.
"""
import struct

import numpy as np


def vertex_values(ctx, slot=None):
    """(n_tet, 4, 3) displacement at the vertices of every tet, from `slot`
    (default: the current state u^n)."""
    from . import _lib
    m = ctx.mesh
    pts = m.xyz[m.tet_vert].reshape(-1, 3)
    tets = np.repeat(np.arange(m.n_tet, dtype=np.int32), 4)
    ctx.set_receivers(pts, max_records=0, tets=tets)
    vals = ctx.sample_receivers(_lib.SLOT_U_CUR if slot is None else slot)
    return vals.reshape(m.n_tet, 4, 3)


def write_vtu(path, mesh, values, name="DISPLACEMENT"):
    """Write tets with discontinuous per-vertex vectors values[(n_tet, 4, 3)].
    Each tet gets its own four points so jumps across faces are visible."""
    nt = mesh.n_tet
    pts = np.ascontiguousarray(mesh.xyz[mesh.tet_vert].reshape(-1, 3), np.float64)
    conn = np.arange(4 * nt, dtype=np.int64)
    offs = 4 * np.arange(1, nt + 1, dtype=np.int64)
    types = np.full(nt, 10, np.uint8)  # VTK_TETRA
    vals = np.ascontiguousarray(np.asarray(values, np.float64).reshape(-1, 3))
    poly = np.ascontiguousarray(mesh.tet_poly, np.int32)
    arrays = [pts, conn, offs, types, vals, poly]
    offsets, pos = [], 0
    for a in arrays:
        offsets.append(pos)
        pos += 8 + a.nbytes
    hdr = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
        "<UnstructuredGrid>\n"
        f'<Piece NumberOfPoints="{4 * nt}" NumberOfCells="{nt}">\n'
        f'<PointData Vectors="{name}"><DataArray type="Float64" Name="{name}" NumberOfComponents="3" '
        f'format="appended" offset="{offsets[4]}"/></PointData>\n'
        f'<CellData Scalars="polyhedron"><DataArray type="Int32" Name="polyhedron" format="appended" '
        f'offset="{offsets[5]}"/></CellData>\n'
        f'<Points><DataArray type="Float64" NumberOfComponents="3" format="appended" offset="{offsets[0]}"/></Points>\n'
        "<Cells>\n"
        f'<DataArray type="Int64" Name="connectivity" format="appended" offset="{offsets[1]}"/>\n'
        f'<DataArray type="Int64" Name="offsets" format="appended" offset="{offsets[2]}"/>\n'
        f'<DataArray type="UInt8" Name="types" format="appended" offset="{offsets[3]}"/>\n'
        "</Cells>\n</Piece>\n</UnstructuredGrid>\n"
        '<AppendedData encoding="raw">\n_'
    )
    with open(path, "wb") as f:
        f.write(hdr.encode())
        for a in arrays:
            f.write(struct.pack("<Q", a.nbytes))
            f.write(a.tobytes())
        f.write(b"\n</AppendedData>\n</VTKFile>\n")


def read_vtu_arrays(path):
    """Minimal reader of files written by write_vtu (for round-trip tests):
    returns (points, values, polyhedron)."""
    import re
    raw = open(path, "rb").read()
    head, data = raw.split(b'<AppendedData encoding="raw">\n_', 1)
    h = head.decode()
    npts = int(re.search(r'NumberOfPoints="(\d+)"', h).group(1))
    ncell = int(re.search(r'NumberOfCells="(\d+)"', h).group(1))

    def block(off, dtype, shape):
        n = struct.unpack("<Q", data[off:off + 8])[0]
        return np.frombuffer(data[off + 8:off + 8 + n], dtype).reshape(shape)

    offs = [int(x) for x in re.findall(r'offset="(\d+)"', h)]
    vals = block(offs[0], np.float64, (npts, 3))
    poly = block(offs[1], np.int32, (ncell,))
    pts = block(offs[2], np.float64, (npts, 3))
    return pts, vals, poly

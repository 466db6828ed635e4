"""Solution export: the modal displacement evaluated at the four vertices of
every tetrahedron (discontinuous nodal values, as EXPORT_SOLUTION computes
them, src/solution_processing.f90:100-180), written as a VTK unstructured
grid (.vtu, binary appended data, readable by ParaView) or as the reference's
EnSight Gold ASCII monitor files (mesh.geo, poly.sca, boundaries.case,
sol_tet_NNNNNN.vec, solution.case; src/export_file_formats.f90:580-988).

The vertex values are computed on the GPU by the receiver kernel (one
receiver per tet vertex, located in its own tet); this module only writes
the file.
"""
import math
import os
import struct

import numpy as np

from .mesh import FACE_SETS


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


# ----------------------------------------------------------------------------
# EnSight Gold ASCII, as the reference's monitors write it: reals one per line
# as Fortran E12.5, integers one per line as I10; tets written with their own
# four vertices (per-tet, discontinuous), vertex order as in the mesh.

def fortran_e(x, w=12, d=5):
    """Fortran Ew.d edit of one real: 0.ddddd mantissa, E+ee exponent (the E
    is dropped for |exponent| > 99, as the standard prescribes), right-aligned
    in w columns; the d digits are correctly rounded to nearest."""
    x = float(x)
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        return ("-Infinity" if x < 0 else "Infinity").rjust(w)
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return (sign + "0." + "0" * d + "E+00").rjust(w)
    mant, ex = f"{abs(x):.{d - 1}e}".split("e")
    e = int(ex) + 1
    tail = f"E{e:+03d}" if abs(e) <= 99 else f"{e:+04d}"
    return (sign + "0." + mant.replace(".", "") + tail).rjust(w)


def _i10(k):
    return f"{int(k):10d}"


def _ensight_header(strict):
    # the reference declares "id assign" but writes the ids all the same;
    # "id given" makes the written ids part of the format
    ids = "assign" if strict else "given"
    return ["Ensight Gold", "Unstructured mesh part", f"node id {ids}", f"element id {ids}"]


def _reals(vals):
    return [fortran_e(v) for v in np.asarray(vals, np.float64).ravel()]


def write_ensight_mesh(folder, mesh, strict=False):
    """mesh.geo (one tetra4 part, 4 points per tet) and poly.sca (1-based
    polyhedron id per tet): ENSIGHT_WRITE_MESH, src/export_file_formats.f90:
    580-680.  strict=True keeps the reference's "id assign" header lines."""
    os.makedirs(folder, exist_ok=True)
    nt = mesh.n_tet
    v = mesh.xyz[mesh.tet_vert]  # (nt, 4, 3), tet-major, vertex inner
    lines = _ensight_header(strict) + ["part", _i10(1), "Static mesh file", "coordinates",
                                       _i10(4 * nt)]
    lines += [_i10(i) for i in range(1, 4 * nt + 1)]
    for d in range(3):
        lines += _reals(v[:, :, d])
    lines += ["tetra4", _i10(nt)]
    lines += [_i10(i) for i in range(1, nt + 1)]
    lines += ["".join(_i10(4 * i + j) for j in range(1, 5)) for i in range(nt)]
    with open(os.path.join(folder, "mesh.geo"), "w") as f:
        f.write("\n".join(lines) + "\n")
    lines = ["scalar per element", "part", _i10(1), "tetra4"]
    lines += [_i10(q) for q in mesh.tet_poly + 1]
    with open(os.path.join(folder, "poly.sca"), "w") as f:
        f.write("\n".join(lines) + "\n")


def boundary_faces(mesh):
    """(tet, face) of every boundary face, tet-major then face 1..4 (the
    reference's loop order), with the boundary tag of its triangle."""
    nn = np.int64(mesh.xyz.shape[0])

    def key(tri):
        s = np.sort(tri, axis=-1).astype(np.int64)
        return (s[..., 0] * nn + s[..., 1]) * nn + s[..., 2]

    tet, face = np.nonzero(mesh.neigh < 0)
    fk = key(mesh.tet_vert[tet][np.arange(tet.size)[:, None], FACE_SETS[face]])
    tk = key(mesh.tria_vert)
    order = np.argsort(tk, kind="stable")
    pos = np.searchsorted(tk[order], fk)
    return tet, face, mesh.tria_tag[order][pos]


def write_ensight_boundary(folder, mesh, face_vertices=True, strict=False):
    """boundaries.case: a tria3 part of every boundary face with its boundary
    tag as a per-element scalar (ENSIGHT_WRITE_BOUNDARY, :682-833).  The
    reference gives every boundary face the coordinates of the first three
    vertices of its tet, whatever the face (:738-748); face_vertices=False
    reproduces that, True (default) writes the face's own vertices."""
    os.makedirs(folder, exist_ok=True)
    tet, face, tag = boundary_faces(mesh)
    nb = tet.size
    if face_vertices:
        vid = mesh.tet_vert[tet][np.arange(nb)[:, None], FACE_SETS[face]]
    else:
        vid = mesh.tet_vert[tet, :3]
    v = mesh.xyz[vid]  # (nb, 3, 3)
    lines = _ensight_header(strict) + ["part", _i10(1), "boundary faces", "coordinates",
                                       _i10(3 * nb)]
    lines += [_i10(i) for i in range(1, 3 * nb + 1)]
    for d in range(3):
        lines += _reals(v[:, :, d])
    lines += ["tria3", _i10(nb)]
    lines += [_i10(i) for i in range(1, nb + 1)]
    lines += ["".join(_i10(3 * i + j) for j in range(1, 4)) for i in range(nb)]
    lines += ["scalar per element", "part", _i10(1), "tria3"]
    lines += [_i10(t) for t in tag]
    with open(os.path.join(folder, "boundaries.case"), "w") as f:
        f.write("\n".join(lines) + "\n")


def ensight_solution_name(num_dt):
    return f"sol_tet_{int(num_dt):06d}.vec"


def write_ensight_solution(folder, values, num_dt):
    """sol_tet_NNNNNN.vec: the vertex displacements values[(n_tet, 4, 3)] as a
    vector per node (ENSIGHT_WRITE_SOLUTION, :837-921): all x components
    (tet-major, vertex inner), then all y, then all z."""
    os.makedirs(folder, exist_ok=True)
    u = np.asarray(values, np.float64).reshape(-1, 4, 3)
    lines = ["vector per node", "part", _i10(1), "coordinates"]
    for d in range(3):
        lines += _reals(u[:, :, d])
    path = os.path.join(folder, ensight_solution_name(num_dt))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def write_ensight_case(folder, time_step, stop_time, every):
    """solution.case (ENSIGHT_WRITE_CASE, :925-988): geometry mesh.geo, the
    vector u over sol_tet_******.vec and the scalar poly; the time set has
    floor(stop_time/time_step)/every + 1 steps, filenames from 0 in
    increments of `every`, times i*every*time_step as F10.4, ten per line."""
    os.makedirs(folder, exist_ok=True)
    ndt = int(math.floor(stop_time / time_step)) // int(every) + 1
    out = ["FORMAT", "type: ensight gold", "GEOMETRY", "model: mesh.geo", "VARIABLE",
           "vector per node: " + _i10(1) + " u sol_tet_******.vec",
           "scalar per element: " + _i10(1) + " poly poly.sca",
           "TIME", "time set: " + _i10(1), "number of steps: " + _i10(ndt),
           "filename start number: " + _i10(0), "filename increment: " + _i10(every),
           "time values:"]
    vals = [f"{i * every * time_step:10.4f}" for i in range(ndt)]
    out += ["".join(vals[k:k + 10]) for k in range(0, ndt, 10)]
    path = os.path.join(folder, "solution.case")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return path


def read_ensight_reals(path, skip):
    """The reals of an EnSight ASCII file written above, after `skip` header
    lines (for round-trip tests)."""
    with open(path) as f:
        lines = f.read().splitlines()[skip:]
    return np.array([float(x) for x in lines])

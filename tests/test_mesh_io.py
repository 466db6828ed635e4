"""Deck, partition-file and EnSight writers: round trips through the readers,
the reference's formats (src/Poly_readfile.f90, src/MAKE_PARTITION_AND_MPI_FILES.f90,
src/MESH_CORRECTION.f90, src/export_file_formats.f90:580-988) and Fortran's
E edit descriptor."""
import os

import numpy as np
import pytest

from lymph3d_apsc_amd import export as X
from lymph3d_apsc_amd import mesh as M

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def same_mesh(a, b):
    assert np.array_equal(a.xyz, b.xyz)
    assert np.array_equal(a.tet_vert, b.tet_vert)
    assert np.array_equal(a.tet_poly, b.tet_poly)
    assert np.array_equal(a.tria_vert, b.tria_vert) and np.array_equal(a.tria_tag, b.tria_tag)
    assert np.array_equal(a.neigh, b.neigh) and np.array_equal(a.bc_space, b.bc_space)
    assert np.array_equal(a.bbox, b.bbox) and np.array_equal(a.hk, b.hk)


def test_deck_round_trip_agglomerated_heterogeneous(tmp_path):
    m = M.cube_mesh(3, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 12), seed=2)
    rng = np.random.default_rng(0)
    # three materials, so the .mate file carries three MATE lines
    k = rng.integers(0, 3, m.n_tet)
    m.rho = np.array([1.0, 1.3, 0.8])[k]
    m.lam = np.array([2.0, 1.1, 3.7])[k]
    m.mu = np.array([1.0, 0.6, 1.9])[k]
    M.write_deck(tmp_path, m, degree=3, dt=1e-3, stop_time=0.5)
    r = M.read_deck(tmp_path, force_unit_material=False)
    same_mesh(r, m)
    assert r.degree == 3 and r.dt == 1e-3 and r.stop_time == 0.5
    for f in ("rho", "lam", "mu"):
        assert np.allclose(getattr(r, f), getattr(m, f), rtol=1e-14, atol=0)
    mate = open(tmp_path / "mat.mate").read()
    assert mate.count("MATE") == 3 and "DIRI    2" in mate and "NEUM    5" in mate
    # every line of the .mesh fits the reference's 100-character read buffer
    assert max(len(s) for s in open(tmp_path / "grid.mesh").read().splitlines()) < 100


def test_deck_round_trip_reference_deck(tmp_path):
    a = M.read_deck(os.path.join(DECKS, "tet_25"))
    M.write_deck(tmp_path, a, degree=a.degree, dt=a.dt, stop_time=a.stop_time)
    b = M.read_deck(tmp_path)
    same_mesh(a, b)
    # a per-tet deck carries no poly lines
    assert " poly " not in open(tmp_path / "grid.mesh").read()


def test_partition_files(tmp_path):
    m = M.cube_mesh(2, agglomerate=True)
    path = M.write_elem_in_poly(tmp_path, m)
    assert os.path.basename(path) == "elem_in_poly_000000.mpi"
    assert np.array_equal(M.read_elem_in_poly(path), m.tet_poly)
    assert np.array_equal(M.read_elem_in_poly(path, n_poly=m.n_poly), m.tet_poly)
    lines = open(M.write_poly_global(tmp_path, m)).read().split()
    assert int(lines[0]) == m.n_tet and [int(x) - 1 for x in lines[1:]] == m.tet_poly.tolist()


def test_mesh_correction_fills_empty_polyhedra(tmp_path):
    # ids 1..5 with 3 and 5 empty: scanning from polyhedron 1, the first tet of
    # each polyhedron with more than one tet goes to the last missing id
    # (MESH_CORRECTION): tet 1 -> 5, then tet 4 -> 3
    p = tmp_path / "elem_in_poly_000000.mpi"
    ids = [1, 1, 1, 2, 2, 4, 4, 4]
    p.write_text(f" {len(ids)}\n" + "".join(f" {i + 1} {q}\n" for i, q in enumerate(ids)))
    assert M.read_elem_in_poly(p, n_poly=5).tolist() == [4, 0, 0, 2, 1, 3, 3, 3]
    p.write_text(" 2\n 1 1\n 2 1\n")
    with pytest.raises(M.MeshError):
        M.read_elem_in_poly(p, n_poly=4)  # one spare tet cannot fill three ids


@pytest.mark.parametrize("x,s", [
    (1.0, " 0.10000E+01"), (-1.5e-3, "-0.15000E-02"), (0.0, " 0.00000E+00"),
    (123456.7, " 0.12346E+06"), (999999.9, " 0.10000E+07"), (1e100, " 0.10000+101"),
    (-2.5e-120, "-0.25000-119"), (0.1, " 0.10000E+00"),
])
def test_fortran_e12_5(x, s):
    assert X.fortran_e(x) == s


def test_ensight_files(tmp_path):
    m = M.cube_mesh(2, agglomerate=True, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"})
    nt = m.n_tet
    X.write_ensight_mesh(tmp_path, m)
    geo = open(tmp_path / "mesh.geo").read().splitlines()
    assert geo[:8] == ["Ensight Gold", "Unstructured mesh part", "node id given", "element id given",
                       "part", f"{1:10d}", "Static mesh file", "coordinates"]
    assert int(geo[8]) == 4 * nt
    xyz = np.array([float(v) for v in geo[9 + 4 * nt:9 + 16 * nt]]).reshape(3, nt, 4)
    v = m.xyz[m.tet_vert]
    assert np.allclose(xyz, np.moveaxis(v, 2, 0), rtol=5e-5, atol=1e-12)
    k = 9 + 16 * nt
    assert geo[k] == "tetra4" and int(geo[k + 1]) == nt
    conn = [list(map(int, s.split())) for s in geo[k + 2 + nt:]]
    assert conn == [[4 * i + 1, 4 * i + 2, 4 * i + 3, 4 * i + 4] for i in range(nt)]
    sca = open(tmp_path / "poly.sca").read().split()
    assert [int(q) - 1 for q in sca[6:]] == m.tet_poly.tolist()
    # strict keeps the reference's header
    X.write_ensight_mesh(tmp_path / "s", m, strict=True)
    assert open(tmp_path / "s" / "mesh.geo").read().splitlines()[2] == "node id assign"
    # boundary faces: one tria3 per boundary triangle, with its tag
    X.write_ensight_boundary(tmp_path, m)
    bd = open(tmp_path / "boundaries.case").read().splitlines()
    nb = m.tria_vert.shape[0]
    assert int(bd[8]) == 3 * nb
    tags = [int(t) for t in bd[-nb:]]
    assert sorted(tags) == sorted(m.tria_tag.tolist())
    pts = np.array([float(v) for v in bd[9 + 3 * nb:9 + 12 * nb]]).reshape(3, nb, 3)
    # each written triangle lies on a box face: one coordinate constant 0 or 1
    on_face = [(np.ptp(pts[d], axis=1) == 0) & np.isin(pts[d][:, 0], (0.0, 1.0)) for d in range(3)]
    assert np.all(on_face[0] | on_face[1] | on_face[2])
    # solution file: x components of all tet vertices, then y, then z
    vals = np.random.default_rng(1).standard_normal((nt, 4, 3))
    path = X.write_ensight_solution(tmp_path, vals, 300)
    assert os.path.basename(path) == "sol_tet_000300.vec"
    got = X.read_ensight_reals(path, 4).reshape(3, nt, 4)
    assert np.allclose(got, np.moveaxis(vals, 2, 0), rtol=5e-5, atol=1e-12)


def test_ensight_case_file(tmp_path):
    # stop 1, dt 1e-3, every 100: floor(1/1e-3)/100 + 1 = 11 time values
    X.write_ensight_case(tmp_path, 1e-3, 1.0, 100)
    c = open(tmp_path / "solution.case").read().splitlines()
    assert c[:5] == ["FORMAT", "type: ensight gold", "GEOMETRY", "model: mesh.geo", "VARIABLE"]
    assert c[5] == "vector per node:          1 u sol_tet_******.vec"
    assert c[6] == "scalar per element:          1 poly poly.sca"
    assert c[9] == "number of steps:         11" and c[11] == "filename increment:        100"
    assert c[12] == "time values:"
    assert c[13] == "".join(f"{0.1 * i:10.4f}" for i in range(10)) and c[14] == "    1.0000"

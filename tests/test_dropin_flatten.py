"""The reference-named Fortran drop-in's mesh flattening (LYMPH_FLATTEN_MESH in
lymph3d_apsc_amd/fortran/lymph_matrix_free.f90): a Mesh_Structure and
Data_Structure filled by oracle/_ref/dropin_harness in the reference's
conventions (1-based ids, neigh_el(face, 0:5), normal(face, d), b_box(3, 2),
prop_mat(mat, 1:3)) must flatten to exactly the LymphMesh arrays
lymph3d_apsc_amd.mesh builds for the same mesh.  No GPU is used; skips where
the harness is not built (it needs the reference tree)."""
import os
import subprocess

import numpy as np
import pytest

from lymph3d_apsc_amd import mesh as M
from tests import dropin_files

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_harness")
DECKS = os.path.join(ROOT, "tests", "golden", "decks")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/dropin_harness not built")


def _mesh(name):
    if name.startswith("tet_"):
        return M.read_deck(os.path.join(DECKS, name))
    if name == "grown":
        m = M.cube_mesh(4, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 12), seed=3)
        rng = np.random.default_rng(3)
        m.lam = np.exp(rng.normal(0, 0.5, m.n_tet))
        m.mu = np.exp(rng.normal(0, 0.5, m.n_tet))
        m.rho = rng.uniform(0.8, 1.2, m.n_tet)
        return m
    return M.cube_mesh(3, agglomerate=True)


@pytest.mark.parametrize("name", ["tet_25", "tet_100", "grown", "cube3agg"])
def test_flattened_mesh_equals_the_pipelines(name, tmp_path):
    m = _mesh(name)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    dropin_files.write(str(fin), m, 2, 1e-3, 1.0)
    r = subprocess.run([EXE, "flatten", str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = dropin_files.read_flat(str(fout))
    _, arrs = m.to_c()
    names = ["xyz", "tet_vert", "tet_poly", "neigh", "bc_space", "normal", "area", "bbox", "hk",
             "rho", "lam", "mu"]
    assert (got["n_node"], got["n_tet"], got["n_poly"]) == (m.xyz.shape[0], m.n_tet, m.n_poly)
    for k, a in zip(names, arrs):
        want = np.asarray(a).ravel()
        if k == "tet_poly" and m.tet_poly is None:
            want = np.arange(m.n_tet)
        assert np.array_equal(got[k], want.astype(got[k].dtype)), k

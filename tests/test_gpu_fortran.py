"""The drop-in boundary exercised from Fortran: the reference's time loop
(src/Lymph3D.f90:303-518) in lymph3d_apsc_amd/fortran/drive_mesh.f90 calls
liblymph_gpu through the ISO_C_BINDING module with the reference's
column-major arrays; its printed errors must match the reference's known
answers like the Python path's."""
import json
import os
import subprocess

import pytest

import lymph3d_apsc_amd as L
from lymph3d_apsc_amd import mesh as M

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "lymph3d_apsc_amd", "fortran", "drive_mesh")


@pytest.mark.parametrize("name,mode", [("tet_5", ""), ("tet_25", ""), ("tet_5", "procedures"),
                                       ("tet_25", "procedures"), ("tet_25", "symmetric")])
def test_fortran_reference_loop(name, mode, tmp_path):
    """mode "procedures": the data sets come from Fortran functions passed as
    callbacks (lymph_build_rhs_fn / lymph_project_fn), as the reference passes
    its procedure arguments; mode "symmetric": face pairs stored once
    (lymph_set_symmetric_storage), errors within four times the tolerance."""
    if L.device_count() == 0:
        pytest.skip("no HIP device")
    if not os.path.exists(EXE):
        pytest.fail("Fortran driver not built")
    print(f"[fortran drop-in] {name} {mode}", flush=True)
    ka = json.load(open(os.path.join(GOLD, "known_answers.json")))[name]
    m = M.read_deck(os.path.join(GOLD, "decks", name))
    f = tmp_path / "mesh.bin"
    m.write_flat(str(f), m.degree, m.dt, m.stop_time)
    r = subprocess.run([EXE, str(f)] + ([mode] if mode else []), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = {}
    for line in r.stdout.splitlines():
        if line.startswith("ERROR IN NORM L2:"):
            vals["l2"] = float(line.split(":")[1])
        if line.startswith("ERROR IN NORM DG:"):
            vals["dg"] = float(line.split(":")[1])
        if line.startswith("steps"):
            vals["steps"] = int(line.split()[1])
    assert vals["steps"] == ka["steps"]
    tol = 2e-12 if mode == "symmetric" else 5e-13
    assert abs(vals["l2"] / ka["err_l2"] - 1) < tol
    assert abs(vals["dg"] / ka["err_dg"] - 1) < tol
    if mode == "symmetric":
        assert "face pairs stored once" in r.stdout

"""Host logic of the monitored deck driver (driver.run_monitored) on the CPU:
the chunked device loop must visit the reference's loop iterations
(src/Lymph3D.f90:389-418: t from dt, `while (t <= stop_time)`, t accumulated
by repeated addition, export after iteration num_dt when num_dt % every == 0)
and end on the same step count and time.  The context is a stand-in that
records the stepping calls; no GPU is used."""
import os

import numpy as np
import pytest

from lymph3d_apsc_amd import _lib, driver
from lymph3d_apsc_amd import mesh as M


class FakeCtx:
    def __init__(self, mesh):
        self.mesh, self.built, self.n_poly, self.n3 = mesh, False, mesh.n_poly, 12
        self.steps = []

    def make_matrices_free(self):
        self.built = True

    def make_rhs_free(self, problem, slot):
        pass

    def vec(self):
        return np.zeros(self.n_poly * self.n3)

    def compute_modal_coefficients_free(self, fn, slot=None):
        return self.vec()

    def set_slot(self, slot, v):
        pass

    def get_slot(self, slot):
        return self.vec()

    def first_time_step_matrix_free(self, g, dt, u0, v0):
        return self.vec()

    def step_n(self, problem, k, t, dt):
        self.steps.append(k)
        for _ in range(k):
            t = t + dt
        return t

    def set_receivers(self, pts, max_records=0, tets=None):
        self.npts = len(pts)

    def sample_receivers(self, slot):
        return np.zeros((self.npts, 3))

    def compute_error_l2_matrix_free(self, a, b):
        return 0.0

    compute_error_dg_matrix_free = compute_error_l2_matrix_free


def reference_loop(dt, stop, every, max_steps=-1):
    t, nsteps, num_dt, exports = 0.0 + dt, 1, 0, []
    while t <= stop and (max_steps < 0 or nsteps < max_steps):
        nsteps += 1
        if num_dt % every == 0:
            exports.append(num_dt)
        num_dt += 1
        t = t + dt
    return t, nsteps, exports


@pytest.mark.parametrize("dt,stop,every,max_steps", [
    (1e-3, 1.0, 100, -1), (1e-3, 0.05, 7, -1), (0.1, 1.0, 3, -1), (1e-3, 1.0, 100, 250),
    (1e-3, 5e-4, 10, -1), (3e-3, 0.5, 1, -1)])
def test_monitored_loop_matches_reference_loop(tmp_path, dt, stop, every, max_steps):
    m = M.cube_mesh(1)
    ctx = FakeCtx(m)
    r = driver.run_monitored(ctx, _lib.PROB_REF_MMS, dt, stop, every, str(tmp_path), max_steps)
    t, nsteps, exports = reference_loop(dt, stop, every, max_steps)
    assert r["t_final"] == t and r["nsteps"] == nsteps
    assert sum(ctx.steps) == nsteps - 1
    names = sorted(f for f in os.listdir(tmp_path) if f.endswith(".vec"))
    assert names == [f"sol_tet_{k:06d}.vec" for k in exports]
    assert os.path.exists(tmp_path / "solution.case")
    if exports:
        assert os.path.exists(tmp_path / "mesh.geo") and os.path.exists(tmp_path / "boundaries.case")

"""The deck driver with EnSight monitors (the reference's SAVEOUT path): the
chunked device loop gives the plain run's final state and errors, and the
exported vertex values are those of the state at each export."""
import os
import sys

import numpy as np
import pytest

import lymph3d_apsc_amd as L
from lymph3d_apsc_amd import driver, export as X, mesh as M

pytestmark = pytest.mark.gpu

DECK = os.path.join(os.path.dirname(__file__), "golden", "decks", "tet_5")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if L.device_count() == 0:
        pytest.skip("no HIP device")


def test_monitored_run(tmp_path):
    print("[monitor] tet_5 p=1 every 100", flush=True)
    sys.stdout.flush()
    plain = driver.run_deck(DECK, verbose=False)
    mon = driver.run_deck(DECK, verbose=False, monitor=str(tmp_path), every=100)
    assert mon["nsteps"] == plain["nsteps"] == 1000 and mon["t_final"] == plain["t_final"]
    assert abs(mon["err_l2"] / plain["err_l2"] - 1) < 1e-13
    assert abs(mon["err_dg"] / plain["err_dg"] - 1) < 1e-13
    # loop iterations 0..998: exports after iterations 0, 100, ..., 900
    names = sorted(os.listdir(tmp_path))
    vecs = [n for n in names if n.endswith(".vec")]
    assert vecs == [X.ensight_solution_name(k) for k in range(0, 901, 100)]
    assert {"mesh.geo", "poly.sca", "boundaries.case", "solution.case"} <= set(names)
    # the state after iteration 900 is the one after 902 steps (first step included)
    m = M.read_deck(DECK)
    with L.Context(m, m.degree) as ctx:
        ctx.run(L.PROB_REF_MMS, m.dt, m.stop_time, max_steps=902)
        ref = X.vertex_values(ctx)
    got = X.read_ensight_reals(str(tmp_path / "sol_tet_000900.vec"), 4).reshape(3, m.n_tet, 4)
    assert np.allclose(got, np.moveaxis(ref, 2, 0), rtol=5e-5, atol=1e-12 * np.abs(ref).max())

"""The 48-bit operator apply (lymph_set_apply_precision(ctx, 48)): each entry
of the mass-orthonormal operator kept as its FP32 rounding hi and a 16-bit
correction t = (k - hi)/hi * 2^24, products summed in FP64.

Tolerances: the operator entries carry a relative error of about 2^-36
(1.5e-11), so a single step's acceleration term dt^2 M^-1(F - Ku) is held to
1e-10 of the FP64 oracle's, and to a thousandth of the 32-bit apply's error
on the same step; over 1000 steps of the cube wave on 4^3 agglomerated cubes
at p = 3 the state is held to 5e-8 of the 64-bit run (1.7e-8 measured) and
to a thousandth of the 32-bit run's drift (6.3e-5 measured): the drift scales
with the operator's rounding, 2^-36 against 2^-25.  Variants and the release of the FP64
blocks change no bit; a captured solve matches the uncaptured steps within
1e-12 (its wavelet is evaluated on the device)."""
import numpy as np
import pytest

import lymph3d_apsc_amd as L
from lymph3d_apsc_amd import _lib, mesh as M

from tests.test_gpu_parity import make_mesh, oracle_for, rel, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if L.device_count() == 0:
        pytest.skip("no HIP device")


def _mesh(name):
    if name == "grown4":
        return M.cube_mesh(4, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 40), seed=3)
    return make_mesh(name)


def _prepared(ctx):
    ctx.make_matrices_free()
    ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
    ctx.make_rhs_free(L.PROB_REF_MMS, _lib.SLOT_RHS_DYN)


@pytest.mark.parametrize("name,p", [("cube3agg", 3), ("hetero", 2), ("hetero", 1), ("cube2agg", 4),
                                    ("grown4", 3)])
def test_step_matches_oracle(name, p):
    say(f"[48-bit apply] {name} p={p}")
    m = _mesh(name)
    o = oracle_for(m, p)
    o.build_operator()
    rng = np.random.default_rng(12)
    n = m.n_poly * 3 * o.Np
    u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
    dt, g = 1e-3, 0.37
    ref = o.step(g, dt, u0, u1, o.rhs(L.PROB_NULL), o.rhs(L.PROB_REF_MMS))
    acc = lambda a: a - (2 * u1 - u0)  # noqa: E731
    with L.Context(m, p) as ctx:
        _prepared(ctx)
        ctx.set_apply_precision(32)
        e32 = rel(acc(ctx.time_step_matrix_free(g, dt, u0, u1)), acc(ref))
        before = ctx.device_bytes
        ctx.set_apply_precision(48)
        assert ctx.device_bytes > before  # the corrections
        a48 = ctx.time_step_matrix_free(g, dt, u0, u1)
        f48 = ctx.first_time_step_matrix_free(g, dt, u0, u1)
        ctx.set_apply_precision(64)
        f64 = ctx.first_time_step_matrix_free(g, dt, u0, u1)
    e48 = rel(acc(a48), acc(ref))
    say(f"  acceleration vs oracle: 32-bit {e32:.2e}, 48-bit {e48:.2e}")
    assert e48 < 1e-10, e48
    assert e48 < 1e-3 * e32, (e48, e32)
    assert rel(f48 - u0, f64 - u0) < 1e-10


@pytest.mark.parametrize("name,p", [("hetero", 1), ("hetero", 2), ("cube3agg", 3), ("grown4", 4)])
def test_variants_agree_bitwise(name, p):
    """Every 48-bit step variant (column quads per load group, waves per
    polyhedron) adds the same products in the same order."""
    say(f"[48-bit variants] {name} p={p}")
    m = _mesh(name)
    rng = np.random.default_rng(5)
    with L.Context(m, p) as ctx:
        _prepared(ctx)
        ctx.set_apply_precision(48)
        n = m.n_poly * ctx.n3
        u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
        outs, firsts = [], []
        nv = ctx.n_step_variants()
        assert nv >= 3
        for v in range(nv):
            ctx.set_step_variant(v)
            outs.append(ctx.time_step_matrix_free(0.37, 1e-3, u0, u1))
            firsts.append(ctx.first_time_step_matrix_free(0.37, 1e-3, u0, u1))
        for v in range(1, nv):
            assert np.array_equal(outs[v], outs[0]), v
            assert np.array_equal(firsts[v], firsts[0]), v
        chosen, times = ctx.autotune(reps=2)
        assert len(times) == nv


def test_long_run_drift():
    """1000 steps of the cube wave (projected start, 4^3 agglomerated cubes,
    p = 3, dt at CFL 0.2) with each apply: the 48-bit state's distance from
    the 64-bit one against the 32-bit state's."""
    say("[48-bit apply] 1000-step drift")
    m, p, n = M.cube_mesh(4, agglomerate=True), 3, 4
    dt = 0.2 * (1 / n) / (np.sqrt(3.0) * p * p)
    us = {}
    with L.Context(m, p) as ctx:
        ctx.make_matrices_free()
        for bits in (64, 32, 48):
            ctx.set_apply_precision(bits)
            r = ctx.run(L.PROB_CUBE_WAVE, dt, 1e9, max_steps=1000, want_u=True)
            us[bits] = r["u"]
    d32 = np.linalg.norm(us[32] - us[64]) / np.linalg.norm(us[64])
    d48 = np.linalg.norm(us[48] - us[64]) / np.linalg.norm(us[64])
    say(f"  drift vs 64-bit after 1000 steps: 32-bit {d32:.2e}, 48-bit {d48:.2e}")
    assert d48 < 5e-8, d48
    assert d48 < 1e-3 * d32, (d48, d32)


def test_release_and_capture_keep_bits():
    """After lymph_release_fp64_operator the 48-bit apply steps bit for bit
    as before (and cannot be selected anew without the FP64 blocks on a
    fresh context); a captured 48-bit solve replays to the same state and
    traces as the same steps uncaptured."""
    say("[48-bit apply] release, capture")
    m, p = make_mesh("hetero"), 2
    dt, nsteps = 2e-3, 12
    rng = np.random.default_rng(9)
    rec = np.array([[0.3, 0.4, 0.5], [0.6, 0.2, 0.7]])
    src, force = np.array([[0.45, 0.55, 0.5]]), np.array([[0.0, 0.0, 1.0]])
    with L.Context(m, p) as ctx:
        _prepared(ctx)
        n = m.n_poly * ctx.n3
        u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
        ctx.set_apply_precision(48)
        a = ctx.time_step_matrix_free(0.37, 1e-3, u0, u1)
        before = ctx.device_bytes
        ctx.release_fp64_operator()
        assert ctx.device_bytes < before
        assert np.array_equal(ctx.time_step_matrix_free(0.37, 1e-3, u0, u1), a)
        ctx.set_apply_precision(32)
        ctx.set_apply_precision(48)  # the corrections were kept
        assert np.array_equal(ctx.time_step_matrix_free(0.37, 1e-3, u0, u1), a)
        with pytest.raises(L.LymphError):
            ctx.set_apply_precision(64)
        # captured (below) against uncaptured: from rest, first step, steps
        t = ctx.start_from_rest(dt, src, force, f0=5.0)
        ctx.set_receivers(rec, max_records=nsteps)
        ctx.step_n_record(L.PROB_RICKER, nsteps, t, dt, every=1)
        u_eager, tr_eager = ctx.get_slot(_lib.SLOT_U_CUR), ctx.traces()
    with L.Context(m, p) as ctx:
        _prepared(ctx)
        ctx.set_apply_precision(48)
        ctx.start_from_rest(dt, src, force, f0=5.0)
        ctx.set_receivers(rec, max_records=nsteps)
        ctx.capture_solve(L.PROB_RICKER, dt, nsteps, every=1)
        for _ in range(2):
            ctx.replay()
            ctx.sync()
            tr = ctx.traces()
            assert np.abs(tr).max() > 0
            assert rel(tr, tr_eager) < 1e-12, rel(tr, tr_eager)
            assert rel(ctx.get_slot(_lib.SLOT_U_CUR), u_eager) < 1e-12
    with L.Context(m, p) as ctx:
        _prepared(ctx)
        ctx.set_apply_precision(32)
        ctx.release_fp64_operator()
        with pytest.raises(L.LymphError):
            ctx.set_apply_precision(48)  # the corrections need the FP64 blocks

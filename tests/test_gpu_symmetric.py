"""Face pairs stored once (lymph_set_symmetric_storage): with SIPG and one
material on both traces of every face, K_QP = K_PQ^T to rounding, and the FP64
step that applies K_QP as the transpose of P's copy must match the full-row
step and the oracle.

Tolerances: single steps within 1e-13 of the full-row step and of the oracle
(relative to the vector's max; the two sides' blocks differ by a few ulps);
full runs of the reference's decks within four times the known-answer
tolerance of test_gpu_parity (the factored build's margin) and within 1e-11
of the oracle's final vector.  Operators that are not symmetric (NIPG,
per-tet materials) and p = 4 are refused with LYMPH_ERR_ARG and the full rows
stay in use.  The 32- and 48-bit applies with pairs stored once are held to
their full-row steps' bounds against the oracle (1e-5, 1e-10) and to their
full rows' long-run drift."""
import os

import numpy as np
import pytest

import lymph3d_apsc_amd as L
from lymph3d_apsc_amd import _lib, mesh as M
from oracle import oracle

from tests.test_gpu_parity import DECKS, KA, hetero_mesh, oracle_for, rel, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if L.device_count() == 0:
        pytest.skip("no HIP device")


def _mesh(name):
    if name.startswith("tet_"):
        return M.read_deck(os.path.join(DECKS, name))
    if name == "grown":  # polyhedra of 4-40 tets: uneven rows, so the step runs in a dispatch order
        return M.cube_mesh(4, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 40), seed=3)
    return M.cube_mesh(int(name[4]), agglomerate=True)


def _prepared(ctx):
    ctx.make_matrices_free()
    ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
    ctx.make_rhs_free(L.PROB_REF_MMS, _lib.SLOT_RHS_DYN)


@pytest.mark.parametrize("name,p", [("cube3agg", 3), ("tet_25", 2), ("cube2agg", 1), ("cube4agg", 3),
                                    ("grown", 3), ("grown", 2)])
def test_steps_match_full_rows_and_oracle(name, p):
    say(f"[symmetric] steps {name} p={p}")
    m = _mesh(name)
    o = oracle_for(m, p)
    o.build_operator()
    rng = np.random.default_rng(7)
    n = m.n_poly * 3 * o.Np
    u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
    dt, g = 1e-3, 0.37
    with L.Context(m, p) as ctx:
        _prepared(ctx)
        full = ctx.time_step_matrix_free(g, dt, u0, u1)
        first_full = ctx.first_time_step_matrix_free(g, dt, u0, u1)
        bytes_full = ctx.device_bytes
        asym = ctx.set_symmetric_storage(True)
        assert 0.0 <= asym < 1e-12, asym
        assert ctx.device_bytes > bytes_full
        sym = ctx.time_step_matrix_free(g, dt, u0, u1)
        first_sym = ctx.first_time_step_matrix_free(g, dt, u0, u1)
        ref = o.step(g, dt, u0, u1, o.rhs(L.PROB_NULL), o.rhs(L.PROB_REF_MMS))
        acc = lambda a: a - (2 * u1 - u0)  # noqa: E731
        assert rel(acc(sym), acc(full)) < 1e-13, rel(acc(sym), acc(full))
        assert rel(acc(sym), acc(ref)) < 1e-13, rel(acc(sym), acc(ref))
        assert rel(first_sym - u0, first_full - u0) < 1e-13
        # device loop (slot rotation) agrees with the full rows over 40 steps
        ctx.set_slot(_lib.SLOT_U_PREV, u0)
        ctx.set_slot(_lib.SLOT_U_CUR, u1)
        ctx.step_n(L.PROB_REF_MMS, 40, 0.0, 1e-4)
        a = ctx.get_slot(_lib.SLOT_U_CUR)
        assert ctx.set_symmetric_storage(False) == 0.0
        ctx.set_slot(_lib.SLOT_U_PREV, u0)
        ctx.set_slot(_lib.SLOT_U_CUR, u1)
        ctx.step_n(L.PROB_REF_MMS, 40, 0.0, 1e-4)
        b = ctx.get_slot(_lib.SLOT_U_CUR)
        assert ctx.device_bytes == bytes_full
        assert rel(a, b) < 1e-12, rel(a, b)


@pytest.mark.parametrize("name", ["tet_25", "tet_100"])
def test_reference_decks_with_pairs_stored_once(name):
    say(f"[symmetric] known answer {name}")
    ka = KA[name]
    m = M.read_deck(os.path.join(DECKS, name))
    with L.Context(m, m.degree) as ctx:
        ctx.make_matrices_free()
        ctx.set_symmetric_storage(True)
        r = ctx.run(L.PROB_REF_MMS, m.dt, m.stop_time, want_u=True)
    assert r["nsteps"] == ka["steps"]
    tol = 4 * (5e-12 if name == "tet_100" else 5e-13)
    assert abs(r["err_l2"] / ka["err_l2"] - 1) < tol, r
    assert abs(r["err_dg"] / ka["err_dg"] - 1) < tol, r
    o = oracle_for(m, m.degree)
    ro = o.run(oracle.PROB_REF_MMS, m.dt, m.stop_time, want_u=True)
    assert rel(r["u"], ro["u"]) < 1e-11, rel(r["u"], ro["u"])


@pytest.mark.parametrize("case", ["nipg", "materials", "p4"])
def test_non_symmetric_operators_are_refused(case):
    say(f"[symmetric] refusal {case}")
    if case == "materials":
        m, p, theta = hetero_mesh(), 2, -1.0
    else:
        m, p, theta = M.cube_mesh(2, agglomerate=True), (4 if case == "p4" else 2), \
            (1.0 if case == "nipg" else -1.0)
    rng = np.random.default_rng(3)
    with L.Context(m, p, theta=theta) as ctx:
        _prepared(ctx)
        n = m.n_poly * 3 * ((p + 1) * (p + 2) * (p + 3) // 6)
        u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
        before = ctx.time_step_matrix_free(0.3, 1e-3, u0, u1)
        with pytest.raises(L.LymphError) as e:
            ctx.set_symmetric_storage(True)
        assert e.value.code == _lib.ERR_ARG
        msg = str(e.value)
        assert ("not symmetric" in msg) if case != "p4" else ("p <= 3" in msg), msg
        after = ctx.time_step_matrix_free(0.3, 1e-3, u0, u1)
        assert np.array_equal(before, after)


def test_rebuild_drops_the_stored_pairs():
    say("[symmetric] rebuild")
    m = M.cube_mesh(2, agglomerate=True)
    with L.Context(m, 2) as ctx:
        _prepared(ctx)
        base = ctx.device_bytes
        ctx.set_symmetric_storage(True)
        assert ctx.device_bytes > base
        ctx.make_matrices_free()
        assert ctx.device_bytes == base


def test_storage_change_drops_the_captured_solve():
    """A solve captured with the pairs stored once holds their buffers by
    address: switching the storage off drops it (replay refused), and a new
    capture replays the full rows."""
    say("[symmetric] captured solve")
    m = M.cube_mesh(2, agglomerate=True)
    with L.Context(m, 2) as ctx:
        _prepared(ctx)
        ctx.set_symmetric_storage(True)
        ctx.capture_solve(L.PROB_REF_MMS, 1e-3, 20)
        ctx.replay()
        ctx.sync()
        a = ctx.get_slot(_lib.SLOT_U_CUR)
        ctx.set_symmetric_storage(False)
        with pytest.raises(L.LymphError) as e:
            ctx.replay()
        assert e.value.code == _lib.ERR_STATE
        ctx.capture_solve(L.PROB_REF_MMS, 1e-3, 20)
        ctx.replay()
        ctx.sync()
        b = ctx.get_slot(_lib.SLOT_U_CUR)
        assert rel(a, b) < 1e-12, rel(a, b)


@pytest.mark.parametrize("name,first", [("cube3agg", "pairs"), ("cube3agg", "fp32"), ("grown", "pairs")])
def test_fp32_apply_with_pairs_stored_once(name, first):
    """The 32-bit apply with face pairs stored once (p = 3, asked for with
    fp32=True): whichever option comes first, the step's acceleration term is
    within 1e-5 of the FP64 oracle and of the full-row 32-bit step, and the
    32-bit copy of the stored blocks is allocated (and freed with the stored
    pairs).  The grown mesh has uneven rows and a dispatch order."""
    say(f"[symmetric] fp32 apply {name}, {first} first")
    m, p = _mesh(name), 3
    o = oracle_for(m, p)
    o.build_operator()
    rng = np.random.default_rng(11)
    n = m.n_poly * 3 * o.Np
    u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
    dt, g = 1e-3, 0.37
    ref = o.step(g, dt, u0, u1, o.rhs(L.PROB_NULL), o.rhs(L.PROB_REF_MMS))
    acc = lambda a: a - (2 * u1 - u0)  # noqa: E731
    with L.Context(m, p) as ctx:
        _prepared(ctx)
        ctx.set_apply_precision(32)
        full32 = ctx.time_step_matrix_free(g, dt, u0, u1)
        bytes32 = ctx.device_bytes
        ctx.set_apply_precision(64)
        if first == "pairs":
            ctx.set_symmetric_storage(True, fp32=True)
            ctx.set_apply_precision(32)
        else:
            ctx.set_apply_precision(32)
            ctx.set_symmetric_storage(True, fp32=True)
        rp, _ = ctx.pattern()
        n_st = m.n_poly + (ctx.n_blk - m.n_poly) // 2
        assert ctx.device_bytes - bytes32 >= n_st * (3 * o.Np) ** 2 * (8 + 4)
        sym32 = ctx.time_step_matrix_free(g, dt, u0, u1)
        assert rel(acc(sym32), acc(ref)) < 1e-5, rel(acc(sym32), acc(ref))
        # the transposed products run the full rows' chain over the same rounded
        # blocks: the step equals the full-row 32-bit step (to the rare entry
        # whose two FP64 sides round to different floats)
        assert rel(acc(sym32), acc(full32)) < 1e-12, rel(acc(sym32), acc(full32))
        # device loop and first step through the same kernels
        ctx.set_slot(_lib.SLOT_U_PREV, u0)
        ctx.set_slot(_lib.SLOT_U_CUR, u1)
        ctx.step_n(L.PROB_REF_MMS, 20, 0.0, 1e-4)
        a = ctx.get_slot(_lib.SLOT_U_CUR)
        first_sym = ctx.first_time_step_matrix_free(g, dt, u0, u1)
        ctx.set_symmetric_storage(False)
        assert ctx.device_bytes == bytes32
        ctx.set_slot(_lib.SLOT_U_PREV, u0)
        ctx.set_slot(_lib.SLOT_U_CUR, u1)
        ctx.step_n(L.PROB_REF_MMS, 20, 0.0, 1e-4)
        b = ctx.get_slot(_lib.SLOT_U_CUR)
        first_full = ctx.first_time_step_matrix_free(g, dt, u0, u1)
        assert rel(a, b) < 1e-5, rel(a, b)
        assert rel(first_sym - u0, first_full - u0) < 1e-5


def test_fp32_pairs_once_is_opt_in():
    """Face pairs stored once for the FP64 step only (the default mode) leave
    the 32-bit apply on its full rows: no 32-bit copy of the stored blocks is
    made and its step equals the full-row 32-bit step bit for bit;
    lymph_autotune is refused while the step in use reads stored pairs."""
    say("[symmetric] fp32 pairs once is opt-in")
    m, p = M.cube_mesh(3, agglomerate=True), 3
    rng = np.random.default_rng(5)
    with L.Context(m, p) as ctx:
        _prepared(ctx)
        n = m.n_poly * 3 * ctx.Np
        u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
        ctx.set_apply_precision(32)
        full32 = ctx.time_step_matrix_free(0.37, 1e-3, u0, u1)
        ctx.set_apply_precision(64)
        bytes0 = ctx.device_bytes
        ctx.set_symmetric_storage(True)
        with pytest.raises(L.LymphError) as e:
            ctx.autotune(reps=2)
        assert e.value.code == _lib.ERR_STATE
        bytes_fp64 = ctx.device_bytes
        assert bytes_fp64 > bytes0
        ctx.set_apply_precision(32)  # the 32-bit blocks exist already: nothing is added
        assert ctx.device_bytes == bytes_fp64
        a = ctx.time_step_matrix_free(0.37, 1e-3, u0, u1)
        assert np.array_equal(a, full32)
        ctx.autotune(reps=2)  # the 32-bit full rows are in use: allowed
        # switching to both precisions stores the 32-bit pairs once
        before = ctx.device_bytes
        ctx.set_symmetric_storage(True, fp32=True)
        n_st = m.n_poly + (ctx.n_blk - m.n_poly) // 2
        assert ctx.device_bytes - before >= n_st * (3 * ctx.Np) ** 2 * 4
        ctx.set_symmetric_storage(False)


def test_fp32_pairs_once_long_run_drift():
    """1000 steps of the 32-bit apply from the same start on the 4^3
    agglomerated cube at p = 3: the pairs-once state is as far from the FP64
    state as the full rows' (the products are the full rows' bit for bit but
    for rare entries whose two FP64 sides round apart)."""
    say("[symmetric] fp32 pairs once, 1000-step drift")
    m, p = M.cube_mesh(4, agglomerate=True), 3
    dt = 0.2 * (1 / 4) / (np.sqrt(3.0) * p * p)
    with L.Context(m, p) as ctx:
        ctx.set_build("factored")
        ctx.make_matrices_free()
        res = {}
        for key in ("fp64", "full32", "pairs32"):
            ctx.set_symmetric_storage(False)
            ctx.set_apply_precision(64)
            ctx.run(L.PROB_CUBE_WAVE, dt, 1e30, max_steps=1)
            if key != "fp64":
                ctx.set_apply_precision(32)
            if key == "pairs32":
                ctx.set_symmetric_storage(True, fp32=True)
            ctx.step_n(L.PROB_CUBE_WAVE, 1000, dt, dt)
            res[key] = ctx.get_slot(_lib.SLOT_U_CUR)
        ctx.set_symmetric_storage(False)
    nrm = np.linalg.norm(res["fp64"])
    d_full = np.linalg.norm(res["full32"] - res["fp64"]) / nrm
    d_pairs = np.linalg.norm(res["pairs32"] - res["fp64"]) / nrm
    say(f"  drift vs fp64: full rows {d_full:.3e}, pairs once {d_pairs:.3e}")
    assert d_full < 1e-4 and abs(d_pairs - d_full) < 0.02 * d_full, (d_full, d_pairs)


@pytest.mark.parametrize("name,first", [("cube3agg", "pairs"), ("cube3agg", "fp48"), ("grown", "pairs")])
def test_48bit_apply_with_pairs_stored_once(name, first):
    """The 48-bit apply with face pairs stored once (p = 3, fp32=True): the
    stored FP32 blocks and their 16-bit corrections are allocated, the step's
    acceleration term is within 1e-10 of the FP64 oracle (the 48-bit full
    rows' bound) and within 1e-11 of the full-row 48-bit step (its products
    are the full rows' bit for bit where the two sides' blocks round alike),
    and the device loop and first step agree with the full rows to 1e-10."""
    say(f"[symmetric] 48-bit apply {name}, {first} first")
    m, p = _mesh(name), 3
    o = oracle_for(m, p)
    o.build_operator()
    rng = np.random.default_rng(13)
    n = m.n_poly * 3 * o.Np
    u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
    dt, g = 1e-3, 0.37
    ref = o.step(g, dt, u0, u1, o.rhs(L.PROB_NULL), o.rhs(L.PROB_REF_MMS))
    acc = lambda a: a - (2 * u1 - u0)  # noqa: E731
    with L.Context(m, p) as ctx:
        _prepared(ctx)
        ctx.set_apply_precision(48)
        full48 = ctx.time_step_matrix_free(g, dt, u0, u1)
        bytes48 = ctx.device_bytes
        ctx.set_apply_precision(64)
        if first == "pairs":
            ctx.set_symmetric_storage(True, fp32=True)
            ctx.set_apply_precision(48)
        else:
            ctx.set_apply_precision(48)
            ctx.set_symmetric_storage(True, fp32=True)
        n_st = m.n_poly + (ctx.n_blk - m.n_poly) // 2
        assert ctx.device_bytes - bytes48 >= n_st * (3 * o.Np) ** 2 * (8 + 4 + 2)
        with pytest.raises(L.LymphError) as e:
            ctx.autotune(reps=2)
        assert e.value.code == _lib.ERR_STATE
        sym48 = ctx.time_step_matrix_free(g, dt, u0, u1)
        e_or, e_full = rel(acc(sym48), acc(ref)), rel(acc(sym48), acc(full48))
        say(f"  acceleration: vs oracle {e_or:.2e}, vs full-row 48-bit {e_full:.2e}")
        assert e_or < 1e-10, e_or
        assert e_full < 1e-11, e_full
        ctx.set_slot(_lib.SLOT_U_PREV, u0)
        ctx.set_slot(_lib.SLOT_U_CUR, u1)
        ctx.step_n(L.PROB_REF_MMS, 20, 0.0, 1e-4)
        a = ctx.get_slot(_lib.SLOT_U_CUR)
        first_sym = ctx.first_time_step_matrix_free(g, dt, u0, u1)
        ctx.set_symmetric_storage(False)
        assert ctx.device_bytes == bytes48
        ctx.set_slot(_lib.SLOT_U_PREV, u0)
        ctx.set_slot(_lib.SLOT_U_CUR, u1)
        ctx.step_n(L.PROB_REF_MMS, 20, 0.0, 1e-4)
        b = ctx.get_slot(_lib.SLOT_U_CUR)
        first_full = ctx.first_time_step_matrix_free(g, dt, u0, u1)
        assert rel(a, b) < 1e-10, rel(a, b)
        assert rel(first_sym - u0, first_full - u0) < 1e-10


def test_48bit_pairs_once_long_run_drift_and_capture():
    """1000 steps of the 48-bit apply from the same start on the 4^3
    agglomerated cube at p = 3: with face pairs stored once the state is as
    far from the FP64 state as the full rows' (within 5%, both under 5e-8);
    a captured solve with the stored pairs replays to the uncaptured steps'
    state within 1e-12."""
    say("[symmetric] 48-bit pairs once, 1000-step drift")
    m, p = M.cube_mesh(4, agglomerate=True), 3
    dt = 0.2 * (1 / 4) / (np.sqrt(3.0) * p * p)
    with L.Context(m, p) as ctx:
        ctx.set_build("factored")
        ctx.make_matrices_free()
        res = {}
        for key in ("fp64", "full48", "pairs48"):
            ctx.set_symmetric_storage(False)
            ctx.set_apply_precision(64)
            ctx.run(L.PROB_CUBE_WAVE, dt, 1e30, max_steps=1)
            if key != "fp64":
                ctx.set_apply_precision(48)
            if key == "pairs48":
                ctx.set_symmetric_storage(True, fp32=True)
            ctx.step_n(L.PROB_CUBE_WAVE, 1000, dt, dt)
            res[key] = ctx.get_slot(_lib.SLOT_U_CUR)
        # captured against uncaptured, both with the stored pairs (from rest, Ricker point force)
        src, force = np.array([[0.45, 0.55, 0.5]]), np.array([[0.0, 0.0, 1.0]])
        t = ctx.start_from_rest(dt, src, force, f0=5.0)
        ctx.step_n(L.PROB_RICKER, 30, t, dt)
        u_eager = ctx.get_slot(_lib.SLOT_U_CUR)
        ctx.start_from_rest(dt, src, force, f0=5.0)
        ctx.capture_solve(L.PROB_RICKER, dt, 30, every=1)
        ctx.replay()
        ctx.sync()
        u_graph = ctx.get_slot(_lib.SLOT_U_CUR)
        ctx.set_symmetric_storage(False)
    nrm = np.linalg.norm(res["fp64"])
    d_full = np.linalg.norm(res["full48"] - res["fp64"]) / nrm
    d_pairs = np.linalg.norm(res["pairs48"] - res["fp64"]) / nrm
    say(f"  drift vs fp64: full rows {d_full:.3e}, pairs once {d_pairs:.3e}")
    assert d_full < 5e-8 and d_pairs < 5e-8, (d_full, d_pairs)
    assert abs(d_pairs - d_full) < 0.05 * d_full, (d_full, d_pairs)
    assert np.abs(u_eager).max() > 0
    assert rel(u_graph, u_eager) < 1e-12, rel(u_graph, u_eager)


def test_48bit_pairs_once_survives_release():
    """With the 48-bit apply on stored pairs, lymph_release_fp64_operator
    frees the FP64 blocks and leaves the step's bits unchanged (the stored
    FP32 blocks and corrections are kept); switching the storage off then
    falls back to the 48-bit full rows, which were kept too."""
    say("[symmetric] 48-bit pairs once after releasing the FP64 blocks")
    m, p = M.cube_mesh(3, agglomerate=True), 3
    rng = np.random.default_rng(21)
    with L.Context(m, p) as ctx:
        _prepared(ctx)
        n = m.n_poly * ctx.n3
        u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
        ctx.set_apply_precision(48)
        full48 = ctx.time_step_matrix_free(0.37, 1e-3, u0, u1)
        ctx.set_symmetric_storage(True, fp32=True)
        a = ctx.time_step_matrix_free(0.37, 1e-3, u0, u1)
        before = ctx.device_bytes
        ctx.release_fp64_operator()
        assert ctx.device_bytes < before
        assert np.array_equal(ctx.time_step_matrix_free(0.37, 1e-3, u0, u1), a)
        ctx.set_symmetric_storage(False)
        assert np.array_equal(ctx.time_step_matrix_free(0.37, 1e-3, u0, u1), full48)
        with pytest.raises(L.LymphError):
            ctx.set_symmetric_storage(True, fp32=True)  # the pair check needs the FP64 blocks


@pytest.mark.parametrize("precision,pairs,tol", [(64, True, 1e-10), (48, True, 1e-8), (48, False, 1e-8),
                                                  (32, True, 1e-4)])
def test_driver_precision_and_pairs_once(precision, pairs, tol):
    """The deck driver's --precision and --pairs-once on the reference deck
    tet_25: the same steps, and L2 and DG errors within `tol` (relative) of
    the plain FP64 run's: 1e-10 for the FP64 pairs once, 1e-8 for the 48-bit
    apply (7e-11 measured), 1e-4 for the 32-bit one (1.7e-6 measured)."""
    from lymph3d_apsc_amd import driver
    say(f"[driver] tet_25 precision {precision} pairs_once {pairs}")
    deck = os.path.join(DECKS, "tet_25")
    ref = driver.run_deck(deck, verbose=False)
    r = driver.run_deck(deck, verbose=False, precision=precision, pairs_once=pairs)
    say(f"  err_l2 {r['err_l2']!r} vs {ref['err_l2']!r}, err_dg {r['err_dg']!r} vs {ref['err_dg']!r}")
    assert r["nsteps"] == ref["nsteps"] and r["t_final"] == ref["t_final"]
    assert abs(r["err_l2"] - ref["err_l2"]) <= tol * ref["err_l2"]
    assert abs(r["err_dg"] - ref["err_dg"]) <= tol * ref["err_dg"]

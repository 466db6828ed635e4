"""GPU parity: every kernel of liblymph_gpu.so against the C oracle on the same
inputs, through the C ABI (lymph3d_apsc_amd.matrix_free.Context).

Tolerances (FP64): operator/RHS/projection/step entries within 1e-13 of the
oracle relative to the block's (vector's) max magnitude -- the kernels keep
the reference's summation order with FMA contraction off, so the remaining
differences are ulp-level (device sin/cos/sqrt); final solution vectors after
full runs within 1e-11 relative (the FP64 path's stated bar); printed L2/DG errors of
the reference decks within 5e-13 (p=1) / 5e-12 (tet_100, 3183 steps at p=3) of
the reference's own known answers."""
import json
import os
import sys

import numpy as np
import pytest

import lymph3d_apsc_amd as L
from lymph3d_apsc_amd import _lib, mesh as M
from oracle import oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DECKS = os.path.join(GOLD, "decks")
KA = json.load(open(os.path.join(GOLD, "known_answers.json")))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if L.device_count() == 0:
        pytest.skip("no HIP device")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    s = np.abs(b).max()
    return 0.0 if s == 0 and np.abs(a).max() == 0 else np.abs(a - b).max() / max(s, 1e-300)


def oracle_for(m, p, alpha=10.0, theta=-1.0):
    kinds = np.zeros(8, np.int32)
    kinds[[2, 3, 4]] = -1
    kinds[[5, 6, 7]] = -2
    space = np.zeros(8, np.int32)
    for t, s in m.tag_space.items():
        space[t] = s
    return oracle.Oracle(m.xyz, m.tet_vert, m.tria_vert, m.tria_tag, kinds, space, p,
                         tet_poly=m.tet_poly, rho=m.rho, lam=m.lam, mu=m.mu, alpha=alpha,
                         theta=theta)


def hetero_mesh(n=3, seed=5):
    m = M.cube_mesh(n, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 12), seed=seed)
    rng = np.random.default_rng(seed)
    m.lam = np.exp(rng.normal(0, 0.5, m.n_tet))
    m.mu = np.exp(rng.normal(0, 0.5, m.n_tet))
    m.rho = rng.uniform(0.8, 1.2, m.n_tet)
    return m


CASES = [
    ("tet_5", 1), ("tet_5", 2), ("tet_5", 3), ("tet_5", 4), ("tet_25", 2), ("tet_100", 3),
    ("cube3", 2), ("cube3agg", 3), ("cube2agg", 4), ("hetero", 2), ("hetero", 3),
]


def make_mesh(name):
    if name.startswith("tet_"):
        return M.read_deck(os.path.join(DECKS, name))
    if name == "cube3":
        return M.cube_mesh(3)
    if name == "cube3agg":
        return M.cube_mesh(3, agglomerate=True)
    if name == "cube2agg":
        return M.cube_mesh(2, agglomerate=True)
    if name == "hetero":
        return hetero_mesh()
    raise KeyError(name)


@pytest.mark.parametrize("name,p", CASES)
def test_operator_rhs_projection(name, p):
    say(f"[operator] {name} p={p}")
    m = make_mesh(name)
    o = oracle_for(m, p)
    o.build_operator()
    with L.Context(m, p) as ctx:
        rp, col = ctx.pattern()
        orp, ocol = o.pattern()
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
        ctx.make_matrices_free()
        K, A, Mf, Mm = ctx.export_operator()
        oK, oA = o.K(), o.Adg()
        oM, oMm, _, _ = o.mass()
        for b in range(K.shape[0]):
            assert rel(K[b], oK[b]) < 1e-13, (b, rel(K[b], oK[b]))
            assert rel(A[b], oA[b]) < 1e-13
        assert rel(Mf, oM) < 1e-13 and rel(Mm, oMm) < 1e-13
        for prob in (L.PROB_NULL, L.PROB_REF_MMS, L.PROB_CUBE_WAVE):
            r = ctx.make_rhs_free(prob)
            orr = o.rhs(prob)
            assert rel(r, orr) < 1e-13, (prob, rel(r, orr))
        # projections: compared in the L2 norm of the functions (the mass
        # matrix norm).  Coefficients alone are not a fair measure: the
        # Legendre basis on a tet's bounding box is far from orthogonal on the
        # tet, so ulp-level differences of the right-hand side (device vs
        # glibc sin/cos) move the solve along poorly conditioned directions.
        for fn in (L.FN_ZERO, L.FN_REF_VEL, L.FN_REF_UEX, L.FN_CUBE_U):
            c = ctx.compute_modal_coefficients_free(fn)
            co = o.project(fn)
            d = (c - co).reshape(m.n_poly, 3, -1)
            cc = co.reshape(m.n_poly, 3, -1)
            num = np.einsum("pim,pmn,pin->", d, oMm, d)
            den = np.einsum("pim,pmn,pin->", cc, oMm, cc)
            assert (den == 0 and num == 0) or np.sqrt(num / den) < 1e-13, (fn, num, den)


@pytest.mark.parametrize("name,p", [("tet_25", 2), ("cube3agg", 3), ("hetero", 2), ("cube2agg", 4),
                                    ("tet_5", 1)])
def test_factored_build_matches_oracle(name, p):
    """The factored operator build (per-basis-pair quadrature sums) gives the
    oracle's K and A_dg to 1e-13 of each block's largest entry and the
    reference-order build's mass blocks bit for bit."""
    say(f"[factored build] {name} p={p}")
    m = make_mesh(name)
    o = oracle_for(m, p)
    o.build_operator()
    oK, oA = o.K(), o.Adg()
    with L.Context(m, p) as ref, L.Context(m, p) as fac:
        ref.make_matrices_free()
        fac.set_build("factored")
        fac.make_matrices_free()
        _, _, Mr, Mmr = ref.export_operator()
        K, A, Mf, Mmf = fac.export_operator()
        for b in range(K.shape[0]):
            assert rel(K[b], oK[b]) < 1e-13, (b, rel(K[b], oK[b]))
            assert rel(A[b], oA[b]) < 1e-13
        assert np.array_equal(Mf, Mr) and np.array_equal(Mmf, Mmr)


@pytest.mark.parametrize("name,p", [("tet_25", 2), ("cube3agg", 3), ("hetero", 2), ("cube2agg", 4)])
def test_steps_match_oracle(name, p):
    say(f"[steps] {name} p={p}")
    m = make_mesh(name)
    o = oracle_for(m, p)
    o.build_operator()
    rng = np.random.default_rng(1)
    n = m.n_poly * 3 * o.Np
    with L.Context(m, p) as ctx:
        ctx.make_matrices_free()
        rs = ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
        rd = ctx.make_rhs_free(L.PROB_REF_MMS, _lib.SLOT_RHS_DYN)
        ors, ord_ = o.rhs(L.PROB_NULL), o.rhs(L.PROB_REF_MMS)
        u0, v0, u1 = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
        dt, g = 1e-3, 0.37
        a = ctx.first_time_step_matrix_free(g, dt, u0, v0)
        b = o.first_step(g, dt, u0, v0, ors, ord_)
        assert rel(a, b) < 1e-13
        a = ctx.time_step_matrix_free(g, dt, u0, u1)
        b = o.step(g, dt, u0, u1, ors, ord_)
        assert rel(a, b) < 1e-13
        # the same step through the reference's column-major layout
        with L.Context(m, p, layout="fortran") as cf:
            cf.make_matrices_free()
            cf.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
            cf.make_rhs_free(L.PROB_REF_MMS, _lib.SLOT_RHS_DYN)
            tof = lambda v: v.reshape(m.n_poly, -1).T.ravel()  # noqa: E731
            af = cf.time_step_matrix_free(g, dt, tof(u0), tof(u1))
            assert np.array_equal(af, tof(a))
        # device-resident loop == repeated host steps
        ctx.set_slot(_lib.SLOT_U_PREV, u0)
        ctx.set_slot(_lib.SLOT_U_CUR, u1)
        t = ctx.step_n(L.PROB_REF_MMS, 5, 0.1, dt)
        up, uc = u0, u1
        tt = 0.1
        for _ in range(5):
            un = o.step(oracle.time_function(oracle.PROB_REF_MMS, tt), dt, up, uc, ors, ord_)
            up, uc = uc, un
            tt = tt + dt
        assert t == tt
        assert rel(ctx.get_slot(_lib.SLOT_U_CUR), uc) < 1e-12
        # errors
        l2 = ctx.compute_error_l2_matrix_free(u0, u1)
        dg = ctx.compute_error_dg_matrix_free(u0, u1)
        assert abs(l2 / o.error_l2_sq(u0, u1) - 1) < 1e-13
        assert abs(dg / o.error_dg_sq(u0, u1) - 1) < 1e-12
        assert ctx.time_steps(L.PROB_REF_MMS, 3, 0.0, dt) > 0


@pytest.mark.parametrize("name", ["tet_5", "tet_25", "tet_125", "tet_100"])
@pytest.mark.parametrize("build", ["reference", "factored"])
def test_reference_known_answers(name, build):
    """Full runs of the reference's decks on the GPU: printed errors vs the
    reference's known answers (5e-13, 5e-12 for the 3183-step p=3 deck; four
    times that with the factored build) and the final vector vs the oracle's
    (1e-11), with either operator build."""
    say(f"[known answer] {name} {build}")
    ka = KA[name]
    m = M.read_deck(os.path.join(DECKS, name))
    with L.Context(m, m.degree) as ctx:
        ctx.set_build(build)
        r = ctx.run(L.PROB_REF_MMS, m.dt, m.stop_time, want_u=True)
    assert r["nsteps"] == ka["steps"]
    # the printed errors are small differences (u - Pi_h u_ex); the factored
    # build regroups the operator's sums, which moves them by a few 1e-12
    tol = 5e-12 if name == "tet_100" else 5e-13
    if build == "factored":
        tol *= 4
    assert abs(r["err_l2"] / ka["err_l2"] - 1) < tol, r
    assert abs(r["err_dg"] / ka["err_dg"] - 1) < tol, r
    o = oracle_for(m, m.degree)
    ro = o.run(oracle.PROB_REF_MMS, m.dt, m.stop_time, want_u=True)
    assert ro["t_final"] == r["t_final"]
    assert rel(r["u"], ro["u"]) < 1e-11, rel(r["u"], ro["u"])


def test_largest_reference_deck():
    """tet_625, the largest deck the reference ships (28,606 tets, p = 1,
    1000 steps; no printed answer): the GPU run against the oracle's final
    vector (1e-11) and errors (1e-11), in both operator builds, and the face
    pairs stored once against the full rows."""
    say("[tet_625]")
    m = M.read_deck(os.path.join(DECKS, "tet_625"))
    ro = oracle_for(m, m.degree).run(oracle.PROB_REF_MMS, m.dt, m.stop_time, want_u=True)
    assert ro["nsteps"] == 1000
    for build in ("reference", "factored"):
        with L.Context(m, m.degree) as ctx:
            ctx.set_build(build)
            r = ctx.run(L.PROB_REF_MMS, m.dt, m.stop_time, want_u=True)
            if build == "factored":
                ctx.set_symmetric_storage(True)
                r2 = ctx.run(L.PROB_REF_MMS, m.dt, m.stop_time, want_u=True)
                assert rel(r2["u"], r["u"]) < 1e-11
        assert r["t_final"] == ro["t_final"] and r["nsteps"] == 1000
        assert rel(r["u"], ro["u"]) < 1e-11, (build, rel(r["u"], ro["u"]))
        assert abs(r["err_l2"] / ro["err_l2"] - 1) < 1e-11
        assert abs(r["err_dg"] / ro["err_dg"] - 1) < 1e-11


@pytest.mark.parametrize("n,p,steps,dt", [(4, 2, 300, None), (3, 3, 200, None), (2, 4, 100, None),
                                         (5, 2, 1000, 1e-4)])
def test_agglomerated_cube_wave_run(n, p, steps, dt):
    """Agglomerated polyhedra (one per cube), the standing-wave data set, a
    CFL-derived dt (or the given one: 5^3 cubes, p=2, dt=1e-4, 1000 steps is
    the CPU-runnable configuration of the plain method): final vector and
    errors vs the oracle."""
    say(f"[cube wave] n={n} p={p} steps={steps}")
    m = M.cube_mesh(n, agglomerate=True)
    dt = dt or 0.2 * (1.0 / n) / (np.sqrt(3.0) * p * p)
    with L.Context(m, p) as ctx:
        r = ctx.run(L.PROB_CUBE_WAVE, dt, 1e9, max_steps=steps, want_u=True)
    ro = oracle_for(m, p).run(oracle.PROB_CUBE_WAVE, dt, 1e9, max_steps=steps, want_u=True)
    assert r["nsteps"] == steps == ro["nsteps"]
    assert rel(r["u"], ro["u"]) < 1e-11
    assert abs(r["err_l2"] / ro["err_l2"] - 1) < 1e-9
    assert abs(r["err_dg"] / ro["err_dg"] - 1) < 1e-9
    assert np.isfinite(r["err_l2"]) and r["err_l2"] < 0.5


@pytest.mark.parametrize("theta,alpha", [(0.0, 10.0), (1.0, 10.0), (-1.0, 3.0)])
@pytest.mark.parametrize("build", ["reference", "factored"])
def test_penalty_variants_match_oracle(theta, alpha, build):
    """IIP (theta = 0) and NIPG (theta = 1) face terms and another penalty
    (alpha = 3): operator blocks, RHS and a step against the oracle built with
    the same theta and alpha, for both operator builds."""
    say(f"[penalty] theta={theta} alpha={alpha} {build}")
    m = make_mesh("hetero")
    p = 2
    o = oracle_for(m, p, alpha=alpha, theta=theta)
    o.build_operator()
    oK, oA = o.K(), o.Adg()
    rng = np.random.default_rng(9)
    n = m.n_poly * 3 * o.Np
    u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
    with L.Context(m, p, alpha=alpha, theta=theta) as ctx:
        ctx.set_build(build)
        ctx.make_matrices_free()
        K, A, _, _ = ctx.export_operator()
        for b in range(K.shape[0]):
            assert rel(K[b], oK[b]) < 1e-13, (b, rel(K[b], oK[b]))
            assert rel(A[b], oA[b]) < 1e-13
        rs = ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
        rd = ctx.make_rhs_free(L.PROB_REF_MMS, _lib.SLOT_RHS_DYN)
        assert rel(rd, o.rhs(L.PROB_REF_MMS)) < 1e-13
        a = ctx.time_step_matrix_free(0.37, 1e-3, u0, u1)
        b = o.step(0.37, 1e-3, u0, u1, o.rhs(L.PROB_NULL), o.rhs(L.PROB_REF_MMS))
        assert rel(a, b) < 1e-13


def _single_tet():
    xyz = np.array([[0.1, 0.2, 0.0], [1.0, 0.0, 0.1], [0.2, 0.9, 0.0], [0.3, 0.3, 1.1]])
    tv = np.array([[0, 1, 2, 3]])
    tri = np.sort(tv[0][M.FACE_SETS], axis=1)
    return M.Mesh(xyz, tv, None, tri, np.full(4, 2), {2: 1})


@pytest.mark.parametrize("case,p", [("single_tet", 1), ("single_tet", 4), ("all_neumann", 2)])
def test_edge_meshes_match_oracle(case, p):
    """Edge cases of the mesh: one tetrahedron with every face on the
    (Dirichlet) boundary, and a cube with no Dirichlet face at all (Neumann
    everywhere): operator, RHS and a step against the oracle; time loops
    whose stop time ends before the second step."""
    say(f"[edge mesh] {case} p={p}")
    m = _single_tet() if case == "single_tet" else M.cube_mesh(2, dirichlet_faces=set(), agglomerate=True)
    o = oracle_for(m, p)
    o.build_operator()
    rng = np.random.default_rng(3)
    n = m.n_poly * 3 * o.Np
    u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
    with L.Context(m, p) as ctx:
        ctx.make_matrices_free()
        K, A, _, _ = ctx.export_operator()
        oK, oA = o.K(), o.Adg()
        for b in range(K.shape[0]):
            assert rel(K[b], oK[b]) < 1e-13 and rel(A[b], oA[b]) < 1e-13
        ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
        rd = ctx.make_rhs_free(L.PROB_REF_MMS, _lib.SLOT_RHS_DYN)
        assert rel(rd, o.rhs(L.PROB_REF_MMS)) < 1e-13
        a = ctx.time_step_matrix_free(0.37, 1e-3, u0, u1)
        b = o.step(0.37, 1e-3, u0, u1, o.rhs(L.PROB_NULL), o.rhs(L.PROB_REF_MMS))
        assert rel(a, b) < 1e-13
        # stop time before the second step: only the first step runs
        r = ctx.run(L.PROB_REF_MMS, 1e-3, 5e-4)
        assert r["nsteps"] == 1 and r["t_final"] == 1e-3
        r = ctx.run(L.PROB_REF_MMS, 1e-3, 1.0, max_steps=1)
        assert r["nsteps"] == 1


def _mms_data():
    """The reference's manufactured data set (problem_data_and_properties.f90:
    f_time, gd, gn) written as the vectorised functions a caller would pass."""
    pi = np.pi

    def forcing(lam, mu, rho, x):
        c = (-2.0 * rho + 3.0 * (lam + 2.0 * mu)) * pi * pi
        sx, sy, sz = np.sin(pi * x[:, 0]), np.sin(pi * x[:, 1]), np.sin(pi * x[:, 2])
        cx, cy, cz = np.cos(pi * x[:, 0]), np.cos(pi * x[:, 1]), np.cos(pi * x[:, 2])
        return np.stack([c * cx * sy * sz, c * sx * cy * sz, c * sx * sy * cz], axis=1)

    def dirichlet(x, tag):
        sx, sy, sz = np.sin(pi * x[:, 0]), np.sin(pi * x[:, 1]), np.sin(pi * x[:, 2])
        cx, cy, cz = np.cos(pi * x[:, 0]), np.cos(pi * x[:, 1]), np.cos(pi * x[:, 2])
        on = ((tag >= 1) & (tag <= 3))[:, None]
        return np.where(on, np.stack([cx * sy * sz, sx * cy * sz, sx * sy * cz], axis=1), 0.0)

    def neumann(lam, mu, n, x, tag):
        sx, sy, sz = np.sin(pi * x[:, 0]), np.sin(pi * x[:, 1]), np.sin(pi * x[:, 2])
        cx, cy, cz = np.cos(pi * x[:, 0]), np.cos(pi * x[:, 1]), np.cos(pi * x[:, 2])
        d = -pi * sx * sy * sz * (3.0 * lam + 2.0 * mu)
        s01 = 2.0 * mu * pi * cx * cy * sz
        s02 = 2.0 * mu * pi * cx * cz * sy
        s12 = 2.0 * mu * pi * cy * cz * sx
        r = np.stack([d * n[:, 0] + s01 * n[:, 1] + s02 * n[:, 2],
                      s01 * n[:, 0] + d * n[:, 1] + s12 * n[:, 2],
                      s02 * n[:, 0] + s12 * n[:, 1] + d * n[:, 2]], axis=1)
        return np.where(((tag >= 1) & (tag <= 3))[:, None], r, 0.0)

    return forcing, dirichlet, neumann


@pytest.mark.parametrize("name,p", [("hetero", 2), ("tet_25", 1), ("cube2agg", 3)])
def test_rhs_from_caller_functions(name, p):
    """MAKE_RHS_FREE and COMPUTE_MODAL_COEFFICIENTS_FREE with the data passed
    as functions (the drop-ins for their procedure arguments) equal the
    built-in data set they restate, on meshes with Dirichlet and Neumann
    faces and per-tet materials; no functions at all gives a zero RHS."""
    say(f"[rhs functions] {name} p={p}")
    m = make_mesh(name)
    f, gd, gn = _mms_data()
    with L.Context(m, p) as ctx:
        ref = ctx.make_rhs_free(L.PROB_REF_MMS)
        got = ctx.make_rhs_free_fn(f, gd, gn)
        assert rel(got, ref) < 1e-13, rel(got, ref)
        assert not np.any(ctx.make_rhs_free_fn())
        # the slot written is the one the stepping uses
        assert np.array_equal(ctx.get_slot(_lib.SLOT_RHS_DYN), np.zeros_like(ref))
        # COMPUTE_MODAL_COEFFICIENTS_FREE of a caller function: uex of the same data set
        ctx.make_matrices_free()
        pi = np.pi

        def uex(x):
            sx, sy, sz = np.sin(pi * x[:, 0]), np.sin(pi * x[:, 1]), np.sin(pi * x[:, 2])
            cx, cy, cz = np.cos(pi * x[:, 0]), np.cos(pi * x[:, 1]), np.cos(pi * x[:, 2])
            return np.stack([cx * sy * sz, sx * cy * sz, sx * sy * cz], axis=1)
        pref = ctx.compute_modal_coefficients_free(L.FN_REF_UEX)
        pgot = ctx.compute_modal_coefficients_free_fn(uex)
        # compared in the L2 norm of the functions (the bounding-box basis makes
        # coefficients a poor measure of ulp-level data differences)
        d2 = ctx.compute_error_l2_matrix_free(pgot, pref)
        n2 = ctx.compute_error_l2_matrix_free(pref, np.zeros_like(pref))
        assert np.sqrt(d2 / n2) < 1e-13, np.sqrt(d2 / n2)


def test_singular_mass_reports_exit_code():
    m = M.cube_mesh(1)
    m.rho = -np.ones(m.n_tet)  # host-forced: M = rho * mass is negative definite
    with L.Context(m, 2) as ctx:
        with pytest.raises(L.LymphError) as e:
            ctx.make_matrices_free()
        assert e.value.code == _lib.EXIT_SINGULARMTX
        with pytest.raises(L.LymphError) as e:
            ctx.time_step_matrix_free(0.0, 1e-3, ctx.vec(), ctx.vec())
        assert e.value.code == _lib.ERR_STATE


def test_instability_control():
    """A time step far beyond the CFL limit (forced from the host) makes the
    leapfrog iteration overflow; lymph_run stops with the reference's
    EXIT_INSTAB at its next check instead of running on with NaN.  The slot
    check itself sees a single NaN."""
    say("[instability]")
    m = make_mesh("tet_5")
    with L.Context(m, 1) as ctx:
        with pytest.raises(L.LymphError) as e:
            ctx.run(L.PROB_REF_MMS, 5.0, 1e9, max_steps=2000)
        assert e.value.code == _lib.EXIT_INSTAB
        u = np.ones(ctx.n_poly * ctx.n3)
        ctx.set_slot(_lib.SLOT_U_CUR, u)
        assert ctx.check_finite(_lib.SLOT_U_CUR)
        u[len(u) // 2] = np.nan
        ctx.set_slot(_lib.SLOT_U_CUR, u)
        assert not ctx.check_finite(_lib.SLOT_U_CUR)
        u[len(u) // 2] = -np.inf
        ctx.set_slot(_lib.SLOT_U_CUR, u)
        assert not ctx.check_finite(_lib.SLOT_U_CUR)
        # a stable run passes every check
        r = ctx.run(L.PROB_REF_MMS, m.dt, m.stop_time, max_steps=600)
        assert r["nsteps"] == 600 and np.isfinite(r["err_l2"])


def test_resume_from_slots():
    """Checkpoint/resume: the state after k steps (U_PREV, U_CUR and t) read
    out with get_slot and loaded into a fresh context continues the run bit
    for bit."""
    say("[resume]")
    m = make_mesh("cube3agg")
    p, dt = 3, 2e-3
    with L.Context(m, p) as a:
        a.make_matrices_free()
        a.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
        a.make_rhs_free(L.PROB_CUBE_WAVE, _lib.SLOT_RHS_DYN)
        u0 = a.compute_modal_coefficients_free(L.FN_CUBE_U, _lib.SLOT_U_PREV)
        a.set_slot(_lib.SLOT_U_CUR, a.first_time_step_matrix_free(1.0, dt, u0, np.zeros_like(u0)))
        t = a.step_n(L.PROB_CUBE_WAVE, 40, dt, dt)
        ck = (a.get_slot(_lib.SLOT_U_PREV), a.get_slot(_lib.SLOT_U_CUR), t)
        t_end = a.step_n(L.PROB_CUBE_WAVE, 60, t, dt)
        full = a.get_slot(_lib.SLOT_U_CUR)
    with L.Context(m, p) as b:
        b.make_matrices_free()
        b.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
        b.make_rhs_free(L.PROB_CUBE_WAVE, _lib.SLOT_RHS_DYN)
        b.set_slot(_lib.SLOT_U_PREV, ck[0])
        b.set_slot(_lib.SLOT_U_CUR, ck[1])
        assert b.step_n(L.PROB_CUBE_WAVE, 60, ck[2], dt) == t_end
        assert np.array_equal(b.get_slot(_lib.SLOT_U_CUR), full)


@pytest.mark.parametrize("p", [2, 3])
def test_receivers_and_traces(p):
    """Receiver values (EXPORT_SOLUTION's point evaluation) of the GPU state vs
    the oracle's evaluation of the same vector; recorded traces equal the
    samples taken step by step."""
    say(f"[receivers] p={p}")
    m = M.cube_mesh(4, agglomerate=True)
    rng = np.random.default_rng(3)
    pts = rng.uniform(0.05, 0.95, (64, 3))
    o = oracle_for(m, p)
    dt = 0.2 * 0.25 / (np.sqrt(3.0) * p * p)
    with L.Context(m, p) as ctx:
        r = ctx.run(L.PROB_CUBE_WAVE, dt, 1e9, max_steps=10, want_u=True)
        tets = ctx.set_receivers(pts, max_records=5)
        assert np.array_equal(tets, o.locate(pts))
        got = ctx.sample_receivers(_lib.SLOT_U_CUR)
        ref = o.eval_points(pts, tets, r["u"])
        assert rel(got, ref) < 1e-13
        t = ctx.step_n_record(L.PROB_CUBE_WAVE, 10, r["t_final"], dt, every=2)
        tr = ctx.traces()
        assert tr.shape == (5, 64, 3)
        assert np.array_equal(tr[-1], ctx.sample_receivers(_lib.SLOT_U_CUR))
        assert t > r["t_final"]


@pytest.mark.parametrize("name,p", [("cube3agg", 3), ("hetero", 2), ("cube2agg", 4)])
def test_fp32_apply(name, p):
    """FP32 operator apply: the step's acceleration term dt^2 M^-1(F - Ku)
    within 1e-5 of the FP64 oracle (FP32 rounding of K and of the per-block
    sums), the update itself in FP64."""
    say(f"[fp32 apply] {name} p={p}")
    m = make_mesh(name)
    o = oracle_for(m, p)
    o.build_operator()
    rng = np.random.default_rng(2)
    n = m.n_poly * 3 * o.Np
    with L.Context(m, p) as ctx:
        ctx.make_matrices_free()
        ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
        ctx.make_rhs_free(L.PROB_REF_MMS, _lib.SLOT_RHS_DYN)
        ctx.set_apply_precision(32)
        u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
        dt, g = 1e-3, 0.37
        a = ctx.time_step_matrix_free(g, dt, u0, u1)
        b = o.step(g, dt, u0, u1, o.rhs(L.PROB_NULL), o.rhs(L.PROB_REF_MMS))
        acc_gpu, acc_ref = a - (2 * u1 - u0), b - (2 * u1 - u0)
        assert rel(acc_gpu, acc_ref) < 1e-5
        assert rel(acc_gpu, acc_ref) > 0  # really the FP32 path
        ctx.set_apply_precision(64)
        a64 = ctx.time_step_matrix_free(g, dt, u0, u1)
        assert rel(a64, b) < 1e-13


def test_fp32_operator_beyond_2_32_entries():
    """An operator of more than 2^32 entries (an ensemble-sized context) is
    converted to FP32 in full: every polyhedron's FP32 step stays within
    1e-4 of its FP64 step, the last ones included."""
    say("[fp32 > 2^32 entries]")
    m = M.cube_mesh(56, agglomerate=True)
    p = 3
    rng = np.random.default_rng(8)
    with L.Context(m, p) as ctx:
        assert ctx.n_blk * ctx.n3 * ctx.n3 > 2 ** 32
        ctx.set_keep_adg(False)
        ctx.make_matrices_free()
        ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
        ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_DYN)
        n = m.n_poly * ctx.n3
        u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
        dt, g = 1e-3, 0.0
        a64 = ctx.time_step_matrix_free(g, dt, u0, u1) - (2 * u1 - u0)
        ctx.set_apply_precision(32)
        a32 = ctx.time_step_matrix_free(g, dt, u0, u1) - (2 * u1 - u0)
    d = np.abs(a32 - a64).reshape(m.n_poly, -1).max(axis=1)
    s = np.abs(a64).reshape(m.n_poly, -1).max(axis=1)
    assert (d <= 1e-4 * s).all(), int(np.argmax(d / s))


def test_fp32_apply_on_ill_conditioned_agglomerates():
    """Config 3's shape -- 30^3 cubes x 6 tets grown into polyhedra of 4-40
    tets, p=4, Neumann top, Ricker source, 2000 steps: the bounding-box
    Legendre basis makes the mass blocks ill conditioned (cond up to ~1e8),
    yet the 32-bit apply (in the mass-orthonormal basis) stays within 1e-3 of
    the FP64 receiver traces instead of diverging."""
    say("[fp32 ill-conditioned p=4]")
    from lymph3d_apsc_amd import ensemble as E
    m = M.cube_mesh(30, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 40), seed=3)
    rng = np.random.default_rng(3)
    bc = m.xyz[m.tet_vert].mean(axis=1)
    m.lam, m.mu = E.lognormal_field(rng, bc), E.lognormal_field(rng, bc)
    m.rho = rng.uniform(0.8, 1.2, m.n_tet)
    p, steps = 4, 2000
    dt = 0.2 * (1 / 30) / (np.sqrt(((m.lam + 2 * m.mu) / m.rho).max()) * p * p)
    rec = rng.uniform(0.05, 0.95, (32, 3))
    traces = []
    with L.Context(m, p) as ctx:
        ctx.set_keep_adg(False)
        ctx.make_matrices_free()
        for bits in (64, 32):
            ctx.set_apply_precision(bits)
            t = ctx.start_from_rest(dt, [[0.5, 0.45, 0.55]], [[0.0, 0.0, 1.0]], f0=5.0)
            ctx.set_receivers(rec, max_records=steps // 10)
            ctx.step_n_record(L.PROB_RICKER, steps, t, dt, every=10)
            traces.append(ctx.traces())
    a64, a32 = traces
    assert np.isfinite(a32).all() and np.abs(a64).max() > 0
    assert np.linalg.norm(a32 - a64) < 1e-3 * np.linalg.norm(a64)


@pytest.mark.parametrize("name,p", [("hetero", 1), ("hetero", 2), ("hetero", 3), ("grown4", 4)])
def test_step_variants_agree_bitwise(name, p):
    """Every step-kernel variant -- load group size, neighbour broadcast mode,
    waves per polyhedron, several polyhedra per wave at p <= 2 -- gives bit
    for bit the same step and first step, for the 64-bit apply (which matches
    the oracle) and for the 32-bit apply."""
    say(f"[variants] {name} p={p}")
    if name == "grown4":
        m = M.cube_mesh(4, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 40), seed=3)
    else:
        m = make_mesh(name)
    o = oracle_for(m, p)
    o.build_operator()
    rng = np.random.default_rng(6)
    n = m.n_poly * 3 * o.Np
    u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
    dt, g = 1e-3, 0.37
    ref = o.step(g, dt, u0, u1, o.rhs(L.PROB_NULL), o.rhs(L.PROB_REF_MMS))
    with L.Context(m, p) as ctx:
        ctx.make_matrices_free()
        ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
        ctx.make_rhs_free(L.PROB_REF_MMS, _lib.SLOT_RHS_DYN)
        for bits in (64, 32):
            ctx.set_apply_precision(bits)
            outs, firsts = [], []
            for v in range(ctx.n_step_variants()):
                ctx.set_step_variant(v)
                outs.append(ctx.time_step_matrix_free(g, dt, u0, u1))
                firsts.append(ctx.first_time_step_matrix_free(g, dt, u0, u1))
            for v in range(1, len(outs)):
                assert np.array_equal(outs[v], outs[0]), (bits, v)
                assert np.array_equal(firsts[v], firsts[0]), (bits, v)
            if bits == 64:
                assert rel(outs[0], ref) < 1e-13


def test_memory_options_keep_results():
    """Dropping A_dg and releasing the FP64 blocks changes no result: the
    operator built without A_dg, the 32-bit apply after the release and a
    rebuild all give bit for bit the steps of a context that keeps everything;
    what was dropped reports a state error."""
    say("[memory options]")
    m = make_mesh("hetero")
    p = 2
    rng = np.random.default_rng(4)
    with L.Context(m, p) as full, L.Context(m, p) as lean:
        n = m.n_poly * full.n3
        u0, u1 = rng.standard_normal(n), rng.standard_normal(n)
        dt, g = 1e-3, 0.37
        full.make_matrices_free()
        lean.set_keep_adg(False)
        lean.make_matrices_free()
        assert lean.device_bytes < full.device_bytes
        for c in (full, lean):
            c.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
            c.make_rhs_free(L.PROB_REF_MMS, _lib.SLOT_RHS_DYN)
        a64 = full.time_step_matrix_free(g, dt, u0, u1)
        assert np.array_equal(lean.time_step_matrix_free(g, dt, u0, u1), a64)
        with pytest.raises(L.LymphError):
            lean.compute_error_dg_matrix_free(u0, u1)
        full.set_apply_precision(32)
        lean.set_apply_precision(32)
        before = lean.device_bytes
        lean.release_fp64_operator()
        assert lean.device_bytes < before
        a32 = full.time_step_matrix_free(g, dt, u0, u1)
        assert np.array_equal(lean.time_step_matrix_free(g, dt, u0, u1), a32)
        with pytest.raises(L.LymphError):
            lean.set_apply_precision(64)
        with pytest.raises(L.LymphError):
            lean.export_operator()
        lean.make_matrices_free()  # rebuild: FP64 blocks back, 32-bit copy refreshed
        assert np.array_equal(lean.time_step_matrix_free(g, dt, u0, u1), a32)
        lean.set_apply_precision(64)
        assert np.array_equal(lean.time_step_matrix_free(g, dt, u0, u1), a64)


def test_vertex_export_matches_oracle(tmp_path):
    """Vertex values for the VTK export (receiver kernel at tet vertices) vs
    the oracle's point evaluation of the same vector."""
    say("[export]")
    from lymph3d_apsc_amd import export
    m = M.cube_mesh(3, agglomerate=True)
    p = 2
    o = oracle_for(m, p)
    with L.Context(m, p) as ctx:
        r = ctx.run(L.PROB_CUBE_WAVE, 0.01, 1e9, max_steps=3, want_u=True)
        vals = export.vertex_values(ctx)
    pts = m.xyz[m.tet_vert].reshape(-1, 3)
    ref = o.eval_points(pts, np.repeat(np.arange(m.n_tet, dtype=np.int32), 4), r["u"])
    assert rel(vals.reshape(-1, 3), ref) < 1e-13
    export.write_vtu(str(tmp_path / "u.vtu"), m, vals)


@pytest.mark.parametrize("p,steps", [(3, 60), (4, 40)])
def test_ricker_point_source_heterogeneous(p, steps):
    """Config-3 shaped case at small size: uneven region-grown polyhedra,
    lognormal lambda/mu, rho in [0.8,1.2], traction-free z=1 face, zero
    initial state, Ricker point force (f0=5) at an interior point.  GPU run vs
    the oracle's step loop; receiver traces vs the oracle's evaluation."""
    say(f"[ricker] p={p}")
    m = M.cube_mesh(3, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 12), seed=9)
    rng = np.random.default_rng(9)
    m.lam = np.exp(rng.normal(0, 0.5, m.n_tet))
    m.mu = np.exp(rng.normal(0, 0.5, m.n_tet))
    m.rho = rng.uniform(0.8, 1.2, m.n_tet)
    src = np.array([[0.41, 0.53, 0.47]])
    force = np.array([[0.0, 0.0, 1.0]])
    rec = rng.uniform(0.1, 0.9, (16, 3))
    dt = 0.2 * (1 / 3) / (np.sqrt(3.0) * p * p) / 2
    f0 = 5.0
    o = oracle_for(m, p)
    o.build_operator()
    stet = o.locate(src)
    rs = o.rhs(oracle.PROB_NULL)
    rd = o.rhs(oracle.PROB_NULL) + o.point_sources(src, stet, force)
    z = np.zeros(o.n_poly * o.n3)
    t = 0.0
    un = o.first_step(oracle.ricker(f0, 1.2 / f0, t), dt, z, z, rs, rd)
    up = z
    t = t + dt
    for _ in range(steps):
        up, un = un, o.step(oracle.ricker(f0, 1.2 / f0, t), dt, up, un, rs, rd)
        t = t + dt
    with L.Context(m, p) as ctx:
        t_gpu = ctx.start_from_rest(dt, src, force, f0=f0)
        assert rel(ctx.get_slot(_lib.SLOT_RHS_DYN), rd) < 1e-13
        ctx.set_receivers(rec, max_records=steps)
        t_gpu = ctx.step_n_record(L.PROB_RICKER, steps, t_gpu, dt, every=1)
        u = ctx.get_slot(_lib.SLOT_U_CUR)
        tr = ctx.traces()
    assert t_gpu == t
    assert np.abs(un).max() > 0
    assert rel(u, un) < 1e-11
    assert rel(tr[-1], o.eval_points(rec, o.locate(rec), un)) < 1e-11


def test_ensemble_union_equals_standalone_runs():
    """An ensemble (instances of mixed size, degree and wavelet frequency
    joined by degree into one block-diagonal context, whole solve captured in
    a HIP graph) gives bit for bit the traces of each instance run alone the
    same way, and within 1e-12 those of the plain uncaptured path (wavelet
    evaluated on the host instead of the device); two replays agree."""
    say("[ensemble]")
    from lymph3d_apsc_amd import ensemble as E
    inst = [E.make_instance(s, n_range=(2, 4), n_receivers=8, f0=4.0 + s) for s in range(5)]
    assert len({i["p"] for i in inst}) == 2
    dt, nsteps, every = 2e-3, 30, 3
    ens = E.Ensemble(inst, dt, nsteps, every=every, apply_bits=64, autotune=False)
    tr1 = ens.run()
    tr2 = ens.run()
    for a, b in zip(tr1, tr2):
        assert np.array_equal(a, b)
    for k, ins in enumerate(inst):
        alone = E.Ensemble([ins], dt, nsteps, every=every, apply_bits=64, autotune=False)
        assert np.array_equal(tr1[k], alone.run()[0]), k
        alone.close()
        with L.Context(ins["mesh"], ins["p"]) as ctx:  # plain path, host-evaluated wavelet
            t = ctx.start_from_rest(dt, ins["source"], ins["force"], f0=ins["f0"])
            ctx.set_receivers(ins["receivers"], max_records=nsteps // every)
            ctx.step_n_record(L.PROB_RICKER, nsteps, t, dt, every=every)
            ref = ctx.traces()
        assert np.abs(ref).max() > 0
        assert rel(tr1[k], ref) < 1e-12, k
    ens.close()
    # FP32 apply ensemble stays close to the FP64 traces
    ens32 = E.Ensemble(inst, dt, nsteps, every=every, apply_bits=32, autotune=False)
    tr32 = ens32.run()
    for a, b in zip(tr32, tr1):
        assert rel(a, b) < 1e-4
    ens32.close()


def test_captured_solve_dropped_when_its_buffers_change():
    """A captured solve holds the receivers' and source groups' buffers and
    the wavelet's factors: replacing any of them drops it, and lymph_replay
    then returns LYMPH_ERR_STATE instead of touching freed memory.  After an
    uncaptured step has rotated the state slots, a replay still ends with
    the captured solve's state in U_CUR."""
    say("[capture] dropped on buffer changes")
    m, p = M.cube_mesh(2, agglomerate=True), 2
    dt, nsteps = 2e-3, 12
    rec = np.array([[0.3, 0.4, 0.5], [0.6, 0.2, 0.7]])
    src, force = np.array([[0.45, 0.55, 0.5]]), np.array([[0.0, 0.0, 1.0]])

    def capture(ctx):
        ctx.set_receivers(rec, max_records=nsteps)
        ctx.capture_solve(L.PROB_RICKER, dt, nsteps, every=1)

    with L.Context(m, p) as ctx:
        ctx.start_from_rest(dt, src, force, f0=5.0)
        capture(ctx)
        ctx.replay()
        ctx.sync()
        u_ref = ctx.get_slot(_lib.SLOT_U_CUR)
        tr_ref = ctx.traces()
        assert np.abs(tr_ref).max() > 0
        ctx.step_n(L.PROB_RICKER, 5, 1.0, dt)  # uncaptured steps rotate the slots
        ctx.replay()
        ctx.sync()
        assert np.array_equal(ctx.get_slot(_lib.SLOT_U_CUR), u_ref)
        assert np.array_equal(ctx.traces(), tr_ref)
        changes = [lambda: ctx.set_receivers(rec[:1], max_records=nsteps),
                   lambda: ctx.set_source_groups(np.zeros(m.n_poly, np.int32), [4.0]),
                   lambda: ctx.set_source_groups(None),
                   lambda: ctx.set_ricker(6.0)]
        for change in changes:
            capture(ctx)
            change()
            with pytest.raises(L.LymphError) as e:
                ctx.replay()
            assert e.value.code == _lib.ERR_STATE
        # setting the same wavelet again keeps the captured solve
        ctx.set_ricker(5.0)
        capture(ctx)
        ctx.set_ricker(5.0)
        ctx.replay()
        ctx.sync()


@pytest.mark.parametrize("name,p", [("grown", 2), ("grown", 3), ("tet_25", 1), ("cube3agg", 4)])
def test_double_couple_sources_match_oracle(name, p):
    """The SISM branch of MAKE_RHS_FREE: double couples (several sources, two
    of them in one polyhedron) added to a RHS slot on the device against the
    oracle's o_double_couples (1e-13), then 30 leapfrog steps from rest with a
    Ricker time factor against the oracle's step loop (1e-11)."""
    say(f"[double couple] {name} p={p}")
    if name == "grown":
        m = M.cube_mesh(3, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 12), seed=9)
    elif name.startswith("tet_"):
        m = M.read_deck(os.path.join(DECKS, name))
    else:
        m = M.cube_mesh(3, agglomerate=True)
    rng = np.random.default_rng(21 + p)
    pts = rng.uniform(0.2, 0.8, (4, 3))
    o = oracle_for(m, p)
    o.build_operator()
    tets = o.locate(pts)
    same = np.flatnonzero(m.tet_poly == m.tet_poly[tets[0]]) if m.tet_poly is not None else [tets[0]]
    tets = np.append(tets, np.int32(same[-1]))  # a second source in the first one's polyhedron
    strike, normal = rng.standard_normal((5, 3)), rng.standard_normal((5, 3))
    m0 = 10 ** rng.uniform(-1, 1, 5)
    rs = o.rhs(oracle.PROB_NULL)
    rd = o.rhs(oracle.PROB_NULL) + o.double_couples(tets, strike, normal, m0)
    f0, steps = 5.0, 30
    dt = 0.2 * o.hmax() / (np.sqrt(3.0) * p * p) / 4
    z = np.zeros(o.n_poly * o.n3)
    t = 0.0
    un = o.first_step(oracle.ricker(f0, 1.2 / f0, t), dt, z, z, rs, rd)
    up = z
    t = t + dt
    for _ in range(steps):
        up, un = un, o.step(oracle.ricker(f0, 1.2 / f0, t), dt, up, un, rs, rd)
        t = t + dt
    with L.Context(m, p) as ctx:
        ctx.make_matrices_free()
        ctx.set_ricker(f0)
        ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
        ctx.make_rhs_free(L.PROB_RICKER, _lib.SLOT_RHS_DYN)
        ctx.add_double_couples(None, strike, normal, m0, tets=tets.astype(np.int32))
        assert rel(ctx.get_slot(_lib.SLOT_RHS_DYN), rd) < 1e-13
        u1 = ctx.first_time_step_matrix_free(L.ricker(f0, 1.2 / f0, 0.0), dt, z, z)
        ctx.set_slot(_lib.SLOT_U_PREV, z)
        ctx.set_slot(_lib.SLOT_U_CUR, u1)
        t_gpu = ctx.step_n(L.PROB_RICKER, steps, 0.0 + dt, dt)
        u = ctx.get_slot(_lib.SLOT_U_CUR)
    assert t_gpu == t
    assert np.abs(un).max() > 0
    assert rel(u, un) < 1e-11, rel(u, un)

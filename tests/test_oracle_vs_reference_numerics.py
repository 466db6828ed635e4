"""Pin the C oracle against the reference's own Fortran element routines
(src/basis_function.f90, src/Poly_ref_mappings.f90, src/assemble_element.f90,
src/vet_mat_operations.f90, src/problem_data_and_properties.f90), compiled
from the reference sources into oracle/_ref/ by `make -C oracle ref`.

Inputs are real tetrahedra of the reference's shipped decks, with the
polyhedron bounding boxes, diameters, normals and areas the oracle computes
(those are pinned separately by the decks' known answers).  Skips where the
reference tree (and so oracle/_ref) is absent, e.g. on the GPU machine.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import deck, oracle
from tests import refnum

pytestmark = pytest.mark.skipif(not refnum.available(), reason="oracle/_ref not built")

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    s = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / s


def _tet(ctx_deck, t):
    xyz = ctx_deck["xyz"][ctx_deck["tet_vert"][t]]
    return xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()


def _bbox_f(bb):
    # oracle bbox (xmin,xmax,ymin,ymax,zmin,zmax) -> Fortran b_box(3,2)
    return np.asfortranarray(np.array(bb).reshape(3, 2))


@pytest.mark.parametrize("p", range(0, 7))
def test_quadrature_and_basis_list(p):
    L = oracle.lib()
    n3, n2, Np = (p + 1) ** 3, (p + 1) ** 2, oracle.lib().o_np(p)
    t3, w3, t2, w2 = np.zeros(4 * n3), np.zeros(n3), np.zeros(4 * n2), np.zeros(n2)
    L.o_quadrature(p, t3, w3, t2, w2)
    rt3, rw3, rt2, rw2 = refnum.quadrature(p)
    assert np.array_equal(t3.reshape(n3, 4), rt3.T)
    assert np.array_equal(w3, rw3)
    assert np.array_equal(t2.reshape(n2, 4), rt2.T)
    assert np.array_equal(w2, rw2)
    if p >= 1:  # (p+1)-point rules integrate the Duffy Jacobians exactly
        assert abs(w3.sum() - 1 / 6) < 1e-15 and abs(w2.sum() - 0.5) < 1e-15
    bl = np.zeros(3 * Np, np.int32)
    L.o_basis_list(p, bl)
    assert np.array_equal(bl.reshape(Np, 3), refnum.basis_list(p))


@pytest.mark.parametrize("name,tets", [("tet_5", [0, 7, 29]), ("tet_100", [0, 500, 1355])])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_element_matrices(name, tets, p):
    d = deck.read_deck(os.path.join(DECKS, name))
    o = oracle.from_deck(d, p=p)
    cn = o.conn()
    L = oracle.lib()
    Np, n3q, n2q = o.Np, (p + 1) ** 3, (p + 1) ** 2
    lam, mu, rho = 1.7, 0.6, 1.3
    for t in tets:
        x, y, z = _tet(d, t)
        fk = np.zeros(12)
        jd = C.c_double()
        L.o_jacobians(x, y, z, fk, C.byref(jd))
        rfk, rjd = refnum.jacobians(x, y, z)
        assert np.array_equal(fk.reshape(3, 4), rfk) and jd.value == rjd
        bb = cn["bbox"][t]
        phi, dphi = np.zeros(Np * n3q), np.zeros(3 * Np * n3q)
        L.o_basis(p, bb, fk, phi, dphi)
        rphi, rdphi = refnum.basis(p, _bbox_f(bb), rfk)
        assert rel(phi.reshape(Np, n3q), rphi) < 2e-15
        assert rel(dphi.reshape(3, Np, n3q), rdphi) < 2e-15
        # volume stiffness and mass from the reference's own dphi/phi
        V = np.zeros(9 * Np * Np)
        L.o_stiffness_volume(p, rjd, lam, mu, np.ascontiguousarray(rdphi.transpose(0, 1, 2)).ravel(), V)
        rV = refnum.stiff_vol(p, rjd, lam, mu, rdphi)
        assert rel(V.reshape(3, 3, Np, Np), rV) < 1e-14
        M = np.zeros(9 * Np * Np)
        L.o_mass_volume(p, rjd, np.ascontiguousarray(rphi).ravel(), rho, M)
        rM = refnum.mass_vol(p, rjd, rphi, rho)
        assert rel(M.reshape(3, 3, Np, Np), rM) < 1e-14
        # every face: Dirichlet / Neumann / interior variants
        for f in range(4):
            e2 = int(cn["neigh"][t, f])
            bb2 = cn["bbox"][e2] if e2 >= 0 else cn["bbox"][t]
            phib, gradb = np.zeros(2 * Np * n2q), np.zeros(6 * Np * n2q)
            L.o_basis_boundary(p, f, bb, bb2.ctypes.data_as(C.c_void_p) if e2 >= 0 else None,
                               fk, phib, gradb)
            rphib, rgradb = refnum.basis_boundary(p, f + 1, e2 if e2 < 0 else e2 + 1,
                                                  _bbox_f(bb), _bbox_f(bb2), rfk)
            assert rel(phib.reshape(2, Np, n2q)[0], rphib[:, :, 0]) < 2e-15
            assert rel(gradb.reshape(2, 3, Np, n2q)[0], rgradb[:, :, :, 0]) < 2e-15
            if e2 >= 0:
                assert rel(phib.reshape(2, Np, n2q)[1], rphib[:, :, 1]) < 2e-15
                assert rel(gradb.reshape(2, 3, Np, n2q)[1], rgradb[:, :, :, 1]) < 2e-15
            nrm, area = cn["normal"][t, f].copy(), float(cn["area"][t, f])
            hk1 = float(cn["hk"][t])
            hk2 = float(cn["hk"][e2]) if e2 >= 0 else hk1
            for kind in ([-1, -2] if e2 < 0 else [e2 + 1]):
                ours = [np.zeros(9 * Np * Np) for _ in range(5)]
                # feed the reference's own face basis to both
                pb = np.ascontiguousarray(np.stack([rphib[:, :, 0], rphib[:, :, 1]])).ravel()
                gb = np.ascontiguousarray(np.stack([rgradb[..., 0], rgradb[..., 1]])).ravel()
                L.o_stiffness_face(p, 10.0, kind if kind < 0 else 0, hk1, hk2, nrm, area,
                                   lam, mu, pb, gb, *ours)
                ref = refnum.stiff_face(p, 10.0, kind, hk1, hk2, nrm, area, lam, mu,
                                        rphib, rgradb)
                for a, b in zip(ours, ref):
                    a = a.reshape(3, 3, Np, Np)
                    if np.abs(b).max() == 0:
                        assert np.abs(a).max() == 0
                    else:
                        assert rel(a, b) < 1e-14


def test_rhs_against_reference_routines():
    """o_build_rhs row of a deck tet == reference MAKE_RHS_VOLUME + MAKE_RHS_FACE
    over its boundary faces (src/matrix_free.f90:711-819)."""
    d = deck.read_deck(os.path.join(DECKS, "tet_5"))
    for p in (1, 2, 3):
        o = oracle.from_deck(d, p=p)
        o.build_operator()
        rhs = o.rhs(oracle.PROB_REF_MMS).reshape(o.n_poly, 3, o.Np)
        cn = o.conn()
        for t in range(o.n_tet):
            x, y, z = _tet(d, t)
            rfk, rjd = refnum.jacobians(x, y, z)
            bb = _bbox_f(cn["bbox"][t])
            rphi, _ = refnum.basis(p, bb, rfk)
            acc = refnum.rhs_volume(p, rfk, rjd, 1.0, 1.0, 1.0, rphi).T.copy()
            for f in range(4):
                e2 = int(cn["neigh"][t, f])
                if e2 >= 0:
                    continue
                rphib, rgradb = refnum.basis_boundary(p, f + 1, e2, bb, bb, rfk)
                fr = refnum.rhs_face(p, f + 1, e2, float(cn["hk"][t]), float(cn["hk"][t]),
                                     cn["normal"][t, f].copy(), float(cn["area"][t, f]), rfk,
                                     1.0, 1.0, rphib, rgradb, int(cn["bc_space"][t, f]))
                acc = acc + fr.T
            assert rel(rhs[t], acc.T) < 1e-13, (p, t)


def test_cholesky_solve_and_fields():
    rng = np.random.default_rng(3)
    L = oracle.lib()
    for n in (4, 10, 20, 35):
        B = rng.standard_normal((n, n))
        A = B @ B.T + n * np.eye(n)
        R = np.zeros(n * n)
        assert L.o_cholesky(n, np.ascontiguousarray(A).ravel(), R) == 0
        R = R.reshape(n, n)
        assert np.array_equal(R, refnum.cholesky(A))
        # closed form: R^T R = A and R = L^T of numpy's lower factor
        assert np.allclose(R, np.linalg.cholesky(A).T, rtol=1e-12, atol=1e-12)
        b = rng.standard_normal(n)
        x = np.zeros(n)
        L.o_solve_lu(n, np.ascontiguousarray(R).ravel(), b, x)
        assert np.array_equal(x, refnum.solve_lu(R, b))
        assert np.allclose(A @ x, b, rtol=1e-10, atol=1e-10)
    # non-SPD -> error code 5 (EXIT_SINGULARMTX)
    assert L.o_cholesky(2, np.array([1.0, 2.0, 2.0, 1.0]), np.zeros(4)) == 5
    for t in (0.0, 0.123, 1.0000000000000007):
        assert oracle.time_function(oracle.PROB_REF_MMS, t) == refnum.time_function(t)


@pytest.mark.parametrize("name,tets", [("tet_5", [0, 13, 29]), ("tet_100", [3, 700])])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_double_couple(name, tets, p):
    """The SISM branch: det4, moment_density and MAKE_DOUBLE_COUPLE of the
    oracle against the reference's own routines on real deck tets, then the
    mesh-level loop (o_double_couples) against the element routine."""
    d = deck.read_deck(os.path.join(DECKS, name))
    o = oracle.from_deck(d, p=p)
    cn = o.conn()
    L = oracle.lib()
    Np, n3q = o.Np, (p + 1) ** 3
    rng = np.random.default_rng(p)
    for t in tets:
        x, y, z = _tet(d, t)
        A = np.stack([x, y, z, np.ones(4)], axis=1)
        assert L.o_det4(np.ascontiguousarray(A).ravel()) == refnum.det4(A)
        vol = abs(refnum.det4(A)) / 6.0
        rfk, rjd = refnum.jacobians(x, y, z)
        _, rdphi = refnum.basis(p, _bbox_f(cn["bbox"][t]), rfk)
        s, n, m0 = rng.standard_normal(3), rng.standard_normal(3), 10 ** rng.uniform(-2, 3)
        mom = np.zeros(9)
        L.o_moment_density(m0, vol, s, n, mom)
        out = np.zeros(3 * Np)
        L.o_double_couple(p, rjd, np.ascontiguousarray(rdphi).ravel(), mom, out)
        ref = refnum.double_couple(p, rjd, rdphi, m0, vol, s, n)
        assert rel(out, ref) < 1e-14, rel(out, ref)
        # the mesh-level loop puts the same vector into the tet's row
        rhs = o.double_couples([t], s, n, [m0])
        assert rel(rhs.reshape(o.n_poly, -1)[t], ref) < 1e-14

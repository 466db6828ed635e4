"""ctypes binding of oracle/_ref/libref_numerics.so: the reference's own
Fortran element routines (built from the reference sources by oracle/Makefile,
target `ref`).  Absent on machines without the reference tree; tests using it
skip there.  Arrays are allocated in Fortran order with the Fortran shapes."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "_ref", "libref_numerics.so")
_L = None


def available():
    return os.path.exists(PATH)


def lib():
    global _L
    if _L is None:
        _L = C.CDLL(PATH)
        _L.rh_time_function.restype = C.c_double
        _L.rh_time_function.argtypes = [C.c_double]
    return _L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def F(*shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


def np_(p):
    return (p + 1) * (p + 2) * (p + 3) // 6


def quadrature(p):
    n3, n2 = (p + 1) ** 3, (p + 1) ** 2
    t3, w3, t2, w2 = F(4, n3), F(n3), F(4, n2), F(n2)
    lib().rh_quadrature(C.c_int(p), _p(t3), _p(w3), _p(t2), _p(w2))
    return t3, w3, t2, w2


def basis_list(p):
    n = np_(p)
    b = F(n, 3, dtype=np.int32)
    lib().rh_basis_list(C.c_int(p), C.c_int(n), _p(b))
    return b


def jacobians(x, y, z):
    fk, jd = F(3, 4), np.zeros(1)
    x, y, z = (np.asfortranarray(v, np.float64) for v in (x, y, z))
    lib().rh_jacobians(_p(x), _p(y), _p(z), _p(fk), _p(jd))
    return fk, jd[0]


def basis(p, bbox_f, fk):
    n, n3 = np_(p), (p + 1) ** 3
    phi, dphi = F(n, n3), F(3, n, n3)
    lib().rh_basis(C.c_int(p), C.c_int(n), _p(bbox_f), _p(fk), _p(phi), _p(dphi))
    return phi, dphi


def basis_boundary(p, face1, e2, bb1, bb2, fk):
    n, n2 = np_(p), (p + 1) ** 2
    phib, gradb = F(n, n2, 2), F(3, n, n2, 2)
    lib().rh_basis_boundary(C.c_int(p), C.c_int(n), C.c_int(face1), C.c_int(e2), _p(bb1),
                            _p(bb2), _p(fk), _p(phib), _p(gradb))
    return phib, gradb


def stiff_vol(p, jdet, lam, mu, dphi):
    n = np_(p)
    v = F(3, 3, n, n)
    lib().rh_stiff_vol(C.c_int(p), C.c_int(n), C.c_double(jdet), C.c_double(lam),
                       C.c_double(mu), _p(dphi), _p(v))
    return v


def mass_vol(p, jdet, phi, rho):
    n = np_(p)
    m = F(3, 3, n, n)
    lib().rh_mass_vol(C.c_int(p), C.c_int(n), C.c_double(jdet), _p(phi), C.c_double(rho), _p(m))
    return m


def stiff_face(p, alpha, e2, hk1, hk2, nrm, area, lam, mu, phib, gradb):
    n = np_(p)
    out = [F(3, 3, n, n) for _ in range(5)]
    nrm = np.asfortranarray(nrm, np.float64)
    lib().rh_stiff_face(C.c_int(p), C.c_int(n), C.c_double(alpha), C.c_int(e2),
                        C.c_double(hk1), C.c_double(hk2), _p(nrm), C.c_double(area),
                        C.c_double(lam), C.c_double(mu), _p(phib), _p(gradb),
                        *(_p(o) for o in out))
    return out


def rhs_volume(p, fk, jdet, lam, mu, rho, phi):
    n = np_(p)
    r = F(3, n)
    lib().rh_rhs_volume(C.c_int(p), C.c_int(n), _p(fk), C.c_double(jdet), C.c_double(lam),
                        C.c_double(mu), C.c_double(rho), _p(phi), _p(r))
    return r


def rhs_face(p, face1, e2, hk1, hk2, nrm, area, fk, lam, mu, phib, gradb, tag):
    n = np_(p)
    r = F(3, n)
    nrm = np.asfortranarray(nrm, np.float64)
    lib().rh_rhs_face(C.c_int(p), C.c_int(n), C.c_int(face1), C.c_int(e2), C.c_double(hk1),
                      C.c_double(hk2), _p(nrm), C.c_double(area), _p(fk), C.c_double(lam),
                      C.c_double(mu), _p(phib), _p(gradb), C.c_int(tag), _p(r))
    return r


def cholesky(a):
    n = a.shape[0]
    A = np.asfortranarray(a)
    R = F(n, n)
    lib().rh_cholesky(C.c_int(n), _p(A), _p(R))
    return R


def solve_lu(R, b):
    n = R.shape[0]
    Rf = np.asfortranarray(R)
    bf = np.asfortranarray(b, np.float64)
    x = F(n)
    lib().rh_solve_lu(C.c_int(n), _p(Rf), _p(bf), _p(x))
    return x


def fields(x, lam, mu, rho, nrm):
    out = F(3, 5)
    x = np.asfortranarray(x, np.float64)
    nrm = np.asfortranarray(nrm, np.float64)
    lib().rh_fields(_p(x), C.c_double(lam), C.c_double(mu), C.c_double(rho), _p(nrm), _p(out))
    return out


def time_function(t):
    return lib().rh_time_function(t)


def double_couple(p, jdet, dphi, m0, vol, s, n):
    """MAKE_DOUBLE_COUPLE with moment_density(m0, vol, s, n); dphi (3, Np, nq3)."""
    np1 = np_(p)
    out = F(3 * np1)
    d = np.asfortranarray(dphi, np.float64)
    s_, n_ = (np.asfortranarray(v, np.float64) for v in (s, n))
    lib().rh_double_couple(C.c_int(p), C.c_int(np1), C.c_double(jdet), _p(d), C.c_double(m0),
                           C.c_double(vol), _p(s_), _p(n_), _p(out))
    return out


def det4(a):
    """det4 of a 4x4 array (Fortran order)."""
    lib().rh_det4.restype = C.c_double
    a = np.asfortranarray(a, np.float64)
    return lib().rh_det4(_p(a))

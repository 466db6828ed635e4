"""ctypes binding of the C oracle (liblymph_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PROB_NULL, PROB_REF_MMS, PROB_CUBE_WAVE = 0, 1, 2
FN_ZERO, FN_REF_VEL, FN_REF_UEX, FN_CUBE_U, FN_CUBE_V, FN_POLY = 0, 1, 2, 3, 4, 5

_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblymph_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.o_ctx_create.restype = C.c_int
        L.o_ctx_create.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_void_p, C.c_int,
                                   C.c_void_p, C.c_void_p, C.c_int, _ip, _ip, _dp,
                                   _dp, _dp, C.c_int, C.c_double, C.c_double,
                                   C.POINTER(C.c_void_p)]
        L.o_ctx_destroy.argtypes = [C.c_void_p]
        L.o_num_threads.argtypes = [C.c_int]
        L.o_num_threads.restype = C.c_int
        L.o_sizes.argtypes = [C.c_void_p, _ip]
        L.o_get_conn.argtypes = [C.c_void_p] + [C.c_void_p] * 7
        L.o_get_pattern.argtypes = [C.c_void_p, _ip, _ip]
        L.o_hmax.argtypes = [C.c_void_p]
        L.o_hmax.restype = C.c_double
        L.o_build_operator.argtypes = [C.c_void_p]
        L.o_build_operator.restype = C.c_int
        L.o_get_K.argtypes = [C.c_void_p, _dp]
        L.o_get_Adg.argtypes = [C.c_void_p, _dp]
        L.o_get_mass.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.o_build_rhs.argtypes = [C.c_void_p, C.c_int, _dp]
        L.o_build_rhs.restype = C.c_int
        L.o_project.argtypes = [C.c_void_p, C.c_int, _dp]
        L.o_project.restype = C.c_int
        L.o_first_step.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.o_step.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.o_error_l2_sq.argtypes = [C.c_void_p, _dp, _dp]
        L.o_error_l2_sq.restype = C.c_double
        L.o_error_dg_sq.argtypes = [C.c_void_p, _dp, _dp]
        L.o_error_dg_sq.restype = C.c_double
        L.o_run.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_void_p, _dp]
        L.o_run.restype = C.c_int
        L.o_time_function.argtypes = [C.c_int, C.c_double]
        L.o_time_function.restype = C.c_double
        L.o_locate_points.argtypes = [C.c_void_p, C.c_int, _dp, _ip]
        L.o_eval_points.argtypes = [C.c_void_p, C.c_int, _dp, _ip, _dp, _dp]
        L.o_point_sources.argtypes = [C.c_void_p, C.c_int, _dp, _ip, _dp, _dp]
        L.o_ricker.argtypes = [C.c_double, C.c_double, C.c_double]
        L.o_ricker.restype = C.c_double
        L.o_np.argtypes = [C.c_int]
        L.o_np.restype = C.c_int
        L.o_basis_list.argtypes = [C.c_int, _ip]
        L.o_quadrature.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.o_jacobians.argtypes = [_dp, _dp, _dp, _dp, C.POINTER(C.c_double)]
        L.o_basis.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.o_basis_boundary.argtypes = [C.c_int, C.c_int, _dp, C.c_void_p, _dp, _dp, _dp]
        L.o_stiffness_volume.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.o_mass_volume.argtypes = [C.c_int, C.c_double, _dp, C.c_double, _dp]
        L.o_stiffness_face.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                                       _dp, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                       _dp, _dp, _dp, _dp, _dp]
        L.o_cholesky.argtypes = [C.c_int, _dp, _dp]
        L.o_cholesky.restype = C.c_int
        L.o_solve_lu.argtypes = [C.c_int, _dp, _dp, _dp]
        L.o_problem_data.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double,
                                     C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.o_det4.argtypes = [_dp]
        L.o_det4.restype = C.c_double
        L.o_moment_density.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp]
        L.o_double_couple.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp]
        L.o_double_couples.argtypes = [C.c_void_p, C.c_int, _ip, _dp, _dp, _dp, _dp]
        L.o_set_poly_field.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
        L.o_set_poly_field.restype = C.c_int
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One mesh context of the C oracle."""

    def __init__(self, xyz, tet_vert, tria_vert, tria_tag, tag_kind, tag_space,
                 p, tet_poly=None, rho=None, lam=None, mu=None, alpha=10.0,
                 theta=-1.0):
        L = lib()
        xyz = np.ascontiguousarray(xyz, np.float64)
        tv = np.ascontiguousarray(tet_vert, np.int32)
        nt = tv.shape[0]
        self._keep = []
        tp = None if tet_poly is None else np.ascontiguousarray(tet_poly, np.int32)
        trv = np.ascontiguousarray(tria_vert, np.int32).reshape(-1, 3)
        trt = np.ascontiguousarray(tria_tag, np.int32)
        ones = np.ones(nt)
        rho = ones if rho is None else np.ascontiguousarray(rho, np.float64)
        lam = ones if lam is None else np.ascontiguousarray(lam, np.float64)
        mu = ones if mu is None else np.ascontiguousarray(mu, np.float64)
        tk = np.ascontiguousarray(tag_kind, np.int32)
        ts = np.ascontiguousarray(tag_space, np.int32)
        self._keep = [tp, trv, trt]
        h = C.c_void_p()
        rc = L.o_ctx_create(xyz.shape[0], xyz, nt, tv, _ptr(tp), trv.shape[0], _ptr(trv),
                            _ptr(trt), tk.shape[0], tk, ts, rho, lam, mu, p, alpha,
                            theta, C.byref(h))
        if rc:
            raise RuntimeError(f"o_ctx_create failed with code {rc}")
        self.h = h
        s = np.zeros(4, np.int32)
        L.o_sizes(h, s)
        self.n_tet, self.n_poly, self.Np, self.n_blk = (int(v) for v in s)
        self.p = p
        self.n3 = 3 * self.Np

    def __del__(self):
        if getattr(self, "h", None) is not None and _LIB is not None:
            _LIB.o_ctx_destroy(self.h)
            self.h = None

    def conn(self):
        nt, np_ = self.n_tet, self.n_poly
        out = dict(neigh=np.zeros((nt, 4), np.int32), neigh_face=np.zeros((nt, 4), np.int32),
                   bc_space=np.zeros((nt, 4), np.int32), normal=np.zeros((nt, 4, 3)),
                   area=np.zeros((nt, 4)), bbox=np.zeros((np_, 6)), hk=np.zeros(np_))
        lib().o_get_conn(self.h, *(_ptr(out[k]) for k in
                                   ("neigh", "neigh_face", "bc_space", "normal", "area", "bbox", "hk")))
        return out

    def pattern(self):
        rp = np.zeros(self.n_poly + 1, np.int32)
        col = np.zeros(self.n_blk, np.int32)
        lib().o_get_pattern(self.h, rp, col)
        return rp, col

    def hmax(self):
        return lib().o_hmax(self.h)

    def build_operator(self):
        rc = lib().o_build_operator(self.h)
        if rc:
            raise RuntimeError(f"o_build_operator failed with code {rc}")

    def K(self):
        k = np.zeros((self.n_blk, self.n3, self.n3))
        lib().o_get_K(self.h, k)
        return k

    def Adg(self):
        a = np.zeros((self.n_blk, self.n3, self.n3))
        lib().o_get_Adg(self.h, a)
        return a

    def mass(self):
        shp = (self.n_poly, self.Np, self.Np)
        M, Mm, RM, RMm = (np.zeros(shp) for _ in range(4))
        lib().o_get_mass(self.h, _ptr(M), _ptr(Mm), _ptr(RM), _ptr(RMm))
        return M, Mm, RM, RMm

    def rhs(self, problem):
        r = np.zeros(self.n_poly * self.n3)
        lib().o_build_rhs(self.h, problem, r)
        return r

    def project(self, fn):
        r = np.zeros(self.n_poly * self.n3)
        rc = lib().o_project(self.h, fn, r)
        if rc:
            raise RuntimeError("operator not built")
        return r

    def first_step(self, g0, dt, u0, v0, rs, rd):
        out = np.zeros_like(u0)
        lib().o_first_step(self.h, g0, dt, u0, v0, rs, rd, out)
        return out

    def step(self, g, dt, um1, u, rs, rd):
        out = np.zeros_like(u)
        lib().o_step(self.h, g, dt, um1, u, rs, rd, out)
        return out

    def error_l2_sq(self, uh, uex):
        return lib().o_error_l2_sq(self.h, uh, uex)

    def error_dg_sq(self, uh, uex):
        return lib().o_error_dg_sq(self.h, uh, uex)

    def locate(self, pts):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        tet = np.zeros(pts.shape[0], np.int32)
        lib().o_locate_points(self.h, pts.shape[0], pts, tet)
        return tet

    def eval_points(self, pts, tet, u):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        out = np.zeros((pts.shape[0], 3))
        lib().o_eval_points(self.h, pts.shape[0], pts, np.ascontiguousarray(tet, np.int32),
                            np.ascontiguousarray(u, np.float64), out)
        return out

    def point_sources(self, pts, tets, forces):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        out = np.zeros(self.n_poly * self.n3)
        lib().o_point_sources(self.h, pts.shape[0], pts, np.ascontiguousarray(tets, np.int32),
                              np.ascontiguousarray(forces, np.float64).reshape(-1, 3), out)
        return out

    def double_couples(self, tets, strike, normal, m0):
        """RHS entries of double-couple sources in the given tets (the SISM
        branch of MAKE_RHS_FREE)."""
        tets = np.ascontiguousarray(tets, np.int32)
        out = np.zeros(self.n_poly * self.n3)
        lib().o_double_couples(self.h, tets.size, tets,
                               np.ascontiguousarray(strike, np.float64).reshape(-1),
                               np.ascontiguousarray(normal, np.float64).reshape(-1),
                               np.ascontiguousarray(m0, np.float64).reshape(-1), out)
        return out

    def set_poly_field(self, exps, coef):
        """Field FN_POLY: component i = sum_k coef[k, i] x^a y^b z^c with
        (a, b, c) = exps[k]."""
        e = np.ascontiguousarray(exps, np.int32).reshape(-1, 3)
        cf = np.ascontiguousarray(coef, np.float64).reshape(-1, 3)
        if lib().o_set_poly_field(self.h, e.shape[0], e, cf):
            raise ValueError("too many polynomial terms")

    def run(self, problem, dt, stop_time, max_steps=-1, want_u=False):
        out = np.zeros(4)
        u = np.zeros(self.n_poly * self.n3) if want_u else None
        rc = lib().o_run(self.h, problem, dt, stop_time, max_steps, _ptr(u), out)
        if rc:
            raise RuntimeError(f"o_run failed with code {rc}")
        res = dict(t_final=out[0], nsteps=int(out[1]), err_l2=out[2], err_dg=out[3])
        if want_u:
            res["u"] = u
        return res


def time_function(problem, t):
    return lib().o_time_function(problem, t)


def from_deck(deck, p=None, **kw):
    return Oracle(deck["xyz"], deck["tet_vert"], deck["tria_vert"], deck["tria_tag"],
                  deck["tag_kind"], deck["tag_space"], deck["degree"] if p is None else p, **kw)


def ricker(f0, t0, t):
    return lib().o_ricker(f0, t0, t)


def problem_data(problem, xyz, normal=None, lam=1.0, mu=1.0, rho=1.0, tag=1):
    """Point values (f, gd, gn), each (n, 3), of a data set at xyz (n, 3)."""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    nrm = np.zeros_like(xyz) if normal is None else np.ascontiguousarray(normal, np.float64).reshape(-1, 3)
    f, gd, gn = (np.zeros((n, 3)) for _ in range(3))
    lib().o_problem_data(problem, n, xyz, nrm, lam, mu, rho, tag, _ptr(f), _ptr(gd), _ptr(gn))
    return f, gd, gn

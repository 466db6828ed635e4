"""Host-side mirror of the reference's `matrix_free` module
(src/matrix_free.f90) on top of liblymph_gpu.so.

`Context` owns one GPU context (one mesh, one degree).  Its methods carry the
reference routine names and argument meaning:

  make_matrices_free()                 MAKE_MATRICES_FREE + mass Cholesky
                                       (src/matrix_free.f90:27-464, src/Lymph3D.f90:264-271)
  make_rhs_free(problem, slot)         MAKE_RHS_FREE (:467-843)
  compute_modal_coefficients_free(fn)  COMPUTE_MODAL_COEFFICIENTS_FREE (:847-997)
  first_time_step_matrix_free(...)     FIRST_TIME_STEP_MATRIX_FREE (:1100-1227)
  time_step_matrix_free(...)           TIME_STEP_MATRIX_FREE (:1230-1385)
  compute_error_l2/dg_matrix_free(...) COMPUTE_ERROR_*_MATRIX_FREE (:1388-1504)

Errors raise LymphError whose .code reuses the reference's exit codes
(EXIT_SINGULARMTX=5 for a non-SPD mass block, EXIT_SETUP=4, ...).  Vectors are
numpy float64 arrays of n_poly*3*Np entries in the layout selected with
`layout` (polyhedron-major by default; "fortran" = the reference's column-major
u(npoly, 3Np)).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr


class Context:
    def __init__(self, mesh, p, alpha=10.0, theta=-1.0, layout="poly"):
        L = lib()
        self.mesh = mesh
        self.p = int(p)
        self.Np = (p + 1) * (p + 2) * (p + 3) // 6
        self.n3 = 3 * self.Np
        self.n_poly = mesh.n_poly
        cm, self._keep = mesh.to_c()
        h = C.c_void_p()
        check(L.lymph_ctx_create(C.byref(cm), self.p, alpha, theta, C.byref(h)), "lymph_ctx_create")
        self.h = h
        self.built = False
        if layout == "fortran":
            check(L.lymph_ctx_set_layout(h, _lib.LAYOUT_FORTRAN), "lymph_ctx_set_layout")
        info = np.zeros(8, np.int64)
        check(L.lymph_ctx_info(h, ptr(info)), "lymph_ctx_info")
        (self.n_tet, _, _, self.n_blk, self.operator_bytes, _,
         self.step_bytes, self.step_flops) = (int(v) for v in info)

    def close(self):
        if getattr(self, "h", None) is not None and _lib._LIB is not None:
            _lib._LIB.lymph_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- helpers --------------------------------------------------------
    def vec(self):
        return np.zeros(self.n_poly * self.n3)

    def _c(self, a):
        a = np.ascontiguousarray(a, np.float64)
        if a.size != self.n_poly * self.n3:
            raise ValueError(f"vector of {a.size} entries, expected {self.n_poly * self.n3}")
        return a

    def pattern(self):
        rp = np.zeros(self.n_poly + 1, np.int32)
        col = np.zeros(self.n_blk, np.int32)
        check(lib().lymph_get_pattern(self.h, ptr(rp), ptr(col)), "lymph_get_pattern")
        return rp, col

    def set_stream(self, stream_handle):
        check(lib().lymph_ctx_set_stream(self.h, C.c_void_p(stream_handle)), "lymph_ctx_set_stream")

    def stream(self):
        return lib().lymph_ctx_get_stream(self.h)

    # ---- the reference routines ----------------------------------------
    def make_matrices_free(self):
        check(lib().lymph_build_operator(self.h), "lymph_build_operator")
        self.built = True

    def export_operator(self):
        """(K, A_dg) as (n_blk, 3Np, 3Np) row-major blocks, (M, M_modal) as
        (n_poly, Np, Np)."""
        K = np.zeros((self.n_blk, self.n3, self.n3))
        A = np.zeros_like(K)
        M = np.zeros((self.n_poly, self.Np, self.Np))
        Mm = np.zeros_like(M)
        check(lib().lymph_export_operator(self.h, ptr(K), ptr(A), ptr(M), ptr(Mm)),
              "lymph_export_operator")
        return K, A, M, Mm

    def make_rhs_free(self, problem, slot=_lib.SLOT_RHS_DYN):
        out = self.vec()
        check(lib().lymph_build_rhs(self.h, problem, slot, ptr(out)), "lymph_build_rhs")
        return out

    def make_rhs_free_fn(self, forcing=None, dirichlet=None, neumann=None, slot=_lib.SLOT_RHS_DYN):
        """MAKE_RHS_FREE with caller data, the procedure arguments of
        src/matrix_free.f90:474-493 as vectorised numpy functions:
        forcing(lam, mu, rho, xyz), dirichlet(xyz, tag), neumann(lam, mu,
        normal, xyz, tag), each returning an (n, 3) array for n points (xyz
        and normal (n, 3); lam, mu, rho, tag (n,)); None = zero data."""
        import numpy.ctypeslib as npc

        def vec(p, n, k=3):
            return npc.as_array(p, shape=(n, k)) if k > 1 else npc.as_array(p, shape=(n,))

        keep = []
        if forcing is not None:
            def f_c(n, lam, mu, rho, xyz, out, user):
                vec(out, n)[:] = forcing(vec(lam, n, 1), vec(mu, n, 1), vec(rho, n, 1), vec(xyz, n))
            keep.append(_lib.FORCING_FN(f_c))
        else:
            keep.append(None)
        if dirichlet is not None:
            def d_c(n, xyz, tag, out, user):
                vec(out, n)[:] = dirichlet(vec(xyz, n), vec(tag, n, 1))
            keep.append(_lib.DIRICHLET_FN(d_c))
        else:
            keep.append(None)
        if neumann is not None:
            def n_c(n, lam, mu, nrm, xyz, tag, out, user):
                vec(out, n)[:] = neumann(vec(lam, n, 1), vec(mu, n, 1), vec(nrm, n), vec(xyz, n),
                                         vec(tag, n, 1))
            keep.append(_lib.NEUMANN_FN(n_c))
        else:
            keep.append(None)
        fps = [C.cast(k, C.c_void_p) if k is not None else None for k in keep]
        out = self.vec()
        check(lib().lymph_build_rhs_fn(self.h, fps[0], fps[1], fps[2], None, slot, ptr(out)),
              "lymph_build_rhs_fn")
        return out

    def compute_modal_coefficients_free_fn(self, field, slot=_lib.SLOT_UEX):
        """COMPUTE_MODAL_COEFFICIENTS_FREE of a caller function (the
        reference's f_analytic argument) as a vectorised numpy function
        field(xyz) -> (n, 3) at the (n, 3) volume quadrature points."""
        import numpy.ctypeslib as npc

        def f_c(n, xyz, out, user):
            npc.as_array(out, shape=(n, 3))[:] = field(npc.as_array(xyz, shape=(n, 3)))
        cb = _lib.FIELD_FN(f_c)
        out = self.vec()
        check(lib().lymph_project_fn(self.h, C.cast(cb, C.c_void_p), None, slot, ptr(out)),
              "lymph_project_fn")
        return out

    def compute_modal_coefficients_free(self, fn, slot=_lib.SLOT_UEX):
        out = self.vec()
        check(lib().lymph_project(self.h, fn, slot, ptr(out)), "lymph_project")
        return out

    def set_slot(self, slot, v):
        v = self._c(v)
        check(lib().lymph_set_slot(self.h, slot, ptr(v)), "lymph_set_slot")

    def get_slot(self, slot):
        out = self.vec()
        check(lib().lymph_get_slot(self.h, slot, ptr(out)), "lymph_get_slot")
        return out

    def first_time_step_matrix_free(self, g0, dt, u0, v0):
        out = self.vec()
        u0, v0 = self._c(u0), self._c(v0)
        check(lib().lymph_first_step(self.h, g0, dt, ptr(u0), ptr(v0), ptr(out)), "lymph_first_step")
        return out

    def time_step_matrix_free(self, g, dt, u0, un):
        """u^{n+1} from u^{n-1}=u0 and u^n=un (argument order of the reference)."""
        out = self.vec()
        u0, un = self._c(u0), self._c(un)
        check(lib().lymph_step(self.h, g, dt, ptr(u0), ptr(un), ptr(out)), "lymph_step")
        return out

    def step_n(self, problem, nsteps, t0, dt):
        t = C.c_double()
        check(lib().lymph_step_n(self.h, problem, nsteps, t0, dt, C.byref(t)), "lymph_step_n")
        return t.value

    def time_steps(self, problem, nsteps, t0, dt):
        ms = C.c_double()
        check(lib().lymph_time_steps(self.h, problem, nsteps, t0, dt, C.byref(ms)), "lymph_time_steps")
        return ms.value

    def autotune(self, reps=20):
        """Time the step-kernel variants and keep the fastest; returns
        (chosen (group size, broadcast mode, waves per polyhedron),
        {(group size, mode, waves): ms/step}) -- ms -1 for variants not run."""
        nv = C.c_int()
        unr = np.zeros(16, np.int32)
        xm = np.zeros(16, np.int32)
        wp = np.zeros(16, np.int32)
        ms = np.zeros(16)
        ch = C.c_int()
        check(lib().lymph_autotune(self.h, reps, C.byref(nv), ptr(unr), ptr(xm), ptr(wp), ptr(ms),
                                   C.byref(ch)), "lymph_autotune")
        table = {(int(unr[i]), int(xm[i]), int(wp[i])): float(ms[i]) for i in range(nv.value)}
        c = ch.value
        self.chosen_variant = c  # the index set_step_variant takes
        return (int(unr[c]), int(xm[c]), int(wp[c])), table

    def set_step_variant(self, index):
        """Select step-kernel variant `index` (numbered as in autotune);
        returns the number of variants."""
        nv = C.c_int()
        check(lib().lymph_set_step_variant(self.h, int(index), C.byref(nv)), "lymph_set_step_variant")
        return nv.value

    def n_step_variants(self):
        nv = C.c_int()
        check(lib().lymph_set_step_variant(self.h, 0, C.byref(nv)), "lymph_set_step_variant")
        return nv.value

    def set_apply_precision(self, bits):
        """Precision of the step's operator apply: 64 (default), 32 (FP32
        blocks in the mass-orthonormal basis, FP32 sums) or 48 (those FP32
        blocks plus a 16-bit correction per entry, FP64 sums)."""
        check(lib().lymph_set_apply_precision(self.h, bits), "lymph_set_apply_precision")
        self.apply_bits = bits

    def set_keep_adg(self, keep):
        """False: do not build (or free) the A_dg blocks; the DG-norm error
        is then unavailable.  Saves one operator's worth of device memory."""
        check(lib().lymph_ctx_set_keep_adg(self.h, int(bool(keep))), "lymph_ctx_set_keep_adg")

    def set_build(self, mode):
        """"reference" (default: the reference's operand order) or "factored"
        (per-basis-pair quadrature sums, ~10x fewer flops, same operator to
        rounding) for the next make_matrices_free."""
        code = {"reference": 0, "factored": 1}[mode]
        check(lib().lymph_ctx_set_build(self.h, code), "lymph_ctx_set_build")

    def release_fp64_operator(self):
        """Free the FP64 K blocks once the 32-bit apply is selected."""
        check(lib().lymph_release_fp64_operator(self.h), "lymph_release_fp64_operator")

    def set_symmetric_storage(self, on=True, fp32=False):
        """Keep every face pair's block once (K_QP applied as the transpose of
        K_PQ) for the FP64 step; the operator must pass the device's pair
        check (SIPG, one material per face).  With fp32=True (p = 3) the
        32- and 48-bit applies keep their blocks (and the 48-bit corrections)
        once too; otherwise they keep their full rows.  Returns the largest max|K_PQ - K_QP^T| / max|K_PQ| found (0
        when switching off)."""
        mode = 0 if not on else (2 if fp32 else 1)
        asym = C.c_double()
        check(lib().lymph_set_symmetric_storage(self.h, mode, C.byref(asym)),
              "lymph_set_symmetric_storage")
        return asym.value

    @property
    def device_bytes(self):
        info = np.zeros(8, np.int64)
        check(lib().lymph_ctx_info(self.h, ptr(info)), "lymph_ctx_info")
        return int(info[5])

    def step_bytes_for(self, bits):
        """Algorithmic bytes of one step with the operator in `bits`
        precision (64, 32, or 48: 4 bytes per entry plus a 2-byte
        correction): the blocks, five vectors, and for the 64-bit apply the
        packed mass factors (the 32- and 48-bit applies work in the
        mass-orthonormal basis and solve nothing)."""
        n3, Np = self.n3, self.Np
        factors = 8 * self.n_poly * Np * (Np + 1) // 2 if bits == 64 else 0
        return (bits // 8) * self.n_blk * n3 * n3 + factors + 8 * 5 * self.n_poly * n3

    # ---- point sources (Ricker) -----------------------------------------
    def add_point_sources(self, points, forces, slot=_lib.SLOT_RHS_DYN, tets=None):
        pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        f = np.ascontiguousarray(forces, np.float64).reshape(-1, 3)
        tets = self.mesh.locate(pts) if tets is None else np.ascontiguousarray(tets, np.int32)
        check(lib().lymph_add_point_sources(self.h, pts.shape[0], ptr(pts), ptr(tets), ptr(f), slot),
              "lymph_add_point_sources")
        return tets

    def add_double_couples(self, points, strike, normal, m0, slot=_lib.SLOT_RHS_DYN, tets=None):
        """Double-couple sources (the SISM branch of MAKE_RHS_FREE) at
        `points` (n, 3), located like FIND_ELEM_FROM_POINT unless `tets` is
        given; strike and fault normal (n, 3), moment m0 (n,)."""
        if tets is None:
            tets = self.mesh.locate(np.ascontiguousarray(points, np.float64).reshape(-1, 3))
        tets = np.ascontiguousarray(tets, np.int32)
        s = np.ascontiguousarray(strike, np.float64).reshape(-1, 3)
        nrm = np.ascontiguousarray(normal, np.float64).reshape(-1, 3)
        m0 = np.ascontiguousarray(m0, np.float64).reshape(-1)
        if not (s.shape[0] == nrm.shape[0] == m0.size == tets.size):
            raise ValueError("one strike, normal and moment per source")
        check(lib().lymph_add_double_couples(self.h, tets.size, ptr(tets), ptr(s), ptr(nrm), ptr(m0), slot),
              "lymph_add_double_couples")
        return tets

    def set_ricker(self, f0, t0=None):
        self.ricker = (f0, 1.2 / f0 if t0 is None else t0)
        check(lib().lymph_set_ricker(self.h, *self.ricker), "lymph_set_ricker")

    def set_source_groups(self, poly_group, f0=None, t0=None):
        """Per-group Ricker wavelets: polyhedron i uses (f0[g], t0[g]) with
        g = poly_group[i]; t0 defaults to 1.2/f0.  None clears."""
        if poly_group is None:
            check(lib().lymph_set_source_groups(self.h, 0, None, None, None), "lymph_set_source_groups")
            return
        pg = np.ascontiguousarray(poly_group, np.int32)
        f0 = np.ascontiguousarray(f0, np.float64)
        t0 = 1.2 / f0 if t0 is None else np.ascontiguousarray(t0, np.float64)
        self._groups = (pg, f0, t0)
        check(lib().lymph_set_source_groups(self.h, f0.size, ptr(pg), ptr(f0), ptr(t0)),
              "lymph_set_source_groups")

    def start_from_rest(self, dt, sources=None, forces=None, f0=5.0, t0=None):
        """Operator (if needed), zero static RHS, point-force dynamic RHS, zero
        initial state, first step at t=0; returns t after the first step."""
        if not self.built:
            self.make_matrices_free()
        self.set_ricker(f0, t0)
        self.make_rhs_free(_lib.PROB_NULL, _lib.SLOT_RHS_STAT)
        self.make_rhs_free(_lib.PROB_RICKER, _lib.SLOT_RHS_DYN)
        if sources is not None:
            self.add_point_sources(sources, forces)
        self.compute_modal_coefficients_free(_lib.FN_ZERO, _lib.SLOT_U_PREV)
        self.compute_modal_coefficients_free(_lib.FN_ZERO, _lib.SLOT_V0)
        g0 = lib().lymph_ricker(self.ricker[0], self.ricker[1], 0.0)
        check(lib().lymph_first_step_dev(self.h, g0, dt, lib().lymph_slot_ptr(self.h, _lib.SLOT_U_PREV),
                                         lib().lymph_slot_ptr(self.h, _lib.SLOT_V0),
                                         lib().lymph_slot_ptr(self.h, _lib.SLOT_U_CUR)),
              "lymph_first_step_dev")
        return 0.0 + dt

    # ---- captured solves (HIP graphs) -------------------------------------
    def capture_solve(self, problem, dt, nsteps, every=1):
        check(lib().lymph_capture_solve(self.h, problem, dt, nsteps, every), "lymph_capture_solve")

    def replay(self):
        check(lib().lymph_replay(self.h), "lymph_replay")

    def sync(self):
        check(lib().lymph_sync(self.h), "lymph_sync")

    # ---- receivers ------------------------------------------------------
    def set_receivers(self, points, max_records=0, tets=None):
        """Receivers at `points` (n, 3), located like FIND_ELEM_FROM_POINT
        (nearest tet barycentre) unless `tets` is given."""
        pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        tets = self.mesh.locate(pts) if tets is None else np.ascontiguousarray(tets, np.int32)
        self.n_recv = pts.shape[0]
        check(lib().lymph_set_receivers(self.h, self.n_recv, ptr(pts), ptr(tets), max_records),
              "lymph_set_receivers")
        return tets

    def sample_receivers(self, slot=_lib.SLOT_U_CUR):
        out = np.zeros((self.n_recv, 3))
        check(lib().lymph_sample_receivers(self.h, slot, ptr(out)), "lymph_sample_receivers")
        return out

    def step_n_record(self, problem, nsteps, t0, dt, every=1):
        t = C.c_double()
        check(lib().lymph_step_n_record(self.h, problem, nsteps, t0, dt, every, C.byref(t)),
              "lymph_step_n_record")
        return t.value

    def traces(self):
        n = C.c_int()
        check(lib().lymph_get_traces(self.h, None, C.byref(n)), "lymph_get_traces")
        out = np.zeros((n.value, self.n_recv, 3))
        check(lib().lymph_get_traces(self.h, ptr(out), C.byref(n)), "lymph_get_traces")
        return out

    def compute_error_l2_matrix_free(self, uh, uex):
        e = C.c_double()
        uh, uex = self._c(uh), self._c(uex)
        check(lib().lymph_error_l2(self.h, ptr(uh), ptr(uex), C.byref(e)), "lymph_error_l2")
        return e.value

    def compute_error_dg_matrix_free(self, uh, uex):
        e = C.c_double()
        uh, uex = self._c(uh), self._c(uex)
        check(lib().lymph_error_dg(self.h, ptr(uh), ptr(uex), C.byref(e)), "lymph_error_dg")
        return e.value

    def check_finite(self, slot=_lib.SLOT_U_CUR):
        """True if slot `slot` holds only finite values (instability control)."""
        ok = C.c_int()
        check(lib().lymph_check_finite(self.h, slot, C.byref(ok)), "lymph_check_finite")
        return bool(ok.value)

    def error_slots(self, slot_uh, slot_uex):
        l2, dg = C.c_double(), C.c_double()
        check(lib().lymph_error_slots(self.h, slot_uh, slot_uex, C.byref(l2), C.byref(dg)),
              "lymph_error_slots")
        return l2.value, dg.value

    def run(self, problem, dt, stop_time, max_steps=-1, want_u=False):
        """The whole path of src/Lymph3D.f90:262-518 on the device."""
        out = np.zeros(4)
        u = self.vec() if want_u else None
        check(lib().lymph_run(self.h, problem, dt, stop_time, max_steps, ptr(u), ptr(out)), "lymph_run")
        self.built = True
        res = dict(t_final=out[0], nsteps=int(out[1]), err_l2=out[2], err_dg=out[3])
        if want_u:
            res["u"] = u
        return res

"""Driver of the reference's program Lymph3D (src/Lymph3D.f90:81-562) for the
matrix-free path, on the GPU:

  read deck -> mesh set-up -> hmax / CFL check -> material override
  (rho=lambda=mu=1, :196-203) -> operator + Cholesky -> RHS (static with null
  data, dynamic with the manufactured data) -> IC projections -> first step ->
  `do while (t <= stop_time)` loop -> exact-solution projection scaled by
  time_function(t) -> L2 / DG errors (printed as :515-517).

With monitors (the deck's SAVEOUT path, :400-406 and :441) the loop also
writes the EnSight monitor files every TMONITOR steps.

With --precision 32 or 48 the operator is applied by the 32- or 48-bit
apply (FP32 blocks in the mass-orthonormal basis; 48: plus a 16-bit
correction per entry, FP64 sums); with --pairs-once each face pair's block is
stored once (a symmetric operator: SIPG and the deck's forced single
material; at p = 3 also for the 32- and 48-bit applies).

Usage: python -m lymph3d_apsc_amd.driver <deck folder> [--max-steps N]
           [--monitor DIR [--every N]] [--precision 64|32|48] [--pairs-once]
"""
import argparse
import sys
import time

from . import _lib
from .matrix_free import Context
from .mesh import read_deck

EXIT_CFL = 1


def compute_cfl(hmax, dt, speed=1.0):
    """COMPUTE_CFL (src/checks.f90:14-40): fail if dt > hmax/speed."""
    return not (dt > hmax / speed)


def _fields(problem):
    """Initial displacement, initial velocity and exact solution of a data set
    (the choice lymph_run makes)."""
    u0 = _lib.FN_CUBE_U if problem == _lib.PROB_CUBE_WAVE else _lib.FN_ZERO
    v0 = _lib.FN_REF_VEL if problem == _lib.PROB_REF_MMS else _lib.FN_ZERO
    ex = {_lib.PROB_REF_MMS: _lib.FN_REF_UEX, _lib.PROB_CUBE_WAVE: _lib.FN_CUBE_U}.get(problem, _lib.FN_ZERO)
    return u0, v0, ex


def run_monitored(ctx, problem, dt, stop_time, every, out, max_steps=-1, strict=False):
    """The loop of src/Lymph3D.f90:325-441 with SAVEOUT monitors: after loop
    iteration num_dt (0-based) the state is exported when num_dt % every == 0
    (EXPORT_SOLUTION: vertex values of every tet; the mesh, polyhedron and
    boundary files with the first export), and solution.case is written at
    the end.  The steps run on the device in chunks between exports, with
    t accumulated as the reference does, so the final state and errors are
    those of Context.run bit for bit."""
    from . import export as X
    u0f, v0f, exf = _fields(problem)
    if not ctx.built:
        ctx.make_matrices_free()
    ctx.make_rhs_free(_lib.PROB_NULL, _lib.SLOT_RHS_STAT)
    ctx.make_rhs_free(problem, _lib.SLOT_RHS_DYN)
    u0 = ctx.compute_modal_coefficients_free(u0f, _lib.SLOT_U_PREV)
    v0 = ctx.compute_modal_coefficients_free(v0f, _lib.SLOT_V0)
    ctx.set_slot(_lib.SLOT_U_CUR, ctx.first_time_step_matrix_free(
        _lib.time_function(problem, 0.0), dt, u0, v0))
    t = 0.0 + dt
    nsteps, num_dt, files = 1, 0, []
    while t <= stop_time and (max_steps < 0 or nsteps < max_steps):
        # steps up to and including the next export, stopping where the loop would
        k, tt = 0, t
        while True:
            k += 1
            tt = tt + dt
            if (num_dt + k - 1) % every == 0 or not tt <= stop_time or \
                    (max_steps >= 0 and nsteps + k >= max_steps):
                break
        t = ctx.step_n(problem, k, t, dt)
        nsteps += k
        num_dt += k
        if (num_dt - 1) % every == 0:
            vals = X.vertex_values(ctx)
            if not files:
                X.write_ensight_mesh(out, ctx.mesh, strict)
                X.write_ensight_boundary(out, ctx.mesh, strict=strict)
            files.append(X.write_ensight_solution(out, vals, num_dt - 1))
    X.write_ensight_case(out, dt, stop_time, every)
    uex = ctx.compute_modal_coefficients_free(exf) * _lib.time_function(problem, t)
    l2 = ctx.compute_error_l2_matrix_free(ctx.get_slot(_lib.SLOT_U_CUR), uex)
    dg = ctx.compute_error_dg_matrix_free(ctx.get_slot(_lib.SLOT_U_CUR), uex)
    return dict(t_final=t, nsteps=nsteps, err_l2=l2 ** 0.5, err_dg=dg ** 0.5, files=files)


def run_deck(folder, max_steps=-1, p=None, verbose=True, problem=_lib.PROB_REF_MMS, monitor=None,
             every=None, precision=64, pairs_once=False):
    """Run a deck; with `monitor` a folder, also write the EnSight monitor
    files there every `every` steps (default: the deck's TMONITOR).
    `precision` 32 or 48 selects that operator apply and `pairs_once` the
    storage of each face pair's block once, both set up after the operator
    build; the error norms stay FP64."""
    mesh = read_deck(folder)
    hmax = mesh.hmax()
    if not compute_cfl(hmax, mesh.dt):
        raise SystemExit(EXIT_CFL)
    deg = mesh.degree if p is None else p
    t0 = time.perf_counter()
    with Context(mesh, deg) as ctx:
        if precision != 64 or pairs_once:
            ctx.make_matrices_free()
            if pairs_once:
                ctx.set_symmetric_storage(True, fp32=precision != 64)
            if precision != 64:
                ctx.set_apply_precision(precision)
        if monitor:
            res = run_monitored(ctx, problem, mesh.dt, mesh.stop_time,
                                int(every or mesh.monitor_every or 1), monitor, max_steps)
        else:
            res = ctx.run(problem, mesh.dt, mesh.stop_time, max_steps)
    res["wall_s"] = time.perf_counter() - t0
    res["hmax"] = hmax
    res["dofs"] = mesh.n_dof(deg)
    res["precision"], res["pairs_once"] = precision, bool(pairs_once)
    if verbose:
        print(" GRID SIZE: ", repr(hmax))
        print(" Global number of degrees of freedom: ", res["dofs"])
        print(" ERROR IN NORM L2: ", repr(res["err_l2"]))
        print(" ERROR IN NORM DG: ", repr(res["err_dg"]))
        print(f" steps {res['nsteps']}  t_final {res['t_final']!r}  wall {res['wall_s']:.3f}s")
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("deck")
    ap.add_argument("--max-steps", type=int, default=-1)
    ap.add_argument("--degree", type=int, default=None)
    ap.add_argument("--monitor", default=None, help="folder for the EnSight monitor files")
    ap.add_argument("--every", type=int, default=None, help="steps between exports (default TMONITOR)")
    ap.add_argument("--precision", type=int, choices=(64, 32, 48), default=64,
                    help="operator apply: FP64 (default), 32-bit, or 48-bit (FP32 + 16-bit corrections)")
    ap.add_argument("--pairs-once", action="store_true",
                    help="store each face pair's block once (symmetric operators)")
    a = ap.parse_args(argv)
    run_deck(a.deck, a.max_steps, a.degree, monitor=a.monitor, every=a.every, precision=a.precision,
             pairs_once=a.pairs_once)
    return 0


if __name__ == "__main__":
    sys.exit(main())

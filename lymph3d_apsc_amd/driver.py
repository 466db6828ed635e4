"""Driver of the reference's program Lymph3D (src/Lymph3D.f90:81-562) for the
matrix-free path, on the GPU:

  read deck -> mesh set-up -> hmax / CFL check -> material override
  (rho=lambda=mu=1, :196-203) -> operator + Cholesky -> RHS (static with null
  data, dynamic with the manufactured data) -> IC projections -> first step ->
  `do while (t <= stop_time)` loop -> exact-solution projection scaled by
  time_function(t) -> L2 / DG errors (printed as :515-517).

Usage: python -m lymph3d_apsc_amd.driver <deck folder> [--max-steps N]
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


def run_deck(folder, max_steps=-1, p=None, verbose=True, problem=_lib.PROB_REF_MMS):
    mesh = read_deck(folder)
    hmax = mesh.hmax()
    if not compute_cfl(hmax, mesh.dt):
        raise SystemExit(EXIT_CFL)
    deg = mesh.degree if p is None else p
    t0 = time.perf_counter()
    with Context(mesh, deg) as ctx:
        res = ctx.run(problem, mesh.dt, mesh.stop_time, max_steps)
    res["wall_s"] = time.perf_counter() - t0
    res["hmax"] = hmax
    res["dofs"] = mesh.n_dof(deg)
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
    a = ap.parse_args(argv)
    run_deck(a.deck, a.max_steps, a.degree)
    return 0


if __name__ == "__main__":
    sys.exit(main())

/*
 * lymph_oracle.c -- CPU restatement of LYMPH3D's matrix-free PolyDG leapfrog
 * path (FP64, scalar, OpenMP over polyhedra; results do not depend on the
 * thread count because every polyhedron's work is independent and reductions
 * are done serially).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Each
 * routine cites the reference routine it restates; expressions keep the
 * reference's operand order so that rounding matches it as closely as C
 * without FMA contraction allows (build with -ffp-contract=off).
 */
#include "lymph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DIMC 3
static const double PI_ = 3.141592653589793238462643383279502884; /* 4*atan(1) */

static double o_pi(void) { return 4.0 * atan(1.0); }
static double o_sqrt2(void) { return sqrt(2.0); }

struct OCtx {
  int p, Np, nq3, nq2, n3;
  double alpha, theta;
  int n_node, n_tet, n_poly;
  double* xyz;
  int* tet_vert;
  int* tet_poly;
  double *rho, *lam, *mu;
  int *neigh, *neigh_face, *bc_space;
  double *normal, *area;
  double *bbox, *hk;
  int *ptet_ptr, *ptet;
  int n_blk;
  int *row_ptr, *col;
  int* face_blk;
  double *nodtet3, *weitet3, *nodtria2, *weitria2;
  int* blist;
  double *K, *A, *M, *Mm, *RM, *RMm;
  int built;
  /* O_FN_POLY: u_i = sum_k poly_coef[k*3+i] x^a y^b z^c, (a,b,c) = poly_exp[k] */
  int poly_n;
  int poly_exp[O_MAX_POLY_TERMS * 3];
  double poly_coef[O_MAX_POLY_TERMS * 3];
};

int o_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int o_np(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

/* ---------------- basis list: src/basis_function.f90:12-44 ---------------- */
void o_basis_list(int p, int* blist) {
  int ii = 0;
  for (int q1 = p; q1 >= 0; --q1)
    for (int q2 = p - q1; q2 >= 0; --q2)
      for (int q3 = p - q1 - q2; q3 >= 0; --q3) {
        blist[ii * 3 + 0] = q1;
        blist[ii * 3 + 1] = q2;
        blist[ii * 3 + 2] = q3;
        ++ii;
      }
}

/* ------------- 1D Gauss-Legendre, hard-coded node order for n<=7 ----------
 * src/basis_function.f90:77-124.  Returns 0 on success, -1 if n > 7. */
static int gl1d(int n, double* x, double* w) {
  switch (n) {
    case 1:
      x[0] = 0.0; w[0] = 2.0;
      return 0;
    case 2:
      x[0] = sqrt(1.0 / 3.0); x[1] = -sqrt(1.0 / 3.0);
      w[0] = 1.0; w[1] = 1.0;
      return 0;
    case 3:
      x[0] = 0.0; x[1] = sqrt(3.0 / 5.0); x[2] = -sqrt(3.0 / 5.0);
      w[0] = 8.0 / 9.0; w[1] = 5.0 / 9.0; w[2] = 5.0 / 9.0;
      return 0;
    case 4:
      x[0] = 0.3399810435848563; x[1] = -0.3399810435848563;
      x[2] = 0.8611363115940526; x[3] = -0.8611363115940526;
      w[0] = (18 + sqrt(30.0)) / 36.0; w[1] = w[0];
      w[2] = (18 - sqrt(30.0)) / 36.0; w[3] = w[2];
      return 0;
    case 5:
      x[0] = 0.0; x[1] = 0.5384693101056831; x[2] = -0.5384693101056831;
      x[3] = 0.9061798459386640; x[4] = -0.9061798459386640;
      w[0] = 128.0 / 225.0;
      w[1] = (322 + 13 * sqrt(70.0)) / 900.0; w[2] = w[1];
      w[3] = (322 - 13 * sqrt(70.0)) / 900.0; w[4] = w[3];
      return 0;
    case 6:
      x[0] = 0.6612093864662645; x[1] = -0.6612093864662645;
      x[2] = 0.2386191860831969; x[3] = -0.2386191860831969;
      x[4] = 0.9324695142031521; x[5] = -0.9324695142031521;
      w[0] = 0.3607615730481386; w[1] = 0.3607615730481386;
      w[2] = 0.4679139345726910; w[3] = 0.4679139345726910;
      w[4] = 0.1713244923791704; w[5] = 0.1713244923791704;
      return 0;
    case 7:
      x[0] = 0.0; x[1] = 0.4058451513773972; x[2] = -0.4058451513773972;
      x[3] = 0.7415311855993945; x[4] = -0.7415311855993945;
      x[5] = 0.9491079123427585; x[6] = -0.9491079123427585;
      w[0] = 0.4179591836734694; w[1] = 0.3818300505051189;
      w[2] = 0.3818300505051189; w[3] = 0.2797053914892766;
      w[4] = 0.2797053914892766; w[5] = 0.1294849661688697;
      w[6] = 0.1294849661688697;
      return 0;
    default:
      return -1;
  }
}

/* quadrature (src/basis_function.f90:131-159) followed by the Duffy map
 * (src/Poly_ref_mappings.f90:137-162) and the square->triangle map (:165-186).
 * nodtet3 is stored q-major: nodtet3[q*4 + k]. */
void o_quadrature(int p, double* nodtet3, double* weitet3, double* nodtria2,
                  double* weitria2) {
  int n = p + 1;
  double x[8], w[8];
  gl1d(n, x, w);
  int c = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) {
        double a = 0.5 * x[i] + 0.5, b = 0.5 * x[j] + 0.5, cc = 0.5 * x[k] + 0.5;
        double wq = w[i] * w[j] * w[k];
        wq = 0.125 * wq;
        nodtet3[c * 4 + 0] = a;
        nodtet3[c * 4 + 1] = (1 - a) * b;
        nodtet3[c * 4 + 2] = (1 - a) * (1 - b) * cc;
        nodtet3[c * 4 + 3] = 1.0;
        double jt = (1 - b) * ((1 - a) * (1 - a));
        weitet3[c] = fabs(jt) * wq;
        ++c;
      }
  c = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = x[i], t = x[j];
      double wq = w[i] * w[j];
      nodtria2[c * 4 + 0] = 0.5 * (1 + s);
      nodtria2[c * 4 + 1] = 0.25 * (1 - s) * (1 + t);
      nodtria2[c * 4 + 2] = 1.0;
      nodtria2[c * 4 + 3] = 1.0;
      weitria2[c] = fabs(0.125 * (1 - s)) * wq;
      ++c;
    }
}

/* node_maps(i,j,k) of src/Poly_ref_mappings.f90:79-134 as nm[k][i][j]. */
static void node_maps(double nm[4][4][4]) {
  memset(nm, 0, sizeof(double) * 64);
  nm[0][0][1] = 1.0; nm[0][1][0] = 1.0; nm[0][3][3] = 1.0;
  nm[1][0][0] = 1.0; nm[1][2][1] = 1.0; nm[1][3][3] = 1.0;
  nm[2][0][0] = 1.0; nm[2][1][1] = 1.0; nm[2][2][2] = 1.0; nm[2][3][3] = 1.0;
  nm[2][2][0] = -1.0; nm[2][2][1] = -1.0;
  nm[3][1][1] = 1.0; nm[3][2][0] = 1.0; nm[3][3][3] = 1.0;
}

/* jacobians: src/Poly_ref_mappings.f90:18-75; det3: vet_mat_operations.f90:38-49.
 * Fk row-major 3x4: Fk[d*4 + j]. */
void o_jacobians(const double* x, const double* y, const double* z, double* Fk,
                 double* Jdet) {
  static const int mat[4][4] = {{-1, 1, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 0, 1},
                                {1, 0, 0, 0}};
  for (int j = 0; j < 4; ++j) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int l = 0; l < 4; ++l) {
      a = a + mat[j][l] * x[l];
      b = b + mat[j][l] * y[l];
      c = c + mat[j][l] * z[l];
    }
    Fk[0 * 4 + j] = a;
    Fk[1 * 4 + j] = b;
    Fk[2 * 4 + j] = c;
  }
#define A_(i, j) Fk[((i)-1) * 4 + ((j)-1)]
  *Jdet = A_(1, 1) * A_(2, 2) * A_(3, 3) - A_(1, 1) * A_(2, 3) * A_(3, 2) -
          A_(1, 2) * A_(2, 1) * A_(3, 3) + A_(1, 2) * A_(2, 3) * A_(3, 1) +
          A_(1, 3) * A_(2, 1) * A_(3, 2) - A_(2, 2) * A_(1, 3) * A_(3, 1);
#undef A_
}

/* Legendre polynomial and derivative, non-normalised, at xp
 * (src/basis_function.f90:268-340, 370-437).  Degree <= 7; P'_6 keeps the
 * reference's constant term -5/16 (src/basis_function.f90:413-414). */
static double leg(int k, double xp) {
  double x2 = xp * xp;
  switch (k) {
    case 0: return 1.0;
    case 1: return xp;
    case 2: return (3 * x2 - 1) / 2;
    case 3: return (5 * (x2 * xp) - 3 * xp) / 2;
    case 4: return (35 * (x2 * x2) - 30 * x2 + 3) / 8;
    case 5: return (63 * (x2 * x2 * xp) - 70 * (x2 * xp) + 15 * xp) / 8;
    case 6: return (231 * (x2 * x2 * x2) - 315 * (x2 * x2) + 105 * x2 - 5) / 16;
    case 7:
      return (429 * (x2 * x2 * x2 * xp) - 693 * (x2 * x2 * xp) +
              315 * (x2 * xp) - 35 * xp) / 16;
  }
  return 0.0;
}
static double dleg(int k, double xp) {
  double x2 = xp * xp;
  switch (k) {
    case 0: return 0.0;
    case 1: return 1.0;
    case 2: return 3 * xp;
    case 3: return (15 * x2 - 3) / 2;
    case 4: return (35 * (x2 * xp) - 15 * xp) / 2;
    case 5: return (315 * (x2 * x2) - 210 * x2 + 15) / 8;
    case 6: return (1386 * (x2 * x2 * xp) - 1260 * (x2 * xp) + 210 * xp - 5) / 16;
    case 7: return (3003 * (x2 * x2 * x2) - 3465 * (x2 * x2) + 945 * x2 - 35) / 16;
  }
  return 0.0;
}

/* Evaluate phi and grad at physical points pts[np][3] for bbox (xmin,xmax,...)
 * (LegendreP / GradLegendreP normalisation, basis :440-517). */
static void eval_basis_pts(int Np, const int* blist, const double* bbox, int npts,
                           const double* pts, double* phi, double* dphi) {
  double hb[3], mb[3], sq[3];
  for (int d = 0; d < 3; ++d) {
    hb[d] = (bbox[2 * d + 1] - bbox[2 * d]) / 2.0;
    mb[d] = (bbox[2 * d + 1] + bbox[2 * d]) / 2.0;
    sq[d] = sqrt(hb[d]);
  }
  for (int f = 0; f < Np; ++f) {
    for (int q = 0; q < npts; ++q) {
      double v[3], dv[3];
      for (int d = 0; d < 3; ++d) {
        double xp = (pts[q * 3 + d] - mb[d]) / hb[d];
        int k = blist[f * 3 + d];
        v[d] = leg(k, xp) / sq[d];
        dv[d] = dleg(k, xp) / (hb[d] * sq[d]);
      }
      phi[f * npts + q] = v[0] * v[1] * v[2];
      dphi[(0 * Np + f) * npts + q] = dv[0] * v[1] * v[2];
      dphi[(1 * Np + f) * npts + q] = v[0] * dv[1] * v[2];
      dphi[(2 * Np + f) * npts + q] = v[0] * v[1] * dv[2];
    }
  }
}

static void tet_points(int nq3, const double* Fk, const double* nodtet3,
                       double* pts) {
  for (int q = 0; q < nq3; ++q)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int l = 0; l < 4; ++l) s = s + Fk[j * 4 + l] * nodtet3[q * 4 + l];
      pts[q * 3 + j] = s;
    }
}

/* basis(): src/basis_function.f90:440-517 */
void o_basis(int p, const double* bbox, const double* Fk, double* phi,
             double* dphi) {
  int Np = o_np(p), n = p + 1, nq3 = n * n * n, nq2 = n * n;
  double* nodtet3 = malloc(sizeof(double) * 4 * nq3);
  double* weitet3 = malloc(sizeof(double) * nq3);
  double* nodtria2 = malloc(sizeof(double) * 4 * nq2);
  double* weitria2 = malloc(sizeof(double) * nq2);
  int* blist = malloc(sizeof(int) * 3 * Np);
  double* pts = malloc(sizeof(double) * 3 * nq3);
  o_quadrature(p, nodtet3, weitet3, nodtria2, weitria2);
  o_basis_list(p, blist);
  tet_points(nq3, Fk, nodtet3, pts);
  eval_basis_pts(Np, blist, bbox, nq3, pts, phi, dphi);
  free(nodtet3); free(weitet3); free(nodtria2); free(weitria2); free(blist);
  free(pts);
}

/* face quadrature points: basis_boundary (src/basis_function.f90:557-575) */
static void face_points(int nq2, int face, const double* Fk,
                        const double* nodtria2, double* pts) {
  double nm[4][4][4];
  node_maps(nm);
  double temp[3][4];
  for (int j = 0; j < 3; ++j)
    for (int q = 0; q < 4; ++q) {
      double s = 0.0;
      for (int l = 0; l < 4; ++l) s = s + Fk[j * 4 + l] * nm[face][l][q];
      temp[j][q] = s;
    }
  for (int j = 0; j < 3; ++j)
    for (int q = 0; q < nq2; ++q) {
      double s = 0.0;
      for (int l = 0; l < 4; ++l) s = s + temp[j][l] * nodtria2[q * 4 + l];
      pts[q * 3 + j] = s;
    }
}

/* face quadrature points as in MAKE_RHS_FACE (src/assemble_element.f90:321-328) */
static void face_points_rhs(int nq2, int face, const double* Fk,
                            const double* nodtria2, double* pts) {
  double nm[4][4][4];
  node_maps(nm);
  for (int q = 0; q < nq2; ++q)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k)
        for (int t = 0; t < 4; ++t)
          s = s + Fk[j * 4 + k] * nm[face][k][t] * nodtria2[q * 4 + t];
      pts[q * 3 + j] = s;
    }
}

void o_basis_boundary(int p, int face, const double* bbox1, const double* bbox2,
                      const double* Fk, double* phi_b, double* grad_b) {
  int Np = o_np(p), n = p + 1, nq3 = n * n * n, nq2 = n * n;
  double* nodtet3 = malloc(sizeof(double) * 4 * nq3);
  double* weitet3 = malloc(sizeof(double) * nq3);
  double* nodtria2 = malloc(sizeof(double) * 4 * nq2);
  double* weitria2 = malloc(sizeof(double) * nq2);
  int* blist = malloc(sizeof(int) * 3 * Np);
  double* pts = malloc(sizeof(double) * 3 * nq2);
  o_quadrature(p, nodtet3, weitet3, nodtria2, weitria2);
  o_basis_list(p, blist);
  face_points(nq2, face, Fk, nodtria2, pts);
  eval_basis_pts(Np, blist, bbox1, nq2, pts, phi_b, grad_b);
  if (bbox2)
    eval_basis_pts(Np, blist, bbox2, nq2, pts, phi_b + Np * nq2,
                   grad_b + 3 * Np * nq2);
  else {
    memset(phi_b + Np * nq2, 0, sizeof(double) * Np * nq2);
    memset(grad_b + 3 * Np * nq2, 0, sizeof(double) * 3 * Np * nq2);
  }
  free(nodtet3); free(weitet3); free(nodtria2); free(weitria2); free(blist);
  free(pts);
}

#define EIDX(i, j, m, n) ((((i)*3 + (j)) * Np + (m)) * Np + (n))

/* MAKE_STIFFNESS_VOLUME: src/assemble_element.f90:24-77 */
static void stiff_vol(int Np, int nq3, double Jdet, const double* weitet3,
                      double lam, double mu, const double* dphi, double* V) {
  memset(V, 0, sizeof(double) * 9 * Np * Np);
#define D(l, f, q) dphi[((l)*Np + (f)) * nq3 + (q)]
  for (int q = 0; q < nq3; ++q) {
    double wj = weitet3[q] * fabs(Jdet);
    for (int m = 0; m < Np; ++m)
      for (int n = 0; n < Np; ++n) {
        double t[3][3];
        double l2m = lam + 2 * mu;
        t[0][0] = l2m * D(0, m, q) * D(0, n, q) + mu * D(1, m, q) * D(1, n, q) +
                  mu * D(2, m, q) * D(2, n, q);
        t[1][1] = l2m * D(1, m, q) * D(1, n, q) + mu * D(0, m, q) * D(0, n, q) +
                  mu * D(2, m, q) * D(2, n, q);
        t[2][2] = l2m * D(2, m, q) * D(2, n, q) + mu * D(0, m, q) * D(0, n, q) +
                  mu * D(1, m, q) * D(1, n, q);
        for (int i = 0; i < 2; ++i)
          for (int j = i + 1; j < 3; ++j) {
            t[i][j] = lam * D(i, m, q) * D(j, n, q) + mu * D(j, m, q) * D(i, n, q);
            t[j][i] = lam * D(j, m, q) * D(i, n, q) + mu * D(i, m, q) * D(j, n, q);
          }
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            V[EIDX(i, j, m, n)] = V[EIDX(i, j, m, n)] + wj * t[i][j];
      }
  }
#undef D
}

void o_stiffness_volume(int p, double Jdet, double lam, double mu,
                        const double* dphi, double* V) {
  int Np = o_np(p), n = p + 1, nq3 = n * n * n, nq2 = n * n;
  double* nodtet3 = malloc(sizeof(double) * 4 * nq3);
  double* weitet3 = malloc(sizeof(double) * nq3);
  double* nodtria2 = malloc(sizeof(double) * 4 * nq2);
  double* weitria2 = malloc(sizeof(double) * nq2);
  o_quadrature(p, nodtet3, weitet3, nodtria2, weitria2);
  stiff_vol(Np, nq3, Jdet, weitet3, lam, mu, dphi, V);
  free(nodtet3); free(weitet3); free(nodtria2); free(weitria2);
}

/* MAKE_MASS_VOLUME (src/assemble_element.f90:81-113), one component block */
static void mass_vol(int Np, int nq3, double Jdet, const double* weitet3,
                     const double* phi, double rho, double* Mb) {
  memset(Mb, 0, sizeof(double) * Np * Np);
  for (int q = 0; q < nq3; ++q)
    for (int m = 0; m < Np; ++m)
      for (int n = 0; n < Np; ++n)
        Mb[m * Np + n] = Mb[m * Np + n] + rho * weitet3[q] * fabs(Jdet) *
                                              phi[m * nq3 + q] * phi[n * nq3 + q];
}

void o_mass_volume(int p, double Jdet, const double* phi, double rho,
                   double* Mfull) {
  int Np = o_np(p), n = p + 1, nq3 = n * n * n, nq2 = n * n;
  double* nodtet3 = malloc(sizeof(double) * 4 * nq3);
  double* weitet3 = malloc(sizeof(double) * nq3);
  double* nodtria2 = malloc(sizeof(double) * 4 * nq2);
  double* weitria2 = malloc(sizeof(double) * nq2);
  double* Mb = malloc(sizeof(double) * Np * Np);
  o_quadrature(p, nodtet3, weitet3, nodtria2, weitria2);
  mass_vol(Np, nq3, Jdet, weitet3, phi, rho, Mb);
  memset(Mfull, 0, sizeof(double) * 9 * Np * Np);
  for (int i = 0; i < 3; ++i)
    for (int m = 0; m < Np; ++m)
      for (int nn = 0; nn < Np; ++nn) Mfull[EIDX(i, i, m, nn)] = Mb[m * Np + nn];
  free(nodtet3); free(weitet3); free(nodtria2); free(weitria2); free(Mb);
}

/* MAKE_STIFFNESS_FACE: src/assemble_element.f90:404-661 (absorbing branch
 * omitted: never reached by the solver, src/matrix_free.f90:279-287).
 * kind: -1 Dirichlet, -2 Neumann, otherwise interior. */
static void stiff_face(int Np, int nq2, int p, double alpha, int kind,
                       double hk1, double hk2, const double* nrm, double area,
                       const double* weitria2, double lam, double mu,
                       const double* phi_b, const double* grad_b, double* S1,
                       double* I1, double* S2, double* I2, double* IT2) {
  size_t sz = sizeof(double) * 9 * Np * Np;
  memset(S1, 0, sz); memset(I1, 0, sz); memset(S2, 0, sz); memset(I2, 0, sz);
  memset(IT2, 0, sz);
  if (kind == -2) return;
  double Dbar = lam + 2 * mu;
  double sigma;
  if (kind == -1) sigma = alpha * (p * p) / hk1 * Dbar;
  else sigma = alpha * (p * p) / (hk1 < hk2 ? hk1 : hk2) * Dbar;
  const double* ph1 = phi_b;
  const double* ph2 = phi_b + Np * nq2;
  const double* g1 = grad_b;
  const double* g2 = grad_b + 3 * Np * nq2;
#define P1(f, q) ph1[(f)*nq2 + (q)]
#define P2(f, q) ph2[(f)*nq2 + (q)]
#define G1(l, f, q) g1[((l)*Np + (f)) * nq2 + (q)]
#define G2(l, f, q) g2[((l)*Np + (f)) * nq2 + (q)]
  double n1 = nrm[0], n2 = nrm[1], n3 = nrm[2];
  double d11 = fabs(n1) * fabs(n1) + 0.5 * (fabs(n2) * fabs(n2)) + 0.5 * (fabs(n3) * fabs(n3));
  double d22 = fabs(n2) * fabs(n2) + 0.5 * (fabs(n1) * fabs(n1)) + 0.5 * (fabs(n3) * fabs(n3));
  double d33 = fabs(n3) * fabs(n3) + 0.5 * (fabs(n1) * fabs(n1)) + 0.5 * (fabs(n2) * fabs(n2));
  double l2m = lam + 2 * mu;
  for (int q = 0; q < nq2; ++q) {
    double w = weitria2[q];
    for (int m = 0; m < Np; ++m)
      for (int n = 0; n < Np; ++n) {
        double t[3][3];
        double pp = P1(m, q) * P1(n, q);
        t[0][0] = pp * d11;
        t[1][1] = pp * d22;
        t[2][2] = pp * d33;
        for (int i = 0; i < 2; ++i)
          for (int j = i + 1; j < 3; ++j) {
            t[i][j] = 0.5 * P1(m, q) * P1(n, q) * nrm[i] * nrm[j];
            t[j][i] = 0.5 * P1(m, q) * P1(n, q) * nrm[i] * nrm[j];
          }
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            S1[EIDX(i, j, m, n)] = S1[EIDX(i, j, m, n)] + sigma * w * t[i][j] * area;
      }
    if (kind == -1) {
      for (int m = 0; m < Np; ++m)
        for (int n = 0; n < Np; ++n) {
          double t[3][3];
          double ph = P1(n, q);
          t[0][0] = l2m * G1(0, m, q) * ph * n1 + mu * G1(1, m, q) * ph * n2 +
                    mu * G1(2, m, q) * ph * n3;
          t[1][1] = l2m * G1(1, m, q) * ph * n2 + mu * G1(0, m, q) * ph * n1 +
                    mu * G1(2, m, q) * ph * n3;
          t[2][2] = l2m * G1(2, m, q) * ph * n3 + mu * G1(1, m, q) * ph * n2 +
                    mu * G1(0, m, q) * ph * n1;
          for (int i = 0; i < 2; ++i)
            for (int j = i + 1; j < 3; ++j) {
              t[i][j] = lam * G1(i, m, q) * ph * nrm[j] + mu * G1(j, m, q) * ph * nrm[i];
              t[j][i] = lam * G1(j, m, q) * ph * nrm[i] + mu * G1(i, m, q) * ph * nrm[j];
            }
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              I1[EIDX(i, j, m, n)] = I1[EIDX(i, j, m, n)] + w * area * t[i][j];
        }
    } else {
      for (int m = 0; m < Np; ++m)
        for (int n = 0; n < Np; ++n) {
          double t[3][3], t1[3][3], t2[3][3], t3[3][3];
          double ph = P1(n, q), phn2 = P2(n, q);
          /* I_E1 */
          t[0][0] = l2m * G1(0, m, q) * ph * n1 + mu * G1(1, m, q) * ph * n2 +
                    mu * G1(2, m, q) * ph * n3;
          t[1][1] = l2m * G1(1, m, q) * ph * n2 + mu * G1(0, m, q) * ph * n1 +
                    mu * G1(2, m, q) * ph * n3;
          t[2][2] = l2m * G1(2, m, q) * ph * n3 + mu * G1(1, m, q) * ph * n2 +
                    mu * G1(0, m, q) * ph * n1;
          for (int i = 0; i < 2; ++i)
            for (int j = i + 1; j < 3; ++j) {
              t[i][j] = lam * G1(i, m, q) * ph * nrm[j] + mu * G1(j, m, q) * ph * nrm[i];
              t[j][i] = lam * G1(j, m, q) * ph * nrm[i] + mu * G1(i, m, q) * ph * nrm[j];
            }
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              I1[EIDX(i, j, m, n)] = I1[EIDX(i, j, m, n)] + 0.5 * w * area * t[i][j];
          /* I_E2 */
          t1[0][0] = l2m * G1(0, m, q) * phn2 * n1 + mu * G1(1, m, q) * phn2 * n2 +
                     mu * G1(2, m, q) * phn2 * n3;
          t1[1][1] = l2m * G1(1, m, q) * phn2 * n2 + mu * G1(0, m, q) * phn2 * n1 +
                     mu * G1(2, m, q) * phn2 * n3;
          t1[2][2] = l2m * G1(2, m, q) * phn2 * n3 + mu * G1(1, m, q) * phn2 * n2 +
                     mu * G1(0, m, q) * phn2 * n1;
          for (int i = 0; i < 2; ++i)
            for (int j = i + 1; j < 3; ++j) {
              t1[i][j] = lam * G1(i, m, q) * phn2 * nrm[j] + mu * G1(j, m, q) * phn2 * nrm[i];
              t1[j][i] = lam * G1(j, m, q) * phn2 * nrm[i] + mu * G1(i, m, q) * phn2 * nrm[j];
            }
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              I2[EIDX(i, j, m, n)] = I2[EIDX(i, j, m, n)] - 0.5 * w * area * t1[i][j];
          /* IT_E2 (stored transposed in (m,n)) */
          t2[0][0] = l2m * G2(0, m, q) * ph * n1 + mu * G2(1, m, q) * ph * n2 +
                     mu * G2(2, m, q) * ph * n3;
          t2[1][1] = l2m * G2(1, m, q) * ph * n2 + mu * G2(0, m, q) * ph * n1 +
                     mu * G2(2, m, q) * ph * n3;
          t2[2][2] = l2m * G2(2, m, q) * ph * n3 + mu * G2(1, m, q) * ph * n2 +
                     mu * G2(0, m, q) * ph * n1;
          for (int i = 0; i < 2; ++i)
            for (int j = i + 1; j < 3; ++j) {
              t2[i][j] = lam * G2(i, m, q) * ph * nrm[j] + mu * G2(j, m, q) * ph * nrm[i];
              t2[j][i] = lam * G2(j, m, q) * ph * nrm[i] + mu * G2(i, m, q) * ph * nrm[j];
            }
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              IT2[EIDX(i, j, n, m)] = IT2[EIDX(i, j, n, m)] + 0.5 * w * area * t2[i][j];
          /* S_E2 */
          double pq = P1(m, q) * P2(n, q);
          t3[0][0] = pq * d11;
          t3[1][1] = pq * d22;
          t3[2][2] = pq * d33;
          for (int i = 0; i < 2; ++i)
            for (int j = i + 1; j < 3; ++j) {
              t3[i][j] = 0.5 * P1(m, q) * P2(n, q) * nrm[i] * nrm[j];
              t3[j][i] = 0.5 * P1(m, q) * P2(n, q) * nrm[i] * nrm[j];
            }
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              S2[EIDX(i, j, m, n)] = S2[EIDX(i, j, m, n)] - sigma * w * t3[i][j] * area;
        }
    }
  }
#undef P1
#undef P2
#undef G1
#undef G2
}

void o_stiffness_face(int p, double alpha, int kind, double hk1, double hk2,
                      const double* normal, double area, double lam, double mu,
                      const double* phi_b, const double* grad_b, double* S1,
                      double* I1, double* S2, double* I2, double* IT2) {
  int Np = o_np(p), n = p + 1, nq3 = n * n * n, nq2 = n * n;
  double* nodtet3 = malloc(sizeof(double) * 4 * nq3);
  double* weitet3 = malloc(sizeof(double) * nq3);
  double* nodtria2 = malloc(sizeof(double) * 4 * nq2);
  double* weitria2 = malloc(sizeof(double) * nq2);
  o_quadrature(p, nodtet3, weitet3, nodtria2, weitria2);
  stiff_face(Np, nq2, p, alpha, kind, hk1, hk2, normal, area, weitria2, lam, mu,
             phi_b, grad_b, S1, I1, S2, I2, IT2);
  free(nodtet3); free(weitet3); free(nodtria2); free(weitria2);
}

/* cholesky: src/vet_mat_operations.f90:84-120.  A and R row-major n x n,
 * R upper with A = R^T R.  Returns 5 (EXIT_SINGULARMTX) if not SPD. */
int o_cholesky(int n, const double* A, double* R) {
  memset(R, 0, sizeof(double) * n * n);
  for (int j = 0; j < n; ++j) {
    double s = A[j * n + j];
    for (int k = 0; k < j; ++k) s = s - R[k * n + j] * R[k * n + j];
    if (s <= 0.0) return 5;
    R[j * n + j] = sqrt(s);
    for (int i = j + 1; i < n; ++i) {
      s = A[j * n + i];
      for (int k = 0; k < j; ++k) s = s - R[k * n + j] * R[k * n + i];
      R[j * n + i] = s / R[j * n + j];
    }
  }
  return 0;
}

/* solve_LU(R^T, R, b): src/vet_mat_operations.f90:123-153 */
void o_solve_lu(int n, const double* R, const double* b, double* x) {
  double y[84];
  y[0] = b[0] / R[0];
  for (int i = 1; i < n; ++i) {
    double s = b[i];
    for (int j = 0; j < i; ++j) s = s - R[j * n + i] * y[j];
    y[i] = s / R[i * n + i];
  }
  x[n - 1] = y[n - 1] / R[(n - 1) * n + (n - 1)];
  for (int i = n - 2; i >= 0; --i) {
    double s = y[i];
    for (int j = i + 1; j < n; ++j) s = s - R[i * n + j] * x[j];
    x[i] = s / R[i * n + i];
  }
}

/* ---------------------- problem data ------------------------------------ */
double o_time_function(int problem, double t) {
  if (problem == O_PROB_CUBE_WAVE) return cos(sqrt(3.0) * o_pi() * t);
  return sin(o_sqrt2() * o_pi() * t); /* time_function, problem_data :14-23 */
}

/* forcing: f_null / f_time (problem_data :48-81) / cube wave */
static void forcing(int problem, double lam, double mu, const double* p,
                    double rho, double* r) {
  double PI = o_pi();
  if (problem == O_PROB_REF_MMS) {
    double c = (-2.0 * rho + 3.0 * (lam + 2.0 * mu));
    r[0] = c * PI * PI * cos(PI * p[0]) * sin(PI * p[1]) * sin(PI * p[2]);
    r[1] = c * PI * PI * sin(PI * p[0]) * cos(PI * p[1]) * sin(PI * p[2]);
    r[2] = c * PI * PI * sin(PI * p[0]) * sin(PI * p[1]) * cos(PI * p[2]);
  } else if (problem == O_PROB_CUBE_WAVE) {
    double sx = sin(PI * p[0]), sy = sin(PI * p[1]), sz = sin(PI * p[2]);
    double cx = cos(PI * p[0]), cy = cos(PI * p[1]), cz = cos(PI * p[2]);
    double S = sx * sy * sz;
    double a = (-3.0 * rho + 3.0 * mu + (lam + mu)) * S;
    double lm = lam + mu;
    r[0] = PI * PI * (a - lm * (cx * cy * sz + cx * sy * cz));
    r[1] = PI * PI * (a - lm * (cx * cy * sz + sx * cy * cz));
    r[2] = PI * PI * (a - lm * (cx * sy * cz + sx * cy * cz));
  } else {
    r[0] = r[1] = r[2] = 0.0;
  }
}

/* Dirichlet data: gd_null / gd (problem_data :130-155) / cube wave */
static void dirichlet(int problem, const double* p, int tag, double* r) {
  double PI = o_pi();
  r[0] = r[1] = r[2] = 0.0;
  if (problem == O_PROB_REF_MMS) {
    if (tag >= 1 && tag <= 3) {
      r[0] = cos(PI * p[0]) * sin(PI * p[1]) * sin(PI * p[2]);
      r[1] = sin(PI * p[0]) * cos(PI * p[1]) * sin(PI * p[2]);
      r[2] = sin(PI * p[0]) * sin(PI * p[1]) * cos(PI * p[2]);
    }
  } else if (problem == O_PROB_CUBE_WAVE) {
    double S = sin(PI * p[0]) * sin(PI * p[1]) * sin(PI * p[2]);
    r[0] = r[1] = r[2] = S;
  }
}

/* Neumann data: gn_null / gn (problem_data :186-243) / cube wave */
static void neumann(int problem, double lam, double mu, const double* n,
                    const double* p, int tag, double* r) {
  double PI = o_pi();
  double st[3][3];
  r[0] = r[1] = r[2] = 0.0;
  if (problem == O_PROB_REF_MMS) {
    if (tag < 1 || tag > 3) return;
    st[0][0] = -PI * sin(PI * p[0]) * sin(PI * p[1]) * sin(PI * p[2]) * (3.0 * lam + 2.0 * mu);
    st[1][1] = st[0][0];
    st[2][2] = st[0][0];
    st[0][1] = 2.0 * mu * PI * cos(PI * p[0]) * cos(PI * p[1]) * sin(PI * p[2]);
    st[1][0] = st[0][1];
    st[0][2] = 2.0 * mu * PI * cos(PI * p[0]) * cos(PI * p[2]) * sin(PI * p[1]);
    st[2][0] = st[0][2];
    st[1][2] = 2.0 * mu * PI * cos(PI * p[1]) * cos(PI * p[2]) * sin(PI * p[0]);
    st[2][1] = st[1][2];
  } else if (problem == O_PROB_CUBE_WAVE) {
    double sx = sin(PI * p[0]), sy = sin(PI * p[1]), sz = sin(PI * p[2]);
    double cx = cos(PI * p[0]), cy = cos(PI * p[1]), cz = cos(PI * p[2]);
    double g[3] = {PI * cx * sy * sz, PI * sx * cy * sz, PI * sx * sy * cz};
    double div = g[0] + g[1] + g[2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        st[i][j] = (i == j ? lam * div : 0.0) + mu * (g[j] + g[i]);
  } else {
    return;
  }
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int j = 0; j < 3; ++j) s = s + st[i][j] * n[j];
    r[i] = s;
  }
}

/* fields for projection */
static void field(const OCtx* c, int fn, const double* p, double* r) {
  double PI = o_pi();
  r[0] = r[1] = r[2] = 0.0;
  if (fn == O_FN_POLY) {
    for (int k = 0; k < c->poly_n; ++k) {
      double m = 1.0;
      for (int d = 0; d < 3; ++d)
        for (int e = 0; e < c->poly_exp[k * 3 + d]; ++e) m = m * p[d];
      for (int i = 0; i < 3; ++i) r[i] = r[i] + c->poly_coef[k * 3 + i] * m;
    }
    return;
  }
  if (fn == O_FN_REF_VEL) {
    double c = o_sqrt2() * PI;
    r[0] = c * cos(PI * p[0]) * sin(PI * p[1]) * sin(PI * p[2]);
    r[1] = c * sin(PI * p[0]) * cos(PI * p[1]) * sin(PI * p[2]);
    r[2] = c * sin(PI * p[0]) * sin(PI * p[1]) * cos(PI * p[2]);
  } else if (fn == O_FN_REF_UEX) {
    r[0] = cos(PI * p[0]) * sin(PI * p[1]) * sin(PI * p[2]);
    r[1] = sin(PI * p[0]) * cos(PI * p[1]) * sin(PI * p[2]);
    r[2] = sin(PI * p[0]) * sin(PI * p[1]) * cos(PI * p[2]);
  } else if (fn == O_FN_CUBE_U) {
    double S = sin(PI * p[0]) * sin(PI * p[1]) * sin(PI * p[2]);
    r[0] = r[1] = r[2] = S;
  }
}

/* ------------------------------ mesh ------------------------------------ */
typedef struct {
  int v[3];
  int tet, face;
} FaceKey;

static int cmp_face(const void* a, const void* b) {
  const FaceKey* x = a;
  const FaceKey* y = b;
  for (int i = 0; i < 3; ++i)
    if (x->v[i] != y->v[i]) return x->v[i] < y->v[i] ? -1 : 1;
  if (x->tet != y->tet) return x->tet < y->tet ? -1 : 1;
  return x->face - y->face;
}

static void sort3(int* v) {
  int t;
  if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
  if (v[1] > v[2]) { t = v[1]; v[1] = v[2]; v[2] = t; }
  if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
}

/* face vertex triplets for normals (CREATE_LOCAL_MESH,
 * src/MAKE_PARTITION_AND_MPI_FILES.f90:1087-1095) */
static const int FTRI[4][3] = {{2, 1, 0}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}};

void o_ctx_destroy(OCtx* c) {
  if (!c) return;
  free(c->xyz); free(c->tet_vert); free(c->tet_poly); free(c->rho); free(c->lam);
  free(c->mu); free(c->neigh); free(c->neigh_face); free(c->bc_space);
  free(c->normal); free(c->area); free(c->bbox); free(c->hk); free(c->ptet_ptr);
  free(c->ptet); free(c->row_ptr); free(c->col); free(c->face_blk);
  free(c->nodtet3); free(c->weitet3); free(c->nodtria2); free(c->weitria2);
  free(c->blist); free(c->K); free(c->A); free(c->M); free(c->Mm); free(c->RM);
  free(c->RMm);
  free(c);
}

int o_ctx_create(int n_node, const double* xyz, int n_tet, const int* tet_vert,
                 const int* tet_poly, int n_tria, const int* tria_vert,
                 const int* tria_tag, int n_tags, const int* tag_kind,
                 const int* tag_space, const double* rho, const double* lam,
                 const double* mu, int p, double alpha, double theta,
                 OCtx** out) {
  *out = NULL;
  if (p < 0 || p > 6 || n_tet <= 0) return 4;
  OCtx* c = calloc(1, sizeof(OCtx));
  c->p = p;
  c->Np = o_np(p);
  c->n3 = 3 * c->Np;
  c->nq3 = (p + 1) * (p + 1) * (p + 1);
  c->nq2 = (p + 1) * (p + 1);
  c->alpha = alpha;
  c->theta = theta;
  c->n_node = n_node;
  c->n_tet = n_tet;
  c->xyz = malloc(sizeof(double) * 3 * n_node);
  memcpy(c->xyz, xyz, sizeof(double) * 3 * n_node);
  c->tet_vert = malloc(sizeof(int) * 4 * n_tet);
  memcpy(c->tet_vert, tet_vert, sizeof(int) * 4 * n_tet);
  c->tet_poly = malloc(sizeof(int) * n_tet);
  int np = 0;
  for (int t = 0; t < n_tet; ++t) {
    c->tet_poly[t] = tet_poly ? tet_poly[t] : t;
    if (c->tet_poly[t] + 1 > np) np = c->tet_poly[t] + 1;
  }
  c->n_poly = np;
  c->rho = malloc(sizeof(double) * n_tet);
  c->lam = malloc(sizeof(double) * n_tet);
  c->mu = malloc(sizeof(double) * n_tet);
  memcpy(c->rho, rho, sizeof(double) * n_tet);
  memcpy(c->lam, lam, sizeof(double) * n_tet);
  memcpy(c->mu, mu, sizeof(double) * n_tet);

  /* poly -> tets, ascending (FIND_TET_IN_POLY order) */
  c->ptet_ptr = calloc(np + 1, sizeof(int));
  c->ptet = malloc(sizeof(int) * n_tet);
  for (int t = 0; t < n_tet; ++t) c->ptet_ptr[c->tet_poly[t] + 1]++;
  for (int i = 0; i < np; ++i) c->ptet_ptr[i + 1] += c->ptet_ptr[i];
  {
    int* fill = calloc(np, sizeof(int));
    for (int t = 0; t < n_tet; ++t) {
      int P = c->tet_poly[t];
      c->ptet[c->ptet_ptr[P] + fill[P]++] = t;
    }
    free(fill);
  }
  for (int P = 0; P < np; ++P)
    if (c->ptet_ptr[P + 1] == c->ptet_ptr[P]) { o_ctx_destroy(c); return 4; }

  /* face matching (CREATE_NEIGH_EL_TRIA :2030-2145): interior pairs first,
   * then boundary triangles of the mesh file, tag -> BC kind. */
  c->neigh = malloc(sizeof(int) * 4 * n_tet);
  c->neigh_face = malloc(sizeof(int) * 4 * n_tet);
  c->bc_space = calloc(4 * n_tet, sizeof(int));
  FaceKey* fk = malloc(sizeof(FaceKey) * 4 * n_tet);
  static const int FSET[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}};
  for (int t = 0; t < n_tet; ++t)
    for (int f = 0; f < 4; ++f) {
      FaceKey* k = &fk[t * 4 + f];
      for (int i = 0; i < 3; ++i) k->v[i] = tet_vert[t * 4 + FSET[f][i]];
      sort3(k->v);
      k->tet = t;
      k->face = f;
      c->neigh[t * 4 + f] = -100;
      c->neigh_face[t * 4 + f] = -1;
    }
  qsort(fk, 4 * (size_t)n_tet, sizeof(FaceKey), cmp_face);
  for (int i = 0; i + 1 < 4 * n_tet; ++i) {
    FaceKey *a = &fk[i], *b = &fk[i + 1];
    if (a->v[0] == b->v[0] && a->v[1] == b->v[1] && a->v[2] == b->v[2]) {
      c->neigh[a->tet * 4 + a->face] = b->tet;
      c->neigh_face[a->tet * 4 + a->face] = b->face;
      c->neigh[b->tet * 4 + b->face] = a->tet;
      c->neigh_face[b->tet * 4 + b->face] = a->face;
    }
  }
  /* boundary */
  FaceKey* bk = malloc(sizeof(FaceKey) * (n_tria > 0 ? n_tria : 1));
  for (int i = 0; i < n_tria; ++i) {
    for (int j = 0; j < 3; ++j) bk[i].v[j] = tria_vert[i * 3 + j];
    sort3(bk[i].v);
    bk[i].tet = tria_tag[i];
    bk[i].face = i;
  }
  qsort(bk, n_tria, sizeof(FaceKey), cmp_face);
  int err = 0;
  for (int t = 0; t < n_tet && !err; ++t)
    for (int f = 0; f < 4; ++f) {
      if (c->neigh[t * 4 + f] != -100) continue;
      FaceKey key;
      for (int i = 0; i < 3; ++i) key.v[i] = tet_vert[t * 4 + FSET[f][i]];
      sort3(key.v);
      int lo = 0, hi = n_tria - 1, found = -1;
      while (lo <= hi) {
        int mid = (lo + hi) / 2;
        FaceKey* m = &bk[mid];
        int cmpv = 0;
        for (int i = 0; i < 3 && !cmpv; ++i)
          cmpv = (key.v[i] > m->v[i]) - (key.v[i] < m->v[i]);
        if (cmpv == 0) { found = mid; break; }
        if (cmpv < 0) hi = mid - 1; else lo = mid + 1;
      }
      if (found < 0) { err = 18; break; } /* EXIT_NEIGHBOUR_EL_ERROR */
      int tag = bk[found].tet;
      int kind = (tag >= 0 && tag < n_tags) ? tag_kind[tag] : 0;
      if (kind != -1 && kind != -2) { err = 18; break; }
      c->neigh[t * 4 + f] = kind;
      c->neigh_face[t * 4 + f] = f;
      c->bc_space[t * 4 + f] = tag_space[tag];
    }
  free(fk);
  free(bk);
  if (err) { o_ctx_destroy(c); return err; }

  /* normals, areas (CREATE_NORMAL_FACE :1429-1504) */
  c->normal = malloc(sizeof(double) * 12 * n_tet);
  c->area = malloc(sizeof(double) * 4 * n_tet);
  for (int t = 0; t < n_tet; ++t)
    for (int f = 0; f < 4; ++f) {
      const double* A = &xyz[3 * tet_vert[t * 4 + FTRI[f][0]]];
      const double* B = &xyz[3 * tet_vert[t * 4 + FTRI[f][1]]];
      const double* C = &xyz[3 * tet_vert[t * 4 + FTRI[f][2]]];
      double Qx = A[0] - B[0], Qy = A[1] - B[1], Qz = A[2] - B[2];
      double Px = C[0] - B[0], Py = C[1] - B[1], Pz = C[2] - B[2];
      double n1 = Py * Qz - Pz * Qy;
      double n2 = -(Px * Qz) + Pz * Qx;
      double n3 = Px * Qy - Py * Qx;
      double nn = sqrt(n1 * n1 + n2 * n2 + n3 * n3);
      c->normal[(t * 4 + f) * 3 + 0] = n1 / nn;
      c->normal[(t * 4 + f) * 3 + 1] = n2 / nn;
      c->normal[(t * 4 + f) * 3 + 2] = n3 / nn;
      c->area[t * 4 + f] = nn;
    }

  /* bounding boxes and diameters (CREATE_BBOX_EL :1510-1580) */
  c->bbox = malloc(sizeof(double) * 6 * np);
  c->hk = malloc(sizeof(double) * np);
  for (int P = 0; P < np; ++P) {
    int nt = c->ptet_ptr[P + 1] - c->ptet_ptr[P];
    int nv = 4 * nt;
    double* vx = malloc(sizeof(double) * 3 * nv);
    int k = 0;
    for (int a = c->ptet_ptr[P]; a < c->ptet_ptr[P + 1]; ++a)
      for (int v = 0; v < 4; ++v, ++k)
        for (int d = 0; d < 3; ++d)
          vx[k * 3 + d] = xyz[3 * tet_vert[c->ptet[a] * 4 + v] + d];
    for (int d = 0; d < 3; ++d) {
      double mn = vx[d], mx = vx[d];
      for (int i = 1; i < nv; ++i) {
        if (vx[i * 3 + d] < mn) mn = vx[i * 3 + d];
        if (vx[i * 3 + d] > mx) mx = vx[i * 3 + d];
      }
      c->bbox[P * 6 + 2 * d] = mn;
      c->bbox[P * 6 + 2 * d + 1] = mx;
    }
    double h = 0.0;
    for (int i = 0; i < nv; ++i)
      for (int l = i + 1; l < nv; ++l) {
        double dx = vx[i * 3] - vx[l * 3], dy = vx[i * 3 + 1] - vx[l * 3 + 1],
               dz = vx[i * 3 + 2] - vx[l * 3 + 2];
        double dist = sqrt(dx * dx + dy * dy + dz * dz);
        if (dist > h) h = dist;
      }
    c->hk[P] = h;
    free(vx);
  }

  /* block pattern: self first, then neighbouring polyhedra in first-appearance
   * order over (tet ascending, face 1..4). */
  c->row_ptr = malloc(sizeof(int) * (np + 1));
  c->face_blk = malloc(sizeof(int) * 4 * n_tet);
  int cap = np * 8 + 16, nb = 0;
  c->col = malloc(sizeof(int) * cap);
  for (int P = 0; P < np; ++P) {
    c->row_ptr[P] = nb;
    if (nb + 1 > cap) { cap *= 2; c->col = realloc(c->col, sizeof(int) * cap); }
    c->col[nb++] = P;
    for (int a = c->ptet_ptr[P]; a < c->ptet_ptr[P + 1]; ++a) {
      int t = c->ptet[a];
      for (int f = 0; f < 4; ++f) {
        int e2 = c->neigh[t * 4 + f];
        c->face_blk[t * 4 + f] = -1;
        if (e2 < 0) continue;
        int Q = c->tet_poly[e2];
        if (Q == P) continue;
        int found = -1;
        for (int b = c->row_ptr[P] + 1; b < nb; ++b)
          if (c->col[b] == Q) { found = b; break; }
        if (found < 0) {
          if (nb + 1 > cap) { cap *= 2; c->col = realloc(c->col, sizeof(int) * cap); }
          c->col[nb] = Q;
          found = nb++;
        }
        c->face_blk[t * 4 + f] = found;
      }
    }
  }
  c->row_ptr[np] = nb;
  c->n_blk = nb;

  c->nodtet3 = malloc(sizeof(double) * 4 * c->nq3);
  c->weitet3 = malloc(sizeof(double) * c->nq3);
  c->nodtria2 = malloc(sizeof(double) * 4 * c->nq2);
  c->weitria2 = malloc(sizeof(double) * c->nq2);
  o_quadrature(p, c->nodtet3, c->weitet3, c->nodtria2, c->weitria2);
  c->blist = malloc(sizeof(int) * 3 * c->Np);
  o_basis_list(p, c->blist);
  *out = c;
  return 0;
}

void o_sizes(const OCtx* c, int* o) {
  o[0] = c->n_tet; o[1] = c->n_poly; o[2] = c->Np; o[3] = c->n_blk;
}

void o_get_conn(const OCtx* c, int* neigh, int* neigh_face, int* bc_space,
                double* normal, double* area, double* bbox, double* hk) {
  int nt = c->n_tet, np = c->n_poly;
  if (neigh) memcpy(neigh, c->neigh, sizeof(int) * 4 * nt);
  if (neigh_face) memcpy(neigh_face, c->neigh_face, sizeof(int) * 4 * nt);
  if (bc_space) memcpy(bc_space, c->bc_space, sizeof(int) * 4 * nt);
  if (normal) memcpy(normal, c->normal, sizeof(double) * 12 * nt);
  if (area) memcpy(area, c->area, sizeof(double) * 4 * nt);
  if (bbox) memcpy(bbox, c->bbox, sizeof(double) * 6 * np);
  if (hk) memcpy(hk, c->hk, sizeof(double) * np);
}

void o_get_pattern(const OCtx* c, int* row_ptr, int* col) {
  memcpy(row_ptr, c->row_ptr, sizeof(int) * (c->n_poly + 1));
  memcpy(col, c->col, sizeof(int) * c->n_blk);
}

/* compute_hmax: src/solution_processing.f90:197-214 */
double o_hmax(const OCtx* c) {
  double h = c->hk[0];
  for (int P = 1; P < c->n_poly; ++P)
    if (c->hk[P] > h) h = c->hk[P];
  return h;
}

static void tet_xyz(const OCtx* c, int t, double* x, double* y, double* z) {
  for (int v = 0; v < 4; ++v) {
    const double* X = &c->xyz[3 * c->tet_vert[t * 4 + v]];
    x[v] = X[0]; y[v] = X[1]; z[v] = X[2];
  }
}

/* MAKE_MATRICES_FREE (src/matrix_free.f90:27-464) + Cholesky of the mass
 * blocks (src/Lymph3D.f90:264-271), one block row per polyhedron. */
int o_build_operator(OCtx* c) {
  int Np = c->Np, n3 = c->n3, nq3 = c->nq3, nq2 = c->nq2, np = c->n_poly;
  size_t bs = (size_t)n3 * n3;
  free(c->K); free(c->A); free(c->M); free(c->Mm); free(c->RM); free(c->RMm);
  c->K = calloc((size_t)c->n_blk * bs, sizeof(double));
  c->A = calloc((size_t)c->n_blk * bs, sizeof(double));
  c->M = calloc((size_t)np * Np * Np, sizeof(double));
  c->Mm = calloc((size_t)np * Np * Np, sizeof(double));
  c->RM = calloc((size_t)np * Np * Np, sizeof(double));
  c->RMm = calloc((size_t)np * Np * Np, sizeof(double));
  int err = 0;
#pragma omp parallel
  {
    double* phi = malloc(sizeof(double) * Np * nq3);
    double* dphi = malloc(sizeof(double) * 3 * Np * nq3);
    double* pts = malloc(sizeof(double) * 3 * nq3);
    double* mb = malloc(sizeof(double) * Np * Np);
    double* V = malloc(sizeof(double) * 9 * Np * Np);
    double* phib = malloc(sizeof(double) * 2 * Np * nq2);
    double* gradb = malloc(sizeof(double) * 6 * Np * nq2);
    double* fpts = malloc(sizeof(double) * 3 * nq2);
    double *S1 = malloc(sizeof(double) * 9 * Np * Np),
           *I1 = malloc(sizeof(double) * 9 * Np * Np),
           *S2 = malloc(sizeof(double) * 9 * Np * Np),
           *I2 = malloc(sizeof(double) * 9 * Np * Np),
           *IT2 = malloc(sizeof(double) * 9 * Np * Np);
#pragma omp for schedule(dynamic, 16)
    for (int P = 0; P < np; ++P) {
      double* Ks = c->K + (size_t)c->row_ptr[P] * bs;
      double* As = c->A + (size_t)c->row_ptr[P] * bs;
      double* MP = c->M + (size_t)P * Np * Np;
      double* MmP = c->Mm + (size_t)P * Np * Np;
      for (int a = c->ptet_ptr[P]; a < c->ptet_ptr[P + 1]; ++a) {
        int t = c->ptet[a];
        double x[4], y[4], z[4], Fk[12], Jdet;
        tet_xyz(c, t, x, y, z);
        o_jacobians(x, y, z, Fk, &Jdet);
        tet_points(nq3, Fk, c->nodtet3, pts);
        eval_basis_pts(Np, c->blist, &c->bbox[P * 6], nq3, pts, phi, dphi);
        mass_vol(Np, nq3, Jdet, c->weitet3, phi, 1.0, mb);
        for (int i = 0; i < Np * Np; ++i) {
          MmP[i] = MmP[i] + mb[i];
          MP[i] = MP[i] + mb[i] * c->rho[t];
        }
        stiff_vol(Np, nq3, Jdet, c->weitet3, c->lam[t], c->mu[t], dphi, V);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int m = 0; m < Np; ++m)
              for (int n = 0; n < Np; ++n) {
                size_t r = (size_t)(i * Np + m) * n3 + (j * Np + n);
                Ks[r] = Ks[r] + V[EIDX(i, j, m, n)];
                As[r] = As[r] + V[EIDX(i, j, m, n)];
              }
        for (int f = 0; f < 4; ++f) {
          int e2 = c->neigh[t * 4 + f];
          int Q = -1;
          if (e2 >= 0) {
            Q = c->tet_poly[e2];
            if (Q == P) continue; /* intra-polyhedron face */
          }
          face_points(nq2, f, Fk, c->nodtria2, fpts);
          eval_basis_pts(Np, c->blist, &c->bbox[P * 6], nq2, fpts, phib, gradb);
          double hk2 = c->hk[P];
          if (e2 >= 0) {
            eval_basis_pts(Np, c->blist, &c->bbox[Q * 6], nq2, fpts,
                           phib + Np * nq2, gradb + 3 * Np * nq2);
            hk2 = c->hk[Q];
          } else {
            memset(phib + Np * nq2, 0, sizeof(double) * Np * nq2);
            memset(gradb + 3 * Np * nq2, 0, sizeof(double) * 3 * Np * nq2);
          }
          stiff_face(Np, nq2, c->p, c->alpha, e2, c->hk[P], hk2,
                     &c->normal[(t * 4 + f) * 3], c->area[t * 4 + f], c->weitria2,
                     c->lam[t], c->mu[t], phib, gradb, S1, I1, S2, I2, IT2);
          if (e2 == -2) continue;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              for (int m = 0; m < Np; ++m)
                for (int n = 0; n < Np; ++n) {
                  size_t r = (size_t)(i * Np + m) * n3 + (j * Np + n);
                  Ks[r] = Ks[r] + S1[EIDX(i, j, m, n)];
                  As[r] = As[r] + S1[EIDX(i, j, m, n)];
                }
          double th = c->theta;
          if (e2 == -1) {
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j)
                for (int m = 0; m < Np; ++m)
                  for (int n = 0; n < Np; ++n) {
                    size_t r = (size_t)(i * Np + m) * n3 + (j * Np + n);
                    Ks[r] = Ks[r] + th * I1[EIDX(i, j, m, n)] - I1[EIDX(j, i, n, m)];
                  }
          } else {
            double* Kb = c->K + (size_t)c->face_blk[t * 4 + f] * bs;
            double* Ab = c->A + (size_t)c->face_blk[t * 4 + f] * bs;
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j)
                for (int m = 0; m < Np; ++m)
                  for (int n = 0; n < Np; ++n) {
                    size_t r = (size_t)(i * Np + m) * n3 + (j * Np + n);
                    Ks[r] = Ks[r] + th * I1[EIDX(i, j, m, n)] - I1[EIDX(j, i, n, m)];
                    Kb[r] = Kb[r] + th * I2[EIDX(i, j, m, n)] + S2[EIDX(i, j, m, n)];
                    Ab[r] = Ab[r] + S2[EIDX(i, j, m, n)];
                    Kb[r] = Kb[r] - IT2[EIDX(i, j, m, n)];
                  }
          }
        }
      }
      int e1 = o_cholesky(Np, MP, c->RM + (size_t)P * Np * Np);
      int e2c = o_cholesky(Np, MmP, c->RMm + (size_t)P * Np * Np);
      if (e1 || e2c) {
#pragma omp atomic write
        err = 5;
      }
    }
    free(phi); free(dphi); free(pts); free(mb); free(V); free(phib); free(gradb);
    free(fpts); free(S1); free(I1); free(S2); free(I2); free(IT2);
  }
  c->built = err == 0;
  return err;
}

void o_get_K(const OCtx* c, double* K) {
  memcpy(K, c->K, sizeof(double) * (size_t)c->n_blk * c->n3 * c->n3);
}
void o_get_Adg(const OCtx* c, double* A) {
  memcpy(A, c->A, sizeof(double) * (size_t)c->n_blk * c->n3 * c->n3);
}
void o_get_mass(const OCtx* c, double* M, double* Mm, double* RM, double* RMm) {
  size_t s = sizeof(double) * (size_t)c->n_poly * c->Np * c->Np;
  if (M) memcpy(M, c->M, s);
  if (Mm) memcpy(Mm, c->Mm, s);
  if (RM) memcpy(RM, c->RM, s);
  if (RMm) memcpy(RMm, c->RMm, s);
}

/* MAKE_RHS_FREE (src/matrix_free.f90:467-843) with MAKE_RHS_VOLUME
 * (assemble_element.f90:171-237) and MAKE_RHS_FACE (:241-390). */
int o_build_rhs(const OCtx* c, int problem, double* rhs) {
  int Np = c->Np, n3 = c->n3, nq3 = c->nq3, nq2 = c->nq2, np = c->n_poly;
  memset(rhs, 0, sizeof(double) * (size_t)np * n3);
#pragma omp parallel
  {
    double* phi = malloc(sizeof(double) * Np * nq3);
    double* dphi = malloc(sizeof(double) * 3 * Np * nq3);
    double* pts = malloc(sizeof(double) * 3 * nq3);
    double* phib = malloc(sizeof(double) * Np * nq2);
    double* gradb = malloc(sizeof(double) * 3 * Np * nq2);
    double* fpts = malloc(sizeof(double) * 3 * nq2);
    double* fpr = malloc(sizeof(double) * 3 * nq2);
    double* acc = malloc(sizeof(double) * n3);
#pragma omp for schedule(dynamic, 16)
    for (int P = 0; P < np; ++P) {
      double* r = rhs + (size_t)P * n3;
      for (int a = c->ptet_ptr[P]; a < c->ptet_ptr[P + 1]; ++a) {
        int t = c->ptet[a];
        double x[4], y[4], z[4], Fk[12], Jdet;
        double lam = c->lam[t], mu = c->mu[t], rho = c->rho[t];
        tet_xyz(c, t, x, y, z);
        o_jacobians(x, y, z, Fk, &Jdet);
        tet_points(nq3, Fk, c->nodtet3, pts);
        eval_basis_pts(Np, c->blist, &c->bbox[P * 6], nq3, pts, phi, dphi);
        memset(acc, 0, sizeof(double) * n3);
        for (int q = 0; q < nq3; ++q)
          for (int m = 0; m < Np; ++m) {
            double f[3];
            forcing(problem, lam, mu, &pts[q * 3], rho, f);
            for (int i = 0; i < 3; ++i)
              acc[i * Np + m] = acc[i * Np + m] + fabs(Jdet) * c->weitet3[q] * f[i] * phi[m * nq3 + q];
          }
        for (int i = 0; i < n3; ++i) r[i] = r[i] + acc[i];
        for (int f = 0; f < 4; ++f) {
          int e2 = c->neigh[t * 4 + f];
          if (e2 >= 0) continue; /* only boundary faces contribute */
          face_points(nq2, f, Fk, c->nodtria2, fpts);
          face_points_rhs(nq2, f, Fk, c->nodtria2, fpr);
          eval_basis_pts(Np, c->blist, &c->bbox[P * 6], nq2, fpts, phib, gradb);
          const double* nv = &c->normal[(t * 4 + f) * 3];
          double area = c->area[t * 4 + f];
          int tag = c->bc_space[t * 4 + f];
          double Dbar = lam + 2 * mu;
          double sigma = c->alpha * (c->p * c->p) / c->hk[P] * Dbar;
          double n1 = nv[0], n2 = nv[1], n3v = nv[2];
          memset(acc, 0, sizeof(double) * n3);
          for (int q = 0; q < nq2; ++q) {
            double w = c->weitria2[q];
            if (e2 == -1) {
              for (int m = 0; m < Np; ++m) {
                double dd[3];
                dirichlet(problem, &fpr[q * 3], tag, dd);
                const double g1 = gradb[(0 * Np + m) * nq2 + q];
                const double g2 = gradb[(1 * Np + m) * nq2 + q];
                const double g3 = gradb[(2 * Np + m) * nq2 + q];
                const double ph = phib[m * nq2 + q];
                double l2m = lam + 2 * mu;
                double tm[3], t2[3];
                tm[0] = l2m * g1 * dd[0] * n1 + mu * g2 * dd[0] * n2 + mu * g3 * dd[0] * n3v +
                        lam * g1 * dd[1] * n2 + mu * g2 * dd[1] * n1 +
                        lam * g1 * dd[2] * n3v + mu * g3 * dd[2] * n1;
                tm[1] = l2m * g2 * dd[1] * n2 + mu * g1 * dd[1] * n1 + mu * g3 * dd[1] * n3v +
                        lam * g2 * dd[0] * n1 + mu * g1 * dd[0] * n2 +
                        lam * g2 * dd[2] * n3v + mu * g3 * dd[2] * n2;
                tm[2] = l2m * g3 * dd[2] * n3v + mu * g1 * dd[2] * n1 + mu * g2 * dd[2] * n2 +
                        lam * g3 * dd[0] * n1 + mu * g1 * dd[0] * n3v +
                        lam * g3 * dd[1] * n2 + mu * g2 * dd[1] * n3v;
                t2[0] = ph * dd[0] * (fabs(n1) * fabs(n1) + 0.5 * (fabs(n2) * fabs(n2)) + 0.5 * (fabs(n3v) * fabs(n3v))) +
                        0.5 * ph * dd[1] * n1 * n2 + 0.5 * ph * dd[2] * n1 * n3v;
                t2[1] = ph * dd[1] * (fabs(n2) * fabs(n2) + 0.5 * (fabs(n1) * fabs(n1)) + 0.5 * (fabs(n3v) * fabs(n3v))) +
                        0.5 * ph * dd[0] * n1 * n2 + 0.5 * ph * dd[2] * n2 * n3v;
                t2[2] = ph * dd[2] * (fabs(n3v) * fabs(n3v) + 0.5 * (fabs(n1) * fabs(n1)) + 0.5 * (fabs(n2) * fabs(n2))) +
                        0.5 * ph * dd[0] * n1 * n3v + 0.5 * ph * dd[1] * n2 * n3v;
                for (int i = 0; i < 3; ++i)
                  acc[i * Np + m] = acc[i * Np + m] + c->theta * w * tm[i] * area +
                                    sigma * w * t2[i] * area;
              }
            } else {
              for (int m = 0; m < Np; ++m) {
                double gn[3];
                neumann(problem, lam, mu, nv, &fpr[q * 3], tag, gn);
                for (int i = 0; i < 3; ++i)
                  acc[i * Np + m] = acc[i * Np + m] + w * phib[m * nq2 + q] * gn[i] * area;
              }
            }
          }
          for (int i = 0; i < n3; ++i)
            if (acc[i] != 0.0) r[i] = r[i] + acc[i];
        }
      }
    }
    free(phi); free(dphi); free(pts); free(phib); free(gradb); free(fpts);
    free(fpr); free(acc);
  }
  return 0;
}

/* COMPUTE_MODAL_COEFFICIENTS_FREE (src/matrix_free.f90:847-997) */
int o_project(const OCtx* c, int fn, double* coeff) {
  int Np = c->Np, n3 = c->n3, nq3 = c->nq3, np = c->n_poly;
  if (!c->built) return 4;
#pragma omp parallel
  {
    double* phi = malloc(sizeof(double) * Np * nq3);
    double* dphi = malloc(sizeof(double) * 3 * Np * nq3);
    double* pts = malloc(sizeof(double) * 3 * nq3);
    double* acc = malloc(sizeof(double) * n3);
#pragma omp for schedule(dynamic, 16)
    for (int P = 0; P < np; ++P) {
      memset(acc, 0, sizeof(double) * n3);
      for (int a = c->ptet_ptr[P]; a < c->ptet_ptr[P + 1]; ++a) {
        int t = c->ptet[a];
        double x[4], y[4], z[4], Fk[12], Jdet;
        tet_xyz(c, t, x, y, z);
        o_jacobians(x, y, z, Fk, &Jdet);
        tet_points(nq3, Fk, c->nodtet3, pts);
        eval_basis_pts(Np, c->blist, &c->bbox[P * 6], nq3, pts, phi, dphi);
        for (int q = 0; q < nq3; ++q)
          for (int m = 0; m < Np; ++m) {
            double e[3];
            field(c, fn, &pts[q * 3], e);
            for (int i = 0; i < 3; ++i)
              acc[i * Np + m] = acc[i * Np + m] + fabs(Jdet) * c->weitet3[q] * e[i] * phi[m * nq3 + q];
          }
      }
      for (int i = 0; i < 3; ++i)
        o_solve_lu(Np, c->RMm + (size_t)P * Np * Np, acc + i * Np,
                   coeff + (size_t)P * n3 + i * Np);
    }
    free(phi); free(dphi); free(pts); free(acc);
  }
  return 0;
}

/* block-row apply as the reference does it: one matmul per block, the first
 * assigned, the others added (src/matrix_free.f90:1292,1315) */
static void row_apply(const OCtx* c, const double* Kall, int P, const double* u,
                      double* tmp) {
  int n3 = c->n3;
  size_t bs = (size_t)n3 * n3;
  for (int b = c->row_ptr[P]; b < c->row_ptr[P + 1]; ++b) {
    const double* Kb = Kall + (size_t)b * bs;
    const double* x = u + (size_t)c->col[b] * n3;
    for (int r = 0; r < n3; ++r) {
      double s = 0.0;
      for (int cc = 0; cc < n3; ++cc) s = s + Kb[r * n3 + cc] * x[cc];
      tmp[r] = (b == c->row_ptr[P]) ? s : tmp[r] + s;
    }
  }
}

/* FIRST_TIME_STEP_MATRIX_FREE: src/matrix_free.f90:1100-1227 */
void o_first_step(const OCtx* c, double g0, double dt, const double* u0,
                  const double* v0, const double* rs, const double* rd,
                  double* u1) {
  int Np = c->Np, n3 = c->n3, np = c->n_poly;
  double dt2 = dt * dt, half_dt2 = 0.5 * dt2;
#pragma omp parallel
  {
    double tmp[3 * 84], x[84];
#pragma omp for schedule(static)
    for (int P = 0; P < np; ++P) {
      row_apply(c, c->K, P, u0, tmp);
      size_t o = (size_t)P * n3;
      for (int r = 0; r < n3; ++r) tmp[r] = -tmp[r] + rs[o + r] + rd[o + r] * g0;
      for (int i = 0; i < 3; ++i) {
        o_solve_lu(Np, c->RM + (size_t)P * Np * Np, tmp + i * Np, x);
        memcpy(tmp + i * Np, x, sizeof(double) * Np);
      }
      for (int r = 0; r < n3; ++r) u1[o + r] = half_dt2 * tmp[r] + u0[o + r] + dt * v0[o + r];
    }
  }
}

/* TIME_STEP_MATRIX_FREE: src/matrix_free.f90:1230-1385 */
void o_step(const OCtx* c, double g, double dt, const double* um1,
            const double* u, const double* rs, const double* rd, double* up1) {
  int Np = c->Np, n3 = c->n3, np = c->n_poly;
  double dt2 = dt * dt;
#pragma omp parallel
  {
    double tmp[3 * 84], x[84];
#pragma omp for schedule(static)
    for (int P = 0; P < np; ++P) {
      row_apply(c, c->K, P, u, tmp);
      size_t o = (size_t)P * n3;
      for (int r = 0; r < n3; ++r) tmp[r] = -tmp[r] + rs[o + r] + rd[o + r] * g;
      for (int i = 0; i < 3; ++i) {
        o_solve_lu(Np, c->RM + (size_t)P * Np * Np, tmp + i * Np, x);
        memcpy(up1 + o + i * Np, x, sizeof(double) * Np);
      }
      for (int r = 0; r < n3; ++r)
        up1[o + r] = dt2 * up1[o + r] + 2.0 * u[o + r] - um1[o + r];
    }
  }
}

/* COMPUTE_ERROR_L2_MATRIX_FREE: src/matrix_free.f90:1388-1418 (squared) */
double o_error_l2_sq(const OCtx* c, const double* uh, const double* uex) {
  int Np = c->Np, n3 = c->n3;
  double err = 0.0;
  double du[84], tmp[84];
  for (int P = 0; P < c->n_poly; ++P) {
    const double* Mm = c->Mm + (size_t)P * Np * Np;
    for (int i = 0; i < 3; ++i) {
      size_t o = (size_t)P * n3 + i * Np;
      for (int m = 0; m < Np; ++m) du[m] = uh[o + m] - uex[o + m];
      for (int m = 0; m < Np; ++m) {
        double s = 0.0;
        for (int n = 0; n < Np; ++n) s = s + Mm[m * Np + n] * du[n];
        tmp[m] = s;
      }
      double el = 0.0;
      for (int m = 0; m < Np; ++m) el = el + du[m] * tmp[m];
      err = err + el;
    }
  }
  return err;
}

/* COMPUTE_ERROR_DG_MATRIX_FREE: src/matrix_free.f90:1421-1504 (squared) */
double o_error_dg_sq(const OCtx* c, const double* uh, const double* uex) {
  int n3 = c->n3;
  size_t N = (size_t)c->n_poly * n3;
  double* du = malloc(sizeof(double) * N);
  double tmp[3 * 84];
  for (size_t i = 0; i < N; ++i) du[i] = uh[i] - uex[i];
  double err = 0.0;
  for (int P = 0; P < c->n_poly; ++P) {
    row_apply(c, c->A, P, du, tmp);
    double el = 0.0;
    for (int r = 0; r < n3; ++r) el = el + du[(size_t)P * n3 + r] * tmp[r];
    err = err + el;
  }
  free(du);
  return err;
}

static int ic_disp_fn(int problem) {
  return problem == O_PROB_CUBE_WAVE ? O_FN_CUBE_U : O_FN_ZERO;
}
static int ic_vel_fn(int problem) {
  return problem == O_PROB_REF_MMS ? O_FN_REF_VEL : O_FN_ZERO;
}
static int exact_fn(int problem) {
  if (problem == O_PROB_REF_MMS) return O_FN_REF_UEX;
  if (problem == O_PROB_CUBE_WAVE) return O_FN_CUBE_U;
  return O_FN_ZERO;
}

/* Driver: src/Lymph3D.f90:262-518 */
int o_run(OCtx* c, int problem, double dt, double stop_time, int max_steps,
          double* u_final, double* out_f) {
  int e = o_build_operator(c);
  if (e) return e;
  size_t N = (size_t)c->n_poly * c->n3;
  double* rs = malloc(sizeof(double) * N);
  double* rd = malloc(sizeof(double) * N);
  double* u0 = malloc(sizeof(double) * N);
  double* un = malloc(sizeof(double) * N);
  double* v0 = malloc(sizeof(double) * N);
  o_build_rhs(c, O_PROB_NULL, rs);
  o_build_rhs(c, problem, rd);
  o_project(c, ic_disp_fn(problem), u0);
  o_project(c, ic_vel_fn(problem), v0);
  double t = 0.0;
  int nsteps = 0;
  o_first_step(c, o_time_function(problem, t), dt, u0, v0, rs, rd, un);
  nsteps++;
  t = t + dt;
  while (t <= stop_time && (max_steps < 0 || nsteps < max_steps)) {
    o_step(c, o_time_function(problem, t), dt, u0, un, rs, rd, v0);
    double* tmp = u0;
    u0 = un;
    un = v0;
    v0 = tmp;
    nsteps++;
    t = t + dt;
  }
  /* exact solution: projection scaled by the time function (:459-467) */
  o_project(c, exact_fn(problem), u0);
  double g = o_time_function(problem, t);
  for (size_t i = 0; i < N; ++i) u0[i] = 0.0 * 0.0 + u0[i] * g;
  double eL2 = sqrt(o_error_l2_sq(c, un, u0));
  double eDG = sqrt(o_error_dg_sq(c, un, u0));
  if (u_final) memcpy(u_final, un, sizeof(double) * N);
  out_f[0] = t;
  out_f[1] = nsteps;
  out_f[2] = eL2;
  out_f[3] = eDG;
  free(rs); free(rd); free(u0); free(un); free(v0);
  (void)PI_;
  return 0;
}

/* FIND_ELEM_FROM_POINT (src/find_tet.f90:82-124): the tet whose barycentre
 * (COMPUTE_BARYCENTER :151-181) is nearest, first minimum on ties. */
void o_locate_points(const OCtx* c, int n, const double* pts, int* tet) {
  for (int k = 0; k < n; ++k) {
    double dmin = 1.0e15;
    int best = -1;
    for (int t = 0; t < c->n_tet; ++t) {
      double b[3];
      for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        for (int v = 0; v < 4; ++v) s = s + c->xyz[3 * c->tet_vert[t * 4 + v] + d];
        b[d] = s / 4;
      }
      double dx = pts[k * 3] - b[0], dy = pts[k * 3 + 1] - b[1], dz = pts[k * 3 + 2] - b[2];
      double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < dmin) { dmin = d2; best = t; }
    }
    tet[k] = best;
  }
}

/* Point value of the modal field, as EXPORT_SOLUTION evaluates it at nodes
 * (src/solution_processing.f90:150-170): sum_i u_i Lx Ly Lz on the
 * polyhedron's bounding box.  out[k*3 + comp]. */
void o_eval_points(const OCtx* c, int n, const double* pts, const int* tet, const double* u,
                   double* out) {
  int Np = c->Np;
  for (int k = 0; k < n; ++k) {
    int P = c->tet_poly[tet[k]];
    const double* bb = &c->bbox[P * 6];
    for (int comp = 0; comp < 3; ++comp) {
      double s = 0.0;
      for (int i = 0; i < Np; ++i) {
        double v[3];
        for (int d = 0; d < 3; ++d) {
          double hb = (bb[2 * d + 1] - bb[2 * d]) / 2.0, mb = (bb[2 * d + 1] + bb[2 * d]) / 2.0;
          v[d] = leg(c->blist[i * 3 + d], (pts[k * 3 + d] - mb) / hb) / sqrt(hb);
        }
        s = s + u[(size_t)P * c->n3 + comp * Np + i] * v[0] * v[1] * v[2];
      }
      out[k * 3 + comp] = s;
    }
  }
}

/* Point force F(P_s)[i*Np+m] += f_i phi_m(x_s) for every source (the exact
 * load of f delta(x - x_s)), phi on the bounding box of the polyhedron of the
 * source's tet; rhs is accumulated in source order. */
void o_point_sources(const OCtx* c, int n, const double* pts, const int* tet, const double* f,
                     double* rhs) {
  int Np = c->Np;
  for (int k = 0; k < n; ++k) {
    int P = c->tet_poly[tet[k]];
    const double* bb = &c->bbox[P * 6];
    for (int i = 0; i < 3; ++i)
      for (int m = 0; m < Np; ++m) {
        double v[3];
        for (int d = 0; d < 3; ++d) {
          double hb = (bb[2 * d + 1] - bb[2 * d]) / 2.0, mb = (bb[2 * d + 1] + bb[2 * d]) / 2.0;
          v[d] = leg(c->blist[m * 3 + d], (pts[k * 3 + d] - mb) / hb) / sqrt(hb);
        }
        size_t o = (size_t)P * c->n3 + i * Np + m;
        rhs[o] = rhs[o] + f[k * 3 + i] * v[0] * v[1] * v[2];
      }
  }
}

/* Ricker wavelet (1 - 2 a) exp(-a), a = (pi f0 (t - t0))^2 */
double o_ricker(double f0, double t0, double t) {
  double a = o_pi() * f0 * (t - t0);
  a = a * a;
  return (1.0 - 2.0 * a) * exp(-a);
}

/* ---------------- test access to the data sets and the polynomial field ---- */
void o_problem_data(int problem, int n, const double* xyz, const double* normal, double lam,
                    double mu, double rho, int tag, double* f, double* gd, double* gn) {
  for (int k = 0; k < n; ++k) {
    if (f) forcing(problem, lam, mu, xyz + 3 * k, rho, f + 3 * k);
    if (gd) dirichlet(problem, xyz + 3 * k, tag, gd + 3 * k);
    if (gn) neumann(problem, lam, mu, normal + 3 * k, xyz + 3 * k, tag, gn + 3 * k);
  }
}

int o_set_poly_field(OCtx* c, int nterms, const int* exps, const double* coef) {
  if (nterms < 0 || nterms > O_MAX_POLY_TERMS) return 1;
  c->poly_n = nterms;
  for (int k = 0; k < 3 * nterms; ++k) {
    c->poly_exp[k] = exps[k];
    c->poly_coef[k] = coef[k];
  }
  return 0;
}

/* ------------------------- double-couple sources ------------------------ */
/* 4x4 determinant by cofactors along the first row, in the term order of
 * det4 (src/vet_mat_operations.f90:52-69); A row-major. */
double o_det4(const double* A) {
#define A_(i, j) A[((i) - 1) * 4 + ((j) - 1)]
  return A_(1, 1) * (A_(2, 2) * (A_(3, 3) * A_(4, 4) - A_(3, 4) * A_(4, 3)) -
                     A_(2, 3) * (A_(3, 2) * A_(4, 4) - A_(3, 4) * A_(4, 2)) +
                     A_(2, 4) * (A_(3, 2) * A_(4, 3) - A_(3, 3) * A_(4, 2))) -
         A_(1, 2) * (A_(2, 1) * (A_(3, 3) * A_(4, 4) - A_(3, 4) * A_(4, 3)) -
                     A_(2, 3) * (A_(3, 1) * A_(4, 4) - A_(3, 4) * A_(4, 1)) +
                     A_(2, 4) * (A_(3, 1) * A_(4, 3) - A_(3, 3) * A_(4, 1))) +
         A_(1, 3) * (A_(2, 1) * (A_(3, 2) * A_(4, 4) - A_(3, 4) * A_(4, 2)) -
                     A_(2, 2) * (A_(3, 1) * A_(4, 4) - A_(3, 4) * A_(4, 1)) +
                     A_(2, 4) * (A_(3, 1) * A_(4, 2) - A_(3, 2) * A_(4, 1))) -
         A_(1, 4) * (A_(2, 1) * (A_(3, 2) * A_(4, 3) - A_(3, 3) * A_(4, 2)) -
                     A_(2, 2) * (A_(3, 1) * A_(4, 3) - A_(3, 3) * A_(4, 1)) +
                     A_(2, 3) * (A_(3, 1) * A_(4, 2) - A_(3, 2) * A_(4, 1)));
#undef A_
}

/* moment_density (src/problem_data_and_properties.f90:317-335):
 * M_ij = ((s_i n_j + s_j n_i) M0) / V; mom row-major 3x3. */
void o_moment_density(double M0, double V, const double* s, const double* n, double* mom) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) mom[i * 3 + j] = ((s[i] * n[j] + s[j] * n[i]) * M0) / V;
}

/* MAKE_DOUBLE_COUPLE (src/assemble_element.f90:116-167): out[i*Np+m] =
 * sum_q (w_q |J|) sum_j M_ij d_j phi_m(x_q); dphi[(d*Np + m)*nq3 + q]. */
void o_double_couple(int p, double Jdet, const double* dphi, const double* mom, double* out) {
  int Np = o_np(p), nq3 = (p + 1) * (p + 1) * (p + 1);
  double t3[4 * 125], w3[125], t2[4 * 25], w2[25];
  o_quadrature(p, t3, w3, t2, w2);
  for (int r = 0; r < 3 * Np; ++r) out[r] = 0.0;
  for (int q = 0; q < nq3; ++q)
    for (int m = 0; m < Np; ++m)
      for (int i = 0; i < 3; ++i) {
        double reduct = 0.0;
        for (int j = 0; j < 3; ++j) reduct = reduct + mom[i * 3 + j] * dphi[(j * Np + m) * nq3 + q];
        out[i * Np + m] = out[i * Np + m] + w3[q] * fabs(Jdet) * reduct;
      }
}

/* The SISM loop of MAKE_RHS_FREE (src/matrix_free.f90:600-664): for every
 * source, in order, the couple of its tet (volume |det4|/6 of the tet,
 * find_tet.f90:184-210; basis on the polyhedron's bounding box) is added to
 * the polyhedron's entries of rhs.  strike/normal[3n], m0[n] per source. */
void o_double_couples(const OCtx* c, int n, const int* tet, const double* strike, const double* normal,
                      const double* m0, double* rhs) {
  int Np = c->Np, nq3 = c->nq3, n3 = c->n3;
  double* phi = malloc(sizeof(double) * Np * nq3);
  double* dphi = malloc(sizeof(double) * 3 * Np * nq3);
  double* pts = malloc(sizeof(double) * 3 * nq3);
  double* cpl = malloc(sizeof(double) * n3);
  for (int k = 0; k < n; ++k) {
    int t = tet[k], P = c->tet_poly[t];
    double x[4], y[4], z[4], Fk[12], Jdet, A[16], mom[9];
    tet_xyz(c, t, x, y, z);
    o_jacobians(x, y, z, Fk, &Jdet);
    for (int v = 0; v < 4; ++v) {
      A[v * 4 + 0] = x[v];
      A[v * 4 + 1] = y[v];
      A[v * 4 + 2] = z[v];
      A[v * 4 + 3] = 1.0;
    }
    double vol = fabs(o_det4(A)) / 6.0;
    o_moment_density(m0[k], vol, strike + 3 * k, normal + 3 * k, mom);
    tet_points(nq3, Fk, c->nodtet3, pts);
    eval_basis_pts(Np, c->blist, &c->bbox[P * 6], nq3, pts, phi, dphi);
    o_double_couple(c->p, Jdet, dphi, mom, cpl);
    for (int r = 0; r < n3; ++r) rhs[(size_t)P * n3 + r] = rhs[(size_t)P * n3 + r] + cpl[r];
  }
  free(phi); free(dphi); free(pts); free(cpl);
}

// lymph_math.hpp -- host/device scalar math of the PolyDG path: Gauss-Legendre
// tables, Duffy and triangle maps, tensor Legendre basis on a bounding box,
// affine tet maps and the built-in problem data.  Expressions keep the operand
// order of the reference routines they restate (cited per function) so that
// the FP64 results agree with the reference to rounding; every translation
// unit including this header is compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#define LYMPH_HD __host__ __device__ __forceinline__

namespace lymph {

constexpr int kMaxP = 4;

LYMPH_HD constexpr int np_of(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

// pi = 4*atan(1) and sqrt(2), sqrt(3) correctly rounded
// (global_parameters, src/Poly_global.f90:416-417).
constexpr double kPi = 3.141592653589793115997963468544185161590576171875;
constexpr double kSqrt2 = 1.4142135623730951454746218587388284504413604736328125;
constexpr double kSqrt3 = 1.7320508075688771931766041234368458390235900878906250;

// 1-D Gauss-Legendre nodes/weights on (-1,1) in the reference's hard-coded
// order (src/basis_function.f90:77-124); n = p+1 <= 5 here.
inline void gl_nodes(int n, double* x, double* w) {
  switch (n) {
    case 1: x[0] = 0.0; w[0] = 2.0; break;
    case 2:
      x[0] = std::sqrt(1.0 / 3.0); x[1] = -std::sqrt(1.0 / 3.0);
      w[0] = 1.0; w[1] = 1.0;
      break;
    case 3:
      x[0] = 0.0; x[1] = std::sqrt(3.0 / 5.0); x[2] = -std::sqrt(3.0 / 5.0);
      w[0] = 8.0 / 9.0; w[1] = 5.0 / 9.0; w[2] = 5.0 / 9.0;
      break;
    case 4:
      x[0] = 0.3399810435848563; x[1] = -0.3399810435848563;
      x[2] = 0.8611363115940526; x[3] = -0.8611363115940526;
      w[0] = w[1] = (18 + std::sqrt(30.0)) / 36.0;
      w[2] = w[3] = (18 - std::sqrt(30.0)) / 36.0;
      break;
    case 5:
      x[0] = 0.0; x[1] = 0.5384693101056831; x[2] = -0.5384693101056831;
      x[3] = 0.9061798459386640; x[4] = -0.9061798459386640;
      w[0] = 128.0 / 225.0;
      w[1] = w[2] = (322 + 13 * std::sqrt(70.0)) / 900.0;
      w[3] = w[4] = (322 - 13 * std::sqrt(70.0)) / 900.0;
      break;
  }
}

// Tet rule = tensor GL on the cube (i outer, then j, k) through the Duffy map
// (src/basis_function.f90:142-154, src/Poly_ref_mappings.f90:137-162); triangle
// rule = tensor GL on the square through the collapsed map (:165-186).
// Layout: tet[q*4 + k] (x^, y^, z^, 1), wt[q]; tri[q*4 + k] (s, t, 1, 1), wtri[q].
inline void quad_tables(int p, double* tet, double* wt, double* tri, double* wtri) {
  const int n = p + 1;
  double x[8], w[8];
  gl_nodes(n, x, w);
  int c = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k, ++c) {
        const double a = 0.5 * x[i] + 0.5, b = 0.5 * x[j] + 0.5, cc = 0.5 * x[k] + 0.5;
        const double wc = 0.125 * (w[i] * w[j] * w[k]);
        tet[c * 4 + 0] = a;
        tet[c * 4 + 1] = (1 - a) * b;
        tet[c * 4 + 2] = (1 - a) * (1 - b) * cc;
        tet[c * 4 + 3] = 1.0;
        wt[c] = std::fabs((1 - b) * ((1 - a) * (1 - a))) * wc;
      }
  c = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j, ++c) {
      const double s = x[i], t = x[j];
      tri[c * 4 + 0] = 0.5 * (1 + s);
      tri[c * 4 + 1] = 0.25 * (1 - s) * (1 + t);
      tri[c * 4 + 2] = 1.0;
      tri[c * 4 + 3] = 1.0;
      wtri[c] = std::fabs(0.125 * (1 - s)) * (w[i] * w[j]);
    }
}

// Total-degree index list, q1 from p down, then q2, q3 (basis_list,
// src/basis_function.f90:12-44).
inline void basis_list(int p, int* bl) {
  int ii = 0;
  for (int q1 = p; q1 >= 0; --q1)
    for (int q2 = p - q1; q2 >= 0; --q2)
      for (int q3 = p - q1 - q2; q3 >= 0; --q3, ++ii) {
        bl[ii * 3 + 0] = q1;
        bl[ii * 3 + 1] = q2;
        bl[ii * 3 + 2] = q3;
      }
}

// Legendre P_k and P'_k on (-1,1), k <= 4 (LegendreP_nonnorm / GradLegendreP,
// src/basis_function.f90:297-310, 398-409).
LYMPH_HD double leg(int k, double xp) {
  const double x2 = xp * xp;
  switch (k) {
    case 0: return 1.0;
    case 1: return xp;
    case 2: return (3 * x2 - 1) / 2;
    case 3: return (5 * (x2 * xp) - 3 * xp) / 2;
    case 4: return (35 * (x2 * x2) - 30 * x2 + 3) / 8;
  }
  return 0.0;
}
LYMPH_HD double dleg(int k, double xp) {
  const double x2 = xp * xp;
  switch (k) {
    case 0: return 0.0;
    case 1: return 1.0;
    case 2: return 3 * xp;
    case 3: return (15 * x2 - 3) / 2;
    case 4: return (35 * (x2 * xp) - 15 * xp) / 2;
  }
  return 0.0;
}

// phi, dphi/dx_d of basis f at point x for bounding box bb (basis,
// src/basis_function.f90:440-517: L_k((x-m)/h)/sqrt(h), L'_k(.)/(h sqrt(h))).
LYMPH_HD void basis_at(const int* bl, const double* bb, const double* x,
                       double& phi, double& d0, double& d1, double& d2) {
  double v[3], dv[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double hb = (bb[2 * d + 1] - bb[2 * d]) / 2.0;
    const double mb = (bb[2 * d + 1] + bb[2 * d]) / 2.0;
    const double sq = sqrt(hb);
    const double xp = (x[d] - mb) / hb;
    v[d] = leg(bl[d], xp) / sq;
    dv[d] = dleg(bl[d], xp) / (hb * sq);
  }
  phi = v[0] * v[1] * v[2];
  d0 = dv[0] * v[1] * v[2];
  d1 = v[0] * dv[1] * v[2];
  d2 = v[0] * v[1] * dv[2];
}

// Affine map x = J x^ + x1 with J = [x2-x1, x3-x1, x4-x1] and det3 in the
// reference's term order (jacobians, src/Poly_ref_mappings.f90:18-75;
// det3, src/vet_mat_operations.f90:38-49).  v[vertex*3 + d]; Fk[d*4 + j].
LYMPH_HD void affine_map(const double* v, double* Fk, double& Jdet) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double x1 = v[d], x2 = v[3 + d], x3 = v[6 + d], x4 = v[9 + d];
    Fk[d * 4 + 0] = ((0.0 + -1 * x1) + 1 * x2) + 0 * x3 + 0 * x4;
    Fk[d * 4 + 1] = ((0.0 + -1 * x1) + 0 * x2) + 1 * x3 + 0 * x4;
    Fk[d * 4 + 2] = ((0.0 + -1 * x1) + 0 * x2) + 0 * x3 + 1 * x4;
    Fk[d * 4 + 3] = ((0.0 + 1 * x1) + 0 * x2) + 0 * x3 + 0 * x4;
  }
  const double a11 = Fk[0], a12 = Fk[1], a13 = Fk[2];
  const double a21 = Fk[4], a22 = Fk[5], a23 = Fk[6];
  const double a31 = Fk[8], a32 = Fk[9], a33 = Fk[10];
  Jdet = a11 * a22 * a33 - a11 * a23 * a32 - a12 * a21 * a33 + a12 * a23 * a31 +
         a13 * a21 * a32 - a22 * a13 * a31;
}

// Physical point of a tet quadrature node (basis :471-478).
LYMPH_HD void map_point(const double* Fk, const double* nod4, double* x) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < 4; ++l) s = s + Fk[j * 4 + l] * nod4[l];
    x[j] = s;
  }
}

// node_maps(:,:,face) (tria2tetfaces_maps, src/Poly_ref_mappings.f90:79-134):
// row i (ref-tet coordinate), column j (s, t, 1, 1).
LYMPH_HD double node_map(int face, int i, int j) {
  switch (face) {
    case 0: return (i == 0 && j == 1) || (i == 1 && j == 0) || (i == 3 && j == 3) ? 1.0 : 0.0;
    case 1: return (i == 0 && j == 0) || (i == 2 && j == 1) || (i == 3 && j == 3) ? 1.0 : 0.0;
    case 2:
      if (i == 2 && (j == 0 || j == 1)) return -1.0;
      return (i == j) ? 1.0 : 0.0;
    default: return (i == 1 && j == 1) || (i == 2 && j == 0) || (i == 3 && j == 3) ? 1.0 : 0.0;
  }
}

// Face quadrature point as basis_boundary computes it (:557-575).
LYMPH_HD void face_point(const double* Fk, int face, const double* tri4, double* x) {
  double temp[3][4];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < 4; ++l) s = s + Fk[j * 4 + l] * node_map(face, l, q);
      temp[j][q] = s;
    }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < 4; ++l) s = s + temp[j][l] * tri4[l];
    x[j] = s;
  }
}

// Face quadrature point as MAKE_RHS_FACE computes it (assemble_element :321-328).
LYMPH_HD void face_point_rhs(const double* Fk, int face, const double* tri4, double* x) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int t = 0; t < 4; ++t) s = s + Fk[j * 4 + k] * node_map(face, k, t) * tri4[t];
    x[j] = s;
  }
}

// ---- problem data (src/problem_data_and_properties.f90) ----
enum { kProbNull = 0, kProbRefMms = 1, kProbCubeWave = 2, kProbRicker = 3 };
enum { kFnZero = 0, kFnRefVel = 1, kFnRefUex = 2, kFnCubeU = 3 };

// f_null / f_time (:48-81) / cube standing wave.
LYMPH_HD void forcing(int prob, double lam, double mu, const double* p, double rho, double* r) {
  if (prob == kProbRefMms) {
    const double c = (-2.0 * rho + 3.0 * (lam + 2.0 * mu));
    r[0] = c * kPi * kPi * cos(kPi * p[0]) * sin(kPi * p[1]) * sin(kPi * p[2]);
    r[1] = c * kPi * kPi * sin(kPi * p[0]) * cos(kPi * p[1]) * sin(kPi * p[2]);
    r[2] = c * kPi * kPi * sin(kPi * p[0]) * sin(kPi * p[1]) * cos(kPi * p[2]);
  } else if (prob == kProbCubeWave) {
    // f = rho u_tt - div sigma(u) for u = S(x)(1,1,1) cos(sqrt3 pi t)
    const double sx = sin(kPi * p[0]), sy = sin(kPi * p[1]), sz = sin(kPi * p[2]);
    const double cx = cos(kPi * p[0]), cy = cos(kPi * p[1]), cz = cos(kPi * p[2]);
    const double S = sx * sy * sz;
    const double a = (-3.0 * rho + 3.0 * mu + (lam + mu)) * S;
    const double lm = lam + mu;
    r[0] = kPi * kPi * (a - lm * (cx * cy * sz + cx * sy * cz));
    r[1] = kPi * kPi * (a - lm * (cx * cy * sz + sx * cy * cz));
    r[2] = kPi * kPi * (a - lm * (cx * sy * cz + sx * cy * cz));
  } else {
    r[0] = r[1] = r[2] = 0.0;
  }
}

// gd_null / gd (:130-155) / cube wave.
LYMPH_HD void dirichlet(int prob, const double* p, int tag, double* r) {
  r[0] = r[1] = r[2] = 0.0;
  if (prob == kProbRefMms) {
    if (tag >= 1 && tag <= 3) {
      r[0] = cos(kPi * p[0]) * sin(kPi * p[1]) * sin(kPi * p[2]);
      r[1] = sin(kPi * p[0]) * cos(kPi * p[1]) * sin(kPi * p[2]);
      r[2] = sin(kPi * p[0]) * sin(kPi * p[1]) * cos(kPi * p[2]);
    }
  } else if (prob == kProbCubeWave) {
    const double S = sin(kPi * p[0]) * sin(kPi * p[1]) * sin(kPi * p[2]);
    r[0] = r[1] = r[2] = S;
  }
}

// gn_null / gn (:186-243) / cube wave: sigma(u) n.
LYMPH_HD void neumann(int prob, double lam, double mu, const double* n, const double* p, int tag,
                      double* r) {
  double st[3][3];
  r[0] = r[1] = r[2] = 0.0;
  if (prob == kProbRefMms) {
    if (tag < 1 || tag > 3) return;
    st[0][0] = -kPi * sin(kPi * p[0]) * sin(kPi * p[1]) * sin(kPi * p[2]) * (3.0 * lam + 2.0 * mu);
    st[1][1] = st[0][0];
    st[2][2] = st[0][0];
    st[0][1] = 2.0 * mu * kPi * cos(kPi * p[0]) * cos(kPi * p[1]) * sin(kPi * p[2]);
    st[1][0] = st[0][1];
    st[0][2] = 2.0 * mu * kPi * cos(kPi * p[0]) * cos(kPi * p[2]) * sin(kPi * p[1]);
    st[2][0] = st[0][2];
    st[1][2] = 2.0 * mu * kPi * cos(kPi * p[1]) * cos(kPi * p[2]) * sin(kPi * p[0]);
    st[2][1] = st[1][2];
  } else if (prob == kProbCubeWave) {
    const double sx = sin(kPi * p[0]), sy = sin(kPi * p[1]), sz = sin(kPi * p[2]);
    const double cx = cos(kPi * p[0]), cy = cos(kPi * p[1]), cz = cos(kPi * p[2]);
    const double g[3] = {kPi * cx * sy * sz, kPi * sx * cy * sz, kPi * sx * sy * cz};
    const double div = g[0] + g[1] + g[2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) st[i][j] = (i == j ? lam * div : 0.0) + mu * (g[j] + g[i]);
  } else {
    return;
  }
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int j = 0; j < 3; ++j) s = s + st[i][j] * n[j];
    r[i] = s;
  }
}

// ic_velocity (:292-302), uex (:245-255), cube-wave displacement.
LYMPH_HD void field(int fn, const double* p, double* r) {
  r[0] = r[1] = r[2] = 0.0;
  if (fn == kFnRefVel) {
    const double c = kSqrt2 * kPi;
    r[0] = c * cos(kPi * p[0]) * sin(kPi * p[1]) * sin(kPi * p[2]);
    r[1] = c * sin(kPi * p[0]) * cos(kPi * p[1]) * sin(kPi * p[2]);
    r[2] = c * sin(kPi * p[0]) * sin(kPi * p[1]) * cos(kPi * p[2]);
  } else if (fn == kFnRefUex) {
    r[0] = cos(kPi * p[0]) * sin(kPi * p[1]) * sin(kPi * p[2]);
    r[1] = sin(kPi * p[0]) * cos(kPi * p[1]) * sin(kPi * p[2]);
    r[2] = sin(kPi * p[0]) * sin(kPi * p[1]) * cos(kPi * p[2]);
  } else if (fn == kFnCubeU) {
    const double S = sin(kPi * p[0]) * sin(kPi * p[1]) * sin(kPi * p[2]);
    r[0] = r[1] = r[2] = S;
  }
}

// Ricker wavelet (1 - 2a) exp(-a), a = (pi f0 (t - t0))^2.
LYMPH_HD double ricker(double f0, double t0, double t) {
  double a = kPi * f0 * (t - t0);
  a = a * a;
  return (1.0 - 2.0 * a) * exp(-a);
}

inline double time_factor(int prob, double t) {
  if (prob == kProbCubeWave) return std::cos(kSqrt3 * kPi * t);
  return std::sin(kSqrt2 * kPi * t);  // time_function (:14-23)
}

}  // namespace lymph

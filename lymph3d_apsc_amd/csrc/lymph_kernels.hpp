// lymph_kernels.hpp -- gfx950 kernels of the PolyDG leapfrog path.
//
// Data layout in HBM (one context):
//  * operator K and A_dg: one block row per polyhedron, blocks of 3Np x 3Np
//    FP64 in "column-pair" order: entry (r, c) of block b sits at
//      b*N3*N3C + ((c>>1)*N3 + r)*2 + (c&1),     N3 = 3Np, N3C = N3 rounded up
//    to even (padding column zero).  A 64-lane wave with lane = row r reads one
//    column pair of a block as a single 16-byte load per lane: N3*16 contiguous
//    bytes per wave instruction, every byte of the block read exactly once.
//  * mass factors R (upper, M = R^T R): packed row-major upper triangle,
//    Np(Np+1)/2 doubles per polyhedron (the bytes the step must read).
//  * vectors: 3Np doubles per polyhedron, component-major inside.
#pragma once
#include "lymph_math.hpp"

namespace lymph {

template <int P>
struct Dims {
  static constexpr int NP = np_of(P);
  static constexpr int N3 = 3 * NP;
  static constexpr int N3C = (N3 + 1) & ~1;
  static constexpr int NPAIR = N3C / 2;
  static constexpr int NQ1 = P + 1;
  static constexpr int NQ3 = NQ1 * NQ1 * NQ1;
  static constexpr int NQ2 = NQ1 * NQ1;
  static constexpr int NPK = NP * (NP + 1) / 2;
  static constexpr size_t BS = (size_t)N3 * N3C;
};

// Operator blocks are read exactly once per step and never again before the
// next step has streamed ~all of HBM past the caches: non-temporal loads keep
// them from displacing the coefficient vectors (which neighbours re-read).
typedef double lymph_dv2 __attribute__((ext_vector_type(2)));
typedef float lymph_fv4 __attribute__((ext_vector_type(4)));
typedef float lymph_fv2 __attribute__((ext_vector_type(2)));
__device__ inline double2 load_stream(const double* p) {
  const lymph_dv2 t = __builtin_nontemporal_load(reinterpret_cast<const lymph_dv2*>(p));
  return make_double2(t.x, t.y);
}
__device__ inline float4 load_stream(const float* p) {
  const lymph_fv4 t = __builtin_nontemporal_load(reinterpret_cast<const lymph_fv4*>(p));
  return make_float4(t.x, t.y, t.z, t.w);
}
// the tail columns of a 32-bit block row (N3 % 4 of them; 1 or 2 for p <= 4)
template <int T>
__device__ inline void load_tail(const float* p, float (&v)[T]) {
  static_assert(T == 1 || T == 2, "tail of 1 or 2 columns");
  if constexpr (T == 2) {
    const lymph_fv2 t = __builtin_nontemporal_load(reinterpret_cast<const lymph_fv2*>(p));
    v[0] = t.x;
    v[1] = t.y;
  } else {
    v[0] = __builtin_nontemporal_load(p);
  }
}

LYMPH_HD size_t kpos(int N3, int r, int c) {
  return ((size_t)(c >> 1) * N3 + r) * 2 + (c & 1);
}

// Layout of a 32-bit block: the first 4*NQ columns as column quads (entry
// (r, c) at ((c>>2)*N3 + r)*4 + (c&3)), then the T = N3 % 4 tail columns
// row by row (at 4*NQ*N3 + r*T + c - 4*NQ); no padding columns, so the block
// stream is exactly N3^2 floats (the block stride is rounded up to 4 floats
// to keep every block 16-byte aligned).
template <int P>
struct Dims32 {
  static constexpr int N3 = 3 * np_of(P);
  static constexpr int NQ = N3 / 4, T = N3 % 4;
  static constexpr size_t BS = ((size_t)N3 * N3 + 3) / 4 * 4;
};
// The 48-bit apply's 16-bit corrections: column quads taken in pairs, a
// row's 8 values of quads 2k, 2k+1 contiguous (one 16-byte load); the last
// quad of an odd count row by row at 4 values per row (one 8-byte load, no
// half-empty slots to come in with the lines); then the N3 % 4 tail columns
// row by row; blocks padded to 8 values (none at p = 1, 3).
template <int P>
struct Dims16 {
  static constexpr int N3 = 3 * np_of(P);
  static constexpr int NQ = N3 / 4, T = N3 % 4, NPF = NQ / 2, NPR = (NQ + 1) / 2;
  static constexpr size_t HALF0 = (size_t)NPF * N3 * 8;
  static constexpr size_t TAIL0 = HALF0 + (size_t)(NQ % 2) * N3 * 4;
  static constexpr size_t BS = (TAIL0 + (size_t)N3 * T + 7) / 8 * 8;
};
LYMPH_HD size_t kpos16(int N3, int r, int c) {
  const int c0 = N3 / 4 * 4, nq = N3 / 4, npf = nq / 2;
  const size_t half0 = (size_t)npf * N3 * 8;
  if (c < c0) {
    const int q = c >> 2;
    if (q < 2 * npf) return ((size_t)(q >> 1) * N3 + r) * 8 + (q & 1) * 4 + (c & 3);
    return half0 + (size_t)r * 4 + (c & 3);
  }
  return half0 + (size_t)(nq % 2) * N3 * 4 + (size_t)r * (N3 % 4) + (c - c0);
}
// The corrections' loads: 8 (a quad pair's row) or 4 (the last quad's row)
// halves, non-temporal like the operator's own.
typedef _Float16 lymph_hv4 __attribute__((ext_vector_type(4)));
typedef _Float16 lymph_hv8 __attribute__((ext_vector_type(8)));
__device__ inline lymph_hv8 load_stream_h8(const _Float16* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const lymph_hv8*>(p));
}
__device__ inline lymph_hv4 load_stream_h4(const _Float16* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const lymph_hv4*>(p));
}
LYMPH_HD size_t kpos32(int N3, int r, int c) {
  const int c0 = N3 / 4 * 4;
  if (c < c0) return ((size_t)(c >> 2) * N3 + r) * 4 + (c & 3);
  return (size_t)c0 * N3 + (size_t)r * (N3 % 4) + (c - c0);
}
LYMPH_HD int pk_off(int NP, int i) { return i * NP - i * (i - 1) / 2; }
// Doubles per block product of the pairs-once steps: whole 128-byte lines, so
// that each product's store covers its lines entirely (the padding lanes
// write zeros) instead of leaving partial lines that other waves complete.
LYMPH_HD constexpr int prod_stride(int n3) { return (n3 + 15) & ~15; }

struct MeshDev {
  const double* tet_xyz;  // n_tet*12
  const int* tet_poly;
  const int* neigh;       // n_tet*4
  const int* bc_space;    // n_tet*4
  const double* normal;   // n_tet*12
  const double* area;     // n_tet*4
  const double* rho;
  const double* lam;
  const double* mu;
  const double* bbox;     // n_poly*6
  const double* hk;       // n_poly
  const int* ptet_ptr;    // n_poly+1
  const int* ptet;        // n_tet
  const int* row_ptr;     // n_poly+1
  const int* col;         // n_blk
  const int* face_blk;    // n_tet*4 block index of the neighbour block, -1 if none
  const double* qtet;     // NQ3*4
  const double* wtet;     // NQ3
  const double* qtri;     // NQ2*4
  const double* wtri;     // NQ2
  const int* blist;       // NP*3
};

// ---------------------------------------------------------------------------
// Operator build: MAKE_MATRICES_FREE (src/matrix_free.f90:27-464).
// One 256-thread workgroup per polyhedron; its tets are processed in order
// (volume term, then faces 1..4), each output entry owned by one thread for
// the whole kernel so every entry accumulates in the reference's order.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void k_build_operator(MeshDev m, double alpha, double theta,
                                                         double* __restrict__ K,
                                                         double* __restrict__ A,
                                                         double* __restrict__ M,
                                                         double* __restrict__ Mm, int poly0) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3, NQ3 = D::NQ3, NQ2 = D::NQ2;
  constexpr int VOLSZ = 4 * NP * NQ3, FACESZ = 8 * NP * NQ2;
  constexpr int SHSZ = VOLSZ > FACESZ ? VOLSZ : FACESZ;
  __shared__ double sh[SHSZ];
  __shared__ double spt[NQ3 * 3];
  __shared__ int sbl[NP * 3];
  const int tid = threadIdx.x;
  const int Pp = poly0 + blockIdx.x;
  for (int i = tid; i < NP * 3; i += blockDim.x) sbl[i] = m.blist[i];
  double bbP[6];
  for (int d = 0; d < 6; ++d) bbP[d] = m.bbox[Pp * 6 + d];
  const double hkP = m.hk[Pp];
  double* Ks = K + (size_t)m.row_ptr[Pp] * D::BS;
  const bool keepA = A != nullptr;  // A_dg may be dropped for batched runs
  double* As = keepA ? A + (size_t)m.row_ptr[Pp] * D::BS : nullptr;
  double* MP = M + (size_t)Pp * NP * NP;
  double* MmP = Mm + (size_t)Pp * NP * NP;
  // volume-phase views
  double* phi = sh;                // [NP][NQ3]
  double* dphi = sh + NP * NQ3;    // [3][NP][NQ3]
  // face-phase views
  double* p1 = sh;                      // [NP][NQ2]
  double* p2 = sh + NP * NQ2;           // [NP][NQ2]
  double* g1 = sh + 2 * NP * NQ2;       // [3][NP][NQ2]
  double* g2 = sh + 5 * NP * NQ2;       // [3][NP][NQ2]

  for (int a = m.ptet_ptr[Pp]; a < m.ptet_ptr[Pp + 1]; ++a) {
    const int t = m.ptet[a];
    double v[12], Fk[12], Jdet;
    for (int i = 0; i < 12; ++i) v[i] = m.tet_xyz[(size_t)t * 12 + i];
    affine_map(v, Fk, Jdet);
    const double aJ = fabs(Jdet);
    const double rho = m.rho[t], lam = m.lam[t], mu = m.mu[t];
    const double l2m = lam + 2 * mu;
    __syncthreads();
    for (int q = tid; q < NQ3; q += blockDim.x) map_point(Fk, &m.qtet[q * 4], &spt[q * 3]);
    __syncthreads();
    for (int e = tid; e < NP * NQ3; e += blockDim.x) {
      const int f = e / NQ3, q = e % NQ3;
      double ph, d0, d1, d2;
      basis_at(&sbl[f * 3], bbP, &spt[q * 3], ph, d0, d1, d2);
      phi[e] = ph;
      dphi[(0 * NP + f) * NQ3 + q] = d0;
      dphi[(1 * NP + f) * NQ3 + q] = d1;
      dphi[(2 * NP + f) * NQ3 + q] = d2;
    }
    __syncthreads();
    // mass (MAKE_MATRICES_FREE :224-233 with MAKE_MASS_VOLUME)
    for (int e = tid; e < NP * NP; e += blockDim.x) {
      const int mm = e / NP, nn = e % NP;
      double s = 0.0;
      for (int q = 0; q < NQ3; ++q)
        s = s + 1.0 * m.wtet[q] * aJ * phi[mm * NQ3 + q] * phi[nn * NQ3 + q];
      MmP[e] = MmP[e] + s;
      MP[e] = MP[e] + s * rho;
    }
    // volume stiffness (MAKE_STIFFNESS_VOLUME, assemble_element :24-77)
    for (int e = tid; e < 9 * NP * NP; e += blockDim.x) {
      const int i = e / (3 * NP * NP), j = (e / (NP * NP)) % 3;
      const int mm = (e / NP) % NP, nn = e % NP;
      const int oi = i == 0 ? 1 : 0, oi2 = i == 2 ? 1 : 2;  // the other two directions
      double s = 0.0;
      const double* Dm = dphi + mm * NQ3;
      const double* Dn = dphi + nn * NQ3;
      constexpr int SD = NP * NQ3;
      for (int q = 0; q < NQ3; ++q) {
        const double wj = m.wtet[q] * aJ;
        double tv;
        if (i == j)
          tv = l2m * Dm[i * SD + q] * Dn[i * SD + q] + mu * Dm[oi * SD + q] * Dn[oi * SD + q] +
               mu * Dm[oi2 * SD + q] * Dn[oi2 * SD + q];
        else
          tv = lam * Dm[i * SD + q] * Dn[j * SD + q] + mu * Dm[j * SD + q] * Dn[i * SD + q];
        s = s + wj * tv;
      }
      const size_t o = kpos(N3, i * NP + mm, j * NP + nn);
      Ks[o] = Ks[o] + s;
      if (keepA) As[o] = As[o] + s;
    }
    // faces (MAKE_STIFFNESS_FACE, assemble_element :404-661; insertion
    // src/matrix_free.f90:371-441)
    for (int f = 0; f < 4; ++f) {
      const int e2 = m.neigh[t * 4 + f];
      int Q = -1;
      if (e2 >= 0) {
        Q = m.tet_poly[e2];
        if (Q == Pp) continue;  // intra-polyhedron face
      }
      if (e2 == -2) continue;  // Neumann: no stiffness contribution
      __syncthreads();
      for (int q = tid; q < NQ2; q += blockDim.x) face_point(Fk, f, &m.qtri[q * 4], &spt[q * 3]);
      __syncthreads();
      double bbQ[6];
      if (Q >= 0)
        for (int d = 0; d < 6; ++d) bbQ[d] = m.bbox[Q * 6 + d];
      for (int e = tid; e < NP * NQ2; e += blockDim.x) {
        const int fb = e / NQ2, q = e % NQ2;
        double ph, d0, d1, d2;
        basis_at(&sbl[fb * 3], bbP, &spt[q * 3], ph, d0, d1, d2);
        p1[e] = ph;
        g1[(0 * NP + fb) * NQ2 + q] = d0;
        g1[(1 * NP + fb) * NQ2 + q] = d1;
        g1[(2 * NP + fb) * NQ2 + q] = d2;
        if (Q >= 0) {
          basis_at(&sbl[fb * 3], bbQ, &spt[q * 3], ph, d0, d1, d2);
          p2[e] = ph;
          g2[(0 * NP + fb) * NQ2 + q] = d0;
          g2[(1 * NP + fb) * NQ2 + q] = d1;
          g2[(2 * NP + fb) * NQ2 + q] = d2;
        }
      }
      __syncthreads();
      const double* nrm = &m.normal[(t * 4 + f) * 3];
      const double n1 = nrm[0], n2 = nrm[1], n3 = nrm[2];
      const double ar = m.area[t * 4 + f];
      const double Dbar = lam + 2 * mu;
      const double hmin = (Q >= 0) ? (hkP < m.hk[Q] ? hkP : m.hk[Q]) : hkP;
      const double sigma = alpha * (P * P) / hmin * Dbar;
      const double d11 = fabs(n1) * fabs(n1) + 0.5 * (fabs(n2) * fabs(n2)) + 0.5 * (fabs(n3) * fabs(n3));
      const double d22 = fabs(n2) * fabs(n2) + 0.5 * (fabs(n1) * fabs(n1)) + 0.5 * (fabs(n3) * fabs(n3));
      const double d33 = fabs(n3) * fabs(n3) + 0.5 * (fabs(n1) * fabs(n1)) + 0.5 * (fabs(n2) * fabs(n2));
      const double dd[3] = {d11, d22, d33};
      const bool interior = Q >= 0;
      double* Kb = interior ? K + (size_t)m.face_blk[t * 4 + f] * D::BS : nullptr;
      double* Ab = interior && keepA ? A + (size_t)m.face_blk[t * 4 + f] * D::BS : nullptr;
      constexpr int SG = NP * NQ2;
      for (int e = tid; e < 9 * NP * NP; e += blockDim.x) {
        const int i = e / (3 * NP * NP), j = (e / (NP * NP)) % 3;
        const int mm = (e / NP) % NP, nn = e % NP;
        const double ni = nrm[i < j ? i : j], nj = nrm[i < j ? j : i];  // reference order n(min)*n(max)
        double s1 = 0.0, i1 = 0.0, i1t = 0.0, i2 = 0.0, it2 = 0.0, s2 = 0.0;
        for (int q = 0; q < NQ2; ++q) {
          const double w = m.wtri[q];
          // S_E1 (:476-490)
          double ts = (i == j) ? (p1[mm * NQ2 + q] * p1[nn * NQ2 + q]) * dd[i]
                               : 0.5 * p1[mm * NQ2 + q] * p1[nn * NQ2 + q] * ni * nj;
          s1 = s1 + sigma * w * ts * ar;
          // I(i,j,m,n) = sigma(grad_m) : (phi_n (x) n) with the component pattern
          // of :504-515; a = grad side, b = phi side.
          auto itemp = [&](const double* G, int gm, double phv, int ii, int jj) -> double {
            if (ii == jj) {
              if (ii == 0)
                return l2m * G[0 * SG + gm * NQ2 + q] * phv * n1 + mu * G[1 * SG + gm * NQ2 + q] * phv * n2 +
                       mu * G[2 * SG + gm * NQ2 + q] * phv * n3;
              if (ii == 1)
                return l2m * G[1 * SG + gm * NQ2 + q] * phv * n2 + mu * G[0 * SG + gm * NQ2 + q] * phv * n1 +
                       mu * G[2 * SG + gm * NQ2 + q] * phv * n3;
              return l2m * G[2 * SG + gm * NQ2 + q] * phv * n3 + mu * G[1 * SG + gm * NQ2 + q] * phv * n2 +
                     mu * G[0 * SG + gm * NQ2 + q] * phv * n1;
            }
            return lam * G[ii * SG + gm * NQ2 + q] * phv * nrm[jj] +
                   mu * G[jj * SG + gm * NQ2 + q] * phv * nrm[ii];
          };
          const double tI = itemp(g1, mm, p1[nn * NQ2 + q], i, j);
          const double tIt = itemp(g1, nn, p1[mm * NQ2 + q], j, i);
          if (!interior) {
            i1 = i1 + w * ar * tI;
            i1t = i1t + w * ar * tIt;
          } else {
            i1 = i1 + 0.5 * w * ar * tI;
            i1t = i1t + 0.5 * w * ar * tIt;
            const double tI2 = itemp(g1, mm, p2[nn * NQ2 + q], i, j);
            i2 = i2 - 0.5 * w * ar * tI2;
            const double tIT2 = itemp(g2, nn, p1[mm * NQ2 + q], i, j);
            it2 = it2 + 0.5 * w * ar * tIT2;
            double t3 = (i == j) ? (p1[mm * NQ2 + q] * p2[nn * NQ2 + q]) * dd[i]
                                 : 0.5 * p1[mm * NQ2 + q] * p2[nn * NQ2 + q] * ni * nj;
            s2 = s2 - sigma * w * t3 * ar;
          }
        }
        const size_t o = kpos(N3, i * NP + mm, j * NP + nn);
        Ks[o] = Ks[o] + s1;
        if (keepA) As[o] = As[o] + s1;
        Ks[o] = Ks[o] + theta * i1 - i1t;
        if (interior) {
          Kb[o] = Kb[o] + theta * i2 + s2;
          if (keepA) Ab[o] = Ab[o] + s2;
          Kb[o] = Kb[o] - it2;
        }
      }
    }
  }
}

// Operator build, factored form (same operator, fewer flops).  The face and
// volume terms of MAKE_STIFFNESS_VOLUME / MAKE_STIFFNESS_FACE are bilinear in
// the basis values, so each is a combination of a few quadrature sums per
// basis pair (m, n):
//   volume:  D_ab = sum_q w|J| d_a phi_m d_b phi_n  (a, b = x, y, z)
//     V_ii = (l+2m) D_ii + m sum_{a != i} D_aa,  V_ij = l D_ij + m D_ji;
//   face:    PP = sum w phiP_m phiP_n,  PQ = sum w phiP_m phiQ_n,
//            BP_a = sum w d_a phiP_m phiP_n,  BPt_a = sum w d_a phiP_n phiP_m,
//            CP_a = sum w d_a phiP_m phiQ_n,  DQ_a = sum w d_a phiQ_n phiP_m,
//     with T(i,j;X) = i==j ? (l+2m) n_i X_i + m sum_{a != i} n_a X_a
//                          : l n_j X_i + m n_i X_j,
//     S+ = s ar PP c_ij, I+ = k ar T(i,j;BP), I+^T = k ar T(j,i;BPt),
//     I- = -ar/2 T(i,j;CP), I-^T = ar/2 T(i,j;DQ), S- = -s ar PQ c_ij,
//     c_ii = dd_i, c_ij = n_i n_j / 2, k = 1 (Dirichlet) or 1/2 (interior);
// inserted like k_build_operator (per tet: volume, then faces 1..4).  Each
// thread owns basis pairs (m, n) for the whole polyhedron and keeps its
// self-block entries in registers, so every entry accumulates in one fixed
// order; sums are regrouped relative to the reference, so entries agree with
// it to rounding (1e-15 relative), not bit for bit.  The mass blocks are
// summed exactly as k_build_operator does.
// Threads per workgroup of k_build_factored: at p <= 2 the basis pairs
// rounded up to whole waves (one pair per thread: 64 and 128 threads, the
// build 3x and 1.6x faster than with 256); from p = 3 on 256 threads of two
// or five pairs, two workgroups per CU at p = 3 (448 threads of one pair left
// room for one and measured 126 ms against 113 ms at 64,000 polyhedra) and no
// register spills at p = 4.
template <int P>
constexpr int factored_threads() {
  constexpr int pairs = np_of(P) * np_of(P);
  return pairs <= 128 ? (pairs + 63) / 64 * 64 : 256;
}

template <int P>
__global__ __launch_bounds__(factored_threads<P>()) void k_build_factored(MeshDev m, double alpha, double theta,
                                                         double* __restrict__ K,
                                                         double* __restrict__ A,
                                                         double* __restrict__ M,
                                                         double* __restrict__ Mm, int poly0) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3, NQ3 = D::NQ3, NQ2 = D::NQ2;
  constexpr int VOLSZ = 4 * NP * NQ3, FACESZ = 8 * NP * NQ2;
  constexpr int SHSZ = VOLSZ > FACESZ ? VOLSZ : FACESZ;
  constexpr int PAIRS = NP * NP;
  constexpr int TPB = factored_threads<P>();
  constexpr int PPT = (PAIRS + TPB - 1) / TPB;  // pairs per thread
  __shared__ double sh[SHSZ];
  __shared__ double spt[NQ3 * 3];
  __shared__ double swq[NQ3];
  __shared__ int sbl[NP * 3];
  const int tid = threadIdx.x;
  const int Pp = poly0 + blockIdx.x;
  const bool keepA = A != nullptr;
  for (int i = tid; i < NP * 3; i += blockDim.x) sbl[i] = m.blist[i];
  double bbP[6];
  for (int d = 0; d < 6; ++d) bbP[d] = m.bbox[Pp * 6 + d];
  const double hkP = m.hk[Pp];
  double* Ks = K + (size_t)m.row_ptr[Pp] * D::BS;
  double* As = keepA ? A + (size_t)m.row_ptr[Pp] * D::BS : nullptr;
  double* phi = sh;               // [NP][NQ3]
  double* dphi = sh + NP * NQ3;   // [3][NP][NQ3]
  double* p1 = sh;                // [NP][NQ2]
  double* p2 = sh + NP * NQ2;
  double* g1 = sh + 2 * NP * NQ2; // [3][NP][NQ2]
  double* g2 = sh + 5 * NP * NQ2;
  double* MP = M + (size_t)Pp * NP * NP;
  double* MmP = Mm + (size_t)Pp * NP * NP;
  for (int a = m.ptet_ptr[Pp]; a < m.ptet_ptr[Pp + 1]; ++a) {
    const int t = m.ptet[a];
    double v[12], Fk[12], Jdet;
    for (int i = 0; i < 12; ++i) v[i] = m.tet_xyz[(size_t)t * 12 + i];
    affine_map(v, Fk, Jdet);
    const double aJ = fabs(Jdet);
    const double rho = m.rho[t], lam = m.lam[t], mu = m.mu[t];
    const double l2m = lam + 2 * mu;
    __syncthreads();
    for (int q = tid; q < NQ3; q += blockDim.x) {
      map_point(Fk, &m.qtet[q * 4], &spt[q * 3]);
      swq[q] = m.wtet[q] * aJ;
    }
    __syncthreads();
    for (int e = tid; e < NP * NQ3; e += blockDim.x) {
      const int f = e / NQ3, q = e % NQ3;
      double ph, d0, d1, d2;
      basis_at(&sbl[f * 3], bbP, &spt[q * 3], ph, d0, d1, d2);
      phi[e] = ph;
      dphi[(0 * NP + f) * NQ3 + q] = d0;
      dphi[(1 * NP + f) * NQ3 + q] = d1;
      dphi[(2 * NP + f) * NQ3 + q] = d2;
    }
    __syncthreads();
    constexpr int SD = NP * NQ3;
#pragma unroll 1
    for (int k = 0; k < PPT; ++k) {
      const int pr = tid + TPB * k;
      if (pr >= PAIRS) continue;
      const int mm = pr / NP, nn = pr % NP;
      double sm = 0.0, Dv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll 2
      for (int q = 0; q < NQ3; ++q) {
        // mass exactly as k_build_operator sums it
        sm = sm + 1.0 * m.wtet[q] * aJ * phi[mm * NQ3 + q] * phi[nn * NQ3 + q];
        const double wq = swq[q];
        const double gm0 = dphi[mm * NQ3 + q], gm1 = dphi[SD + mm * NQ3 + q],
                     gm2 = dphi[2 * SD + mm * NQ3 + q];
        const double gn0 = wq * dphi[nn * NQ3 + q], gn1 = wq * dphi[SD + nn * NQ3 + q],
                     gn2 = wq * dphi[2 * SD + nn * NQ3 + q];
        Dv[0][0] += gm0 * gn0; Dv[0][1] += gm0 * gn1; Dv[0][2] += gm0 * gn2;
        Dv[1][0] += gm1 * gn0; Dv[1][1] += gm1 * gn1; Dv[1][2] += gm1 * gn2;
        Dv[2][0] += gm2 * gn0; Dv[2][1] += gm2 * gn1; Dv[2][2] += gm2 * gn2;
      }
      MmP[pr] = MmP[pr] + sm;
      MP[pr] = MP[pr] + sm * rho;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double val;
          if (i == j)
            val = l2m * Dv[i][i] + mu * Dv[(i + 1) % 3][(i + 1) % 3] + mu * Dv[(i + 2) % 3][(i + 2) % 3];
          else
            val = lam * Dv[i][j] + mu * Dv[j][i];
          const size_t o = kpos(N3, i * NP + mm, j * NP + nn);
          Ks[o] = Ks[o] + val;
          if (keepA) As[o] = As[o] + val;
        }
    }
#pragma unroll 1
    for (int f = 0; f < 4; ++f) {
      const int e2 = m.neigh[t * 4 + f];
      int Q = -1;
      if (e2 >= 0) {
        Q = m.tet_poly[e2];
        if (Q == Pp) continue;  // intra-polyhedron face
      }
      if (e2 == -2) continue;  // Neumann: no stiffness contribution
      __syncthreads();
      for (int q = tid; q < NQ2; q += blockDim.x) face_point(Fk, f, &m.qtri[q * 4], &spt[q * 3]);
      __syncthreads();
      double bbQ[6];
      if (Q >= 0)
        for (int d = 0; d < 6; ++d) bbQ[d] = m.bbox[Q * 6 + d];
      for (int e = tid; e < NP * NQ2; e += blockDim.x) {
        const int fb = e / NQ2, q = e % NQ2;
        double ph, d0, d1, d2;
        basis_at(&sbl[fb * 3], bbP, &spt[q * 3], ph, d0, d1, d2);
        p1[e] = ph;
        g1[(0 * NP + fb) * NQ2 + q] = d0;
        g1[(1 * NP + fb) * NQ2 + q] = d1;
        g1[(2 * NP + fb) * NQ2 + q] = d2;
        if (Q >= 0) {
          basis_at(&sbl[fb * 3], bbQ, &spt[q * 3], ph, d0, d1, d2);
          p2[e] = ph;
          g2[(0 * NP + fb) * NQ2 + q] = d0;
          g2[(1 * NP + fb) * NQ2 + q] = d1;
          g2[(2 * NP + fb) * NQ2 + q] = d2;
        }
      }
      __syncthreads();
      const double* nrm = &m.normal[(t * 4 + f) * 3];
      const double n1 = nrm[0], n2 = nrm[1], n3 = nrm[2];
      const double nv[3] = {n1, n2, n3};
      const double ar = m.area[t * 4 + f];
      const double Dbar = lam + 2 * mu;
      const double hmin = (Q >= 0) ? (hkP < m.hk[Q] ? hkP : m.hk[Q]) : hkP;
      const double sigma = alpha * (P * P) / hmin * Dbar;
      const double d11 = fabs(n1) * fabs(n1) + 0.5 * (fabs(n2) * fabs(n2)) + 0.5 * (fabs(n3) * fabs(n3));
      const double d22 = fabs(n2) * fabs(n2) + 0.5 * (fabs(n1) * fabs(n1)) + 0.5 * (fabs(n3) * fabs(n3));
      const double d33 = fabs(n3) * fabs(n3) + 0.5 * (fabs(n1) * fabs(n1)) + 0.5 * (fabs(n2) * fabs(n2));
      const double dd[3] = {d11, d22, d33};
      const bool interior = Q >= 0;
      const double kk = interior ? 0.5 : 1.0;
      double* Kb = interior ? K + (size_t)m.face_blk[t * 4 + f] * D::BS : nullptr;
      double* Ab = interior && keepA ? A + (size_t)m.face_blk[t * 4 + f] * D::BS : nullptr;
      constexpr int SG = NP * NQ2;
      // T(i,j;X): the traction pattern of itemp
      auto Tij = [&](int i, int j, const double* X) -> double {
        if (i == j) return l2m * nv[i] * X[i] + mu * nv[(i + 1) % 3] * X[(i + 1) % 3] +
                           mu * nv[(i + 2) % 3] * X[(i + 2) % 3];
        return lam * nv[j] * X[i] + mu * nv[i] * X[j];
      };
#pragma unroll 1
      for (int k = 0; k < PPT; ++k) {
        const int pr = tid + TPB * k;
        if (pr >= PAIRS) continue;
        const int mm = pr / NP, nn = pr % NP;
        double PPs = 0.0, PQs = 0.0, BP[3] = {0, 0, 0}, BPt[3] = {0, 0, 0}, CP[3] = {0, 0, 0},
               DQ[3] = {0, 0, 0};
#pragma unroll 2
        for (int q = 0; q < NQ2; ++q) {
          const double w = m.wtri[q];
          const double pm = p1[mm * NQ2 + q], pn = p1[nn * NQ2 + q];
          const double wpm = w * pm, wpn = w * pn;
          PPs += wpm * pn;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            BP[d] += g1[d * SG + mm * NQ2 + q] * wpn;
            BPt[d] += g1[d * SG + nn * NQ2 + q] * wpm;
          }
          if (interior) {
            const double qn = p2[nn * NQ2 + q];
            const double wqn = w * qn;
            PQs += wpm * qn;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              CP[d] += g1[d * SG + mm * NQ2 + q] * wqn;
              DQ[d] += g2[d * SG + nn * NQ2 + q] * wpm;
            }
          }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const double cij = (i == j) ? dd[i] : 0.5 * nv[i] * nv[j];
            const double s1 = sigma * ar * PPs * cij;
            const double i1 = kk * ar * Tij(i, j, BP);
            const double i1t = kk * ar * Tij(j, i, BPt);
            const size_t o = kpos(N3, i * NP + mm, j * NP + nn);
            Ks[o] = Ks[o] + s1;
            if (keepA) As[o] = As[o] + s1;
            Ks[o] = Ks[o] + theta * i1 - i1t;
            if (interior) {
              const double i2 = -0.5 * ar * Tij(i, j, CP);
              const double it2 = 0.5 * ar * Tij(i, j, DQ);
              const double s2 = -(sigma * ar * PQs * cij);
              Kb[o] = Kb[o] + theta * i2 + s2;
              if (keepA) Ab[o] = Ab[o] + s2;
              Kb[o] = Kb[o] - it2;
            }
          }
      }
    }
  }

}

// Cholesky of both mass blocks of a polyhedron (cholesky,
// src/vet_mat_operations.f90:84-120), one wave per polyhedron: the block is
// staged in LDS by coalesced loads, each row j of the packed upper factor is
// computed across the wave (every lane forms the pivot, lane i > j the entry
// (j, i)) with each entry's sum over k in the serial order, and the factor
// leaves as one coalesced store.  status[0] set to 1 if a block is not SPD.
template <int P>
__global__ __launch_bounds__(64) void k_cholesky(const double* __restrict__ M,
                                                 const double* __restrict__ Mm,
                                                 double* __restrict__ RM, double* __restrict__ RMm,
                                                 int n_poly, int* __restrict__ status) {
  using D = Dims<P>;
  constexpr int NP = D::NP, NPK = D::NPK;
  static_assert(NP <= 64, "one lane per row of the factor");
  __shared__ double sA[NP * NP];
  __shared__ double sR[NPK];
  const int lane = threadIdx.x;
  const int Pp = blockIdx.x;
  if (Pp >= n_poly) return;
  for (int which = 0; which < 2; ++which) {
    const double* A = (which ? Mm : M) + (size_t)Pp * NP * NP;
    double* R = (which ? RMm : RM) + (size_t)Pp * NPK;
    __syncthreads();  // the previous block's LDS reads are done
    for (int e = lane; e < NP * NP; e += 64) sA[e] = A[e];
    __syncthreads();
    for (int j = 0; j < NP; ++j) {
      double s = sA[j * NP + j];
      for (int k = 0; k < j; ++k) {
        const double rkj = sR[pk_off(NP, k) + (j - k)];
        s = s - rkj * rkj;
      }
      if (!(s > 0.0)) {  // the same value on every lane: the wave leaves together
        if (lane == 0) atomicOr(status, 1);
        return;
      }
      const double rjj = sqrt(s);
      double rji = 0.0;
      if (lane > j && lane < NP) {
        double s2 = sA[j * NP + lane];
        for (int k = 0; k < j; ++k)
          s2 = s2 - sR[pk_off(NP, k) + (j - k)] * sR[pk_off(NP, k) + (lane - k)];
        rji = s2 / rjj;
      }
      // row j is read from the next step on (this step reads rows k < j only)
      if (lane == j) sR[pk_off(NP, j)] = rjj;
      if (lane > j && lane < NP) sR[pk_off(NP, j) + (lane - j)] = rji;
      __syncthreads();
    }
    for (int e = lane; e < NPK; e += 64) R[e] = sR[e];
  }
}

// ---------------------------------------------------------------------------
// RHS (MAKE_RHS_FREE, src/matrix_free.f90:467-843) and L2 projection
// (COMPUTE_MODAL_COEFFICIENTS_FREE, :847-997).  One 128-thread workgroup per
// polyhedron, entry (i, m) owned by one thread.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(128) void k_rhs(MeshDev m, int prob, double alpha, double theta,
                                             double* __restrict__ rhs,
                                             const double* __restrict__ tabv = nullptr,
                                             const double* __restrict__ tabf = nullptr) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3, NQ3 = D::NQ3, NQ2 = D::NQ2;
  __shared__ double phi[NP * NQ3];
  __shared__ double gb[3 * NP * NQ2];
  __shared__ double spt[NQ3 * 3];
  __shared__ double sval[NQ3 * 3];
  __shared__ int sbl[NP * 3];
  const int tid = threadIdx.x;
  const int Pp = blockIdx.x;
  for (int i = tid; i < NP * 3; i += blockDim.x) sbl[i] = m.blist[i];
  double bbP[6];
  for (int d = 0; d < 6; ++d) bbP[d] = m.bbox[Pp * 6 + d];
  double* r = rhs + (size_t)Pp * N3;
  double racc = 0.0;  // running rhs entry of this thread
  for (int a = m.ptet_ptr[Pp]; a < m.ptet_ptr[Pp + 1]; ++a) {
    const int t = m.ptet[a];
    double v[12], Fk[12], Jdet;
    for (int i = 0; i < 12; ++i) v[i] = m.tet_xyz[(size_t)t * 12 + i];
    affine_map(v, Fk, Jdet);
    const double aJ = fabs(Jdet);
    const double rho = m.rho[t], lam = m.lam[t], mu = m.mu[t];
    __syncthreads();
    for (int q = tid; q < NQ3; q += blockDim.x) {
      map_point(Fk, &m.qtet[q * 4], &spt[q * 3]);
      if (tabv) {  // values of the caller's forcing at k_rhs_points' volume points
        for (int c = 0; c < 3; ++c) sval[q * 3 + c] = tabv[((size_t)t * NQ3 + q) * 3 + c];
      } else {
        forcing(prob, lam, mu, &spt[q * 3], rho, &sval[q * 3]);
      }
    }
    __syncthreads();
    for (int e = tid; e < NP * NQ3; e += blockDim.x) {
      const int f = e / NQ3, q = e % NQ3;
      double d0, d1, d2;
      basis_at(&sbl[f * 3], bbP, &spt[q * 3], phi[e], d0, d1, d2);
    }
    __syncthreads();
    if (tid < N3) {
      const int i = tid / NP, mm = tid % NP;
      double s = 0.0;
      for (int q = 0; q < NQ3; ++q) s = s + aJ * m.wtet[q] * sval[q * 3 + i] * phi[mm * NQ3 + q];
      racc = racc + s;
    }
    for (int f = 0; f < 4; ++f) {
      const int e2 = m.neigh[t * 4 + f];
      if (e2 >= 0) continue;  // only boundary faces carry data
      __syncthreads();
      const double* nrm = &m.normal[(t * 4 + f) * 3];
      const int tag = m.bc_space[t * 4 + f];
      for (int q = tid; q < NQ2; q += blockDim.x) {
        double xr[3];
        face_point(Fk, f, &m.qtri[q * 4], &spt[q * 3]);
        face_point_rhs(Fk, f, &m.qtri[q * 4], xr);
        if (tabf) {  // the caller's Dirichlet / Neumann data at k_rhs_points' face points
          for (int c = 0; c < 3; ++c) sval[q * 3 + c] = tabf[(((size_t)t * 4 + f) * NQ2 + q) * 3 + c];
        } else if (e2 == -1) {
          dirichlet(prob, xr, tag, &sval[q * 3]);
        } else {
          neumann(prob, lam, mu, nrm, xr, tag, &sval[q * 3]);
        }
      }
      __syncthreads();
      for (int e = tid; e < NP * NQ2; e += blockDim.x) {
        const int fb = e / NQ2, q = e % NQ2;
        basis_at(&sbl[fb * 3], bbP, &spt[q * 3], phi[e], gb[(0 * NP + fb) * NQ2 + q],
                 gb[(1 * NP + fb) * NQ2 + q], gb[(2 * NP + fb) * NQ2 + q]);
      }
      __syncthreads();
      if (tid < N3) {
        const int i = tid / NP, mm = tid % NP;
        const double n1 = nrm[0], n2 = nrm[1], n3 = nrm[2];
        const double ar = m.area[t * 4 + f];
        const double l2m = lam + 2 * mu;
        const double sigma = alpha * (P * P) / m.hk[Pp] * (lam + 2 * mu);
        double s = 0.0;
        for (int q = 0; q < NQ2; ++q) {
          const double w = m.wtri[q];
          const double* dd = &sval[q * 3];
          const double ph = phi[mm * NQ2 + q];
          if (e2 == -1) {
            const double gg1 = gb[(0 * NP + mm) * NQ2 + q], gg2 = gb[(1 * NP + mm) * NQ2 + q],
                         gg3 = gb[(2 * NP + mm) * NQ2 + q];
            double tm, t2;
            if (i == 0) {
              tm = l2m * gg1 * dd[0] * n1 + mu * gg2 * dd[0] * n2 + mu * gg3 * dd[0] * n3 +
                   lam * gg1 * dd[1] * n2 + mu * gg2 * dd[1] * n1 + lam * gg1 * dd[2] * n3 +
                   mu * gg3 * dd[2] * n1;
              t2 = ph * dd[0] * (fabs(n1) * fabs(n1) + 0.5 * (fabs(n2) * fabs(n2)) + 0.5 * (fabs(n3) * fabs(n3))) +
                   0.5 * ph * dd[1] * n1 * n2 + 0.5 * ph * dd[2] * n1 * n3;
            } else if (i == 1) {
              tm = l2m * gg2 * dd[1] * n2 + mu * gg1 * dd[1] * n1 + mu * gg3 * dd[1] * n3 +
                   lam * gg2 * dd[0] * n1 + mu * gg1 * dd[0] * n2 + lam * gg2 * dd[2] * n3 +
                   mu * gg3 * dd[2] * n2;
              t2 = ph * dd[1] * (fabs(n2) * fabs(n2) + 0.5 * (fabs(n1) * fabs(n1)) + 0.5 * (fabs(n3) * fabs(n3))) +
                   0.5 * ph * dd[0] * n1 * n2 + 0.5 * ph * dd[2] * n2 * n3;
            } else {
              tm = l2m * gg3 * dd[2] * n3 + mu * gg1 * dd[2] * n1 + mu * gg2 * dd[2] * n2 +
                   lam * gg3 * dd[0] * n1 + mu * gg1 * dd[0] * n3 + lam * gg3 * dd[1] * n2 +
                   mu * gg2 * dd[1] * n3;
              t2 = ph * dd[2] * (fabs(n3) * fabs(n3) + 0.5 * (fabs(n1) * fabs(n1)) + 0.5 * (fabs(n2) * fabs(n2))) +
                   0.5 * ph * dd[0] * n1 * n3 + 0.5 * ph * dd[1] * n2 * n3;
            }
            s = s + theta * w * tm * ar + sigma * w * t2 * ar;
          } else {
            s = s + w * ph * dd[i] * ar;
          }
        }
        if (s != 0.0) racc = racc + s;
      }
    }
  }
  if (tid < N3) r[tid] = racc;
}

// Physical points where MAKE_RHS_FREE evaluates its data (as k_rhs computes
// them): volume quadrature points vpts[(t*NQ3 + q)*3 + d] (forcing) and the
// data points of every tet face fpts[((t*4 + f)*NQ2 + q)*3 + d] (Dirichlet /
// Neumann data; src/matrix_free.f90:711, :781-801).  One thread per point.
template <int P>
__global__ void k_rhs_points(MeshDev m, int n_tet, double* __restrict__ vpts,
                             double* __restrict__ fpts) {
  using D = Dims<P>;
  constexpr int NQ3 = D::NQ3, NQ2 = D::NQ2, PER = NQ3 + 4 * NQ2;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n_tet * PER) return;
  const int t = (int)(i / PER), k = (int)(i % PER);
  double v[12], Fk[12], Jdet;
  for (int j = 0; j < 12; ++j) v[j] = m.tet_xyz[(size_t)t * 12 + j];
  affine_map(v, Fk, Jdet);
  if (k < NQ3) {
    map_point(Fk, &m.qtet[k * 4], vpts + ((size_t)t * NQ3 + k) * 3);
  } else {
    const int f = (k - NQ3) / NQ2, q = (k - NQ3) % NQ2;
    face_point_rhs(Fk, f, &m.qtri[q * 4], fpts + (((size_t)t * 4 + f) * NQ2 + q) * 3);
  }
}

// Forward/backward substitution with the packed factor in the reference's
// order (solve_LU with L = R^T, src/vet_mat_operations.f90:123-153); x in place.
// M^-1 w = R^-1 R^-T w across one wave, lane = (component cc, row ii) of a
// 3Np <= 64 vector held one entry per lane (lanes >= 3Np pass through): the
// forward sweep broadcasts y_j = s_j / R_jj from lane (cc, j) and the rows
// below subtract R_ji y_j (ascending j, as solve_packed); the backward sweep
// runs the columns from the last, the rows above subtracting R_ij x_j.  NP
// dependent steps per sweep instead of Np(Np+1)/2 on one lane.
template <int NP>
__device__ __forceinline__ double solve_wave(const double* __restrict__ R, double w, int lane) {
  const bool row = lane < 3 * NP;
  const int cc = lane / NP, ii = lane % NP;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int rj = pk_off(NP, j);
    const double yj = __shfl(w, row ? cc * NP + j : j, 64) / R[rj];
    if (row && ii == j)
      w = yj;
    else if (row && ii > j)
      w = w - R[rj + (ii - j)] * yj;
  }
  const int ri = pk_off(NP, row ? ii : 0);
#pragma unroll
  for (int j = NP - 1; j >= 0; --j) {
    const double xj = __shfl(w, row ? cc * NP + j : j, 64) / R[pk_off(NP, j)];
    if (row && ii == j)
      w = xj;
    else if (row && ii < j)
      w = w - R[ri + (j - ii)] * xj;
  }
  return w;
}

template <int NP>
__device__ __forceinline__ void solve_packed(const double* __restrict__ R, double* x) {
  // loops kept rolled: fully unrolled, the compiler hoists all Np(Np+1)/2
  // factor entries into registers and the step kernel drops to one wave/SIMD
  x[0] = x[0] / R[0];
#pragma unroll 1
  for (int i = 1; i < NP; ++i) {
    double s = x[i];
#pragma unroll 1
    for (int j = 0; j < i; ++j) s = s - R[pk_off(NP, j) + (i - j)] * x[j];
    x[i] = s / R[pk_off(NP, i)];
  }
  x[NP - 1] = x[NP - 1] / R[pk_off(NP, NP - 1)];
#pragma unroll 1
  for (int i = NP - 2; i >= 0; --i) {
    double s = x[i];
#pragma unroll 1
    for (int j = i + 1; j < NP; ++j) s = s - R[pk_off(NP, i) + (j - i)] * x[j];
    x[i] = s / R[pk_off(NP, i)];
  }
}

template <int P>
__global__ __launch_bounds__(128) void k_project(MeshDev m, int fn, const double* __restrict__ RMm,
                                                 double* __restrict__ out,
                                                 const double* __restrict__ tabv) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3, NQ3 = D::NQ3;
  __shared__ double phi[NP * NQ3];
  __shared__ double spt[NQ3 * 3];
  __shared__ double sval[NQ3 * 3];
  __shared__ double sb[N3];
  __shared__ int sbl[NP * 3];
  const int tid = threadIdx.x;
  const int Pp = blockIdx.x;
  for (int i = tid; i < NP * 3; i += blockDim.x) sbl[i] = m.blist[i];
  double bbP[6];
  for (int d = 0; d < 6; ++d) bbP[d] = m.bbox[Pp * 6 + d];
  double acc = 0.0;
  for (int a = m.ptet_ptr[Pp]; a < m.ptet_ptr[Pp + 1]; ++a) {
    const int t = m.ptet[a];
    double v[12], Fk[12], Jdet;
    for (int i = 0; i < 12; ++i) v[i] = m.tet_xyz[(size_t)t * 12 + i];
    affine_map(v, Fk, Jdet);
    const double aJ = fabs(Jdet);
    __syncthreads();
    for (int q = tid; q < NQ3; q += blockDim.x) {
      map_point(Fk, &m.qtet[q * 4], &spt[q * 3]);
      if (tabv) {  // the caller's field at k_rhs_points' volume points
        for (int c = 0; c < 3; ++c) sval[q * 3 + c] = tabv[((size_t)t * NQ3 + q) * 3 + c];
      } else {
        field(fn, &spt[q * 3], &sval[q * 3]);
      }
    }
    __syncthreads();
    for (int e = tid; e < NP * NQ3; e += blockDim.x) {
      const int f = e / NQ3, q = e % NQ3;
      double d0, d1, d2;
      basis_at(&sbl[f * 3], bbP, &spt[q * 3], phi[e], d0, d1, d2);
    }
    __syncthreads();
    if (tid < N3) {
      const int i = tid / NP, mm = tid % NP;
      for (int q = 0; q < NQ3; ++q) acc = acc + aJ * m.wtet[q] * sval[q * 3 + i] * phi[mm * NQ3 + q];
    }
  }
  if (tid < N3) sb[tid] = acc;
  __syncthreads();
  if (tid < 3) solve_packed<NP>(RMm + (size_t)Pp * D::NPK, sb + tid * NP);
  __syncthreads();
  if (tid < N3) out[(size_t)Pp * N3 + tid] = sb[tid];
}

// ---------------------------------------------------------------------------
// The leapfrog step (TIME_STEP_MATRIX_FREE, src/matrix_free.f90:1230-1385;
// FIRST_TIME_STEP_MATRIX_FREE :1100-1227) fused in one kernel: one wave per
// polyhedron, lane r = row r of the block row.  Per block the wave streams
// the block's column pairs (16 B per lane per load), with the neighbour's
// coefficients as wave-uniform (scalar) loads: the halo exchange of
// src/exchange_data.f90 becomes the on-device gather col[b].  Then
// w = -Ku + F_s + F_d g, three triangular-solve pairs with the packed mass
// factor, and the update.
//   MODE 0: out = (c1*x + 2*u) - wa            (c1 = dt^2, u = x_apply, wa = u^{n-1})
//   MODE 1: out = (c1*x + u) + c2*wa           (c1 = dt^2/2, u = u^0, wa = v^0, c2 = dt)
//   MODE 2: part[P] = sum_r u_r (K u)_r        (DG error: K = A_dg, u = uh - uex)
// ---------------------------------------------------------------------------
// Broadcast lane `l`'s double to the whole wave (v_readlane on both halves).
__device__ __forceinline__ double lane_bcast(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// Polyhedron of this workgroup.  Workgroups are dealt to the 8 XCDs round
// robin (workgroup b runs on XCD b % 8), each XCD with its own L2; the grid is
// padded to a multiple of 8 and XCD x is given the x-th eighth of the
// polyhedra, so neighbours -- whose coefficient blocks a row re-reads -- mostly
// share an L2.  A bijection of [0, grid); ids >= n_poly are padding.
__device__ inline int step_poly() {
  const int per = (int)(gridDim.x >> 3);
  return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}
inline unsigned step_grid(int n_poly) { return 8u * (unsigned)((n_poly + 7) / 8); }

// With a dispatch order (order[i] = polyhedron of slot i, step_grid entries,
// padding slots holding n_poly): each XCD's share of the polyhedra is walked
// longest block row first, so on meshes of uneven polyhedra the long rows
// start early instead of forming the tail of the launch.  Every polyhedron is
// still computed by one workgroup exactly as without it.
__device__ inline int step_poly(const int* __restrict__ order) {
  const int i = step_poly();
  return order ? order[i] : i;
}

// UNR = column pairs whose loads are issued together (register footprint vs
// loads in flight).  XS = how the neighbour's coefficients reach the lanes:
// 0 = wave-uniform scalar loads per pair, 1 = one coalesced vector load per
// block (lane c holds x[c]) broadcast with v_readlane, so the inner loop waits
// only on the block's own HBM loads.  Chosen per degree by lymph_autotune.
template <int P, int MODE, int UNR, int XS, int WPP = 1>
__global__ __launch_bounds__(64 * WPP) void k_leapfrog(const double* __restrict__ K,
                                                 const int* __restrict__ row_ptr,
                                                 const int* __restrict__ col,
                                                 const double* __restrict__ Rp,
                                                 const double* __restrict__ rs,
                                                 const double* __restrict__ rd,
                                                 const double* __restrict__ u,
                                                 const double* __restrict__ wa,
                                                 double* __restrict__ out, double g, double c1,
                                                 double c2, const int* __restrict__ gsel,
                                                 const double* __restrict__ gtab, int n_poly,
    const int* __restrict__ order) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3, NPAIR = D::NPAIR;
  constexpr int RPL = (N3 + 63) / 64;
  __shared__ double sw[N3];
  __shared__ double sR[D::NPK];
  extern __shared__ double sblk[];  // WPP > 1: per-block products [block][row]
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int Pp = step_poly(order);
  if (Pp >= n_poly) return;  // padding workgroups of the last XCD share
  if (MODE != 2)
    for (int k = threadIdx.x; k < D::NPK; k += 64 * WPP) sR[k] = Rp[(size_t)Pp * D::NPK + k];
  double acc[RPL];
#pragma unroll
  for (int k = 0; k < RPL; ++k) acc[k] = 0.0;
  const int b0 = row_ptr[Pp], b1 = row_ptr[Pp + 1];
  // WPP waves share a polyhedron's blocks round robin; each block's product
  // goes to LDS and wave 0 adds them in block order, the reference's order
  for (int b = b0 + wv; b < b1; b += WPP) {
    const int cb = __builtin_amdgcn_readfirstlane(col[b]);
    const double* __restrict__ xv = u + (size_t)cb * N3;
    const double* __restrict__ Kb = K + (size_t)b * D::BS;
    double xl[RPL];
    if (XS == 1) {
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int r = lane + 64 * k;
        xl[k] = r < N3 ? xv[r] : 0.0;
      }
    }
    double s[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) s[k] = 0.0;
#pragma unroll 1
    for (int c0 = 0; c0 < NPAIR; c0 += UNR) {
      double2 kv[UNR][RPL];
#pragma unroll
      for (int uu = 0; uu < UNR; ++uu) {
        const int c2i = c0 + uu;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const int r = lane + 64 * k;
          if ((NPAIR % UNR == 0 || c2i < NPAIR) && r < N3)
            kv[uu][k] = load_stream(Kb + ((size_t)c2i * N3 + r) * 2);
        }
      }
#pragma unroll
      for (int uu = 0; uu < UNR; ++uu) {
        const int c2i = c0 + uu;
        if (NPAIR % UNR == 0 || c2i < NPAIR) {
          const bool odd_tail = (N3 % 2 == 1) && (c2i == NPAIR - 1);
          double x0, x1;
          if (XS == 1) {
            const int ca = 2 * c2i, cbb = 2 * c2i + 1;
            const double xa = (RPL > 1 && (ca >> 6)) ? xl[RPL - 1] : xl[0];
            const double xb = (RPL > 1 && (cbb >> 6)) ? xl[RPL - 1] : xl[0];
            x0 = lane_bcast(xa, ca & 63);
            x1 = odd_tail ? 0.0 : lane_bcast(xb, cbb & 63);
          } else {
            x0 = xv[2 * c2i];
            x1 = odd_tail ? 0.0 : xv[2 * c2i + 1];
          }
#pragma unroll
          for (int k = 0; k < RPL; ++k) {
            const int r = lane + 64 * k;
            if (r < N3) {
              s[k] = s[k] + kv[uu][k].x * x0;
              if (!odd_tail) s[k] = s[k] + kv[uu][k].y * x1;
            }
          }
        }
      }
    }
    if constexpr (WPP == 1) {
#pragma unroll
      for (int k = 0; k < RPL; ++k) acc[k] = (b == b0) ? s[k] : acc[k] + s[k];
    } else {
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int r = lane + 64 * k;
        if (r < N3) sblk[(size_t)(b - b0) * N3 + r] = s[k];
      }
    }
  }
  if constexpr (WPP > 1) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int r = lane + 64 * k;
      if (r < N3 && wv == 0) {
        double a = sblk[r];
        for (int j = 1; j < b1 - b0; ++j) a = a + sblk[(size_t)j * N3 + r];
        acc[k] = a;
      }
    }
  }
  const size_t o = (size_t)Pp * N3;
  if (MODE == 2) {
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int r = lane + 64 * k;
      if (r < N3 && wv == 0) sw[r] = acc[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double el = 0.0;
      for (int r = 0; r < N3; ++r) el = el + u[o + r] * sw[r];
      out[Pp] = el;
    }
    return;
  }
  // time factor: one scalar, or per source group (ensembles whose instances
  // carry their own wavelet; the table is filled for this step)
  const double gp = gsel ? gtab[gsel[Pp]] : g;
  if constexpr (P == 3) {
    // one row per lane: wave 0 solves across its lanes (p = 3 only: at p <= 2
    // the packed variants keep solve_packed, and every variant of a degree
    // must give the same bits)
    __syncthreads();  // sR
    if (wv != 0) return;  // no barrier follows
    double w = lane < N3 ? -acc[0] + rs[o + lane] + rd[o + lane] * gp : 0.0;
    w = solve_wave<NP>(sR, w, lane);
    if (lane < N3) {
      if (MODE == 0)
        out[o + lane] = c1 * w + 2.0 * u[o + lane] - wa[o + lane];
      else
        out[o + lane] = c1 * w + u[o + lane] + c2 * wa[o + lane];
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    const int r = lane + 64 * k;
    if (r < N3 && wv == 0) sw[r] = -acc[k] + rs[o + r] + rd[o + r] * gp;
  }
  __syncthreads();
  if (threadIdx.x < 3) solve_packed<NP>(sR, sw + threadIdx.x * NP);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    const int r = lane + 64 * k;
    if (r < N3 && wv == 0) {
      if (MODE == 0)
        out[o + r] = c1 * sw[r] + 2.0 * u[o + r] - wa[o + r];
      else
        out[o + r] = c1 * sw[r] + u[o + r] + c2 * wa[o + r];
    }
  }
}

// Step with every face pair's block stored once, in two launches.  With SIPG
// and one material on both traces of each face, K_QP = K_PQ^T to rounding, so
// only the self blocks and K_PQ for Q > P are kept (Ks; sidx[b] >= 0 the
// stored block of full-row block b, < 0 a block of the neighbour's row), and
// of each self block (K_PP = K_PP^T to rounding) only its upper triangle:
// column pair c2i's rows 0..2c2i+1 packed from entry 2 c2i (c2i + 1)
// (self_upper_off), half the bytes of the full block.
// k_sym_products: one wave per stored block K_PQ, the self blocks first
// (stored index = polyhedron, ssrc[P] its full-row slot) and then the pair
// blocks (pinfo: full-row slot b in P's row, P, Q, slot of (Q, P)), each
// XCD taking a contiguous share of both, so that the block's loads depend on
// nothing loaded: P's product K_PQ u_Q (FMA-free chain over ascending columns,
// as k_leapfrog) into prod[b], and for a pair block the neighbour's product
// K_QP u_P = K_PQ^T u_P into prod[tdst[b]], the slot of block (Q, P) of Q's
// row.  Lane i holds row i of K_PQ and u_P[i]; the products K_PQ[i][c] u_P[i]
// go to LDS in three chunks of CW columns and lane c adds column c's
// products over ascending rows: the same chain Q's own row would run over
// K_QP's columns, so with K_QP = K_PQ^T exactly (the 32-bit copies) the
// transposed product is the full-row product bit for bit.  A self block's
// product is its upper triangle's row product plus the transposed product
// of its strict upper triangle, both with u_P.  Each wave's stores come after
// all of its loads.  k_sym_update: per polyhedron, the row's products summed
// in block order, then M^-1 = S S^T with S = R^-1 (packed like R) as two
// triangular products through LDS, and the update of k_leapfrog.  p <= 3
// (3Np <= 64 rows, whole column pairs).
__host__ __device__ constexpr int self_upper_off(int c2i) { return 2 * c2i * (c2i + 1); }

// LDS ordering within one wave (the rocPRIM wave barrier: fences at
// wavefront scope around a code-motion barrier); enough for a workgroup of
// one wave, and it does not wait on the wave's global loads and stores.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One stored block of k_sym_products (self: a self block's packed upper
// triangle; otherwise a pair block K_PQ with its transposed product).  Kb is
// the block's base plus the lane's offset.  A wave holds its registers until
// its stores are acknowledged, so a pair block's row product is stored before
// the LDS transpose (u_P is loaded with the block: no load follows a store).
template <int P>
__device__ __forceinline__ void sym_block(bool self, const double* __restrict__ Kb, int b, int Pp, int Qq,
                                          int td, const double* __restrict__ u,
                                          double* __restrict__ prod, double* __restrict__ sT,
                                          int lane) {
  using D = Dims<P>;
  constexpr int N3 = D::N3, NPAIR = D::NPAIR, PS = prod_stride(N3);
  constexpr int CW = N3 / 3, CP = CW / 2, LDT = N3 + 1;
  const bool row = lane < N3;
  double2 kv[NPAIR];
#pragma unroll
  for (int c2i = 0; c2i < NPAIR; ++c2i) {
    const bool ld = row && (!self || lane <= 2 * c2i + 1);
    kv[c2i] = ld ? load_stream(Kb + (self ? self_upper_off(c2i) : c2i * N3 * 2))
                 : make_double2(0.0, 0.0);
    if (self && lane == 2 * c2i + 1) kv[c2i].x = 0.0;  // K_PP[lane][lane - 1]: below the diagonal
  }
  const double xl = row ? u[(size_t)Qq * N3 + lane] : 0.0;
  const double up = self ? xl : (row ? u[(size_t)Pp * N3 + lane] : 0.0);  // P's own coefficients
  double s = 0.0;
#pragma unroll
  for (int c2i = 0; c2i < NPAIR; ++c2i) {
    const double x0 = lane_bcast(xl, 2 * c2i), x1 = lane_bcast(xl, 2 * c2i + 1);
    s = s + kv[c2i].x * x0;
    s = s + kv[c2i].y * x1;
  }
  if (!self && lane < PS) prod[(size_t)b * PS + lane] = s;  // lanes past the rows hold zeros
  double st = 0.0;
#pragma unroll
  for (int h = 0; h < 3; ++h) {
    wave_lds_sync();  // the previous chunk's reads are done
    if (row) {
#pragma unroll
      for (int k = 0; k < CP; ++k) {
        sT[(2 * k) * LDT + lane] = kv[h * CP + k].x * up;
        sT[(2 * k + 1) * LDT + lane] = kv[h * CP + k].y * up;
      }
      // the self block's diagonal is in its row product only (a wave's LDS
      // writes land in program order)
      if (self && lane >= h * CW && lane < (h + 1) * CW) sT[(lane - h * CW) * LDT + lane] = 0.0;
    }
    wave_lds_sync();
    if (lane >= h * CW && lane < (h + 1) * CW) {
      const double* __restrict__ colp = sT + (lane - h * CW) * LDT;
      double v = 0.0;
#pragma unroll
      for (int r = 0; r < N3; ++r) v = v + colp[r];
      st = v;
    }
  }
  if (lane < PS) {
    if (self)
      prod[(size_t)b * PS + lane] = s + st;
    else
      prod[(size_t)td * PS + lane] = st;
  }
}

template <int P>
__global__ __launch_bounds__(64) void k_sym_products(const double* __restrict__ Ks,
                                                     const int* __restrict__ ssrc,
                                                     const int4* __restrict__ pinfo,
                                                     const double* __restrict__ u,
                                                     double* __restrict__ prod, int n_poly,
                                                     int n_pair, int self_per, int pair_per) {
  using D = Dims<P>;
  constexpr int N3 = D::N3, NPAIR = D::NPAIR;
  static_assert(N3 <= 64 && N3 % 2 == 0, "one lane per row, whole column pairs");
  static_assert((N3 / 3) % 2 == 0, "whole column pairs per chunk");
  static_assert(self_upper_off(NPAIR) <= D::BS, "packed self block fits its slot");
  __shared__ double sT[(N3 / 3) * (N3 + 1)];  // odd stride: a column read spreads over the banks
  const int lane = threadIdx.x;
  // XCD-aware: workgroup i of XCD x takes self block x self_per + i, or pair
  // x pair_per + (i - self_per); padding workgroups of the last shares leave
  const int xcd = (int)(blockIdx.x & 7), i = (int)(blockIdx.x >> 3);
  const bool self = i < self_per;
  int sb;
  if (self) {
    sb = xcd * self_per + i;
    if (sb >= n_poly) return;
  } else {
    const int j = xcd * pair_per + (i - self_per);
    if (j >= n_pair) return;
    sb = n_poly + j;
  }
  int4 q;
  if (self)
    q = make_int4(ssrc[sb], sb, sb, -1);
  else
    q = pinfo[sb - n_poly];
  sym_block<P>(self, Ks + (size_t)sb * D::BS + 2 * lane, q.x, q.y, q.z, q.w, u, prod, sT, lane);
}

// k_sym_products at small degree (3Np <= 32: p = 1, 2): one wave carries
// G = 64 / 3Np stored blocks, lane = (block, row), so short rows do not leave
// lanes idle.  Per block the sums of sym_block in its order: the row product
// K_PQ u_Q (FMA-free chain over ascending columns, u_Q broadcast from LDS),
// then the block's products K_PQ[i][c] u_P[i] in LDS at once (a self block's
// diagonal zeroed) and lane (block, c) adding column c over ascending rows;
// so the two kernels agree bit for bit.  Stores cover the rows only (the
// update reads no padding).
template <int P>
__global__ __launch_bounds__(64) void k_sym_products_packed(const double* __restrict__ Ks,
                                                            const int* __restrict__ ssrc,
                                                            const int4* __restrict__ pinfo,
                                                            const double* __restrict__ u,
                                                            double* __restrict__ prod, int n_poly,
                                                            int n_stored) {
  using D = Dims<P>;
  constexpr int N3 = D::N3, NPAIR = D::NPAIR, PS = prod_stride(N3);
  static_assert(N3 <= 32 && N3 % 2 == 0, "packed stored blocks need 3Np <= 32, whole column pairs");
  static_assert(self_upper_off(NPAIR) <= D::BS, "packed self block fits its slot");
  constexpr int G = 64 / N3;
  __shared__ double sX[G][N3];          // u_Q of each block
  __shared__ double sT[G][N3][N3 + 1];  // K_PQ[i][c] u_P[i]; odd row stride
  const int lane = threadIdx.x, sub = lane / N3, r = lane % N3;
  // XCD-aware: workgroups are dealt to the XCDs round robin; XCD x takes a
  // contiguous eighth of the waves, hence of the stored blocks
  const int per = (int)(gridDim.x >> 3);
  const int w = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  const int sb = w * G + sub;
  const bool on = sub < G && sb < n_stored;
  const bool self = sb < n_poly;
  int b = 0, Pp = 0, Qq = 0, td = 0;
  if (on) {
    if (self) {
      b = ssrc[sb];
      Pp = Qq = sb;
    } else {
      const int4 q = pinfo[sb - n_poly];
      b = q.x;
      Pp = q.y;
      Qq = q.z;
      td = q.w;
    }
  }
  const double* __restrict__ Kb = Ks + (size_t)(on ? sb : 0) * D::BS;
  double2 kv[NPAIR];
#pragma unroll
  for (int c2i = 0; c2i < NPAIR; ++c2i) {
    const bool ld = on && (!self || r <= 2 * c2i + 1);
    kv[c2i] = ld ? load_stream(Kb + (self ? self_upper_off(c2i) + 2 * r : ((size_t)c2i * N3 + r) * 2))
                 : make_double2(0.0, 0.0);
    if (self && r == 2 * c2i + 1) kv[c2i].x = 0.0;  // K_PP[r][r - 1]: below the diagonal
  }
  const double xl = on ? u[(size_t)Qq * N3 + r] : 0.0;
  const double up = on ? u[(size_t)Pp * N3 + r] : 0.0;
  if (on) sX[sub][r] = xl;
  __syncthreads();
  double s = 0.0;
  if (on) {
#pragma unroll
    for (int c2i = 0; c2i < NPAIR; ++c2i) {
      s = s + kv[c2i].x * sX[sub][2 * c2i];
      s = s + kv[c2i].y * sX[sub][2 * c2i + 1];
    }
  }
  if (on) {
#pragma unroll
    for (int c2i = 0; c2i < NPAIR; ++c2i) {
      sT[sub][r][2 * c2i] = kv[c2i].x * up;
      sT[sub][r][2 * c2i + 1] = kv[c2i].y * up;
    }
    if (self) sT[sub][r][r] = 0.0;  // the diagonal is in the row product only
  }
  __syncthreads();
  double st = 0.0;
  if (on) {
    double v = 0.0;  // column r of this block over ascending rows
#pragma unroll
    for (int i = 0; i < N3; ++i) v = v + sT[sub][i][r];
    st = v;
  }
  if (on) {
    if (self) {
      prod[(size_t)b * PS + r] = s + st;
    } else {
      prod[(size_t)b * PS + r] = s;
      prod[(size_t)td * PS + r] = st;
    }
  }
}

// K_PP's upper triangle in the packed layout of k_sym_products, one thread
// per stored entry: the self blocks listed in self_s (stored index s, taken
// from full-row slot src[s]).
template <int P>
__global__ void k_pack_self_upper(const double* __restrict__ K, double* __restrict__ Ks,
                                  const int* __restrict__ src, const int* __restrict__ self_s,
                                  int n_self) {
  using D = Dims<P>;
  constexpr int N3 = D::N3, NPAIR = D::NPAIR, NE = self_upper_off(NPAIR);
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)n_self * NE) return;
  const int k = (int)(id / NE), e = (int)(id % NE);
  int c2i = 0;  // column pair of packed entry e: self_upper_off(c2i) <= e < self_upper_off(c2i + 1)
  while (self_upper_off(c2i + 1) <= e) ++c2i;
  const int r = (e - self_upper_off(c2i)) >> 1, h = e & 1;
  const int s = self_s[k];
  Ks[(size_t)s * D::BS + e] = K[(size_t)src[s] * D::BS + ((size_t)c2i * N3 + r) * 2 + h];
}

template <int P, int MODE>
__global__ __launch_bounds__(64) void k_sym_update(const double* __restrict__ prod,
                                                   const int* __restrict__ row_ptr,
                                                   const double* __restrict__ Sp,
                                                   const double* __restrict__ rs,
                                                   const double* __restrict__ rd,
                                                   const double* __restrict__ u,
                                                   const double* __restrict__ wa,
                                                   double* __restrict__ out, double g, double c1,
                                                   double c2, const int* __restrict__ gsel,
                                                   const double* __restrict__ gtab, int n_poly,
                                                   const int* __restrict__ order) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3;
  static_assert(N3 <= 64, "one lane per row");
  static_assert(MODE == 0 || MODE == 1, "leapfrog and first step");
  __shared__ double sS[D::NPK];
  __shared__ double sv[N3];
  const int lane = threadIdx.x;
  const bool row = lane < N3;
  const int Pp = step_poly(order);
  if (Pp >= n_poly) return;
  for (int k = lane; k < D::NPK; k += 64) sS[k] = Sp[(size_t)Pp * D::NPK + k];
  const int b0 = row_ptr[Pp], b1 = row_ptr[Pp + 1];
  const size_t o = (size_t)Pp * N3;
  double vs = 0.0, vd = 0.0, vu = 0.0, va = 0.0;
  if (row) {
    vs = rs[o + lane];
    vd = rd[o + lane];
    vu = u[o + lane];
    va = wa[o + lane];
  }
  // the row's products, eight loads in flight at a time, added in block order
  double acc = 0.0;
  if (row) {
    for (int b = b0; b < b1; b += 8) {
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        v[k] = b + k < b1 ? prod[(size_t)(b + k) * prod_stride(N3) + lane] : 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (b + k < b1) acc = (b + k == b0) ? v[k] : acc + v[k];
    }
  }
  const double gp = gsel ? gtab[gsel[Pp]] : g;
  const int cc = row ? lane / NP : 0, ii = lane % NP;
  const double* __restrict__ vc = sv + cc * NP;  // this lane's component
  if (row) sv[lane] = -acc + vs + vd * gp;
  __syncthreads();  // sS, sv
  // z = S^T w: z_i = sum_{j <= i} S_ji w_j
  double z = 0.0;
  if (row) {
#pragma unroll
    for (int j = 0; j < NP; ++j)
      if (j <= ii) z = z + sS[pk_off(NP, j) + (ii - j)] * vc[j];
  }
  __syncthreads();
  if (row) sv[lane] = z;
  __syncthreads();
  // M^-1 w = S z: x_i = sum_{j >= i} S_ij z_j
  double w = 0.0;
  if (row) {
#pragma unroll
    for (int j = 0; j < NP; ++j)
      if (j >= ii) w = w + sS[pk_off(NP, ii) + (j - ii)] * vc[j];
    if (MODE == 0)
      out[o + lane] = c1 * w + 2.0 * vu - va;
    else
      out[o + lane] = c1 * w + vu + c2 * va;
  }
}

// k_sym_update at small degree (3Np <= 32: p = 1, 2): G = 64 / 3Np
// polyhedra per wave, lane = (polyhedron, row), each with its own S and
// vector in LDS; the sums of k_sym_update in its order, so the two agree bit
// for bit.
template <int P, int MODE>
__global__ __launch_bounds__(64) void k_sym_update_packed(const double* __restrict__ prod,
                                                          const int* __restrict__ row_ptr,
                                                          const double* __restrict__ Sp,
                                                          const double* __restrict__ rs,
                                                          const double* __restrict__ rd,
                                                          const double* __restrict__ u,
                                                          const double* __restrict__ wa,
                                                          double* __restrict__ out, double g,
                                                          double c1, double c2,
                                                          const int* __restrict__ gsel,
                                                          const double* __restrict__ gtab, int n_poly) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3;
  static_assert(N3 <= 32, "packed update needs 3Np <= 32");
  static_assert(MODE == 0 || MODE == 1, "leapfrog and first step");
  constexpr int G = 64 / N3;
  __shared__ double sS[G][D::NPK];
  __shared__ double sv[G][N3];
  const int lane = threadIdx.x, sub = lane / N3, r = lane % N3;
  const int P0 = step_poly() * G;
  const int Pp = P0 + sub;
  const bool act = sub < G && Pp < n_poly;
  for (int k = lane; k < G * D::NPK; k += 64) {
    const int sp = k / D::NPK, e = k % D::NPK;
    if (P0 + sp < n_poly) sS[sp][e] = Sp[(size_t)(P0 + sp) * D::NPK + e];
  }
  const size_t o = (size_t)Pp * N3 + r;
  double vs = 0.0, vd = 0.0, vu = 0.0, va = 0.0, acc = 0.0;
  if (act) {
    vs = rs[o];
    vd = rd[o];
    vu = u[o];
    va = wa[o];
    const int b0 = row_ptr[Pp], b1 = row_ptr[Pp + 1];
    for (int b = b0; b < b1; b += 8) {  // the row's products, added in block order
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = b + k < b1 ? prod[(size_t)(b + k) * prod_stride(N3) + r] : 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (b + k < b1) acc = (b + k == b0) ? v[k] : acc + v[k];
    }
  }
  const int cc = r / NP, ii = r % NP;
  const double* __restrict__ vc = &sv[act ? sub : 0][cc * NP];  // this lane's component
  if (act) {
    const double gp = gsel ? gtab[gsel[Pp]] : g;
    sv[sub][r] = -acc + vs + vd * gp;
  }
  __syncthreads();  // sS, sv
  double z = 0.0;  // z = S^T w: z_i = sum_{j <= i} S_ji w_j
  if (act) {
#pragma unroll
    for (int j = 0; j < NP; ++j)
      if (j <= ii) z = z + sS[sub][pk_off(NP, j) + (ii - j)] * vc[j];
  }
  __syncthreads();
  if (act) sv[sub][r] = z;
  __syncthreads();
  if (act) {  // M^-1 w = S z: x_i = sum_{j >= i} S_ij z_j
    double w = 0.0;
#pragma unroll
    for (int j = 0; j < NP; ++j)
      if (j >= ii) w = w + sS[sub][pk_off(NP, ii) + (j - ii)] * vc[j];
    if (MODE == 0)
      out[o] = c1 * w + 2.0 * vu - va;
    else
      out[o] = c1 * w + vu + c2 * va;
  }
}

// The pairs-once steps of the 32- and 48-bit applies (mass-orthonormal
// basis, no solves): ~K_QP = R_Q^-T K_PQ^T R_P^-1 = ~K_PQ^T, so the stored
// list of the FP64 step applies to the FP32 blocks (column quads, 3Np a
// multiple of 12 so that each of three chunks is whole quads: p = 3).  A
// self block keeps its upper triangle by column quads: quad q's rows
// 0..4q+3 from float self32_off(q), row r at self32_off(q) + 4r (and the
// 48-bit corrections by quad pairs: pair k's rows 0..min(8k+7, 3Np-1) from
// half self16_off(k), row r at self16_off(k) + self16_w(k) r, 8 halves per
// row, 4 for the last quad of an odd count); entries below the diagonal that
// a quad's rows carry are zeroed on load.
//
// One wave per polyhedron walks its row's stored blocks.  Lane i holds row i
// of the block; the block's entries go to LDS in three chunks of CW columns,
// and lane c runs column c's chain over ascending rows r with u_P[r] -- the
// chain Q's own full row runs over block (Q, P), so where the two rounded
// blocks are exact transposes the transposed product equals the full rows'
// bit for bit.  Then each lane's row chain over ascending columns: from zero
// for a pair block, and for a self block continuing its own column chain
// (the strict upper triangle's column r holds row r's entries left of the
// diagonal, in the full row's order), so a self block's product is also the
// full rows' product bit for bit.  (Summing the transposed products in three
// row groups instead made the 32-bit state drift from FP64 faster over long
// runs, increasingly with the mesh size: 7x the full rows' drift after 1000
// steps at 40^3.)
//
// A wave adds the row products of its stored blocks in FP64 (block order)
// and stores that sum once, rowp[P]; only a pair block's transposed product
// is stored per block, in the slot of block (Q, P) of Q's row.
// k_sym_update32 adds to rowp[P] the transposed products of the blocks kept
// in the neighbours' rows (sidx < 0), in block order, in FP64, and applies
// k_leapfrog32's update.
//
// 32-bit: products and per-block sums in FP32 as k_leapfrog32, transposed
// products stored as floats (prod_stride(3Np) per block, half the bytes of a
// 64-bit product; PT = double: the 48-bit apply's, k_sym_products48).
LYMPH_HD constexpr int self32_off(int q) { return 8 * q * (q + 1); }
LYMPH_HD constexpr int self16_rows(int n3, int k) { return 8 * k + 8 < n3 ? 8 * k + 8 : n3; }
LYMPH_HD constexpr int self16_w(int n3, int k) { return 2 * k + 1 < n3 / 4 ? 8 : 4; }
LYMPH_HD constexpr int self16_off(int n3, int k) {
  int o = 0;
  for (int i = 0; i < k; ++i) o += self16_w(n3, i) * self16_rows(n3, i);
  return o;
}

// A self block's upper triangle in the packed layouts above, from the
// full-row block of slot src[s] (s < n_self): one thread per packed entry.
template <int P>
__global__ void k_pack_self32(const float* __restrict__ K, float* __restrict__ Ks,
                              const int* __restrict__ src, int n_self) {
  constexpr int N3 = Dims32<P>::N3, NE = self32_off(Dims32<P>::NQ);
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)n_self * NE) return;
  const int s = (int)(id / NE), e = (int)(id % NE);
  int q = 0;
  while (self32_off(q + 1) <= e) ++q;
  const int r = (e - self32_off(q)) >> 2, c = 4 * q + (e & 3);
  Ks[(size_t)s * Dims32<P>::BS + e] = K[(size_t)src[s] * Dims32<P>::BS + kpos32(N3, r, c)];
}
template <int P>
__global__ void k_pack_self16(const _Float16* __restrict__ K, _Float16* __restrict__ Ks,
                              const int* __restrict__ src, int n_self) {
  constexpr int N3 = Dims32<P>::N3, NE = self16_off(N3, Dims16<P>::NPR);
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)n_self * NE) return;
  const int s = (int)(id / NE), e = (int)(id % NE);
  int k = 0;
  while (self16_off(N3, k + 1) <= e) ++k;
  const int w = self16_w(N3, k), r = (e - self16_off(N3, k)) / w, h = (e - self16_off(N3, k)) % w;
  const int c = 8 * k + h;  // quad 2k + h / 4, column h % 4 of it
  Ks[(size_t)s * Dims16<P>::BS + e] = K[(size_t)src[s] * Dims16<P>::BS + kpos16(N3, r, c)];
}

template <int P>
__global__ __launch_bounds__(64) void k_sym_products32(const float* __restrict__ Ks,
                                                       const int* __restrict__ row_ptr,
                                                       const int* __restrict__ col,
                                                       const int* __restrict__ sidx,
                                                       const int* __restrict__ tdst,
                                                       const double* __restrict__ u,
                                                       float* __restrict__ prod,
                                                       double* __restrict__ rowp, int n_poly,
                                                       const int* __restrict__ order) {
  constexpr int N3 = Dims32<P>::N3, NQ = Dims32<P>::NQ;
  static_assert(N3 <= 64 && N3 % 12 == 0, "one lane per row, whole quads per chunk");
  static_assert(self32_off(NQ) <= (int)Dims32<P>::BS, "packed self block fits its slot");
  constexpr int CW = N3 / 3, CQ = CW / 4;
  // one chunk row by row (lanes past the rows write padding rows): a lane
  // stores its quads with 16-byte writes, and a column read over 20 lanes
  // touches 20 distinct banks
  __shared__ __align__(16) float sT[64 * CW];
  __shared__ __align__(16) float sX[64];  // u_Q of the current block, read by the row chains
  const int lane = threadIdx.x;
  const bool row = lane < N3;
  const int rl = row ? lane : N3 - 1;  // lanes past the rows load the last row's (cached) lines
  const int Pp = step_poly(order);
  if (Pp >= n_poly) return;  // padding workgroups of the last XCD share
  const int b0 = row_ptr[Pp], b1 = row_ptr[Pp + 1];
  const int lcb = lane < b1 - b0 ? col[b0 + lane] : 0;
  const int lsi = lane < b1 - b0 ? sidx[b0 + lane] : 0;
  const int ltd = lane < b1 - b0 ? tdst[b0 + lane] : 0;
  const float up = (float)u[(size_t)Pp * N3 + rl];
  double racc = 0.0;  // the row products of this wave's stored blocks, in block order
  for (int b = b0; b < b1; ++b) {
    const int j = b - b0;
    const int si = j < 64 ? __builtin_amdgcn_readlane(lsi, j) : __builtin_amdgcn_readfirstlane(sidx[b]);
    if (si < 0) continue;  // written by the neighbour's wave
    const int cb = j < 64 ? __builtin_amdgcn_readlane(lcb, j) : __builtin_amdgcn_readfirstlane(col[b]);
    const int td = j < 64 ? __builtin_amdgcn_readlane(ltd, j) : __builtin_amdgcn_readfirstlane(tdst[b]);
    const bool pair = cb != Pp;
    const float* __restrict__ Kb = Ks + (size_t)si * Dims32<P>::BS;
    const float xl = (float)u[(size_t)cb * N3 + rl];
    // the self blocks' row clamps and masks are formed per block: hoisted
    // out of the loop they would hold 60 masks in scalar registers
    int rb = rl;
    asm volatile("" : "+v"(rb));
    // one load path for both kinds (an address select per quad), so that the
    // chunks wait on their own quads only
    float4 kq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      kq[q] = load_stream(Kb + (pair ? (q * N3 + rb) * 4 : self32_off(q) + 4 * (rb < 4 * q + 3 ? rb : 4 * q + 3)));
    if (!pair) {  // the upper triangle's rows only
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        if (4 * q + 0 < rb) kq[q].x = 0.0f;
        if (4 * q + 1 < rb) kq[q].y = 0.0f;
        if (4 * q + 2 < rb) kq[q].z = 0.0f;
        if (4 * q + 3 < rb) kq[q].w = 0.0f;
      }
    }
    float st = 0.0f;
#pragma unroll
    for (int h = 0; h < 3; ++h) {
      wave_lds_sync();  // the previous chunk's (and block's) reads are done
      if (h == 0) sX[lane] = xl;  // lanes past the rows write padding
#pragma unroll
      for (int k = 0; k < CQ; ++k) {
        const float4 v = kq[h * CQ + k];
        *reinterpret_cast<float4*>(sT + lane * CW + 4 * k) =
            make_float4(v.x * up, v.y * up, v.z * up, v.w * up);
      }
      // a self block's diagonal belongs to its rows only (a wave's LDS
      // writes land in program order)
      if (!pair && lane >= h * CW && lane < (h + 1) * CW) sT[lane * CW + (lane - h * CW)] = 0.0f;
      wave_lds_sync();
      if (lane >= h * CW && lane < (h + 1) * CW) {  // column's chain over ascending rows
        const float* __restrict__ colp = sT + (lane - h * CW);
        float v = 0.0f;
#pragma unroll
        for (int r = 0; r < N3; ++r) v = v + colp[r * CW];
        st = v;
      }
    }
    float s = pair ? 0.0f : st;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const float4 xq = *reinterpret_cast<const float4*>(sX + 4 * q);  // a broadcast
      s = s + kq[q].x * xq.x;
      s = s + kq[q].y * xq.y;
      s = s + kq[q].z * xq.z;
      s = s + kq[q].w * xq.w;
    }
    // the block's row chain ends here: left free, the compiler interleaves it
    // with the next block's loads and holds both in registers
    asm volatile("" : "+v"(s));
    racc = racc + (double)s;
    if (pair && lane < prod_stride(N3)) prod[(size_t)td * prod_stride(N3) + lane] = row ? st : 0.0f;
  }
  if (lane < prod_stride(N3)) rowp[(size_t)Pp * prod_stride(N3) + lane] = row ? racc : 0.0;
}

template <int P, int MODE, typename PT = float>
__global__ __launch_bounds__(64) void k_sym_update32(const PT* __restrict__ prod,
                                                     const double* __restrict__ rowp,
                                                     const int* __restrict__ sidx,
                                                     const int* __restrict__ row_ptr,
                                                     const double* __restrict__ rs,
                                                     const double* __restrict__ rd,
                                                     const double* __restrict__ u,
                                                     const double* __restrict__ wa,
                                                     double* __restrict__ out, double g, double c1,
                                                     double c2, const int* __restrict__ gsel,
                                                     const double* __restrict__ gtab, int n_poly,
                                                     const int* __restrict__ order) {
  constexpr int N3 = Dims32<P>::N3;
  static_assert(N3 <= 64, "one lane per row");
  static_assert(MODE == 0 || MODE == 1, "leapfrog and first step");
  const int lane = threadIdx.x;
  if (lane >= N3) return;  // no barrier in this kernel
  const int Pp = step_poly(order);
  if (Pp >= n_poly) return;
  const int b0 = row_ptr[Pp], b1 = row_ptr[Pp + 1];
  const size_t o = (size_t)Pp * N3;
  const double vs = rs[o + lane], vd = rd[o + lane], vu = u[o + lane], va = wa[o + lane];
  // the wave's own row products (summed by the products pass), then the
  // neighbours' transposed products in block order
  double acc = rowp[(size_t)Pp * prod_stride(N3) + lane];
  for (int b = b0; b < b1; b += 8) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const bool theirs = b + k < b1 && sidx[b + k] < 0;
      v[k] = theirs ? (double)prod[(size_t)(b + k) * prod_stride(N3) + lane] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (b + k < b1 && sidx[b + k] < 0) acc = acc + v[k];
  }
  const double gp = gsel ? gtab[gsel[Pp]] : g;
  const double w = -acc + vs + vd * gp;
  if (MODE == 0)
    out[o + lane] = c1 * w + 2.0 * vu - va;
  else
    out[o + lane] = c1 * w + vu + c2 * va;
}

// Symmetry of the operator, pair by pair: ratio[w] = max |K_PQ - K_QP^T| /
// max |K_PQ| for the blocks pb[w] = K_PQ and pt[w] = K_QP (column-pair
// layout, whole pairs: 3Np even).
template <int P>
__global__ __launch_bounds__(64) void k_pair_asym(const double* __restrict__ K,
                                                  const int* __restrict__ pb,
                                                  const int* __restrict__ pt, int n_pairs,
                                                  double* __restrict__ ratio) {
  using D = Dims<P>;
  constexpr int N3 = D::N3;
  static_assert(N3 % 2 == 0, "whole column pairs");
  const int w = blockIdx.x;
  if (w >= n_pairs) return;
  const double* A = K + (size_t)pb[w] * D::BS;
  const double* B = K + (size_t)pt[w] * D::BS;
  double dmax = 0.0, amax = 0.0;
  for (int e = threadIdx.x; e < N3 * N3; e += 64) {
    const int r = e % N3, c = e / N3;
    const double a = A[((size_t)(c >> 1) * N3 + r) * 2 + (c & 1)];   // K_PQ[r][c]
    const double bt = B[((size_t)(r >> 1) * N3 + c) * 2 + (r & 1)];  // K_QP[c][r]
    const double d = fabs(a - bt);
    dmax = (d > dmax || d != d) ? d : dmax;  // a NaN makes the pair fail
    amax = fabs(a) > amax ? fabs(a) : amax;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double od = __shfl_xor(dmax, off, 64), oa = __shfl_xor(amax, off, 64);
    dmax = (od > dmax || od != od) ? od : dmax;
    amax = oa > amax ? oa : amax;
  }
  if (threadIdx.x == 0) ratio[w] = dmax == 0.0 ? 0.0 : dmax / amax;
}

// Ks[s] = K[src[s]], whole blocks of bs entries.
template <class T>
__global__ void k_gather_blocks(const T* __restrict__ K, T* __restrict__ Ks,
                                const int* __restrict__ src, int n, size_t bs) {
  const size_t total = (size_t)n * bs;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t s = i / bs, k = i - s * bs;
    Ks[i] = K[(size_t)src[s] * bs + k];
  }
}

// Barrier of the packed kernels: a whole workgroup for one-wave workgroups;
// with several waves per workgroup each wave owns its own LDS region and
// polyhedra (and its own loop trip count), so only the wave's own LDS
// traffic is ordered (the rocPRIM wave barrier: fences at wavefront scope
// around a code-motion barrier; a wave's LDS operations complete in order).
template <int WPW>
__device__ inline void lsync() {
  if constexpr (WPW == 1) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Packed variant of the step for small blocks (3Np <= 32: p = 1, 2): one
// wave carries G = 64 / 3Np polyhedra, lane = (polyhedron, row), so lanes are
// not left idle by short rows.  Each polyhedron walks its own block row; the
// neighbour's coefficients go through LDS (one segment per polyhedron); the
// block products are added in block order and the solves and update are the
// ones of k_leapfrog, so results are bit-identical to it.  TB = 64 (FP64
// operator in column pairs), 32 (mass-orthonormal FP32 operator in column
// quads, no solves) or 48 (the FP32 operator plus its 16-bit corrections Kr,
// with k_leapfrog48's arithmetic in its order: bit-identical to it).
template <int P, int MODE, int TB, int WPW = 1>
__global__ __launch_bounds__(64 * WPW) void k_leapfrog_packed(const void* __restrict__ Kv,
                                                        const _Float16* __restrict__ Kr,
                                                        const int* __restrict__ row_ptr,
                                                        const int* __restrict__ col,
                                                        const double* __restrict__ Rp,
                                                        const double* __restrict__ rs,
                                                        const double* __restrict__ rd,
                                                        const double* __restrict__ u,
                                                        const double* __restrict__ wa,
                                                        double* __restrict__ out, double g,
                                                        double c1, double c2,
                                                        const int* __restrict__ gsel,
                                                        const double* __restrict__ gtab,
                                                        int n_poly) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3, NPAIR = D::NPAIR;
  static_assert(N3 <= 32, "packed step needs 3Np <= 32");
  constexpr int G = 64 / N3;
  constexpr int NQUAD = Dims32<P>::NQ, TAIL = Dims32<P>::T;
  // neighbour coefficients, double-buffered by block parity: one barrier per block
  __shared__ double sxa[WPW][2][G][N3 + 1];
  __shared__ double swa[WPW][G][N3];
  __shared__ double sRa[WPW][G][D::NPK];
  const int wv = threadIdx.x >> 6;
  double(&sx)[2][G][N3 + 1] = sxa[wv];
  double(&sw)[G][N3] = swa[wv];
  double(&sR)[G][D::NPK] = sRa[wv];
  const int lane = threadIdx.x & 63;
  const int sub = lane / N3, r = lane % N3;
  const int P0 = (step_poly() * WPW + wv) * G;
  const int Pp = P0 + sub;
  const bool act = sub < G && Pp < n_poly;
  if (MODE != 2 && TB == 64)
    for (int k = lane; k < G * D::NPK; k += 64) {
      const int sp = k / D::NPK, e = k % D::NPK, Q = P0 + sp;
      if (Q < n_poly) sR[sp][e] = Rp[(size_t)Q * D::NPK + e];
    }
  const int b0 = act ? row_ptr[Pp] : 0, cnt = act ? row_ptr[Pp + 1] - b0 : 0;
  double xn = cnt > 0 ? u[(size_t)col[b0] * N3 + r] : 0.0;  // neighbour value of the next block
  double acc = 0.0;
  for (int j = 0; __any(j < cnt); ++j) {
    const bool on = j < cnt;
    const int b = b0 + j;
    const int buf = j & 1;
    // the block's loads and the next neighbour value are issued before the
    // barrier, so they overlap the exchange of this block's coefficients
    double2 kd[TB == 64 ? NPAIR : 1];
    float4 kf[TB != 64 ? NQUAD : 1];
    float kt[TB != 64 && TAIL > 0 ? TAIL : 1];
    lymph_hv8 kr[TB == 48 ? (NQUAD + 1) / 2 : 1];
    float krt[TB == 48 && TAIL > 0 ? TAIL : 1];
    if (on) {
      if constexpr (TB == 64) {
        const double* Kb = static_cast<const double*>(Kv) + (size_t)b * D::BS;
#pragma unroll
        for (int c2i = 0; c2i < NPAIR; ++c2i) kd[c2i] = load_stream(Kb + ((size_t)c2i * N3 + r) * 2);
      } else {
        const float* Kb = static_cast<const float*>(Kv) + (size_t)b * Dims32<P>::BS;
#pragma unroll
        for (int q = 0; q < NQUAD; ++q) kf[q] = load_stream(Kb + ((size_t)q * N3 + r) * 4);
        if constexpr (TAIL > 0) load_tail<TAIL>(Kb + (size_t)4 * NQUAD * N3 + (size_t)r * TAIL, kt);
        if constexpr (TB == 48) {
          const _Float16* Rb = Kr + (size_t)b * Dims16<P>::BS;
#pragma unroll
          for (int k = 0; k < (NQUAD + 1) / 2; ++k) {
            if (2 * k + 1 < NQUAD) {
              kr[k] = load_stream_h8(Rb + ((size_t)k * N3 + r) * 8);
            } else {  // the last quad of an odd count: 4 values per row
              const lymph_hv4 h = load_stream_h4(Rb + Dims16<P>::HALF0 + (size_t)r * 4);
              kr[k] = lymph_hv8{h.x, h.y, h.z, h.w, 0, 0, 0, 0};
            }
          }
          if constexpr (TAIL > 0) {
#pragma unroll
            for (int j = 0; j < TAIL; ++j) krt[j] = (float)Rb[Dims16<P>::TAIL0 + (size_t)r * TAIL + j];
          }
        }
      }
      sx[buf][sub][r] = xn;
      if (j + 1 < cnt) xn = u[(size_t)col[b + 1] * N3 + r];
    }
    lsync<WPW>();
    if (on) {
      double s = 0.0;
      if constexpr (TB == 64) {
#pragma unroll
        for (int c2i = 0; c2i < NPAIR; ++c2i) {
          const bool odd_tail = (N3 % 2 == 1) && (c2i == NPAIR - 1);
          s = s + kd[c2i].x * sx[buf][sub][2 * c2i];
          if (!odd_tail) s = s + kd[c2i].y * sx[buf][sub][2 * c2i + 1];
        }
      } else if constexpr (TB == 48) {  // k_leapfrog48's chains
        double sd = 0.0;
        float sc = 0.0f;
#pragma unroll
        for (int q = 0; q < NQUAD; ++q) {
          double xq[4];
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) xq[jj] = (4 * q + jj < N3) ? sx[buf][sub][4 * q + jj] : 0.0;
          const lymph_hv8 t8 = kr[q >> 1];
          const int o4 = (q & 1) * 4;
          sd = fma((double)kf[q].x, xq[0], sd);
          sd = fma((double)kf[q].y, xq[1], sd);
          sd = fma((double)kf[q].z, xq[2], sd);
          sd = fma((double)kf[q].w, xq[3], sd);
          sc = fmaf(kf[q].x * (float)t8[o4 + 0], (float)xq[0], sc);
          sc = fmaf(kf[q].y * (float)t8[o4 + 1], (float)xq[1], sc);
          sc = fmaf(kf[q].z * (float)t8[o4 + 2], (float)xq[2], sc);
          sc = fmaf(kf[q].w * (float)t8[o4 + 3], (float)xq[3], sc);
        }
        if constexpr (TAIL > 0) {
#pragma unroll
          for (int j = 0; j < TAIL; ++j) {
            const double xt = sx[buf][sub][4 * NQUAD + j];
            sd = fma((double)kt[j], xt, sd);
            sc = fmaf(kt[j] * krt[j], (float)xt, sc);
          }
        }
        s = sd + (double)sc * 0x1p-24;
      } else {
        float sf = 0.0f;
#pragma unroll
        for (int q = 0; q < NQUAD; ++q) {
          float xq[4];
#pragma unroll
          for (int jj = 0; jj < 4; ++jj)
            xq[jj] = (4 * q + jj < N3) ? (float)sx[buf][sub][4 * q + jj] : 0.0f;
          sf = sf + kf[q].x * xq[0];
          sf = sf + kf[q].y * xq[1];
          sf = sf + kf[q].z * xq[2];
          sf = sf + kf[q].w * xq[3];
        }
        if constexpr (TAIL > 0) {
#pragma unroll
          for (int j = 0; j < TAIL; ++j) sf = sf + kt[j] * (float)sx[buf][sub][4 * NQUAD + j];
        }
        s = (double)sf;
      }
      acc = (j == 0) ? s : acc + s;
    }
  }
  const size_t o = (size_t)Pp * N3 + r;
  if (MODE == 2) {
    if (act) sw[sub][r] = acc;
    lsync<WPW>();
    if (act && r == 0) {
      double el = 0.0;
      for (int k = 0; k < N3; ++k) el = el + u[(size_t)Pp * N3 + k] * sw[sub][k];
      out[Pp] = el;
    }
    return;
  }
  double w = 0.0;
  if (act) {
    const double gp = gsel ? gtab[gsel[Pp]] : g;
    w = -acc + rs[o] + rd[o] * gp;
  }
  if constexpr (TB == 64) {
    if (act) sw[sub][r] = w;
    lsync<WPW>();
    if (lane < 3 * G) {
      const int sp = lane / 3, comp = lane % 3;
      if (P0 + sp < n_poly) solve_packed<NP>(sR[sp], sw[sp] + comp * NP);
    }
    lsync<WPW>();
    if (act) w = sw[sub][r];
  }
  if (act) {
    if (MODE == 0)
      out[o] = c1 * w + 2.0 * u[o] - wa[o];
    else
      out[o] = c1 * w + u[o] + c2 * wa[o];
  }
}

// Packed step with a joint block walk (3Np <= 32): the G polyhedra of a wave
// own the contiguous block range row_ptr[P0] .. row_ptr[P0 + G]; at loop
// iteration j lane group s takes block B0 + j G + s, so each iteration
// streams G consecutive blocks and no cache line is left half read for a
// later iteration (32-bit blocks of 576 B at p = 1 end mid-line).  Each
// block's product goes to LDS; each row then adds its own blocks in block
// order, so results are bit-identical to k_leapfrog (TB = 48: to
// k_leapfrog48, with Kr its 16-bit corrections).  Dynamic LDS: the products
// of the wave's longest block range (G x longest row x 3Np doubles).
template <int P, int MODE, int TB, int WPW = 1>
__global__ __launch_bounds__(64 * WPW) void k_leapfrog_joint(const void* __restrict__ Kv,
                                                       const _Float16* __restrict__ Kr,
                                                       const int* __restrict__ row_ptr,
                                                       const int* __restrict__ col,
                                                       const double* __restrict__ Rp,
                                                       const double* __restrict__ rs,
                                                       const double* __restrict__ rd,
                                                       const double* __restrict__ u,
                                                       const double* __restrict__ wa,
                                                       double* __restrict__ out, double g,
                                                       double c1, double c2,
                                                       const int* __restrict__ gsel,
                                                       const double* __restrict__ gtab,
                                                       int n_poly, int prod_stride) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3, NPAIR = D::NPAIR;
  static_assert(N3 <= 32, "joint packed step needs 3Np <= 32");
  constexpr int G = 64 / N3;
  constexpr int NQUAD = Dims32<P>::NQ, TAIL = Dims32<P>::T;
  __shared__ double sxa[WPW][2][G][N3 + 1];
  __shared__ double swa[WPW][G][N3];
  __shared__ double sRa[WPW][G][D::NPK];
  const int wv = threadIdx.x >> 6;
  double(&sx)[2][G][N3 + 1] = sxa[wv];
  double(&sw)[G][N3] = swa[wv];
  double(&sR)[G][D::NPK] = sRa[wv];
  extern __shared__ double prod_all[];  // per wave: [block - B0][row]
  const int lane = threadIdx.x & 63;
  const int sub = lane / N3, r = lane % N3;
  const int P0 = (step_poly() * WPW + wv) * G;
  const int Pp = P0 + sub;
  const bool act = sub < G && Pp < n_poly;
  if (MODE != 2 && TB == 64)
    for (int k = lane; k < G * D::NPK; k += 64) {
      const int sp = k / D::NPK, e = k % D::NPK, Q = P0 + sp;
      if (Q < n_poly) sR[sp][e] = Rp[(size_t)Q * D::NPK + e];
    }
  if (P0 >= n_poly) return;  // padding group: the wave exits before any barrier
  double* prod = prod_all + (size_t)wv * prod_stride;
  const int B0 = row_ptr[P0], B1 = row_ptr[min(P0 + G, n_poly)];
  const int nblk = B1 - B0;
  int bj = B0 + sub;  // this lane group's block at iteration 0
  double xn = (sub < G && bj < B1) ? u[(size_t)col[bj] * N3 + r] : 0.0;
  for (int j = 0; j * G < nblk; ++j, bj += G) {
    const bool on = sub < G && bj < B1;
    const int buf = j & 1;
    double2 kd[TB == 64 ? NPAIR : 1];
    float4 kf[TB != 64 ? NQUAD : 1];
    float kt[TB != 64 && TAIL > 0 ? TAIL : 1];
    lymph_hv8 kr[TB == 48 ? (NQUAD + 1) / 2 : 1];
    float krt[TB == 48 && TAIL > 0 ? TAIL : 1];
    if (on) {
      if constexpr (TB == 64) {
        const double* Kb = static_cast<const double*>(Kv) + (size_t)bj * D::BS;
#pragma unroll
        for (int c2i = 0; c2i < NPAIR; ++c2i) kd[c2i] = load_stream(Kb + ((size_t)c2i * N3 + r) * 2);
      } else {
        const float* Kb = static_cast<const float*>(Kv) + (size_t)bj * Dims32<P>::BS;
#pragma unroll
        for (int q = 0; q < NQUAD; ++q) kf[q] = load_stream(Kb + ((size_t)q * N3 + r) * 4);
        if constexpr (TAIL > 0) load_tail<TAIL>(Kb + (size_t)4 * NQUAD * N3 + (size_t)r * TAIL, kt);
        if constexpr (TB == 48) {
          const _Float16* Rb = Kr + (size_t)bj * Dims16<P>::BS;
#pragma unroll
          for (int kk = 0; kk < (NQUAD + 1) / 2; ++kk) {
            if (2 * kk + 1 < NQUAD) {
              kr[kk] = load_stream_h8(Rb + ((size_t)kk * N3 + r) * 8);
            } else {  // the last quad of an odd count: 4 values per row
              const lymph_hv4 h = load_stream_h4(Rb + Dims16<P>::HALF0 + (size_t)r * 4);
              kr[kk] = lymph_hv8{h.x, h.y, h.z, h.w, 0, 0, 0, 0};
            }
          }
          if constexpr (TAIL > 0) {
#pragma unroll
            for (int jj = 0; jj < TAIL; ++jj) krt[jj] = (float)Rb[Dims16<P>::TAIL0 + (size_t)r * TAIL + jj];
          }
        }
      }
      sx[buf][sub][r] = xn;
      if (bj + G < B1) xn = u[(size_t)col[bj + G] * N3 + r];
    }
    lsync<WPW>();
    if (on) {
      double s = 0.0;
      if constexpr (TB == 64) {
#pragma unroll
        for (int c2i = 0; c2i < NPAIR; ++c2i) {
          const bool odd_tail = (N3 % 2 == 1) && (c2i == NPAIR - 1);
          s = s + kd[c2i].x * sx[buf][sub][2 * c2i];
          if (!odd_tail) s = s + kd[c2i].y * sx[buf][sub][2 * c2i + 1];
        }
      } else if constexpr (TB == 48) {  // k_leapfrog48's chains
        double sd = 0.0;
        float sc = 0.0f;
#pragma unroll
        for (int q = 0; q < NQUAD; ++q) {
          double xq[4];
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) xq[jj] = sx[buf][sub][4 * q + jj];
          const lymph_hv8 t8 = kr[q >> 1];
          const int o4 = (q & 1) * 4;
          sd = fma((double)kf[q].x, xq[0], sd);
          sd = fma((double)kf[q].y, xq[1], sd);
          sd = fma((double)kf[q].z, xq[2], sd);
          sd = fma((double)kf[q].w, xq[3], sd);
          sc = fmaf(kf[q].x * (float)t8[o4 + 0], (float)xq[0], sc);
          sc = fmaf(kf[q].y * (float)t8[o4 + 1], (float)xq[1], sc);
          sc = fmaf(kf[q].z * (float)t8[o4 + 2], (float)xq[2], sc);
          sc = fmaf(kf[q].w * (float)t8[o4 + 3], (float)xq[3], sc);
        }
        if constexpr (TAIL > 0) {
#pragma unroll
          for (int jj = 0; jj < TAIL; ++jj) {
            const double xt = sx[buf][sub][4 * NQUAD + jj];
            sd = fma((double)kt[jj], xt, sd);
            sc = fmaf(kt[jj] * krt[jj], (float)xt, sc);
          }
        }
        s = sd + (double)sc * 0x1p-24;
      } else {
        float sf = 0.0f;
#pragma unroll
        for (int q = 0; q < NQUAD; ++q) {
          float xq[4];
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) xq[jj] = (float)sx[buf][sub][4 * q + jj];
          sf = sf + kf[q].x * xq[0];
          sf = sf + kf[q].y * xq[1];
          sf = sf + kf[q].z * xq[2];
          sf = sf + kf[q].w * xq[3];
        }
        if constexpr (TAIL > 0) {
#pragma unroll
          for (int jj = 0; jj < TAIL; ++jj) sf = sf + kt[jj] * (float)sx[buf][sub][4 * NQUAD + jj];
        }
        s = (double)sf;
      }
      prod[(size_t)(bj - B0) * N3 + r] = s;
    }
  }
  lsync<WPW>();
  double acc = 0.0;
  if (act) {  // this row's blocks, in block order
    const int a0 = row_ptr[Pp] - B0, a1 = row_ptr[Pp + 1] - B0;
    acc = prod[(size_t)a0 * N3 + r];
    for (int k = a0 + 1; k < a1; ++k) acc = acc + prod[(size_t)k * N3 + r];
  }
  const size_t o = (size_t)Pp * N3 + r;
  if (MODE == 2) {
    if (act) sw[sub][r] = acc;
    lsync<WPW>();
    if (act && r == 0) {
      double el = 0.0;
      for (int k = 0; k < N3; ++k) el = el + u[(size_t)Pp * N3 + k] * sw[sub][k];
      out[Pp] = el;
    }
    return;
  }
  double w = 0.0;
  if (act) {
    const double gp = gsel ? gtab[gsel[Pp]] : g;
    w = -acc + rs[o] + rd[o] * gp;
  }
  if constexpr (TB == 64) {
    if (act) sw[sub][r] = w;
    lsync<WPW>();
    if (lane < 3 * G) {
      const int sp = lane / 3, comp = lane % 3;
      if (P0 + sp < n_poly) solve_packed<NP>(sR[sp], sw[sp] + comp * NP);
    }
    lsync<WPW>();
    if (act) w = sw[sub][r];
  }
  if (act) {
    if (MODE == 0)
      out[o] = c1 * w + 2.0 * u[o] - wa[o];
    else
      out[o] = c1 * w + u[o] + c2 * wa[o];
  }
}

// ---------------------------------------------------------------------------
// FP32-apply variant of the step: the operator blocks, taken to the
// mass-orthonormal basis in FP64 (k_ortho_operator below), are held in FP32 in
// "column-quad" order (entry (r, c) at ((c>>2)*N3 + r)*4 + (c&3); a lane
// reads four columns of its row as one 16-byte load, half the bytes of the
// FP64 stream); products and the per-block sum are FP32; the sum over blocks,
// the RHS, the basis transforms and the update stay FP64.
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// 32-bit apply in the mass-orthonormal basis.  With M_P = R_P^T R_P per
// polyhedron, the variables ~u = R u turn the leapfrog step into
//   ~u^{n+1} = 2 ~u^n - ~u^{n-1} + dt^2 (~F - ~K ~u^n),
//   ~K_PQ = R_P^-T K_PQ R_Q^-1,   ~F = R^-T F,
// the same iteration with M^-1 turned into the identity.  ~K is scaled like
// M^-1 K whatever the conditioning of the bounding-box Legendre basis on the
// polyhedron, so rounding it to FP32 perturbs the iteration by FP32 rounding
// relative to itself.  Rounding K in the original basis perturbs M^-1 K by
// that rounding times cond(M_P) -- up to 2e8 at p = 4 on agglomerates of
// 4-40 tets, enough to make the iteration diverge.  The 32-bit step also
// needs no triangular solves.
// ---------------------------------------------------------------------------

// S = R^-1 (upper triangular, packed like R), one thread per (polyhedron, column)
template <int P>
__global__ void k_inv_upper(const double* __restrict__ RM, double* __restrict__ S, int n_poly) {
  using D = Dims<P>;
  constexpr int NP = D::NP;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n_poly * NP) return;
  const int Pp = id / NP, j = id % NP;
  const double* R = RM + (size_t)Pp * D::NPK;
  double* Sp = S + (size_t)Pp * D::NPK;
  double x[NP];  // column j of R^-1: R x = e_j, x[i] = 0 for i > j
#pragma unroll
  for (int i = 0; i < NP; ++i) x[i] = 0.0;
  for (int i = j; i >= 0; --i) {
    double s = (i == j) ? 1.0 : 0.0;
    for (int k = i + 1; k <= j; ++k) s = s - R[pk_off(NP, i) + (k - i)] * x[k];
    x[i] = s / R[pk_off(NP, i)];
  }
  for (int i = 0; i <= j; ++i) Sp[pk_off(NP, i) + (j - i)] = x[i];
}

// The 48-bit apply keeps an operator entry k as its FP32 rounding hi and a
// 16-bit correction t = (k - hi) / hi * 2^24 (|t| <= 1 since hi is k
// rounded to nearest; FP16 keeps 11 significant bits of it), so that
// hi (1 + t 2^-24) carries about 35 bits: a relative error of 2^-36 where
// the FP32 rounding leaves 2^-25.
__device__ __forceinline__ _Float16 split_residual(double k, float hi) {
  if (hi == 0.0f) return (_Float16)0.0f;
  return (_Float16)(float)((k - (double)hi) / (double)hi * 0x1p24);
}

// ~K_PQ = S_P^T K_PQ S_Q rounded to FP32 in the column-quad layout; one
// workgroup per block, one component strip (Np rows) at a time:
// T = X S_Q (per column component), Z = S_P^T T.
template <int P>
__global__ __launch_bounds__(256) void k_ortho_operator(const double* __restrict__ K,
                                                        const double* __restrict__ S,
                                                        const int* __restrict__ blk_row,
                                                        const int* __restrict__ col,
                                                        float* __restrict__ K32,
                                                        _Float16* __restrict__ K16, size_t b0) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3, NPK = D::NPK;
  __shared__ double SP[NPK], SQ[NPK];
  __shared__ double X[NP * N3], T[NP * N3];
  const int tid = threadIdx.x;
  const size_t b = b0 + blockIdx.x;
  const int Pp = blk_row[b], Q = col[b];
  for (int k = tid; k < NPK; k += blockDim.x) {
    SP[k] = S[(size_t)Pp * NPK + k];
    SQ[k] = S[(size_t)Q * NPK + k];
  }
  const double* Kb = K + b * D::BS;
  float* Ob = K32 + b * Dims32<P>::BS;
  _Float16* Rb = K16 ? K16 + b * Dims16<P>::BS : nullptr;  // unused slots zeroed by the host
  for (int i = 0; i < 3; ++i) {
    __syncthreads();
    for (int e = tid; e < NP * N3; e += blockDim.x) {
      const int r = e / N3, c = e % N3;
      X[e] = Kb[kpos(N3, i * NP + r, c)];
    }
    __syncthreads();
    for (int e = tid; e < NP * N3; e += blockDim.x) {  // T = X S_Q, S_Q upper
      const int r = e / N3, c = e % N3, j = c / NP, cc = c % NP;
      double s = 0.0;
      for (int k = 0; k <= cc; ++k) s = s + X[r * N3 + j * NP + k] * SQ[pk_off(NP, k) + (cc - k)];
      T[e] = s;
    }
    __syncthreads();
    for (int e = tid; e < NP * N3; e += blockDim.x) {  // Z = S_P^T T
      const int r = e / N3, c = e % N3;
      double s = 0.0;
      for (int k = 0; k <= r; ++k) s = s + SP[pk_off(NP, k) + (r - k)] * T[k * N3 + c];
      const float hi = (float)s;
      const size_t o = kpos32(N3, i * NP + r, c);
      Ob[o] = hi;
      if (Rb) Rb[kpos16(N3, i * NP + r, c)] = split_residual(s, hi);
    }
  }
  // the floats that round the block stride up to a multiple of 4
  constexpr int PAD = (int)(Dims32<P>::BS - (size_t)N3 * N3);
  if constexpr (PAD > 0)
    if (tid < PAD) Ob[(size_t)N3 * N3 + tid] = 0.0f;
}

// Mass-orthonormal transforms with one wave per polyhedron, lane i on row i
// of all three components: MODE 0 ~u = R u (row i sums R_ik x_k over k >= i
// in ascending k, x_k broadcast from lane k); MODE 2 ~F = R^-T F (forward
// substitution swept by columns: y_j = s_j / R_jj from lane j, then every row
// below subtracts R_ji y_j, so row i subtracts in ascending k as a serial
// loop would).  `list` (n entries) restricts the transform to the listed
// polyhedra; null = the first n.
template <int P, int MODE>
__global__ __launch_bounds__(64) void k_tilde(const double* __restrict__ RM,
                                              const double* __restrict__ in,
                                              double* __restrict__ out,
                                              const int* __restrict__ list, int n) {
  using D = Dims<P>;
  constexpr int NP = D::NP;
  static_assert(NP <= 64, "one lane per row");
  static_assert(MODE == 0 || MODE == 2, "mode 1 is k_tilde_back");
  const int w = blockIdx.x;
  if (w >= n) return;
  const int Pp = list ? list[w] : w;
  const int i = threadIdx.x;
  const bool row = i < NP;
  const double* R = RM + (size_t)Pp * D::NPK;
  const double* x = in + (size_t)Pp * D::N3;
  double* y = out + (size_t)Pp * D::N3;
  double x0 = 0.0, x1 = 0.0, x2 = 0.0;
  if (row) {
    x0 = x[i];
    x1 = x[NP + i];
    x2 = x[2 * NP + i];
  }
  if constexpr (MODE == 0) {
    const int ri = row ? pk_off(NP, i) : 0;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const double a0 = __shfl(x0, k, 64), a1 = __shfl(x1, k, 64), a2 = __shfl(x2, k, 64);
      if (row && i <= k) {
        const double r = R[ri + (k - i)];
        s0 = s0 + r * a0;
        s1 = s1 + r * a1;
        s2 = s2 + r * a2;
      }
    }
    if (row) {
      y[i] = s0;
      y[NP + i] = s1;
      y[2 * NP + i] = s2;
    }
  } else {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const int rj = pk_off(NP, j);
      const double rjj = R[rj];
      const double y0 = __shfl(x0, j, 64) / rjj;
      const double y1 = __shfl(x1, j, 64) / rjj;
      const double y2 = __shfl(x2, j, 64) / rjj;
      if (i == j) {
        y[i] = y0;
        y[NP + i] = y1;
        y[2 * NP + i] = y2;
      } else if (row && i > j) {
        const double r = R[rj + (i - j)];
        x0 = x0 - r * y0;
        x1 = x1 - r * y1;
        x2 = x2 - r * y2;
      }
    }
  }
}

// u = R^-1 ~u (mode 1 of the transforms above) with one wave per polyhedron: lane i keeps
// row i's partial sums of the three components, and the columns are swept
// from the last: y_j = s_j / R_jj is broadcast from lane j, then every row
// above subtracts R_ij y_j.  NP steps instead of NP(NP+1)/2 dependent ones
// per thread; row i subtracts its terms in descending column order.
template <int P>
__global__ __launch_bounds__(64) void k_tilde_back(const double* __restrict__ RM,
                                                   const double* __restrict__ in,
                                                   double* __restrict__ out,
                                                   const int* __restrict__ list, int n) {
  using D = Dims<P>;
  constexpr int NP = D::NP;
  static_assert(NP <= 64, "one lane per row");
  const int w = blockIdx.x;
  if (w >= n) return;
  const int Pp = list ? list[w] : w;
  const int i = threadIdx.x;
  const bool row = i < NP;
  const double* R = RM + (size_t)Pp * D::NPK;
  const double* x = in + (size_t)Pp * D::N3;
  double* y = out + (size_t)Pp * D::N3;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (row) {
    s0 = x[i];
    s1 = x[NP + i];
    s2 = x[2 * NP + i];
  }
  const int ri = row ? pk_off(NP, i) : 0;
#pragma unroll
  for (int j = NP - 1; j >= 0; --j) {
    const double rjj = R[pk_off(NP, j)];
    const double y0 = __shfl(s0, j, 64) / rjj;
    const double y1 = __shfl(s1, j, 64) / rjj;
    const double y2 = __shfl(s2, j, 64) / rjj;
    if (i == j) {
      y[i] = y0;
      y[NP + i] = y1;
      y[2 * NP + i] = y2;
    } else if (i < j) {
      const double r = R[ri + (j - i)];
      s0 = s0 - r * y0;
      s1 = s1 - r * y1;
      s2 = s2 - r * y2;
    }
  }
}

template <int P, int MODE, int UNR, int WPP = 1>
__global__ __launch_bounds__(64 * WPP) void k_leapfrog32(const float* __restrict__ K,
                                                   const int* __restrict__ row_ptr,
                                                   const int* __restrict__ col,
                                                   const double* __restrict__ Rp,
                                                   const double* __restrict__ rs,
                                                   const double* __restrict__ rd,
                                                   const double* __restrict__ u,
                                                   const double* __restrict__ wa,
                                                   double* __restrict__ out, double g, double c1,
                                                   double c2, const int* __restrict__ gsel,
                                                   const double* __restrict__ gtab, int n_poly,
    const int* __restrict__ order) {
  // all vectors in the mass-orthonormal basis (K is ~K); Rp is unused
  using D = Dims<P>;
  constexpr int N3 = D::N3, NQUAD = Dims32<P>::NQ, TAIL = Dims32<P>::T;
  constexpr int RPL = (N3 + 63) / 64;
  extern __shared__ double sblk[];  // WPP > 1: per-block products [block][row]
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int Pp = step_poly(order);
  if (Pp >= n_poly) return;  // padding workgroups of the last XCD share
  double acc[RPL];
#pragma unroll
  for (int k = 0; k < RPL; ++k) acc[k] = 0.0;
  const int b0 = row_ptr[Pp], b1 = row_ptr[Pp + 1];
  for (int b = b0 + wv; b < b1; b += WPP) {
    const int cb = __builtin_amdgcn_readfirstlane(col[b]);
    const double* __restrict__ xv = u + (size_t)cb * N3;
    const float* __restrict__ Kb = K + (size_t)b * Dims32<P>::BS;
    float xl[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int r = lane + 64 * k;
      xl[k] = r < N3 ? (float)xv[r] : 0.0f;
    }
    float s[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) s[k] = 0.0f;
#pragma unroll 1
    for (int q0 = 0; q0 < NQUAD; q0 += UNR) {
      float4 kv[UNR][RPL];
#pragma unroll
      for (int uu = 0; uu < UNR; ++uu) {
        const int q = q0 + uu;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const int r = lane + 64 * k;
          if ((NQUAD % UNR == 0 || q < NQUAD) && r < N3)
            kv[uu][k] = load_stream(Kb + ((size_t)q * N3 + r) * 4);
        }
      }
#pragma unroll
      for (int uu = 0; uu < UNR; ++uu) {
        const int q = q0 + uu;
        if (NQUAD % UNR == 0 || q < NQUAD) {
          float xq[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int c = 4 * q + j;
            const float src = (RPL > 1 && (c >> 6)) ? xl[RPL - 1] : xl[0];
            xq[j] = (c < N3) ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(src), c & 63))
                             : 0.0f;
          }
#pragma unroll
          for (int k = 0; k < RPL; ++k) {
            const int r = lane + 64 * k;
            if (r < N3) {
              s[k] = s[k] + kv[uu][k].x * xq[0];
              s[k] = s[k] + kv[uu][k].y * xq[1];
              s[k] = s[k] + kv[uu][k].z * xq[2];
              s[k] = s[k] + kv[uu][k].w * xq[3];
            }
          }
        }
      }
    }
    if constexpr (TAIL > 0) {  // last N3 % 4 columns, stored row by row
      constexpr int C0 = 4 * NQUAD;
      float xt[TAIL];
#pragma unroll
      for (int j = 0; j < TAIL; ++j) {
        const int c = C0 + j;
        const float src = (RPL > 1 && (c >> 6)) ? xl[RPL - 1] : xl[0];
        xt[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(src), c & 63));
      }
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int r = lane + 64 * k;
        if (r < N3) {
          float kt[TAIL];
          load_tail<TAIL>(Kb + (size_t)C0 * N3 + (size_t)r * TAIL, kt);
#pragma unroll
          for (int j = 0; j < TAIL; ++j) s[k] = s[k] + kt[j] * xt[j];
        }
      }
    }
    if constexpr (WPP == 1) {
#pragma unroll
      for (int k = 0; k < RPL; ++k) acc[k] = (b == b0) ? (double)s[k] : acc[k] + (double)s[k];
    } else {
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int r = lane + 64 * k;
        if (r < N3) sblk[(size_t)(b - b0) * N3 + r] = (double)s[k];
      }
    }
  }
  if constexpr (WPP > 1) {
    __syncthreads();
    if (wv != 0) return;  // no barrier follows in this kernel
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int r = lane + 64 * k;
      if (r < N3) {
        double a = sblk[r];
        for (int j = 1; j < b1 - b0; ++j) a = a + sblk[(size_t)j * N3 + r];
        acc[k] = a;
      }
    }
  }
  const size_t o = (size_t)Pp * N3;
  // time factor: one scalar, or per source group (ensembles whose instances
  // carry their own wavelet; the table is filled for this step)
  const double gp = gsel ? gtab[gsel[Pp]] : g;
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    const int r = lane + 64 * k;
    if (r < N3) {
      const double w = -acc[k] + rs[o + r] + rd[o + r] * gp;
      if (MODE == 0)
        out[o + r] = c1 * w + 2.0 * u[o + r] - wa[o + r];
      else
        out[o + r] = c1 * w + u[o + r] + c2 * wa[o + r];
    }
  }
}

// The 48-bit apply: k_leapfrog32's walk over ~K (mass-orthonormal basis, no
// solves) with each entry hi (1 + t 2^-24) (split_residual): hi x summed in
// FP64 fused multiply-adds from the FP64 coefficients, the correction
// (hi t) x in FP32 and scaled by 2^-24 once per block, then the block
// products added in block order in FP64 and k_leapfrog32's update.  6 bytes
// per operator entry; the state stays within about 1e-9 of the FP64 step's
// over 1000 steps where the 32-bit apply drifts to 1e-5.

template <int P, int MODE, int UNR, int WPP>
__global__ __launch_bounds__(64 * WPP) void k_leapfrog48(const float* __restrict__ K,
                                                   const _Float16* __restrict__ Kr,
                                                   const int* __restrict__ row_ptr,
                                                   const int* __restrict__ col,
                                                   const double* __restrict__ rs,
                                                   const double* __restrict__ rd,
                                                   const double* __restrict__ u,
                                                   const double* __restrict__ wa,
                                                   double* __restrict__ out, double g, double c1,
                                                   double c2, const int* __restrict__ gsel,
                                                   const double* __restrict__ gtab, int n_poly,
                                                   const int* __restrict__ order) {
  using D = Dims<P>;
  constexpr int N3 = D::N3, NQUAD = Dims32<P>::NQ, TAIL = Dims32<P>::T;
  constexpr int RPL = (N3 + 63) / 64;
  static_assert(UNR % 2 == 0, "whole quad pairs per load group");
  extern __shared__ double sblk[];  // WPP > 1: per-block products [block][row]
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int Pp = step_poly(order);
  if (Pp >= n_poly) return;  // padding workgroups of the last XCD share
  double acc[RPL];
#pragma unroll
  for (int k = 0; k < RPL; ++k) acc[k] = 0.0;
  const int b0 = row_ptr[Pp], b1 = row_ptr[Pp + 1];
  for (int b = b0 + wv; b < b1; b += WPP) {
    const int cb = __builtin_amdgcn_readfirstlane(col[b]);
    const double* __restrict__ xv = u + (size_t)cb * N3;
    const float* __restrict__ Kb = K + (size_t)b * Dims32<P>::BS;
    const _Float16* __restrict__ Rb = Kr + (size_t)b * Dims16<P>::BS;
    double xl[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int r = lane + 64 * k;
      xl[k] = r < N3 ? xv[r] : 0.0;
    }
    double s[RPL];
    float sc[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      s[k] = 0.0;
      sc[k] = 0.0f;
    }
#pragma unroll 1
    for (int q0 = 0; q0 < NQUAD; q0 += UNR) {  // UNR even: whole quad pairs
      float4 kv[UNR][RPL];
      lymph_hv8 tv[UNR / 2][RPL];
#pragma unroll
      for (int uu = 0; uu < UNR; ++uu) {
        const int q = q0 + uu;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const int r = lane + 64 * k;
          if (q < NQUAD && r < N3) {
            kv[uu][k] = load_stream(Kb + ((size_t)q * N3 + r) * 4);
            if ((uu & 1) == 0) {
              if (q + 1 < NQUAD) {
                tv[uu / 2][k] = load_stream_h8(Rb + ((size_t)(q >> 1) * N3 + r) * 8);
              } else {  // the last quad of an odd count: 4 values per row
                const lymph_hv4 h = load_stream_h4(Rb + Dims16<P>::HALF0 + (size_t)r * 4);
                tv[uu / 2][k] = lymph_hv8{h.x, h.y, h.z, h.w, 0, 0, 0, 0};
              }
            }
          }
        }
      }
#pragma unroll
      for (int uu = 0; uu < UNR; ++uu) {
        const int q = q0 + uu;
        if (q < NQUAD) {
          double xq[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int c = 4 * q + j;
            const double src = (RPL > 1 && (c >> 6)) ? xl[RPL - 1] : xl[0];
            xq[j] = (c < N3) ? lane_bcast(src, c & 63) : 0.0;
          }
#pragma unroll
          for (int k = 0; k < RPL; ++k) {
            const int r = lane + 64 * k;
            if (r < N3) {
              const float4 h = kv[uu][k];
              const lymph_hv8 t8 = tv[uu / 2][k];
              const int o4 = (uu & 1) * 4;
              s[k] = fma((double)h.x, xq[0], s[k]);
              s[k] = fma((double)h.y, xq[1], s[k]);
              s[k] = fma((double)h.z, xq[2], s[k]);
              s[k] = fma((double)h.w, xq[3], s[k]);
              sc[k] = fmaf(h.x * (float)t8[o4 + 0], (float)xq[0], sc[k]);
              sc[k] = fmaf(h.y * (float)t8[o4 + 1], (float)xq[1], sc[k]);
              sc[k] = fmaf(h.z * (float)t8[o4 + 2], (float)xq[2], sc[k]);
              sc[k] = fmaf(h.w * (float)t8[o4 + 3], (float)xq[3], sc[k]);
            }
          }
        }
      }
    }
    if constexpr (TAIL > 0) {  // last N3 % 4 columns, stored row by row
      constexpr int C0 = 4 * NQUAD;
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int r = lane + 64 * k;
#pragma unroll
        for (int j = 0; j < TAIL; ++j) {
          const int c = C0 + j;
          const double src = (RPL > 1 && (c >> 6)) ? xl[RPL - 1] : xl[0];
          const double xt = lane_bcast(src, c & 63);
          if (r < N3) {
            const float h = Kb[(size_t)C0 * N3 + (size_t)r * TAIL + j];
            const float t = (float)Rb[Dims16<P>::TAIL0 + (size_t)r * TAIL + j];
            s[k] = fma((double)h, xt, s[k]);
            sc[k] = fmaf(h * t, (float)xt, sc[k]);
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < RPL; ++k) s[k] = s[k] + (double)sc[k] * 0x1p-24;
    if constexpr (WPP == 1) {
#pragma unroll
      for (int k = 0; k < RPL; ++k) acc[k] = (b == b0) ? s[k] : acc[k] + s[k];
    } else {
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int r = lane + 64 * k;
        if (r < N3) sblk[(size_t)(b - b0) * N3 + r] = s[k];
      }
    }
  }
  if constexpr (WPP > 1) {
    __syncthreads();
    if (wv != 0) return;  // no barrier follows in this kernel
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int r = lane + 64 * k;
      if (r < N3) {
        double a = sblk[r];
        for (int j = 1; j < b1 - b0; ++j) a = a + sblk[(size_t)j * N3 + r];
        acc[k] = a;
      }
    }
  }
  const size_t o = (size_t)Pp * N3;
  const double gp = gsel ? gtab[gsel[Pp]] : g;
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    const int r = lane + 64 * k;
    if (r < N3) {
      const double w = -acc[k] + rs[o + r] + rd[o + r] * gp;
      if (MODE == 0)
        out[o + r] = c1 * w + 2.0 * u[o + r] - wa[o + r];
      else
        out[o + r] = c1 * w + u[o + r] + c2 * wa[o + r];
    }
  }
}

// The pairs-once step of the 48-bit apply: k_sym_products32's walk over the
// stored list and its packed self blocks, with k_leapfrog48's arithmetic
// (hi x in FP64 fused multiply-adds, (hi t) x in FP32, joined with the 2^-24
// scale per block), products stored as doubles (prod_stride(3Np) per block)
// and summed by k_sym_update32<P, MODE, double>.  Here the row chain of each
// chunk runs first and writes the chunk's hi and hi t to LDS as it forms
// them; the chunk's columns then run their chains over ascending rows with
// u_P[r], the chain Q's own row runs over block (Q, P), so a pair block's
// products are the full rows' bit for bit where the two stored blocks round
// alike.  A self block's product is its upper triangle's row chain (diagonal
// included) plus its strict upper triangle's column chain: the full row's
// terms in another order, which moves its FP64 sum by rounding only.  p = 3
// (3Np = 60: whole quads per chunk, no tail columns).
template <int P>
__global__ __launch_bounds__(64) void k_sym_products48(const float* __restrict__ Ks,
                                                       const _Float16* __restrict__ Kr,
                                                       const int* __restrict__ row_ptr,
                                                       const int* __restrict__ col,
                                                       const int* __restrict__ sidx,
                                                       const int* __restrict__ tdst,
                                                       const double* __restrict__ u,
                                                       double* __restrict__ prod,
                                                       double* __restrict__ rowp, int n_poly,
                                                       const int* __restrict__ order) {
  constexpr int N3 = Dims32<P>::N3, NQ = Dims32<P>::NQ, NPR = Dims16<P>::NPR;
  constexpr int PS = prod_stride(N3);
  static_assert(N3 <= 64 && N3 % 12 == 0 && Dims32<P>::T == 0, "one lane per row, whole quads per chunk");
  static_assert(self32_off(NQ) <= (int)Dims32<P>::BS, "packed self block fits its slot");
  static_assert(self16_off(N3, NPR) <= (int)Dims16<P>::BS, "packed self corrections fit their slot");
  constexpr int CW = N3 / 3, CQ = CW / 4;
  // one chunk's hi and hi t row by row (lanes past the rows write padding
  // rows): 16-byte writes, and a column read over 20 lanes hits 20 banks
  __shared__ __align__(16) float sH[64 * CW], sC[64 * CW];
  __shared__ __align__(16) double sU[N3];  // u_P, read by the column chains
  __shared__ __align__(16) double sX[64];  // u_Q of the current block, read by the row chains
  const int lane = threadIdx.x;
  const bool row = lane < N3;
  const int rl = row ? lane : N3 - 1;  // lanes past the rows load the last row's (cached) lines
  const int Pp = step_poly(order);
  if (Pp >= n_poly) return;  // padding workgroups of the last XCD share
  const int b0 = row_ptr[Pp], b1 = row_ptr[Pp + 1];
  const int lcb = lane < b1 - b0 ? col[b0 + lane] : 0;
  const int lsi = lane < b1 - b0 ? sidx[b0 + lane] : 0;
  const int ltd = lane < b1 - b0 ? tdst[b0 + lane] : 0;
  if (row) sU[lane] = u[(size_t)Pp * N3 + lane];  // ordered before its reads by the chunks' syncs
  double racc = 0.0;  // the row products of this wave's stored blocks, in block order
  for (int b = b0; b < b1; ++b) {
    const int j = b - b0;
    const int si = j < 64 ? __builtin_amdgcn_readlane(lsi, j) : __builtin_amdgcn_readfirstlane(sidx[b]);
    if (si < 0) continue;  // written by the neighbour's wave
    const int cb = j < 64 ? __builtin_amdgcn_readlane(lcb, j) : __builtin_amdgcn_readfirstlane(col[b]);
    const int td = j < 64 ? __builtin_amdgcn_readlane(ltd, j) : __builtin_amdgcn_readfirstlane(tdst[b]);
    const float* __restrict__ Kb = Ks + (size_t)si * Dims32<P>::BS;
    const _Float16* __restrict__ Rb = Kr + (size_t)si * Dims16<P>::BS;
    const double xl = u[(size_t)cb * N3 + rl];
    const bool pair = cb != Pp;
    // the self blocks' row clamps and masks are formed per block: hoisted
    // out of the loop they would hold 60 masks in scalar registers
    int rb = rl;
    asm volatile("" : "+v"(rb));
    // one load path for both kinds (an address select per load), so that the
    // chunks wait on their own loads only
    float4 kq[NQ];
    lymph_hv8 tq[NPR];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      kq[q] = load_stream(Kb + (pair ? (q * N3 + rb) * 4 : self32_off(q) + 4 * (rb < 4 * q + 3 ? rb : 4 * q + 3)));
#pragma unroll
    for (int k = 0; k < NPR; ++k) {
      const int last = self16_rows(N3, k) - 1;
      const int self_at = self16_off(N3, k) + self16_w(N3, k) * (rb < last ? rb : last);
      if (2 * k + 1 < NQ) {
        tq[k] = load_stream_h8(Rb + (pair ? (k * N3 + rb) * 8 : self_at));
      } else {  // the last quad of an odd count: 4 values per row
        const lymph_hv4 h = load_stream_h4(Rb + (pair ? (int)Dims16<P>::HALF0 + rb * 4 : self_at));
        tq[k] = lymph_hv8{h.x, h.y, h.z, h.w, 0, 0, 0, 0};
      }
    }
    if (!pair) {  // the upper triangle's rows only
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        if (4 * q + 0 < rb) kq[q].x = 0.0f;
        if (4 * q + 1 < rb) kq[q].y = 0.0f;
        if (4 * q + 2 < rb) kq[q].z = 0.0f;
        if (4 * q + 3 < rb) kq[q].w = 0.0f;
      }
    }
    // three chunks of CW columns: the row chain continues over the chunk's
    // quads, each entry's hi and hi t going to LDS as they are formed, then
    // the chunk's columns run their chains
    wave_lds_sync();  // the previous block's reads of sX are done
    sX[lane] = xl;  // lanes past the rows write padding
    wave_lds_sync();
    double s = 0.0, st = 0.0;
    float sc = 0.0f;
#pragma unroll
    for (int h = 0; h < 3; ++h) {
      wave_lds_sync();  // the previous chunk's reads are done
#pragma unroll
      for (int k = 0; k < CQ; ++k) {
        const int q = h * CQ + k;
        const double2 xa = *reinterpret_cast<const double2*>(sX + 4 * q);  // broadcasts
        const double2 xb = *reinterpret_cast<const double2*>(sX + 4 * q + 2);
        const double xq[4] = {xa.x, xa.y, xb.x, xb.y};
        const float hv[4] = {kq[q].x, kq[q].y, kq[q].z, kq[q].w};
        const lymph_hv8 t8 = tq[q >> 1];
        const int o4 = (q & 1) * 4;
        float hc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          hc[j] = hv[j] * (float)t8[o4 + j];
          s = fma((double)hv[j], xq[j], s);
          sc = fmaf(hc[j], (float)xq[j], sc);
        }
        *reinterpret_cast<float4*>(sH + lane * CW + 4 * k) = kq[q];
        *reinterpret_cast<float4*>(sC + lane * CW + 4 * k) = make_float4(hc[0], hc[1], hc[2], hc[3]);
      }
      // the chunk's row chain is complete here: otherwise the compiler defers
      // it past the column chains and keeps every broadcast u_Q live
      asm volatile("" : "+v"(s), "+v"(sc));
      // a self block's diagonal belongs to its rows only (a wave's LDS
      // writes land in program order)
      if (!pair && lane >= h * CW && lane < (h + 1) * CW) {
        sH[lane * CW + (lane - h * CW)] = 0.0f;
        sC[lane * CW + (lane - h * CW)] = 0.0f;
      }
      wave_lds_sync();
      if (lane >= h * CW && lane < (h + 1) * CW) {  // column's chain over ascending rows
        const float* __restrict__ hp = sH + (lane - h * CW);
        const float* __restrict__ cp = sC + (lane - h * CW);
        double v = 0.0;
        float vc = 0.0f;
#pragma unroll
        for (int r = 0; r < N3; ++r) {
          const double ur = sU[r];  // the same address in every lane: a broadcast
          v = fma((double)hp[r * CW], ur, v);
          vc = fmaf(cp[r * CW], (float)ur, vc);
        }
        st = v + (double)vc * 0x1p-24;
      }
    }
    s = s + (double)sc * 0x1p-24;
    if (!pair) s = s + st;  // the strict upper triangle's transposed product
    asm volatile("" : "+v"(s));  // as in k_sym_products32
    racc = racc + s;
    if (pair && lane < PS) prod[(size_t)td * PS + lane] = row ? st : 0.0;
  }
  if (lane < PS) rowp[(size_t)Pp * PS + lane] = row ? racc : 0.0;  // zeros past the rows
}

// L2 error terms du_i^T M_modal du_i per (polyhedron, component)
// (COMPUTE_ERROR_L2_MATRIX_FREE, src/matrix_free.f90:1406-1416).
template <int P>
__global__ void k_l2_terms(const double* __restrict__ Mm, const double* __restrict__ uh,
                           const double* __restrict__ ue, double* __restrict__ part, int n_poly) {
  using D = Dims<P>;
  constexpr int NP = D::NP;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= 3 * n_poly) return;
  const int Pp = id / 3, i = id % 3;
  const size_t o = (size_t)Pp * D::N3 + i * NP;
  double du[NP];
  for (int mm = 0; mm < NP; ++mm) du[mm] = uh[o + mm] - ue[o + mm];
  const double* A = Mm + (size_t)Pp * NP * NP;
  double el = 0.0;
  for (int mm = 0; mm < NP; ++mm) {
    double s = 0.0;
    for (int nn = 0; nn < NP; ++nn) s = s + A[mm * NP + nn] * du[nn];
    el = el + du[mm] * s;
  }
  part[id] = el;
}

// Sequential sum in index order (the reference accumulates the error terms
// element by element); one thread: called once per error evaluation.
__global__ void k_seq_sum(const double* __restrict__ part, int n, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + part[i];
    out[0] = s;
  }
}

__global__ void k_sub(const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ c, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = a[i] - b[i];
}

// flag[0] = 1 if some a[i] is NaN or infinite (flag cleared by k_set_int
// first; every writer stores the same value, so no atomics are needed)
__global__ void k_set_int(int* __restrict__ flag, int v) { flag[0] = v; }
__global__ void k_nonfinite(const double* __restrict__ a, size_t n, int* __restrict__ flag) {
  bool bad = false;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    bad |= !isfinite(a[i]);
  if (__any(bad) && bad) flag[0] = 1;
}

__global__ void k_scale(double* __restrict__ a, double g, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0 * 0.0 + a[i] * g;
}

// poly-major u[P*N3 + d]  <->  reference column-major u(npoly, N3): u[d*npoly + P]
__global__ void k_to_poly(const double* __restrict__ f, double* __restrict__ p, int n_poly, int n3) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)n_poly * n3) {
    const int P = (int)(i / n3), d = (int)(i % n3);
    p[i] = f[(size_t)d * n_poly + P];
  }
}
__global__ void k_to_fortran(const double* __restrict__ p, double* __restrict__ f, int n_poly, int n3) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)n_poly * n3) {
    const int d = (int)(i / n_poly), P = (int)(i % n_poly);
    f[i] = p[(size_t)P * n3 + d];
  }
}

// Receivers: value of the modal field at a point, sum_i u_i Lx Ly Lz on the
// polyhedron's bounding box, as EXPORT_SOLUTION evaluates it
// (src/solution_processing.f90:150-170).  One thread per (receiver, comp);
// out[r*3 + comp] of record row `row`.
template <int P>
__global__ void k_sample(const double* __restrict__ u, const int* __restrict__ rpoly,
                         const double* __restrict__ rxyz, const double* __restrict__ bbox,
                         const int* __restrict__ blist, int n, double* __restrict__ out) {
  using D = Dims<P>;
  constexpr int NP = D::NP;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= 3 * n) return;
  const int r = id / 3, comp = id % 3;
  const int Pp = rpoly[r];
  const double* bb = bbox + (size_t)Pp * 6;
  double xp[3], sq[3];
  for (int d = 0; d < 3; ++d) {
    const double hb = (bb[2 * d + 1] - bb[2 * d]) / 2.0, mb = (bb[2 * d + 1] + bb[2 * d]) / 2.0;
    xp[d] = (rxyz[r * 3 + d] - mb) / hb;
    sq[d] = sqrt(hb);
  }
  const double* uu = u + (size_t)Pp * D::N3 + comp * NP;
  double s = 0.0;
  for (int i = 0; i < NP; ++i) {
    const double vx = leg(blist[i * 3 + 0], xp[0]) / sq[0];
    const double vy = leg(blist[i * 3 + 1], xp[1]) / sq[1];
    const double vz = leg(blist[i * 3 + 2], xp[2]) / sq[2];
    s = s + uu[i] * vx * vy * vz;
  }
  out[id] = s;
}

// Point forces: rhs[P_k][i*Np+m] += f_k,i phi_m(x_k), sources applied in order
// (one workgroup, a barrier between sources so two sources in one
// polyhedron accumulate in order).
template <int P>
__global__ __launch_bounds__(128) void k_point_sources(const double* __restrict__ sxyz,
                                                       const int* __restrict__ spoly,
                                                       const double* __restrict__ sf, int n,
                                                       const double* __restrict__ bbox,
                                                       const int* __restrict__ blist,
                                                       double* __restrict__ rhs,
                                                       const int* __restrict__ tet_poly) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3;
  for (int k = 0; k < n; ++k) {
    // spoly holds polyhedra, or tets when their polyhedron map is given
    const int Pp = tet_poly ? tet_poly[spoly[k]] : spoly[k];
    const double* bb = bbox + (size_t)Pp * 6;
    for (int e = threadIdx.x; e < N3; e += blockDim.x) {
      const int i = e / NP, m = e % NP;
      double v[3];
      for (int d = 0; d < 3; ++d) {
        const double hb = (bb[2 * d + 1] - bb[2 * d]) / 2.0, mb = (bb[2 * d + 1] + bb[2 * d]) / 2.0;
        v[d] = leg(blist[m * 3 + d], (sxyz[k * 3 + d] - mb) / hb) / sqrt(hb);
      }
      const size_t o = (size_t)Pp * N3 + e;
      rhs[o] = rhs[o] + sf[k * 3 + i] * v[0] * v[1] * v[2];
    }
    __syncthreads();
  }
}


// Double-couple sources, the SISM branch of MAKE_RHS_FREE
// (src/matrix_free.f90:600-664): for source k in order, in tet stet[k] of
// polyhedron P, the moment density M = ((s n^T + n s^T) M0) / V with V =
// |det4| / 6 of the tet (moment_density, src/problem_data_and_properties.f90:
// 317-335; volume, src/find_tet.f90:184-210), and MAKE_DOUBLE_COUPLE
// (src/assemble_element.f90:116-167): couple[i*Np+m] = sum_q (w_q |J|)
// sum_j M_ij d_j phi_m(x_q) over the tet's quadrature points with P's basis,
// then rhs[P] += couple.  One workgroup, thread = entry (i, m), sources in
// order so that sources sharing a polyhedron add as the reference adds them.
__device__ inline double det4_rows(const double* v) {  // rows (x_r, y_r, z_r, 1)
#define A_(i, j) ((j) == 4 ? 1.0 : v[((i) - 1) * 3 + ((j) - 1)])
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

template <int P>
__global__ __launch_bounds__(128) void k_double_couples(MeshDev m, const int* __restrict__ stet,
                                                        const double* __restrict__ strike,
                                                        const double* __restrict__ normal,
                                                        const double* __restrict__ m0, int n,
                                                        double* __restrict__ rhs) {
  using D = Dims<P>;
  constexpr int NP = D::NP, N3 = D::N3, NQ3 = D::NQ3;
  __shared__ double spt[NQ3 * 3];
  __shared__ int sbl[NP * 3];
  const int tid = threadIdx.x;
  for (int i = tid; i < NP * 3; i += blockDim.x) sbl[i] = m.blist[i];
  for (int k = 0; k < n; ++k) {
    const int t = stet[k];
    const int Pp = m.tet_poly[t];
    double v[12], Fk[12], Jdet;
    for (int i = 0; i < 12; ++i) v[i] = m.tet_xyz[(size_t)t * 12 + i];
    affine_map(v, Fk, Jdet);
    const double vol = fabs(det4_rows(v)) / 6.0;
    double mom[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        mom[i * 3 + j] = ((strike[3 * k + i] * normal[3 * k + j] + strike[3 * k + j] * normal[3 * k + i]) *
                          m0[k]) / vol;
    __syncthreads();  // the previous source's points are no longer read
    for (int q = tid; q < NQ3; q += blockDim.x) map_point(Fk, &m.qtet[q * 4], &spt[q * 3]);
    __syncthreads();
    double bb[6];
    for (int d = 0; d < 6; ++d) bb[d] = m.bbox[(size_t)Pp * 6 + d];
    for (int e = tid; e < N3; e += blockDim.x) {
      const int i = e / NP, mm = e % NP;
      double acc = 0.0;
      for (int q = 0; q < NQ3; ++q) {
        double ph, g[3];
        basis_at(&sbl[mm * 3], bb, &spt[q * 3], ph, g[0], g[1], g[2]);
        double reduct = 0.0;
        for (int j = 0; j < 3; ++j) reduct = reduct + mom[i * 3 + j] * g[j];
        acc = acc + m.wtet[q] * fabs(Jdet) * reduct;
      }
      const size_t o = (size_t)Pp * N3 + e;
      rhs[o] = rhs[o] + acc;
    }
  }
}

// Per-group Ricker time factors of one step: gtab[i] = ricker(f0[i], t0[i], t).
__global__ void k_ricker_table(const double* __restrict__ f0, const double* __restrict__ t0,
                               int n, double t, double* __restrict__ gtab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) gtab[i] = ricker(f0[i], t0[i], t);
}

}  // namespace lymph

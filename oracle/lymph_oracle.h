/*
 * lymph_oracle.h -- CPU restatement of LYMPH3D's matrix-free PolyDG leapfrog path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It is never called by the product library.
 *
 * Every function follows one routine of the reference (paths relative to the
 * reference repository root), term by term and in the same summation order:
 *   quadrature / GL nodes      src/basis_function.f90:47-161
 *   Duffy & triangle maps      src/Poly_ref_mappings.f90:137-186
 *   affine map, det3           src/Poly_ref_mappings.f90:18-75, src/vet_mat_operations.f90:38-49
 *   face -> ref-tet maps       src/Poly_ref_mappings.f90:79-134
 *   basis list / Legendre      src/basis_function.f90:12-44,268-648
 *   volume stiffness, mass     src/assemble_element.f90:24-113
 *   face SIPG/NIPG blocks      src/assemble_element.f90:404-661
 *   RHS volume / face          src/assemble_element.f90:171-390
 *   operator assembly          src/matrix_free.f90:27-464
 *   RHS assembly               src/matrix_free.f90:467-843
 *   L2 projection              src/matrix_free.f90:847-997
 *   first / regular step       src/matrix_free.f90:1100-1385
 *   L2 / DG error              src/matrix_free.f90:1388-1504
 *   Cholesky, solve_LU         src/vet_mat_operations.f90:84-153
 *   time loop                  src/Lymph3D.f90:303-418, errors :450-518
 *   normals/areas/bbox/hk      src/MAKE_PARTITION_AND_MPI_FILES.f90:1086-1095,1429-1580
 *   face matching, BC tags     src/MAKE_PARTITION_AND_MPI_FILES.f90:1914-2200
 *   problem data               src/problem_data_and_properties.f90
 *   double couples (SISM)      src/matrix_free.f90:600-664, src/assemble_element.f90:116-167,
 *                              moment_density src/problem_data_and_properties.f90:317-335,
 *                              det4 src/vet_mat_operations.f90:52-69, volume src/find_tet.f90:184-210
 *
 * Polyhedra: the reference stores one block row per tetrahedron.  Here each
 * polyhedron owns one block row (self block first, then one block per
 * neighbouring polyhedron in first-appearance order over its tets' faces).
 * With one tet per polyhedron this is exactly the reference's K_loc(e,1..5).
 *
 * All arrays are 0-based, row-major.  Element DOF index: comp*Np + mode.
 */
#ifndef LYMPH_ORACLE_H
#define LYMPH_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OCtx OCtx;

/* Problem data sets (src/problem_data_and_properties.f90 and the cube wave). */
enum {
  O_PROB_NULL = 0,       /* f_null, gd_null, gn_null */
  O_PROB_REF_MMS = 1,    /* f_time, gd, gn; g(t) = sin(sqrt2*pi*t) */
  O_PROB_CUBE_WAVE = 2   /* u_i = sin(pi x)sin(pi y)sin(pi z) cos(sqrt3*pi*t) */
};
/* Fields for L2 projection. */
enum {
  O_FN_ZERO = 0,         /* ic_displacement / uex_stat */
  O_FN_REF_VEL = 1,      /* ic_velocity */
  O_FN_REF_UEX = 2,      /* uex */
  O_FN_CUBE_U = 3,       /* sin sin sin (1,1,1) */
  O_FN_CUBE_V = 4,       /* zero */
  O_FN_POLY = 5          /* the polynomial field set by o_set_poly_field */
};
#define O_MAX_POLY_TERMS 64

/* Build the mesh context.
 *  xyz[n_node*3], tet_vert[n_tet*4] (0-based node ids, file order v1..v4),
 *  tet_poly[n_tet] (0-based, may be NULL => one poly per tet),
 *  tria_vert[n_tria*3], tria_tag[n_tria]: boundary triangles with their tag,
 *  tag_kind[tag] in {0 (ignored), -1 Dirichlet, -2 Neumann}, tag_space[tag]
 *  space-function tag, both indexed 0..n_tags-1,
 *  rho/lam/mu per tet.  Returns 0 or an error code. */
int o_ctx_create(int n_node, const double* xyz, int n_tet, const int* tet_vert,
                 const int* tet_poly, int n_tria, const int* tria_vert,
                 const int* tria_tag, int n_tags, const int* tag_kind,
                 const int* tag_space, const double* rho, const double* lam,
                 const double* mu, int p, double alpha, double theta,
                 OCtx** out);
void o_ctx_destroy(OCtx* c);
int o_num_threads(int n);

/* sizes: [n_tet, n_poly, Np, n_blk] */
void o_sizes(const OCtx* c, int* out4);
void o_get_conn(const OCtx* c, int* neigh, int* neigh_face, int* bc_space,
                double* normal, double* area, double* bbox, double* hk);
void o_get_pattern(const OCtx* c, int* row_ptr, int* col);
double o_hmax(const OCtx* c);

int o_build_operator(OCtx* c);
/* copy out assembled data (row-major 3Np x 3Np blocks; Np x Np mass blocks) */
void o_get_K(const OCtx* c, double* K);
void o_get_Adg(const OCtx* c, double* A);
void o_get_mass(const OCtx* c, double* M, double* Mmodal, double* RM,
                double* RMmodal);

int o_build_rhs(const OCtx* c, int problem, double* rhs /* n_poly*3Np */);
int o_project(const OCtx* c, int fn, double* coeff /* n_poly*3Np */);

void o_first_step(const OCtx* c, double g0, double dt, const double* u0,
                  const double* v0, const double* rhs_s, const double* rhs_d,
                  double* u1);
void o_step(const OCtx* c, double g, double dt, const double* um1,
            const double* u, const double* rhs_s, const double* rhs_d,
            double* up1);
double o_error_l2_sq(const OCtx* c, const double* uh, const double* uex);
double o_error_dg_sq(const OCtx* c, const double* uh, const double* uex);

/* Whole driver (src/Lymph3D.f90:262-518): operator, RHS, IC projection, first
 * step, "while (t <= stop)" loop, exact-solution projection and errors.
 * max_steps < 0 means unlimited.  out_f[0..3] = t_final, nsteps, errL2, errDG.
 * u_final may be NULL. */
int o_run(OCtx* c, int problem, double dt, double stop_time, int max_steps,
          double* u_final, double* out_f);

double o_time_function(int problem, double t);

/* receivers: nearest-barycentre location and point evaluation */
void o_locate_points(const OCtx* c, int n, const double* pts, int* tet);
void o_eval_points(const OCtx* c, int n, const double* pts, const int* tet, const double* u,
                   double* out);

/* point force sources and the Ricker wavelet */
void o_point_sources(const OCtx* c, int n, const double* pts, const int* tet, const double* f,
                     double* rhs);
double o_ricker(double f0, double t0, double t);

/* Double-couple sources (the SISM branch of MAKE_RHS_FREE): det4, the
 * moment density, MAKE_DOUBLE_COUPLE on one tet, and the loop over sources
 * adding each couple to its polyhedron's entries of rhs[n_poly*3Np]. */
double o_det4(const double* A /* 4x4 row-major */);
void o_moment_density(double M0, double V, const double* s, const double* n, double* mom /* 3x3 */);
void o_double_couple(int p, double Jdet, const double* dphi /* [3][Np][nq3] */, const double* mom,
                     double* out /* 3Np */);
void o_double_couples(const OCtx* c, int n, const int* tet, const double* strike, const double* normal,
                      const double* m0, double* rhs);

/* Point values of a data set (the reference's f_forcing, f_dirichlet,
 * f_neumann of src/problem_data_and_properties.f90, or the cube wave):
 * f[3n], gd[3n], gn[3n] (any may be NULL) at xyz[3n], normals normal[3n]. */
void o_problem_data(int problem, int n, const double* xyz, const double* normal, double lam,
                    double mu, double rho, int tag, double* f, double* gd, double* gn);
/* Polynomial field for o_project(c, O_FN_POLY, ...): component i is
 * sum_k coef[3k+i] x^exps[3k] y^exps[3k+1] z^exps[3k+2]; nterms <= 64. */
int o_set_poly_field(OCtx* c, int nterms, const int* exps, const double* coef);

/* ---- element-level restatements (checked against the reference's own
 *      Fortran routines by the tests) ---- */
int o_np(int p);
void o_basis_list(int p, int* blist /* Np*3 */);
/* nodtet3[4*nq3] (node-major: q*4+k), weitet3[nq3], nodtria2[4*nq2], weitria2[nq2] */
void o_quadrature(int p, double* nodtet3, double* weitet3, double* nodtria2,
                  double* weitria2);
void o_jacobians(const double* x4, const double* y4, const double* z4,
                 double* Fk /* 3x4 row-major */, double* Jdet);
/* phi[Np*nq3], dphi[3*Np*nq3] at tet quadrature points, basis on bbox[6] */
void o_basis(int p, const double* bbox, const double* Fk, double* phi,
             double* dphi);
/* phi_b[2*Np*nq2] (side-major), grad_b[2*3*Np*nq2]; bbox2 may be NULL */
void o_basis_boundary(int p, int face, const double* bbox1,
                      const double* bbox2, const double* Fk, double* phi_b,
                      double* grad_b);
/* element blocks [3][3][Np][Np] */
void o_stiffness_volume(int p, double Jdet, double lam, double mu,
                        const double* dphi, double* V);
void o_mass_volume(int p, double Jdet, const double* phi, double rho,
                   double* Mfull /* [3][3][Np][Np] */);
/* kind: -1 Dirichlet, -2 Neumann, >=0 interior */
void o_stiffness_face(int p, double alpha, int kind, double hk1, double hk2,
                      const double* normal, double area, double lam, double mu,
                      const double* phi_b, const double* grad_b, double* S1,
                      double* I1, double* S2, double* I2, double* IT2);
int o_cholesky(int n, const double* A, double* R);
void o_solve_lu(int n, const double* R, const double* b, double* x);

#ifdef __cplusplus
}
#endif
#endif

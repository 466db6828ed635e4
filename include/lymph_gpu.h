/*
 * lymph_gpu.h -- C ABI of liblymph_gpu.so, the MI355X (gfx950) implementation
 * of LYMPH3D's matrix-free PolyDG leapfrog path for 3-D linear elastodynamics.
 *
 * The Fortran host keeps its mesh pipeline (READ_INPUT_FILES, partitioning,
 * agglomeration, normals/bounding boxes) and hands the flattened mesh to
 * lymph_ctx_create; every numerical routine of the path below then runs as
 * hand-written HIP kernels on one GPU.  Each entry point cites the reference
 * routine it replaces (paths relative to the reference repository root).
 *
 * Conventions
 *  - Every entry point returns 0 on success or a nonzero code; codes reuse the
 *    reference's exit codes (src/Poly_global.f90:130-173) where one exists:
 *      LYMPH_EXIT_SETUP 4, LYMPH_EXIT_SINGULARMTX 5, LYMPH_EXIT_NEIGHBOUR 18;
 *    plus LYMPH_ERR_ARG 101 (bad argument), LYMPH_ERR_HIP 102 (HIP failure),
 *    LYMPH_ERR_STATE 103 (call order).  lymph_last_error() gives a message.
 *    Nothing aborts across the ABI.
 *  - All arrays are 0-based.  Solution-like vectors hold 3*Np doubles per
 *    polyhedron, component-major inside a polyhedron (index comp*Np + mode,
 *    the reference's (i-1)*Np+m, src/matrix_free.f90:247-252).  Host vectors
 *    use the layout chosen with lymph_ctx_set_layout: LYMPH_LAYOUT_POLY
 *    (u[P*3Np + dof], default) or LYMPH_LAYOUT_FORTRAN (the reference's
 *    column-major u(npoly, 3Np), u[dof*npoly + P]).
 *  - Ownership: the library owns all device memory of a context for its
 *    lifetime; host arrays are only read/written during a call.
 *  - Threading: one host thread per context; calls on one context are not
 *    re-entrant.  Contexts are independent.  All work of a context runs on its
 *    HIP stream (its own, or the one given to lymph_ctx_set_stream); entry
 *    points taking host pointers synchronise that stream before returning,
 *    lymph_step_n / lymph_*_dev entry points do not.
 *  - Time and the time factor g(t) are evaluated by the caller in FP64 and
 *    passed in (the reference's time_function, problem_data :14-23).
 */
#ifndef LYMPH_GPU_H
#define LYMPH_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LYMPH_EXIT_INSTAB 2 /* src/Poly_global.f90 EXIT_INSTAB */
#define LYMPH_EXIT_SETUP 4
#define LYMPH_EXIT_SINGULARMTX 5
#define LYMPH_EXIT_NEIGHBOUR 18
#define LYMPH_ERR_ARG 101
#define LYMPH_ERR_HIP 102
#define LYMPH_ERR_STATE 103

#define LYMPH_LAYOUT_POLY 0
#define LYMPH_LAYOUT_FORTRAN 1

/* Built-in data sets replacing MAKE_RHS_FREE's procedure arguments
 * (src/matrix_free.f90:474-493; src/problem_data_and_properties.f90). */
#define LYMPH_PROB_NULL 0      /* f_null, gd_null, gn_null */
#define LYMPH_PROB_REF_MMS 1   /* f_time, gd, gn: the reference's built-in manufactured solution */
#define LYMPH_PROB_CUBE_WAVE 2 /* u_i = sin(pi x)sin(pi y)sin(pi z) cos(sqrt3 pi t), all components */
#define LYMPH_PROB_RICKER 3    /* no volume/boundary data; point forces (lymph_add_point_sources)
                                  with the Ricker time factor set by lymph_set_ricker */

/* Fields for lymph_project (COMPUTE_MODAL_COEFFICIENTS_FREE's f_analytic). */
#define LYMPH_FN_ZERO 0    /* ic_displacement, uex_stat */
#define LYMPH_FN_REF_VEL 1 /* ic_velocity */
#define LYMPH_FN_REF_UEX 2 /* uex */
#define LYMPH_FN_CUBE_U 3  /* sin(pi x)sin(pi y)sin(pi z) on every component */

/* Device-resident vector slots of a context (used by the *_dev / step_n calls). */
#define LYMPH_SLOT_RHS_STAT 0
#define LYMPH_SLOT_RHS_DYN 1
#define LYMPH_SLOT_U_PREV 2 /* u^{n-1} */
#define LYMPH_SLOT_U_CUR 3  /* u^n     */
#define LYMPH_SLOT_U_NEXT 4 /* u^{n+1} (scratch for the rotation) */
#define LYMPH_SLOT_V0 5
#define LYMPH_SLOT_UEX 6
#define LYMPH_NUM_SLOTS 7

typedef struct LymphCtx LymphCtx;

/* Flattened mesh: the data contract of the reference's Mesh_Structure
 * (src/Poly_mesh.f90:13-104) restricted to what the path reads.
 *  neigh[t*4+f]: neighbouring tet across face f (0-based), or -1 (Dirichlet,
 *   BC_DIRI) / -2 (Neumann, BC_NEUM) -- Element%neigh_el(face,2);
 *  faces f=0..3 are the reference's faces 1..4: (v3,v2,v1), (v1,v2,v4),
 *   (v2,v3,v4), (v3,v1,v4) (src/MAKE_PARTITION_AND_MPI_FILES.f90:1087-1095);
 *  normal/area as CREATE_NORMAL_FACE (:1429-1504): area = |P x Q| (twice the
 *   triangle area), bbox/hk of the polyhedron as CREATE_BBOX_EL (:1510-1580). */
typedef struct LymphMesh {
  int32_t n_node;
  int32_t n_tet;
  int32_t n_poly;
  const double* xyz;        /* n_node*3 */
  const int32_t* tet_vert;  /* n_tet*4 */
  const int32_t* tet_poly;  /* n_tet, or NULL: polyhedron == tet */
  const int32_t* neigh;     /* n_tet*4 */
  const int32_t* bc_space;  /* n_tet*4, space_fun_tag of boundary faces */
  const double* normal;     /* n_tet*12 */
  const double* area;       /* n_tet*4 */
  const double* bbox;       /* n_poly*6: xmin,xmax,ymin,ymax,zmin,zmax */
  const double* hk;         /* n_poly */
  const double* rho;        /* n_tet */
  const double* lam;        /* n_tet */
  const double* mu;         /* n_tet */
} LymphMesh;

/* Message of the last failing call on this thread ("" if none). */
const char* lymph_last_error(void);
/* Library version string. */
const char* lymph_version(void);
/* Number of visible HIP devices (0 without a GPU); never fails. */
int lymph_device_count(void);

/* Create a context on the current HIP device: uploads the mesh, builds the
 * block-row pattern (self block + one block per neighbouring polyhedron) and
 * the quadrature tables.  Replaces the allocations and set-up of
 * src/Lymph3D.f90:219-252 (p from Elem_loc(1)%Degree; alpha, theta from
 * set_properties, src/problem_data_and_properties.f90:304-315). */
int lymph_ctx_create(const LymphMesh* mesh, int p, double alpha, double theta,
                     LymphCtx** out);
int lymph_ctx_destroy(LymphCtx* ctx);
/* info[0..7] = n_tet, n_poly, Np, n_blk, operator bytes, device bytes total,
 *              algorithmic bytes of one FP64 leapfrog step, flops per step. */
int lymph_ctx_info(const LymphCtx* ctx, int64_t* info);
int lymph_ctx_set_layout(LymphCtx* ctx, int layout);
/* Run all work on `stream` (a hipStream_t) instead of the context's own. */
int lymph_ctx_set_stream(LymphCtx* ctx, void* stream);
void* lymph_ctx_get_stream(const LymphCtx* ctx);
/* Block pattern: row_ptr[n_poly+1], col[n_blk]. */
int lymph_get_pattern(const LymphCtx* ctx, int32_t* row_ptr, int32_t* col);

/* MAKE_MATRICES_FREE (src/matrix_free.f90:27-464) followed by the mass-block
 * Cholesky factorisations (src/Lymph3D.f90:264-271), on the GPU.
 * Returns LYMPH_EXIT_SINGULARMTX if a mass block is not SPD. */
int lymph_build_operator(LymphCtx* ctx);
/* Copy the assembled data to the host for checking: K and A_dg as row-major
 * 3Np x 3Np blocks (n_blk of them, pattern order); M, M_modal as full Np x Np
 * per polyhedron; any pointer may be NULL. */
int lymph_export_operator(const LymphCtx* ctx, double* K, double* A_dg,
                          double* M, double* M_modal);

/* MAKE_RHS_FREE (src/matrix_free.f90:467-843) for a built-in data set into a
 * device slot (LYMPH_SLOT_RHS_STAT / _DYN); copied to rhs_out if non-NULL. */
int lymph_build_rhs(LymphCtx* ctx, int problem, int slot, double* rhs_out);
/* MAKE_RHS_FREE with caller-supplied data, the drop-in for the procedure
 * arguments f_forcing, f_dirichlet, f_neumann (src/matrix_free.f90:474-493):
 * the library computes the physical points where the reference evaluates
 * them (volume quadrature points of every tet; data points of every
 * Dirichlet and Neumann face) and calls each function once with all its
 * points; the values then enter the same RHS kernel as the built-in data
 * sets.  Arrays are n x 3 row-major (xyz[3*i+d], normal, outputs) or n
 * entries (material of the point's tet, space-function tag of the face).
 * A NULL function stands for zero data.  Functions run on the calling host
 * thread before this call returns; `user` is passed through. */
typedef void (*lymph_forcing_fn)(int n, const double* lambda, const double* mu, const double* rho,
                                 const double* xyz, double* f, void* user);
typedef void (*lymph_dirichlet_fn)(int n, const double* xyz, const int32_t* space_tag, double* gd,
                                   void* user);
typedef void (*lymph_neumann_fn)(int n, const double* lambda, const double* mu,
                                 const double* normal, const double* xyz,
                                 const int32_t* space_tag, double* gn, void* user);
int lymph_build_rhs_fn(LymphCtx* ctx, lymph_forcing_fn f, lymph_dirichlet_fn gd,
                       lymph_neumann_fn gn, void* user, int slot, double* rhs_out);
/* COMPUTE_MODAL_COEFFICIENTS_FREE (src/matrix_free.f90:847-997): L2
 * projection of field `fn` into a device slot; copied to out if non-NULL. */
int lymph_project(LymphCtx* ctx, int fn, int slot, double* out);

/* COMPUTE_MODAL_COEFFICIENTS_FREE with a caller function, the drop-in for
 * its procedure argument f_analytic: fld(n, xyz, values, user) gets the n
 * volume quadrature points (n x 3 row-major) where the reference evaluates
 * f_analytic and writes n x 3 values; then the same projection kernel as the
 * built-in fields.  Needs the mass factors (lymph_build_operator). */
typedef void (*lymph_field_fn)(int n, const double* xyz, double* values, void* user);
int lymph_project_fn(LymphCtx* ctx, lymph_field_fn fld, void* user, int slot, double* out);

/* Host <-> device slot copies (layout as set). */
int lymph_set_slot(LymphCtx* ctx, int slot, const double* host);
int lymph_get_slot(LymphCtx* ctx, int slot, double* host);
/* Raw device pointer of a slot (poly layout), for device-resident callers. */
double* lymph_slot_ptr(LymphCtx* ctx, int slot);

/* FIRST_TIME_STEP_MATRIX_FREE (src/matrix_free.f90:1100-1227):
 * u1 = u0 + dt v0 + dt^2/2 M^-1 (F_s + F_d g0 - K u0); host pointers, uses
 * the RHS slots built before. */
int lymph_first_step(LymphCtx* ctx, double g0, double dt, const double* u0,
                     const double* v0, double* u1);
/* TIME_STEP_MATRIX_FREE (src/matrix_free.f90:1230-1385) -- the drop-in:
 * u^{n+1} = dt^2 M^-1 (F_s + F_d g - K u^n) + 2u^n - u^{n-1}; host pointers. */
int lymph_step(LymphCtx* ctx, double g, double dt, const double* um1,
               const double* u, double* up1);
/* Same two steps on device pointers (poly layout), asynchronous on the
 * context stream. */
int lymph_first_step_dev(LymphCtx* ctx, double g0, double dt, const double* u0,
                         const double* v0, double* u1);
int lymph_step_dev(LymphCtx* ctx, double g, double dt, const double* um1,
                   const double* u, double* up1);
/* Device-resident time loop body of src/Lymph3D.f90:389-418: `nsteps` steps
 * starting at time t0 (g(t) = time factor of `problem`, evaluated on the host
 * per step with t accumulated as t = t + dt) on slots U_PREV/U_CUR, rotating
 * them; asynchronous.  On return *t_out holds the accumulated time. */
int lymph_step_n(LymphCtx* ctx, int problem, int nsteps, double t0, double dt,
                 double* t_out);
/* Time `nsteps` device-resident steps with HIP events recorded on the context
 * stream around the launches: *ms = elapsed milliseconds (synchronises). */
int lymph_time_steps(LymphCtx* ctx, int problem, int nsteps, double t0,
                     double dt, double* ms);

/* Time every variant of the step kernel for `reps` steps on copies of the
 * state slots, keep the fastest for this context and report them: for
 * variant i, unr[i] = column pairs (quads for the 32- and 48-bit applies)
 * loaded per unrolled group, xmode[i] = how the neighbour's coefficients
 * reach the lanes (0 scalar loads, 1 vector load + lane broadcast, 2 several
 * polyhedra per wave with the coefficients exchanged through LDS, p <= 2, 3
 * the same with the wave's polyhedra walking their blocks jointly), wpp[i] =
 * waves sharing one polyhedron's blocks (their block products are added in
 * block order, so results do not depend on it; for xmode 2 and 3, waves per
 * workgroup, each with its own polyhedra), ms[i] = milliseconds per step (-1: not run,
 * the longest block row does not fit the variant's LDS); arrays of at least
 * 16 entries, any may be NULL; *chosen = index of the variant kept.  No
 * reference counterpart (the reference has one CPU loop); results do not
 * depend on the variant. */
int lymph_autotune(LymphCtx* ctx, int reps, int* n_variants, int* unr, int* xmode, int* wpp,
                   double* ms, int* chosen);
/* Select step-kernel variant `index` (as numbered by lymph_autotune for the
 * current apply precision) without timing; *n_variants (may be NULL) = count.
 * No reference counterpart. */
int lymph_set_step_variant(LymphCtx* ctx, int index, int* n_variants);

/* Operator precision of the step's block apply: 64 (default), 32 or 48.
 * With 32 the steps run in the mass-orthonormal basis u~ = R u (M = R^T R per
 * polyhedron): the blocks R_P^-T K_PQ R_Q^-1 are formed in FP64 and rounded
 * once to FP32 (a second copy, half the bytes per step), products and
 * per-block sums are FP32, the sum over blocks, the RHS, the basis transforms
 * at the start and end of every stepping call and the update stay FP64, and
 * the step needs no triangular solves.  Rounding K itself would perturb
 * M^-1 K by the rounding times cond(M_P), which diverges on ill-conditioned
 * agglomerates.  48 keeps, beside each FP32 entry hi of that operator, the
 * 16-bit correction t = (k - hi) / hi * 2^24 (FP16), so that hi (1 + t 2^-24)
 * carries about 35 bits (6 bytes per entry), and sums the products in FP64:
 * its state stays within about 1e-9 of the 64-bit step's over 1000 steps
 * where the 32-bit one drifts to 1e-5.  Selecting 48 needs the FP64 blocks
 * once (the corrections are formed from them).  No reference counterpart
 * (the reference is FP64 only).  Call after lymph_build_operator; resets the
 * autotuned variant. */
int lymph_set_apply_precision(LymphCtx* ctx, int bits);

/* Device-memory options for large batched runs (no reference counterpart:
 * the reference keeps K_loc and A_dg_loc in host memory per MPI rank).
 * lymph_ctx_set_keep_adg(ctx, 0) before lymph_build_operator skips the
 * A_dg blocks (needed only by lymph_error_dg, which then returns
 * LYMPH_ERR_STATE); after a build it frees them.
 * lymph_release_fp64_operator frees the FP64 K blocks once the 32- or 48-bit
 * apply is selected (lymph_set_apply_precision(ctx, 32 or 48)); 64-bit apply,
 * K export and a first selection of the 48-bit apply then return
 * LYMPH_ERR_STATE until lymph_build_operator runs again. */
int lymph_ctx_set_keep_adg(LymphCtx* ctx, int keep);

/* How lymph_build_operator evaluates MAKE_MATRICES_FREE's sums:
 * LYMPH_BUILD_REFERENCE (default) term by term in the reference's operand
 * order (entries agree with it to the last ulps); LYMPH_BUILD_FACTORED as
 * combinations of per-basis-pair quadrature sums (D_ab = sum w|J| d_a phi_m
 * d_b phi_n, face sums of phi phi and d_a phi phi), about 10x fewer flops,
 * entries equal to rounding (~1e-15 relative); mass blocks are identical in
 * both.  No reference counterpart (a device-side evaluation choice). */
#define LYMPH_BUILD_REFERENCE 0
#define LYMPH_BUILD_FACTORED 1
int lymph_ctx_set_build(LymphCtx* ctx, int mode);
int lymph_release_fp64_operator(LymphCtx* ctx);

/* Modes of lymph_set_symmetric_storage. */
#define LYMPH_SYM_OFF 0       /* full block rows (default) */
#define LYMPH_SYM_FP64 1      /* FP64 step reads each face pair once */
#define LYMPH_SYM_FP64_FP32 2 /* and, at p = 3, so do the 32- and 48-bit applies */

/* Face pairs stored once (no reference counterpart; an operator storage
 * choice behind TIME_STEP_MATRIX_FREE, src/matrix_free.f90:1230-1385).  With
 * SIPG (theta = -1) and one material on both traces of every face -- the
 * reference forces rho = lambda = mu = 1, src/Lymph3D.f90:196-203 -- the
 * blocks satisfy K_QP = K_PQ^T up to the rounding of the two sides' face
 * quadratures.  on = 1 checks every pair and self block on the device
 * (largest max |K_PQ - K_QP^T| / max |K_PQ| to *asym when asym is not null)
 * and, when it is at most 1e-12, keeps K_PQ for Q > P and the upper triangle
 * of each self block only; the FP64 step (lymph_step*, lymph_first_step*,
 * lymph_step_n, lymph_run, captured solves) then reads about half the
 * operator bytes at p = 3 and applies K_QP as the transpose of P's copy.  An operator that fails the check
 * returns LYMPH_ERR_ARG and the full rows stay in use.  Needs p <= 3 and the
 * FP64 blocks (not released); lymph_build_operator drops the stored pairs;
 * on = LYMPH_SYM_OFF frees them.  Only with on = LYMPH_SYM_FP64_FP32 (p = 3)
 * do the 32- and 48-bit applies (lymph_set_apply_precision) keep their blocks
 * once too (the 48-bit apply its 16-bit corrections as well); with
 * LYMPH_SYM_FP64 they keep their full rows.  A failing call leaves no stored
 * pairs switched on.  While the step in use reads stored pairs,
 * lymph_autotune returns LYMPH_ERR_STATE (the full-row variants are not run).
 * The DG norm is unaffected. */
int lymph_set_symmetric_storage(LymphCtx* ctx, int on, double* asym);

/* Point forces f_k delta(x - x_k) added to a RHS slot (normally
 * LYMPH_SLOT_RHS_DYN): rhs[P][i*Np+m] += f_k,i phi_m(x_k) in the polyhedron of
 * tet[k].  No reference counterpart: the reference's point source is the
 * double couple of MAKE_DOUBLE_COUPLE (src/assemble_element.f90:116-167) and
 * its Ricker keyword is parsed but unused (src/Poly_readfile.f90:589-680). */
int lymph_add_point_sources(LymphCtx* ctx, int n, const double* xyz, const int32_t* tet,
                            const double* force, int slot);
/* Double-couple point sources, the SISM branch of MAKE_RHS_FREE
 * (src/matrix_free.f90:600-664; MAKE_DOUBLE_COUPLE, src/assemble_element.f90:
 * 116-167): for source k, in order, located in tet[k] (as FIND_ELEM_FROM_POINT
 * locates it, nearest barycentre), the moment density
 * M = (s n^T + n s^T) M0 / V with strike s = strike[3k..], fault normal
 * n = normal[3k..], moment M0 = m0[k] and V the tet's volume, integrated
 * against the gradients of the polyhedron's basis over the tet:
 * rhs[P][i*Np+m] += sum_q w_q |J| sum_j M_ij d_j phi_m(x_q).  The reference
 * takes the strike, normal and moment of its first source for every source
 * and adds the couples in every MAKE_RHS_FREE call (static and dynamic RHS);
 * its drop-in passes them that way. */
int lymph_add_double_couples(LymphCtx* ctx, int n, const int32_t* tet, const double* strike,
                             const double* normal, const double* m0, int slot);
/* lymph_add_point_sources without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The sources are staged in one file-scope device array of the
 * library (n <= 16384 sources per call, 832 KiB, held by every GPU the
 * library runs on until the process ends) instead of per-call device
 * buffers; it calls no hipMalloc/hipFree.  The context's device must be
 * current.  n > 16384 returns LYMPH_ERR_ARG.  Every call on one GPU shares
 * that GPU's array, so calls that overlap on one GPU (other host threads,
 * other contexts) race; calls return only once the update is done.  Same
 * result as lymph_add_point_sources, bit for bit. */
int lymph_add_point_sources_noalloc(LymphCtx* ctx, int n, const double* xyz, const int32_t* tet,
                                    const double* force, int slot);
/* Time factor of LYMPH_PROB_RICKER for this context: Ricker wavelet of peak
 * frequency f0 centred at t0 (default f0 = 5, t0 = 1.2/f0). */
int lymph_set_ricker(LymphCtx* ctx, double f0, double t0);
/* (1 - 2a) exp(-a), a = (pi f0 (t - t0))^2 */
double lymph_ricker(double f0, double t0, double t);
/* Per-group wavelets for batched instances: polyhedron i takes the Ricker
 * factor of group poly_group[i] (f0[g], t0[g]) in the stepping entry points
 * that evaluate the time factor themselves (lymph_step_n, lymph_time_steps,
 * lymph_step_n_record, lymph_capture_solve) with LYMPH_PROB_RICKER; the
 * factors are computed on the device per step.  n_groups = 0 restores the
 * single lymph_set_ricker wavelet.  No reference counterpart. */
int lymph_set_source_groups(LymphCtx* ctx, int n_groups, const int32_t* poly_group,
                            const double* f0, const double* t0);

/* Receivers: point values of the displacement, evaluated like EXPORT_SOLUTION
 * (src/solution_processing.f90:150-170) in the tet given for each point
 * (e.g. located like FIND_ELEM_FROM_POINT, src/find_tet.f90:82-124, nearest
 * barycentre).  xyz[n*3], tet[n]; room for max_records trace rows. */
int lymph_set_receivers(LymphCtx* ctx, int n, const double* xyz, const int32_t* tet,
                        int max_records);
/* out[n*3]: receiver values of the vector in `slot`. */
int lymph_sample_receivers(LymphCtx* ctx, int slot, double* out);
/* lymph_step_n that also records the receiver values of u^{n+1} after every
 * `every`-th step into the trace buffer (device-resident). */
int lymph_step_n_record(LymphCtx* ctx, int problem, int nsteps, double t0, double dt,
                        int every, double* t_out);
/* out[n_records][n][3] (may be NULL to query the count). */
int lymph_get_traces(LymphCtx* ctx, double* out, int* n_records);

/* Capture one whole solve from rest into a HIP graph on the context stream:
 * zero initial displacement and velocity (projection kernels), first step at
 * t=0, then `nsteps` steps (recording receivers every `every` steps when
 * receivers are set).  The time factors (single Ricker wavelet or the data
 * set's g(t)) are those at capture time; the RHS slots are read at replay.  No reference counterpart (ensembles are new); the
 * captured work is the same kernels the uncaptured calls launch. */
int lymph_capture_solve(LymphCtx* ctx, int problem, double dt, int nsteps, int every);
/* Launch the captured solve (asynchronous; traces overwritten in place).
 * Changing the apply precision, the face-pair storage, the receivers, the
 * source groups or the single Ricker wavelet, or releasing the FP64 blocks,
 * drops the captured solve; lymph_replay then returns LYMPH_ERR_STATE until
 * lymph_capture_solve runs again.  The RHS slots are read when the replay
 * runs, so point sources added or RHS slots rebuilt after the capture take
 * effect in the next replay.  After a replay the state slots hold the
 * captured solve's last two states, whatever uncaptured calls ran between. */
int lymph_replay(LymphCtx* ctx);
/* Wait for all work of the context's stream. */
int lymph_sync(LymphCtx* ctx);

/* COMPUTE_ERROR_L2_MATRIX_FREE / _DG_ (src/matrix_free.f90:1388-1504):
 * squared norms of uh - uex; host pointers or slots. */
int lymph_error_l2(LymphCtx* ctx, const double* uh, const double* uex,
                   double* err2);
int lymph_error_dg(LymphCtx* ctx, const double* uh, const double* uex,
                   double* err2);
int lymph_error_slots(LymphCtx* ctx, int slot_uh, int slot_uex, double* l2sq,
                      double* dgsq);

/* Whole path of src/Lymph3D.f90:262-518 on the device: operator, RHS,
 * IC projections, first step, `while (t <= stop_time)` loop (at most
 * max_steps steps if max_steps >= 0), exact-solution projection scaled by
 * g(t_final), errors.  out[0..3] = t_final, steps, err_L2, err_DG (the
 * square roots printed at src/Lymph3D.f90:516-517); u_final may be NULL. */
int lymph_run(LymphCtx* ctx, int problem, double dt, double stop_time,
              int max_steps, double* u_final, double* out);

/* Instability control: *finite = 1 if slot `slot` holds only finite values,
 * 0 if it holds a NaN or an infinity; synchronous.  lymph_run applies it to
 * u^n every 256 steps and after the last one and then returns
 * LYMPH_EXIT_INSTAB.  The reference parses an instability-control flag
 * (IS_instabilitycontrol, src/Poly_global.f90; exit code EXIT_INSTAB) but
 * never tests the solution. */
int lymph_check_finite(LymphCtx* ctx, int slot, int* finite);

/* Time factor g(t) of a data set: sin(sqrt2 pi t) for LYMPH_PROB_REF_MMS
 * (time_function), cos(sqrt3 pi t) for LYMPH_PROB_CUBE_WAVE. */
double lymph_time_function(int problem, double t);

#ifdef __cplusplus
}
#endif
#endif

! lymph_gpu.f90 -- ISO_C_BINDING interface to liblymph_gpu.so for LYMPH3D's
! Fortran host (declarations of include/lymph_gpu.h, one to one).
!
! A Fortran caller keeps its own mesh set-up and passes the flattened mesh
! (0-based ids, see type lymph_mesh) to lymph_ctx_create; the drop-ins for
! MAKE_MATRICES_FREE, MAKE_RHS_FREE, COMPUTE_MODAL_COEFFICIENTS_FREE,
! FIRST_TIME_STEP_MATRIX_FREE, TIME_STEP_MATRIX_FREE and
! COMPUTE_ERROR_*_MATRIX_FREE (src/matrix_free.f90) are the functions below.
! With lymph_ctx_set_layout(ctx, LYMPH_LAYOUT_FORTRAN) the solution arrays are
! the reference's own u(npoly, 3*Np) column-major arrays, passed as they are.
module lymph_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LYMPH_EXIT_INSTAB = 2
  integer(c_int), parameter :: LYMPH_EXIT_SETUP = 4, LYMPH_EXIT_SINGULARMTX = 5
  integer(c_int), parameter :: LYMPH_EXIT_NEIGHBOUR = 18
  integer(c_int), parameter :: LYMPH_ERR_ARG = 101, LYMPH_ERR_HIP = 102, LYMPH_ERR_STATE = 103
  integer(c_int), parameter :: LYMPH_LAYOUT_POLY = 0, LYMPH_LAYOUT_FORTRAN = 1
  integer(c_int), parameter :: LYMPH_SYM_OFF = 0, LYMPH_SYM_FP64 = 1, LYMPH_SYM_FP64_FP32 = 2
  integer(c_int), parameter :: LYMPH_PROB_NULL = 0, LYMPH_PROB_REF_MMS = 1, LYMPH_PROB_CUBE_WAVE = 2
  integer(c_int), parameter :: LYMPH_FN_ZERO = 0, LYMPH_FN_REF_VEL = 1, LYMPH_FN_REF_UEX = 2
  integer(c_int), parameter :: LYMPH_FN_CUBE_U = 3
  integer(c_int), parameter :: LYMPH_SLOT_RHS_STAT = 0, LYMPH_SLOT_RHS_DYN = 1
  integer(c_int), parameter :: LYMPH_SLOT_U_PREV = 2, LYMPH_SLOT_U_CUR = 3, LYMPH_SLOT_U_NEXT = 4
  integer(c_int), parameter :: LYMPH_SLOT_V0 = 5, LYMPH_SLOT_UEX = 6
  integer(c_int), parameter :: LYMPH_PROB_RICKER = 3
  integer(c_int), parameter :: LYMPH_BUILD_REFERENCE = 0, LYMPH_BUILD_FACTORED = 1

  ! struct LymphMesh
  type, bind(C) :: lymph_mesh
    integer(c_int32_t) :: n_node, n_tet, n_poly
    type(c_ptr) :: xyz, tet_vert, tet_poly, neigh, bc_space, normal, area, bbox, hk
    type(c_ptr) :: rho, lam, mu
  end type lymph_mesh

  interface
    function lymph_last_error() bind(C, name='lymph_last_error') result(s)
      import :: c_ptr
      type(c_ptr) :: s
    end function
    function lymph_device_count() bind(C, name='lymph_device_count') result(n)
      import :: c_int
      integer(c_int) :: n
    end function
    function lymph_ctx_create(mesh, p, alpha, theta, ctx) bind(C, name='lymph_ctx_create') result(rc)
      import :: c_int, c_double, c_ptr, lymph_mesh
      type(lymph_mesh), intent(in) :: mesh
      integer(c_int), value :: p
      real(c_double), value :: alpha, theta
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function
    function lymph_ctx_destroy(ctx) bind(C, name='lymph_ctx_destroy') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function lymph_ctx_set_layout(ctx, layout) bind(C, name='lymph_ctx_set_layout') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: layout
      integer(c_int) :: rc
    end function
    function lymph_build_operator(ctx) bind(C, name='lymph_build_operator') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function lymph_build_rhs(ctx, problem, slot, rhs) bind(C, name='lymph_build_rhs') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: problem, slot
      real(c_double) :: rhs(*)
      integer(c_int) :: rc
    end function
    function lymph_project(ctx, fn, slot, coeff) bind(C, name='lymph_project') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: fn, slot
      real(c_double) :: coeff(*)
      integer(c_int) :: rc
    end function
    function lymph_first_step(ctx, g0, dt, u0, v0, u1) bind(C, name='lymph_first_step') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: g0, dt
      real(c_double), intent(in) :: u0(*), v0(*)
      real(c_double), intent(out) :: u1(*)
      integer(c_int) :: rc
    end function
    function lymph_step(ctx, g, dt, um1, u, up1) bind(C, name='lymph_step') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: g, dt
      real(c_double), intent(in) :: um1(*), u(*)
      real(c_double), intent(out) :: up1(*)
      integer(c_int) :: rc
    end function
    function lymph_error_l2(ctx, uh, uex, err2) bind(C, name='lymph_error_l2') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uh(*), uex(*)
      real(c_double), intent(out) :: err2
      integer(c_int) :: rc
    end function
    function lymph_error_dg(ctx, uh, uex, err2) bind(C, name='lymph_error_dg') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uh(*), uex(*)
      real(c_double), intent(out) :: err2
      integer(c_int) :: rc
    end function
    function lymph_run(ctx, problem, dt, stop_time, max_steps, u_final, out) &
        bind(C, name='lymph_run') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: problem, max_steps
      real(c_double), value :: dt, stop_time
      real(c_double) :: u_final(*)
      real(c_double), intent(out) :: out(4)
      integer(c_int) :: rc
    end function
    function lymph_version() bind(C, name='lymph_version') result(s)
      import :: c_ptr
      type(c_ptr) :: s
    end function
    function lymph_ctx_info(ctx, info) bind(C, name='lymph_ctx_info') result(rc)
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: info(8)
      integer(c_int) :: rc
    end function
    function lymph_ctx_set_stream(ctx, stream) bind(C, name='lymph_ctx_set_stream') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream
      integer(c_int) :: rc
    end function
    function lymph_ctx_get_stream(ctx) bind(C, name='lymph_ctx_get_stream') result(s)
      import :: c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr) :: s
    end function
    function lymph_get_pattern(ctx, row_ptr, col) bind(C, name='lymph_get_pattern') result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: row_ptr(*), col(*)
      integer(c_int) :: rc
    end function
    function lymph_export_operator(ctx, K, A_dg, M, M_modal) bind(C, name='lymph_export_operator') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, K, A_dg, M, M_modal   ! c_loc of the host arrays, or c_null_ptr
      integer(c_int) :: rc
    end function
    function lymph_set_slot(ctx, slot, host) bind(C, name='lymph_set_slot') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      real(c_double), intent(in) :: host(*)
      integer(c_int) :: rc
    end function
    function lymph_get_slot(ctx, slot, host) bind(C, name='lymph_get_slot') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      real(c_double), intent(out) :: host(*)
      integer(c_int) :: rc
    end function
    function lymph_slot_ptr(ctx, slot) bind(C, name='lymph_slot_ptr') result(d)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      type(c_ptr) :: d
    end function
    function lymph_first_step_dev(ctx, g0, dt, u0, v0, u1) bind(C, name='lymph_first_step_dev') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, u0, v0, u1   ! device pointers (lymph_slot_ptr)
      real(c_double), value :: g0, dt
      integer(c_int) :: rc
    end function
    function lymph_step_dev(ctx, g, dt, um1, u, up1) bind(C, name='lymph_step_dev') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, um1, u, up1   ! device pointers (lymph_slot_ptr)
      real(c_double), value :: g, dt
      integer(c_int) :: rc
    end function
    function lymph_step_n(ctx, problem, nsteps, t0, dt, t_out) bind(C, name='lymph_step_n') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: problem, nsteps
      real(c_double), value :: t0, dt
      real(c_double), intent(out) :: t_out
      integer(c_int) :: rc
    end function
    function lymph_time_steps(ctx, problem, nsteps, t0, dt, ms) bind(C, name='lymph_time_steps') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: problem, nsteps
      real(c_double), value :: t0, dt
      real(c_double), intent(out) :: ms
      integer(c_int) :: rc
    end function
    function lymph_autotune(ctx, reps, n_variants, unr, xmode, wpp, ms, chosen) &
        bind(C, name='lymph_autotune') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: reps
      integer(c_int), intent(out) :: n_variants, unr(16), xmode(16), wpp(16), chosen
      real(c_double), intent(out) :: ms(16)
      integer(c_int) :: rc
    end function
    function lymph_set_step_variant(ctx, idx, n_variants) bind(C, name='lymph_set_step_variant') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: idx
      integer(c_int), intent(out) :: n_variants
      integer(c_int) :: rc
    end function
    function lymph_set_apply_precision(ctx, bits) bind(C, name='lymph_set_apply_precision') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: bits
      integer(c_int) :: rc
    end function
    function lymph_ctx_set_keep_adg(ctx, keep) bind(C, name='lymph_ctx_set_keep_adg') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: keep
      integer(c_int) :: rc
    end function
    function lymph_ctx_set_build(ctx, mode) bind(C, name='lymph_ctx_set_build') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      integer(c_int) :: rc
    end function
    function lymph_release_fp64_operator(ctx) bind(C, name='lymph_release_fp64_operator') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function lymph_set_symmetric_storage(ctx, on, asym) &
        bind(C, name='lymph_set_symmetric_storage') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
      real(c_double), intent(out) :: asym
      integer(c_int) :: rc
    end function
    function lymph_add_point_sources(ctx, n, xyz, tet, force, slot) &
        bind(C, name='lymph_add_point_sources') result(rc)
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, slot
      real(c_double), intent(in) :: xyz(*), force(*)
      integer(c_int32_t), intent(in) :: tet(*)
      integer(c_int) :: rc
    end function
    function lymph_add_double_couples(ctx, n, tet, strike, normal, m0, slot) &
        bind(C, name='lymph_add_double_couples') result(rc)
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, slot
      integer(c_int32_t), intent(in) :: tet(*)
      real(c_double), intent(in) :: strike(*), normal(*), m0(*)
      integer(c_int) :: rc
    end function
    function lymph_add_point_sources_noalloc(ctx, n, xyz, tet, force, slot) &
        bind(C, name='lymph_add_point_sources_noalloc') result(rc)
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, slot
      real(c_double), intent(in) :: xyz(*), force(*)
      integer(c_int32_t), intent(in) :: tet(*)
      integer(c_int) :: rc
    end function
    function lymph_set_ricker(ctx, f0, t0) bind(C, name='lymph_set_ricker') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: f0, t0
      integer(c_int) :: rc
    end function
    function lymph_ricker(f0, t0, t) bind(C, name='lymph_ricker') result(g)
      import :: c_double
      real(c_double), value :: f0, t0, t
      real(c_double) :: g
    end function
    function lymph_set_source_groups(ctx, n_groups, poly_group, f0, t0) &
        bind(C, name='lymph_set_source_groups') result(rc)
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_groups
      integer(c_int32_t), intent(in) :: poly_group(*)
      real(c_double), intent(in) :: f0(*), t0(*)
      integer(c_int) :: rc
    end function
    function lymph_set_receivers(ctx, n, xyz, tet, max_records) bind(C, name='lymph_set_receivers') result(rc)
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, max_records
      real(c_double), intent(in) :: xyz(*)
      integer(c_int32_t), intent(in) :: tet(*)
      integer(c_int) :: rc
    end function
    function lymph_sample_receivers(ctx, slot, out) bind(C, name='lymph_sample_receivers') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function
    function lymph_step_n_record(ctx, problem, nsteps, t0, dt, every, t_out) &
        bind(C, name='lymph_step_n_record') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: problem, nsteps, every
      real(c_double), value :: t0, dt
      real(c_double), intent(out) :: t_out
      integer(c_int) :: rc
    end function
    function lymph_get_traces(ctx, out, n_records) bind(C, name='lymph_get_traces') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, out   ! c_loc of the host array, or c_null_ptr for the count
      integer(c_int), intent(out) :: n_records
      integer(c_int) :: rc
    end function
    function lymph_capture_solve(ctx, problem, dt, nsteps, every) bind(C, name='lymph_capture_solve') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: problem, nsteps, every
      real(c_double), value :: dt
      integer(c_int) :: rc
    end function
    function lymph_replay(ctx) bind(C, name='lymph_replay') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function lymph_sync(ctx) bind(C, name='lymph_sync') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function lymph_error_slots(ctx, slot_uh, slot_uex, l2sq, dgsq) bind(C, name='lymph_error_slots') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, l2sq, dgsq   ! c_loc of reals, or c_null_ptr
      integer(c_int), value :: slot_uh, slot_uex
      integer(c_int) :: rc
    end function
    ! f_forcing / f_dirichlet / f_neumann as C functions of n points at once:
    ! the host wraps its own procedures in bind(C) routines of these shapes
    ! and passes c_funloc of them (c_null_funptr = zero data).
    function lymph_build_rhs_fn(ctx, f, gd, gn, user, slot, rhs_out) &
        bind(C, name='lymph_build_rhs_fn') result(rc)
      import :: c_int, c_ptr, c_funptr
      type(c_ptr), value :: ctx, user, rhs_out
      type(c_funptr), value :: f, gd, gn
      integer(c_int), value :: slot
      integer(c_int) :: rc
    end function
    function lymph_project_fn(ctx, fld, user, slot, out) bind(C, name='lymph_project_fn') result(rc)
      import :: c_int, c_ptr, c_funptr
      type(c_ptr), value :: ctx, user, out
      type(c_funptr), value :: fld
      integer(c_int), value :: slot
      integer(c_int) :: rc
    end function
    function lymph_check_finite(ctx, slot, finite) bind(C, name='lymph_check_finite') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      integer(c_int), intent(out) :: finite
      integer(c_int) :: rc
    end function
    function lymph_time_function(problem, t) bind(C, name='lymph_time_function') result(g)
      import :: c_int, c_double
      integer(c_int), value :: problem
      real(c_double), value :: t
      real(c_double) :: g
    end function
  end interface

contains

  ! Message of the last failing call, as a Fortran string.
  function lymph_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    interface
      function c_strlen(s) bind(C, name='strlen') result(n)
        import :: c_ptr, c_size_t
        type(c_ptr), value :: s
        integer(c_size_t) :: n
      end function
    end interface
    type(c_ptr) :: s
    s = lymph_last_error()
    n = int(c_strlen(s))
    call c_f_pointer(s, p, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function

end module lymph_gpu

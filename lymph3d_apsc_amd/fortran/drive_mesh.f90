! drive_mesh.f90 -- the reference's time loop (src/Lymph3D.f90:303-518) with
! the matrix-free routines replaced by the liblymph_gpu drop-ins, called from
! Fortran through module lymph_gpu with the reference's own column-major
! arrays u(npoly, 3*Np).  The flattened mesh comes from a stream file written
! by lymph3d_apsc_amd.mesh.Mesh.write_flat (it plays the part of the
! reference's Mesh_Structure after MAKE_PARTITION_AND_MPI_FILES).
!
! usage: drive_mesh <mesh.bin>
! exit status: 0 ok, 3 no HIP device, 1 any other failure.
program drive_mesh
  use iso_c_binding
  use lymph_gpu
  implicit none
  integer(c_int32_t) :: hdr(5)
  real(c_double) :: dt, stop_time, t, e2l, e2d, half
  real(c_double), allocatable, target :: xyz(:), normal(:), area(:), bbox(:), hk(:)
  real(c_double), allocatable, target :: rho(:), lam(:), mu(:)
  integer(c_int32_t), allocatable, target :: tet_vert(:), tet_poly(:), neigh(:), bc(:)
  real(c_double), allocatable :: u0(:,:), un(:,:), v0(:,:), uex(:,:), rhs_s(:,:), rhs_d(:,:)
  type(lymph_mesh) :: m
  type(c_ptr) :: ctx
  integer(c_int) :: rc, p
  integer :: nn, nt, npoly, np3, nsteps
  character(len=1024) :: path

  call get_command_argument(1, path)
  open(10, file=trim(path), access='stream', form='unformatted', status='old')
  read(10) hdr
  read(10) dt, stop_time
  nn = hdr(1); nt = hdr(2); npoly = hdr(3); p = hdr(4)
  allocate(xyz(3*nn), tet_vert(4*nt), tet_poly(nt), neigh(4*nt), bc(4*nt), normal(12*nt), &
           area(4*nt), bbox(6*npoly), hk(npoly), rho(nt), lam(nt), mu(nt))
  read(10) xyz, tet_vert, tet_poly, neigh, bc, normal, area, bbox, hk, rho, lam, mu
  close(10)
  m%n_node = nn; m%n_tet = nt; m%n_poly = npoly
  m%xyz = c_loc(xyz); m%tet_vert = c_loc(tet_vert); m%tet_poly = c_loc(tet_poly)
  m%neigh = c_loc(neigh); m%bc_space = c_loc(bc); m%normal = c_loc(normal)
  m%area = c_loc(area); m%bbox = c_loc(bbox); m%hk = c_loc(hk)
  m%rho = c_loc(rho); m%lam = c_loc(lam); m%mu = c_loc(mu)

  rc = lymph_ctx_create(m, p, 10.0d0, -1.0d0, ctx)
  if (rc == LYMPH_ERR_HIP .and. lymph_device_count() == 0) then
    print '(A)', 'no HIP device: ' // lymph_error_message()
    call exit(3)
  end if
  call ok(rc, 'lymph_ctx_create')
  call ok(lymph_ctx_set_layout(ctx, LYMPH_LAYOUT_FORTRAN), 'lymph_ctx_set_layout')
  np3 = 3 * (p + 1) * (p + 2) * (p + 3) / 6
  allocate(u0(npoly, np3), un(npoly, np3), v0(npoly, np3), uex(npoly, np3), &
           rhs_s(npoly, np3), rhs_d(npoly, np3))

  ! MAKE_MATRICES_FREE + cholesky (src/Lymph3D.f90:262-271)
  call ok(lymph_build_operator(ctx), 'lymph_build_operator')
  ! MAKE_RHS_FREE static (null data) and dynamic (f_time, gd, gn) (:284-286)
  call ok(lymph_build_rhs(ctx, LYMPH_PROB_NULL, LYMPH_SLOT_RHS_STAT, rhs_s), 'rhs static')
  call ok(lymph_build_rhs(ctx, LYMPH_PROB_REF_MMS, LYMPH_SLOT_RHS_DYN, rhs_d), 'rhs dynamic')
  ! COMPUTE_MODAL_COEFFICIENTS_FREE of the initial conditions (:328-329)
  call ok(lymph_project(ctx, LYMPH_FN_ZERO, LYMPH_SLOT_UEX, u0), 'ic displacement')
  call ok(lymph_project(ctx, LYMPH_FN_REF_VEL, LYMPH_SLOT_UEX, v0), 'ic velocity')
  ! FIRST_TIME_STEP_MATRIX_FREE (:368)
  t = 0.0d0
  nsteps = 1
  call ok(lymph_first_step(ctx, lymph_time_function(LYMPH_PROB_REF_MMS, t), dt, u0, v0, un), &
          'first step')
  t = t + dt
  ! time loop (:389-418): TIME_STEP_MATRIX_FREE, then u0 = un; un = v0
  do while (t <= stop_time)
    call ok(lymph_step(ctx, lymph_time_function(LYMPH_PROB_REF_MMS, t), dt, u0, un, v0), 'step')
    u0 = un
    un = v0
    nsteps = nsteps + 1
    t = t + dt
  end do
  ! exact solution projection scaled by time_function(t) (:459-467), errors (:490-517)
  call ok(lymph_project(ctx, LYMPH_FN_REF_UEX, LYMPH_SLOT_UEX, uex), 'uex')
  half = lymph_time_function(LYMPH_PROB_REF_MMS, t)
  uex = uex * half
  call ok(lymph_error_l2(ctx, un, uex, e2l), 'error L2')
  call ok(lymph_error_dg(ctx, un, uex, e2d), 'error DG')
  print '(A,I0,A,ES24.16)', 'steps ', nsteps, ' t_final ', t
  print '(A,ES24.16)', 'ERROR IN NORM L2: ', sqrt(e2l)
  print '(A,ES24.16)', 'ERROR IN NORM DG: ', sqrt(e2d)
  rc = lymph_ctx_destroy(ctx)

contains

  subroutine ok(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      print '(A,A,I0,A,A)', what, ' failed with code ', code, ': ', lymph_error_message()
      call exit(1)
    end if
  end subroutine

end program drive_mesh

! drive_mesh.f90 -- the reference's time loop (src/Lymph3D.f90:303-518) with
! the matrix-free routines replaced by the liblymph_gpu drop-ins, called from
! Fortran through module lymph_gpu with the reference's own column-major
! arrays u(npoly, 3*Np).  The flattened mesh comes from a stream file written
! by lymph3d_apsc_amd.mesh.Mesh.write_flat (it plays the part of the
! reference's Mesh_Structure after MAKE_PARTITION_AND_MPI_FILES).
!
! usage: drive_mesh <mesh.bin> [procedures | symmetric]
!   With "procedures" the data of MAKE_RHS_FREE and the fields of
!   COMPUTE_MODAL_COEFFICIENTS_FREE come from the Fortran functions of module
!   mms_procedures (the manufactured solution of
!   src/problem_data_and_properties.f90, evaluated on the host) through
!   lymph_build_rhs_fn / lymph_project_fn, as a host passes its own
!   procedure arguments; otherwise from the built-in data set.
! exit status: 0 ok, 3 no HIP device, 1 any other failure.

! The host's pointwise data functions, wrapped as the batched C callbacks of
! lymph_build_rhs_fn / lymph_project_fn (points and values 3 x n).
module mms_procedures
  use iso_c_binding
  implicit none
  real(c_double), parameter :: pi = 3.141592653589793238462643383279502884d0
  real(c_double), parameter :: sqrt2 = 1.414213562373095048801688724209698079d0
contains
  subroutine disp(x, r)  ! uex / gd pattern
    real(c_double), intent(in) :: x(3)
    real(c_double), intent(out) :: r(3)
    r(1) = cos(pi * x(1)) * sin(pi * x(2)) * sin(pi * x(3))
    r(2) = sin(pi * x(1)) * cos(pi * x(2)) * sin(pi * x(3))
    r(3) = sin(pi * x(1)) * sin(pi * x(2)) * cos(pi * x(3))
  end subroutine

  subroutine forcing_c(n, lam, mu, rho, xyz, f, user) bind(C)
    integer(c_int), value :: n
    real(c_double), intent(in) :: lam(n), mu(n), rho(n), xyz(3, n)
    real(c_double), intent(out) :: f(3, n)
    type(c_ptr), value :: user
    integer :: i
    real(c_double) :: c
    do i = 1, n
      c = (-2.0d0 * rho(i) + 3.0d0 * (lam(i) + 2.0d0 * mu(i)))
      f(1, i) = c * pi * pi * cos(pi * xyz(1, i)) * sin(pi * xyz(2, i)) * sin(pi * xyz(3, i))
      f(2, i) = c * pi * pi * sin(pi * xyz(1, i)) * cos(pi * xyz(2, i)) * sin(pi * xyz(3, i))
      f(3, i) = c * pi * pi * sin(pi * xyz(1, i)) * sin(pi * xyz(2, i)) * cos(pi * xyz(3, i))
    end do
  end subroutine

  subroutine dirichlet_c(n, xyz, tag, gd, user) bind(C)
    integer(c_int), value :: n
    real(c_double), intent(in) :: xyz(3, n)
    integer(c_int32_t), intent(in) :: tag(n)
    real(c_double), intent(out) :: gd(3, n)
    type(c_ptr), value :: user
    integer :: i
    do i = 1, n
      gd(:, i) = 0.0d0
      if (tag(i) >= 1 .and. tag(i) <= 3) call disp(xyz(:, i), gd(:, i))
    end do
  end subroutine

  subroutine neumann_c(n, lam, mu, nrm, xyz, tag, gn, user) bind(C)
    integer(c_int), value :: n
    real(c_double), intent(in) :: lam(n), mu(n), nrm(3, n), xyz(3, n)
    integer(c_int32_t), intent(in) :: tag(n)
    real(c_double), intent(out) :: gn(3, n)
    type(c_ptr), value :: user
    integer :: i, j
    real(c_double) :: st(3, 3), x(3)
    do i = 1, n
      gn(:, i) = 0.0d0
      if (tag(i) < 1 .or. tag(i) > 3) cycle
      x = xyz(:, i)
      st(1, 1) = -pi * sin(pi * x(1)) * sin(pi * x(2)) * sin(pi * x(3)) * (3.0d0 * lam(i) + 2.0d0 * mu(i))
      st(2, 2) = st(1, 1)
      st(3, 3) = st(1, 1)
      st(1, 2) = 2.0d0 * mu(i) * pi * cos(pi * x(1)) * cos(pi * x(2)) * sin(pi * x(3))
      st(2, 1) = st(1, 2)
      st(1, 3) = 2.0d0 * mu(i) * pi * cos(pi * x(1)) * cos(pi * x(3)) * sin(pi * x(2))
      st(3, 1) = st(1, 3)
      st(2, 3) = 2.0d0 * mu(i) * pi * cos(pi * x(2)) * cos(pi * x(3)) * sin(pi * x(1))
      st(3, 2) = st(2, 3)
      do j = 1, 3
        gn(j, i) = st(j, 1) * nrm(1, i) + st(j, 2) * nrm(2, i) + st(j, 3) * nrm(3, i)
      end do
    end do
  end subroutine

  subroutine zero_c(n, xyz, v, user) bind(C)
    integer(c_int), value :: n
    real(c_double), intent(in) :: xyz(3, n)
    real(c_double), intent(out) :: v(3, n)
    type(c_ptr), value :: user
    v = 0.0d0
  end subroutine

  subroutine velocity_c(n, xyz, v, user) bind(C)  ! ic_velocity
    integer(c_int), value :: n
    real(c_double), intent(in) :: xyz(3, n)
    real(c_double), intent(out) :: v(3, n)
    type(c_ptr), value :: user
    integer :: i
    do i = 1, n
      call disp(xyz(:, i), v(:, i))
      v(:, i) = sqrt2 * pi * v(:, i)
    end do
  end subroutine

  subroutine uex_c(n, xyz, v, user) bind(C)
    integer(c_int), value :: n
    real(c_double), intent(in) :: xyz(3, n)
    real(c_double), intent(out) :: v(3, n)
    type(c_ptr), value :: user
    integer :: i
    do i = 1, n
      call disp(xyz(:, i), v(:, i))
    end do
  end subroutine
end module mms_procedures

program drive_mesh
  use iso_c_binding
  use lymph_gpu
  use mms_procedures
  implicit none
  integer(c_int32_t) :: hdr(5)
  real(c_double) :: dt, stop_time, t, e2l, e2d, half
  real(c_double), allocatable, target :: xyz(:), normal(:), area(:), bbox(:), hk(:)
  real(c_double), allocatable, target :: rho(:), lam(:), mu(:)
  integer(c_int32_t), allocatable, target :: tet_vert(:), tet_poly(:), neigh(:), bc(:)
  real(c_double), allocatable, target :: u0(:,:), un(:,:), v0(:,:), uex(:,:), rhs_s(:,:), rhs_d(:,:)
  type(lymph_mesh) :: m
  type(c_ptr) :: ctx
  integer(c_int) :: rc, p
  integer :: nn, nt, npoly, np3, nsteps
  character(len=1024) :: path, mode
  logical :: procs, sym
  real(c_double) :: asym

  call get_command_argument(1, path)
  call get_command_argument(2, mode)
  procs = trim(mode) == 'procedures'
  sym = trim(mode) == 'symmetric'
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
  ! optionally keep each face pair's block once (SIPG, one material per face)
  if (sym) then
    call ok(lymph_set_symmetric_storage(ctx, 1, asym), 'lymph_set_symmetric_storage')
    print '(A,ES10.3)', ' face pairs stored once, largest pair asymmetry ', asym
  end if
  ! MAKE_RHS_FREE static (null data) and dynamic (f_time, gd, gn) (:284-286)
  ! COMPUTE_MODAL_COEFFICIENTS_FREE of the initial conditions (:328-329)
  if (procs) then
    call ok(lymph_build_rhs_fn(ctx, c_null_funptr, c_null_funptr, c_null_funptr, c_null_ptr, &
                               LYMPH_SLOT_RHS_STAT, c_loc(rhs_s)), 'rhs static')
    call ok(lymph_build_rhs_fn(ctx, c_funloc(forcing_c), c_funloc(dirichlet_c), c_funloc(neumann_c), &
                               c_null_ptr, LYMPH_SLOT_RHS_DYN, c_loc(rhs_d)), 'rhs dynamic')
    call ok(lymph_project_fn(ctx, c_funloc(zero_c), c_null_ptr, LYMPH_SLOT_UEX, c_loc(u0)), &
            'ic displacement')
    call ok(lymph_project_fn(ctx, c_funloc(velocity_c), c_null_ptr, LYMPH_SLOT_UEX, c_loc(v0)), &
            'ic velocity')
  else
    call ok(lymph_build_rhs(ctx, LYMPH_PROB_NULL, LYMPH_SLOT_RHS_STAT, rhs_s), 'rhs static')
    call ok(lymph_build_rhs(ctx, LYMPH_PROB_REF_MMS, LYMPH_SLOT_RHS_DYN, rhs_d), 'rhs dynamic')
    call ok(lymph_project(ctx, LYMPH_FN_ZERO, LYMPH_SLOT_UEX, u0), 'ic displacement')
    call ok(lymph_project(ctx, LYMPH_FN_REF_VEL, LYMPH_SLOT_UEX, v0), 'ic velocity')
  end if
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
  if (procs) then
    call ok(lymph_project_fn(ctx, c_funloc(uex_c), c_null_ptr, LYMPH_SLOT_UEX, c_loc(uex)), 'uex')
  else
    call ok(lymph_project(ctx, LYMPH_FN_REF_UEX, LYMPH_SLOT_UEX, uex), 'uex')
  end if
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

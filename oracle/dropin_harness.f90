! dropin_harness.f90 -- TEST INFRASTRUCTURE ONLY.  Drives the reference-named
! Fortran drop-in (lymph3d_apsc_amd/fortran/lymph_matrix_free.f90, module
! matrix_free_gpu) the way src/Lymph3D.f90 does, with the reference's own
! modules for the data types (Poly_global, Poly_mesh, Poly_data), the host
! Cholesky (vet_mat_operations) and the data functions
! (problem_data_and_properties), compiled from the reference sources where
! they lie (INCLUDE lines; nothing is copied).  Built by `make -C oracle
! dropin` into oracle/_ref/.
!
! It fills one rank's Mesh_Structure and Data_Structure from a stream file
! holding the fields MAKE_PARTITION_AND_MPI_FILES and READ_INPUT_FILES would
! set (written by tests/dropin_files.py), then
!   dropin_harness flatten <in> <out>  writes what LYMPH_FLATTEN_MESH makes of
!                                      them (for the CPU test);
!   dropin_harness rhs <in> <out>      writes rhs_stat_loc, rhs_dyn_loc and M_loc
!                                      after MAKE_MATRICES_FREE and both
!                                      MAKE_RHS_FREE calls;
!   dropin_harness run <in>            runs src/Lymph3D.f90:262-518 (matrices,
!                                      host Cholesky, RHS, initial conditions,
!                                      first step, the t <= stop_time loop,
!                                      exact-solution projection, errors)
!                                      through the drop-ins and prints the
!                                      steps and both errors.
! Exit status 3 when no HIP device is visible.
include '/root/reference/src/Poly_global.f90'
include '/root/reference/src/Poly_mesh.f90'
include '/root/reference/src/Poly_data.f90'
include '/root/reference/src/vet_mat_operations.f90'
include '/root/reference/src/problem_data_and_properties.f90'
include '../lymph3d_apsc_amd/fortran/lymph_gpu.f90'
include '../lymph3d_apsc_amd/fortran/lymph_matrix_free.f90'

program dropin_harness
  use iso_c_binding
  use Poly_global
  use global_parameters, only: DIM
  use Poly_mesh
  use Poly_data
  use vet_mat_operations, only: cholesky
  use problem_data_and_properties
  use lymph_gpu, only: lymph_device_count
  use matrix_free_gpu
  implicit none
  type(Mesh_Structure) :: PolyMesh
  type(Data_Structure) :: PolyData
  type(lymph_flat_mesh) :: fm
  character(len=1024) :: mode, fin, fout
  integer(c_int32_t) :: hdr(6)
  integer :: nn, ne, npoly, p, nmat, nsism, ie, ip, iface, Np, i, num_dt, n_neigh
  integer(c_int32_t), allocatable :: con(:,:), nel(:,:,:), eip(:), matp(:)
  real(c_double), allocatable :: cx(:), cy(:), cz(:), nrm(:,:,:), ar(:,:), bb(:,:,:), hkk(:)
  real(kind=8), allocatable :: K_loc(:,:,:,:), A_dg_loc(:,:,:,:), M_loc(:,:,:,:), M_modal_loc(:,:,:,:)
  real(kind=8), allocatable :: R_M_loc(:,:,:,:), R_M_modal_loc(:,:,:,:), M_tmp(:,:)
  real(kind=8), allocatable :: rhs_stat_loc(:,:), rhs_dyn_loc(:,:), u0_loc(:,:), un_loc(:,:), v0_loc(:,:)
  real(kind=8), allocatable :: un_mpi(:,:), uex_mpi(:,:)
  real(kind=8) :: t, err_L2, err_DG

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  nn = hdr(1); ne = hdr(2); npoly = hdr(3); p = hdr(4); nmat = hdr(5); nsism = hdr(6)
  read(10) time_step, stop_time
  allocate(cx(nn), cy(nn), cz(nn), con(4, ne), nel(4, 0:5, ne), nrm(4, 3, ne), ar(4, ne), &
           eip(ne), matp(ne), bb(3, 2, npoly), hkk(npoly))
  read(10) cx, cy, cz, con, nel, nrm, ar, eip, matp, bb, hkk
  ! Data_Structure: materials (rho, lambda, mu) and SISM lines
  PolyData%nmat = nmat
  allocate(PolyData%prop_mat(nmat, 3))
  read(10) PolyData%prop_mat
  PolyData%nload_sism_el = nsism
  allocate(PolyData%val_sism_el(max(nsism, 1), 21))
  if (nsism > 0) read(10) PolyData%val_sism_el
  close(10)

  ! Mesh_Structure of the only rank: local numbering = global numbering
  PolyMesh%num_node = nn; PolyMesh%num_node_loc = nn
  PolyMesh%num_tet = ne; PolyMesh%num_elem = ne; PolyMesh%num_elem_loc = ne
  PolyMesh%num_poly = npoly; PolyMesh%num_poly_loc = npoly
  allocate(PolyMesh%coord_x(nn), PolyMesh%coord_y(nn), PolyMesh%coord_z(nn))
  PolyMesh%coord_x = cx; PolyMesh%coord_y = cy; PolyMesh%coord_z = cz
  allocate(PolyMesh%node_loc2glo(nn), PolyMesh%elem_loc2glo(ne), PolyMesh%elem_glo2loc(ne), &
           PolyMesh%poly_loc2glo(npoly), PolyMesh%elem_in_poly(ne))
  PolyMesh%node_loc2glo = [(i, i = 1, nn)]
  PolyMesh%elem_loc2glo = [(i, i = 1, ne)]
  PolyMesh%elem_glo2loc = [(i, i = 1, ne)]
  PolyMesh%poly_loc2glo = [(i, i = 1, npoly)]
  PolyMesh%elem_in_poly = eip
  allocate(PolyMesh%Elem_loc(ne), PolyMesh%Poly(npoly))
  do ie = 1, ne
    PolyMesh%Elem_loc(ie)%el_type = 'TET'
    PolyMesh%Elem_loc(ie)%num_vert = 4
    PolyMesh%Elem_loc(ie)%num_faces = 4
    PolyMesh%Elem_loc(ie)%Degree = p
    PolyMesh%Elem_loc(ie)%mat_prop = matp(ie)
    allocate(PolyMesh%Elem_loc(ie)%vert(4), PolyMesh%Elem_loc(ie)%neigh_el(4, 0:5), &
             PolyMesh%Elem_loc(ie)%normal(4, 3), PolyMesh%Elem_loc(ie)%area(4))
    PolyMesh%Elem_loc(ie)%vert = con(:, ie)
    do iface = 1, 4
      PolyMesh%Elem_loc(ie)%neigh_el(iface, 0:5) = nel(iface, 0:5, ie)
    end do
    PolyMesh%Elem_loc(ie)%normal = nrm(:, :, ie)
    PolyMesh%Elem_loc(ie)%area = ar(:, ie)
  end do
  do ip = 1, npoly
    PolyMesh%Poly(ip)%b_box = bb(:, :, ip)
    PolyMesh%Poly(ip)%hk = hkk(ip)
  end do

  if (trim(mode) == 'flatten') then
    call LYMPH_FLATTEN_MESH(PolyMesh, PolyData, fm)
    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) fm%n_node, fm%n_tet, fm%n_poly
    write(11) fm%xyz, fm%tet_vert, fm%tet_poly, fm%neigh, fm%bc_space, fm%normal, fm%area, &
              fm%bbox, fm%hk, fm%rho, fm%lam, fm%mu
    close(11)
    stop
  end if

  if (lymph_device_count() == 0) then
    print '(A)', 'no HIP device'
    call exit(3)
  end if
  ! src/Lymph3D.f90:219-252: the arrays of the time loop
  Np = (p + 1) * (p + 2) * (p + 3) / 6
  allocate(K_loc(ne, 5, DIM * Np, DIM * Np), A_dg_loc(ne, 5, DIM * Np, DIM * Np), &
           M_loc(ne, DIM, Np, Np), M_modal_loc(ne, DIM, Np, Np), R_M_loc(ne, DIM, Np, Np), &
           R_M_modal_loc(ne, DIM, Np, Np), M_tmp(Np, Np))
  allocate(rhs_stat_loc(ne, DIM * Np), rhs_dyn_loc(ne, DIM * Np), u0_loc(ne, DIM * Np), &
           un_loc(ne, DIM * Np), v0_loc(ne, DIM * Np), un_mpi(1, DIM * Np), uex_mpi(1, DIM * Np))
  rhs_stat_loc = 0.0d0; rhs_dyn_loc = 0.0d0; un_loc = 0.0d0; un_mpi = 0.0d0; uex_mpi = 0.0d0
  ! :262-271
  call MAKE_MATRICES_FREE(PolyMesh, PolyData, ne, Np, K_loc, A_dg_loc, M_loc, M_modal_loc, n_neigh)
  do ie = 1, ne
    do i = 1, DIM
      M_tmp = M_loc(ie, i, :, :)
      R_M_loc(ie, i, :, :) = cholesky(M_tmp, Np)
      M_tmp = M_modal_loc(ie, i, :, :)
      R_M_modal_loc(ie, i, :, :) = cholesky(M_tmp, Np)
    end do
  end do
  ! :284-286
  call MAKE_RHS_FREE(PolyMesh, PolyData, ne, Np, rhs_stat_loc, f_null, gd_null, gn_null)
  call MAKE_RHS_FREE(PolyMesh, PolyData, ne, Np, rhs_dyn_loc, f_time, gd, gn)
  if (trim(mode) == 'rhs') then  ! both RHS arrays and M_loc, as the host holds them
    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) rhs_stat_loc, rhs_dyn_loc, M_loc
    close(11)
    call lymph_dropin_release()
    stop
  end if
  ! :303-316
  t = 0.0d0
  num_dt = 0
  dt2 = time_step * time_step
  half_dt2 = 0.5d0 * dt2
  ! :328-329
  call COMPUTE_MODAL_COEFFICIENTS_FREE(PolyMesh, Np, R_M_modal_loc, ic_displacement, u0_loc)
  call COMPUTE_MODAL_COEFFICIENTS_FREE(PolyMesh, Np, R_M_modal_loc, ic_velocity, v0_loc)
  num_dt = num_dt + 1
  ! :368-381
  call FIRST_TIME_STEP_MATRIX_FREE(PolyMesh, Np, t, K_loc, R_M_loc, rhs_stat_loc, rhs_dyn_loc, &
                                   u0_loc, v0_loc, un_loc, un_mpi)
  num_dt = num_dt + 1
  t = t + time_step
  ! :389-420
  do while (t <= stop_time)
    call TIME_STEP_MATRIX_FREE(PolyMesh, Np, t, K_loc, R_M_loc, rhs_stat_loc, rhs_dyn_loc, &
                               u0_loc, un_loc, v0_loc, un_mpi)
    u0_loc = un_loc
    un_loc = v0_loc
    num_dt = num_dt + 1
    t = t + time_step
  end do
  num_dt = num_dt - 1
  ! :459-467
  call COMPUTE_MODAL_COEFFICIENTS_FREE(PolyMesh, Np, R_M_modal_loc, uex, u0_loc)
  call COMPUTE_MODAL_COEFFICIENTS_FREE(PolyMesh, Np, R_M_modal_loc, uex_stat, v0_loc)
  u0_loc = v0_loc * 0.0d0 + u0_loc * time_function(t)
  ! :490-517
  call COMPUTE_ERROR_L2_MATRIX_FREE(PolyMesh, Np, M_modal_loc, un_loc, u0_loc, err_L2)
  call COMPUTE_ERROR_DG_MATRIX_FREE(PolyMesh, Np, A_dg_loc, un_loc, u0_loc, un_mpi, uex_mpi, err_DG)
  err_L2 = dsqrt(err_L2)
  err_DG = dsqrt(err_DG)
  print '(A,I0,A,ES24.16)', 'steps ', num_dt, ' t_final ', t
  print '(A,ES24.16)', 'ERROR IN NORM L2: ', err_L2
  print '(A,ES24.16)', 'ERROR IN NORM DG: ', err_DG
  call lymph_dropin_release()
end program dropin_harness

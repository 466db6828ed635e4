! lymph_matrix_free.f90 -- the matrix-free routines of the reference
! (src/matrix_free.f90) under their own names and argument lists, running on
! one GPU through liblymph_gpu (module lymph_gpu):
!   MAKE_MATRICES_FREE               src/matrix_free.f90:27
!   MAKE_RHS_FREE                    :467  (procedure arguments and SISM sources)
!   COMPUTE_MODAL_COEFFICIENTS_FREE  :847  (procedure argument)
!   FIRST_TIME_STEP_MATRIX_FREE      :1100
!   TIME_STEP_MATRIX_FREE            :1230
!   COMPUTE_ERROR_L2_MATRIX_FREE     :1388
!   COMPUTE_ERROR_DG_MATRIX_FREE     :1421
! It is compiled inside the reference's own build: it uses the reference's
! modules Poly_global (time_step), global_parameters (DIM), Poly_mesh
! (Mesh_Structure), Poly_data (Data_Structure) and problem_data_and_properties
! (time_function, set_properties).  src/Lymph3D.f90 then takes these seven
! routines from here and keeps PREPROCESS/POST_PROCESS_SOLUTION_MATRIX_FREE
! from module matrix_free:
!     use matrix_free, only: PREPROCESS_SOLUTION_MATRIX_FREE, POST_PROCESS_MATRIX_FREE
!     use matrix_free_gpu
!
! One rank, one GPU: the whole local mesh lives on the device and the halo
! arguments (un_mpi, uh_mpi, ...) are not read; the operator and mass factors
! stay on the device, so K_loc, A_dg_loc and R_M_loc passed to the steps are
! not read either.  MAKE_MATRICES_FREE still fills the host arrays it
! returns in the reference (M_loc and M_modal_loc always, which the host
! factors; K_loc and A_dg_loc when each polyhedron is one tet, the only case
! with the reference's (tet, 1 + face) block layout).  Vectors are the
! reference's u(num_elem_loc, DIM*Np) arrays, one row per polyhedron; every
! call moves the vectors it is given between host and device.  The context is
! created by MAKE_MATRICES_FREE; lymph_dropin_release frees it.
module matrix_free_gpu
  use iso_c_binding
  use Poly_global, only: time_step
  use global_parameters, only: DIM
  use Poly_mesh, only: Mesh_Structure
  use Poly_data, only: Data_Structure
  use problem_data_and_properties, only: time_function, set_properties
  use lymph_gpu
  implicit none
  private
  public :: MAKE_MATRICES_FREE, MAKE_RHS_FREE, COMPUTE_MODAL_COEFFICIENTS_FREE, &
            FIRST_TIME_STEP_MATRIX_FREE, TIME_STEP_MATRIX_FREE, &
            COMPUTE_ERROR_L2_MATRIX_FREE, COMPUTE_ERROR_DG_MATRIX_FREE, &
            LYMPH_FLATTEN_MESH, lymph_flat_mesh, lymph_dropin_context, lymph_dropin_release, &
            lymph_dropin_failed

  !> struct LymphMesh's arrays, 0-based ids (see include/lymph_gpu.h)
  type :: lymph_flat_mesh
    integer(c_int32_t) :: n_node = 0, n_tet = 0, n_poly = 0
    real(c_double), allocatable :: xyz(:), normal(:), area(:), bbox(:), hk(:)
    real(c_double), allocatable :: rho(:), lam(:), mu(:)
    integer(c_int32_t), allocatable :: tet_vert(:), tet_poly(:), neigh(:), bc_space(:)
  end type lymph_flat_mesh

  abstract interface
    function forcing_if(lambda, mu, point, rho) result(res)
      import :: DIM
      real(kind=8) :: rho, lambda, mu
      real(kind=8), dimension(DIM) :: point, res
    end function forcing_if
    function dirichlet_if(point, space_fun_tag) result(res)
      import :: DIM
      real(kind=8), dimension(DIM) :: point, res
      integer(kind=4) :: space_fun_tag
    end function dirichlet_if
    function neumann_if(lambda, mu, normal, point, space_fun_tag) result(res)
      import :: DIM
      real(kind=8) :: lambda, mu
      real(kind=8), dimension(DIM) :: point, res, normal
      integer(kind=4) :: space_fun_tag
    end function neumann_if
    function analytic_if(point) result(res)
      import :: DIM
      real(kind=8), dimension(DIM) :: point, res
    end function analytic_if
  end interface

  type(c_ptr), save :: ctx = c_null_ptr
  type(lymph_flat_mesh), target, save :: flat
  ! the procedure arguments of the call in progress, for the C callbacks
  procedure(forcing_if), pointer, save :: pf => null()
  procedure(dirichlet_if), pointer, save :: pd => null()
  procedure(neumann_if), pointer, save :: pn => null()
  procedure(analytic_if), pointer, save :: pa => null()
  integer(c_int), save :: last_rc = 0

contains

  !> The context the drop-ins run on (null before MAKE_MATRICES_FREE).
  function lymph_dropin_context() result(c)
    type(c_ptr) :: c
    c = ctx
  end function

  !> 0, or the code of the last failing library call of a drop-in.
  function lymph_dropin_failed() result(rc)
    integer(c_int) :: rc
    rc = last_rc
  end function

  subroutine lymph_dropin_release()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = lymph_ctx_destroy(ctx)
    ctx = c_null_ptr
  end subroutine

  ! A failing call stops the run with the library's message, as the
  ! reference stops on its own errors.
  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    last_rc = rc
    if (rc /= 0) then
      print '(A,A,I0,A,A)', what, ' failed with code ', rc, ': ', lymph_error_message()
      error stop 1
    end if
  end subroutine

  !> Mesh_Structure (src/Poly_mesh.f90:13-104) of one rank -> LymphMesh:
  !> vertices by local node position (node_loc2glo), tets by local id, the
  !> polyhedron of each tet through elem_loc2glo / elem_in_poly /
  !> poly_loc2glo, neighbours from neigh_el(face, 2) (global id, or -1
  !> Dirichlet / -2 Neumann) through elem_glo2loc, space-function tags from
  !> neigh_el(face, 5), normals and areas per face, b_box and hk per
  !> polyhedron, rho / lambda / mu from PolyData%prop_mat(mat_prop, 1:3).
  subroutine LYMPH_FLATTEN_MESH(PolyMesh, PolyData, fm)
    type(Mesh_Structure), intent(in) :: PolyMesh
    type(Data_Structure), intent(in) :: PolyData
    type(lymph_flat_mesh), intent(out) :: fm
    integer :: nn, ne, np, ie, iv, f, d, ip, e2, gmax, mat
    integer, allocatable :: node_g2l(:), poly_g2l(:)

    nn = PolyMesh%num_node_loc
    ne = PolyMesh%num_elem_loc
    np = PolyMesh%num_poly_loc
    fm%n_node = nn
    fm%n_tet = ne
    fm%n_poly = np
    allocate(fm%xyz(3 * nn), fm%tet_vert(4 * ne), fm%tet_poly(ne), fm%neigh(4 * ne), &
             fm%bc_space(4 * ne), fm%normal(12 * ne), fm%area(4 * ne), fm%bbox(6 * np), &
             fm%hk(np), fm%rho(ne), fm%lam(ne), fm%mu(ne))
    ! global node id -> local position (FIND_POS_LOC_NODE's map)
    gmax = maxval(PolyMesh%node_loc2glo(1:nn))
    allocate(node_g2l(gmax))
    node_g2l = 0
    do iv = 1, nn
      node_g2l(PolyMesh%node_loc2glo(iv)) = iv
      fm%xyz(3 * (iv - 1) + 1) = PolyMesh%coord_x(iv)
      fm%xyz(3 * (iv - 1) + 2) = PolyMesh%coord_y(iv)
      fm%xyz(3 * (iv - 1) + 3) = PolyMesh%coord_z(iv)
    end do
    ! global polyhedron id -> local (GET_EL_LOC_FROM_EL_GLO's map)
    gmax = maxval(PolyMesh%poly_loc2glo(1:np))
    allocate(poly_g2l(gmax))
    poly_g2l = 0
    do ip = 1, np
      poly_g2l(PolyMesh%poly_loc2glo(ip)) = ip
      do d = 1, 3
        fm%bbox(6 * (ip - 1) + 2 * d - 1) = PolyMesh%Poly(ip)%b_box(d, 1)
        fm%bbox(6 * (ip - 1) + 2 * d) = PolyMesh%Poly(ip)%b_box(d, 2)
      end do
      fm%hk(ip) = PolyMesh%Poly(ip)%hk
    end do
    do ie = 1, ne
      do iv = 1, 4
        fm%tet_vert(4 * (ie - 1) + iv) = node_g2l(PolyMesh%Elem_loc(ie)%vert(iv)) - 1
      end do
      fm%tet_poly(ie) = poly_g2l(PolyMesh%elem_in_poly(PolyMesh%elem_loc2glo(ie))) - 1
      do f = 1, 4
        e2 = PolyMesh%Elem_loc(ie)%neigh_el(f, 2)
        if (e2 >= 0) e2 = PolyMesh%elem_glo2loc(e2) - 1
        fm%neigh(4 * (ie - 1) + f) = e2
        fm%bc_space(4 * (ie - 1) + f) = PolyMesh%Elem_loc(ie)%neigh_el(f, 5)
        fm%area(4 * (ie - 1) + f) = PolyMesh%Elem_loc(ie)%area(f)
        do d = 1, 3
          fm%normal(12 * (ie - 1) + 3 * (f - 1) + d) = PolyMesh%Elem_loc(ie)%normal(f, d)
        end do
      end do
      mat = PolyMesh%Elem_loc(ie)%mat_prop
      fm%rho(ie) = PolyData%prop_mat(mat, 1)
      fm%lam(ie) = PolyData%prop_mat(mat, 2)
      fm%mu(ie) = PolyData%prop_mat(mat, 3)
    end do
  end subroutine LYMPH_FLATTEN_MESH

  !> MAKE_MATRICES_FREE (src/matrix_free.f90:27-464) on the device; see the
  !> module header for which host arrays it fills.
  subroutine MAKE_MATRICES_FREE(PolyMesh, PolyData, num_elem_loc, Np, K_loc, A_dg_loc, M_loc, &
                                M_modal_loc, max_faces)
    type(Mesh_Structure), intent(inout) :: PolyMesh
    type(Data_Structure), intent(in) :: PolyData
    integer(kind=4), intent(in) :: Np, num_elem_loc
    real(kind=8), dimension(:,:,:,:), intent(out) :: K_loc, A_dg_loc, M_loc, M_modal_loc
    integer(kind=4), intent(out) :: max_faces
    type(lymph_mesh) :: m
    real(kind=8) :: alpha, theta, c
    real(c_double), allocatable, target :: Mb(:), Mmb(:), Kb(:), Ab(:)
    integer(c_int32_t), allocatable :: rp(:), col(:)
    integer(c_int64_t) :: info(8)
    integer :: np_, n3, nblk, ip, i, r, cc, b, f, q, ie

    call lymph_dropin_release()
    call LYMPH_FLATTEN_MESH(PolyMesh, PolyData, flat)
    m%n_node = flat%n_node
    m%n_tet = flat%n_tet
    m%n_poly = flat%n_poly
    m%xyz = c_loc(flat%xyz)
    m%tet_vert = c_loc(flat%tet_vert)
    m%tet_poly = c_loc(flat%tet_poly)
    m%neigh = c_loc(flat%neigh)
    m%bc_space = c_loc(flat%bc_space)
    m%normal = c_loc(flat%normal)
    m%area = c_loc(flat%area)
    m%bbox = c_loc(flat%bbox)
    m%hk = c_loc(flat%hk)
    m%rho = c_loc(flat%rho)
    m%lam = c_loc(flat%lam)
    m%mu = c_loc(flat%mu)
    call set_properties(alpha, theta, c)
    call check(lymph_ctx_create(m, PolyMesh%Elem_loc(1)%Degree, alpha, theta, ctx), 'lymph_ctx_create')
    call check(lymph_ctx_set_layout(ctx, LYMPH_LAYOUT_FORTRAN), 'lymph_ctx_set_layout')
    call check(lymph_build_operator(ctx), 'lymph_build_operator')
    max_faces = PolyMesh%Elem_loc(1)%num_faces
    np_ = flat%n_poly
    n3 = DIM * Np
    ! mass blocks, one per polyhedron and the same for each component
    allocate(Mb(np_ * Np * Np), Mmb(np_ * Np * Np))
    call check(lymph_export_operator(ctx, c_null_ptr, c_null_ptr, c_loc(Mb), c_loc(Mmb)), &
               'lymph_export_operator')
    M_loc = 0.0d0
    M_modal_loc = 0.0d0
    do ip = 1, min(np_, size(M_loc, 1))
      do i = 1, DIM
        do r = 1, Np
          do cc = 1, Np
            M_loc(ip, i, r, cc) = Mb((ip - 1) * Np * Np + (r - 1) * Np + cc)
            M_modal_loc(ip, i, r, cc) = Mmb((ip - 1) * Np * Np + (r - 1) * Np + cc)
          end do
        end do
      end do
    end do
    K_loc = 0.0d0
    A_dg_loc = 0.0d0
    if (np_ /= flat%n_tet) return  ! agglomerated: no (tet, 1 + face) layout
    call check(lymph_ctx_info(ctx, info), 'lymph_ctx_info')
    nblk = int(info(4))
    allocate(rp(np_ + 1), col(nblk), Kb(int(nblk, 8) * n3 * n3), Ab(int(nblk, 8) * n3 * n3))
    call check(lymph_get_pattern(ctx, rp, col), 'lymph_get_pattern')
    call check(lymph_export_operator(ctx, c_loc(Kb), c_loc(Ab), c_null_ptr, c_null_ptr), &
               'lymph_export_operator')
    do ie = 1, np_
      do b = rp(ie) + 1, rp(ie + 1)
        q = 1  ! self block: slot 1; a neighbour: 1 + the face it lies across
        if (col(b) /= ie - 1) then
          do f = 1, 4
            if (flat%neigh(4 * (ie - 1) + f) == col(b)) q = 1 + f
          end do
        end if
        do r = 1, n3
          do cc = 1, n3
            K_loc(ie, q, r, cc) = Kb(int(b - 1, 8) * n3 * n3 + (r - 1) * n3 + cc)
            A_dg_loc(ie, q, r, cc) = Ab(int(b - 1, 8) * n3 * n3 + (r - 1) * n3 + cc)
          end do
        end do
      end do
    end do
  end subroutine MAKE_MATRICES_FREE

  ! C callbacks of lymph_build_rhs_fn / lymph_project_fn: the procedure
  ! arguments of the call in progress, point by point
  subroutine tramp_forcing(n, lam, mu, rho, xyz, f, user) bind(C)
    integer(c_int), value :: n
    real(c_double), intent(in) :: lam(n), mu(n), rho(n), xyz(3, n)
    real(c_double), intent(out) :: f(3, n)
    type(c_ptr), value :: user
    real(kind=8) :: l, m, r, x(DIM)
    integer :: i
    do i = 1, n
      l = lam(i); m = mu(i); r = rho(i); x = xyz(:, i)
      f(:, i) = pf(l, m, x, r)
    end do
  end subroutine

  subroutine tramp_dirichlet(n, xyz, tag, gd, user) bind(C)
    integer(c_int), value :: n
    real(c_double), intent(in) :: xyz(3, n)
    integer(c_int32_t), intent(in) :: tag(n)
    real(c_double), intent(out) :: gd(3, n)
    type(c_ptr), value :: user
    real(kind=8) :: x(DIM)
    integer(kind=4) :: tg
    integer :: i
    do i = 1, n
      x = xyz(:, i); tg = tag(i)
      gd(:, i) = pd(x, tg)
    end do
  end subroutine

  subroutine tramp_neumann(n, lam, mu, nrm, xyz, tag, gn, user) bind(C)
    integer(c_int), value :: n
    real(c_double), intent(in) :: lam(n), mu(n), nrm(3, n), xyz(3, n)
    integer(c_int32_t), intent(in) :: tag(n)
    real(c_double), intent(out) :: gn(3, n)
    type(c_ptr), value :: user
    real(kind=8) :: l, m, x(DIM), nn(DIM)
    integer(kind=4) :: tg
    integer :: i
    do i = 1, n
      l = lam(i); m = mu(i); x = xyz(:, i); nn = nrm(:, i); tg = tag(i)
      gn(:, i) = pn(l, m, nn, x, tg)
    end do
  end subroutine

  subroutine tramp_analytic(n, xyz, v, user) bind(C)
    integer(c_int), value :: n
    real(c_double), intent(in) :: xyz(3, n)
    real(c_double), intent(out) :: v(3, n)
    type(c_ptr), value :: user
    real(kind=8) :: x(DIM)
    integer :: i
    do i = 1, n
      x = xyz(:, i)
      v(:, i) = pa(x)
    end do
  end subroutine

  !> Nearest tet barycentre, first minimum, as FIND_ELEM_FROM_POINT
  !> (src/find_tet.f90:82-124) on one rank; 0-based.
  function locate_point(point) result(tet)
    real(kind=8), intent(in) :: point(DIM)
    integer(c_int32_t) :: tet
    real(kind=8) :: b(DIM), d2, dmin
    integer :: ie, iv, d
    dmin = 1.0d15
    tet = -1
    do ie = 1, flat%n_tet
      do d = 1, DIM
        b(d) = 0.0d0
        do iv = 1, 4
          b(d) = b(d) + flat%xyz(3 * flat%tet_vert(4 * (ie - 1) + iv) + d)
        end do
        b(d) = b(d) / 4.0d0
      end do
      d2 = (point(1) - b(1))**2 + (point(2) - b(2))**2 + (point(3) - b(3))**2
      if (d2 < dmin) then
        dmin = d2
        tet = ie - 1
      end if
    end do
  end function

  !> MAKE_RHS_FREE (src/matrix_free.f90:467-843): the procedure arguments
  !> evaluated on the host at the points the device returns, then the SISM
  !> double couples (each located like FIND_ELEM_FROM_POINT, with the strike,
  !> normal and moment of the first SISM line as the reference takes them,
  !> :648-653).
  subroutine MAKE_RHS_FREE(PolyMesh, PolyData, num_elem_loc, Np, rhs_loc, f_forcing, f_dirichlet, &
                           f_neumann)
    type(Mesh_Structure), intent(inout) :: PolyMesh
    type(Data_Structure), intent(in) :: PolyData
    integer(kind=4), intent(in) :: Np, num_elem_loc
    real(kind=8), dimension(PolyMesh%num_elem_loc, DIM*Np), intent(inout) :: rhs_loc
    procedure(forcing_if) :: f_forcing
    procedure(dirichlet_if) :: f_dirichlet
    procedure(neumann_if) :: f_neumann
    integer :: ns, k
    integer(c_int32_t), allocatable :: tets(:)
    real(c_double), allocatable :: strike(:), normal(:), m0(:)

    pf => f_forcing
    pd => f_dirichlet
    pn => f_neumann
    call check(lymph_build_rhs_fn(ctx, c_funloc(tramp_forcing), c_funloc(tramp_dirichlet), &
                                  c_funloc(tramp_neumann), c_null_ptr, LYMPH_SLOT_RHS_DYN, c_null_ptr), &
               'lymph_build_rhs_fn')
    ns = PolyData%nload_sism_el
    if (ns > 0) then
      allocate(tets(ns), strike(3 * ns), normal(3 * ns), m0(ns))
      do k = 1, ns
        tets(k) = locate_point(PolyData%val_sism_el(k, 1:3))
        strike(3 * k - 2:3 * k) = PolyData%val_sism_el(1, 7:9)
        normal(3 * k - 2:3 * k) = PolyData%val_sism_el(1, 10:12)
        m0(k) = PolyData%val_sism_el(1, 14)
      end do
      call check(lymph_add_double_couples(ctx, ns, tets, strike, normal, m0, LYMPH_SLOT_RHS_DYN), &
                 'lymph_add_double_couples')
    end if
    call check(lymph_get_slot(ctx, LYMPH_SLOT_RHS_DYN, rhs_loc), 'lymph_get_slot')
  end subroutine MAKE_RHS_FREE

  !> COMPUTE_MODAL_COEFFICIENTS_FREE (src/matrix_free.f90:847-997): L2
  !> projection of f_analytic with the device's modal mass factors
  !> (R_M_modal_loc is not read).
  subroutine COMPUTE_MODAL_COEFFICIENTS_FREE(PolyMesh, Np, R_M_modal_loc, f_analytic, modal_coeff)
    type(Mesh_Structure) :: PolyMesh
    integer(kind=4) :: Np
    real(kind=8), dimension(:,:,:,:), intent(in) :: R_M_modal_loc
    procedure(analytic_if) :: f_analytic
    real(kind=8), dimension(:,:), allocatable, target, intent(out) :: modal_coeff

    allocate(modal_coeff(flat%n_poly, DIM * Np))
    pa => f_analytic
    call check(lymph_project_fn(ctx, c_funloc(tramp_analytic), c_null_ptr, LYMPH_SLOT_UEX, &
                                c_loc(modal_coeff)), 'lymph_project_fn')
  end subroutine COMPUTE_MODAL_COEFFICIENTS_FREE

  !> FIRST_TIME_STEP_MATRIX_FREE (src/matrix_free.f90:1100-1227):
  !> u1 = half_dt2 M^-1 (F_s + F_d time_function(t) - K u0) + u0 + time_step v0.
  subroutine FIRST_TIME_STEP_MATRIX_FREE(PolyMesh, Np, t, K_loc, R_M_loc, rhs_stat_loc, rhs_dyn_loc, &
                                         u0_loc, v0_loc, usol_loc, u0_mpi)
    type(Mesh_Structure), intent(in) :: PolyMesh
    real(kind=8), intent(in) :: t
    integer(kind=4), intent(in) :: Np
    real(kind=8), dimension(:,:,:,:), intent(in) :: R_M_loc, K_loc
    real(kind=8), dimension(PolyMesh%num_elem_loc, DIM*Np), intent(in) :: rhs_stat_loc, rhs_dyn_loc
    real(kind=8), dimension(PolyMesh%num_elem_loc, DIM*Np), intent(in) :: u0_loc, v0_loc
    real(kind=8), dimension(PolyMesh%num_elem_loc, DIM*Np), intent(out) :: usol_loc
    real(kind=8), dimension(:,:), intent(in) :: u0_mpi
    real(kind=8) :: tt

    call check(lymph_set_slot(ctx, LYMPH_SLOT_RHS_STAT, rhs_stat_loc), 'lymph_set_slot')
    call check(lymph_set_slot(ctx, LYMPH_SLOT_RHS_DYN, rhs_dyn_loc), 'lymph_set_slot')
    tt = t
    call check(lymph_first_step(ctx, time_function(tt), time_step, u0_loc, v0_loc, usol_loc), &
               'lymph_first_step')
  end subroutine FIRST_TIME_STEP_MATRIX_FREE

  !> TIME_STEP_MATRIX_FREE (src/matrix_free.f90:1230-1385):
  !> usol = dt2 M^-1 (F_s + F_d time_function(t) - K un) + 2 un - u0.
  subroutine TIME_STEP_MATRIX_FREE(PolyMesh, Np, t, K_loc, R_M_loc, rhs_stat_loc, rhs_dyn_loc, &
                                   u0_loc, un_loc, usol_loc, un_mpi)
    type(Mesh_Structure), intent(in) :: PolyMesh
    real(kind=8), intent(in) :: t
    integer(kind=4), intent(in) :: Np
    real(kind=8), dimension(:,:,:,:), intent(in) :: R_M_loc, K_loc
    real(kind=8), dimension(PolyMesh%num_elem_loc, DIM*Np), intent(in) :: rhs_stat_loc, rhs_dyn_loc
    real(kind=8), dimension(PolyMesh%num_elem_loc, DIM*Np), intent(in) :: u0_loc, un_loc
    real(kind=8), dimension(PolyMesh%num_elem_loc, DIM*Np), intent(out) :: usol_loc
    real(kind=8), dimension(:,:), intent(in) :: un_mpi
    real(kind=8) :: tt

    call check(lymph_set_slot(ctx, LYMPH_SLOT_RHS_STAT, rhs_stat_loc), 'lymph_set_slot')
    call check(lymph_set_slot(ctx, LYMPH_SLOT_RHS_DYN, rhs_dyn_loc), 'lymph_set_slot')
    tt = t
    call check(lymph_step(ctx, time_function(tt), time_step, u0_loc, un_loc, usol_loc), 'lymph_step')
  end subroutine TIME_STEP_MATRIX_FREE

  !> COMPUTE_ERROR_L2_MATRIX_FREE (src/matrix_free.f90:1388-1418): this
  !> rank's sum of du^T M_modal du (M_modal_loc is not read).
  subroutine COMPUTE_ERROR_L2_MATRIX_FREE(PolyMesh, Np, M_modal_loc, uh_loc, uex_loc, err_L2_loc)
    type(Mesh_Structure), intent(in) :: PolyMesh
    integer(kind=4), intent(in) :: Np
    real(kind=8), dimension(:,:,:,:), intent(in) :: M_modal_loc
    real(kind=8), dimension(PolyMesh%num_elem_loc, DIM*Np), intent(in) :: uh_loc, uex_loc
    real(kind=8), intent(out) :: err_L2_loc
    call check(lymph_error_l2(ctx, uh_loc, uex_loc, err_L2_loc), 'lymph_error_l2')
  end subroutine COMPUTE_ERROR_L2_MATRIX_FREE

  !> COMPUTE_ERROR_DG_MATRIX_FREE (src/matrix_free.f90:1421-1504): this
  !> rank's sum of du^T A_dg du (A_dg_loc and the halo are not read).
  subroutine COMPUTE_ERROR_DG_MATRIX_FREE(PolyMesh, Np, A_dg_loc, uh_loc, uex_loc, uh_mpi, uex_mpi, &
                                          err_DG_loc)
    type(Mesh_Structure), intent(in) :: PolyMesh
    integer(kind=4), intent(in) :: Np
    real(kind=8), dimension(:,:,:,:), intent(in) :: A_dg_loc
    real(kind=8), dimension(PolyMesh%num_elem_loc, DIM*Np), intent(in) :: uh_loc, uex_loc
    real(kind=8), dimension(:,:), intent(in) :: uh_mpi, uex_mpi
    real(kind=8), intent(out) :: err_DG_loc
    call check(lymph_error_dg(ctx, uh_loc, uex_loc, err_DG_loc), 'lymph_error_dg')
  end subroutine COMPUTE_ERROR_DG_MATRIX_FREE

end module matrix_free_gpu

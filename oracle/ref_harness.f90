! ref_harness.f90 -- C-callable wrappers around the reference's own numeric
! leaf modules, compiled from the reference sources where they lie (INCLUDE
! lines below; nothing is copied).  TEST INFRASTRUCTURE ONLY: the resulting
! oracle/_ref/libref_numerics.so pins the C oracle element by element.
!
! The included modules need no MPI, METIS or other library; the modules that
! do (matrix_free, exchange_data, the driver) are not built here.
include '/root/reference/src/Poly_global.f90'
include '/root/reference/src/Poly_mesh.f90'
include '/root/reference/src/vet_mat_operations.f90'
include '/root/reference/src/Poly_ref_mappings.f90'
include '/root/reference/src/basis_function.f90'
include '/root/reference/src/problem_data_and_properties.f90'
include '/root/reference/src/assemble_element.f90'

module lymph_ref_harness
  use iso_c_binding
  use global_parameters
  use basis_function
  use Poly_ref_mappings
  use vet_mat_operations
  use problem_data_and_properties
  use assemble_element
  implicit none
contains

  subroutine quad_tables(p, nq3, nq2, nodtet3, weitet3, nodtria2, weitria2)
    integer, intent(in) :: p
    integer, intent(out) :: nq3, nq2
    real(8), allocatable, intent(out) :: nodtet3(:,:), weitet3(:), nodtria2(:,:), weitria2(:)
    real(8), allocatable :: nod2(:,:), wei2(:), nod3(:,:), wei3(:)
    call quadrature(nod2, wei2, nod3, wei3, p, nq3, nq2)
    allocate(nodtet3(4,nq3), weitet3(nq3), nodtria2(4,nq2), weitria2(nq2))
    call mapping_quadrature_3D(nod3, wei3, nq3, nodtet3, weitet3)
    call mapping_quadrature_2D(nod2, wei2, nq2, nodtria2, weitria2)
  end subroutine

  subroutine rh_quadrature(p, t3, w3, t2, w2) bind(C, name='rh_quadrature')
    integer(c_int), value :: p
    real(c_double) :: t3(4,(p+1)**3), w3((p+1)**3), t2(4,(p+1)**2), w2((p+1)**2)
    integer :: nq3, nq2
    real(8), allocatable :: a(:,:), b(:), c(:,:), d(:)
    call quad_tables(p, nq3, nq2, a, b, c, d)
    t3 = a; w3 = b; t2 = c; w2 = d
  end subroutine

  subroutine rh_basis_list(p, np, blist) bind(C, name='rh_basis_list')
    integer(c_int), value :: p, np
    integer(c_int) :: blist(np,3)
    integer(4) :: bl(np,3)
    call basis_list(bl, p, np)
    blist = bl
  end subroutine

  subroutine rh_jacobians(x, y, z, fk, jdet) bind(C, name='rh_jacobians')
    real(c_double) :: x(4), y(4), z(4), fk(3,4), jdet
    real(8) :: jinv(3,3)
    call jacobians(x, y, z, fk, jinv, jdet)
  end subroutine

  subroutine rh_basis(p, np, bbox, fk, phi, dphi) bind(C, name='rh_basis')
    integer(c_int), value :: p, np
    real(c_double) :: bbox(3,2), fk(3,4), phi(np,(p+1)**3), dphi(3,np,(p+1)**3)
    integer :: nq3, nq2
    integer(4) :: bl(np,3)
    real(8), allocatable :: a(:,:), b(:), c(:,:), d(:)
    call quad_tables(p, nq3, nq2, a, b, c, d)
    call basis_list(bl, p, np)
    call basis(phi, dphi, bbox, np, bl, fk, a, nq3)
  end subroutine

  subroutine rh_basis_boundary(p, np, face, e2, bbox1, bbox2, fk, phib, gradb) &
      bind(C, name='rh_basis_boundary')
    integer(c_int), value :: p, np, face, e2
    real(c_double) :: bbox1(3,2), bbox2(3,2), fk(3,4)
    real(c_double) :: phib(np,(p+1)**2,2), gradb(3,np,(p+1)**2,2)
    integer :: nq3, nq2
    integer(4) :: bl(np,3)
    real(8) :: nm(4,4,4), nmi(2,3,4)
    real(8), allocatable :: a(:,:), b(:), c(:,:), d(:)
    call quad_tables(p, nq3, nq2, a, b, c, d)
    call basis_list(bl, p, np)
    call tria2tetfaces_maps(nm, nmi)
    call basis_boundary(phib, gradb, face, e2, bbox1, bbox2, bl, np, fk, nm, c, nq2)
  end subroutine

  subroutine rh_stiff_vol(p, np, jdet, lam, mu, dphi, v) bind(C, name='rh_stiff_vol')
    integer(c_int), value :: p, np
    real(c_double), value :: jdet, lam, mu
    real(c_double) :: dphi(3,np,(p+1)**3), v(3,3,np,np)
    integer :: nq3, nq2
    real(8), allocatable :: a(:,:), b(:), c(:,:), d(:)
    call quad_tables(p, nq3, nq2, a, b, c, d)
    call MAKE_STIFFNESS_VOLUME(np, jdet, b, nq3, lam, mu, dphi, v)
  end subroutine

  subroutine rh_mass_vol(p, np, jdet, phi, rho, m) bind(C, name='rh_mass_vol')
    integer(c_int), value :: p, np
    real(c_double), value :: jdet, rho
    real(c_double) :: phi(np,(p+1)**3), m(3,3,np,np)
    integer :: nq3, nq2
    real(8), allocatable :: a(:,:), b(:), c(:,:), d(:)
    call quad_tables(p, nq3, nq2, a, b, c, d)
    call MAKE_MASS_VOLUME(np, jdet, b, nq3, phi, rho, m)
  end subroutine

  subroutine rh_stiff_face(p, np, alpha, e2, hk1, hk2, nrm, area, lam, mu, phib, gradb, &
      s1, i1, s2, i2, it2) bind(C, name='rh_stiff_face')
    integer(c_int), value :: p, np, e2
    real(c_double), value :: alpha, hk1, hk2, area, lam, mu
    real(c_double) :: nrm(3), phib(np,(p+1)**2,2), gradb(3,np,(p+1)**2,2)
    real(c_double), dimension(3,3,np,np) :: s1, i1, s2, i2, it2
    integer :: nq3, nq2
    real(8), allocatable :: a(:,:), b(:), c(:,:), d(:)
    call quad_tables(p, nq3, nq2, a, b, c, d)
    call MAKE_STIFFNESS_FACE(alpha, p, np, e2, hk1, hk2, nrm, area, d, nq2, lam, mu, &
                             phib, gradb, s1, i1, s2, i2, IT_E2=it2)
  end subroutine

  subroutine rh_rhs_volume(p, np, fk, jdet, lam, mu, rho, phi, rhs) bind(C, name='rh_rhs_volume')
    integer(c_int), value :: p, np
    real(c_double), value :: jdet, lam, mu, rho
    real(c_double) :: fk(3,4), phi(np,(p+1)**3), rhs(3,np)
    integer :: nq3, nq2
    real(8), allocatable :: a(:,:), b(:), c(:,:), d(:)
    call quad_tables(p, nq3, nq2, a, b, c, d)
    call MAKE_RHS_VOLUME(np, fk, jdet, a, b, nq3, lam, mu, phi, rhs, rho, f_time)
  end subroutine

  subroutine rh_rhs_face(p, np, face, e2, hk1, hk2, nrm, area, fk, lam, mu, phib, gradb, &
      tag, rhs) bind(C, name='rh_rhs_face')
    integer(c_int), value :: p, np, face, e2, tag
    real(c_double), value :: hk1, hk2, area, lam, mu
    real(c_double) :: nrm(3), fk(3,4), phib(np,(p+1)**2,2), gradb(3,np,(p+1)**2,2), rhs(3,np)
    integer :: nq3, nq2
    real(8) :: nm(4,4,4), nmi(2,3,4), alpha, theta, cc
    real(8), allocatable :: a(:,:), b(:), c(:,:), d(:)
    call quad_tables(p, nq3, nq2, a, b, c, d)
    call tria2tetfaces_maps(nm, nmi)
    call set_properties(alpha, theta, cc)
    call MAKE_RHS_FACE(theta, alpha, p, np, face, e2, hk1, hk2, nrm, area, fk, c, d, nq2, &
                       lam, mu, nm, phib, gradb, tag, rhs, gd, gn)
  end subroutine

  subroutine rh_cholesky(n, a, r) bind(C, name='rh_cholesky')
    integer(c_int), value :: n
    real(c_double) :: a(n,n), r(n,n)
    r = cholesky(a, n)
  end subroutine

  subroutine rh_solve_lu(n, r, b, x) bind(C, name='rh_solve_lu')
    integer(c_int), value :: n
    real(c_double) :: r(n,n), b(n), x(n)
    real(8) :: rt(n,n)
    rt = transpose(r)
    x = solve_LU(rt, r, b, n)
  end subroutine

  ! out(:,1) f_time, (:,2) gd tag 1, (:,3) gn tag 1, (:,4) uex, (:,5) ic_velocity
  subroutine rh_fields(x, lam, mu, rho, nrm, out) bind(C, name='rh_fields')
    real(c_double) :: x(3), nrm(3), out(3,5)
    real(c_double), value :: lam, mu, rho
    integer(4) :: tag
    real(8) :: l, m, r
    tag = 1
    l = lam; m = mu; r = rho
    out(:,1) = f_time(l, m, x, r)
    out(:,2) = gd(x, tag)
    out(:,3) = gn(l, m, nrm, x, tag)
    out(:,4) = uex(x)
    out(:,5) = ic_velocity(x)
  end subroutine

  function rh_time_function(t) bind(C, name='rh_time_function') result(g)
    real(c_double), value :: t
    real(c_double) :: g
    real(8) :: tt
    tt = t
    g = time_function(tt)
  end function

  ! MAKE_DOUBLE_COUPLE with the moment of moment_density(m0, vol, s, n)
  subroutine rh_double_couple(p, np, jdet, dphi, m0, vol, s, n, rhs) bind(C, name='rh_double_couple')
    integer(c_int), value :: p, np
    real(c_double), value :: jdet, m0, vol
    real(c_double) :: dphi(3,np,(p+1)**3), s(3), n(3), rhs(3*np)
    integer :: nq3, nq2
    real(8), allocatable :: a(:,:), b(:), c(:,:), d(:)
    real(8) :: mom(3,3), mm, vv, ss(3), nn(3)
    call quad_tables(p, nq3, nq2, a, b, c, d)
    mm = m0; vv = vol; ss = s; nn = n
    mom = moment_density(mm, vv, ss, nn)
    call MAKE_DOUBLE_COUPLE(np, jdet, b, nq3, dphi, mom, rhs)
  end subroutine

  function rh_det4(a) bind(C, name='rh_det4') result(d)
    real(c_double) :: a(4,4)
    real(c_double) :: d
    d = det4(a)
  end function

end module lymph_ref_harness

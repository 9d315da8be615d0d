! ref_harness.f90 -- C-callable entry points into the reference modules that
! compile from their own source files (mod_constants, mod_numerical_params,
! mod_nvtx, mod_maths, quadrature_rules, elem_hex, jacobian_oper,
! mod_fluid_viscosity, mod_filters).
!
! Test infrastructure only: the oracle's restatement of the basis, the shape
! functions, the Jacobian, Sutherland's law and the filter tables is checked
! against these.  The element order porder is fixed when the reference modules
! are compiled (-D__PORDER__), so one library is built per order.
module ref_harness
   use iso_c_binding
   use mod_constants
   use mod_maths
   use quadrature_rules
   use elem_hex
   use jacobian_oper
   use mod_fluid_viscosity
   use mod_filters
   implicit none
contains

   function ref_porder() bind(C, name='ref_porder') result(p)
      integer(c_int) :: p
      p = porder
   end function ref_porder

   ! GLL roots and weights (mod_maths.f90:29-39)
   subroutine ref_gll(xi, w) bind(C, name='ref_gll')
      real(c_double), intent(out) :: xi(porder+1), w(porder+1)
      call getGaussLobattoLegendre_weights_and_roots(porder, w, xi)
   end subroutine ref_gll

   ! Lagrange basis values and derivatives at x on a 1-D grid (mod_maths.f90:243-323)
   subroutine ref_lagrange(xi_grid, x, l, dl) bind(C, name='ref_lagrange')
      real(c_double), intent(in) :: xi_grid(porder+1)
      real(c_double), value :: x
      real(c_double), intent(out) :: l(porder+1), dl(porder+1)
      call eval_lagrangePoly(porder, xi_grid, x, l)
      call eval_lagrangePolyDeriv(porder, xi_grid, x, dl)
   end subroutine ref_lagrange

   ! Quadrature points and shape functions, in the sequence of the reference's
   ! set-up (CFDSolverBase.f90:1530-1580): GaussLobattoLegendre_hex, then
   ! hex_highorder at every Gauss point.
   subroutine ref_shape(atoIJK, xgp, wgp, Ngp, dNgp, dlxigp_ip) bind(C, name='ref_shape')
      integer(c_int), intent(in) :: atoIJK(nnode)
      real(c_double), intent(out) :: xgp(ngaus,ndime), wgp(ngaus), Ngp(ngaus,nnode)
      real(c_double), intent(out) :: dNgp(ndime,nnode,ngaus), dlxigp_ip(ngaus,ndime,porder+1)
      real(8) :: N(nnode), dN(ndime,nnode), Nl(nnode), dNl(ndime,nnode), dl(ndime,porder+1)
      integer(4) :: ig
      call GaussLobattoLegendre_hex(porder, ngaus, atoIJK, xgp, wgp)
      do ig = 1, ngaus
         dl(:,:) = 0.0d0
         call hex_highorder(porder, nnode, xgp(ig,1), xgp(ig,2), xgp(ig,3), atoIJK, N, dN, &
                            Nl, dNl, dl)
         Ngp(ig,:) = N(:)
         dNgp(:,:,ig) = dN(:,:)
         dlxigp_ip(ig,:,:) = dl(:,:)
      end do
   end subroutine ref_shape

   ! Jacobian inverse and Gauss-point volumes (jacobian_oper.f90:7-140)
   subroutine ref_elem_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He) &
         bind(C, name='ref_elem_jacobian')
      integer(c_int), value :: nelem, npoin
      integer(c_int), intent(in) :: connec(nelem,nnode)
      real(c_double), intent(in) :: coord(npoin,ndime), dNgp(ndime,nnode,ngaus), wgp(ngaus)
      real(c_double), intent(out) :: gpvol(1,ngaus,nelem), He(ndime,ndime,ngaus,nelem)
      call elem_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He)
   end subroutine ref_elem_jacobian

   ! Sutherland's law (mod_fluid_viscosity.f90:21-41)
   subroutine ref_sutherland(npoin, Tem, mu_factor, mu_fluid) bind(C, name='ref_sutherland')
      integer(c_int), value :: npoin
      real(c_double), intent(in) :: Tem(npoin), mu_factor(npoin)
      real(c_double), intent(inout) :: mu_fluid(npoin)
      call sutherland_viscosity(npoin, Tem, mu_factor, mu_fluid)
   end subroutine ref_sutherland

   ! Filter tables of the spectral sensor (mod_filters.f90:15-55)
   subroutine ref_filters(cijk, al, am, an) bind(C, name='ref_filters')
      integer(c_int), intent(out) :: cijk(porder+3)
      real(c_double), intent(out) :: al(3), am(3), an(3)
      call init_filters()
      cijk(:) = convertIJK(0:porder+2)
      al(:) = al_weights(-1:1)
      am(:) = am_weights(-1:1)
      an(:) = an_weights(-1:1)
      call deallocate_filters()
   end subroutine ref_filters

end module ref_harness

! shim_driver.f90 -- test program for the Fortran layer mod_sod2d_hip: reads a
! case written by tests/test_gpu_fortran_shim.py (raw little-endian FP64 arrays
! in the reference's layouts), maps the arrays to the device, calls the
! operators through their reference interfaces and writes the results back
! as raw FP64 arrays.  The arrays are real(__RP__), the module's rp: with
! -D__RP__=4 the real(4) interfaces are exercised (the live-path and LPS
! operators exist in the real(8) build only).  Usage: shim_driver <directory>
#ifndef __RP__
#define __RP__ 8
#endif
program shim_driver
   use iso_c_binding
   use mod_sod2d_hip
   implicit none
   integer, parameter :: porder = __PORDER__, nnode = (porder+1)**3, ngaus = nnode
   integer, parameter :: rk = __RP__
   character(len=512) :: dir
   integer(4) :: nelem, npoin, npoin_w, u_meta
   integer(4), allocatable :: connec(:,:), lpoin_w(:), inv(:,:,:), gI(:), gJ(:), gK(:)
   real(rk), allocatable :: He(:,:,:,:), gpvol(:,:,:), dl(:,:,:), Ngp(:,:), dNgp(:,:,:), xgp(:,:)
   real(rk), allocatable :: wgp(:), coord(:,:)
   real(rk), allocatable :: u(:,:), q(:,:), rho(:), pr(:), h(:), Tem(:), mu_fluid(:), mu_e(:,:)
   real(rk), allocatable :: mu_sgs(:,:), Ml(:), invMl(:), Reta(:), eta(:), helem_l(:,:), helem(:)
   real(rk), allocatable :: csound(:), Rmass(:), Rmom(:,:), Rener(:), Rconv(:), mue_out(:,:)
   real(rk), allocatable :: mue_l(:,:), rho1(:), q1(:,:), E1(:), rhoS(:), qS(:,:), ES(:)
   real(rk), allocatable :: mu_imex(:,:), mue_l2(:,:), divU(:), curlU(:,:)
   integer(4), allocatable :: lelpn(:), atoIJK(:)
   real(rk), allocatable :: tau(:), PMass(:,:), PEner(:,:), PMX(:,:), PMY(:,:), PMZ(:,:)
   real(rk), allocatable :: Rsm(:), Rsq(:,:), Rse(:)
   real(rk) :: leviCivi(3,3,3)
   real(rk) :: Cp, Prt, gamma_gas, ce, cmax, cglob, Rgas, dt, lam1, lams, gam
   integer(4) :: flag_norm

   call get_command_argument(1, dir)
   open(newunit=u_meta, file=trim(dir)//'/meta.txt', status='old', action='read')
   read(u_meta, *) nelem, npoin, npoin_w
   read(u_meta, *) Cp, Prt, gamma_gas
   read(u_meta, *) ce, cmax, cglob, Rgas, flag_norm
   read(u_meta, *) dt, lam1, lams, gam
   close(u_meta)

   allocate(connec(nelem,nnode), lpoin_w(npoin_w), inv(porder+1,porder+1,porder+1))
   allocate(gI(nnode), gJ(nnode), gK(nnode))
   allocate(He(3,3,ngaus,nelem), gpvol(1,ngaus,nelem), dl(ngaus,3,porder+1))
   allocate(Ngp(ngaus,nnode), dNgp(3,nnode,ngaus), xgp(ngaus,3), wgp(ngaus), coord(npoin,3))
   allocate(u(npoin,3), q(npoin,3), rho(npoin), pr(npoin), h(npoin), Tem(npoin), mu_fluid(npoin))
   allocate(mu_e(nelem,ngaus), mu_sgs(nelem,ngaus), Ml(npoin), invMl(npoin), Reta(npoin))
   allocate(eta(npoin), helem_l(nelem,nnode), helem(nelem), csound(npoin))
   allocate(Rmass(npoin), Rmom(npoin,3), Rener(npoin), Rconv(npoin), mue_out(nelem,ngaus))
   allocate(mue_l(nelem,nnode), rho1(npoin), q1(npoin,3), E1(npoin), rhoS(npoin), qS(npoin,3))
   allocate(ES(npoin), mu_imex(nelem,ngaus), mue_l2(nelem,nnode), divU(npoin), curlU(npoin,3))
   allocate(lelpn(npoin), atoIJK(nnode))
   allocate(tau(nelem), PMass(npoin,3), PEner(npoin,3), PMX(npoin,3), PMY(npoin,3), PMZ(npoin,3))
   allocate(Rsm(npoin), Rsq(npoin,3), Rse(npoin))
   leviCivi = 0.0_rk; atoIJK = 0; mu_imex = 0.0_rk; mue_l2 = 0.0_rk
   Ngp = 0.0_rk; dNgp = 0.0_rk; xgp = 0.0_rk; wgp = 0.0_rk; coord = 0.0_rk; Ml = 0.0_rk
   csound = 0.0_rk; mu_sgs = 0.0_rk; mue_out = 0.0_rk; mue_l = 0.0_rk

   call rd_i(connec, 'connec'); call rd_i(lpoin_w, 'lpoin_w'); call rd_i(inv, 'invAtoIJK')
   call rd_i(gI, 'gI'); call rd_i(gJ, 'gJ'); call rd_i(gK, 'gK')
   call rd_r(He, 'He'); call rd_r(gpvol, 'gpvol'); call rd_r(dl, 'dlxigp_ip')
   call rd_r(u, 'u'); call rd_r(q, 'q'); call rd_r(rho, 'rho'); call rd_r(pr, 'pr')
   call rd_r(h, 'h'); call rd_r(Tem, 'Tem'); call rd_r(mu_fluid, 'mu_fluid')
   call rd_r(mu_e, 'mu_e'); call rd_r(invMl, 'invMl'); call rd_r(Reta, 'Reta')
   call rd_r(eta, 'eta'); call rd_r(helem_l, 'helem_l'); call rd_r(helem, 'helem')
   call rd_r(rho1, 'rho1'); call rd_r(q1, 'q1'); call rd_r(E1, 'E1')
   call rd_i(lelpn, 'lelpn'); call rd_r(csound, 'csound'); call rd_r(Ml, 'Ml')
   rhoS = rho; qS = q; ES = h

   call sod2d_hip_init(0)
   call sod2d_hip_set_visc_params(ce, cmax, cglob, Rgas, flag_norm, .false.)
   call sod2d_hip_enter_data(connec); call sod2d_hip_enter_data(lpoin_w)
   call sod2d_hip_enter_data(inv); call sod2d_hip_enter_data(gI)
   call sod2d_hip_enter_data(gJ); call sod2d_hip_enter_data(gK)
   call sod2d_hip_enter_data(He); call sod2d_hip_enter_data(gpvol); call sod2d_hip_enter_data(dl)
   call sod2d_hip_enter_data(u); call sod2d_hip_enter_data(q); call sod2d_hip_enter_data(rho)
   call sod2d_hip_enter_data(pr); call sod2d_hip_enter_data(h); call sod2d_hip_enter_data(Tem)
   call sod2d_hip_enter_data(mu_fluid); call sod2d_hip_enter_data(mu_e)
   call sod2d_hip_enter_data(mu_sgs); call sod2d_hip_enter_data(invMl)
   call sod2d_hip_enter_data(Reta); call sod2d_hip_enter_data(eta)
   call sod2d_hip_enter_data(helem_l); call sod2d_hip_enter_data(helem)
   call sod2d_hip_enter_data(Rmass, .false.); call sod2d_hip_enter_data(Rmom, .false.)
   call sod2d_hip_enter_data(Rener, .false.); call sod2d_hip_enter_data(Rconv, .false.)
   call sod2d_hip_enter_data(mue_out, .false.); call sod2d_hip_enter_data(mue_l)
   call sod2d_hip_enter_data(rho1); call sod2d_hip_enter_data(q1); call sod2d_hip_enter_data(E1)
   call sod2d_hip_enter_data(rhoS); call sod2d_hip_enter_data(qS); call sod2d_hip_enter_data(ES)
   call sod2d_hip_enter_data(lelpn); call sod2d_hip_enter_data(csound)
   call sod2d_hip_enter_data(Ml); call sod2d_hip_enter_data(mu_imex)
   call sod2d_hip_enter_data(mue_l2); call sod2d_hip_enter_data(divU, .false.)
   call sod2d_hip_enter_data(curlU, .false.)
   call sod2d_hip_enter_data(tau, .false.); call sod2d_hip_enter_data(PMass, .false.)
   call sod2d_hip_enter_data(PEner, .false.); call sod2d_hip_enter_data(PMX, .false.)
   call sod2d_hip_enter_data(PMY, .false.); call sod2d_hip_enter_data(PMZ, .false.)
   call sod2d_hip_enter_data(Rsm, .false.); call sod2d_hip_enter_data(Rsq, .false.)
   call sod2d_hip_enter_data(Rse, .false.)

   ! the RHS of updateF: convection (initialising) then diffusion (accumulating),
   ! both with fact = -1 (time_integ_ls.f90:605-640)
   call full_convec_ijk(nelem,npoin,connec,Ngp,dNgp,He,gpvol,dl,xgp,inv,gI,gJ,gK,u,q,rho,pr,h, &
                        Rmass,Rmom,Rener,.true.,-1.0_rk)
   call full_diffusion_ijk(nelem,npoin,connec,Ngp,He,gpvol,dl,inv,gI,gJ,gK,Cp,Prt,rho,rho,u,Tem, &
                           mu_fluid,mu_e,mu_sgs,Ml,Rmass,Rmom,Rener,.false.,-1.0_rk)
   call sod2d_hip_update_host(Rmass); call sod2d_hip_update_host(Rmom)
   call sod2d_hip_update_host(Rener)
   call wr(Rmass, 'out_Rmass'); call wr(Rmom, 'out_Rmom'); call wr(Rener, 'out_Rener')

   ! stage combine with the fused mass inverse
   call rk_stage_update(npoin,npoin_w,lpoin_w,dt,lam1,lams,gam,rho1,q1,E1,rhoS,qS,ES, &
                        Rmass,Rmom,Rener,invMl)
   call sod2d_hip_update_host(rhoS); call sod2d_hip_update_host(qS); call sod2d_hip_update_host(ES)
   call sod2d_hip_update_host(Rmass)
   call wr(rhoS, 'out_rho'); call wr(qS, 'out_q'); call wr(ES, 'out_E')
   call wr(Rmass, 'out_Rmass_scaled')

   ! scalar transport residual with the saved fact of its last call (none: 1)
   call generic_scalar_convec_ijk(nelem,npoin,connec,Ngp,dNgp,He,gpvol,dl,xgp,inv,gI,gJ,gK,q, &
                                  eta,u,Rconv)
   call lumped_solver_scal_opt(npoin,npoin_w,lpoin_w,invMl,Rconv)
   call sod2d_hip_update_host(Rconv)
   call wr(Rconv, 'out_Rconv')

   ! entropy-viscosity sensor
   call smart_visc_spectral(nelem,npoin,npoin_w,connec,lpoin_w,Reta,Ngp,coord,dNgp,gpvol,wgp, &
                            gamma_gas,rho,u,csound,Tem,eta,helem_l,helem,Ml,mue_out,inv,gI,gJ,gK, &
                            mue_l)
   call sod2d_hip_update_host(mue_out)
   call wr(mue_out, 'out_mu_e')

#if __RP__ == 8
   ! the live path's sensor and limiter (time_integ_ls.f90:455-463)
   call smart_visc_spectral_imex(nelem,npoin,npoin_w,connec,lpoin_w,Reta,Ngp,coord,dNgp,gpvol, &
                                 wgp,gamma_gas,rho,u,csound,Tem,eta,helem_l,helem,Ml,mu_imex,inv, &
                                 gI,gJ,gK,mue_l2)
   call sod2d_hip_update_host(mu_imex)
   call wr(mu_imex, 'out_mu_imex')
   call eval_divergence(nelem,npoin,connec,He,gpvol,dl,inv,gI,gJ,gK,u,divU)
   call compute_vorticity(nelem,npoin,npoin_w,lpoin_w,connec,lelpn,He,dNgp,leviCivi,dl,atoIJK, &
                          inv,gI,gJ,gK,u,curlU,.true.)
   call sod2d_hip_update_host(divU); call sod2d_hip_update_host(curlU)
   call wr(divU, 'out_divU'); call wr(curlU, 'out_curlU')
   call ducros_visc_limit(nelem,npoin,npoin_w,connec,lpoin_w,Ngp,coord,dNgp,gpvol,wgp,rho,u, &
                          divU,curlU,csound,helem_l,helem,Ml,mu_imex,inv,gI,gJ,gK,mue_l2)
   call sod2d_hip_update_host(mu_imex)
   call wr(mu_imex, 'out_mu_ducros')

   ! LPS stabilisation (time_integ_ls.f90:266-268,607-611)
   call comp_tau(nelem,npoin,connec,csound,u,helem,dt,tau)
   call full_proj_ijk(nelem,npoin,npoin_w,connec,lpoin_w,Ngp,He,gpvol,dl,inv,gI,gJ,gK,Cp,Prt, &
                      rho,u,Tem,Ml,invMl,PMass,PEner,PMX,PMY,PMZ)
   call full_diff_stab_ijk(nelem,npoin,connec,Ngp,He,gpvol,dl,inv,gI,gJ,gK,Cp,Prt,rho,rho,u,Tem, &
                           mu_fluid,mu_e,mu_sgs,Ml,PMass,PEner,PMX,PMY,PMZ,tau,Rsm,Rsq,Rse, &
                           .true.,-1.0_rk)
   call sod2d_hip_update_host(tau); call sod2d_hip_update_host(PMass)
   call sod2d_hip_update_host(PMX); call sod2d_hip_update_host(Rsm)
   call sod2d_hip_update_host(Rsq); call sod2d_hip_update_host(Rse)
   call wr(tau, 'out_tau'); call wr(PMass, 'out_PMass'); call wr(PMX, 'out_PMX')
   call wr(Rsm, 'out_Rsm'); call wr(Rsq, 'out_Rsq'); call wr(Rse, 'out_Rse')
#endif

   call sod2d_hip_synchronize()
   call sod2d_hip_finalize()
   write(*,'(a)') 'shim_driver: ok'

contains

   subroutine rd_r(a, name)
      real(rk), intent(out) :: a(*)
      character(*), intent(in) :: name
      integer :: un, n
      real(8), allocatable :: t(:)
      inquire(file=trim(dir)//'/'//name//'.bin', size=n)
      allocate(t(n/8))
      open(newunit=un, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
           status='old', action='read')
      read(un) t
      close(un)
      a(1:n/8) = real(t, rk)
   end subroutine rd_r

   subroutine rd_i(a, name)
      integer(4), intent(out) :: a(*)
      character(*), intent(in) :: name
      integer :: un, n
      inquire(file=trim(dir)//'/'//name//'.bin', size=n)
      open(newunit=un, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
           status='old', action='read')
      read(un) a(1:n/4)
      close(un)
   end subroutine rd_i

   subroutine wr(a, name)
      real(rk), intent(in) :: a(..)
      character(*), intent(in) :: name
      integer :: un
      open(newunit=un, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
           status='replace', action='write')
      select rank (a)
      rank (1)
         write(un) real(a, 8)
      rank (2)
         write(un) real(a, 8)
      end select
      close(un)
   end subroutine wr

end program shim_driver

! tgv_host.f90 -- a Fortran host that runs the reference's explicit time loop
! (CFDSolverBase_evalTimeIteration's level copy, then rk_4_ls_main,
! time_integ_ls.f90:202-487, with updateF :489-691) on a periodic TGV case
! entirely through mod_sod2d_hip: the arrays are mapped once with
! sod2d_hip_enter_data, every step runs on the device, and the conserved fields
! and mu_e of each step are copied back and written out.
!
! The sequencing is the reference's at fixed dt:
!   * first step: Reta(:,1) from generic_scalar_convec_ijk of
!     f_eta = u(:,:,1) eta(:,1) and its lumped solve, smart_visc_spectral, and
!     updateF at y^n that keeps R (time_integ_ls.f90:282-331);
!   * stage 1 with the kept residual (rk_stage_update) on the first step; on
!     later steps the previous step's last stage stored it already
!     (rk_stage_fused with gam_next);
!   * stages 2..S: rk_stage_fused (updateF, lumped inverse and combine);
!   * end of step: eos_entropy_step, entropy_convec_residual and
!     smart_visc_spectral on the new state (:389-473).
! With the LPS stabilisation (flag_lps_stab, meta.txt's optional fifth line
! set to 1) comp_tau runs at the start of every step on the level-1 velocity
! (:266-268) and every updateF is rk_stage_lps_fused (full_proj_ijk and
! full_diff_stab_ijk in place of full_diffusion_ijk).
! The level copy 2 -> 1 of the conserved and primitive fields is an exchange
! of the level index (level 2 of the primitives is rewritten before it is
! read, and stage 1 reads the conserved level 1 only); eta(:,1) = eta(:,2) is
! a device copy.
!
! Usage: tgv_host <case directory>.  The directory holds meta.txt, the
! Runge-Kutta tables rkL.bin / rkG.bin ((S+1) x (S+1), the reference's lambda
! and gamma of rk_4_ls_main) and the mesh, geometry and initial fields as raw
! binary files in Fortran order; step<k>_{rho,q,E,mu_e}.bin are written there.
program tgv_host
   use mod_sod2d_hip
   implicit none
   integer(4), parameter :: p = __PORDER__, nnode = (p+1)**3, ndime = 3
   character(len=512) :: dir
   integer(4) :: nelem, npoin, npoin_w, nsteps, nstage, istep, s, l1, l2, lt
   real(8) :: gamma_gas, Rgas, Cp, Prt, mu_factor, po, rho0, dt, ce, cmax, cglob
   integer(4) :: flag_norm, lps, ios
   integer(4), allocatable :: connec(:,:), lpoin_w(:), invAtoIJK(:,:,:)
   integer(4), allocatable :: gmshAtoI(:), gmshAtoJ(:), gmshAtoK(:)
   real(8), allocatable :: He(:,:,:,:), gpvol(:,:,:), dlxigp_ip(:,:,:), invMl(:)
   real(8), allocatable :: helem_l(:,:), helem(:)
   real(8), allocatable :: rho(:,:), q(:,:,:), E(:,:), u(:,:,:), pr(:,:), Tem(:,:), e_int(:,:)
   real(8), allocatable :: eta(:,:), csound(:), machno(:), mu_fluid(:)
   real(8), allocatable :: mu_e(:,:), mue_l(:,:), Reta(:,:), auxReta(:), f_eta(:,:)
   real(8), allocatable :: Rmass(:), Rmom(:,:), Rener(:)
   real(8), allocatable :: rkL(:,:), rkG(:,:)
   ! the LPS path: tau per element, the projections and the enthalpy work array
   real(8), allocatable :: tau_stab(:), aux_h(:), ProjMass(:,:), ProjEner(:,:)
   real(8), allocatable :: ProjMX(:,:), ProjMY(:,:), ProjMZ(:,:)
   ! arguments of the reference's interfaces that the device path does not read
   real(8) :: Ngp(nnode,nnode), dNgp(ndime,nnode,nnode), xgp(nnode,ndime), wgp(nnode)
   real(8), allocatable :: coord(:,:), Ml(:)
   character(len=16) :: tag
   logical :: first

   call get_command_argument(1, dir)
   if (len_trim(dir) == 0) then
      write(*,'(a)') 'usage: tgv_host <case directory>'
      error stop 2
   end if
   open(10, file=trim(dir)//'/meta.txt', status='old', action='read')
   read(10,*) nelem, npoin, npoin_w, nsteps, nstage
   read(10,*) gamma_gas, Rgas, Cp, Prt, mu_factor, po, rho0
   read(10,*) dt
   read(10,*) ce, cmax, cglob, flag_norm
   read(10,*,iostat=ios) lps
   if (ios /= 0) lps = 0
   close(10)

   allocate(connec(nelem,nnode), lpoin_w(npoin_w), invAtoIJK(p+1,p+1,p+1))
   allocate(gmshAtoI(nnode), gmshAtoJ(nnode), gmshAtoK(nnode))
   allocate(He(ndime,ndime,nnode,nelem), gpvol(1,nnode,nelem), dlxigp_ip(nnode,ndime,p+1))
   allocate(invMl(npoin), helem_l(nelem,nnode), helem(nelem))
   allocate(rho(npoin,2), q(npoin,ndime,2), E(npoin,2), u(npoin,ndime,2), pr(npoin,2))
   allocate(Tem(npoin,2), e_int(npoin,2), eta(npoin,2), csound(npoin), machno(npoin))
   allocate(mu_fluid(npoin), mu_e(nelem,nnode), mue_l(nelem,nnode), Reta(npoin,2))
   allocate(auxReta(npoin), f_eta(npoin,ndime), Rmass(npoin), Rmom(npoin,ndime), Rener(npoin))
   allocate(rkL(0:nstage,0:nstage), rkG(0:nstage,0:nstage), coord(npoin,ndime), Ml(npoin))
   allocate(tau_stab(nelem), aux_h(npoin), ProjMass(npoin,ndime), ProjEner(npoin,ndime))
   allocate(ProjMX(npoin,ndime), ProjMY(npoin,ndime), ProjMZ(npoin,ndime))
   tau_stab = 0.0d0; aux_h = 0.0d0; ProjMass = 0.0d0; ProjEner = 0.0d0
   ProjMX = 0.0d0; ProjMY = 0.0d0; ProjMZ = 0.0d0
   Ngp = 0.0d0; dNgp = 0.0d0; xgp = 0.0d0; wgp = 0.0d0; coord = 0.0d0; Ml = 0.0d0

   call read_i4(dir, 'connec', connec, size(connec))
   call read_i4(dir, 'lpoin_w', lpoin_w, size(lpoin_w))
   call read_i4(dir, 'invAtoIJK', invAtoIJK, size(invAtoIJK))
   call read_i4(dir, 'gmshAtoI', gmshAtoI, nnode)
   call read_i4(dir, 'gmshAtoJ', gmshAtoJ, nnode)
   call read_i4(dir, 'gmshAtoK', gmshAtoK, nnode)
   call read_r8(dir, 'He', He, size(He))
   call read_r8(dir, 'gpvol', gpvol, size(gpvol))
   call read_r8(dir, 'dlxigp_ip', dlxigp_ip, size(dlxigp_ip))
   call read_r8(dir, 'invMl', invMl, npoin)
   call read_r8(dir, 'helem_l', helem_l, size(helem_l))
   call read_r8(dir, 'helem', helem, nelem)
   call read_r8(dir, 'rkL', rkL, size(rkL))
   call read_r8(dir, 'rkG', rkG, size(rkG))
   ! both levels start from the initial fields
   call read_r8(dir, 'rho', rho(:,1), npoin)
   call read_r8(dir, 'q', q(:,:,1), npoin*ndime)
   call read_r8(dir, 'E', E(:,1), npoin)
   call read_r8(dir, 'u', u(:,:,1), npoin*ndime)
   call read_r8(dir, 'pr', pr(:,1), npoin)
   call read_r8(dir, 'Tem', Tem(:,1), npoin)
   call read_r8(dir, 'e_int', e_int(:,1), npoin)
   call read_r8(dir, 'eta', eta(:,1), npoin)
   call read_r8(dir, 'csound', csound, npoin)
   call read_r8(dir, 'machno', machno, npoin)
   call read_r8(dir, 'mu_fluid', mu_fluid, npoin)
   rho(:,2) = rho(:,1); q(:,:,2) = q(:,:,1); E(:,2) = E(:,1); u(:,:,2) = u(:,:,1)
   pr(:,2) = pr(:,1); Tem(:,2) = Tem(:,1); e_int(:,2) = e_int(:,1); eta(:,2) = eta(:,1)
   mu_e = 0.0d0; mue_l = 0.0d0; Reta = 0.0d0; auxReta = 0.0d0
   Rmass = 0.0d0; Rmom = 0.0d0; Rener = 0.0d0
   do s = 1, ndime
      f_eta(:,s) = u(:,s,1) * eta(:,1)
   end do

   call sod2d_hip_init(0)
   call sod2d_hip_set_visc_params(ce, cmax, cglob, Rgas, flag_norm, .false.)
   call sod2d_hip_enter_data(connec)
   call sod2d_hip_enter_data(lpoin_w)
   call sod2d_hip_enter_data(invAtoIJK)
   call sod2d_hip_enter_data(gmshAtoI)
   call sod2d_hip_enter_data(gmshAtoJ)
   call sod2d_hip_enter_data(gmshAtoK)
   call sod2d_hip_enter_data(He)
   call sod2d_hip_enter_data(gpvol)
   call sod2d_hip_enter_data(dlxigp_ip)
   call sod2d_hip_enter_data(invMl)
   call sod2d_hip_enter_data(helem_l)
   call sod2d_hip_enter_data(helem)
   call sod2d_hip_enter_data(rho)
   call sod2d_hip_enter_data(q)
   call sod2d_hip_enter_data(E)
   call sod2d_hip_enter_data(u)
   call sod2d_hip_enter_data(pr)
   call sod2d_hip_enter_data(Tem)
   call sod2d_hip_enter_data(e_int)
   call sod2d_hip_enter_data(eta)
   call sod2d_hip_enter_data(csound)
   call sod2d_hip_enter_data(machno)
   call sod2d_hip_enter_data(mu_fluid)
   call sod2d_hip_enter_data(mu_e)
   call sod2d_hip_enter_data(mue_l)
   call sod2d_hip_enter_data(Reta)
   call sod2d_hip_enter_data(auxReta)
   call sod2d_hip_enter_data(f_eta)
   call sod2d_hip_enter_data(Rmass)
   call sod2d_hip_enter_data(Rmom)
   call sod2d_hip_enter_data(Rener)
   if (lps /= 0) then
      call sod2d_hip_enter_data(tau_stab)
      call sod2d_hip_enter_data(aux_h)
      call sod2d_hip_enter_data(ProjMass)
      call sod2d_hip_enter_data(ProjEner)
      call sod2d_hip_enter_data(ProjMX)
      call sod2d_hip_enter_data(ProjMY)
      call sod2d_hip_enter_data(ProjMZ)
   end if

   l1 = 1; l2 = 2
   first = .true.
   do istep = 1, nsteps
      ! level 2 -> level 1 (an exchange of the level index) and eta(:,1) = eta(:,2)
      lt = l1; l1 = l2; l2 = lt
      call sod2d_hip_copy(eta(:,2), eta(:,1))
      if (lps /= 0) call comp_tau(nelem, npoin, connec, csound, u(:,:,l1), helem, dt, tau_stab)
      if (first) then
         ! Reta(:,1) of f_eta = u(:,:,1) eta(:,1); the sensor; updateF at y^n
         call generic_scalar_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, &
            xgp, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, f_eta, eta(:,1), u(:,:,l1), Reta(:,1))
         call lumped_solver_scal_opt(npoin, npoin_w, lpoin_w, invMl, Reta(:,1))
         call smart_visc_spectral(nelem, npoin, npoin_w, connec, lpoin_w, Reta(:,1), Ngp, coord, &
            dNgp, gpvol, wgp, gamma_gas, rho(:,l1), u(:,:,l1), csound, Tem(:,l1), eta(:,1), &
            helem_l, helem, Ml, mu_e, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mue_l)
         call stage(0.0d0, 1.0d0, 0.0d0, .true., .false.)
         ! stage 1 with the kept residual (time_integ_ls.f90:323-331)
         call rk_stage_update(npoin, npoin_w, lpoin_w, dt, 1.0d0, 0.0d0, rkG(1,0), rho(:,l1), &
            q(:,:,l1), E(:,l1), rho(:,l2), q(:,:,l2), E(:,l2), Rmass, Rmom, Rener)
         first = .false.
      end if
      do s = 2, nstage
         ! the last stage keeps R and stores the next step's stage 1 in level 1
         call stage(rkL(s,0), rkL(s,s-1), rkG(s,s-1), s == nstage, s == nstage)
      end do
      ! end of step: EOS, entropy and Sutherland; the entropy residual; the sensor
      call eos_entropy_step(npoin, npoin_w, lpoin_w, gamma_gas, Rgas, po, rho0, mu_factor, &
         rho(:,l2), q(:,:,l2), E(:,l2), u(:,:,l2), e_int(:,l2), pr(:,l2), csound, machno, &
         Tem(:,l2), eta, mu_fluid)
      call entropy_convec_residual(nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, dlxigp_ip, &
         invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, u(:,:,l1), eta, invMl, 1.0d0, dt, Reta, auxReta)
      call smart_visc_spectral(nelem, npoin, npoin_w, connec, lpoin_w, auxReta, Ngp, coord, &
         dNgp, gpvol, wgp, gamma_gas, rho(:,l2), u(:,:,l2), csound, Tem(:,l2), eta(:,2), &
         helem_l, helem, Ml, mu_e, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mue_l)
      ! this step's conserved fields and mu_e back to the host
      call sod2d_hip_update_host(rho(:,l2))
      call sod2d_hip_update_host(q(:,:,l2))
      call sod2d_hip_update_host(E(:,l2))
      call sod2d_hip_update_host(mu_e)
      write(tag,'(a,i0,a)') 'step', istep, '_'
      call write_r8(dir, trim(tag)//'rho', rho(:,l2), npoin)
      call write_r8(dir, trim(tag)//'q', q(:,:,l2), npoin*ndime)
      call write_r8(dir, trim(tag)//'E', E(:,l2), npoin)
      call write_r8(dir, trim(tag)//'mu_e', mu_e, nelem*nnode)
   end do
   call sod2d_hip_finalize()
   write(*,'(a,i0,a)') 'tgv_host: ', nsteps, ' steps done'

contains

   ! one updateF, lumped solve and stage combine (rk_stage_fused, or
   ! rk_stage_lps_fused with the LPS stabilisation) from level l1 (y^n) and level
   ! l2 (the stage state) into level l2; keep_R stores the residual, next also
   ! writes the next step's stage 1 into level l1
   subroutine stage(lam1, lams, gam, keep_R, next)
      real(8), intent(in) :: lam1, lams, gam
      logical, intent(in) :: keep_R, next
      if (lps /= 0) then
         if (next) then
            call rk_stage_lps_fused(nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, &
               dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mu_e, invMl, tau_stab, &
               gamma_gas, Rgas, Cp, Prt, mu_factor, dt, lam1, lams, gam, rho(:,l1), q(:,:,l1), &
               E(:,l1), rho(:,l2), q(:,:,l2), E(:,l2), u(:,:,l2), e_int(:,l2), pr(:,l2), aux_h, &
               Tem(:,l2), mu_fluid, ProjMass, ProjEner, ProjMX, ProjMY, ProjMZ, Rmass, Rmom, &
               Rener, rkG(1,0), rho(:,l1), q(:,:,l1), E(:,l1))
         else if (keep_R) then
            call rk_stage_lps_fused(nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, &
               dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mu_e, invMl, tau_stab, &
               gamma_gas, Rgas, Cp, Prt, mu_factor, dt, lam1, lams, gam, rho(:,l1), q(:,:,l1), &
               E(:,l1), rho(:,l2), q(:,:,l2), E(:,l2), u(:,:,l2), e_int(:,l2), pr(:,l2), aux_h, &
               Tem(:,l2), mu_fluid, ProjMass, ProjEner, ProjMX, ProjMY, ProjMZ, Rmass, Rmom, Rener)
         else
            call rk_stage_lps_fused(nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, &
               dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mu_e, invMl, tau_stab, &
               gamma_gas, Rgas, Cp, Prt, mu_factor, dt, lam1, lams, gam, rho(:,l1), q(:,:,l1), &
               E(:,l1), rho(:,l2), q(:,:,l2), E(:,l2), u(:,:,l2), e_int(:,l2), pr(:,l2), aux_h, &
               Tem(:,l2), mu_fluid, ProjMass, ProjEner, ProjMX, ProjMY, ProjMZ)
         end if
      else if (next) then
         call rk_stage_fused(nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, dlxigp_ip, &
            invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mu_e, invMl, gamma_gas, Rgas, Cp, Prt, &
            mu_factor, dt, lam1, lams, gam, rho(:,l1), q(:,:,l1), E(:,l1), rho(:,l2), &
            q(:,:,l2), E(:,l2), Rmass, Rmom, Rener, rkG(1,0), rho(:,l1), q(:,:,l1), E(:,l1))
      else if (keep_R) then
         call rk_stage_fused(nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, dlxigp_ip, &
            invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mu_e, invMl, gamma_gas, Rgas, Cp, Prt, &
            mu_factor, dt, lam1, lams, gam, rho(:,l1), q(:,:,l1), E(:,l1), rho(:,l2), &
            q(:,:,l2), E(:,l2), Rmass, Rmom, Rener)
      else
         call rk_stage_fused(nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, dlxigp_ip, &
            invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mu_e, invMl, gamma_gas, Rgas, Cp, Prt, &
            mu_factor, dt, lam1, lams, gam, rho(:,l1), q(:,:,l1), E(:,l1), rho(:,l2), &
            q(:,:,l2), E(:,l2))
      end if
   end subroutine stage

   subroutine read_i4(d, name, a, n)
      character(*), intent(in) :: d, name
      integer(4), intent(in) :: n
      integer(4), intent(out) :: a(n)
      open(11, file=trim(d)//'/'//name//'.bin', access='stream', form='unformatted', &
           status='old', action='read')
      read(11) a
      close(11)
   end subroutine read_i4

   subroutine read_r8(d, name, a, n)
      character(*), intent(in) :: d, name
      integer(4), intent(in) :: n
      real(8), intent(out) :: a(n)
      open(11, file=trim(d)//'/'//name//'.bin', access='stream', form='unformatted', &
           status='old', action='read')
      read(11) a
      close(11)
   end subroutine read_r8

   subroutine write_r8(d, name, a, n)
      character(*), intent(in) :: d, name
      integer(4), intent(in) :: n
      real(8), intent(in) :: a(n)
      open(12, file=trim(d)//'/'//name//'.bin', access='stream', form='unformatted', &
           status='replace', action='write')
      write(12) a
      close(12)
   end subroutine write_r8

end program tgv_host

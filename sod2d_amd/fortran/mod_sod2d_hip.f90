! mod_sod2d_hip.f90 -- ISO_C_BINDING layer over libsod2d_hip (include/sod2d_hip.h).
!
! The subroutines full_convec_ijk, full_diffusion_ijk, generic_scalar_convec_ijk
! (elem_convec.f90:227-594, elem_diffu.f90:12-223), smart_visc_spectral
! (mod_entropy_viscosity.f90:14-111) and lumped_solver_scal_opt /
! lumped_solver_vect_opt (mod_solver.f90:40-104) keep the reference's dummy
! argument lists (names, order, types, kinds, shapes, intents and optional
! arguments), so a caller switches to the device path by using this module in
! place of elem_convec / elem_diffu / mod_entropy_viscosity / mod_solver.
! rk_stage_update is the stage combine of rk_4_ls_main
! (time_integ_ls.f90:323-361), which the reference writes inline, with the
! lumped-mass inverse fused in.
!
! Arrays are host arrays that have been mapped with sod2d_hip_enter_data (the
! equivalent of OpenACC's enter data): each subroutine looks up their device
! mirrors and runs on them; results stay on the device until
! sod2d_hip_update_host.  A missing mapping or a failed launch stops the
! program with the library's message, as the reference stops on errors.
!
! The element order is fixed at compile time (-D__PORDER__, as the reference),
! and so is the real kind rp of the reference's real(rp) arguments
! (-D__RP__=8 or 4, the reference's USE_RP, CMakeLists.txt:30; default 8).
! With rp = 4 the operators, the sensor, the lumped solves and the stage
! combine take real(4) arrays (the C-ABI's _f32 entry points); the live-path
! and LPS operators exist in the rp = 8 build only.  The fused stages keep the
! conserved state, the lumped-mass inverse and the combine in real(8) in both
! builds; their operator arrays (He, gpvol, dlxigp_ip, mu_e) are real(rp).
#ifndef __RP__
#define __RP__ 8
#endif
#if __RP__ == 4
#define RPSFX '_f32'
#define C_RP c_float
#else
#define RPSFX '_f64'
#define C_RP c_double
#endif
module mod_sod2d_hip
   use iso_c_binding
   implicit none
   private

   integer(4), parameter :: rp = __RP__
   integer(4), parameter :: ndime = 3
   integer(4), parameter :: porder = __PORDER__
   integer(4), parameter :: nnode = (porder+1)**3
   integer(4), parameter :: ngaus = nnode

   public :: sod2d_hip_init, sod2d_hip_finalize, sod2d_hip_synchronize
   public :: sod2d_hip_enter_data, sod2d_hip_update_device, sod2d_hip_update_host
   public :: sod2d_hip_exit_data, sod2d_hip_set_visc_params, sod2d_hip_copy
   public :: full_convec_ijk, full_diffusion_ijk, generic_scalar_convec_ijk
   public :: smart_visc_spectral, lumped_solver_scal_opt, lumped_solver_vect_opt
   public :: rk_stage_update
#if __RP__ == 8
   public :: smart_visc_spectral_imex, eval_divergence, compute_vorticity, ducros_visc_limit
   public :: comp_tau, full_proj_ijk, full_diff_stab_ijk
#endif
   public :: sod2d_hip_set_live_params
   public :: rk_stage_fused, rk_stage_lps_fused, eos_entropy_step, entropy_convec_residual
   public :: rk4_stage_fused, rk4_stage_update

   type, bind(C) :: visc_params_t
      real(c_double) :: ce, cmax, cglob, Rgas
      integer(c_int) :: flag_normalise_entropy, flag_drop_c_in_envit
   end type visc_params_t

   ! sodhip_stage_params: the physical constants of a fused stage
   type, bind(C) :: stage_params_t
      real(c_double) :: gamma_gas, Rgas, Cp, Pr, mu_factor, c_rho, c_ener, fact, p_ref
      ! per-element Sutherland factor (device), or none: the uniform mu_factor
      type(c_ptr) :: mu_factor_e = c_null_ptr
   end type stage_params_t

   type(c_ptr), save :: ctx = c_null_ptr
   ! module globals the reference's sensor reads from mod_numerical_params,
   ! with the reference's defaults (mod_numerical_params.f90:29-41)
   type(visc_params_t), save :: vprm = visc_params_t(0.1d0, 0.5d0, 1.0d0, 1.0d0, 1, 0)
   ! the reference's optional `fact` keeps the value of the last call that
   ! passed it (a local with an initialiser is SAVE'd)
   real(rp), save :: fact_convec = 1.0_rp, fact_diffu = 1.0_rp, fact_scalar = 1.0_rp
   ! likewise fact_low_mach of smart_visc_spectral, once dropped, stays dropped
   logical, save :: low_mach_dropped = .false.
   ! module globals of the live-path sensor (mod_numerical_params.f90:36,172-173
   ! and the case's flag_high_mach); smart_visc_spectral_imex's high-Mach
   ! normalisation sets the module's ce to ce_comp for every later call
   real(rp), save :: ce_comp = 1.0_rp, ducros_min_val = 0.1_rp
   logical, save :: flag_high_mach = .false.
   ! LPS constants (mod_numerical_params.f90:31,46): c_lps_comp is default real
   real, save :: c_lps_comp = 0.1
   real(rp), save :: c_species_stab = 0.5_rp
   real(rp), save :: fact_diffstab = 1.0_rp

   interface sod2d_hip_enter_data
      module procedure enter_r8, enter_i4, enter_r4
   end interface
   interface sod2d_hip_update_device
      module procedure upd_dev_r8, upd_dev_i4, upd_dev_r4
   end interface
   interface sod2d_hip_update_host
      module procedure upd_host_r8, upd_host_i4, upd_host_r4
   end interface
   interface sod2d_hip_exit_data
      module procedure exit_r8, exit_i4, exit_r4
   end interface

   interface
      integer(c_int) function c_create(device, c) bind(C, name='sodhip_create')
         import :: c_int, c_ptr
         integer(c_int), value :: device
         type(c_ptr) :: c
      end function
      integer(c_int) function c_destroy(c) bind(C, name='sodhip_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: c
      end function
      integer(c_int) function c_synchronize(c) bind(C, name='sodhip_synchronize')
         import :: c_int, c_ptr
         type(c_ptr), value :: c
      end function
      integer(c_int) function c_enter_data(c, host, bytes, copyin) bind(C, name='sodhip_enter_data')
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: c, host
         integer(c_size_t), value :: bytes
         integer(c_int), value :: copyin
      end function
      integer(c_int) function c_update_device(c, host, bytes) bind(C, name='sodhip_update_device')
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: c, host
         integer(c_size_t), value :: bytes
      end function
      integer(c_int) function c_update_host(c, host, bytes) bind(C, name='sodhip_update_host')
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: c, host
         integer(c_size_t), value :: bytes
      end function
      integer(c_int) function c_exit_data(c, host, copyout) bind(C, name='sodhip_exit_data')
         import :: c_int, c_ptr
         type(c_ptr), value :: c, host
         integer(c_int), value :: copyout
      end function
      integer(c_int) function c_copy_present(c, src, dst, bytes) bind(C, name='sodhip_copy_present')
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: c, src, dst
         integer(c_size_t), value :: bytes
      end function
      type(c_ptr) function c_deviceptr(c, host) bind(C, name='sodhip_deviceptr')
         import :: c_ptr
         type(c_ptr), value :: c
         type(*), dimension(*), intent(in) :: host
      end function
      integer(c_int) function c_last_error(buf, n) bind(C, name='sodhip_last_error')
         import :: c_int, c_char
         character(kind=c_char), dimension(*) :: buf
         integer(c_int), value :: n
      end function
      integer(c_int) function c_full_convec(c, p, nelem, npoin, connec, He, gpvol, dl, inv, gI, &
            gJ, gK, u, q, rho, pr, E, Rmass, Rmom, Rener, init, fact) &
            bind(C, name='sodhip_full_convec_ijk'//RPSFX)
         import :: c_int, c_ptr, c_float, c_double
         type(c_ptr), value :: c, connec, He, gpvol, dl, inv, gI, gJ, gK, u, q, rho, pr, E
         type(c_ptr), value :: Rmass, Rmom, Rener
         integer(c_int), value :: p, nelem, npoin, init
         real(C_RP), value :: fact
      end function
      integer(c_int) function c_full_diffusion(c, p, nelem, npoin, connec, He, gpvol, dl, inv, &
            gI, gJ, gK, Cp, Pr, rho_n, rho, u, Tem, mu_fluid, mu_e, mu_sgs, c_rho, c_ener, &
            Rmass, Rmom, Rener, init, fact) bind(C, name='sodhip_full_diffusion_ijk'//RPSFX)
         import :: c_int, c_ptr, c_float, c_double
         type(c_ptr), value :: c, connec, He, gpvol, dl, inv, gI, gJ, gK, rho_n, rho, u, Tem
         type(c_ptr), value :: mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener
         integer(c_int), value :: p, nelem, npoin, init
         real(C_RP), value :: Cp, Pr, c_rho, c_ener, fact
      end function
      integer(c_int) function c_comp_tau(c, p, nelem, npoin, connec, csound, u, helem_k, &
            c_species, drop_c, tau) bind(C, name='sodhip_comp_tau_f64')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c, connec, csound, u, helem_k, tau
         integer(c_int), value :: p, nelem, npoin, drop_c
         real(c_double), value :: c_species
      end function
      integer(c_int) function c_full_proj(c, p, nelem, npoin, npoin_w, connec, lpoin_w, He, &
            gpvol, dl, inv, gI, gJ, gK, rho, u, Tem, invMl, PMass, PEner, PMX, PMY, PMZ) &
            bind(C, name='sodhip_full_proj_ijk_f64')
         import :: c_int, c_ptr
         type(c_ptr), value :: c, connec, lpoin_w, He, gpvol, dl, inv, gI, gJ, gK, rho, u, Tem
         type(c_ptr), value :: invMl, PMass, PEner, PMX, PMY, PMZ
         integer(c_int), value :: p, nelem, npoin, npoin_w
      end function
      integer(c_int) function c_diff_stab(c, p, nelem, npoin, connec, He, gpvol, dl, inv, gI, &
            gJ, gK, Cp, Pr, rho_n, rho, u, Tem, mu_fluid, mu_e, mu_sgs, PMass, PEner, PMX, PMY, &
            PMZ, tau, c_rho, c_ener, c_lps, Rmass, Rmom, Rener, init, fact) &
            bind(C, name='sodhip_full_diff_stab_ijk_f64')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c, connec, He, gpvol, dl, inv, gI, gJ, gK, rho_n, rho, u, Tem
         type(c_ptr), value :: mu_fluid, mu_e, mu_sgs, PMass, PEner, PMX, PMY, PMZ, tau
         type(c_ptr), value :: Rmass, Rmom, Rener
         integer(c_int), value :: p, nelem, npoin, init
         real(c_double), value :: Cp, Pr, c_rho, c_ener, c_lps, fact
      end function
      integer(c_int) function c_scalar_convec(c, p, nelem, npoin, connec, He, gpvol, dl, inv, gI, &
            gJ, gK, q, eta, u, Rconvec, init, fact) &
            bind(C, name='sodhip_generic_scalar_convec_ijk'//RPSFX)
         import :: c_int, c_ptr, c_float, c_double
         type(c_ptr), value :: c, connec, He, gpvol, dl, inv, gI, gJ, gK, q, eta, u, Rconvec
         integer(c_int), value :: p, nelem, npoin, init
         real(C_RP), value :: fact
      end function
      integer(c_int) function c_smart_visc(c, p, nelem, npoin, npoin_w, connec, lpoin_w, Reta, &
            gamma_gas, rho, u, Tem, eta, helem, helem_k, inv, prm, mu_e, mue_l) &
            bind(C, name='sodhip_smart_visc_spectral'//RPSFX)
         import :: c_int, c_ptr, c_float, c_double, visc_params_t
         type(c_ptr), value :: c, connec, lpoin_w, Reta, rho, u, Tem, eta, helem, helem_k, inv
         type(c_ptr), value :: mu_e, mue_l
         integer(c_int), value :: p, nelem, npoin, npoin_w
         real(C_RP), value :: gamma_gas
         type(visc_params_t), intent(in) :: prm
      end function
      integer(c_int) function c_lumped_scal(c, npoin, npoin_w, lpoin_w, invMl, R) &
            bind(C, name='sodhip_lumped_solver_scal'//RPSFX)
         import :: c_int, c_ptr
         type(c_ptr), value :: c, lpoin_w, invMl, R
         integer(c_int), value :: npoin, npoin_w
      end function
      integer(c_int) function c_lumped_vect(c, npoin, npoin_w, lpoin_w, invMl, R) &
            bind(C, name='sodhip_lumped_solver_vect'//RPSFX)
         import :: c_int, c_ptr
         type(c_ptr), value :: c, lpoin_w, invMl, R
         integer(c_int), value :: npoin, npoin_w
      end function
      integer(c_int) function c_rk_update(c, npoin, npoin_w, lpoin_w, invMl, dt, lam1, lams, gam, &
            rho1, q1, E1, rho, q, E, Rmass, Rmom, Rener) &
            bind(C, name='sodhip_rk_stage_update'//RPSFX)
         import :: c_int, c_ptr, c_float, c_double
         type(c_ptr), value :: c, lpoin_w, invMl, rho1, q1, E1, rho, q, E, Rmass, Rmom, Rener
         integer(c_int), value :: npoin, npoin_w
         real(C_RP), value :: dt, lam1, lams, gam
      end function
      integer(c_int) function c_visc_imex(c, p, nelem, npoin, npoin_w, connec, lpoin_w, Reta, &
            rho, u, csound, eta, helem_k, prm, ce_comp, high_mach, mu_e, mue_l) &
            bind(C, name='sodhip_smart_visc_spectral_imex_f64')
         import :: c_int, c_ptr, c_double, visc_params_t
         type(c_ptr), value :: c, connec, lpoin_w, Reta, rho, u, csound, eta, helem_k, mu_e, mue_l
         integer(c_int), value :: p, nelem, npoin, npoin_w, high_mach
         real(c_double), value :: ce_comp
         type(visc_params_t), intent(inout) :: prm
      end function
      integer(c_int) function c_eval_div(c, p, nelem, npoin, connec, He, gpvol, dl, inv, gI, gJ, &
            gK, u, divU) bind(C, name='sodhip_eval_divergence_f64')
         import :: c_int, c_ptr
         type(c_ptr), value :: c, connec, He, gpvol, dl, inv, gI, gJ, gK, u, divU
         integer(c_int), value :: p, nelem, npoin
      end function
      integer(c_int) function c_vorticity(c, p, nelem, npoin, npoin_w, lpoin_w, connec, lelpn, &
            He, dl, inv, gI, gJ, gK, u, curlU) bind(C, name='sodhip_compute_vorticity_f64')
         import :: c_int, c_ptr
         type(c_ptr), value :: c, lpoin_w, connec, lelpn, He, dl, inv, gI, gJ, gK, u, curlU
         integer(c_int), value :: p, nelem, npoin, npoin_w
      end function
      integer(c_int) function c_ducros(c, p, nelem, npoin, connec, u, divU, vorti, Ml, helem, &
            dmin, mu_e) bind(C, name='sodhip_ducros_visc_limit_f64')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c, connec, u, divU, vorti, Ml, helem, mu_e
         integer(c_int), value :: p, nelem, npoin
         real(c_double), value :: dmin
      end function
      integer(c_int) function c_rk_stage(c, p, nelem, npoin, npoin_w, connec, lpoin_w, opp, He, &
            gpvol, dl, inv, gI, gJ, gK, mu_e, invMl, prm, dt, lam1, lams, gam, rho1, q1, E1, &
            rho, q, E, Rmass, Rmom, Rener) bind(C, name='sodhip_rk_stage_f64')
         import :: c_int, c_ptr, c_double, stage_params_t
         type(c_ptr), value :: c, connec, lpoin_w, He, gpvol, dl, inv, gI, gJ, gK, mu_e, invMl
         type(c_ptr), value :: rho1, q1, E1, rho, q, E, Rmass, Rmom, Rener
         integer(c_int), value :: p, nelem, npoin, npoin_w, opp
         type(stage_params_t), intent(in) :: prm
         real(c_double), value :: dt, lam1, lams, gam
      end function
      integer(c_int) function c_rk_stage_next(c, p, nelem, npoin, npoin_w, connec, lpoin_w, opp, &
            He, gpvol, dl, inv, gI, gJ, gK, mu_e, invMl, prm, dt, lam1, lams, gam, rho1, q1, E1, &
            rho, q, E, Rmass, Rmom, Rener, gam_next, rhoN, qN, EN) &
            bind(C, name='sodhip_rk_stage_next_f64')
         import :: c_int, c_ptr, c_double, stage_params_t
         type(c_ptr), value :: c, connec, lpoin_w, He, gpvol, dl, inv, gI, gJ, gK, mu_e, invMl
         type(c_ptr), value :: rho1, q1, E1, rho, q, E, Rmass, Rmom, Rener, rhoN, qN, EN
         integer(c_int), value :: p, nelem, npoin, npoin_w, opp
         type(stage_params_t), intent(in) :: prm
         real(c_double), value :: dt, lam1, lams, gam, gam_next
      end function
      integer(c_int) function c_rk_stage_lps(c, p, nelem, npoin, npoin_w, connec, lpoin_w, He, &
            gpvol, dl, inv, gI, gJ, gK, mu_e, invMl, tau, prm, c_lps, dt, lam1, lams, gam, rho1, &
            q1, E1, rho, q, E, u, e_int, pr, aux_h, Tem, mu_fluid, PM, PE, PX, PY, PZ, Rmass, &
            Rmom, Rener, gam_next, rhoN, qN, EN) bind(C, name='sodhip_rk_stage_lps_f64')
         import :: c_int, c_ptr, c_double, stage_params_t
         type(c_ptr), value :: c, connec, lpoin_w, He, gpvol, dl, inv, gI, gJ, gK, mu_e, invMl, tau
         type(c_ptr), value :: rho1, q1, E1, rho, q, E, u, e_int, pr, aux_h, Tem, mu_fluid
         type(c_ptr), value :: PM, PE, PX, PY, PZ, Rmass, Rmom, Rener, rhoN, qN, EN
         integer(c_int), value :: p, nelem, npoin, npoin_w
         type(stage_params_t), intent(in) :: prm
         real(c_double), value :: c_lps, dt, lam1, lams, gam, gam_next
      end function
      integer(c_int) function c_eos_entropy_step(c, npoin, npoin_w, lpoin_w, g, Rgas, p_inf, &
            rho_inf, mu_factor, rho, q, E, u, e_int, pr, csound, machno, Tem, eta, mu_fluid) &
            bind(C, name='sodhip_eos_entropy_step_f64')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c, lpoin_w, rho, q, E, u, e_int, pr, csound, machno, Tem, eta
         type(c_ptr), value :: mu_fluid
         integer(c_int), value :: npoin, npoin_w
         real(c_double), value :: g, Rgas, p_inf, rho_inf, mu_factor
      end function
      integer(c_int) function c_entropy_convec_residual(c, p, nelem, npoin, npoin_w, connec, &
            lpoin_w, He, gpvol, dl, inv, gI, gJ, gK, u, eta, invMl, factor_comp, dt, Reta, &
            auxReta) bind(C, name='sodhip_entropy_convec_residual_f64')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c, connec, lpoin_w, He, gpvol, dl, inv, gI, gJ, gK, u, eta, invMl
         type(c_ptr), value :: Reta, auxReta
         integer(c_int), value :: p, nelem, npoin, npoin_w
         real(c_double), value :: factor_comp, dt
      end function
      integer(c_int) function c_rk4_stage(c, p, nelem, npoin, npoin_w, connec, lpoin_w, opp, He, &
            gpvol, dl, inv, gI, gJ, gK, mu_e, invMl, prm, dt, a_next, b, first, last, rho, q, E, &
            srho, sq, sE, arho, aq, aE, Smass, Smom, Sener, Rmass, Rmom, Rener) &
            bind(C, name='sodhip_rk4_stage_f64')
         import :: c_int, c_ptr, c_double, stage_params_t
         type(c_ptr), value :: c, connec, lpoin_w, He, gpvol, dl, inv, gI, gJ, gK, mu_e, invMl
         type(c_ptr), value :: rho, q, E, srho, sq, sE, arho, aq, aE, Smass, Smom, Sener
         type(c_ptr), value :: Rmass, Rmom, Rener
         integer(c_int), value :: p, nelem, npoin, npoin_w, opp, first, last
         type(stage_params_t), intent(in) :: prm
         real(c_double), value :: dt, a_next, b
      end function
      integer(c_int) function c_rk4_update(c, npoin, npoin_w, lpoin_w, invMl, dt, a_next, b, &
            first, last, rho, q, E, arho, aq, aE, Rmass, Rmom, Rener, Smass, Smom, Sener) &
            bind(C, name='sodhip_rk4_stage_update_f64')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: c, lpoin_w, invMl, rho, q, E, arho, aq, aE, Rmass, Rmom, Rener
         type(c_ptr), value :: Smass, Smom, Sener
         integer(c_int), value :: npoin, npoin_w, first, last
         real(c_double), value :: dt, a_next, b
      end function
   end interface

contains

   ! ------------------------------------------------------------ context
   subroutine sod2d_hip_init(device)
      integer(4), intent(in), optional :: device
      integer(c_int) :: dev
      dev = 0
      if (present(device)) dev = device
      if (c_associated(ctx)) return
      call check(c_create(dev, ctx), 'sodhip_create')
   end subroutine sod2d_hip_init

   subroutine sod2d_hip_finalize()
      if (.not. c_associated(ctx)) return
      call check(c_destroy(ctx), 'sodhip_destroy')
      ctx = c_null_ptr
   end subroutine sod2d_hip_finalize

   subroutine sod2d_hip_synchronize()
      call check(c_synchronize(ctx), 'sodhip_synchronize')
   end subroutine sod2d_hip_synchronize

   ! Module globals of mod_numerical_params read by smart_visc_spectral:
   ! ce, cmax, cglob, nscbc_Rgas_inf, flag_normalise_entropy, flag_drop_c_in_envit
   subroutine sod2d_hip_set_visc_params(ce, cmax, cglob, Rgas, flag_normalise_entropy, &
         flag_drop_c_in_envit)
      real(rp), intent(in) :: ce, cmax, cglob, Rgas
      integer(4), intent(in) :: flag_normalise_entropy
      logical, intent(in) :: flag_drop_c_in_envit
      vprm%ce = ce
      vprm%cmax = cmax
      vprm%cglob = cglob
      vprm%Rgas = Rgas
      vprm%flag_normalise_entropy = flag_normalise_entropy
      if (flag_drop_c_in_envit) low_mach_dropped = .true.
   end subroutine sod2d_hip_set_visc_params

   ! ------------------------------------------------------------ present table
   subroutine enter_r8(a, copyin)
      real(c_double), dimension(..), intent(in), target :: a
      logical, intent(in), optional :: copyin
      integer(c_int) :: ci
      ci = 1
      if (present(copyin)) then
         if (.not. copyin) ci = 0
      end if
      call check(c_enter_data(ctx, c_loc(a), size(a, kind=c_size_t) * 8_c_size_t, ci), &
                 'sodhip_enter_data')
   end subroutine enter_r8

   subroutine enter_r4(a, copyin)
      real(c_float), dimension(..), intent(in), target :: a
      logical, intent(in), optional :: copyin
      integer(c_int) :: ci
      ci = 1
      if (present(copyin)) then
         if (.not. copyin) ci = 0
      end if
      call check(c_enter_data(ctx, c_loc(a), size(a, kind=c_size_t) * 4_c_size_t, ci), &
                 'sodhip_enter_data')
   end subroutine enter_r4

   subroutine enter_i4(a, copyin)
      integer(c_int), dimension(..), intent(in), target :: a
      logical, intent(in), optional :: copyin
      integer(c_int) :: ci
      ci = 1
      if (present(copyin)) then
         if (.not. copyin) ci = 0
      end if
      call check(c_enter_data(ctx, c_loc(a), size(a, kind=c_size_t) * 4_c_size_t, ci), &
                 'sodhip_enter_data')
   end subroutine enter_i4

   subroutine upd_dev_r8(a)
      real(c_double), dimension(..), intent(in), target :: a
      call check(c_update_device(ctx, c_loc(a), size(a, kind=c_size_t) * 8_c_size_t), &
                 'sodhip_update_device')
   end subroutine upd_dev_r8

   subroutine upd_dev_r4(a)
      real(c_float), dimension(..), intent(in), target :: a
      call check(c_update_device(ctx, c_loc(a), size(a, kind=c_size_t) * 4_c_size_t), &
                 'sodhip_update_device')
   end subroutine upd_dev_r4

   subroutine upd_dev_i4(a)
      integer(c_int), dimension(..), intent(in), target :: a
      call check(c_update_device(ctx, c_loc(a), size(a, kind=c_size_t) * 4_c_size_t), &
                 'sodhip_update_device')
   end subroutine upd_dev_i4

   subroutine upd_host_r8(a)
      real(c_double), dimension(..), intent(inout), target :: a
      call check(c_update_host(ctx, c_loc(a), size(a, kind=c_size_t) * 8_c_size_t), &
                 'sodhip_update_host')
   end subroutine upd_host_r8

   ! dst = src on the device for mapped arrays of equal size (an `!$acc
   ! kernels` assignment on present data, e.g. eta(:,1) = eta(:,2))
   subroutine sod2d_hip_copy(src, dst)
      real(c_double), dimension(..), intent(in), target :: src
      real(c_double), dimension(..), intent(inout), target :: dst
      if (size(src) /= size(dst)) then
         write(*,'(a)') 'mod_sod2d_hip: sod2d_hip_copy of arrays of different sizes'
         error stop 1
      end if
      call check(c_copy_present(ctx, c_loc(src), c_loc(dst), size(src, kind=c_size_t) * 8_c_size_t), &
                 'sodhip_copy_present')
   end subroutine sod2d_hip_copy

   subroutine upd_host_r4(a)
      real(c_float), dimension(..), intent(inout), target :: a
      call check(c_update_host(ctx, c_loc(a), size(a, kind=c_size_t) * 4_c_size_t), &
                 'sodhip_update_host')
   end subroutine upd_host_r4

   subroutine upd_host_i4(a)
      integer(c_int), dimension(..), intent(inout), target :: a
      call check(c_update_host(ctx, c_loc(a), size(a, kind=c_size_t) * 4_c_size_t), &
                 'sodhip_update_host')
   end subroutine upd_host_i4

   subroutine exit_r8(a, copyout)
      real(c_double), dimension(..), intent(inout), target :: a
      logical, intent(in), optional :: copyout
      integer(c_int) :: co
      co = 0
      if (present(copyout)) then
         if (copyout) co = 1
      end if
      call check(c_exit_data(ctx, c_loc(a), co), 'sodhip_exit_data')
   end subroutine exit_r8

   subroutine exit_r4(a, copyout)
      real(c_float), dimension(..), intent(inout), target :: a
      logical, intent(in), optional :: copyout
      integer(c_int) :: co
      co = 0
      if (present(copyout)) then
         if (copyout) co = 1
      end if
      call check(c_exit_data(ctx, c_loc(a), co), 'sodhip_exit_data')
   end subroutine exit_r4

   subroutine exit_i4(a, copyout)
      integer(c_int), dimension(..), intent(inout), target :: a
      logical, intent(in), optional :: copyout
      integer(c_int) :: co
      co = 0
      if (present(copyout)) then
         if (copyout) co = 1
      end if
      call check(c_exit_data(ctx, c_loc(a), co), 'sodhip_exit_data')
   end subroutine exit_i4

   ! ------------------------------------------------------------ helpers
   subroutine check(rc, what)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: what
      character(kind=c_char) :: buf(512)
      integer(c_int) :: n, i
      character(len=512) :: msg
      if (rc == 0) return
      n = c_last_error(buf, 512)
      msg = ' '
      do i = 1, 512
         if (buf(i) == c_null_char) exit
         msg(i:i) = buf(i)
      end do
      write(*,'(a,a,a,i0,a,a)') 'mod_sod2d_hip: ', what, ' failed (', rc, '): ', trim(msg)
      error stop 1
   end subroutine check

   ! Module globals of the live-path sensor: ce_comp, ducros_min_val
   ! (mod_numerical_params.f90:36,173) and the case's flag_high_mach
   subroutine sod2d_hip_set_live_params(ce_comp_in, flag_high_mach_in, ducros_min_val_in)
      real(rp), intent(in) :: ce_comp_in, ducros_min_val_in
      logical, intent(in) :: flag_high_mach_in
      ce_comp = ce_comp_in
      flag_high_mach = flag_high_mach_in
      ducros_min_val = ducros_min_val_in
   end subroutine sod2d_hip_set_live_params

   type(c_ptr) function dev(host, what)
      type(*), dimension(*), intent(in) :: host
      character(*), intent(in) :: what
      if (.not. c_associated(ctx)) then
         write(*,'(a)') 'mod_sod2d_hip: sod2d_hip_init was not called'
         error stop 1
      end if
      dev = c_deviceptr(ctx, host)
      if (.not. c_associated(dev)) then
         write(*,'(a,a,a)') 'mod_sod2d_hip: argument ', what, ' is not mapped to the device'
         error stop 1
      end if
   end function dev

   integer(c_int) function init_flag(initialze)
      logical, optional, intent(in) :: initialze
      init_flag = 1
      if (present(initialze)) then
         if (.not. initialze) init_flag = 0
      end if
   end function init_flag

   ! ------------------------------------------------------------ operators
   ! elem_convec.f90:227-470.  Ngp, dNgp and xgp are part of the reference's
   ! interface and are not read.
   subroutine full_convec_ijk(nelem,npoin,connec,Ngp,dNgp,He,gpvol,dlxigp_ip,xgp,invAtoIJK, &
         gmshAtoI,gmshAtoJ,gmshAtoK,u,q,rho,pr,E,Rmass,Rmom,Rener,initialze,fact)
      integer(4), intent(in)  :: nelem, npoin
      integer(4), intent(in)  :: connec(nelem,nnode)
      real(rp),    intent(in)  :: Ngp(ngaus,nnode), dNgp(ndime,nnode,ngaus)
      real(rp),    intent(in)  :: He(ndime,ndime,ngaus,nelem),xgp(ngaus,ndime)
      real(rp),    intent(in)  :: dlxigp_ip(ngaus,ndime,porder+1)
      real(rp),    intent(in)  :: gpvol(1,ngaus,nelem)
      integer(4), intent(in)  :: invAtoIJK(porder+1,porder+1,porder+1),gmshAtoI(nnode)
      integer(4), intent(in)  :: gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp),    intent(in)  :: q(npoin,ndime), u(npoin,ndime), rho(npoin),pr(npoin), E(npoin)
      real(rp),    intent(inout) :: Rmass(npoin)
      real(rp),    intent(inout) :: Rmom(npoin,ndime)
      real(rp),    intent(inout) :: Rener(npoin)
      logical, optional, intent(in)    :: initialze
      real(rp), optional, intent(in)  :: fact
      if (present(fact)) fact_convec = fact
      call check(c_full_convec(ctx, porder, nelem, npoin, dev(connec,'connec'), dev(He,'He'), &
         dev(gpvol,'gpvol'), dev(dlxigp_ip,'dlxigp_ip'), dev(invAtoIJK,'invAtoIJK'), &
         dev(gmshAtoI,'gmshAtoI'), dev(gmshAtoJ,'gmshAtoJ'), dev(gmshAtoK,'gmshAtoK'), &
         dev(u,'u'), dev(q,'q'), dev(rho,'rho'), dev(pr,'pr'), dev(E,'E'), dev(Rmass,'Rmass'), &
         dev(Rmom,'Rmom'), dev(Rener,'Rener'), init_flag(initialze), fact_convec), &
         'full_convec_ijk')
   end subroutine full_convec_ijk

   ! elem_diffu.f90:12-223.  Ngp and Ml are part of the reference's interface
   ! and are not read; c_rho and c_ener take the reference's defaults (1).
   subroutine full_diffusion_ijk(nelem,npoin,connec,Ngp,He,gpvol,dlxigp_ip,invAtoIJK,gmshAtoI, &
         gmshAtoJ,gmshAtoK,Cp,Pr,rho_n,rho,u,Tem,mu_fluid,mu_e,mu_sgs,Ml,Rmass,Rmom,Rener, &
         initialze,fact)
      integer(4), intent(in)  :: nelem, npoin
      integer(4), intent(in)  :: connec(nelem,nnode)
      real(rp),   intent(in)  :: Ngp(ngaus,nnode)
      real(rp),   intent(in)  :: He(ndime,ndime,ngaus,nelem),dlxigp_ip(ngaus,ndime,porder+1)
      real(rp),   intent(in)  :: gpvol(1,ngaus,nelem)
      integer(4), intent(in)  :: invAtoIJK(porder+1,porder+1,porder+1),gmshAtoI(nnode)
      integer(4), intent(in)  :: gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp),   intent(in)  :: Cp,Pr,rho_n(npoin),rho(npoin),u(npoin,ndime), Tem(npoin)
      real(rp),   intent(in)  :: mu_e(nelem,ngaus), mu_sgs(nelem,ngaus),Ml(npoin)
      real(rp),   intent(in)  :: mu_fluid(npoin)
      real(rp),   intent(inout) :: Rmass(npoin)
      real(rp),   intent(inout) :: Rmom(npoin,ndime)
      real(rp),   intent(inout) :: Rener(npoin)
      logical, optional, intent(in)    :: initialze
      real(rp), optional, intent(in)  :: fact
      if (present(fact)) fact_diffu = fact
      call check(c_full_diffusion(ctx, porder, nelem, npoin, dev(connec,'connec'), dev(He,'He'), &
         dev(gpvol,'gpvol'), dev(dlxigp_ip,'dlxigp_ip'), dev(invAtoIJK,'invAtoIJK'), &
         dev(gmshAtoI,'gmshAtoI'), dev(gmshAtoJ,'gmshAtoJ'), dev(gmshAtoK,'gmshAtoK'), Cp, Pr, &
         dev(rho_n,'rho_n'), dev(rho,'rho'), dev(u,'u'), dev(Tem,'Tem'), &
         dev(mu_fluid,'mu_fluid'), dev(mu_e,'mu_e'), dev(mu_sgs,'mu_sgs'), 1.0_rp, 1.0_rp, &
         dev(Rmass,'Rmass'), dev(Rmom,'Rmom'), dev(Rener,'Rener'), init_flag(initialze), &
         fact_diffu), 'full_diffusion_ijk')
   end subroutine full_diffusion_ijk

#if __RP__ == 8
   ! elem_stab.f90:589-616.  dt is part of the reference's interface and is not
   ! read; c_species_stab and flag_drop_c_in_envit are module globals.
   subroutine comp_tau(nelem,npoin,connec,csound,u,helem_k,dt,tau)
      integer(4), intent(in)   :: nelem, npoin,connec(nelem,nnode)
      real(rp),    intent(in)  :: u(npoin,ndime),helem_k(nelem),dt,csound(npoin)
      real(rp),    intent(out) :: tau(nelem)
      call check(c_comp_tau(ctx, porder, nelem, npoin, dev(connec,'connec'), &
         dev(csound,'csound'), dev(u,'u'), dev(helem_k,'helem_k'), c_species_stab, &
         vprm%flag_drop_c_in_envit, dev(tau,'tau')), 'comp_tau')
   end subroutine comp_tau

   ! elem_stab.f90:441-586.  Ngp, Cp, Pr and Ml are part of the reference's
   ! interface and are not read.
   subroutine full_proj_ijk(nelem,npoin,npoin_w,connec,lpoin_w,Ngp,He,gpvol,dlxigp_ip,invAtoIJK, &
         gmshAtoI,gmshAtoJ,gmshAtoK,Cp,Pr,rho,u,Tem,Ml,invMl,ProjMass,ProjEner,ProjMX,ProjMY, &
         ProjMZ)
      integer(4), intent(in)  :: nelem, npoin,npoin_w
      integer(4), intent(in)  :: connec(nelem,nnode),lpoin_w(npoin_w)
      real(rp),   intent(in)  :: Ngp(ngaus,nnode)
      real(rp),   intent(in)  :: He(ndime,ndime,ngaus,nelem),dlxigp_ip(ngaus,ndime,porder+1)
      real(rp),   intent(in)  :: gpvol(1,ngaus,nelem)
      integer(4), intent(in)  :: invAtoIJK(porder+1,porder+1,porder+1),gmshAtoI(nnode)
      integer(4), intent(in)  :: gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp),   intent(in)  :: Cp,Pr,rho(npoin),u(npoin,ndime),Tem(npoin),Ml(npoin)
      real(rp),   intent(in)  :: invMl(npoin)
      real(rp),   intent(inout) :: ProjMass(npoin,ndime),ProjEner(npoin,ndime)
      real(rp),   intent(inout) :: ProjMX(npoin,ndime),ProjMY(npoin,ndime),ProjMZ(npoin,ndime)
      call check(c_full_proj(ctx, porder, nelem, npoin, npoin_w, dev(connec,'connec'), &
         dev(lpoin_w,'lpoin_w'), dev(He,'He'), dev(gpvol,'gpvol'), dev(dlxigp_ip,'dlxigp_ip'), &
         dev(invAtoIJK,'invAtoIJK'), dev(gmshAtoI,'gmshAtoI'), dev(gmshAtoJ,'gmshAtoJ'), &
         dev(gmshAtoK,'gmshAtoK'), dev(rho,'rho'), dev(u,'u'), dev(Tem,'Tem'), &
         dev(invMl,'invMl'), dev(ProjMass,'ProjMass'), dev(ProjEner,'ProjEner'), &
         dev(ProjMX,'ProjMX'), dev(ProjMY,'ProjMY'), dev(ProjMZ,'ProjMZ')), 'full_proj_ijk')
   end subroutine full_proj_ijk

   ! elem_stab.f90:207-438.  Ngp and Ml are part of the reference's interface
   ! and are not read; c_rho and c_ener take the reference's defaults (1).
   subroutine full_diff_stab_ijk(nelem,npoin,connec,Ngp,He,gpvol,dlxigp_ip,invAtoIJK,gmshAtoI, &
         gmshAtoJ,gmshAtoK,Cp,Pr,rho_n,rho,u,Tem,mu_fluid,mu_e,mu_sgs,Ml,ProjMass,ProjEner, &
         ProjMX,ProjMY,ProjMZ,tau_stab,Rmass,Rmom,Rener,initialze,fact)
      integer(4), intent(in)  :: nelem, npoin
      integer(4), intent(in)  :: connec(nelem,nnode)
      real(rp),   intent(in)  :: Ngp(ngaus,nnode)
      real(rp),   intent(in)  :: He(ndime,ndime,ngaus,nelem),dlxigp_ip(ngaus,ndime,porder+1)
      real(rp),   intent(in)  :: gpvol(1,ngaus,nelem)
      integer(4), intent(in)  :: invAtoIJK(porder+1,porder+1,porder+1),gmshAtoI(nnode)
      integer(4), intent(in)  :: gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp),   intent(in)  :: Cp,Pr,rho_n(npoin),rho(npoin),u(npoin,ndime), Tem(npoin)
      real(rp),   intent(in)  :: Ml(npoin), mu_e(nelem,ngaus), mu_sgs(nelem,ngaus)
      real(rp),   intent(in)  :: tau_stab(nelem),ProjMass(npoin,ndime),ProjEner(npoin,ndime)
      real(rp),   intent(in)  :: ProjMX(npoin,ndime),ProjMY(npoin,ndime),ProjMZ(npoin,ndime)
      real(rp),   intent(in)  :: mu_fluid(npoin)
      real(rp),   intent(inout) :: Rmass(npoin)
      real(rp),   intent(inout) :: Rmom(npoin,ndime)
      real(rp),   intent(inout) :: Rener(npoin)
      logical, optional, intent(in)    :: initialze
      real(rp), optional, intent(in)  :: fact
      if (present(fact)) fact_diffstab = fact
      call check(c_diff_stab(ctx, porder, nelem, npoin, dev(connec,'connec'), dev(He,'He'), &
         dev(gpvol,'gpvol'), dev(dlxigp_ip,'dlxigp_ip'), dev(invAtoIJK,'invAtoIJK'), &
         dev(gmshAtoI,'gmshAtoI'), dev(gmshAtoJ,'gmshAtoJ'), dev(gmshAtoK,'gmshAtoK'), Cp, Pr, &
         dev(rho_n,'rho_n'), dev(rho,'rho'), dev(u,'u'), dev(Tem,'Tem'), &
         dev(mu_fluid,'mu_fluid'), dev(mu_e,'mu_e'), dev(mu_sgs,'mu_sgs'), &
         dev(ProjMass,'ProjMass'), dev(ProjEner,'ProjEner'), dev(ProjMX,'ProjMX'), &
         dev(ProjMY,'ProjMY'), dev(ProjMZ,'ProjMZ'), dev(tau_stab,'tau_stab'), 1.0_rp, 1.0_rp, &
         real(c_lps_comp, rp), dev(Rmass,'Rmass'), dev(Rmom,'Rmom'), dev(Rener,'Rener'), &
         init_flag(initialze), fact_diffstab), 'full_diff_stab_ijk')
   end subroutine full_diff_stab_ijk
#endif

   ! elem_convec.f90:472-594
   subroutine generic_scalar_convec_ijk(nelem,npoin,connec,Ngp,dNgp,He,gpvol,dlxigp_ip,xgp, &
         invAtoIJK,gmshAtoI,gmshAtoJ,gmshAtoK,q,eta,u,Rconvec,initialze,fact)
      integer(4), intent(in)  :: nelem, npoin
      integer(4), intent(in)  :: connec(nelem,nnode)
      real(rp),    intent(in)  :: Ngp(ngaus,nnode), dNgp(ndime,nnode,ngaus)
      real(rp),    intent(in)  :: He(ndime,ndime,ngaus,nelem),xgp(ngaus,ndime)
      real(rp),    intent(in)  :: dlxigp_ip(ngaus,ndime,porder+1)
      real(rp),    intent(in)  :: gpvol(1,ngaus,nelem)
      integer(4), intent(in)  :: invAtoIJK(porder+1,porder+1,porder+1),gmshAtoI(nnode)
      integer(4), intent(in)  :: gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp),    intent(in)  :: q(npoin,ndime)
      real(rp),    intent(in)  :: eta(npoin)
      real(rp),    intent(in)  :: u(npoin,ndime)
      real(rp),    intent(inout) :: Rconvec(npoin)
      logical, optional, intent(in)    :: initialze
      real(rp), optional, intent(in)  :: fact
      if (present(fact)) fact_scalar = fact
      call check(c_scalar_convec(ctx, porder, nelem, npoin, dev(connec,'connec'), dev(He,'He'), &
         dev(gpvol,'gpvol'), dev(dlxigp_ip,'dlxigp_ip'), dev(invAtoIJK,'invAtoIJK'), &
         dev(gmshAtoI,'gmshAtoI'), dev(gmshAtoJ,'gmshAtoJ'), dev(gmshAtoK,'gmshAtoK'), &
         dev(q,'q'), dev(eta,'eta'), dev(u,'u'), dev(Rconvec,'Rconvec'), init_flag(initialze), &
         fact_scalar), 'generic_scalar_convec_ijk')
   end subroutine generic_scalar_convec_ijk

   ! mod_entropy_viscosity.f90:14-111.  Ngp, coord, dNgp, gpvol, wgp, csound,
   ! Ml and gmshAtoI/J/K are part of the reference's interface and are not read.
   subroutine smart_visc_spectral(nelem,npoin,npoin_w,connec,lpoin_w,Reta,Ngp,coord,dNgp,gpvol, &
         wgp,gamma_gas,rho,u,csound,Tem,eta,helem,helem_k,Ml,mu_e,invAtoIJK,gmshAtoI,gmshAtoJ, &
         gmshAtoK,mue_l)
      integer(4), intent(in)  :: nelem, npoin,npoin_w, connec(nelem,nnode),lpoin_w(npoin_w)
      real(rp),   intent(in)  :: Reta(npoin), Ngp(ngaus,nnode),gamma_gas
      real(rp),   intent(in)  :: rho(npoin), u(npoin,ndime),csound(npoin), Tem(npoin), eta(npoin)
      real(rp),   intent(in)  :: helem(nelem,nnode),helem_k(nelem),Ml(npoin)
      real(rp),   intent(out) :: mu_e(nelem,ngaus)
      real(rp),   intent(in)  :: coord(npoin,ndime), dNgp(ndime,nnode,ngaus), wgp(ngaus)
      real(rp),   intent(in)  :: gpvol(1,ngaus,nelem)
      integer(4), intent(in)  :: invAtoIJK(porder+1,porder+1,porder+1),gmshAtoI(nnode)
      integer(4), intent(in)  :: gmshAtoJ(nnode),gmshAtoK(nnode)
      real(rp),intent(inout)  :: mue_l(nelem,nnode)
      type(visc_params_t) :: prm
      prm = vprm
      prm%flag_drop_c_in_envit = merge(1, 0, low_mach_dropped)
      call check(c_smart_visc(ctx, porder, nelem, npoin, npoin_w, dev(connec,'connec'), &
         dev(lpoin_w,'lpoin_w'), dev(Reta,'Reta'), gamma_gas, dev(rho,'rho'), dev(u,'u'), &
         dev(Tem,'Tem'), dev(eta,'eta'), dev(helem,'helem'), dev(helem_k,'helem_k'), &
         dev(invAtoIJK,'invAtoIJK'), prm, dev(mu_e,'mu_e'), dev(mue_l,'mue_l')), &
         'smart_visc_spectral')
   end subroutine smart_visc_spectral

#if __RP__ == 8
   ! mod_entropy_viscosity.f90:113-227, the live path's sensor.  Ngp, coord,
   ! dNgp, gpvol, wgp, Tem, helem, Ml, invAtoIJK and gmshAtoI/J/K are part of
   ! the reference's interface and are not read.
   subroutine smart_visc_spectral_imex(nelem,npoin,npoin_w,connec,lpoin_w,Reta,Ngp,coord,dNgp, &
         gpvol,wgp,gamma_gas,rho,u,csound,Tem,eta,helem,helem_k,Ml,mu_e,invAtoIJK,gmshAtoI, &
         gmshAtoJ,gmshAtoK,mue_l)
      integer(4), intent(in)  :: nelem, npoin,npoin_w, connec(nelem,nnode),lpoin_w(npoin_w)
      real(rp),   intent(in)  :: Reta(npoin), Ngp(ngaus,nnode),gamma_gas
      real(rp),   intent(in)  :: rho(npoin), u(npoin,ndime),csound(npoin), Tem(npoin), eta(npoin)
      real(rp),   intent(in)  :: helem(nelem,nnode),helem_k(nelem),Ml(npoin)
      real(rp),   intent(out) :: mu_e(nelem,ngaus)
      real(rp),   intent(in)  :: coord(npoin,ndime), dNgp(ndime,nnode,ngaus), wgp(ngaus)
      real(rp),   intent(in)  :: gpvol(1,ngaus,nelem)
      integer(4), intent(in)  :: invAtoIJK(porder+1,porder+1,porder+1),gmshAtoI(nnode)
      integer(4), intent(in)  :: gmshAtoJ(nnode),gmshAtoK(nnode)
      real(rp),intent(inout)  :: mue_l(nelem,nnode)
      vprm%flag_drop_c_in_envit = merge(1, 0, low_mach_dropped)
      call check(c_visc_imex(ctx, porder, nelem, npoin, npoin_w, dev(connec,'connec'), &
         dev(lpoin_w,'lpoin_w'), dev(Reta,'Reta'), dev(rho,'rho'), dev(u,'u'), &
         dev(csound,'csound'), dev(eta,'eta'), dev(helem_k,'helem_k'), vprm, ce_comp, &
         merge(1, 0, flag_high_mach), dev(mu_e,'mu_e'), dev(mue_l,'mue_l')), &
         'smart_visc_spectral_imex')
   end subroutine smart_visc_spectral_imex

   ! mod_operators.f90:485-570
   subroutine eval_divergence(nelem,npoin,connec,He,gpvol,dlxigp_ip,invAtoIJK,gmshAtoI,gmshAtoJ, &
         gmshAtoK,u,Rmom)
      integer(4), intent(in)  :: nelem, npoin
      integer(4), intent(in)  :: connec(nelem,nnode)
      real(rp),   intent(in)  :: He(ndime,ndime,ngaus,nelem),dlxigp_ip(ngaus,ndime,porder+1)
      real(rp),   intent(in)  :: gpvol(1,ngaus,nelem)
      integer(4), intent(in)  :: invAtoIJK(porder+1,porder+1,porder+1),gmshAtoI(nnode)
      integer(4), intent(in)  :: gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp),   intent(in)  :: u(npoin,ndime)
      real(rp),   intent(out) :: Rmom(npoin)
      call check(c_eval_div(ctx, porder, nelem, npoin, dev(connec,'connec'), dev(He,'He'), &
         dev(gpvol,'gpvol'), dev(dlxigp_ip,'dlxigp_ip'), dev(invAtoIJK,'invAtoIJK'), &
         dev(gmshAtoI,'gmshAtoI'), dev(gmshAtoJ,'gmshAtoJ'), dev(gmshAtoK,'gmshAtoK'), &
         dev(u,'u'), dev(Rmom,'Rmom')), 'eval_divergence')
   end subroutine eval_divergence

   ! mod_operators.f90:684-775.  dNgp, leviCivi (CFDSolverBase.f90:1605-1611's
   ! Levi-Civita symbol) and atoIJK are part of the reference's interface and
   ! are not read.
   subroutine compute_vorticity(nelem,npoin,npoin_w,lpoin_w,connec,lelpn,He,dNgp,leviCivi, &
         dlxigp_ip,atoIJK,invAtoIJK,gmshAtoI,gmshAtoJ,gmshAtoK,u,curlU,makeConti)
      integer(4), intent(in)  :: nelem, npoin, npoin_w, lpoin_w(npoin), connec(nelem,nnode)
      integer(4), intent(in)  :: lelpn(npoin)
      real(rp),   intent(in)  :: He(ndime,ndime,ngaus,nelem), dNgp(ndime,nnode,ngaus)
      real(rp),   intent(in)  :: u(npoin,ndime)
      real(rp),   intent(in)  :: leviCivi(ndime,ndime,ndime)
      real(rp),   intent(in)  :: dlxigp_ip(ngaus,ndime,porder+1)
      integer(4), intent(in)  :: atoIJK(nnode),invAtoIJK(porder+1,porder+1,porder+1)
      integer(4), intent(in)  :: gmshAtoI(nnode), gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp),   intent(out) :: curlU(npoin,ndime)
      logical,    intent(in)  :: makeConti
      type(c_ptr) :: plel
      plel = c_null_ptr
      if (makeConti) plel = dev(lelpn,'lelpn')
      call check(c_vorticity(ctx, porder, nelem, npoin, npoin_w, dev(lpoin_w,'lpoin_w'), &
         dev(connec,'connec'), plel, dev(He,'He'), dev(dlxigp_ip,'dlxigp_ip'), &
         dev(invAtoIJK,'invAtoIJK'), dev(gmshAtoI,'gmshAtoI'), dev(gmshAtoJ,'gmshAtoJ'), &
         dev(gmshAtoK,'gmshAtoK'), dev(u,'u'), dev(curlU,'curlU')), 'compute_vorticity')
   end subroutine compute_vorticity

   ! mod_entropy_viscosity.f90:277-314.  Ngp, coord, dNgp, gpvol, wgp, rho,
   ! csound, helem_k, invAtoIJK, gmshAtoI/J/K and mue_l are part of the
   ! reference's interface and do not enter the result.
   subroutine ducros_visc_limit(nelem,npoin,npoin_w,connec,lpoin_w,Ngp,coord,dNgp,gpvol,wgp, &
         rho,u,divU,vorti,csound,helem,helem_k,Ml,mu_e,invAtoIJK,gmshAtoI,gmshAtoJ,gmshAtoK,mue_l)
      integer(4), intent(in)  :: nelem, npoin,npoin_w, connec(nelem,nnode),lpoin_w(npoin_w)
      real(rp),   intent(in)  :: Ngp(ngaus,nnode),helem(nelem,nnode)
      real(rp),   intent(in)  :: rho(npoin), u(npoin,ndime),csound(npoin),helem_k(nelem)
      real(rp),   intent(in)  :: Ml(npoin),divU(npoin),vorti(npoin,ndime)
      real(rp),   intent(inout) :: mu_e(nelem,ngaus)
      real(rp),   intent(in)  :: coord(npoin,ndime), dNgp(ndime,nnode,ngaus), wgp(ngaus)
      real(rp),   intent(in)  :: gpvol(1,ngaus,nelem)
      integer(4), intent(in)  :: invAtoIJK(porder+1,porder+1,porder+1),gmshAtoI(nnode)
      integer(4), intent(in)  :: gmshAtoJ(nnode),gmshAtoK(nnode)
      real(rp),intent(inout)  :: mue_l(nelem,nnode)
      call check(c_ducros(ctx, porder, nelem, npoin, dev(connec,'connec'), dev(u,'u'), &
         dev(divU,'divU'), dev(vorti,'vorti'), dev(Ml,'Ml'), dev(helem,'helem'), &
         ducros_min_val, dev(mu_e,'mu_e')), 'ducros_visc_limit')
   end subroutine ducros_visc_limit
#endif

   ! mod_solver.f90:40-59
   subroutine lumped_solver_scal_opt(npoin,npoin_w,lpoin_w,invMl,R)
      integer(4), intent(in)    :: npoin, npoin_w, lpoin_w(npoin_w)
      real(rp),    intent(in)    :: invMl(npoin)
      real(rp),    intent(inout) :: R(npoin)
      call check(c_lumped_scal(ctx, npoin, npoin_w, dev(lpoin_w,'lpoin_w'), dev(invMl,'invMl'), &
         dev(R,'R')), 'lumped_solver_scal_opt')
   end subroutine lumped_solver_scal_opt

   ! mod_solver.f90:81-104
   subroutine lumped_solver_vect_opt(npoin,npoin_w,lpoin_w,invMl,R)
      integer(4), intent(in)    :: npoin, npoin_w, lpoin_w(npoin_w)
      real(rp),    intent(in)    :: invMl(npoin)
      real(rp),    intent(inout) :: R(npoin,ndime)
      call check(c_lumped_vect(ctx, npoin, npoin_w, dev(lpoin_w,'lpoin_w'), dev(invMl,'invMl'), &
         dev(R,'R')), 'lumped_solver_vect_opt')
   end subroutine lumped_solver_vect_opt

   ! Stage combine of rk_4_ls_main (time_integ_ls.f90:323-361) over working
   ! nodes with the lumped solve fused in: R <- R*invMl (kept, as the in-place
   ! lumped solve of the reference), then y <- lam1*y1 + lams*y + gam*dt*R for
   ! rho, E and q.  With invMl absent R already carries the mass inverse.
   subroutine rk_stage_update(npoin,npoin_w,lpoin_w,dt,lam1,lams,gam,rho1,q1,E1,rho,q,E, &
         Rmass,Rmom,Rener,invMl)
      integer(4), intent(in)    :: npoin, npoin_w, lpoin_w(npoin_w)
      real(rp),   intent(in)    :: dt, lam1, lams, gam
      real(rp),   intent(in)    :: rho1(npoin), q1(npoin,ndime), E1(npoin)
      real(rp),   intent(inout) :: rho(npoin), q(npoin,ndime), E(npoin)
      real(rp),   intent(inout) :: Rmass(npoin), Rmom(npoin,ndime), Rener(npoin)
      real(rp),   intent(in), optional :: invMl(npoin)
      type(c_ptr) :: pinv
      pinv = c_null_ptr
      if (present(invMl)) pinv = dev(invMl,'invMl')
      call check(c_rk_update(ctx, npoin, npoin_w, dev(lpoin_w,'lpoin_w'), pinv, dt, lam1, lams, &
         gam, dev(rho1,'rho1'), dev(q1,'q1'), dev(E1,'E1'), dev(rho,'rho'), dev(q,'q'), &
         dev(E,'E'), dev(Rmass,'Rmass'), dev(Rmom,'Rmom'), dev(Rener,'Rener')), 'rk_stage_update')
   end subroutine rk_stage_update

   ! ------------------------------------------------------------ fused stages
   ! One stage of rk_4_ls_main (time_integ_ls.f90:339-361) with updateF
   ! (:489-691, flag_lps_stab = .false.) fused in: primitives and Sutherland
   ! at (rho, q, E), R = -(full_convec_ijk + full_diffusion_ijk), the lumped
   ! inverse, y <- lam1*y1 + lams*y + gam*dt*R.  R*invMl goes to Rmass/Rmom/
   ! Rener when they are present.  With gam_next and rho_next/q_next/E_next
   ! the next step's first stage at fixed dt is stored as well
   ! (sodhip_rk_stage_next_f64).  Operator precision follows rp.
   subroutine rk_stage_fused(nelem,npoin,npoin_w,connec,lpoin_w,He,gpvol,dlxigp_ip,invAtoIJK, &
         gmshAtoI,gmshAtoJ,gmshAtoK,mu_e,invMl,gamma_gas,Rgas,Cp,Prt,mu_factor,dt,lam1,lams, &
         gam,rho1,q1,E1,rho,q,E,Rmass,Rmom,Rener,gam_next,rho_next,q_next,E_next)
      integer(4), intent(in)    :: nelem, npoin, npoin_w
      integer(4), intent(in)    :: connec(nelem,nnode), lpoin_w(npoin_w)
      real(rp),   intent(in)    :: He(ndime,ndime,ngaus,nelem), gpvol(1,ngaus,nelem)
      real(rp),   intent(in)    :: dlxigp_ip(ngaus,ndime,porder+1)
      integer(4), intent(in)    :: invAtoIJK(porder+1,porder+1,porder+1), gmshAtoI(nnode)
      integer(4), intent(in)    :: gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp),   intent(in)    :: mu_e(nelem,ngaus)
      real(8),    intent(in)    :: invMl(npoin)
      real(rp),   intent(in)    :: gamma_gas, Rgas, Cp, Prt, mu_factor, dt, lam1, lams, gam
      real(8),    intent(in)    :: rho1(npoin), q1(npoin,ndime), E1(npoin)
      real(8),    intent(inout) :: rho(npoin), q(npoin,ndime), E(npoin)
      real(8),    intent(inout), optional :: Rmass(npoin), Rmom(npoin,ndime), Rener(npoin)
      real(rp),   intent(in),    optional :: gam_next
      real(8),    intent(inout), optional :: rho_next(npoin), q_next(npoin,ndime), E_next(npoin)
      type(stage_params_t) :: prm
      type(c_ptr) :: pm, pq, pe
      prm = stage_params_t(real(gamma_gas,8), real(Rgas,8), real(Cp,8), real(Prt,8), &
         real(mu_factor,8), 1.0d0, 1.0d0, -1.0d0, 0.0d0)
      pm = c_null_ptr; pq = c_null_ptr; pe = c_null_ptr
      if (present(Rmass)) then
         pm = dev(Rmass,'Rmass'); pq = dev(Rmom,'Rmom'); pe = dev(Rener,'Rener')
      end if
      if (present(gam_next)) then
         call check(c_rk_stage_next(ctx, porder, nelem, npoin, npoin_w, dev(connec,'connec'), &
            dev(lpoin_w,'lpoin_w'), rp, dev(He,'He'), dev(gpvol,'gpvol'), &
            dev(dlxigp_ip,'dlxigp_ip'), dev(invAtoIJK,'invAtoIJK'), dev(gmshAtoI,'gmshAtoI'), &
            dev(gmshAtoJ,'gmshAtoJ'), dev(gmshAtoK,'gmshAtoK'), dev(mu_e,'mu_e'), &
            dev(invMl,'invMl'), prm, real(dt,8), real(lam1,8), real(lams,8), real(gam,8), &
            dev(rho1,'rho1'), dev(q1,'q1'), dev(E1,'E1'), dev(rho,'rho'), dev(q,'q'), &
            dev(E,'E'), pm, pq, pe, real(gam_next,8), dev(rho_next,'rho_next'), &
            dev(q_next,'q_next'), dev(E_next,'E_next')), 'rk_stage_fused (next)')
      else
         call check(c_rk_stage(ctx, porder, nelem, npoin, npoin_w, dev(connec,'connec'), &
            dev(lpoin_w,'lpoin_w'), rp, dev(He,'He'), dev(gpvol,'gpvol'), &
            dev(dlxigp_ip,'dlxigp_ip'), dev(invAtoIJK,'invAtoIJK'), dev(gmshAtoI,'gmshAtoI'), &
            dev(gmshAtoJ,'gmshAtoJ'), dev(gmshAtoK,'gmshAtoK'), dev(mu_e,'mu_e'), &
            dev(invMl,'invMl'), prm, real(dt,8), real(lam1,8), real(lams,8), real(gam,8), &
            dev(rho1,'rho1'), dev(q1,'q1'), dev(E1,'E1'), dev(rho,'rho'), dev(q,'q'), &
            dev(E,'E'), pm, pq, pe), 'rk_stage_fused')
      end if
   end subroutine rk_stage_fused

   ! One stage of the LPS path (flag_lps_stab, time_integ_ls.f90:607-611):
   ! updateF with full_proj_ijk and full_diff_stab_ijk, the lumped solve and
   ! the stage combine (sodhip_rk_stage_lps_f64, FP64).
   subroutine rk_stage_lps_fused(nelem,npoin,npoin_w,connec,lpoin_w,He,gpvol,dlxigp_ip, &
         invAtoIJK,gmshAtoI,gmshAtoJ,gmshAtoK,mu_e,invMl,tau_stab,gamma_gas,Rgas,Cp,Prt, &
         mu_factor,dt,lam1,lams,gam,rho1,q1,E1,rho,q,E,u,e_int,pr,aux_h,Tem,mu_fluid,ProjMass, &
         ProjEner,ProjMX,ProjMY,ProjMZ,Rmass,Rmom,Rener,gam_next,rho_next,q_next,E_next)
      integer(4), intent(in)    :: nelem, npoin, npoin_w
      integer(4), intent(in)    :: connec(nelem,nnode), lpoin_w(npoin_w)
      real(8),    intent(in)    :: He(ndime,ndime,ngaus,nelem), gpvol(1,ngaus,nelem)
      real(8),    intent(in)    :: dlxigp_ip(ngaus,ndime,porder+1)
      integer(4), intent(in)    :: invAtoIJK(porder+1,porder+1,porder+1), gmshAtoI(nnode)
      integer(4), intent(in)    :: gmshAtoJ(nnode), gmshAtoK(nnode)
      real(8),    intent(in)    :: mu_e(nelem,ngaus), invMl(npoin), tau_stab(nelem)
      real(8),    intent(in)    :: gamma_gas, Rgas, Cp, Prt, mu_factor, dt, lam1, lams, gam
      real(8),    intent(in)    :: rho1(npoin), q1(npoin,ndime), E1(npoin)
      real(8),    intent(inout) :: rho(npoin), q(npoin,ndime), E(npoin)
      real(8),    intent(inout) :: u(npoin,ndime), e_int(npoin), pr(npoin), aux_h(npoin)
      real(8),    intent(inout) :: Tem(npoin), mu_fluid(npoin)
      real(8),    intent(inout) :: ProjMass(npoin,ndime), ProjEner(npoin,ndime)
      real(8),    intent(inout) :: ProjMX(npoin,ndime), ProjMY(npoin,ndime), ProjMZ(npoin,ndime)
      real(8),    intent(inout), optional :: Rmass(npoin), Rmom(npoin,ndime), Rener(npoin)
      real(8),    intent(in),    optional :: gam_next
      real(8),    intent(inout), optional :: rho_next(npoin), q_next(npoin,ndime), E_next(npoin)
      type(stage_params_t) :: prm
      type(c_ptr) :: pm, pq, pe, nr, nq, ne
      real(8) :: gn
      prm = stage_params_t(gamma_gas, Rgas, Cp, Prt, mu_factor, 1.0d0, 1.0d0, -1.0d0, 0.0d0)
      pm = c_null_ptr; pq = c_null_ptr; pe = c_null_ptr
      nr = c_null_ptr; nq = c_null_ptr; ne = c_null_ptr; gn = 0.0d0
      if (present(Rmass)) then
         pm = dev(Rmass,'Rmass'); pq = dev(Rmom,'Rmom'); pe = dev(Rener,'Rener')
      end if
      if (present(gam_next)) then
         gn = gam_next
         nr = dev(rho_next,'rho_next'); nq = dev(q_next,'q_next'); ne = dev(E_next,'E_next')
      end if
      call check(c_rk_stage_lps(ctx, porder, nelem, npoin, npoin_w, dev(connec,'connec'), &
         dev(lpoin_w,'lpoin_w'), dev(He,'He'), dev(gpvol,'gpvol'), dev(dlxigp_ip,'dlxigp_ip'), &
         dev(invAtoIJK,'invAtoIJK'), dev(gmshAtoI,'gmshAtoI'), dev(gmshAtoJ,'gmshAtoJ'), &
         dev(gmshAtoK,'gmshAtoK'), dev(mu_e,'mu_e'), dev(invMl,'invMl'), &
         dev(tau_stab,'tau_stab'), prm, real(c_lps_comp,8), dt, lam1, lams, gam, &
         dev(rho1,'rho1'), dev(q1,'q1'), dev(E1,'E1'), dev(rho,'rho'), dev(q,'q'), dev(E,'E'), &
         dev(u,'u'), dev(e_int,'e_int'), dev(pr,'pr'), dev(aux_h,'aux_h'), dev(Tem,'Tem'), &
         dev(mu_fluid,'mu_fluid'), dev(ProjMass,'ProjMass'), dev(ProjEner,'ProjEner'), &
         dev(ProjMX,'ProjMX'), dev(ProjMY,'ProjMY'), dev(ProjMZ,'ProjMZ'), pm, pq, pe, gn, nr, &
         nq, ne), 'rk_stage_lps_fused')
   end subroutine rk_stage_lps_fused

   ! rk_4_ls_main's end-of-step node updates in one pass (time_integ_ls.f90:
   ! 391-407, thermo_entropy mod_entropy_viscosity.f90:411-433 without the
   ! flux, sutherland_viscosity with a uniform factor): eta is eta(npoin,2),
   ! level 2 is written.
   subroutine eos_entropy_step(npoin,npoin_w,lpoin_w,gamma_gas,Rgas,p_inf,rho_inf,mu_factor, &
         rho,q,E,u,e_int,pr,csound,machno,Tem,eta,mu_fluid)
      integer(4), intent(in)    :: npoin, npoin_w, lpoin_w(npoin_w)
      real(8),    intent(in)    :: gamma_gas, Rgas, p_inf, rho_inf, mu_factor
      real(8),    intent(in)    :: rho(npoin), q(npoin,ndime), E(npoin)
      real(8),    intent(inout) :: u(npoin,ndime), e_int(npoin), pr(npoin), csound(npoin)
      real(8),    intent(inout) :: machno(npoin), Tem(npoin), eta(npoin,2), mu_fluid(npoin)
      call check(c_eos_entropy_step(ctx, npoin, npoin_w, dev(lpoin_w,'lpoin_w'), gamma_gas, &
         Rgas, p_inf, rho_inf, mu_factor, dev(rho,'rho'), dev(q,'q'), dev(E,'E'), dev(u,'u'), &
         dev(e_int,'e_int'), dev(pr,'pr'), dev(csound,'csound'), dev(machno,'machno'), &
         dev(Tem,'Tem'), dev(eta,'eta'), dev(mu_fluid,'mu_fluid')), 'eos_entropy_step')
   end subroutine eos_entropy_step

   ! The end-of-step entropy residual (time_integ_ls.f90:429-443): Reta(:,2) of
   ! f_eta = u*eta(:,1), its lumped solve, auxReta = 1.5 Reta2 - 0.5 Reta1 +
   ! factor_comp (eta2-eta1)/dt, then Reta(:,1) <- Reta(:,2).
   subroutine entropy_convec_residual(nelem,npoin,npoin_w,connec,lpoin_w,He,gpvol,dlxigp_ip, &
         invAtoIJK,gmshAtoI,gmshAtoJ,gmshAtoK,u,eta,invMl,factor_comp,dt,Reta,auxReta)
      integer(4), intent(in)    :: nelem, npoin, npoin_w
      integer(4), intent(in)    :: connec(nelem,nnode), lpoin_w(npoin_w)
      real(8),    intent(in)    :: He(ndime,ndime,ngaus,nelem), gpvol(1,ngaus,nelem)
      real(8),    intent(in)    :: dlxigp_ip(ngaus,ndime,porder+1)
      integer(4), intent(in)    :: invAtoIJK(porder+1,porder+1,porder+1), gmshAtoI(nnode)
      integer(4), intent(in)    :: gmshAtoJ(nnode), gmshAtoK(nnode)
      real(8),    intent(in)    :: u(npoin,ndime), eta(npoin,2), invMl(npoin)
      real(8),    intent(in)    :: factor_comp, dt
      real(8),    intent(inout) :: Reta(npoin,2), auxReta(npoin)
      call check(c_entropy_convec_residual(ctx, porder, nelem, npoin, npoin_w, &
         dev(connec,'connec'), dev(lpoin_w,'lpoin_w'), dev(He,'He'), dev(gpvol,'gpvol'), &
         dev(dlxigp_ip,'dlxigp_ip'), dev(invAtoIJK,'invAtoIJK'), dev(gmshAtoI,'gmshAtoI'), &
         dev(gmshAtoJ,'gmshAtoJ'), dev(gmshAtoK,'gmshAtoK'), dev(u,'u'), dev(eta,'eta'), &
         dev(invMl,'invMl'), factor_comp, dt, dev(Reta,'Reta'), dev(auxReta,'auxReta')), &
         'entropy_convec_residual')
   end subroutine entropy_convec_residual

   ! One stage of the classical RK4 rk_4_main (time_integ.f90:250-470,
   ! compiled out in the reference) with the RHS fused: R at the stage state
   ! (srho, sq, sE), operators without fact, lumped inverse, Rsum <- Rsum + b R,
   ! then aux = y - dt a_next R, or y <- y - dt Rsum at the last stage.
   subroutine rk4_stage_fused(nelem,npoin,npoin_w,connec,lpoin_w,He,gpvol,dlxigp_ip,invAtoIJK, &
         gmshAtoI,gmshAtoJ,gmshAtoK,mu_e,invMl,gamma_gas,Rgas,Cp,Prt,mu_factor,dt,a_next,b, &
         first,last,rho,q,E,srho,sq,sE,aux_rho,aux_q,aux_E,Rmass_sum,Rmom_sum,Rener_sum)
      integer(4), intent(in)    :: nelem, npoin, npoin_w
      integer(4), intent(in)    :: connec(nelem,nnode), lpoin_w(npoin_w)
      real(rp),   intent(in)    :: He(ndime,ndime,ngaus,nelem), gpvol(1,ngaus,nelem)
      real(rp),   intent(in)    :: dlxigp_ip(ngaus,ndime,porder+1)
      integer(4), intent(in)    :: invAtoIJK(porder+1,porder+1,porder+1), gmshAtoI(nnode)
      integer(4), intent(in)    :: gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp),   intent(in)    :: mu_e(nelem,ngaus)
      real(8),    intent(in)    :: invMl(npoin)
      real(rp),   intent(in)    :: gamma_gas, Rgas, Cp, Prt, mu_factor, dt, a_next, b
      logical,    intent(in)    :: first, last
      real(8),    intent(inout) :: rho(npoin), q(npoin,ndime), E(npoin)
      real(8),    intent(in)    :: srho(npoin), sq(npoin,ndime), sE(npoin)
      real(8),    intent(inout) :: aux_rho(npoin), aux_q(npoin,ndime), aux_E(npoin)
      real(8),    intent(inout) :: Rmass_sum(npoin), Rmom_sum(npoin,ndime), Rener_sum(npoin)
      type(stage_params_t) :: prm
      integer(c_int) :: fi, la
      prm = stage_params_t(real(gamma_gas,8), real(Rgas,8), real(Cp,8), real(Prt,8), &
         real(mu_factor,8), 1.0d0, 1.0d0, 1.0d0, 0.0d0)
      fi = 0; la = 0
      if (first) fi = 1
      if (last) la = 1
      call check(c_rk4_stage(ctx, porder, nelem, npoin, npoin_w, dev(connec,'connec'), &
         dev(lpoin_w,'lpoin_w'), rp, dev(He,'He'), dev(gpvol,'gpvol'), &
         dev(dlxigp_ip,'dlxigp_ip'), dev(invAtoIJK,'invAtoIJK'), dev(gmshAtoI,'gmshAtoI'), &
         dev(gmshAtoJ,'gmshAtoJ'), dev(gmshAtoK,'gmshAtoK'), dev(mu_e,'mu_e'), &
         dev(invMl,'invMl'), prm, real(dt,8), real(a_next,8), real(b,8), fi, la, dev(rho,'rho'), &
         dev(q,'q'), dev(E,'E'), dev(srho,'srho'), dev(sq,'sq'), dev(sE,'sE'), &
         dev(aux_rho,'aux_rho'), dev(aux_q,'aux_q'), dev(aux_E,'aux_E'), &
         dev(Rmass_sum,'Rmass_sum'), dev(Rmom_sum,'Rmom_sum'), dev(Rener_sum,'Rener_sum'), &
         c_null_ptr, c_null_ptr, c_null_ptr), 'rk4_stage_fused')
   end subroutine rk4_stage_fused

   ! The node loops of an rk_4_main stage on assembled residuals (time_integ.f90:
   ! 266-275, 431-451, 460-468), with the lumped inverse when invMl is present.
   subroutine rk4_stage_update(npoin,npoin_w,lpoin_w,dt,a_next,b,first,last,rho,q,E,aux_rho, &
         aux_q,aux_E,Rmass,Rmom,Rener,Rmass_sum,Rmom_sum,Rener_sum,invMl)
      integer(4), intent(in)    :: npoin, npoin_w, lpoin_w(npoin_w)
      real(8),    intent(in)    :: dt, a_next, b
      logical,    intent(in)    :: first, last
      real(8),    intent(inout) :: rho(npoin), q(npoin,ndime), E(npoin)
      real(8),    intent(inout) :: aux_rho(npoin), aux_q(npoin,ndime), aux_E(npoin)
      real(8),    intent(inout) :: Rmass(npoin), Rmom(npoin,ndime), Rener(npoin)
      real(8),    intent(inout) :: Rmass_sum(npoin), Rmom_sum(npoin,ndime), Rener_sum(npoin)
      real(8),    intent(in), optional :: invMl(npoin)
      type(c_ptr) :: pinv
      integer(c_int) :: fi, la
      pinv = c_null_ptr
      if (present(invMl)) pinv = dev(invMl,'invMl')
      fi = 0; la = 0
      if (first) fi = 1
      if (last) la = 1
      call check(c_rk4_update(ctx, npoin, npoin_w, dev(lpoin_w,'lpoin_w'), pinv, dt, a_next, b, &
         fi, la, dev(rho,'rho'), dev(q,'q'), dev(E,'E'), dev(aux_rho,'aux_rho'), &
         dev(aux_q,'aux_q'), dev(aux_E,'aux_E'), dev(Rmass,'Rmass'), dev(Rmom,'Rmom'), &
         dev(Rener,'Rener'), dev(Rmass_sum,'Rmass_sum'), dev(Rmom_sum,'Rmom_sum'), &
         dev(Rener_sum,'Rener_sum')), 'rk4_stage_update')
   end subroutine rk4_stage_update

end module mod_sod2d_hip

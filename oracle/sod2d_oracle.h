/*
 * sod2d_oracle.h -- CPU FP64 restatement of SOD2D's compressible SEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP path: it
 * may be used by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py, and by nothing else.  The product never links it.
 *
 * Conventions (identical to the reference's Fortran dummy arguments):
 *   - arrays are column-major with the reference's dimension order, passed as
 *     plain pointers; e.g. He(ndime,ndime,ngaus,nelem) is He[i+3*(j+3*(g+ngaus*e))];
 *   - integer index tables (connec, lpoin_w, invAtoIJK, gmshAto*) hold the
 *     reference's 1-based values;
 *   - p (porder) is a runtime argument here (the reference fixes it at compile
 *     time through __PORDER__); nnode = ngaus = (p+1)^3.
 *
 * Each function cites the reference file:line it restates.  The pinning of each
 * function (golden vectors, closed forms, invariants) lives in tests/.
 */
#ifndef SOD2D_ORACLE_H
#define SOD2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- basis / quadrature (mod_maths.f90, quadrature_rules.f90, elem_hex.f90) ---- */
void or_valepo(int n, double x, double *y, double *dy, double *d2y);
void or_zelegl(int n, double *et, double *vn);
void or_welegl(int n, const double *et, const double *vn, double *wt);
void or_gll(int p, double *xi, double *w);
void or_lagrange(int p, const double *xi, double xp, double *l);
void or_lagrange_deriv(int p, const double *xi, double xp, double *dl);
/* gmsh high-order hexahedron (i,j,k) table, (p+1)^3 x 3 column-major */
void or_gmsh_hex_table(int p, int *table);
/* a2ijk = gmsh2ijk (mod_ijk_indices.f90:61-114; mod_meshConversorTool.f90:160-161) */
void or_a2ijk(int p, int *a2ijk);
void or_atoijk_inverse(int p, const int *a2ijk, int *invAtoIJK, int *gI, int *gJ, int *gK);
/* xgp(ngaus,3), wgp(ngaus), Ngp(ngaus,nnode), dNgp(3,nnode,ngaus), dlxigp_ip(ngaus,3,p+1) */
void or_shape(int p, const int *a2ijk, double *xgp, double *wgp, double *Ngp,
              double *dNgp, double *dlxigp_ip);

/* ---- geometry (jacobian_oper.f90, mass_matrix.f90, mod_geom.f90) ---- */
void or_elem_jacobian(int p, int nelem, int npoin, const int *connec, const double *coord,
                      const double *dNgp, const double *wgp, double *gpvol, double *He);
void or_lumped_mass(int p, int nelem, int npoin, const int *connec, const double *gpvol, double *Ml);
void or_inv_ml(int npoin, int npoin_w, const int *lpoin_w, const double *Ml, double *invMl);
void or_char_length(int p, int nelem, int npoin, const int *connec, const double *coord, double *helem);
void or_char_length_spectral(int p, int nelem, int npoin, const int *connec, const double *coord,
                             double *helem_l);

/* ---- operators (elem_convec.f90, elem_diffu.f90) ----
 * initialize: 1 = zero outputs first; fact: multiplier of the element residual.
 * nthreads > 1 runs the element loop with OpenMP and atomic assembly (used only
 * to time a multi-core CPU baseline; the sum order then varies). */
void or_full_convec_ijk(int p, int nelem, int npoin, const int *connec, const double *He,
                        const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                        const int *gI, const int *gJ, const int *gK, const double *u,
                        const double *q, const double *rho, const double *pr, const double *E,
                        double *Rmass, double *Rmom, double *Rener, int initialize, double fact,
                        int nthreads);
void or_full_diffusion_ijk(int p, int nelem, int npoin, const int *connec, const double *He,
                           const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                           const int *gI, const int *gJ, const int *gK, double Cp, double Pr,
                           const double *rho_n, const double *rho, const double *u,
                           const double *Tem, const double *mu_fluid, const double *mu_e,
                           const double *mu_sgs, double c_rho, double c_ener, double *Rmass,
                           double *Rmom, double *Rener, int initialize, double fact, int nthreads);
void or_comp_tau(int p, int nelem, int npoin, const int *connec, const double *csound,
                 const double *u, const double *helem_k, double c_species_stab, int drop_c,
                 double *tau);
void or_full_proj_ijk(int p, int nelem, int npoin, int npoin_w, const int *connec,
                      const int *lpoin_w, const double *He, const double *gpvol,
                      const double *dlxigp_ip, const int *invAtoIJK, const int *gI, const int *gJ,
                      const int *gK, const double *rho, const double *u, const double *Tem,
                      const double *invMl, double *ProjMass, double *ProjEner, double *ProjMX,
                      double *ProjMY, double *ProjMZ, int nthreads);
void or_full_diff_stab_ijk(int p, int nelem, int npoin, const int *connec, const double *He,
                           const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                           const int *gI, const int *gJ, const int *gK, double Cp, double Pr,
                           const double *rho_n, const double *rho, const double *u,
                           const double *Tem, const double *mu_fluid, const double *mu_e,
                           const double *mu_sgs, const double *ProjMass, const double *ProjEner,
                           const double *ProjMX, const double *ProjMY, const double *ProjMZ,
                           const double *tau_stab, double c_rho, double c_ener, double c_lps,
                           double *Rmass, double *Rmom, double *Rener, int initialize,
                           double fact, int nthreads);
void or_generic_scalar_convec_ijk(int p, int nelem, int npoin, const int *connec, const double *He,
                                  const double *gpvol, const double *dlxigp_ip,
                                  const int *invAtoIJK, const int *gI, const int *gJ,
                                  const int *gK, const double *q, const double *eta,
                                  const double *u, double *Rconvec, int initialize, double fact,
                                  int nthreads);

/* ---- entropy viscosity (mod_entropy_viscosity.f90, mod_filters.f90) ---- */
void or_convert_ijk(int p, int *convertIJK /* size p+3, index 0..p+2 */);
typedef struct {
    double ce, cmax, cglob, Rgas, c_rho, c_ener, ce_comp, factor_comp;
    int flag_normalise_entropy, flag_drop_c_in_envit;
} or_visc_params;
void or_smart_visc_spectral(int p, int nelem, int npoin, int npoin_w, const int *connec,
                            const int *lpoin_w, const double *Reta, double gamma_gas,
                            const double *rho, const double *u, const double *Tem,
                            const double *eta, const double *helem, const double *helem_k,
                            const int *invAtoIJK, const or_visc_params *prm, double *mu_e,
                            double *mue_l);
/* ---- live-path sensor (mod_entropy_viscosity.f90:113-314, mod_operators.f90) ---- */
void or_smart_visc_spectral_imex(int p, int nelem, int npoin, int npoin_w, const int *connec,
                                 const int *lpoin_w, const double *Reta, const double *rho,
                                 const double *u, const double *csound, const double *eta,
                                 const double *helem_k, or_visc_params *prm, int flag_high_mach,
                                 double *mu_e, double *mue_l);
void or_eval_divergence(int p, int nelem, int npoin, const int *connec, const double *He,
                        const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                        const int *gI, const int *gJ, const int *gK, const double *u,
                        double *divU);
void or_compute_vorticity(int p, int nelem, int npoin, int npoin_w, const int *lpoin_w,
                          const int *connec, const int *lelpn, const double *He,
                          const double *dlxigp_ip, const int *invAtoIJK, const int *gI,
                          const int *gJ, const int *gK, const double *u, double *curlU);
void or_ducros_visc_limit(int p, int nelem, int npoin, const int *connec, const double *u,
                          const double *divU, const double *vorti, const double *Ml,
                          const double *helem, double ducros_min_val, double *mu_e);
void or_elem_per_node(int nelem, int nnode, int npoin, const int *connec, int *lelpn);
void or_thermo_entropy(int npoin, int npoin_w, const int *lpoin_w, double gamma_gas,
                       double p_inf, double rho_inf, const double *rho, const double *pr,
                       const double *u, double *eta /* (npoin,2) */, double *feta);

/* ---- nodal updates (mod_solver.f90, mod_fluid_viscosity.f90, time_integ_ls.f90) ---- */
void or_lumped_solver_scal(int npoin, int npoin_w, const int *lpoin_w, const double *invMl,
                           double *R);
void or_lumped_solver_vect(int npoin, int npoin_w, const int *lpoin_w, const double *invMl,
                           double *R);
void or_sutherland(int npoin, const double *Tem, const double *mu_factor, double *mu_fluid);
/* updateF's EOS block; outputs u, e_int, pr, aux_h, Tem at the current level */
void or_eos_stage(int npoin, int npoin_w, const int *lpoin_w, double gamma_gas, double Rgas,
                  const double *rho, const double *q, const double *E, double *u, double *e_int,
                  double *pr, double *aux_h, double *Tem);
/* rk_4_ls_main's end-of-step EOS block (adds csound and machno) */
void or_eos_step(int npoin, int npoin_w, const int *lpoin_w, double gamma_gas, double Rgas,
                 const double *rho, const double *q, const double *E, double *u, double *e_int,
                 double *pr, double *csound, double *machno, double *Tem);
/* lambda_ij/gamma_ij (6x6 column-major, 1-based rows/cols stored at [i-1 + 6*(j-1)]) */
int or_rk_ls_coeffs(int stages, int variant, double *lambda_ij, double *gamma_ij);
void or_rk4_coeffs(double *a, double *b, double *c);
/* y(pos) = l1*y(1) + ls*y(pos) + g*dt*R over working nodes, for rho, E and q(:,1:3) */
void or_rk_ls_update(int npoin, int npoin_w, const int *lpoin_w, double l1, double ls, double gdt,
                     const double *rho1, const double *q1, const double *E1, double *rho,
                     double *q, double *E, const double *Rmass, const double *Rmom,
                     const double *Rener);

/* ---- diagnostics and dt (mod_analysis.f90, mod_time_ops.f90) ---- */
double or_volavg_ek(int p, int nelem, int npoin, const int *connec, const double *gpvol,
                    const double *Ngp, double rho0, const double *rho, const double *u);
void or_visc_dissipation(int p, int nelem, int npoin, const int *connec, double rho0,
                         const double *mu_fluid, const double *mu_e, const double *u, double volT,
                         const double *gpvol, const double *He, const double *dNgp,
                         double *eps_S, double *eps_D, double *eps_T);
double or_max_mach(int npoin, int npoin_w, const int *lpoin_w, const double *machno);
double or_adapt_dt_cfl(int p, int nelem, int npoin, const int *connec, const double *helem,
                       const double *u, const double *csound, double cfl_conv, double cfl_diff,
                       const double *mu_fluid, const double *mu_sgs, const double *mu_e,
                       const double *rho);

/* ---- TGV case (TGVSolver.f90) ---- */
typedef struct {
    double Cp, Prt, M, Re, rho0, gamma_gas;
    /* derived */
    double Rgas, to, po, mu_factor;
} or_tgv_params;
void or_tgv_params_derive(or_tgv_params *t);
void or_tgv_ic(int npoin, const double *coord, const or_tgv_params *t, double *u, double *q,
               double *rho, double *pr, double *E, double *e_int, double *Tem, double *csound,
               double *eta, double *machno);

#ifdef __cplusplus
}
#endif
#endif

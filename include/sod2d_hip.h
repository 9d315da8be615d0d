/*
 * sod2d_hip.h -- C-ABI of the MI355X (gfx950) implementation of SOD2D's
 * compressible spectral-element right-hand side and Runge-Kutta stage.
 *
 * Drop-in boundary.  Each operator entry point replaces one Fortran subroutine
 * of the reference and takes the same arguments with the same meaning, layout
 * (column-major, the reference's dimension order) and 1-based index values;
 * the Fortran shim sod2d_amd/fortran/mod_sod2d_hip.f90 keeps the reference's
 * subroutine interfaces and forwards to these functions.  Reference citations
 * are relative to the LIUJUNLE97/sod2d source tree.
 *
 * Pointer arguments of the operator entry points are DEVICE pointers (the
 * Fortran shim maps host arrays to their device mirrors with sodhip_deviceptr,
 * the equivalent of OpenACC's host_data use_device).  All work is enqueued on
 * the context's stream; entry points return after enqueueing.  Functions that
 * hand data to the host synchronise.
 *
 * Every function returns 0 on success or a SODHIP_ERR_* code; the message of
 * the last error of the calling thread is available from sodhip_last_error.
 * The reference aborts (stop 1 / MPI_Abort) where this layer returns an error.
 */
#ifndef SOD2D_HIP_H
#define SOD2D_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SODHIP_OK 0
#define SODHIP_ERR_ARG 1
#define SODHIP_ERR_HIP 2
#define SODHIP_ERR_NOT_PRESENT 3
#define SODHIP_ERR_UNSUPPORTED 4
#define SODHIP_ERR_ALLOC 5

typedef struct sodhip_ctx_s *sodhip_ctx;

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */
int sodhip_version(void);
/* copies the last error message (NUL-terminated, truncated to len) */
int sodhip_last_error(char *buf, int len);

/* ------------------------------------------------------------------------ */
/* Host-side case setup (CPU only; no device needed).  These restate the     */
/* reference's set-up routines that produce the operators' inputs.           */
/* ------------------------------------------------------------------------ */

/* Element tables of order porder (2 <= porder <= 8):
 *   a2ijk(nnode)                 mesh_a2ijk = gmsh2ijk (mod_ijk_indices.f90:61-114)
 *   invAtoIJK(p+1,p+1,p+1), gmshAtoI/J/K(nnode)   atoIJKInverse (mod_geom.f90:127-148)
 *   xgp(ngaus,3), wgp(ngaus)     GaussLobattoLegendre_hex (quadrature_rules.f90:51-78)
 *   dNgp(3,nnode,ngaus), dlxigp_ip(ngaus,3,p+1)
 *                                hex_highorder (elem_hex.f90:22-43) as driven by
 *                                CFDSolverBase_evalShapeFunctions (CFDSolverBase.f90:1468-1619)
 * Any output pointer may be NULL. */
int sodhip_basis_tables(int porder, int *a2ijk, int *invAtoIJK, int *gmshAtoI, int *gmshAtoJ,
                        int *gmshAtoK, double *xgp, double *wgp, double *dNgp, double *dlxigp_ip);

/* The /order tables of a SOD2D mesh file of order porder, as the partitioner
 * writes them: a2ijk = gmsh2ijk and vtk2ijk (nnode), a2ij = gmsh2ij and vtk2ij
 * (npbou = (p+1)^2) (set_allocate_hexahedronHO_ijk_indices and
 * set_allocate_quadrilateralHO_ij_indices, mod_ijk_indices.f90:61-154;
 * mod_meshConversorTool.f90:161-162).  Any output pointer may be NULL. */
int sodhip_order_tables(int porder, int *a2ijk, int *gmsh2ijk, int *vtk2ijk, int *a2ij,
                        int *gmsh2ij, int *vtk2ij);

/* Structured box of nx*ny*nz HEX((p+1)^3) elements on [0,L]^3 in SOD2D's partition
 * format (mod_mpi_mesh.f90:15-48): npoin nodes including periodic slaves,
 * connecParOrig/connecParWork (nelem,nnode) 1-based, workingNodesPar (npoin_w),
 * coordPar (npoin,3).  Element-local node a follows the gmsh high-order order.
 * periodic != 0 folds the x=L, y=L, z=L planes onto x=0, y=0, z=0 (slave ->
 * master, as create_working_lists_parallel, mod_meshConversorTool.f90:4917-5008).
 * jitter > 0 moves every interior-free node (non-periodic meshes only) by
 * U(-jitter,jitter) times the local minimum GLL spacing per axis, from seed. */
int sodhip_box_mesh_sizes(int porder, int nx, int ny, int nz, int periodic, int *nelem, int *npoin,
                          int *npoin_w);
int sodhip_box_mesh(int porder, int nx, int ny, int nz, int periodic, const double *lengths,
                    double jitter, unsigned long long seed, int *connec_orig, int *connec_work,
                    int *lpoin_w, double *coord);

/* elem_jacobian (jacobian_oper.f90:7-140): He(3,3,ngaus,nelem), gpvol(1,ngaus,nelem)
 * from connecParOrig and coordPar.  Returns SODHIP_ERR_ARG on a non-positive
 * Jacobian determinant. */
int sodhip_elem_jacobian(int porder, int nelem, int npoin, const int *connec, const double *coord,
                         double *He, double *gpvol);
/* lumped_mass_spectral + invMl (mass_matrix.f90:11-42; CFDSolverBase.f90:1771-1782) */
int sodhip_lumped_mass(int porder, int nelem, int npoin, const int *connec_work,
                       const double *gpvol, int npoin_w, const int *lpoin_w, double *Ml,
                       double *invMl);
/* char_length (mod_geom.f90:13-37) -> helem(nelem); char_length_spectral
 * (mod_geom.f90:39-66) -> helem_l(nelem,nnode); both from connecParOrig */
int sodhip_char_length(int porder, int nelem, int npoin, const int *connec_orig,
                       const double *coord, double *helem, double *helem_l);

/* Element clusters of a p=3 mesh (host arrays; set-up, no device needed), the
 * tables of the brick-local assembly the fused FP64 stage uses: elements are
 * grouped greedily into clusters of up to 8 (2x2x2 bricks on a box mesh
 * numbered along i, j, k), each with at most 344 distinct nodes.  No
 * counterpart in the reference (its assembly is an atomic scatter,
 * elem_convec.f90:451-465).  connec(nelem,64) 1-based column-major; lpoin_w
 * 1-based.  sizes[0..3] receive ncl, nloc (local nodes over all clusters),
 * ns (working nodes finished by the surface gather) and the number of their
 * records; each non-NULL array receives its table: cel(ncl*8) element ids
 * (0-based, -1 = empty slot, ascending within a cluster), cloc(ncl*8*64)
 * the local index of the node at each lexicographic position of each slot's
 * element, cnb(ncl+1) local-node offsets, cnode(nloc) 0-based node ids
 * (ascending within a cluster), ckind(nloc) 0 interior working node / 1
 * surface working node / 2 not working, snode(ns) 1-based nodes in working-list
 * order, srow(ns+1), ssrc (records cnb[c]+l, ascending cluster per node).
 * SODHIP_ERR_UNSUPPORTED when an element lists a node twice. */
int sodhip_element_clusters(int porder, int nelem, int npoin, const int *connec, int npoin_w,
                            const int *lpoin_w, int *sizes, int *cel, int *cloc, int *cnb,
                            int *cnode, int *ckind, int *snode, int *srow, int *ssrc);

/* ------------------------------------------------------------------------ */
/* Device context                                                            */
/* ------------------------------------------------------------------------ */
int sodhip_create(int device, sodhip_ctx *ctx);
int sodhip_destroy(sodhip_ctx ctx);
/* stream (hipStream_t) on which all work of ctx is enqueued; NULL = null stream */
int sodhip_set_stream(sodhip_ctx ctx, void *stream);
int sodhip_synchronize(sodhip_ctx ctx);

/* Lifetime of the cached connectivity tables.  The first call with a given
 * connec (device pointer, porder, nelem, npoin) builds element-major and
 * node-ordered copies of it, and the first call with a given lpoin_w pointer
 * builds its working-node incidence; later calls with the same pointers reuse
 * them without reading the arrays again.  A caller that frees such an array, or
 * rewrites it with another mesh, calls sodhip_forget_mesh with its device
 * address first (sodhip_exit_data does this for the range it releases).  No
 * counterpart in the reference: OpenACC's present table has no derived data. */
int sodhip_forget_mesh(sodhip_ctx ctx, const void *connec_or_lpoin_w);

/* Scratch buffers of ctx that a captured HIP graph records by address: the
 * per element-node residual partials and the step's reduction partials.
 * sodhip_reserve_scratch grows them to at least the given sizes; while pinned
 * (sodhip_pin_scratch(ctx, 1)) they are never reallocated, and a call needing
 * more fails with SODHIP_ERR_ARG instead of moving them under the graph.  Host-
 * read diagnostics (dt, EK, dissipation, max Mach) use a buffer of their own.
 * No counterpart in the reference. */
int sodhip_reserve_scratch(sodhip_ctx ctx, size_t scratch_bytes, size_t scratch2_bytes);
int sodhip_pin_scratch(sodhip_ctx ctx, int pin);

/* OpenACC-style present table for callers that hold host arrays (the Fortran
 * shim): enter data create/copyin, update device/host, exit data, and the
 * host -> device address lookup (NULL when not present). */
int sodhip_enter_data(sodhip_ctx ctx, const void *host, size_t bytes, int copyin);
int sodhip_update_device(sodhip_ctx ctx, const void *host, size_t bytes);
int sodhip_update_host(sodhip_ctx ctx, void *host, size_t bytes);
int sodhip_exit_data(sodhip_ctx ctx, const void *host, int copyout);
void *sodhip_deviceptr(sodhip_ctx ctx, const void *host);
/* Device-side copy of `bytes` from the mirror of src_host to the mirror of
 * dst_host, on the context stream: the array assignments the reference runs
 * as `!$acc kernels` on present data, such as the level copy
 * eta(:,1) = eta(:,2) of CFDSolverBase_evalTimeIteration
 * (CFDSolverBase.f90:1929-2164). */
int sodhip_copy_present(sodhip_ctx ctx, const void *src_host, void *dst_host, size_t bytes);

/* Kernel timing with HIP events on the context stream: when enabled, every
 * launch of kernel class `kind` is bracketed by events; sodhip_prof_get returns
 * the summed duration (ms) and the number of launches since the last reset.
 * kind: 0 element operator kernel, 1 node gather/update kernel, 2 sensor,
 * 3 other nodal kernels, 4 the LPS projections (full_proj_ijk's element
 * kernel and gather). */
/* Launch shape of the fused RK stage on this context: blocks per CU of the
 * element kernel (1..64; with a grid-stride walk this sets the elements each
 * workgroup takes), whether the geometry moves through LDS in 16-byte lane
 * rows when He and gpvol are aligned for it (1) or in 4-byte rows (0), and
 * the gather/update block size (64, 128 or 256).  Defaults 16, 1, 256.  The
 * results do not depend on the shape.  No counterpart in the reference: the
 * autotuner (sod2d_amd/autotune.py) sets it. */
int sodhip_set_stage_config(sodhip_ctx ctx, int elem_blocks_per_cu, int geo_rows16,
                            int gather_threads);
int sodhip_get_stage_config(sodhip_ctx ctx, int *elem_blocks_per_cu, int *geo_rows16,
                            int *gather_threads);

int sodhip_prof_enable(sodhip_ctx ctx, int enable);
int sodhip_prof_get(sodhip_ctx ctx, int kind, double *total_ms, long long *count);
int sodhip_prof_reset(sodhip_ctx ctx);

/* ------------------------------------------------------------------------ */
/* Joined instances                                                          */
/* ------------------------------------------------------------------------ */
/* Independent meshes joined into one disconnected mesh (an ensemble of runs:
 * elements, nodes and working-node lists concatenated with offsets), so that
 * one launch per kernel covers every instance.  No counterpart in the
 * reference, which runs one mesh.  The fields that differ between instances
 * (device arrays):
 *   wseg(ninst+1)      instance i owns the lpoin_w positions wseg[i]..wseg[i+1]-1
 *                      (its entropy normalisation runs over them);
 *   elem_inst(nelem)   the instance of each element;
 *   p_inf, rho_inf, mu_factor (npoin): the reference state of the entropy and
 *                      the Sutherland factor of each node's instance (the
 *                      reference's mu_factor(npoin)).
 * The fused stage takes the Sutherland factor per element through
 * sodhip_stage_params.mu_factor_e. */
typedef struct {
    int ninst;
    const int *wseg;
    const int *elem_inst;
    const double *p_inf, *rho_inf, *mu_factor;
} sodhip_instances;

/* ------------------------------------------------------------------------ */
/* Drop-in operators (device pointers, reference layouts)                    */
/* ------------------------------------------------------------------------ */

/* full_convec_ijk (elem_convec.f90:227-470).  initialize != 0 zeroes Rmass,
 * Rmom, Rener first; fact multiplies the element residual (the reference's
 * optional `fact`; the caller resolves its SAVE'd default).  connec(nelem,nnode),
 * He(3,3,ngaus,nelem), gpvol(1,ngaus,nelem), dlxigp_ip(ngaus,3,p+1),
 * invAtoIJK(p+1,p+1,p+1), gmshAto*(nnode), nodal u(npoin,3), q(npoin,3), rho, pr,
 * E (the energy slot; updateF passes the enthalpy aux_h, time_integ_ls.f90:605-606),
 * Rmass(npoin), Rmom(npoin,3), Rener(npoin).  Assembly is a deterministic
 * node-ordered gather (no atomics): results do not depend on scheduling. */
int sodhip_full_convec_ijk_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
                               const double *He, const double *gpvol, const double *dlxigp_ip,
                               const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
                               const int *gmshAtoK, const double *u, const double *q,
                               const double *rho, const double *pr, const double *E,
                               double *Rmass, double *Rmom, double *Rener, int initialize,
                               double fact);
int sodhip_full_convec_ijk_f32(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
                               const float *He, const float *gpvol, const float *dlxigp_ip,
                               const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
                               const int *gmshAtoK, const float *u, const float *q,
                               const float *rho, const float *pr, const float *E, float *Rmass,
                               float *Rmom, float *Rener, int initialize, float fact);

/* full_diffusion_ijk (elem_diffu.f90:12-223).  mu_e and mu_sgs are (nelem,ngaus);
 * mu_sgs may be NULL (identically zero).  Ml is not used by the reference and
 * is not an argument here.  c_rho and c_ener are the module globals of
 * mod_numerical_params.f90 (default 1). */
int sodhip_full_diffusion_ijk_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                  const int *connec, const double *He, const double *gpvol,
                                  const double *dlxigp_ip, const int *invAtoIJK,
                                  const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                                  double Cp, double Pr, const double *rho_n, const double *rho,
                                  const double *u, const double *Tem, const double *mu_fluid,
                                  const double *mu_e, const double *mu_sgs, double c_rho,
                                  double c_ener, double *Rmass, double *Rmom, double *Rener,
                                  int initialize, double fact);
int sodhip_full_diffusion_ijk_f32(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                  const int *connec, const float *He, const float *gpvol,
                                  const float *dlxigp_ip, const int *invAtoIJK,
                                  const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                                  float Cp, float Pr, const float *rho_n, const float *rho,
                                  const float *u, const float *Tem, const float *mu_fluid,
                                  const float *mu_e, const float *mu_sgs, float c_rho,
                                  float c_ener, float *Rmass, float *Rmom, float *Rener,
                                  int initialize, float fact);

/* LPS stabilisation (elem_stab.f90), the reference's default viscous path
 * (flag_lps_stab, time_integ_ls.f90:266-268,607-611).
 *
 * comp_tau (elem_stab.f90:589-616): tau(e) = max over the element's nodes of
 * helem_k(e) * (c_species_stab/p) * (c + |u|); the sound speed is dropped when
 * flag_drop_c_in_envit is set.  c_species_stab is the module constant of
 * mod_numerical_params.f90 (0.5).  The reference's dt argument is unused. */
int sodhip_comp_tau_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
                        const double *csound, const double *u, const double *helem_k,
                        double c_species_stab, int flag_drop_c_in_envit, double *tau);

/* full_proj_ijk (elem_stab.f90:441-586): lumped L2 projections of grad rho
 * (ProjMass), grad T (ProjEner) and of the rows of 2 S - 2/3 div u I
 * (ProjMX/MY/MZ), each (npoin,3); overwritten, then multiplied by invMl over
 * the working nodes.  The reference's Ngp, Cp, Pr and Ml are unused. */
int sodhip_full_proj_ijk_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                             const int *connec, const int *lpoin_w, const double *He,
                             const double *gpvol, const double *dlxigp_ip,
                             const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
                             const int *gmshAtoK, const double *rho, const double *u,
                             const double *Tem, const double *invMl, double *ProjMass,
                             double *ProjEner, double *ProjMX, double *ProjMY, double *ProjMZ);

/* full_diff_stab_ijk (elem_stab.f90:207-438): the viscous operator with the
 * entropy-viscosity coefficients applied per quadrature point and the LPS
 * fluxes -c_lps tau_stab(e) (Proj - grad); c_lps is c_lps_comp
 * (mod_numerical_params.f90:31, a default-real 0.1).  initialize/fact as
 * full_diffusion_ijk; mu_sgs may be NULL (zero). */
int sodhip_full_diff_stab_ijk_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                  const int *connec, const double *He, const double *gpvol,
                                  const double *dlxigp_ip, const int *invAtoIJK,
                                  const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                                  double Cp, double Pr, const double *rho_n, const double *rho,
                                  const double *u, const double *Tem, const double *mu_fluid,
                                  const double *mu_e, const double *mu_sgs,
                                  const double *ProjMass, const double *ProjEner,
                                  const double *ProjMX, const double *ProjMY,
                                  const double *ProjMZ, const double *tau_stab, double c_rho,
                                  double c_ener, double c_lps, double *Rmass, double *Rmom,
                                  double *Rener, int initialize, double fact);

/* generic_scalar_convec_ijk (elem_convec.f90:472-594): q is the flux f(npoin,3),
 * eta the transported scalar, u the velocity; output Rconvec(npoin).  q may be
 * NULL: the flux is then f = u * eta, formed in the kernel (the time loop's
 * f_eta = u(:,:,1) * eta(:,1), time_integ_ls.f90:277-281 and thermo_entropy). */
int sodhip_generic_scalar_convec_ijk_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                         const int *connec, const double *He, const double *gpvol,
                                         const double *dlxigp_ip, const int *invAtoIJK,
                                         const int *gmshAtoI, const int *gmshAtoJ,
                                         const int *gmshAtoK, const double *q, const double *eta,
                                         const double *u, double *Rconvec, int initialize,
                                         double fact);
/* The same operator on real(4) arrays (the reference's default USE_RP=4 build,
 * CMakeLists.txt:30): FP32 arithmetic. */
int sodhip_generic_scalar_convec_ijk_f32(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                         const int *connec, const float *He, const float *gpvol,
                                         const float *dlxigp_ip, const int *invAtoIJK,
                                         const int *gmshAtoI, const int *gmshAtoJ,
                                         const int *gmshAtoK, const float *q, const float *eta,
                                         const float *u, float *Rconvec, int initialize,
                                         float fact);

/* Module globals read by smart_visc_spectral (mod_numerical_params.f90:1-175). */
typedef struct {
    double ce, cmax, cglob, Rgas; /* Rgas = nscbc_Rgas_inf */
    int flag_normalise_entropy, flag_drop_c_in_envit;
} sodhip_visc_params;

/* smart_visc_spectral (mod_entropy_viscosity.f90:14-111) with the filter tables
 * of init_filters (mod_filters.f90:15-55).  helem(nelem,nnode) and
 * helem_k(nelem) as in the reference's dummy arguments (callers pass helem_l and
 * helem).  Outputs mu_e(nelem,ngaus) and mue_l(nelem,nnode).  The entropy
 * normalisation is a deterministic two-level FP64 reduction. */
int sodhip_smart_visc_spectral_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                                   const int *connec, const int *lpoin_w, const double *Reta,
                                   double gamma_gas, const double *rho, const double *u,
                                   const double *Tem, const double *eta, const double *helem,
                                   const double *helem_k, const int *invAtoIJK,
                                   const sodhip_visc_params *prm, double *mu_e, double *mue_l);
/* The same sensor on joined instances: each instance's entropy normalisation
 * over its own working nodes, with the blocking of a single mesh (the bits of
 * the instance run alone). */
int sodhip_smart_visc_spectral_ens_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                       int npoin_w, const int *connec, const int *lpoin_w,
                                       const sodhip_instances *ens, const double *Reta,
                                       double gamma_gas, const double *rho, const double *u,
                                       const double *Tem, const double *eta,
                                       const double *helem, const double *helem_k,
                                       const int *invAtoIJK, const sodhip_visc_params *prm,
                                       double *mu_e, double *mue_l);
/* The same sensor on real(4) arrays (USE_RP=4): FP32 inputs and outputs, FP64
 * arithmetic and entropy normalisation. */
int sodhip_smart_visc_spectral_f32(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                                   const int *connec, const int *lpoin_w, const float *Reta,
                                   float gamma_gas, const float *rho, const float *u,
                                   const float *Tem, const float *eta, const float *helem,
                                   const float *helem_k, const int *invAtoIJK,
                                   const sodhip_visc_params *prm, float *mu_e, float *mue_l);

/* smart_visc_spectral_imex (mod_entropy_viscosity.f90:113-227), the sensor of the
 * reference's live LS-RK path (time_integ_ls.f90:292,455): per element
 * beta_e = max rho helem_k (cmax/p)(|u| + c) and mu_e = cglob min(ce rho |Reta| /
 * norm (helem_k/p)^2, beta_e), no filter; mue_l receives the incoming mu_e.
 * helem_k is the caller's helem(nelem) (the reference's helem(nelem,nnode)
 * argument is not read).  With flag_normalise_entropy and flag_high_mach the
 * norm is the minimum density over the working nodes and prm->ce is set to
 * ce_comp, as the reference assigns its module global. */
int sodhip_smart_visc_spectral_imex_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                        int npoin_w, const int *connec, const int *lpoin_w,
                                        const double *Reta, const double *rho, const double *u,
                                        const double *csound, const double *eta,
                                        const double *helem_k, sodhip_visc_params *prm,
                                        double ce_comp, int flag_high_mach, double *mu_e,
                                        double *mue_l);

/* eval_divergence (mod_operators.f90:485-570): divU(npoin) = assembled gpvol * div u
 * (zeroed first). */
int sodhip_eval_divergence_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                               const int *connec, const double *He, const double *gpvol,
                               const double *dlxigp_ip, const int *invAtoIJK, const int *gmshAtoI,
                               const int *gmshAtoJ, const int *gmshAtoK, const double *u,
                               double *divU);

/* compute_vorticity with makeConti (mod_operators.f90:684-775): curlU(npoin,3) =
 * assembled eps_ijk du_j/dx_k (the reference's leviCivi of CFDSolverBase.f90:
 * 1605-1611, i.e. minus the curl), divided by lelpn on the working nodes
 * (makeConti; with lelpn NULL the sums are left undivided).  The reference's
 * dNgp, leviCivi and atoIJK arguments are not needed. */
int sodhip_compute_vorticity_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                                 const int *lpoin_w, const int *connec, const int *lelpn,
                                 const double *He, const double *dlxigp_ip,
                                 const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
                                 const int *gmshAtoK, const double *u, double *curlU);

/* eval_divergence and compute_vorticity of the same u in one element pass (the
 * pair the live path calls back to back, time_integ_ls.f90:459-460). */
int sodhip_div_curl_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                        const int *lpoin_w, const int *connec, const int *lelpn, const double *He,
                        const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                        const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                        const double *u, double *divU, double *curlU);

/* ducros_visc_limit (mod_entropy_viscosity.f90:277-314): mu_e(nelem,ngaus) scaled
 * in place by max(phi - ducros_min_val, 0) phi2 + ducros_min_val with divUl =
 * divU/Ml, phi = min(4/3 divUl^2/(divUl^2 + |vorti|^2 + eps), 1), phi2 =
 * divUl^2/(divUl^2 + (|u|/helem)^2 + eps); helem(nelem,nnode) is the caller's
 * helem_l. */
int sodhip_ducros_visc_limit_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                 const int *connec, const double *u, const double *divU,
                                 const double *vorti, const double *Ml, const double *helem,
                                 double ducros_min_val, double *mu_e);

/* elemPerNode (mod_geom.f90:68-99), host: lelpn(npoin) = incidences of each node
 * in connec(nelem,nnode). */
int sodhip_elem_per_node(int nelem, int nnode, int npoin, const int *connec, int *lelpn);

/* lumped_solver_scal_opt / lumped_solver_vect_opt (mod_solver.f90:40-104) */
int sodhip_lumped_solver_scal_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                  const double *invMl, double *R);
int sodhip_lumped_solver_vect_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                  const double *invMl, double *R);
/* real(4) arrays (USE_RP=4), FP32 arithmetic */
int sodhip_lumped_solver_scal_f32(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                  const float *invMl, float *R);
int sodhip_lumped_solver_vect_f32(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                  const float *invMl, float *R);

/* RK stage update with the lumped-mass inverse fused in (the reference has these
 * as inline loops: mod_solver.f90:40-104 then time_integ_ls.f90:323-331,350-358):
 * over working nodes, R <- R*invMl (written back, as the reference's in-place
 * lumped solve), then y <- lam1*y1 + lams*y + gam*dt*R for rho, E and q(:,1:3).
 * invMl == NULL: R already carries the mass inverse and is left unchanged (the
 * first stage of rk_4_ls_main, time_integ_ls.f90:323-331). */
int sodhip_rk_stage_update_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                               const double *invMl, double dt, double lam1, double lams,
                               double gam, const double *rho1, const double *q1,
                               const double *E1, double *rho, double *q, double *E,
                               double *Rmass, double *Rmom, double *Rener);
/* real(4) arrays (USE_RP=4), FP32 arithmetic */
int sodhip_rk_stage_update_f32(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                               const float *invMl, float dt, float lam1, float lams, float gam,
                               const float *rho1, const float *q1, const float *E1, float *rho,
                               float *q, float *E, float *Rmass, float *Rmom, float *Rener);

/* The node update of one stage of the classical RK4 integrator rk_4_main
 * (time_integ.f90; compiled out in the reference by #if 0, restated from its
 * loops), with the lumped-mass inverse fused in.  Over working nodes:
 * R <- R*invMl, written back (lumped_solver_*, time_integ.f90:431-435; skipped
 * when invMl is NULL); Rsum <- Rsum + b*R (:442-451; first != 0: Rsum starts
 * from zero, as the step's initialisation :234-240 leaves it); then, unless
 * last, the next stage's state aux = y - dt*a_next*R (:266-275), or, at the
 * last stage, y <- y - dt*Rsum (:460-468).  R has the sign of the operators
 * called without fact (aux_fact = 1): y advances by -dt*R.  aux must not alias
 * y; aux may be NULL at the last stage. */
int sodhip_rk4_stage_update_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                const double *invMl, double dt, double a_next, double b,
                                int first, int last, double *rho, double *q, double *E,
                                double *aux_rho, double *aux_q, double *aux_E, double *Rmass,
                                double *Rmom, double *Rener, double *Rmass_sum,
                                double *Rmom_sum, double *Rener_sum);

/* ------------------------------------------------------------------------ */
/* Nodal pieces of the time step (FP64, device pointers)                    */
/* ------------------------------------------------------------------------ */

/* updateF "Update u and EOS" (time_integ_ls.f90:574-587) over working nodes:
 * u = q/rho, e_int, pr, aux_h = gamma/(gamma-1) pr/rho (may be NULL), Tem. */
int sodhip_eos_stage_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                         double gamma_gas, double Rgas, const double *rho, const double *q,
                         const double *E, double *u, double *e_int, double *pr, double *aux_h,
                         double *Tem);
/* rk_4_ls_main's end-of-step EOS (time_integ_ls.f90:391-407): adds csound, machno */
int sodhip_eos_step_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                        double gamma_gas, double Rgas, const double *rho, const double *q,
                        const double *E, double *u, double *e_int, double *pr, double *csound,
                        double *machno, double *Tem);
/* sutherland_viscosity (mod_fluid_viscosity.f90:21-41) over all npoin; mu_factor
 * may be NULL, then the scalar mu_factor_scalar applies to every node. */
int sodhip_sutherland_f64(sodhip_ctx ctx, int npoin, const double *Tem, const double *mu_factor,
                          double mu_factor_scalar, double *mu_fluid);
/* thermo_entropy (mod_entropy_viscosity.f90:411-433); eta(npoin,2), feta(npoin,3).
 * feta may be NULL (then u is not read): the flux is not stored, for callers
 * that pass q = NULL to sodhip_generic_scalar_convec_ijk_f64. */
int sodhip_thermo_entropy_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                              double gamma_gas, double p_inf, double rho_inf, const double *rho,
                              const double *pr, const double *u, double *eta, double *feta);
/* entropy residual of rk_4_ls_main (time_integ_ls.f90:437-443): Reta(npoin,2),
 * eta(npoin,2); auxReta = 1.5 Reta2 - 0.5 Reta1 + factor_comp (eta2-eta1)/dt,
 * then Reta1 <- Reta2. */
int sodhip_entropy_residual_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                double factor_comp, double dt, double *Reta, const double *eta,
                                double *auxReta);
/* The node updates at the end of rk_4_ls_main in one pass over the working
 * nodes: sodhip_eos_step_f64 (time_integ_ls.f90:391-407), then
 * sodhip_thermo_entropy_f64 with feta = NULL, then sutherland_viscosity
 * (mod_fluid_viscosity.f90:21-41) with the uniform mu_factor.  Each output
 * has the bits those three calls give on the working nodes; mu_fluid is not
 * written elsewhere (the reference's loop covers all npoin, where nodes outside
 * lpoin_w keep their temperature).  No counterpart in the reference. */
int sodhip_eos_entropy_step_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                double gamma_gas, double Rgas, double p_inf, double rho_inf,
                                double mu_factor, const double *rho, const double *q,
                                const double *E, double *u, double *e_int, double *pr,
                                double *csound, double *machno, double *Tem, double *eta,
                                double *mu_fluid);
/* sodhip_eos_entropy_step_f64 on joined instances: p_inf, rho_inf and
 * mu_factor per node from ens. */
int sodhip_eos_entropy_step_ens_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                    double gamma_gas, double Rgas, const sodhip_instances *ens,
                                    const double *rho, const double *q, const double *E,
                                    double *u, double *e_int, double *pr, double *csound,
                                    double *machno, double *Tem, double *eta, double *mu_fluid);
/* The entropy residual of rk_4_ls_main (time_integ_ls.f90:429-443) in one
 * call: sodhip_generic_scalar_convec_ijk_f64 with q = NULL (f = u * eta(:,1)),
 * initialize = 1 and fact = 1 into Reta(:,2), sodhip_lumped_solver_scal_f64 on
 * it, then sodhip_entropy_residual_f64 (auxReta, Reta(:,1) <- Reta(:,2)), the
 * last two on the gathered values of each working node.  Reta and eta are
 * (npoin,2).  Each output has the bits of the three calls on the working
 * nodes; Reta(:,2) is not written elsewhere (the separate gather writes 0 at
 * nodes no element references).  No counterpart in the reference. */
int sodhip_entropy_convec_residual_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                       int npoin_w, const int *connec, const int *lpoin_w,
                                       const double *He, const double *gpvol,
                                       const double *dlxigp_ip, const int *invAtoIJK,
                                       const int *gmshAtoI, const int *gmshAtoJ,
                                       const int *gmshAtoK, const double *u, const double *eta,
                                       const double *invMl, double factor_comp, double dt,
                                       double *Reta, double *auxReta);

/* ------------------------------------------------------------------------ */
/* MLP surrogate of the entropy-viscosity sensor (no counterpart in the       */
/* reference; it stands in for smart_visc_spectral, which stays its          */
/* per-element fallback)                                                     */
/* ------------------------------------------------------------------------ */
/* The MLP 12 -> 64 -> 64 -> 1 with SiLU activations.  w: the 5057 FP32
 * parameters (W1 12x64 row = feature, b1 64, W2 64x64 row = input, b2 64,
 * w3 64, b3); wq: their BF16 operand copy (5184 values: W1 padded to 16 rows,
 * W2, w3; 16-byte aligned), refreshed by sodhip_mlp_quantize and by every
 * training step; m, v:
 * Adam moments (5057 each, training only); lo/hi: the range of each feature
 * over the training data (the deployed sensor's fallback test).  Device
 * pointers except lo/hi. */
typedef struct {
    float *w;
    void *wq;
    float *m, *v;
    float lo[12], hi[12];
} sodhip_mlp;
/* sizes of w and wq, and the number of features (host) */
int sodhip_mlp_sizes(int *nparams, int *nq, int *nfeatures);
int sodhip_mlp_quantize(sodhip_ctx ctx, const sodhip_mlp *model);
/* The twelve features of every element (p = 3) from the inputs of
 * smart_visc_spectral (same arguments; the entropy normalisation as that call
 * forms it): X(nelem,16) row-major FP32, features 0..11 then zeros, and
 * beta(nelem) the element's cap estimate.  The features read 16 of the
 * element's 64 nodes, its 8 vertices and its 8 interior nodes.  Over these,
 * with b = rho helem_k cmax/p (|u| + c), beta_e = max b, r = ce rho|Reta|/norm
 * h^2 / beta_e, m = min(r,1) and the log scale g(x) = 1 + log10(x + 1e-10)/10:
 * g of the max, mean, min and standard deviation of m and of its mean over
 * the vertices and over the interior nodes; share of nodes with r >= 1;
 * log10(1 + max r)/4; largest Mach number; mean of |u|/(|u|+c);
 * 10 (max-min)/mean of rho and of T. */
int sodhip_sensor_features_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                               const int *connec, const int *lpoin_w, const double *Reta,
                               double gamma_gas, const double *rho, const double *u,
                               const double *Tem, const double *eta, const double *helem,
                               const double *helem_k, const int *invAtoIJK,
                               const sodhip_visc_params *prm, float *X, double *beta);
/* The same features on joined instances (each instance's entropy
 * normalisation), for dumping training data from an ensemble run. */
int sodhip_sensor_features_ens_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                   int npoin_w, const int *connec, const int *lpoin_w,
                                   const sodhip_instances *ens, const double *Reta,
                                   double gamma_gas, const double *rho, const double *u,
                                   const double *Tem, const double *eta, const double *helem,
                                   const double *helem_k, const int *invAtoIJK,
                                   const sodhip_visc_params *prm, float *X, double *beta);
/* y(n) = MLP(X(n,16)) on the matrix cores (BF16 operands, FP32 accumulation) */
int sodhip_mlp_forward(sodhip_ctx ctx, const sodhip_mlp *model, int n, const float *X, float *y);
/* Gradient of L = (1/n) sum (y - t)^2 over the batch: forward (y, and the
 * pre-activations into Z(2*n*64)), then the backward pass; gpart receives
 * nwaves partial gradients of 5057 values (nwaves a multiple of 4), whose sum
 * in order is the gradient. */
int sodhip_mlp_gradient(sodhip_ctx ctx, const sodhip_mlp *model, int n, const float *X,
                        const float *t, float *y, float *Z, float *gpart, int nwaves);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8; step = 1, 2, ...) on the
 * gradient of sodhip_mlp_gradient, then the BF16 copy. */
int sodhip_mlp_train_step(sodhip_ctx ctx, sodhip_mlp *model, int n, const float *X,
                          const float *t, int step, float lr, float *y, float *Z, float *gpart,
                          int nwaves);
/* The surrogate sensor: smart_visc_spectral's arguments and outputs (p = 3).
 * Per element mu_e(e,:) = cglob * beta_e * (10^(10 (y_e - 1)) - 1e-10) with
 * y_e the MLP of the element's features (the mean of mu_e/(cglob beta_e) over
 * the element on the log scale g), unless a feature lies outside [lo, hi] or
 * y_e outside [0, 1]: such an element is flagged and takes smart_visc_spectral's values
 * (mu_e and mue_l; mue_l is written for those elements only).  One kernel
 * forms features, MLP output and mu_e; a second one runs the plain sensor on
 * the listed fallback elements.  flag(nelem) (device) receives 1 for a
 * fallback element, else 0; X(nelem,16), beta(nelem) and y(nelem) (device,
 * each may be NULL) receive the features, cap estimates and MLP outputs as
 * sodhip_sensor_features_f64 and sodhip_mlp_forward form them; nfallback
 * (device, may be NULL) receives the number of fallback elements. */
int sodhip_surrogate_visc_spectral_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                       int npoin_w, const int *connec, const int *lpoin_w,
                                       const double *Reta, double gamma_gas, const double *rho,
                                       const double *u, const double *Tem, const double *eta,
                                       const double *helem, const double *helem_k,
                                       const int *invAtoIJK, const sodhip_visc_params *prm,
                                       const sodhip_mlp *model, double *mu_e, double *mue_l,
                                       float *X, double *beta, float *y, int *flag,
                                       int *nfallback);

/* ------------------------------------------------------------------------ */
/* Fused Runge-Kutta stage                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    double gamma_gas, Rgas, Cp, Pr; /* Pr: the Prt passed by updateF */
    double mu_factor;               /* Sutherland factor (flag_mu_factor), uniform */
    double c_rho, c_ener;           /* mod_numerical_params.f90 globals (default 1) */
    double fact;                    /* residual sign: updateF passes -1 */
    double p_ref;                   /* uniform pressure removed before the operator
                                       arithmetic (only grad p enters); 0 = none */
    const double *mu_factor_e;      /* NULL, or a Sutherland factor per element (device,
                                       nelem): the reference's mu_factor(npoin) when it is
                                       constant over each element, as on a mesh joining
                                       instances of different Reynolds numbers */
} sodhip_stage_params;

/* One stage of rk_4_ls_main (time_integ_ls.f90:339-361) with updateF
 * (time_integ_ls.f90:489-691, flag_lps_stab = .false., no sources, one
 * partition) fused in: at the current state (rho, q, E) it forms the primitives
 * and Sutherland's viscosity, evaluates R = fact*(full_convec_ijk with the
 * enthalpy in the energy slot + full_diffusion_ijk with mu_sgs = 0), applies the
 * lumped inverse, and updates y <- lam1*y1 + lams*y + gam*dt*R over working
 * nodes.  op_precision 8: He, gpvol, dlxigp_ip and mu_e are double; 4: they are
 * float and the operator arithmetic runs in FP32, while the state, the gather,
 * the mass inverse and the stage combine stay FP64.  Rmass/Rmom/Rener (all or
 * none) receive R*invMl (the reference keeps it for the next step's stage 1). */
int sodhip_rk_stage_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                        const int *connec, const int *lpoin_w, int op_precision, const void *He,
                        const void *gpvol, const void *dlxigp_ip, const int *invAtoIJK,
                        const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                        const void *mu_e, const double *invMl, const sodhip_stage_params *prm,
                        double dt, double lam1, double lams, double gam, const double *rho1,
                        const double *q1, const double *E1, double *rho, double *q, double *E,
                        double *Rmass, double *Rmom, double *Rener);

/* sodhip_rk_stage_f64 for the last stage of a step at fixed dt, fused with the
 * first stage of the next step of the reference sequencing: besides the new
 * state y (rho, q, E), stores y_next = 1 * y + 0 * 0 + gam_next * dt * R over
 * the working nodes, the stage update of time_integ_ls.f90:323-331 with the
 * residual this stage keeps (R as stored by sodhip_rk_stage_f64).  No
 * counterpart in the reference, which evaluates that update at the start of
 * the next step; valid while dt is unchanged between the two steps.  y_next
 * must not alias y; it may alias y1 (each node reads y1 before writing). */
int sodhip_rk_stage_next_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                             const int *connec, const int *lpoin_w, int op_precision,
                             const void *He, const void *gpvol, const void *dlxigp_ip,
                             const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
                             const int *gmshAtoK, const void *mu_e, const double *invMl,
                             const sodhip_stage_params *prm, double dt, double lam1, double lams,
                             double gam, const double *rho1, const double *q1, const double *E1,
                             double *rho, double *q, double *E, double *Rmass, double *Rmom,
                             double *Rener, double gam_next, double *rho_next, double *q_next,
                             double *E_next);

/* One stage of the classical RK4 integrator rk_4_main (time_integ.f90:250-470,
 * compiled out in the reference by #if 0) with the RHS fused in: at the stage
 * state (srho, sq, sE) -- y itself at the first stage, where a_1 = 0 -- it forms
 * the primitives and Sutherland's viscosity, evaluates R = fact*(full_convec_ijk
 * with the enthalpy slot + full_diffusion_ijk, mu_sgs = 0) (the reference's
 * rk_4_main calls them without fact: prm->fact = 1), applies the lumped
 * inverse, and performs the node update of sodhip_rk4_stage_update_f64 with
 * the stage's a_next (a_{i+1}) and b (b_i).  The stage state may be the aux
 * arrays themselves (the update reads every node's state in an earlier
 * launch).  op_precision as sodhip_rk_stage_f64.  Rmass/Rmom/Rener (all or
 * none) receive R*invMl. */
int sodhip_rk4_stage_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                         const int *connec, const int *lpoin_w, int op_precision, const void *He,
                         const void *gpvol, const void *dlxigp_ip, const int *invAtoIJK,
                         const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                         const void *mu_e, const double *invMl, const sodhip_stage_params *prm,
                         double dt, double a_next, double b, int first, int last, double *rho,
                         double *q, double *E, const double *srho, const double *sq,
                         const double *sE, double *aux_rho, double *aux_q, double *aux_E,
                         double *Rmass_sum, double *Rmom_sum, double *Rener_sum, double *Rmass,
                         double *Rmom, double *Rener);

/* sodhip_rk_stage_f64 without any device allocation (same arguments, same
 * kernels, same results).  No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops).
 * The residual records live in a library-owned device array of 64 MiB per GPU,
 * mapped by sodhip_create and held until the process ends, instead of the
 * context's scratch; a mesh whose records exceed it (nelem*nnode*40 bytes with
 * FP64 operators, *24 with FP32 operators) is rejected with SODHIP_ERR_ARG.
 * The mesh and working-node tables are looked up, never built: an allocating
 * call on the same connec and lpoin_w (sodhip_rk_stage_f64) must come first,
 * else SODHIP_ERR_ARG.  Launch timing is off during the call.  Every call on
 * one GPU shares the array, which keeps the last call's records: calls that
 * overlap on one GPU (several contexts or streams of it) race. */
int sodhip_rk_stage_noalloc_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                                const int *connec, const int *lpoin_w, int op_precision,
                                const void *He, const void *gpvol, const void *dlxigp_ip,
                                const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
                                const int *gmshAtoK, const void *mu_e, const double *invMl,
                                const sodhip_stage_params *prm, double dt, double lam1,
                                double lams, double gam, const double *rho1, const double *q1,
                                const double *E1, double *rho, double *q, double *E,
                                double *Rmass, double *Rmom, double *Rener);

/* One LS-RK stage of the reference's default viscous path (flag_lps_stab):
 * updateF (time_integ_ls.f90:489-691) with full_proj_ijk and
 * full_diff_stab_ijk in place of full_diffusion_ijk, the lumped solve and the
 * stage combine y <- lam1*y1 + lams*y + gam*dt*R, as sodhip_rk_stage_f64 does
 * for the plain path (FP64).  u, e_int, pr, aux_h, Tem receive the EOS of the
 * current state (working nodes), mu_fluid Sutherland's viscosity (all nodes),
 * the Proj arrays (all or none; NULL keeps the projections inside the call,
 * where the stabilised operator reads each node's 15 values as one 128-byte
 * record) the projections; tau_stab is comp_tau's.  mu_sgs is zero.
 * The convection and the stabilised viscous terms are summed per element
 * before the node-ordered gather.  Rmass/Rmom/Rener (all or none) receive
 * R*invMl; rho_next/q_next/E_next (all or none) the next step's first stage
 * as in sodhip_rk_stage_next_f64. */
int sodhip_rk_stage_lps_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                            const int *connec, const int *lpoin_w, const double *He,
                            const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                            const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                            const double *mu_e, const double *invMl, const double *tau_stab,
                            const sodhip_stage_params *prm, double c_lps, double dt, double lam1,
                            double lams, double gam, const double *rho1, const double *q1,
                            const double *E1, double *rho, double *q, double *E, double *u,
                            double *e_int, double *pr, double *aux_h, double *Tem,
                            double *mu_fluid, double *ProjMass, double *ProjEner, double *ProjMX,
                            double *ProjMY, double *ProjMZ, double *Rmass, double *Rmom,
                            double *Rener, double gam_next, double *rho_next, double *q_next,
                            double *E_next);

/* ------------------------------------------------------------------------ */
/* Time step and diagnostics (results returned to the host; synchronising)   */
/* ------------------------------------------------------------------------ */

/* adapt_dt_cfl (mod_time_ops.f90:11-63); mu_sgs/mu_e are (nelem,nnode);
 * mu_fluid == NULL disables the diffusive limit, mu_sgs may be NULL (zero). */
int sodhip_adapt_dt_cfl_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
                            const double *helem, const double *u, const double *csound,
                            double cfl_conv, double cfl_diff, const double *mu_fluid,
                            const double *mu_sgs, const double *mu_e, const double *rho,
                            double *dt);
/* volAvg_EK (mod_analysis.f90:10-44) */
int sodhip_volavg_ek_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
                         const double *gpvol, double rho0, const double *rho, const double *u,
                         double *EK);
/* visc_dissipationRate (mod_analysis.f90:46-144) */
int sodhip_visc_dissipation_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                const int *connec, double rho0, const double *mu_fluid,
                                const double *mu_e, const double *u, double volT,
                                const double *gpvol, const double *He, const double *dlxigp_ip,
                                const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
                                const int *gmshAtoK, double *eps_S, double *eps_D, double *eps_T);
/* maxMach (mod_analysis.f90:146-168) */
int sodhip_max_mach_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                        const double *machno, double *maxmach);

#ifdef __cplusplus
}
#endif
#endif

// ops_stage.hip -- the fused Runge-Kutta stage (updateF + lumped solve + stage
// combine), the CFL time step and the Taylor-Green diagnostics.
//
// Fused stage = two launches:
//   1. k_rhs<CONV,DIFF,CONS> (ops_element.hip): per element, the primitives of
//      updateF's EOS block and Sutherland's law are formed from the conserved
//      FP64 state in registers, then the split-form convection and the viscous
//      operator are evaluated together (they share grad(rho) and grad(u)) and
//      written as per element-node partials;
//   2. k_gather_update: per working node, the partials are gathered in the
//      reference's serial (element, node) order, scaled by fact and by the
//      lumped inverse, and the Shu-Osher combine y = lam1*y1 + lams*y + gam*dt*R
//      is applied -- the RHS never round-trips through HBM as an assembled array
//      unless the caller asks for it (the reference keeps it for the next step's
//      first stage).
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "element.hpp"

namespace sodhip {

template <typename T>
int rhs_fused_elements(sodhip_ctx ctx, int porder, const MeshCache *m, const T *He,
                       const T *gpvol, const T *dl, const int *inv, const int *gI, const int *gJ,
                       const int *gK, const double *rho, const double *q, const double *E,
                       const T *mu_e, double gamma, double Rgas, double Cp, double Pr,
                       double mu_factor, const double *mu_factor_e, double c_rho, double c_ener,
                       double p_ref, T *out);
int fused_record_width(int porder, int op_bytes);

// RD1 / RDY: the base state y1 / the current state y enters the combine (its
// coefficient lam1 / lams is non-zero); otherwise the term is 0 * 0, the same
// value as 0 * (finite y), without reading the field.
// NX: the new y also yields the first stage of the next step of the
// reference sequencing, yn = nl1 * y + 0 * 0 + ngdt * R (the stage update with
// lam1 = 1, lams = 0 and the kept residual, time_integ_ls.f90:323-331), which
// is stored to rhoN/qN/EN.
// RW: values per record, 5 or 6 (five residuals and a pad; the record then
// starts on a 2*sizeof(TS) boundary and is read with three 2-wide loads; an
// FP64 record of 5 is read with three aligned loads as well).
template <typename TS> struct Pair2;
template <> struct Pair2<double> {
    using V = double2;
};
template <> struct Pair2<float> {
    using V = float2;
};
template <typename TS, int RW>
__device__ __forceinline__ void load_record(const TS *__restrict__ scr, int sj, double v[5])
{
    if constexpr (RW == 6) {
        using V = typename Pair2<TS>::V;
        const V *r = (const V *)(scr + 6 * (size_t)sj);
        const V a = r[0], b = r[1], c = r[2];
        v[0] = (double)a.x;
        v[1] = (double)a.y;
        v[2] = (double)b.x;
        v[3] = (double)b.y;
        v[4] = (double)c.x;
    } else if constexpr (sizeof(TS) == 8) {
        load_rec5((const double *)scr, sj, v);
    } else {
#pragma unroll
        for (int f = 0; f < 5; ++f) v[f] = (double)scr[5 * (size_t)sj + f];
    }
}

template <typename TS, bool RD1, bool RDY, bool NX, int RW>
__global__ __launch_bounds__(256) void k_gather_update(
    int npoin, int npoin_w, const int *__restrict__ lpoin_w, const int *__restrict__ wrow,
    const int *__restrict__ wsrc, const TS *__restrict__ scr, size_t EN,
    const double *__restrict__ invMl, double fact, double lam1, double lams, double gdt,
    const double *rho1, const double *q1, const double *E1, double *__restrict__ rho,
    double *__restrict__ q, double *__restrict__ E, double *__restrict__ Rmass,
    double *__restrict__ Rmom, double *__restrict__ Rener, NextStage nx)
{
    // y1 may share memory with the next-stage output (each node's y1 is read
    // before its next-stage value is written), so y1 carries no __restrict__
    // XCD-aware: the blocks of one XCD (blockIdx % 8) take one contiguous
    // eighth of the nodes, whose partials share that XCD's L2 lines
    const int per = gridDim.x >> 3; // grid is a multiple of 8
    const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const int j0 = wrow[i], j1 = wrow[i + 1];
    const int *__restrict__ src = wsrc;
    // the node's state is loaded before the slots, so that its latency overlaps
    // the slot-index -> partial chain
    const size_t n1 = n + (size_t)npoin, n2 = n + 2 * (size_t)npoin;
    const double m = invMl[n];
    const double b0 = RD1 ? rho1[n] : 0.0, b1 = RD1 ? q1[n] : 0.0, b2 = RD1 ? q1[n1] : 0.0,
                 b3 = RD1 ? q1[n2] : 0.0, b4 = RD1 ? E1[n] : 0.0;
    const double y0 = RDY ? rho[n] : 0.0, y1 = RDY ? q[n] : 0.0, y2 = RDY ? q[n1] : 0.0,
                 y3 = RDY ? q[n2] : 0.0, y4 = RDY ? E[n] : 0.0;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // four slots per step, all loads issued before the (in-order) additions
    for (int j = j0; j < j1; j += 4) {
        int sj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) sj[u] = j + u < j1 ? src[j + u] : -1;
        double v[4][5];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            // an absent slot reads record 0 (always present) and is not added
            load_record<TS, RW>(scr, max(sj[u], 0), v[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (sj[u] >= 0) {
#pragma unroll
                for (int f = 0; f < 5; ++f) acc[f] += fact * v[u][f];
            }
    }
    const double rm = acc[0] * m, r0 = acc[1] * m, r1 = acc[2] * m, r2 = acc[3] * m,
                 re = acc[4] * m;
    const double z0 = lam1 * b0 + lams * y0 + gdt * rm;
    const double z1 = lam1 * b1 + lams * y1 + gdt * r0;
    const double z2 = lam1 * b2 + lams * y2 + gdt * r1;
    const double z3 = lam1 * b3 + lams * y3 + gdt * r2;
    const double z4 = lam1 * b4 + lams * y4 + gdt * re;
    rho[n] = z0;
    q[n] = z1;
    q[n1] = z2;
    q[n2] = z3;
    E[n] = z4;
    if constexpr (NX) {
        const double zero = 0.0, nls = 0.0;
        nx.rho[n] = nx.nl1 * z0 + nls * zero + nx.ngdt * rm;
        nx.q[n] = nx.nl1 * z1 + nls * zero + nx.ngdt * r0;
        nx.q[n1] = nx.nl1 * z2 + nls * zero + nx.ngdt * r1;
        nx.q[n2] = nx.nl1 * z3 + nls * zero + nx.ngdt * r2;
        nx.E[n] = nx.nl1 * z4 + nls * zero + nx.ngdt * re;
    }
    if (Rmass) {
        Rmass[n] = rm;
        Rmom[n] = r0;
        Rmom[n1] = r1;
        Rmom[n2] = r2;
        Rener[n] = re;
    }
}

// Classical RK4 stage of rk_4_main (time_integ.f90, compiled out in the
// reference by #if 0): per working node the partials are gathered as in
// k_gather_update, scaled by fact and the lumped inverse (R, lumped_solver_*,
// :431-435), accumulated Rsum <- Rsum + b_i R (:442-451; FIRST: Rsum was
// zeroed at the start of the step, :234-240), and then either the next
// stage's state aux = y - dt a_{i+1} R (:266-275) or, at the LAST stage, the
// step update y <- y - dt Rsum (:460-468).
template <typename TS, int RW, bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void k_gather_rk4(
    int npoin, int npoin_w, const int *__restrict__ lpoin_w, const int *__restrict__ wrow,
    const int *__restrict__ wsrc, const TS *__restrict__ scr, const double *__restrict__ invMl,
    double fact, double dta, double bi, double dt, double *__restrict__ rho,
    double *__restrict__ q, double *__restrict__ E, double *__restrict__ arho,
    double *__restrict__ aq, double *__restrict__ aE, double *__restrict__ Smass,
    double *__restrict__ Smom, double *__restrict__ Sener, double *__restrict__ Rmass,
    double *__restrict__ Rmom, double *__restrict__ Rener)
{
    const int per = gridDim.x >> 3; // grid is a multiple of 8
    const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const int j0 = wrow[i], j1 = wrow[i + 1];
    const size_t k[5] = {(size_t)n, (size_t)n, n + (size_t)npoin, n + 2 * (size_t)npoin,
                         (size_t)n};
    double *yv[5] = {rho, q, q, q, E};
    double *av[5] = {arho, aq, aq, aq, aE};
    double *sv[5] = {Smass, Smom, Smom, Smom, Sener};
    double *rv[5] = {Rmass, Rmom, Rmom, Rmom, Rener};
    const double m = invMl[n];
    double y[5], sold[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        y[f] = yv[f][k[f]];
        sold[f] = FIRST ? 0.0 : sv[f][k[f]];
    }
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = j0; j < j1; j += 4) {
        int sj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) sj[u] = j + u < j1 ? wsrc[j + u] : -1;
        double v[4][5];
#pragma unroll
        for (int u = 0; u < 4; ++u) load_record<TS, RW>(scr, max(sj[u], 0), v[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (sj[u] >= 0) {
#pragma unroll
                for (int f = 0; f < 5; ++f) acc[f] += fact * v[u][f];
            }
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        const double r = acc[f] * m;
        const double sn = sold[f] + bi * r;
        sv[f][k[f]] = sn;
        if (Rmass) rv[f][k[f]] = r;
        if constexpr (LAST)
            yv[f][k[f]] = y[f] - dt * sn;
        else
            av[f][k[f]] = y[f] - dta * r;
    }
}

// the gather/update of one stage from the records scr (rw values per record,
// type TS) of the working nodes list[0..nw) (1-based), node i owning the
// records src[row[i]..row[i+1])
template <typename TS>
int gather_update_on(sodhip_ctx ctx, int npoin, int nw, const int *list, const int *row,
                     const int *src, const TS *scr, int rw, const double *invMl, double fact,
                     double dt, double lam1, double lams, double gam, const double *rho1,
                     const double *q1, const double *E1, double *rho, double *q, double *E,
                     double *Rmass, double *Rmom, double *Rener, const NextStage *next)
{
    if (rw != 5 && rw != 6) return set_error(SODHIP_ERR_ARG, "record width %d", rw);
    if (nw <= 0) return SODHIP_OK;
    int tok = prof_begin(ctx, 1);
    const int gt = ctx->gather_threads;
    const int grid = ((nw + gt - 1) / gt + 7) / 8 * 8;
    const NextStage nx = next ? *next : NextStage{0.0, 0.0, nullptr, nullptr, nullptr};
#define SODHIP_GU3(R1, RY, NX, RWV)                                                            \
    k_gather_update<TS, R1, RY, NX, RWV><<<grid, gt, 0, ctx->stream>>>(                        \
        npoin, nw, list, row, src, scr, 0, invMl, fact, lam1, lams, gam * dt,                   \
        rho1, q1, E1, rho, q, E, Rmass, Rmom, Rener, nx)
#define SODHIP_GU2(R1, RY, NX)                                                                 \
    if (rw == 6)                                                                               \
        SODHIP_GU3(R1, RY, NX, 6);                                                             \
    else                                                                                       \
        SODHIP_GU3(R1, RY, NX, 5)
#define SODHIP_GU(R1, RY)                                                                      \
    if (next)                                                                                  \
        SODHIP_GU2(R1, RY, true);                                                              \
    else                                                                                       \
        SODHIP_GU2(R1, RY, false)
    if (lam1 != 0.0 && lams != 0.0)
        SODHIP_GU(true, true);
    else if (lam1 != 0.0)
        SODHIP_GU(true, false);
    else if (lams != 0.0)
        SODHIP_GU(false, true);
    else
        SODHIP_GU(false, false);
#undef SODHIP_GU
#undef SODHIP_GU2
#undef SODHIP_GU3
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}
template int gather_update_on<double>(sodhip_ctx, int, int, const int *, const int *, const int *,
                                      const double *, int, const double *, double, double,
                                      double, double, double, const double *, const double *,
                                      const double *, double *, double *, double *, double *,
                                      double *, double *, const NextStage *);

// the gather/update of one stage from the element records scr (5 values per
// element-node, type TS) over the working incidence of m
template <typename TS>
int gather_update(sodhip_ctx ctx, const MeshCache *m, int npoin, int npoin_w,
                  const int *lpoin_w, const TS *scr, int rw, size_t EN, const double *invMl,
                  double fact, double dt, double lam1, double lams, double gam,
                  const double *rho1, const double *q1, const double *E1, double *rho, double *q,
                  double *E, double *Rmass, double *Rmom, double *Rener, const NextStage *next)
{
    (void)EN;
    return gather_update_on<TS>(ctx, npoin, npoin_w, lpoin_w, m->wrow, m->wsrc, scr, rw, invMl,
                                fact, dt, lam1, lams, gam, rho1, q1, E1, rho, q, E, Rmass, Rmom,
                                Rener, next);
}
template int gather_update<double>(sodhip_ctx, const MeshCache *, int, int, const int *,
                                   const double *, int, size_t, const double *, double, double,
                                   double, double, double, const double *, const double *,
                                   const double *, double *, double *, double *, double *,
                                   double *, double *, const NextStage *);

// One fused stage on the mesh tables m (with its working incidence built) and
// the record scratch scr of scr_bytes
template <typename T>
static int rk_stage_on(sodhip_ctx ctx, int porder, const MeshCache *m, int npoin, int npoin_w,
                       const int *lpoin_w, const T *He, const T *gpvol, const T *dl,
                       const int *inv, const int *gI, const int *gJ, const int *gK, const T *mu_e,
                       const double *invMl, const sodhip_stage_params *prm, double dt,
                       double lam1, double lams, double gam, const double *rho1, const double *q1,
                       const double *E1, double *rho, double *q, double *E, double *Rmass,
                       double *Rmom, double *Rener, const NextStage *next, void *scr_mem,
                       size_t scr_bytes, const ClusterTables *cl)
{
    const size_t EN = (size_t)m->nelem * nnode_of(porder);
    if constexpr (sizeof(T) == 8) {
        if (cl && porder == 3) {
            // brick-local assembly: interior nodes finished in the element
            // kernel, surface nodes by the gather over their partial records
            if ((size_t)cl->nloc * 5 * sizeof(double) > scr_bytes)
                return set_error(SODHIP_ERR_ARG,
                                 "the stage needs %zu bytes of record scratch, %zu available",
                                 (size_t)cl->nloc * 5 * sizeof(double), scr_bytes);
            double *rec = (double *)scr_mem;
            int rc = rhs_cluster_stage(ctx, m, cl, He, gpvol, dl, inv, gI, gJ, gK, mu_e,
                                       prm->gamma_gas, prm->Rgas, prm->Cp, prm->Pr,
                                       prm->mu_factor, prm->mu_factor_e, prm->c_rho,
                                       prm->c_ener, prm->p_ref, invMl,
                                       prm->fact, dt, lam1, lams, gam, rho1, q1, E1, rho, q, E,
                                       Rmass, Rmom, Rener, next, rec);
            if (rc) return rc;
            return gather_update_on<double>(ctx, npoin, cl->ns, cl->snode, cl->srow, cl->ssrc,
                                            rec, 5, invMl, prm->fact, dt, lam1, lams, gam, rho1,
                                            q1, E1, rho, q, E, Rmass, Rmom, Rener, next);
        }
    }
    (void)npoin_w;
    const int rw = fused_record_width(porder, (int)sizeof(T));
    if (rw * EN * sizeof(T) > scr_bytes)
        return set_error(SODHIP_ERR_ARG, "the stage needs %zu bytes of record scratch, %zu available",
                         rw * EN * sizeof(T), scr_bytes);
    T *scr = (T *)scr_mem;
    int rc = rhs_fused_elements<T>(ctx, porder, m, He, gpvol, dl, inv, gI, gJ, gK, rho, q, E, mu_e,
                                   prm->gamma_gas, prm->Rgas, prm->Cp, prm->Pr, prm->mu_factor, prm->mu_factor_e,
                                   prm->c_rho, prm->c_ener, prm->p_ref, scr);
    if (rc) return rc;
    SODHIP_TRY_HIP(hipGetLastError());
    return gather_update<T>(ctx, m, npoin, npoin_w, lpoin_w, scr, rw, EN, invMl, prm->fact, dt,
                            lam1, lams, gam, rho1, q1, E1, rho, q, E, Rmass, Rmom, Rener, next);
}

template <typename T>
static int rk_stage(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                    const int *connec, const int *lpoin_w, const T *He, const T *gpvol,
                    const T *dl, const int *inv, const int *gI, const int *gJ, const int *gK,
                    const T *mu_e, const double *invMl, const sodhip_stage_params *prm, double dt,
                    double lam1, double lams, double gam, const double *rho1, const double *q1,
                    const double *E1, double *rho, double *q, double *E, double *Rmass,
                    double *Rmom, double *Rener, const NextStage *next = nullptr)
{
    MeshCache *m = nullptr;
    int rc = get_mesh(ctx, porder, nelem, npoin, connec, &m);
    if (rc) return rc;
    if ((rc = get_working(ctx, m, npoin_w, lpoin_w))) return rc;
    const size_t EN = (size_t)nelem * nnode_of(porder);
    const int rw = fused_record_width(porder, (int)sizeof(T));
    const ClusterTables *cl = nullptr;
    if (sizeof(T) == 8 && porder == 3 && ctx->cluster_stage) {
        // a mesh without clusters (an element listing a node twice) keeps
        // the node-ordered gather
        if (get_clusters(ctx, m, &cl) != SODHIP_OK) cl = nullptr;
    }
    const size_t need = cl ? (size_t)cl->nloc * 5 * sizeof(double) : rw * EN * sizeof(T);
    if ((rc = ensure_buf(ctx->scratch, need))) return rc;
    return rk_stage_on<T>(ctx, porder, m, npoin, npoin_w, lpoin_w, He, gpvol, dl, inv, gI, gJ, gK,
                          mu_e, invMl, prm, dt, lam1, lams, gam, rho1, q1, E1, rho, q, E, Rmass,
                          Rmom, Rener, next, ctx->scratch.ptr, ctx->scratch.bytes, cl);
}

// One classical RK4 stage (see k_gather_rk4): the fused RHS at the stage state
// (srho, sq, sE), then the gather with the RK4 node update
template <typename T>
static int rk4_stage(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                     const int *connec, const int *lpoin_w, const T *He, const T *gpvol,
                     const T *dl, const int *inv, const int *gI, const int *gJ, const int *gK,
                     const T *mu_e, const double *invMl, const sodhip_stage_params *prm, double dt,
                     double a_next, double b, int first, int last, double *rho, double *q,
                     double *E, const double *srho, const double *sq, const double *sE,
                     double *arho, double *aq, double *aE, double *Smass, double *Smom,
                     double *Sener, double *Rmass, double *Rmom, double *Rener)
{
    MeshCache *m = nullptr;
    int rc = get_mesh(ctx, porder, nelem, npoin, connec, &m);
    if (rc) return rc;
    if ((rc = get_working(ctx, m, npoin_w, lpoin_w))) return rc;
    const size_t EN = (size_t)nelem * nnode_of(porder);
    const int rw = fused_record_width(porder, (int)sizeof(T));
    if ((rc = ensure_buf(ctx->scratch, rw * EN * sizeof(T)))) return rc;
    T *scr = (T *)ctx->scratch.ptr;
    rc = rhs_fused_elements<T>(ctx, porder, m, He, gpvol, dl, inv, gI, gJ, gK, srho, sq, sE, mu_e,
                               prm->gamma_gas, prm->Rgas, prm->Cp, prm->Pr, prm->mu_factor, prm->mu_factor_e,
                               prm->c_rho, prm->c_ener, prm->p_ref, scr);
    if (rc) return rc;
    int tok = prof_begin(ctx, 1);
    const int grid = ((npoin_w + 255) / 256 + 7) / 8 * 8;
#define SODHIP_RK4G(RWV, FI, LA)                                                               \
    k_gather_rk4<T, RWV, FI, LA><<<grid, 256, 0, ctx->stream>>>(                               \
        npoin, npoin_w, lpoin_w, m->wrow, m->wsrc, scr, invMl, prm->fact, dt * a_next, b, dt,   \
        rho, q, E, arho, aq, aE, Smass, Smom, Sener, Rmass, Rmom, Rener)
#define SODHIP_RK4F(RWV)                                                                       \
    if (first && last)                                                                         \
        SODHIP_RK4G(RWV, true, true);                                                          \
    else if (first)                                                                            \
        SODHIP_RK4G(RWV, true, false);                                                         \
    else if (last)                                                                             \
        SODHIP_RK4G(RWV, false, true);                                                         \
    else                                                                                       \
        SODHIP_RK4G(RWV, false, false)
    if (rw == 6)
        SODHIP_RK4F(6);
    else
        SODHIP_RK4F(5);
#undef SODHIP_RK4F
#undef SODHIP_RK4G
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

// ---------------------------------------------------------------- allocation-free stage
// The allocation-free stage keeps its residual records in this file-scope
// device array: one copy per GPU, created when the library's code object is
// loaded on that GPU and held until the process ends.  In host code the name
// `scratch` is a host-side stand-in, not a device address; noalloc_init, run
// by sodhip_create with the context's device current, obtains each GPU's
// address of its copy once, and only that address is ever passed to kernels.
// Every call on one GPU shares the array and it keeps the last call's records
// (nothing in it has to start at zero: every record is written before it is
// read); calls that overlap on one GPU race on it.
constexpr size_t kNoAllocScratchBytes = (size_t)SODHIP_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_noalloc_mu;
static std::map<int, void *> g_noalloc_ptr; // device -> address of its copy of scratch

int noalloc_init(int device)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr.count(device)) return SODHIP_OK;
    void *ptr = nullptr;
    SODHIP_TRY_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    SODHIP_REQUIRE(ptr, "no device address for the allocation-free scratch on device %d", device);
    g_noalloc_ptr[device] = ptr;
    return SODHIP_OK;
}

static void *noalloc_ptr(int device)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    auto it = g_noalloc_ptr.find(device);
    return it == g_noalloc_ptr.end() ? nullptr : it->second;
}

// ---------------------------------------------------------------- CFL time step
// adapt_dt_cfl (mod_time_ops.f90:11-63): one thread per element over the
// reference's element-fastest connec and mu_e (for a node position the lanes
// of a wave read consecutive entries), the nodes visited an i-line at a time
// (build_alex) with all loads of a line issued before its arithmetic (fmax is
// exact, so the order does not change the result); then the min over blocks.
__global__ __launch_bounds__(256) void k_dt_cfl(int nelem, int nnode, int porder,
                                                const int *__restrict__ connec,
                                                const int *__restrict__ perm,
                                                const double *__restrict__ helem,
                                                const double *__restrict__ u,
                                                const double *__restrict__ csound, int npoin,
                                                double cfl_conv, double cfl_diff,
                                                const double *__restrict__ mu_fluid,
                                                const double *__restrict__ mu_sgs,
                                                const double *__restrict__ mu_e,
                                                const double *__restrict__ rho,
                                                double *__restrict__ part)
{
    __shared__ double sh[256];
    __shared__ int alex[125];
    build_alex(perm, nnode, alex);
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    const int KA = porder + 1; // one i-line per group
    double dtl = 1.0e14;
    if (e < nelem) {
        double L3 = 0.0, maxmu = 0.0;
        for (int l0 = 0; l0 < nnode; l0 += KA) {
            int n[5];
            size_t ea[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                const int a = alex[min(l0 + min(j, KA - 1), nnode - 1)];
                ea[j] = e + (size_t)nelem * a;
                n[j] = connec[ea[j]] - 1;
            }
            double v0[5], v1[5], v2[5], cs[5], mf[5], me[5], sg[5], rh[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                v0[j] = u[n[j]];
                v1[j] = u[n[j] + npoin];
                v2[j] = u[n[j] + 2 * (size_t)npoin];
                cs[j] = csound[n[j]];
                if (mu_fluid) {
                    mf[j] = mu_fluid[n[j]];
                    me[j] = mu_e[ea[j]];
                    sg[j] = mu_sgs ? mu_sgs[ea[j]] : 0.0;
                    rh[j] = rho[n[j]];
                }
            }
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                double um = fabs(v0[j]);
                um = fmax(um, fabs(v1[j]));
                um = fmax(um, fabs(v2[j]));
                L3 = fmax(L3, um + cs[j]);
                if (mu_fluid) maxmu = fmax(maxmu, (sg[j] + mf[j] + me[j]) / rh[j]);
            }
        }
        const double hh = helem[e] / (double)(2 * porder + 1);
        dtl = cfl_conv * hh / L3;
        if (mu_fluid) dtl = fmin(dtl, cfl_diff * (hh * hh) / maxmu);
    }
    sh[threadIdx.x] = dtl;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ---------------------------------------------------------------- diagnostics
// volAvg_EK (mod_analysis.f90:10-44): per element sum_g gpvol 1/2 rho |u|^2 in
// FP64 (the shape functions are the identity at the GLL points), one lane per
// node (gpvol and connecT rows read whole), the element's sum by a fixed tree
// into elK; then fixed-order block sums.
template <int P>
__global__ __launch_bounds__(((P + 1) * (P + 1) * (P + 1) + 63) / 64 * 64) void k_ek(
    int nelem, int npoin, const int *__restrict__ connecT, const double *__restrict__ gpvol,
    const double *__restrict__ rho, const double *__restrict__ u, double *__restrict__ elK)
{
    constexpr int NN = (P + 1) * (P + 1) * (P + 1), NT = (NN + 63) / 64 * 64;
    __shared__ double rs[NT];
    const int lane = threadIdx.x;
    const bool act = lane < NN;
    const XcdElems xe = xcd_elems(nelem);
    for (int e = xe.first; e < xe.end; e += xe.stride) {
        double v = 0.0;
        if (act) {
            const size_t ea = (size_t)e * NN + lane;
            const int n = connecT[ea];
            const double w = gpvol[ea], r = rho[n];
            const double u0 = u[n], u1 = u[n + npoin], u2 = u[n + 2 * (size_t)npoin];
            v = w * 0.5 * r * (u0 * u0 + u1 * u1 + u2 * u2);
        }
        rs[lane] = v;
        __syncthreads();
        for (int s = NT / 2; s > 0; s >>= 1) {
            if (lane < s) rs[lane] += rs[lane + s];
            __syncthreads();
        }
        if (lane == 0) elK[e] = rs[0];
        __syncthreads();
    }
}

// visc_dissipationRate (mod_analysis.f90:46-144): per quadrature point the
// velocity gradient (line derivatives, as the operators), enstrophy
// (mu_f + mu_e)|curl u|^2 and dilatation mu_f (div u)^2, weighted by gpvol;
// element sums by a fixed tree, written per element.
template <int P>
__global__ __launch_bounds__(((P + 1) * (P + 1) * (P + 1) + 63) / 64 * 64) void k_dissipation(
    int nelem, int npoin, const int *__restrict__ connecT, const double *__restrict__ He,
    const double *__restrict__ gpvol, const double *__restrict__ dl, const int *__restrict__ inv,
    const int *__restrict__ gI, const int *__restrict__ gJ, const int *__restrict__ gK,
    const double *__restrict__ u, const double *__restrict__ mu_fluid,
    const double *__restrict__ mu_e, double *__restrict__ elS, double *__restrict__ elD)
{
    constexpr int N1 = P + 1, NN = N1 * N1 * N1, NT = (NN + 63) / 64 * 64;
    __shared__ double su[3 * NT];
    __shared__ double rs[NT], rd[NT];
    const int lane = threadIdx.x;
    const bool act = lane < NN;
    const int a = act ? lane : 0;
    const int own[3] = {min(max(gI[a] - 1, 0), P), min(max(gJ[a] - 1, 0), P),
                        min(max(gK[a] - 1, 0), P)};
    const XcdElems xe = xcd_elems(nelem);
    for (int e = xe.first; e < xe.end; e += xe.stride) {
        const size_t ea = (size_t)e * NN + a;
        const int n = act ? connecT[ea] : 0;
        if (act)
            for (int d = 0; d < 3; ++d) su[d * NT + a] = u[n + (size_t)npoin * d];
        __syncthreads();
        double du[3][3]; // du[i][k] = d u_i / d xi_k
        for (int k = 0; k < 3; ++k) {
            for (int i = 0; i < 3; ++i) du[i][k] = 0.0;
            for (int m = 0; m < N1; ++m) {
                int idx[3] = {own[0], own[1], own[2]};
                idx[k] = m;
                const int b = min(max(inv[idx[0] + N1 * (idx[1] + N1 * idx[2])] - 1, 0), NN - 1);
                const double c = dl[a + NN * (k + 3 * m)];
                for (int i = 0; i < 3; ++i) du[i][k] += c * su[i * NT + b];
            }
        }
        double g[3][3]; // g[j][i] = d u_i / d x_j
        const double *hp = He + 9 * ea;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i)
                g[j][i] = act ? hp[j + 0] * du[i][0] + hp[j + 3] * du[i][1] + hp[j + 6] * du[i][2] : 0.0;
        const double div = g[0][0] + g[1][1] + g[2][2];
        const double w0 = g[1][2] - g[2][1], w1 = g[2][0] - g[0][2], w2 = g[0][1] - g[1][0];
        const double curl2 = w0 * w0 + w1 * w1 + w2 * w2;
        const double vol = act ? gpvol[ea] : 0.0;
        const double mf = act ? mu_fluid[n] : 0.0;
        const double me = act ? mu_e[e + (size_t)nelem * a] : 0.0;
        rs[lane] = vol * (mf + me) * curl2;
        rd[lane] = vol * mf * div * div;
        __syncthreads();
        for (int s = NT / 2; s > 0; s >>= 1) {
            if (lane < s) {
                rs[lane] += rs[lane + s];
                rd[lane] += rd[lane + s];
            }
            __syncthreads();
        }
        if (lane == 0) {
            elS[e] = rs[0];
            elD[e] = rd[0];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_sum_blocks(int n, const double *__restrict__ v,
                                                    double *__restrict__ part)
{
    __shared__ double sh[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    sh[threadIdx.x] = i < n ? v[i] : 0.0;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_max_mach(int npoin_w, const int *__restrict__ lpoin_w,
                                                  const double *__restrict__ machno,
                                                  double *__restrict__ part)
{
    __shared__ double sh[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    sh[threadIdx.x] = i < npoin_w ? machno[lpoin_w[i] - 1] : 0.0;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

static int fetch_parts(sodhip_ctx ctx, const double *dev, int n, std::vector<double> &h)
{
    h.resize(n);
    SODHIP_TRY_HIP(hipMemcpyAsync(h.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    return SODHIP_OK;
}

} // namespace sodhip

using namespace sodhip;

extern "C" int sodhip_rk_stage_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                                   const int *connec, const int *lpoin_w, int op_precision,
                                   const void *He, const void *gpvol, const void *dlxigp_ip,
                                   const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
                                   const int *gmshAtoK, const void *mu_e, const double *invMl,
                                   const sodhip_stage_params *prm, double dt, double lam1,
                                   double lams, double gam, const double *rho1, const double *q1,
                                   const double *E1, double *rho, double *q, double *E,
                                   double *Rmass, double *Rmom, double *Rener)
{
    SODHIP_REQUIRE(ctx && prm, "null argument");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_REQUIRE(nelem > 0 && npoin > 0 && npoin_w > 0 && npoin_w <= npoin,
                   "invalid mesh sizes");
    SODHIP_REQUIRE(connec && lpoin_w && He && gpvol && dlxigp_ip && invAtoIJK && gmshAtoI &&
                       gmshAtoJ && gmshAtoK && mu_e && invMl && rho1 && q1 && E1 && rho && q && E,
                   "null array argument");
    SODHIP_REQUIRE(!!Rmass == !!Rmom && !!Rmom == !!Rener, "Rmass/Rmom/Rener: all or none");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    if (op_precision == 8)
        return rk_stage<double>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w,
                                (const double *)He, (const double *)gpvol,
                                (const double *)dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK,
                                (const double *)mu_e, invMl, prm, dt, lam1, lams, gam, rho1, q1,
                                E1, rho, q, E, Rmass, Rmom, Rener);
    if (op_precision == 4)
        return rk_stage<float>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w,
                               (const float *)He, (const float *)gpvol, (const float *)dlxigp_ip,
                               invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, (const float *)mu_e, invMl,
                               prm, dt, lam1, lams, gam, rho1, q1, E1, rho, q, E, Rmass, Rmom,
                               Rener);
    return set_error(SODHIP_ERR_ARG, "op_precision must be 8 or 4 (got %d)", op_precision);
}

extern "C" int sodhip_rk_stage_next_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w, const int *connec,
    const int *lpoin_w, int op_precision, const void *He, const void *gpvol,
    const void *dlxigp_ip, const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
    const int *gmshAtoK, const void *mu_e, const double *invMl, const sodhip_stage_params *prm,
    double dt, double lam1, double lams, double gam, const double *rho1, const double *q1,
    const double *E1, double *rho, double *q, double *E, double *Rmass, double *Rmom,
    double *Rener, double gam_next, double *rho_next, double *q_next, double *E_next)
{
    SODHIP_REQUIRE(ctx && prm, "null argument");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_REQUIRE(nelem > 0 && npoin > 0 && npoin_w > 0 && npoin_w <= npoin,
                   "invalid mesh sizes");
    SODHIP_REQUIRE(connec && lpoin_w && He && gpvol && dlxigp_ip && invAtoIJK && gmshAtoI &&
                       gmshAtoJ && gmshAtoK && mu_e && invMl && rho1 && q1 && E1 && rho && q && E,
                   "null array argument");
    SODHIP_REQUIRE(!!Rmass == !!Rmom && !!Rmom == !!Rener, "Rmass/Rmom/Rener: all or none");
    SODHIP_REQUIRE(rho_next && q_next && E_next, "null next-stage array");
    SODHIP_REQUIRE(rho_next != rho && q_next != q && E_next != E,
                   "the next stage must not overwrite the new state");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    const NextStage nx{1.0, gam_next * dt, rho_next, q_next, E_next};
    if (op_precision == 8)
        return rk_stage<double>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w,
                                (const double *)He, (const double *)gpvol,
                                (const double *)dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK,
                                (const double *)mu_e, invMl, prm, dt, lam1, lams, gam, rho1, q1,
                                E1, rho, q, E, Rmass, Rmom, Rener, &nx);
    if (op_precision == 4)
        return rk_stage<float>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w,
                               (const float *)He, (const float *)gpvol, (const float *)dlxigp_ip,
                               invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, (const float *)mu_e, invMl,
                               prm, dt, lam1, lams, gam, rho1, q1, E1, rho, q, E, Rmass, Rmom,
                               Rener, &nx);
    return set_error(SODHIP_ERR_ARG, "op_precision must be 8 or 4 (got %d)", op_precision);
}

extern "C" int sodhip_rk_stage_noalloc_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w, const int *connec,
    const int *lpoin_w, int op_precision, const void *He, const void *gpvol,
    const void *dlxigp_ip, const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
    const int *gmshAtoK, const void *mu_e, const double *invMl, const sodhip_stage_params *prm,
    double dt, double lam1, double lams, double gam, const double *rho1, const double *q1,
    const double *E1, double *rho, double *q, double *E, double *Rmass, double *Rmom,
    double *Rener)
{
    SODHIP_REQUIRE(ctx && prm, "null argument");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_REQUIRE(nelem > 0 && npoin > 0 && npoin_w > 0 && npoin_w <= npoin,
                   "invalid mesh sizes");
    SODHIP_REQUIRE(connec && lpoin_w && He && gpvol && dlxigp_ip && invAtoIJK && gmshAtoI &&
                       gmshAtoJ && gmshAtoK && mu_e && invMl && rho1 && q1 && E1 && rho && q && E,
                   "null array argument");
    SODHIP_REQUIRE(!!Rmass == !!Rmom && !!Rmom == !!Rener, "Rmass/Rmom/Rener: all or none");
    SODHIP_REQUIRE(op_precision == 8 || op_precision == 4,
                   "op_precision must be 8 or 4 (got %d)", op_precision);
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    void *scr = noalloc_ptr(ctx->device);
    SODHIP_REQUIRE(scr, "no allocation-free scratch on device %d", ctx->device);
    const size_t EN = (size_t)nelem * nnode_of(porder);
    const size_t need = fused_record_width(porder, op_precision) * EN * (size_t)op_precision;
    SODHIP_REQUIRE(need <= kNoAllocScratchBytes,
                   "the allocation-free stage holds at most %zu bytes of records (%d MiB); this "
                   "mesh needs %zu (nelem = %d at porder %d)",
                   kNoAllocScratchBytes, SODHIP_NOALLOC_SCRATCH_MIB, need, nelem, porder);
    const MeshCache *m = nullptr;
    int rc = find_mesh(ctx, porder, nelem, npoin, connec, npoin_w, lpoin_w, &m);
    if (rc) return rc;
    // launch timing may create events: it is off for this call
    const bool prof = ctx->prof;
    ctx->prof = false;
    if (op_precision == 8)
        rc = rk_stage_on<double>(ctx, porder, m, npoin, npoin_w, lpoin_w, (const double *)He,
                                 (const double *)gpvol, (const double *)dlxigp_ip, invAtoIJK,
                                 gmshAtoI, gmshAtoJ, gmshAtoK, (const double *)mu_e, invMl, prm,
                                 dt, lam1, lams, gam, rho1, q1, E1, rho, q, E, Rmass, Rmom, Rener,
                                 nullptr, scr, kNoAllocScratchBytes,
                                 ctx->cluster_stage && porder == 3 ? m->clusters.get() : nullptr);
    else
        rc = rk_stage_on<float>(ctx, porder, m, npoin, npoin_w, lpoin_w, (const float *)He,
                                (const float *)gpvol, (const float *)dlxigp_ip, invAtoIJK,
                                gmshAtoI, gmshAtoJ, gmshAtoK, (const float *)mu_e, invMl, prm, dt,
                                lam1, lams, gam, rho1, q1, E1, rho, q, E, Rmass, Rmom, Rener,
                                nullptr, scr, kNoAllocScratchBytes, nullptr);
    ctx->prof = prof;
    return rc;
}

extern "C" int sodhip_rk4_stage_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w, const int *connec,
    const int *lpoin_w, int op_precision, const void *He, const void *gpvol,
    const void *dlxigp_ip, const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
    const int *gmshAtoK, const void *mu_e, const double *invMl, const sodhip_stage_params *prm,
    double dt, double a_next, double b, int first, int last, double *rho, double *q, double *E,
    const double *srho, const double *sq, const double *sE, double *aux_rho, double *aux_q,
    double *aux_E, double *Rmass_sum, double *Rmom_sum, double *Rener_sum, double *Rmass,
    double *Rmom, double *Rener)
{
    SODHIP_REQUIRE(ctx && prm, "null argument");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_REQUIRE(nelem > 0 && npoin > 0 && npoin_w > 0 && npoin_w <= npoin,
                   "invalid mesh sizes");
    SODHIP_REQUIRE(connec && lpoin_w && He && gpvol && dlxigp_ip && invAtoIJK && gmshAtoI &&
                       gmshAtoJ && gmshAtoK && mu_e && invMl && rho && q && E && srho && sq &&
                       sE && Rmass_sum && Rmom_sum && Rener_sum,
                   "null array argument");
    SODHIP_REQUIRE(last || (aux_rho && aux_q && aux_E), "the next stage's state is needed");
    SODHIP_REQUIRE(last || (aux_rho != rho && aux_q != q && aux_E != E),
                   "the next stage's state must not overwrite y");
    SODHIP_REQUIRE(!!Rmass == !!Rmom && !!Rmom == !!Rener, "Rmass/Rmom/Rener: all or none");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    if (op_precision == 8)
        return rk4_stage<double>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w,
                                 (const double *)He, (const double *)gpvol,
                                 (const double *)dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ,
                                 gmshAtoK, (const double *)mu_e, invMl, prm, dt, a_next, b, first,
                                 last, rho, q, E, srho, sq, sE, aux_rho, aux_q, aux_E, Rmass_sum,
                                 Rmom_sum, Rener_sum, Rmass, Rmom, Rener);
    if (op_precision == 4)
        return rk4_stage<float>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w,
                                (const float *)He, (const float *)gpvol, (const float *)dlxigp_ip,
                                invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, (const float *)mu_e,
                                invMl, prm, dt, a_next, b, first, last, rho, q, E, srho, sq, sE,
                                aux_rho, aux_q, aux_E, Rmass_sum, Rmom_sum, Rener_sum, Rmass,
                                Rmom, Rener);
    return set_error(SODHIP_ERR_ARG, "op_precision must be 8 or 4 (got %d)", op_precision);
}

extern "C" int sodhip_adapt_dt_cfl_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                       const int *connec, const double *helem, const double *u,
                                       const double *csound, double cfl_conv, double cfl_diff,
                                       const double *mu_fluid, const double *mu_sgs,
                                       const double *mu_e, const double *rho, double *dt)
{
    SODHIP_REQUIRE(ctx && connec && helem && u && csound && dt, "null argument");
    SODHIP_REQUIRE(!mu_fluid || (mu_e && rho), "the diffusive limit needs mu_e and rho");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    int rc = get_mesh(ctx, porder, nelem, npoin, connec, &m);
    if (rc) return rc;
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    const int nb = tpe_grid(nelem);
    if ((rc = ensure_buf(ctx->diag, (size_t)nb * sizeof(double)))) return rc;
    double *part = (double *)ctx->diag.ptr;
    int tok = prof_begin(ctx, 3);
    k_dt_cfl<<<nb, 256, 0, ctx->stream>>>(nelem, nnode_of(porder), porder, connec, m->perm, helem,
                                          u, csound, npoin, cfl_conv, cfl_diff, mu_fluid, mu_sgs,
                                          mu_e, rho, part);
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    std::vector<double> h;
    if ((rc = fetch_parts(ctx, part, nb, h))) return rc;
    double v = 1.0e14;
    for (double x : h) v = std::min(v, x);
    *dt = v;
    return SODHIP_OK;
}

extern "C" int sodhip_volavg_ek_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                    const int *connec, const double *gpvol, double rho0,
                                    const double *rho, const double *u, double *EK)
{
    SODHIP_REQUIRE(ctx && connec && gpvol && rho && u && EK, "null argument");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    int rc = get_mesh(ctx, porder, nelem, npoin, connec, &m);
    if (rc) return rc;
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    const int nb = (nelem + 255) / 256;
    if ((rc = ensure_buf(ctx->diag, ((size_t)nelem + nb) * sizeof(double)))) return rc;
    double *elK = (double *)ctx->diag.ptr, *part = elK + nelem;
    const int grid = diag_grid(ctx, nelem);
    switch (porder) {
    case 2: k_ek<2><<<grid, 64, 0, ctx->stream>>>(nelem, npoin, m->connecT, gpvol, rho, u, elK); break;
    case 3: k_ek<3><<<grid, 64, 0, ctx->stream>>>(nelem, npoin, m->connecT, gpvol, rho, u, elK); break;
    case 4: k_ek<4><<<grid, 128, 0, ctx->stream>>>(nelem, npoin, m->connecT, gpvol, rho, u, elK); break;
    }
    k_sum_blocks<<<nb, 256, 0, ctx->stream>>>(nelem, elK, part);
    SODHIP_TRY_HIP(hipGetLastError());
    std::vector<double> h;
    if ((rc = fetch_parts(ctx, part, nb, h))) return rc;
    double s = 0.0;
    for (double x : h) s += x;
    // the reference normalises by (2*3.14159)^3 whatever the box (mod_analysis.f90:37)
    *EK = s / (rho0 * std::pow(2.0 * 3.14159, 3.0));
    return SODHIP_OK;
}

extern "C" int sodhip_visc_dissipation_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                           const int *connec, double rho0,
                                           const double *mu_fluid, const double *mu_e,
                                           const double *u, double volT, const double *gpvol,
                                           const double *He, const double *dlxigp_ip,
                                           const int *invAtoIJK, const int *gmshAtoI,
                                           const int *gmshAtoJ, const int *gmshAtoK,
                                           double *eps_S, double *eps_D, double *eps_T)
{
    SODHIP_REQUIRE(ctx && connec && mu_fluid && mu_e && u && gpvol && He && dlxigp_ip &&
                       invAtoIJK && gmshAtoI && gmshAtoJ && gmshAtoK && eps_S && eps_D && eps_T,
                   "null argument");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    int rc = get_mesh(ctx, porder, nelem, npoin, connec, &m);
    if (rc) return rc;
    const int nb = (nelem + 255) / 256;
    if ((rc = ensure_buf(ctx->diag, (2 * (size_t)nelem + 2 * nb) * sizeof(double)))) return rc;
    double *elS = (double *)ctx->diag.ptr, *elD = elS + nelem, *pS = elD + nelem, *pD = pS + nb;
    const int grid = diag_grid(ctx, nelem, 16); // 114 VGPRs: 4 waves per SIMD
    switch (porder) {
    case 2:
        k_dissipation<2><<<grid, 64, 0, ctx->stream>>>(nelem, npoin, m->connecT, He, gpvol,
                                                      dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ,
                                                      gmshAtoK, u, mu_fluid, mu_e, elS, elD);
        break;
    case 3:
        k_dissipation<3><<<grid, 64, 0, ctx->stream>>>(nelem, npoin, m->connecT, He, gpvol,
                                                      dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ,
                                                      gmshAtoK, u, mu_fluid, mu_e, elS, elD);
        break;
    case 4:
        k_dissipation<4><<<grid, 128, 0, ctx->stream>>>(nelem, npoin, m->connecT, He, gpvol,
                                                       dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ,
                                                       gmshAtoK, u, mu_fluid, mu_e, elS, elD);
        break;
    }
    k_sum_blocks<<<nb, 256, 0, ctx->stream>>>(nelem, elS, pS);
    k_sum_blocks<<<nb, 256, 0, ctx->stream>>>(nelem, elD, pD);
    SODHIP_TRY_HIP(hipGetLastError());
    std::vector<double> h;
    if ((rc = fetch_parts(ctx, pS, 2 * nb, h))) return rc;
    double sS = 0.0, sD = 0.0;
    for (int i = 0; i < nb; ++i) sS += h[i];
    for (int i = 0; i < nb; ++i) sD += h[nb + i];
    const double alpha = 1.0 / (rho0 * volT);
    sS *= alpha;
    sD = (double)(4.0f / 3.0f) * sD * alpha; // default-real literal (mod_analysis.f90:139)
    *eps_S = sS;
    *eps_D = sD;
    *eps_T = sS + sD;
    return SODHIP_OK;
}

extern "C" int sodhip_max_mach_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                   const double *machno, double *maxmach)
{
    SODHIP_REQUIRE(ctx && lpoin_w && machno && maxmach, "null argument");
    SODHIP_REQUIRE(npoin_w > 0 && npoin_w <= npoin, "npoin_w outside 1..npoin");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    const int nb = (npoin_w + 255) / 256;
    int rc = ensure_buf(ctx->diag, (size_t)nb * sizeof(double));
    if (rc) return rc;
    double *part = (double *)ctx->diag.ptr;
    k_max_mach<<<nb, 256, 0, ctx->stream>>>(npoin_w, lpoin_w, machno, part);
    SODHIP_TRY_HIP(hipGetLastError());
    std::vector<double> h;
    if ((rc = fetch_parts(ctx, part, nb, h))) return rc;
    double v = 0.0;
    for (double x : h) v = std::max(v, x);
    *maxmach = v;
    return SODHIP_OK;
}

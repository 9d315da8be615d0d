// ops_node.hip -- nodal kernels: lumped-mass solve and the Runge-Kutta stage
// update with the lumped inverse fused in.
#include "common.hpp"

namespace sodhip {

// lumped_solver_scal_opt / _vect_opt (mod_solver.f90:40-104): R <- R*invMl over
// working nodes, for ncomp components of an (npoin,ncomp) array.
__global__ __launch_bounds__(256) void k_lumped(int npoin, int npoin_w,
                                                const int *__restrict__ lpoin_w,
                                                const double *__restrict__ invMl,
                                                double *__restrict__ R, int ncomp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const double m = invMl[n];
    for (int c = 0; c < ncomp; ++c) R[n + (size_t)npoin * c] *= m;
}

// R <- R*invMl (stored back), y <- lam1*y1 + lams*y + gdt*R over working nodes
// for the five conserved fields (time_integ_ls.f90:323-331,350-358).
__global__ __launch_bounds__(256) void k_rk_update(int npoin, int npoin_w,
                                                   const int *__restrict__ lpoin_w,
                                                   const double *__restrict__ invMl, double lam1,
                                                   double lams, double gdt,
                                                   const double *__restrict__ rho1,
                                                   const double *__restrict__ q1,
                                                   const double *__restrict__ E1,
                                                   double *__restrict__ rho, double *__restrict__ q,
                                                   double *__restrict__ E,
                                                   double *__restrict__ Rmass,
                                                   double *__restrict__ Rmom,
                                                   double *__restrict__ Rener)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const double m = invMl[n];
    const double rm = Rmass[n] * m;
    Rmass[n] = rm;
    rho[n] = lam1 * rho1[n] + lams * rho[n] + gdt * rm;
    const double re = Rener[n] * m;
    Rener[n] = re;
    E[n] = lam1 * E1[n] + lams * E[n] + gdt * re;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const size_t k = n + (size_t)npoin * d;
        const double rq = Rmom[k] * m;
        Rmom[k] = rq;
        q[k] = lam1 * q1[k] + lams * q[k] + gdt * rq;
    }
}

} // namespace sodhip

using namespace sodhip;

static int lumped(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w, const double *invMl,
                  double *R, int ncomp)
{
    SODHIP_REQUIRE(ctx && lpoin_w && invMl && R, "null argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    if (npoin_w == 0) return SODHIP_OK;
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    int tok = prof_begin(ctx, 3);
    k_lumped<<<(npoin_w + 255) / 256, 256, 0, ctx->stream>>>(npoin, npoin_w, lpoin_w, invMl, R,
                                                             ncomp);
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

extern "C" int sodhip_lumped_solver_scal_f64(sodhip_ctx ctx, int npoin, int npoin_w,
                                             const int *lpoin_w, const double *invMl, double *R)
{
    return lumped(ctx, npoin, npoin_w, lpoin_w, invMl, R, 1);
}

extern "C" int sodhip_lumped_solver_vect_f64(sodhip_ctx ctx, int npoin, int npoin_w,
                                             const int *lpoin_w, const double *invMl, double *R)
{
    return lumped(ctx, npoin, npoin_w, lpoin_w, invMl, R, 3);
}

extern "C" int sodhip_rk_stage_update_f64(sodhip_ctx ctx, int npoin, int npoin_w,
                                          const int *lpoin_w, const double *invMl, double dt,
                                          double lam1, double lams, double gam,
                                          const double *rho1, const double *q1, const double *E1,
                                          double *rho, double *q, double *E, double *Rmass,
                                          double *Rmom, double *Rener)
{
    SODHIP_REQUIRE(ctx && lpoin_w && invMl && rho1 && q1 && E1 && rho && q && E && Rmass && Rmom &&
                       Rener,
                   "null argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    if (npoin_w == 0) return SODHIP_OK;
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    int tok = prof_begin(ctx, 1);
    k_rk_update<<<(npoin_w + 255) / 256, 256, 0, ctx->stream>>>(
        npoin, npoin_w, lpoin_w, invMl, lam1, lams, gam * dt, rho1, q1, E1, rho, q, E, Rmass, Rmom,
        Rener);
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

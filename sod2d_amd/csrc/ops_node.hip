// ops_node.hip -- nodal kernels: lumped-mass solve and the Runge-Kutta stage
// update with the lumped inverse fused in.
#include "common.hpp"

namespace sodhip {

// lumped_solver_scal_opt / _vect_opt (mod_solver.f90:40-104): R <- R*invMl over
// working nodes, for ncomp components of an (npoin,ncomp) array, in the
// arrays' precision T.
template <typename T>
__global__ __launch_bounds__(256) void k_lumped(int npoin, int npoin_w,
                                                const int *__restrict__ lpoin_w,
                                                const T *__restrict__ invMl,
                                                T *__restrict__ R, int ncomp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const T m = invMl[n];
    for (int c = 0; c < ncomp; ++c) R[n + (size_t)npoin * c] *= m;
}

// R <- R*invMl (stored back), y <- lam1*y1 + lams*y + gdt*R over working nodes
// for the five conserved fields (time_integ_ls.f90:323-331,350-358).  RD1 / RDY:
// y1 / y enters (non-zero coefficient); otherwise the term is 0 * 0, the same
// value as 0 * (finite y), without reading the field.  Arithmetic in the
// arrays' precision T.
template <typename T, bool RD1, bool RDY>
__global__ __launch_bounds__(256) void k_rk_update(int npoin, int npoin_w,
                                                   const int *__restrict__ lpoin_w,
                                                   const T *__restrict__ invMl, T lam1, T lams,
                                                   T gdt, const T *__restrict__ rho1,
                                                   const T *__restrict__ q1,
                                                   const T *__restrict__ E1, T *__restrict__ rho,
                                                   T *__restrict__ q, T *__restrict__ E,
                                                   T *__restrict__ Rmass, T *__restrict__ Rmom,
                                                   T *__restrict__ Rener)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const T z = T(0);
    if (invMl) {
        const T m = invMl[n];
        const T rm = Rmass[n] * m;
        Rmass[n] = rm;
        rho[n] = lam1 * (RD1 ? rho1[n] : z) + lams * (RDY ? rho[n] : z) + gdt * rm;
        const T re = Rener[n] * m;
        Rener[n] = re;
        E[n] = lam1 * (RD1 ? E1[n] : z) + lams * (RDY ? E[n] : z) + gdt * re;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t k = n + (size_t)npoin * d;
            const T rq = Rmom[k] * m;
            Rmom[k] = rq;
            q[k] = lam1 * (RD1 ? q1[k] : z) + lams * (RDY ? q[k] : z) + gdt * rq;
        }
    } else { // R already carries the mass inverse (the residual kept from updateF)
        rho[n] = lam1 * (RD1 ? rho1[n] : z) + lams * (RDY ? rho[n] : z) + gdt * Rmass[n];
        E[n] = lam1 * (RD1 ? E1[n] : z) + lams * (RDY ? E[n] : z) + gdt * Rener[n];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t k = n + (size_t)npoin * d;
            q[k] = lam1 * (RD1 ? q1[k] : z) + lams * (RDY ? q[k] : z) + gdt * Rmom[k];
        }
    }
}

} // namespace sodhip

using namespace sodhip;

template <typename T>
static int lumped(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w, const T *invMl,
                  T *R, int ncomp)
{
    SODHIP_REQUIRE(ctx && lpoin_w && invMl && R, "null argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    if (npoin_w == 0) return SODHIP_OK;
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    int tok = prof_begin(ctx, 3);
    k_lumped<T><<<(npoin_w + 255) / 256, 256, 0, ctx->stream>>>(npoin, npoin_w, lpoin_w, invMl, R,
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

extern "C" int sodhip_lumped_solver_scal_f32(sodhip_ctx ctx, int npoin, int npoin_w,
                                             const int *lpoin_w, const float *invMl, float *R)
{
    return lumped(ctx, npoin, npoin_w, lpoin_w, invMl, R, 1);
}

extern "C" int sodhip_lumped_solver_vect_f32(sodhip_ctx ctx, int npoin, int npoin_w,
                                             const int *lpoin_w, const float *invMl, float *R)
{
    return lumped(ctx, npoin, npoin_w, lpoin_w, invMl, R, 3);
}

template <typename T>
static int rk_stage_update(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                           const T *invMl, T dt, T lam1, T lams, T gam, const T *rho1, const T *q1,
                           const T *E1, T *rho, T *q, T *E, T *Rmass, T *Rmom, T *Rener)
{
    SODHIP_REQUIRE(ctx && lpoin_w && rho1 && q1 && E1 && rho && q && E && Rmass && Rmom && Rener,
                   "null argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    if (npoin_w == 0) return SODHIP_OK;
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    int tok = prof_begin(ctx, 3);
#define SODHIP_RKU(R1, RY)                                                                     \
    k_rk_update<T, R1, RY><<<(npoin_w + 255) / 256, 256, 0, ctx->stream>>>(                    \
        npoin, npoin_w, lpoin_w, invMl, lam1, lams, gam * dt, rho1, q1, E1, rho, q, E, Rmass,  \
        Rmom, Rener)
    if (lam1 != T(0) && lams != T(0))
        SODHIP_RKU(true, true);
    else if (lam1 != T(0))
        SODHIP_RKU(true, false);
    else if (lams != T(0))
        SODHIP_RKU(false, true);
    else
        SODHIP_RKU(false, false);
#undef SODHIP_RKU
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

extern "C" int sodhip_rk_stage_update_f64(sodhip_ctx ctx, int npoin, int npoin_w,
                                          const int *lpoin_w, const double *invMl, double dt,
                                          double lam1, double lams, double gam,
                                          const double *rho1, const double *q1, const double *E1,
                                          double *rho, double *q, double *E, double *Rmass,
                                          double *Rmom, double *Rener)
{
    return rk_stage_update<double>(ctx, npoin, npoin_w, lpoin_w, invMl, dt, lam1, lams, gam, rho1,
                                   q1, E1, rho, q, E, Rmass, Rmom, Rener);
}

extern "C" int sodhip_rk_stage_update_f32(sodhip_ctx ctx, int npoin, int npoin_w,
                                          const int *lpoin_w, const float *invMl, float dt,
                                          float lam1, float lams, float gam, const float *rho1,
                                          const float *q1, const float *E1, float *rho, float *q,
                                          float *E, float *Rmass, float *Rmom, float *Rener)
{
    return rk_stage_update<float>(ctx, npoin, npoin_w, lpoin_w, invMl, dt, lam1, lams, gam, rho1,
                                  q1, E1, rho, q, E, Rmass, Rmom, Rener);
}

namespace sodhip {
// The node update of a classical RK4 stage of rk_4_main (time_integ.f90,
// compiled out in the reference by #if 0) on assembled residuals: R <- R*invMl
// (lumped_solver_*, :431-435; skipped when invMl is NULL), Rsum <- Rsum + b R
// (:442-451, FIRST: Rsum starts from zero, :234-240), then the next stage's
// state aux = y - dt a_next R (:266-275) or, at the LAST stage,
// y <- y - dt Rsum (:460-468), over the working nodes.
template <bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void k_rk4_update(
    int npoin, int npoin_w, const int *__restrict__ lpoin_w, const double *__restrict__ invMl,
    double dta, double b, double dt, double *__restrict__ rho, double *__restrict__ q,
    double *__restrict__ E, double *__restrict__ arho, double *__restrict__ aq,
    double *__restrict__ aE, double *__restrict__ Rmass, double *__restrict__ Rmom,
    double *__restrict__ Rener, double *__restrict__ Smass, double *__restrict__ Smom,
    double *__restrict__ Sener)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const size_t k[5] = {(size_t)n, (size_t)n, n + (size_t)npoin, n + 2 * (size_t)npoin,
                         (size_t)n};
    double *yv[5] = {rho, q, q, q, E};
    double *av[5] = {arho, aq, aq, aq, aE};
    double *rv[5] = {Rmass, Rmom, Rmom, Rmom, Rener};
    double *sv[5] = {Smass, Smom, Smom, Smom, Sener};
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        double r = rv[f][k[f]];
        if (invMl) {
            r = r * invMl[n];
            rv[f][k[f]] = r;
        }
        const double sn = (FIRST ? 0.0 : sv[f][k[f]]) + b * r;
        sv[f][k[f]] = sn;
        if constexpr (LAST)
            yv[f][k[f]] = yv[f][k[f]] - dt * sn;
        else
            av[f][k[f]] = yv[f][k[f]] - dta * r;
    }
}
} // namespace sodhip

extern "C" int sodhip_rk4_stage_update_f64(sodhip_ctx ctx, int npoin, int npoin_w,
                                           const int *lpoin_w, const double *invMl, double dt,
                                           double a_next, double b, int first, int last,
                                           double *rho, double *q, double *E, double *aux_rho,
                                           double *aux_q, double *aux_E, double *Rmass,
                                           double *Rmom, double *Rener, double *Rmass_sum,
                                           double *Rmom_sum, double *Rener_sum)
{
    SODHIP_REQUIRE(ctx && lpoin_w && rho && q && E && Rmass && Rmom && Rener && Rmass_sum &&
                       Rmom_sum && Rener_sum,
                   "null argument");
    SODHIP_REQUIRE(last || (aux_rho && aux_q && aux_E), "the next stage's state is needed");
    SODHIP_REQUIRE(last || (aux_rho != rho && aux_q != q && aux_E != E),
                   "the next stage's state must not overwrite y");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    if (npoin_w == 0) return SODHIP_OK;
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    int tok = prof_begin(ctx, 3);
#define SODHIP_RK4U(FI, LA)                                                                    \
    k_rk4_update<FI, LA><<<(npoin_w + 255) / 256, 256, 0, ctx->stream>>>(                      \
        npoin, npoin_w, lpoin_w, invMl, dt * a_next, b, dt, rho, q, E, aux_rho, aux_q, aux_E,   \
        Rmass, Rmom, Rener, Rmass_sum, Rmom_sum, Rener_sum)
    if (first && last)
        SODHIP_RK4U(true, true);
    else if (first)
        SODHIP_RK4U(true, false);
    else if (last)
        SODHIP_RK4U(false, true);
    else
        SODHIP_RK4U(false, false);
#undef SODHIP_RK4U
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

// ---------------------------------------------------------------- EOS and friends
namespace sodhip {

// updateF "Update u and EOS" (time_integ_ls.f90:574-587)
__global__ __launch_bounds__(256) void k_eos_stage(int npoin, int npoin_w,
                                                   const int *__restrict__ lpoin_w, double g,
                                                   double Rgas, const double *__restrict__ rho,
                                                   const double *__restrict__ q,
                                                   const double *__restrict__ E,
                                                   double *__restrict__ u, double *__restrict__ e_int,
                                                   double *__restrict__ pr,
                                                   double *__restrict__ aux_h,
                                                   double *__restrict__ Tem)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const double r = rho[n];
    double umag = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double ud = q[n + (size_t)npoin * d] / r;
        u[n + (size_t)npoin * d] = ud;
        umag += ud * ud;
    }
    const double ei = E[n] / r - 0.5 * umag;
    const double p = r * (g - 1.0) * ei;
    e_int[n] = ei;
    pr[n] = p;
    if (aux_h) aux_h[n] = (g / (g - 1.0)) * p / r;
    Tem[n] = p / (r * Rgas);
}

// updateF's EOS block on the working nodes (wflag) followed by
// sutherland_viscosity on every node, in one pass over all npoin: each output
// has the bits of k_eos_stage then k_sutherland (nodes outside the working
// list keep their temperature)
__global__ __launch_bounds__(256) void k_eos_stage_suth(
    int npoin, const unsigned char *__restrict__ wflag, double g, double Rgas,
    double mu_factor, const double *__restrict__ rho, const double *__restrict__ q,
    const double *__restrict__ E, double *__restrict__ u, double *__restrict__ e_int,
    double *__restrict__ pr, double *__restrict__ aux_h, double *__restrict__ Tem,
    double *__restrict__ mu_fluid)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= npoin) return;
    double T;
    if (wflag[n]) {
        const double r = rho[n];
        double umag = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double ud = q[n + (size_t)npoin * d] / r;
            u[n + (size_t)npoin * d] = ud;
            umag += ud * ud;
        }
        const double ei = E[n] / r - 0.5 * umag;
        const double p = r * (g - 1.0) * ei;
        e_int[n] = ei;
        pr[n] = p;
        if (aux_h) aux_h[n] = (g / (g - 1.0)) * p / r;
        T = p / (r * Rgas);
        Tem[n] = T;
    } else {
        T = Tem[n];
    }
    mu_fluid[n] = mu_factor * 0.000001458 * pow(T, 1.5) / (T + 110.4);
}

int eos_stage_sutherland(sodhip_ctx ctx, const MeshCache *m, double gamma_gas, double Rgas,
                         double mu_factor, const double *rho, const double *q, const double *E,
                         double *u, double *e_int, double *pr, double *aux_h, double *Tem,
                         double *mu_fluid)
{
    SODHIP_REQUIRE(m && m->wflag, "working-node flags not built");
    const int npoin = m->npoin;
    int tok = prof_begin(ctx, 3);
    k_eos_stage_suth<<<(npoin + 255) / 256, 256, 0, ctx->stream>>>(
        npoin, m->wflag, gamma_gas, Rgas, mu_factor, rho, q, E, u, e_int, pr, aux_h, Tem,
        mu_fluid);
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

// rk_4_ls_main end-of-step EOS (time_integ_ls.f90:391-407)
__global__ __launch_bounds__(256) void k_eos_step(int npoin, int npoin_w,
                                                  const int *__restrict__ lpoin_w, double g,
                                                  double Rgas, const double *__restrict__ rho,
                                                  const double *__restrict__ q,
                                                  const double *__restrict__ E,
                                                  double *__restrict__ u, double *__restrict__ e_int,
                                                  double *__restrict__ pr,
                                                  double *__restrict__ csound,
                                                  double *__restrict__ machno,
                                                  double *__restrict__ Tem)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const double r = rho[n];
    double umag = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double ud = q[n + (size_t)npoin * d] / r;
        u[n + (size_t)npoin * d] = ud;
        umag += ud * ud;
    }
    const double ei = E[n] / r - 0.5 * umag;
    const double p = r * (g - 1.0) * ei;
    e_int[n] = ei;
    pr[n] = p;
    const double c = sqrt(g * p / r);
    csound[n] = c;
    machno[n] = sqrt(umag) / c;
    Tem[n] = p / (r * Rgas);
}

// end of a step on the working nodes: the EOS of k_eos_step, then
// thermo_entropy (eta(:,1) <- eta(:,2), eta(:,2) from the new rho and p) and
// sutherland_viscosity from the new T, each with the arithmetic of its own
// kernel, so every output has the same bits as the three kernels give
// The per-node arrays p_inf_n, rho_inf_n, mu_factor_n (each may be nullptr)
// replace the scalars on a mesh joining instances that differ in them.
__global__ __launch_bounds__(256) void k_end_state(
    int npoin, int npoin_w, const int *__restrict__ lpoin_w, double g, double Rgas,
    double p_inf_s, double rho_inf_s, double mu_factor_s, const double *__restrict__ p_inf_n,
    const double *__restrict__ rho_inf_n, const double *__restrict__ mu_factor_n,
    const double *__restrict__ rho, const double *__restrict__ q, const double *__restrict__ E,
    double *__restrict__ u, double *__restrict__ e_int, double *__restrict__ pr,
    double *__restrict__ csound, double *__restrict__ machno, double *__restrict__ Tem,
    double *__restrict__ eta, double *__restrict__ mu_fluid)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const double p_inf = p_inf_n ? p_inf_n[n] : p_inf_s;
    const double rho_inf = rho_inf_n ? rho_inf_n[n] : rho_inf_s;
    const double mu_factor = mu_factor_n ? mu_factor_n[n] : mu_factor_s;
    const double r = rho[n];
    const double e2 = eta[n + (size_t)npoin];
    double umag = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double ud = q[n + (size_t)npoin * d] / r;
        u[n + (size_t)npoin * d] = ud;
        umag += ud * ud;
    }
    const double ei = E[n] / r - 0.5 * umag;
    const double p = r * (g - 1.0) * ei;
    e_int[n] = ei;
    pr[n] = p;
    const double c = sqrt(g * p / r);
    csound[n] = c;
    machno[n] = sqrt(umag) / c;
    const double T = p / (r * Rgas);
    Tem[n] = T;
    eta[n] = e2;
    eta[n + (size_t)npoin] =
        (r / (g - 1.0)) * log(fmax(p / p_inf, 2.220446049250313e-16) / pow(r / rho_inf, g));
    mu_fluid[n] = mu_factor * 0.000001458 * pow(T, 1.5) / (T + 110.4);
}

// sutherland_viscosity (mod_fluid_viscosity.f90:21-41), over all npoin
__global__ __launch_bounds__(256) void k_sutherland(int npoin, const double *__restrict__ Tem,
                                                    const double *__restrict__ mu_factor,
                                                    double mu_factor_scalar,
                                                    double *__restrict__ mu_fluid)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= npoin) return;
    const double T = Tem[n];
    const double f = mu_factor ? mu_factor[n] : mu_factor_scalar;
    mu_fluid[n] = f * 0.000001458 * pow(T, 1.5) / (T + 110.4);
}

// thermo_entropy (mod_entropy_viscosity.f90:411-433); eta is (npoin,2)
__global__ __launch_bounds__(256) void k_thermo_entropy(int npoin, int npoin_w,
                                                        const int *__restrict__ lpoin_w, double g,
                                                        double p_inf, double rho_inf,
                                                        const double *__restrict__ rho,
                                                        const double *__restrict__ pr,
                                                        const double *__restrict__ u,
                                                        double *__restrict__ eta,
                                                        double *__restrict__ feta)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const double e1 = eta[n + (size_t)npoin];
    eta[n] = e1;
    const double r = rho[n];
    eta[n + (size_t)npoin] =
        (r / (g - 1.0)) * log(fmax(pr[n] / p_inf, 2.220446049250313e-16) / pow(r / rho_inf, g));
    if (feta) {
#pragma unroll
        for (int d = 0; d < 3; ++d) feta[n + (size_t)npoin * d] = u[n + (size_t)npoin * d] * e1;
    }
}

// entropy residual of rk_4_ls_main (time_integ_ls.f90:437-443); Reta and eta are (npoin,2)
__global__ __launch_bounds__(256) void k_entropy_residual(int npoin, int npoin_w,
                                                          const int *__restrict__ lpoin_w,
                                                          double factor_comp, double dt,
                                                          double *__restrict__ Reta,
                                                          const double *__restrict__ eta,
                                                          double *__restrict__ auxReta)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const double r1 = Reta[n], r2 = Reta[n + (size_t)npoin];
    auxReta[n] = (1.5 * r2 - 0.5 * r1) + factor_comp * (eta[n + (size_t)npoin] - eta[n]) / dt;
    Reta[n] = r2;
}

} // namespace sodhip

#define SODHIP_NODE_LAUNCH(kind, count, kernel, ...)                                        \
    do {                                                                                   \
        if ((count) > 0) {                                                                 \
            int _tok = prof_begin(ctx, kind);                                              \
            kernel<<<((count) + 255) / 256, 256, 0, ctx->stream>>>(__VA_ARGS__);           \
            prof_end(ctx, _tok);                                                           \
            SODHIP_TRY_HIP(hipGetLastError());                                             \
        }                                                                                  \
    } while (0)

extern "C" int sodhip_eos_stage_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                    double gamma_gas, double Rgas, const double *rho,
                                    const double *q, const double *E, double *u, double *e_int,
                                    double *pr, double *aux_h, double *Tem)
{
    SODHIP_REQUIRE(ctx && lpoin_w && rho && q && E && u && e_int && pr && Tem, "null argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    SODHIP_NODE_LAUNCH(3, npoin_w, k_eos_stage, npoin, npoin_w, lpoin_w, gamma_gas, Rgas, rho, q, E,
                       u, e_int, pr, aux_h, Tem);
    return SODHIP_OK;
}

extern "C" int sodhip_eos_step_f64(sodhip_ctx ctx, int npoin, int npoin_w, const int *lpoin_w,
                                   double gamma_gas, double Rgas, const double *rho,
                                   const double *q, const double *E, double *u, double *e_int,
                                   double *pr, double *csound, double *machno, double *Tem)
{
    SODHIP_REQUIRE(ctx && lpoin_w && rho && q && E && u && e_int && pr && csound && machno && Tem,
                   "null argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    SODHIP_NODE_LAUNCH(3, npoin_w, k_eos_step, npoin, npoin_w, lpoin_w, gamma_gas, Rgas, rho, q, E,
                       u, e_int, pr, csound, machno, Tem);
    return SODHIP_OK;
}

extern "C" int sodhip_eos_entropy_step_f64(sodhip_ctx ctx, int npoin, int npoin_w,
                                           const int *lpoin_w, double gamma_gas, double Rgas,
                                           double p_inf, double rho_inf, double mu_factor,
                                           const double *rho, const double *q, const double *E,
                                           double *u, double *e_int, double *pr, double *csound,
                                           double *machno, double *Tem, double *eta,
                                           double *mu_fluid)
{
    SODHIP_REQUIRE(ctx && lpoin_w && rho && q && E && u && e_int && pr && csound && machno && Tem &&
                       eta && mu_fluid,
                   "null argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    SODHIP_NODE_LAUNCH(3, npoin_w, k_end_state, npoin, npoin_w, lpoin_w, gamma_gas, Rgas, p_inf,
                       rho_inf, mu_factor, (const double *)nullptr, (const double *)nullptr,
                       (const double *)nullptr, rho, q, E, u, e_int, pr, csound, machno, Tem, eta,
                       mu_fluid);
    return SODHIP_OK;
}

extern "C" int sodhip_eos_entropy_step_ens_f64(sodhip_ctx ctx, int npoin, int npoin_w,
                                               const int *lpoin_w, double gamma_gas, double Rgas,
                                               const sodhip_instances *ens, const double *rho,
                                               const double *q, const double *E, double *u,
                                               double *e_int, double *pr, double *csound,
                                               double *machno, double *Tem, double *eta,
                                               double *mu_fluid)
{
    SODHIP_REQUIRE(ctx && ens && ens->p_inf && ens->rho_inf && ens->mu_factor && lpoin_w && rho &&
                       q && E && u && e_int && pr && csound && machno && Tem && eta && mu_fluid,
                   "null argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    SODHIP_NODE_LAUNCH(3, npoin_w, k_end_state, npoin, npoin_w, lpoin_w, gamma_gas, Rgas, 0.0, 0.0,
                       0.0, ens->p_inf, ens->rho_inf, ens->mu_factor, rho, q, E, u, e_int, pr,
                       csound, machno, Tem, eta, mu_fluid);
    return SODHIP_OK;
}

extern "C" int sodhip_sutherland_f64(sodhip_ctx ctx, int npoin, const double *Tem,
                                     const double *mu_factor, double mu_factor_scalar,
                                     double *mu_fluid)
{
    SODHIP_REQUIRE(ctx && Tem && mu_fluid, "null argument");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    SODHIP_NODE_LAUNCH(3, npoin, k_sutherland, npoin, Tem, mu_factor, mu_factor_scalar, mu_fluid);
    return SODHIP_OK;
}

extern "C" int sodhip_thermo_entropy_f64(sodhip_ctx ctx, int npoin, int npoin_w,
                                         const int *lpoin_w, double gamma_gas, double p_inf,
                                         double rho_inf, const double *rho, const double *pr,
                                         const double *u, double *eta, double *feta)
{
    SODHIP_REQUIRE(ctx && lpoin_w && rho && pr && eta && (u || !feta), "null argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    SODHIP_NODE_LAUNCH(3, npoin_w, k_thermo_entropy, npoin, npoin_w, lpoin_w, gamma_gas, p_inf,
                       rho_inf, rho, pr, u, eta, feta);
    return SODHIP_OK;
}

extern "C" int sodhip_entropy_residual_f64(sodhip_ctx ctx, int npoin, int npoin_w,
                                           const int *lpoin_w, double factor_comp, double dt,
                                           double *Reta, const double *eta, double *auxReta)
{
    SODHIP_REQUIRE(ctx && lpoin_w && Reta && eta && auxReta, "null argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    SODHIP_NODE_LAUNCH(3, npoin_w, k_entropy_residual, npoin, npoin_w, lpoin_w, factor_comp, dt,
                       Reta, eta, auxReta);
    return SODHIP_OK;
}

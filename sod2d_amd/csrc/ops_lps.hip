// ops_lps.hip -- local-projection stabilisation of the viscous operator
// (elem_stab.f90): the element time scale comp_tau, the lumped L2 projections
// of the gradients full_proj_ijk, and the stabilised viscous operator
// full_diff_stab_ijk, with their drop-in C-ABI entry points (FP64).
//
// Mapping as the other element kernels: one element per workgroup of
// NT = roundup((p+1)^3, 64) lanes, lane a owns element-local node /
// quadrature point a, nodal fields staged field-major in LDS, line sums from
// the per-lane GLL line tables, element records gathered per node in the
// reference's (element, node) order by a second kernel (deterministic, no
// atomics).
#include "common.hpp"
#include "device.hpp"
#include "element.hpp"

namespace sodhip {


struct LpsArgs {
    int nelem, npoin;
    const int *connecT; // element-major, 0-based
    const int *perm;    // record position of each element-local node
    const double *He, *gpvol, *dl;
    const int *inv, *gI, *gJ, *gK;
    const double *rho_n, *rho, *u, *Tem, *mu_fluid, *mu_e, *mu_sgs;
    const double *pMass, *pEner, *pMX, *pMY, *pMZ; // projections (npoin,3)
    // or one record per node (kProjRec values: the upper triangle of the
    // projected tensor, then ProjMass, then ProjEner)
    const double *prj;
    const double *tau;                             // tau_stab(nelem)
    double Cp, Pr, c_rho, c_ener, c_lps;
    double *out;
};

// Physical gradients of rho, T and u at the lane's point from the staged
// element fields s (rho, T, u1, u2, u3, each NT long):
//   gR[i] = sum_j He(i,j) dR/dxi_j,  gU[i][j] = sum_k He(j,k) du_i/dxi_k
// (elem_stab.f90:300-335, :497-531).
template <int P>
__device__ __forceinline__ void lps_gradients(const LineTab<P, double> &lt, const double *s,
                                              int a, const double *hp, double gR[3],
                                              double gT[3], double gU[3][3])
{
    constexpr int NT = EG<P>::NT;
    double g[5][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double c0 = lt.ds[k];
#pragma unroll
        for (int f = 0; f < 5; ++f) g[f][k] = c0 * s[f * NT + a];
#pragma unroll
        for (int o = 0; o < P; ++o) {
            const int b = lt.off[k][o];
            const double c = lt.d[k][o];
#pragma unroll
            for (int f = 0; f < 5; ++f) g[f][k] += c * s[f * NT + b];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        gR[i] = hp[i] * g[0][0] + hp[i + 3] * g[0][1] + hp[i + 6] * g[0][2];
        gT[i] = hp[i] * g[1][0] + hp[i + 3] * g[1][1] + hp[i + 6] * g[1][2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            gU[i][j] = hp[j] * g[2 + i][0] + hp[j + 3] * g[2 + i][1] + hp[j + 6] * g[2 + i][2];
}

// Lane L of an element kernel evaluates the node at lexicographic position L
// (I fastest) and the LDS fields are kept in that order, so the 16 lanes of
// an LDS access phase read 16 distinct positions along every line.  Returns
// the gmsh index of the lane's node and turns the line table's neighbours
// into lexicographic positions; tmp holds NN ints (LDS).
template <int P>
__device__ __forceinline__ int lex_lane(const int *__restrict__ perm, int lane, bool act,
                                        int *tmp)
{
    constexpr int NN = EG<P>::NN;
    if (act) tmp[min(max(perm[lane], 0), NN - 1)] = lane;
    lds_barrier();
    const int a = act ? min(max(tmp[lane], 0), NN - 1) : 0;
    lds_barrier();
    return a;
}

template <int P>
__device__ __forceinline__ void lex_neighbours(const int *__restrict__ perm,
                                               LineTab<P, double> &lt)
{
    constexpr int NN = EG<P>::NN;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int o = 0; o < P; ++o) lt.off[k][o] = min(max(perm[lt.off[k][o]], 0), NN - 1);
}

// full_proj_ijk's element part (elem_stab.f90:489-563): records of 12 values
// per element-node, gpvol * (the upper triangle t11 t12 t13 t22 t23 t33 of the
// symmetric tensor 2 S - 2/3 div u I, grad rho, grad T).  The tensor rows the
// reference projects separately are rebuilt from the triangle by the gather:
// tau(i,j) and tau(j,i) are the same sum gradU(i,j) + gradU(j,i).
template <int P>
__global__ __launch_bounds__(EG<P>::NT) void k_lps_proj(LpsArgs A)
{
    constexpr int NN = EG<P>::NN, NT = EG<P>::NT;
    __shared__ double s[5 * NT];
    const int lane = threadIdx.x;
    const bool act = lane < NN;
    const int a = lex_lane<P>(A.perm, lane, act, (int *)s);
    const int pa = act ? lane : 0;
    LineTab<P, double> lt;
    build_linetab<P, double>(a, A.dl, A.inv, A.gI, A.gJ, A.gK, lt);
    lex_neighbours<P>(A.perm, lt);
    const ElemRange er = elem_range(A.nelem);
    for (int e = er.begin; e < er.end; e += er.stride) {
        const size_t ea = (size_t)e * NN + a;
        const int n = act ? A.connecT[ea] : 0;
        if (act) {
            s[pa] = A.rho[n];
            s[NT + pa] = A.Tem[n];
#pragma unroll
            for (int d = 0; d < 3; ++d) s[(2 + d) * NT + pa] = A.u[n + (size_t)A.npoin * d];
        }
        double hp[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) hp[i] = A.He[9 * ea + i];
        const double w = A.gpvol[ea];
        lds_barrier();
        double gR[3], gT[3], gU[3][3];
        lps_gradients<P>(lt, s, pa, hp, gR, gT, gU);
        const double divU = gU[0][0] + gU[1][1] + gU[2][2];
        if (act) {
            double o[12];
            int f = 0;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int d = i; d < 3; ++d) {
                    double t = gU[i][d] + gU[d][i];
                    if (i == d) t = t - (2.0 / 3.0) * divU;
                    o[f++] = w * t;
                }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                o[6 + d] = w * gR[d];
                o[9 + d] = w * gT[d];
            }
            double2 *dst = (double2 *)(A.out + 12 * ((size_t)e * NN + pa));
#pragma unroll
            for (int v = 0; v < 6; ++v) dst[v] = make_double2(o[2 * v], o[2 * v + 1]);
        }
        lds_barrier();
    }
}

// full_diff_stab_ijk (elem_stab.f90:207-438): per quadrature point the viscous
// fluxes with the LPS terms -c_lps tau_e (Proj - grad) (scaled by rho_n for
// momentum, rho_n Cp/Pr for energy), mapped to contravariant components
// G_k(F) = gpvol sum_d He(d,k) F_d and staged in LDS; then the weak divergence
// sum_k sum_{b on the k-line} dl(b,k,I_k(a)) G_k(F)(b).  Records of 5 values
// (mass, momentum 1..3, energy) per element-node; ACC adds them to the records
// already there (the convection of the same stage).
// AOS (the fused stage): the projections come as node records (A.prj), 96
// contiguous bytes per node, instead of 15 arrays.
template <int P, bool ACC = false, bool AOS = false>
__global__ __launch_bounds__(EG<P>::NT) void k_diff_stab(LpsArgs A)
{
    constexpr int NN = EG<P>::NN, NT = EG<P>::NT;
    __shared__ double s[5 * NT];
    __shared__ double G[15 * NT]; // flux f (mass, mom x, y, z, energy), direction k
    const int lane = threadIdx.x;
    const bool act = lane < NN;
    const int a = lex_lane<P>(A.perm, lane, act, (int *)s);
    const int pa = act ? lane : 0;
    LineTab<P, double> lt;
    build_linetab<P, double>(a, A.dl, A.inv, A.gI, A.gJ, A.gK, lt);
    lex_neighbours<P>(A.perm, lt);
    const double twoThirds = 2.0 / 3.0;
    const ElemRange er = elem_range(A.nelem);
    for (int e = er.begin; e < er.end; e += er.stride) {
        const size_t ea = (size_t)e * NN + a;
        const int n = act ? A.connecT[ea] : 0;
        const size_t np = (size_t)A.npoin;
        double ul[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) ul[d] = A.u[n + np * d];
        if (act) {
            s[pa] = A.rho[n];
            s[NT + pa] = A.Tem[n];
#pragma unroll
            for (int d = 0; d < 3; ++d) s[(2 + d) * NT + pa] = ul[d];
        }
        double hp[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) hp[i] = A.He[9 * ea + i];
        const double w = A.gpvol[ea];
        const double rhon = A.rho_n[n], mufl = A.mu_fluid[n];
        const size_t eg = (size_t)e + (size_t)A.nelem * a; // (nelem,ngaus) element-fastest
        const double mue = A.mu_e[eg];
        const double musgs = A.mu_sgs ? A.mu_sgs[eg] : 0.0;
        const double taus = A.tau[e];
        lds_barrier();
        // the node's projections: MX MY MZ Mass Ener, three components each
        double pj[15];
        if constexpr (AOS) {
            const double2 *r = (const double2 *)(A.prj + kProjRec * (size_t)n);
            double v[12];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double2 x = r[q];
                v[2 * q] = x.x;
                v[2 * q + 1] = x.y;
            }
            // rows of the symmetric tensor, ProjMass, ProjEner
            const int at[15] = {0, 1, 2, 1, 3, 4, 2, 4, 5, 6, 7, 8, 9, 10, 11};
#pragma unroll
            for (int j = 0; j < 15; ++j) pj[j] = v[at[j]];
        }
        double gR[3], gT[3], gU[3][3];
        lps_gradients<P>(lt, s, pa, hp, gR, gT, gU);
        const double divU = gU[0][0] + gU[1][1] + gU[2][2];
        const double nu_e = A.c_rho * mue / rhon;
        const double mu_fgp = mufl + rhon * musgs;
        const double kappa_e =
            mufl * A.Cp / A.Pr + A.c_ener * mue * A.Cp / A.Pr + A.Cp * rhon * musgs / 0.9;
        const double cl = -A.c_lps * taus;
        if constexpr (!AOS) {
            const double *pf[5] = {A.pMX, A.pMY, A.pMZ, A.pMass, A.pEner};
#pragma unroll
            for (int f = 0; f < 5; ++f)
#pragma unroll
                for (int d = 0; d < 3; ++d) pj[3 * f + d] = pf[f][n + np * d];
        }
        double F[5][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double t = gU[i][d] + gU[d][i];
                if (i == d) t = t - twoThirds * divU;
                F[1 + i][d] = cl * (pj[3 * i + d] - t) * rhon;
            }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            F[4][d] = cl * (pj[12 + d] - gT[d]) * rhon * A.Cp / A.Pr;
            F[0][d] = cl * (pj[9 + d] - gR[d]);
        }
        double tauU[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                tauU[i] = tauU[i] + (mu_fgp + mue) * (gU[i][j] + gU[j][i]) * ul[j];
                double t = (mu_fgp + mue) * (gU[i][j] + gU[j][i]);
                if (i == j) t = t - mu_fgp * twoThirds * divU;
                F[1 + i][j] = F[1 + i][j] + t;
            }
            tauU[i] = tauU[i] - mu_fgp * twoThirds * divU * ul[i];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            F[4][d] = F[4][d] + kappa_e * gT[d] + tauU[d];
            F[0][d] = F[0][d] + nu_e * gR[d];
        }
        if (act) {
#pragma unroll
            for (int f = 0; f < 5; ++f)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    G[(3 * f + k) * NT + pa] =
                        w * (hp[3 * k] * F[f][0] + hp[1 + 3 * k] * F[f][1] +
                             hp[2 + 3 * k] * F[f][2]);
        }
        lds_barrier();
        double R[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double *gk = G + (3 * f + k) * NT;
                acc += lt.ds[k] * gk[pa];
#pragma unroll
                for (int o = 0; o < P; ++o) acc += lt.t[k][o] * gk[lt.off[k][o]];
            }
            R[f] = acc;
        }
        if (act) {
            double *o = A.out + 5 * ((size_t)e * NN + pa);
#pragma unroll
            for (int f = 0; f < 5; ++f) o[f] = ACC ? o[f] + R[f] : R[f];
        }
        lds_barrier();
    }
}

// comp_tau (elem_stab.f90:589-616): tau_e = max over the element's nodes of
// h (fact_c c + |u|), h = helem_k (c_species_stab / p); one thread per
// element over the reference's element-fastest connec, the nodes visited an
// i-line at a time (build_alex) with all loads of a line issued together
// (fmax is exact).
__global__ __launch_bounds__(256) void k_comp_tau(int nelem, int nnode, int porder,
                                                  const int *__restrict__ connec,
                                                  const int *__restrict__ perm,
                                                  const double *__restrict__ csound,
                                                  const double *__restrict__ u, int npoin,
                                                  const double *__restrict__ helem_k,
                                                  double c_species_stab, double fact_c,
                                                  double *__restrict__ tau)
{
    __shared__ int alex[125];
    build_alex(perm, nnode, alex);
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    const double h = helem_k[e] * (c_species_stab / (double)porder);
    const int KA = porder + 1;
    double taul = 0.0;
    for (int l0 = 0; l0 < nnode; l0 += KA) {
        int n[5];
#pragma unroll
        for (int j = 0; j < 5; ++j)
            n[j] = connec[e + (size_t)nelem * alex[min(l0 + min(j, KA - 1), nnode - 1)]] - 1;
        double u0[5], u1[5], u2[5], cs[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            u0[j] = u[n[j]];
            u1[j] = u[n[j] + (size_t)npoin];
            u2[j] = u[n[j] + 2 * (size_t)npoin];
            cs[j] = csound[n[j]];
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const double aux1 = fact_c * cs[j] + sqrt(u0[j] * u0[j] + u1[j] * u1[j] + u2[j] * u2[j]);
            taul = fmax(taul, h * aux1);
        }
    }
    tau[e] = taul;
}

// Node-ordered assemblies.  Node n's slots row[n]..row[n+1]-1 are taken four
// at a time with every record load issued before the additions, which stay in
// slot order: R_f(n) = (init ? 0 : R_f(n)) + sum_j fact * rec_j[f].  Blocks are
// mapped XCD-aware (the nodes of one XCD are one contiguous range).
__device__ __forceinline__ int xcd_node(int nthr)
{
    const int per = gridDim.x >> 3; // grid is a multiple of 8
    const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int n = blk * blockDim.x + threadIdx.x;
    return n < nthr ? n : -1;
}

struct Rec5Out {
    double *R[5];
};
__global__ __launch_bounds__(256) void k_gather5(int npoin, const int *__restrict__ row,
                                                 const int *__restrict__ src,
                                                 const double *__restrict__ scr, Rec5Out out,
                                                 int init, double fact)
{
    const int n = xcd_node(npoin);
    if (n < 0) return;
    const int j0 = row[n], j1 = row[n + 1];
    double acc[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) acc[f] = init ? 0.0 : out.R[f][n];
    for (int j = j0; j < j1; j += 4) {
        int sj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) sj[u] = j + u < j1 ? src[j + u] : -1;
        double v[4][5];
#pragma unroll
        for (int u = 0; u < 4; ++u) load_rec5(scr, max(sj[u], 0), v[u]); // absent: not added
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (sj[u] >= 0) {
#pragma unroll
                for (int f = 0; f < 5; ++f) acc[f] += fact * v[u][f];
            }
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) out.R[f][n] = acc[f];
}

// full_proj_ijk's assembly and lumped solve: the 12-value records summed per
// node (the projections start from zero, elem_stab.f90:466-472), scaled by
// invMl on the working nodes (lumped_solver_vect_opt, :576-581), and written
// as ProjMX/MY/MZ (rows of the symmetric tensor), ProjMass, ProjEner.
struct ProjOut {
    double *MX, *MY, *MZ, *Mass, *Ener; // the reference's arrays, or all NULL
    double *prj;                        // one record of kProjRec values per node, or NULL
};
__global__ __launch_bounds__(256) void k_gather_proj(int npoin, const int *__restrict__ row,
                                                     const int *__restrict__ src,
                                                     const double *__restrict__ scr,
                                                     const unsigned char *__restrict__ wflag,
                                                     const double *__restrict__ invMl,
                                                     ProjOut out)
{
    const int n = xcd_node(npoin);
    if (n < 0) return;
    const int j0 = row[n], j1 = row[n + 1];
    double acc[12];
#pragma unroll
    for (int f = 0; f < 12; ++f) acc[f] = 0.0;
    // two slots per step; the loads of an absent slot read record 0 (always
    // present) unconditionally, which is cheaper than masking each load, and
    // are not added
    for (int j = j0; j < j1; j += 2) {
        int sj[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) sj[u] = j + u < j1 ? src[j + u] : -1;
        double2 v[2][6];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const double2 *r = (const double2 *)(scr + 12 * (size_t)max(sj[u], 0));
#pragma unroll
            for (int q = 0; q < 6; ++q) v[u][q] = r[q];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
            if (sj[u] >= 0) {
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    acc[2 * q] += v[u][q].x;
                    acc[2 * q + 1] += v[u][q].y;
                }
            }
    }
    if (wflag[n]) {
        const double m = invMl[n];
#pragma unroll
        for (int f = 0; f < 12; ++f) acc[f] *= m;
    }
    if (out.MX) {
        const size_t n1 = n + (size_t)npoin, n2 = n + 2 * (size_t)npoin;
        out.MX[n] = acc[0];
        out.MX[n1] = acc[1];
        out.MX[n2] = acc[2];
        out.MY[n] = acc[1];
        out.MY[n1] = acc[3];
        out.MY[n2] = acc[4];
        out.MZ[n] = acc[2];
        out.MZ[n1] = acc[4];
        out.MZ[n2] = acc[5];
        out.Mass[n] = acc[6];
        out.Mass[n1] = acc[7];
        out.Mass[n2] = acc[8];
        out.Ener[n] = acc[9];
        out.Ener[n1] = acc[10];
        out.Ener[n2] = acc[11];
    }
    if (out.prj) {
        double2 *r = (double2 *)(out.prj + kProjRec * (size_t)n);
#pragma unroll
        for (int q = 0; q < 6; ++q) r[q] = make_double2(acc[2 * q], acc[2 * q + 1]);
    }
}

static int node_grid(int npoin) { return ((npoin + 255) / 256 + 7) / 8 * 8; }

static int lps_common(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
                      const double *He, const double *gpvol, const double *dl, const int *inv,
                      const int *gI, const int *gJ, const int *gK, MeshCache **m)
{
    SODHIP_REQUIRE(ctx, "null context");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_REQUIRE(nelem > 0 && npoin > 0 && connec, "invalid mesh arguments");
    SODHIP_REQUIRE(He && gpvol && dl && inv && gI && gJ && gK, "null array argument");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    return get_mesh(ctx, porder, nelem, npoin, connec, m);
}

// full_proj_ijk into the caller's projections, records through out (12 x EN)
static int proj_run(sodhip_ctx ctx, int porder, const MeshCache *m, LpsArgs A,
                    const double *invMl, double *ProjMass, double *ProjEner, double *ProjMX,
                    double *ProjMY, double *ProjMZ, double *prj = nullptr);

#define SODHIP_LPS_LAUNCH(KERNEL, GRID, ARGS)                                              \
    switch (porder) {                                                                      \
    case 2: KERNEL<2><<<GRID, EG<2>::NT, 0, ctx->stream>>>(ARGS); break;                   \
    case 3: KERNEL<3><<<GRID, EG<3>::NT, 0, ctx->stream>>>(ARGS); break;                   \
    case 4: KERNEL<4><<<GRID, EG<4>::NT, 0, ctx->stream>>>(ARGS); break;                   \
    default: return set_error(SODHIP_ERR_UNSUPPORTED, "porder %d not supported", porder); \
    }

static int proj_run(sodhip_ctx ctx, int porder, const MeshCache *m, LpsArgs A,
                    const double *invMl, double *ProjMass, double *ProjEner, double *ProjMX,
                    double *ProjMY, double *ProjMZ, double *prj)
{
    // 128 blocks per CU, several times what a CU holds at once: blocks that
    // finish early take more, which evens out the tail (16, one round, took
    // 1.57 ms per projection pass at 64^3 against 1.52 ms)
    const int grid = elem_grid(ctx, m->nelem, 128);
    int tok = prof_begin(ctx, 4);
    SODHIP_LPS_LAUNCH(k_lps_proj, grid, A)
    SODHIP_TRY_HIP(hipGetLastError());
    k_gather_proj<<<node_grid(m->npoin), 256, 0, ctx->stream>>>(
        m->npoin, m->row, m->src, A.out, m->wflag, invMl,
        ProjOut{ProjMX, ProjMY, ProjMZ, ProjMass, ProjEner, prj});
    SODHIP_TRY_HIP(hipGetLastError());
    prof_end(ctx, tok);
    return SODHIP_OK;
}

} // namespace sodhip

using namespace sodhip;

extern "C" int sodhip_comp_tau_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                   const int *connec, const double *csound, const double *u,
                                   const double *helem_k, double c_species_stab,
                                   int flag_drop_c_in_envit, double *tau)
{
    SODHIP_REQUIRE(ctx, "null context");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_REQUIRE(nelem > 0 && npoin > 0, "invalid mesh arguments");
    SODHIP_REQUIRE(connec && csound && u && helem_k && tau, "null array argument");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    const int nn = nnode_of(porder);
    MeshCache *m = nullptr;
    int rc = get_mesh(ctx, porder, nelem, npoin, connec, &m);
    if (rc) return rc;
    k_comp_tau<<<tpe_grid(nelem), 256, 0, ctx->stream>>>(
        nelem, nn, porder, connec, m->perm, csound, u, npoin, helem_k, c_species_stab,
        flag_drop_c_in_envit ? 0.0 : 1.0, tau);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

extern "C" int sodhip_full_proj_ijk_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                        int npoin_w, const int *connec, const int *lpoin_w,
                                        const double *He, const double *gpvol,
                                        const double *dlxigp_ip, const int *invAtoIJK,
                                        const int *gmshAtoI, const int *gmshAtoJ,
                                        const int *gmshAtoK, const double *rho, const double *u,
                                        const double *Tem, const double *invMl,
                                        double *ProjMass, double *ProjEner, double *ProjMX,
                                        double *ProjMY, double *ProjMZ)
{
    MeshCache *m = nullptr;
    int rc = lps_common(ctx, porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                        gmshAtoI, gmshAtoJ, gmshAtoK, &m);
    if (rc) return rc;
    SODHIP_REQUIRE(rho && u && Tem && invMl && lpoin_w && ProjMass && ProjEner && ProjMX &&
                       ProjMY && ProjMZ,
                   "null array argument");
    SODHIP_REQUIRE(npoin_w > 0 && npoin_w <= npoin, "invalid npoin_w");
    const size_t EN = (size_t)nelem * nnode_of(porder);
    if ((rc = ensure_buf(ctx->scratch, 12 * EN * sizeof(double)))) return rc;
    if ((rc = get_working(ctx, m, npoin_w, lpoin_w))) return rc;
    LpsArgs A{};
    A.nelem = nelem;
    A.npoin = npoin;
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dlxigp_ip;
    A.inv = invAtoIJK;
    A.gI = gmshAtoI;
    A.gJ = gmshAtoJ;
    A.gK = gmshAtoK;
    A.rho = rho;
    A.u = u;
    A.Tem = Tem;
    A.out = (double *)ctx->scratch.ptr;
    return proj_run(ctx, porder, m, A, invMl, ProjMass, ProjEner, ProjMX, ProjMY, ProjMZ);
}

extern "C" int sodhip_full_diff_stab_ijk_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec, const double *He,
    const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK, const int *gmshAtoI,
    const int *gmshAtoJ, const int *gmshAtoK, double Cp, double Pr, const double *rho_n,
    const double *rho, const double *u, const double *Tem, const double *mu_fluid,
    const double *mu_e, const double *mu_sgs, const double *ProjMass, const double *ProjEner,
    const double *ProjMX, const double *ProjMY, const double *ProjMZ, const double *tau_stab,
    double c_rho, double c_ener, double c_lps, double *Rmass, double *Rmom, double *Rener,
    int initialize, double fact)
{
    MeshCache *m = nullptr;
    int rc = lps_common(ctx, porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                        gmshAtoI, gmshAtoJ, gmshAtoK, &m);
    if (rc) return rc;
    SODHIP_REQUIRE(rho_n && rho && u && Tem && mu_fluid && mu_e && ProjMass && ProjEner &&
                       ProjMX && ProjMY && ProjMZ && tau_stab && Rmass && Rmom && Rener,
                   "null array argument");
    const size_t EN = (size_t)nelem * nnode_of(porder);
    if ((rc = ensure_buf(ctx->scratch, 5 * EN * sizeof(double)))) return rc;
    LpsArgs A{};
    A.nelem = nelem;
    A.npoin = npoin;
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dlxigp_ip;
    A.inv = invAtoIJK;
    A.gI = gmshAtoI;
    A.gJ = gmshAtoJ;
    A.gK = gmshAtoK;
    A.rho_n = rho_n;
    A.rho = rho;
    A.u = u;
    A.Tem = Tem;
    A.mu_fluid = mu_fluid;
    A.mu_e = mu_e;
    A.mu_sgs = mu_sgs;
    A.pMass = ProjMass;
    A.pEner = ProjEner;
    A.pMX = ProjMX;
    A.pMY = ProjMY;
    A.pMZ = ProjMZ;
    A.tau = tau_stab;
    A.Cp = Cp;
    A.Pr = Pr;
    A.c_rho = c_rho;
    A.c_ener = c_ener;
    A.c_lps = c_lps;
    A.out = (double *)ctx->scratch.ptr;
    const int grid = elem_grid(ctx, nelem, 128);
    int tok = prof_begin(ctx, 0);
    SODHIP_LPS_LAUNCH(k_diff_stab, grid, A)
    SODHIP_TRY_HIP(hipGetLastError());
    prof_end(ctx, tok);
    Rec5Out o{{Rmass, Rmom, Rmom + npoin, Rmom + 2 * (size_t)npoin, Rener}};
    tok = prof_begin(ctx, 1);
    k_gather5<<<node_grid(npoin), 256, 0, ctx->stream>>>(npoin, m->row, m->src, A.out, o,
                                                          initialize, fact);
    SODHIP_TRY_HIP(hipGetLastError());
    prof_end(ctx, tok);
    return SODHIP_OK;
}

extern "C" int sodhip_rk_stage_lps_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w, const int *connec,
    const int *lpoin_w, const double *He, const double *gpvol, const double *dlxigp_ip,
    const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
    const double *mu_e, const double *invMl, const double *tau_stab,
    const sodhip_stage_params *prm, double c_lps, double dt, double lam1, double lams,
    double gam, const double *rho1, const double *q1, const double *E1, double *rho, double *q,
    double *E, double *u, double *e_int, double *pr, double *aux_h, double *Tem,
    double *mu_fluid, double *ProjMass, double *ProjEner, double *ProjMX, double *ProjMY,
    double *ProjMZ, double *Rmass, double *Rmom, double *Rener, double gam_next,
    double *rho_next, double *q_next, double *E_next)
{
    SODHIP_REQUIRE(prm, "null argument");
    MeshCache *m = nullptr;
    int rc = lps_common(ctx, porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                        gmshAtoI, gmshAtoJ, gmshAtoK, &m);
    if (rc) return rc;
    SODHIP_REQUIRE(lpoin_w && npoin_w > 0 && npoin_w <= npoin, "invalid working nodes");
    SODHIP_REQUIRE(mu_e && invMl && tau_stab && rho1 && q1 && E1 && rho && q && E && u &&
                       e_int && pr && aux_h && Tem && mu_fluid,
                   "null array argument");
    SODHIP_REQUIRE(!!ProjMass == !!ProjEner && !!ProjEner == !!ProjMX && !!ProjMX == !!ProjMY &&
                       !!ProjMY == !!ProjMZ,
                   "ProjMass/ProjEner/ProjMX/ProjMY/ProjMZ: all or none");
    SODHIP_REQUIRE(!!Rmass == !!Rmom && !!Rmom == !!Rener, "Rmass/Rmom/Rener: all or none");
    SODHIP_REQUIRE(!!rho_next == !!q_next && !!q_next == !!E_next, "next stage: all or none");
    SODHIP_REQUIRE(!rho_next || (rho_next != rho && q_next != q && E_next != E),
                   "the next stage must not overwrite the new state");
    if ((rc = get_working(ctx, m, npoin_w, lpoin_w))) return rc;
    // updateF's EOS block and Sutherland's law (time_integ_ls.f90:572-599)
    if ((rc = eos_stage_sutherland(ctx, m, prm->gamma_gas, prm->Rgas, prm->mu_factor, rho, q, E, u,
                                   e_int, pr, aux_h, Tem, mu_fluid)))
        return rc;
    const size_t EN = (size_t)nelem * nnode_of(porder);
    // the projection records start on a 256-byte boundary (their gathers read
    // them with 16-byte loads, and 5*EN doubles is odd when EN is)
    const size_t prec_at = (5 * EN + 31) & ~(size_t)31;
    // then the node projection records, on a 256-byte boundary
    const size_t prj_at = (prec_at + 12 * EN + 31) & ~(size_t)31;
    if ((rc = ensure_buf(ctx->scratch, (prj_at + kProjRec * (size_t)npoin) * sizeof(double))))
        return rc;
    double *rec = (double *)ctx->scratch.ptr, *prec = rec + prec_at, *prj = rec + prj_at;
    // at p=3 the convection and the stabilised viscous operator run in one
    // two-wave element kernel after the projections; otherwise the
    // convection records (elem_convec.f90:227-470, E slot = enthalpy) come
    // first and the stabilised operator adds to them
    const bool split = porder == 3;
    if (!split && (rc = conv_elements<double>(ctx, porder, m, He, gpvol, dlxigp_ip, invAtoIJK,
                                              gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, aux_h,
                                              rec)))
        return rc;
    LpsArgs A{};
    A.nelem = nelem;
    A.npoin = npoin;
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dlxigp_ip;
    A.inv = invAtoIJK;
    A.gI = gmshAtoI;
    A.gJ = gmshAtoJ;
    A.gK = gmshAtoK;
    A.rho_n = rho;
    A.rho = rho;
    A.u = u;
    A.Tem = Tem;
    A.mu_fluid = mu_fluid;
    A.mu_e = mu_e;
    A.mu_sgs = nullptr;
    A.prj = prj;
    A.tau = tau_stab;
    A.Cp = prm->Cp;
    A.Pr = prm->Pr;
    A.c_rho = prm->c_rho;
    A.c_ener = prm->c_ener;
    A.c_lps = c_lps;
    // projections (elem_stab.f90:441-586)
    A.out = prec;
    if ((rc = proj_run(ctx, porder, m, A, invMl, ProjMass, ProjEner, ProjMX, ProjMY, ProjMZ,
                       prj)))
        return rc;
    // stabilised viscous operator added to the convection records: both carry
    // the same fact (-1), applied once by the gather
    if (split) {
        if ((rc = rhs_lps_elements(ctx, m, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ,
                                   gmshAtoK, rho, q, E, mu_e, prj, tau_stab, prm->gamma_gas,
                                   prm->Rgas, prm->Cp, prm->Pr, prm->mu_factor, prm->c_rho,
                                   prm->c_ener, c_lps, rec)))
            return rc;
    } else {
        A.out = rec;
        const int grid = elem_grid(ctx, nelem, 128);
        int tok = prof_begin(ctx, 0);
        switch (porder) {
        case 2: k_diff_stab<2, true, true><<<grid, EG<2>::NT, 0, ctx->stream>>>(A); break;
        case 4: k_diff_stab<4, true, true><<<grid, EG<4>::NT, 0, ctx->stream>>>(A); break;
        }
        SODHIP_TRY_HIP(hipGetLastError());
        prof_end(ctx, tok);
    }
    const NextStage nx{1.0, gam_next * dt, rho_next, q_next, E_next};
    return gather_update<double>(ctx, m, npoin, npoin_w, lpoin_w, rec, 5, EN, invMl, prm->fact, dt,
                                 lam1, lams, gam, rho1, q1, E1, rho, q, E, Rmass, Rmom, Rener,
                                 rho_next ? &nx : nullptr);
}

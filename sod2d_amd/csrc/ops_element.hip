// ops_element.hip -- element kernels of the split-form convection, the viscous
// operator and the generic scalar convection, the node-ordered assembly gather,
// and their drop-in C-ABI entry points.
//
// Layout and mapping (gfx950):
//  * one element per workgroup of NT = roundup((p+1)^3, 64) lanes; lane a owns
//    element-local node / quadrature point a (the reference's gmsh order);
//  * the element's nodal fields are staged field-major in LDS; each lane builds
//    once, in registers, the LDS offsets of the p other nodes on its three GLL
//    lines and the matching rows of the derivative table (and of its transpose
//    for the weak divergence of the viscous fluxes);
//  * blocks walk the elements in a grid-stride loop whose block -> element map
//    gives each XCD (blockIdx % 8) one contiguous eighth of the elements, so
//    nodes shared by neighbouring elements are re-read from the same L2;
//  * element residuals go to a scratch array of 5-field records per element-node
//    (whole elements: coalesced writes); a second kernel gathers them per node through
//    the node-ordered incidence table in the reference's serial (element, node)
//    order -- deterministic, no atomics.
//
// Arithmetic restatement: the split forms are evaluated through identities that
// hold for arbitrary nodal inputs (linearity of the discrete derivative and
// reordering of finite sums), which remove redundant work of the reference loop:
//  * the enthalpy-slot and kinetic-energy terms of the energy equation share one
//    derivative of H = E + k and of rho*H;
//  * advective derivatives a.grad(f) use contravariant components
//    sum_j He(j,k) a_j at the quadrature point;
//  * flux divergences sum_j d_j(s u_j) are evaluated from the line values of s
//    times the contravariant velocity sum_j He_g(j,k) u_j(m) at the line node m.
#include <cstdio>
#include <type_traits>

#include "common.hpp"
#include "device.hpp"
#include "element.hpp"

#ifndef SODHIP_SPLIT_WAVES
#define SODHIP_SPLIT_WAVES 2
#endif
#ifndef SODHIP_SPLIT_RHS
#define SODHIP_SPLIT_RHS 1
#endif
#ifndef SODHIP_SPLIT_WAVES_F32
#define SODHIP_SPLIT_WAVES_F32 2
#endif
// ring depth (elements in flight per block) in FP64 and FP32 operators: in
// FP64 two keep four blocks per CU within 160 KB of LDS
// consecutive elements a block evaluates in a row (runs of one block, XCD-
// aware over runs): they share the 128-byte lines and pages of the element-
// fastest mu_e(nelem,ngaus)
#ifndef SODHIP_SPLIT_RUN
#define SODHIP_SPLIT_RUN 4
#endif
#ifndef SODHIP_SPLIT_AHEAD64
#define SODHIP_SPLIT_AHEAD64 2
#endif
#ifndef SODHIP_SPLIT_AHEAD32
#define SODHIP_SPLIT_AHEAD32 3
#endif

namespace sodhip {

// Copies the geometry block of one element (HEB bytes of He, then GVB bytes of
// gpvol) HBM -> LDS with the direct-to-LDS loads of gfx950: one wave
// instruction moves a row of 64 lanes x W bytes; the NW waves of the
// workgroup take alternate rows.  Lanes past the end re-read the last W bytes.
template <int W, int HEB, int GVB, int NW>
__device__ __forceinline__ void geo_copy(const unsigned char *he, const unsigned char *gv,
                                         unsigned char *geo, int wl, int wv)
{
    constexpr int R = 64 * W, ROWS = (HEB + GVB + R - 1) / R;
#pragma unroll
    for (int r0 = 0; r0 < ROWS; r0 += NW) {
        const int r = r0 + wv;
        if (r < ROWS) {
            const int byte = r * R + wl * W;
            const unsigned char *src = byte < HEB ? he + byte : gv + min(byte - HEB, GVB - W);
            // offset applied in the LDS address space (a generic round trip of a
            // constant LDS address miscompiles)
            auto *dst = (__attribute__((address_space(3))) unsigned char *)geo + r * R;
            if constexpr (W == 16)
                __builtin_amdgcn_global_load_lds((const void *)src, dst, 16, 0, 0);
            else
                __builtin_amdgcn_global_load_lds((const void *)src, dst, 4, 0, 0);
        }
    }
}

template <typename T> struct RhsArgs {
    int nelem, npoin;
    size_t EN;
    const int *connecT;
    const T *He, *gpvol, *dl;
    const int *inv, *gI, *gJ, *gK;
    // primitive inputs (drop-in operators)
    const T *u, *q, *rho, *pr, *E;
    const T *rho_n, *Tem, *mu_fluid, *mu_e, *mu_sgs;
    T Cp, Pr, c_rho, c_ener;
    // conserved inputs (fused stage): FP64 state
    const double *cr, *cq, *cE;
    double gamma, Rgas, mu_factor, p_ref;
    // per-element Sutherland factor (an ensemble of meshes whose instances
    // differ in it), or nullptr for the uniform mu_factor
    const double *mu_factor_e;
    const int *perm; // element-local scratch position of each node
    T *out;          // records of 5 (mass, mom1, mom2, mom3, energy) per element-node;
                     // (e,a) at record e*NN + perm[a]
    int geo16;       // He and gpvol blocks of every element are 16-byte aligned
    // brick-local assembly of the fused p=3 stage (see ClusterTables): the
    // element order, local-node tables, the partial records of surface nodes,
    // and the stage update the element kernel applies to interior nodes
    // (the arithmetic of k_gather_update)
    const int *cel, *cloc, *cnb, *cnode, *ckind;
    int ncl;
    double *rec;
    const double *invMl;
    double fact, lam1, lams, gdt;
    int rd1, rdy;
    const double *y1r, *y1q, *y1E;
    double *yr, *yq, *yE;
    double *Rm, *Rq, *Re;
    double nx_l1, nx_gdt;
    double *nxr, *nxq, *nxE;
    // LPS stage (the viscous wave evaluates full_diff_stab_ijk): node
    // projection records (kProjRec values), comp_tau's tau_stab, c_lps
    const double *prj, *tau;
    double c_lps;
};

// Sutherland's law (mod_fluid_viscosity.f90:21-41)
template <typename T> __device__ __forceinline__ T sutherland(T mu_factor, T Tk)
{
    return mu_factor * T(0.000001458) * (Tk * sqrt(Tk)) / (Tk + T(110.4));
}

// CONV: split-form convection (full_convec_ijk); DIFF: viscous operator
// (full_diffusion_ijk); CONS: inputs are the conserved FP64 state (the fused
// stage forms the primitives of updateF's EOS block and Sutherland's viscosity
// in-kernel) instead of the primitive arrays.
//
// Per element (one lane per node / quadrature point), in phases that keep the
// live register set small (values cross phases through LDS):
//   stage   nodal fields -> LDS;
//   conv    velocity lines: d/dxi u_i, the contravariant line weights
//           W_k(m) = D_k(m) sum_j He_g(j,k) u_j(m), div(u_i u), div u; then the
//           density, pressure, H, rho*H and momentum lines against W;
//   diff 1  density, temperature and velocity lines -> grad rho, grad T, grad u
//           -> viscous fluxes mapped to contravariant components
//           G_k(F) = gpvol sum_d He(d,k) F_d, staged in LDS;
//   diff 2  weak divergence sum_k sum_{b on k-line} dl(b,k,I_k(a)) G_k(F)(b)
//           with the transposed coefficients kept per lane in LDS;
//   out     residuals written to the element's scratch block.
template <int P, typename T, bool CONV, bool DIFF, bool CONS>
__global__ __launch_bounds__(EG<P>::NT, 2) void k_rhs(RhsArgs<T> A)
{
    constexpr int N1 = EG<P>::N1, NN = EG<P>::NN, NT = EG<P>::NT;
    constexpr int F_RHO = 0, F_U = 1, F_T = 4;
    constexpr int F_C = DIFF ? 5 : 4; // first convection-only field
    constexpr int F_P = F_C, F_H = F_C + 1, F_RH = F_C + 2, F_Q = F_C + 3;
    constexpr int NF1 = CONV ? F_C + 6 : F_C;
    constexpr int NF2 = DIFF ? 18 : 0;
    // nodal fields (s1) and contravariant viscous fluxes (s2): the fluxes are
    // staged while the convective lines still read the nodal fields
    __shared__ T s1[NF1 * NT];
    __shared__ T s2[(NF2 > 0 ? NF2 : 1) * NT];

    const int lane = threadIdx.x;
    const bool act = lane < NN;
    const int a = act ? lane : 0;
    LineTab<P, T> lt;
    build_linetab<P, T>(a, A.dl, A.inv, A.gI, A.gJ, A.gK, lt);
    const int pa = min(max(A.perm[a], 0), NN - 1); // scratch position within the element
    // line derivative d/dxi_k of LDS field f at the lane's point
    auto dxi = [&](int f, int k, T own) {
        T v = lt.ds[k] * own;
#pragma unroll
        for (int o = 0; o < P; ++o) v += lt.d[k][o] * s1[f * NT + lt.off[k][o]];
        return v;
    };

    // Inputs of one element at the lane's node; the next element's are loaded
    // while the current one is computed (software pipelining of the global
    // loads, whose latency would otherwise be exposed at low occupancy).
    struct In {
        double cr, cq[3], cE;                // conserved state (CONS)
        T rho, u[3], q[3], pr, E, Tem, rhon, muf; // primitive inputs (drop-in)
        T mue, musgs;
    };
    auto load = [&](In &x, int e, int n) {
        const size_t ea = (size_t)e * NN + a;
        if constexpr (CONS) {
            x.cr = A.cr[n];
            x.cq[0] = A.cq[n];
            x.cq[1] = A.cq[n + A.npoin];
            x.cq[2] = A.cq[n + 2 * (size_t)A.npoin];
            x.cE = A.cE[n];
        } else {
            x.rho = A.rho[n];
#pragma unroll
            for (int d = 0; d < 3; ++d) x.u[d] = A.u[n + (size_t)A.npoin * d];
            if constexpr (CONV) {
#pragma unroll
                for (int d = 0; d < 3; ++d) x.q[d] = A.q[n + (size_t)A.npoin * d];
                x.pr = A.pr[n];
                x.E = A.E[n];
            }
            if constexpr (DIFF) {
                x.Tem = A.Tem[n];
                x.rhon = A.rho_n[n];
                x.muf = A.mu_fluid[n];
                x.musgs = A.mu_sgs ? A.mu_sgs[e + (size_t)A.nelem * a] : T(0);
            }
        }
        if constexpr (DIFF) x.mue = A.mu_e[e + (size_t)A.nelem * a];
    };

    // Geometry of one element (He, then gpvol) is copied HBM -> LDS by the
    // direct-to-LDS loads of gfx950 (no registers involved), one element ahead:
    // the copy of element e+stride runs during the arithmetic of element e.  A
    // wave instruction copies one row of 64 lanes x W bytes (W = 16 when the
    // per-element blocks are 16-byte aligned, else 4); the waves of the
    // workgroup take alternate rows.
    constexpr int HE_BYTES = 9 * NN * (int)sizeof(T), GV_BYTES = NN * (int)sizeof(T);
    constexpr int NW = NT / 64;
    constexpr int ROWS16 = (HE_BYTES + GV_BYTES + 1023) / 1024;
    constexpr int ROWS4 = (HE_BYTES + GV_BYTES + 255) / 256;
    __shared__ __attribute__((aligned(16))) unsigned char
        geo[(ROWS16 * 1024 > ROWS4 * 256 ? ROWS16 * 1024 : ROWS4 * 256)];
    const int wl = lane & 63, wv = lane >> 6;
    auto geo_issue = [&](int e) {
        const unsigned char *he = (const unsigned char *)(A.He + (size_t)e * 9 * NN);
        const unsigned char *gv = (const unsigned char *)(A.gpvol + (size_t)e * NN);
        if (A.geo16)
            geo_copy<16, HE_BYTES, GV_BYTES, NW>(he, gv, geo, wl, wv);
        else
            geo_copy<4, HE_BYTES, GV_BYTES, NW>(he, gv, geo, wl, wv);
    };
    auto esync = [&]() { lds_barrier(); };
    constexpr int WAIT_VM0 = kWaitVm0, WAIT_LGKM0 = kWaitLgkm0;

    const ElemRange er = elem_range(A.nelem);
    if (er.begin >= er.end) return;
    // element indices past the block's range are clamped to its last element:
    // the extra loads are harmless and keep the load stream unconditional
    const int elast = er.begin + (er.end - 1 - er.begin) / er.stride * er.stride;
    In cur, nxt;
    geo_issue(er.begin);
    load(cur, er.begin, A.connecT[(size_t)er.begin * NN + a]);
    int n_nx = A.connecT[(size_t)min(er.begin + er.stride, elast) * NN + a];
    for (int e = er.begin; e < er.end; e += er.stride) {
        const int e2 = min(e + er.stride, elast), e3 = min(e + 2 * er.stride, elast);
        // geometry of element e: wait for its copy (and the state loads issued
        // with it), move it to registers, then start the copies of element e2
        __builtin_amdgcn_s_waitcnt(WAIT_VM0);
        esync();
        T He[3][3];
        {
            const T *hp = (const T *)geo + 9 * a;
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i) He[i][j] = hp[i + 3 * j];
        }
        const T vol = ((const T *)geo)[9 * NN + a];
        __builtin_amdgcn_s_waitcnt(WAIT_LGKM0);
        esync();
        geo_issue(e2);
        load(nxt, e2, n_nx);
        const int n_nx2 = A.connecT[(size_t)e3 * NN + a];
        T mue = T(0), musgs = T(0), muf = T(0), rhon = T(0);
        if constexpr (DIFF) mue = cur.mue;
        // ---------------- stage own node values
        {
            T rho, u[3], Tk = T(0), p = T(0), H = T(0), rhoH = T(0), q[3] = {T(0), T(0), T(0)};
            if constexpr (CONS) {
                const double r = cur.cr;
                const double ri = 1.0 / r;
                double ud[3], umag = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    ud[d] = cur.cq[d] * ri;
                    umag += ud[d] * ud[d];
                }
                const double eint = cur.cE * ri - 0.5 * umag;
                const double pd = r * (A.gamma - 1.0) * eint;
                const double hd = (A.gamma / (A.gamma - 1.0)) * pd * ri;
                const double Td = pd * ri / A.Rgas;
                rho = T(r);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    u[d] = T(ud[d]);
                    q[d] = T(cur.cq[d]);
                }
                // only grad p enters: a uniform reference pressure is removed
                // before the (possibly FP32) operator arithmetic
                p = T(pd - A.p_ref);
                H = T(hd + 0.5 * umag);
                rhoH = T(r * (hd + 0.5 * umag));
                Tk = T(Td);
                rhon = rho;
                muf = T(sutherland<double>(A.mu_factor_e ? A.mu_factor_e[e] : A.mu_factor, Td));
            } else {
                rho = cur.rho;
#pragma unroll
                for (int d = 0; d < 3; ++d) u[d] = cur.u[d];
                if constexpr (CONV) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) q[d] = cur.q[d];
                    p = cur.pr;
                    const T kin = u[0] * u[0] * T(0.5) + u[1] * u[1] * T(0.5) + u[2] * u[2] * T(0.5);
                    H = cur.E + kin;
                    rhoH = rho * H;
                }
                if constexpr (DIFF) {
                    Tk = cur.Tem;
                    rhon = cur.rhon;
                    muf = cur.muf;
                    musgs = cur.musgs;
                }
            }
            if (act) {
                s1[F_RHO * NT + a] = rho;
#pragma unroll
                for (int d = 0; d < 3; ++d) s1[(F_U + d) * NT + a] = u[d];
                if constexpr (DIFF) s1[F_T * NT + a] = Tk;
                if constexpr (CONV) {
                    s1[F_P * NT + a] = p;
                    s1[F_H * NT + a] = H;
                    s1[F_RH * NT + a] = rhoH;
#pragma unroll
                    for (int d = 0; d < 3; ++d) s1[(F_Q + d) * NT + a] = q[d];
                }
            }
        }
        esync();

        T Rm = 0, Rq[3] = {0, 0, 0}, Re = 0;
        T gR[3] = {0, 0, 0};                           // DIFF: grad rho
        T gU[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}; // DIFF: gU[i][j] = d u_i / d x_j
        const T rho = s1[F_RHO * NT + a];
        T u[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) u[i] = s1[(F_U + i) * NT + a];
        // ---------------- velocity lines: d/dxi u_i (-> grad u) and, for the
        // convection, the contravariant line weights W and div(u_i u), div u
        T Wo[3], Wx[3][P];
        T divUw = 0;
        {
            T q[3] = {0, 0, 0}, Qt[3] = {0, 0, 0}, Ut[3] = {0, 0, 0}, divFuu[3] = {0, 0, 0};
            if constexpr (CONV) {
#pragma unroll
                for (int i = 0; i < 3; ++i) q[i] = s1[(F_Q + i) * NT + a];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    Ut[k] = He[0][k] * u[0] + He[1][k] * u[1] + He[2][k] * u[2];
                    Qt[k] = He[0][k] * q[0] + He[1][k] * q[1] + He[2][k] * q[2];
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T c0 = lt.ds[k];
                T du[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) du[i] = c0 * u[i];
                if constexpr (CONV) {
                    Wo[k] = c0 * Ut[k];
                    divUw += Wo[k];
#pragma unroll
                    for (int i = 0; i < 3; ++i) divFuu[i] += Wo[k] * u[i];
                }
#pragma unroll
                for (int o = 0; o < P; ++o) {
                    const int b = lt.off[k][o];
                    const T c = lt.d[k][o];
                    T ub[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        ub[i] = s1[(F_U + i) * NT + b];
                        du[i] += c * ub[i];
                    }
                    if constexpr (CONV) {
                        const T w = c * (He[0][k] * ub[0] + He[1][k] * ub[1] + He[2][k] * ub[2]);
                        Wx[k][o] = w;
                        divUw += w;
#pragma unroll
                        for (int i = 0; i < 3; ++i) divFuu[i] += w * ub[i];
                    }
                }
                if constexpr (CONV) {
#pragma unroll
                    for (int i = 0; i < 3; ++i) Rq[i] += Qt[k] * du[i];
                }
                if constexpr (DIFF) {
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) gU[i][j] += He[j][k] * du[i];
                }
            }
            if constexpr (CONV) {
#pragma unroll
                for (int i = 0; i < 3; ++i) Rq[i] = T(0.25) * (Rq[i] + rho * divFuu[i]);
            }
        }
        // ---------------- density line: grad rho (viscous) and u.grad rho terms
        // (convection)
        {
            const T *sf = s1 + F_RHO * NT;
            T d[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                T v = lt.ds[k] * rho;
#pragma unroll
                for (int o = 0; o < P; ++o) v += lt.d[k][o] * sf[lt.off[k][o]];
                d[k] = v;
            }
            if constexpr (DIFF) {
#pragma unroll
                for (int i = 0; i < 3; ++i) gR[i] = He[i][0] * d[0] + He[i][1] * d[1] + He[i][2] * d[2];
            }
            if constexpr (CONV) {
                T Ut[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) Ut[k] = He[0][k] * u[0] + He[1][k] * u[1] + He[2][k] * u[2];
                const T urho = Ut[0] * d[0] + Ut[1] * d[1] + Ut[2] * d[2];
                Rm += T(0.5) * urho;
#pragma unroll
                for (int i = 0; i < 3; ++i) Rq[i] += T(0.25) * u[i] * urho;
                Re += T(0.25) * s1[F_H * NT + a] * urho;
            }
        }

        // ---------------- viscous fluxes, staged as contravariant components
        // G_k(F) = gpvol sum_d He(d,k) F_d for the weak divergence of pass 2
        if constexpr (DIFF) {
            const T mufg = muf + rhon * musgs;
            const T mut = mufg + mue;
            const T divU = gU[0][0] + gU[1][1] + gU[2][2];
            const T tt = mufg * T(2.0 / 3.0) * divU;
            T tau[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) tau[i][j] = mut * (gU[i][j] + gU[j][i]);
#pragma unroll
            for (int i = 0; i < 3; ++i) tau[i][i] -= tt;
            T dT[3];
            const T Tk = s1[F_T * NT + a];
#pragma unroll
            for (int k = 0; k < 3; ++k) dT[k] = dxi(F_T, k, Tk);
            T gt[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) gt[i] = He[i][0] * dT[0] + He[i][1] * dT[1] + He[i][2] * dT[2];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T h0 = vol * He[0][k], h1 = vol * He[1][k], h2 = vol * He[2][k];
                T tuk = 0;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const T g = h0 * tau[r][0] + h1 * tau[r][1] + h2 * tau[r][2];
                    if (act) s2[(9 + 3 * r + k) * NT + a] = g;
                    tuk += g * u[r]; // tau is symmetric: contravariant (tau u)_k
                }
                if (act) {
                    s2[(6 + k) * NT + a] = tuk;
                    s2[(0 + k) * NT + a] = h0 * gR[0] + h1 * gR[1] + h2 * gR[2];
                    s2[(3 + k) * NT + a] = h0 * gt[0] + h1 * gt[1] + h2 * gt[2];
                }
            }
        }

        // ---------------- convective scalar and momentum lines, one field per
        // (rolled) iteration: line derivative d and the W-weighted line sum
        // `div` = div(s u) folded at once into the residuals
        T divQ = 0;
        if constexpr (CONV) {
#pragma unroll 1
            for (int fi = 1; fi < 7; ++fi) { // p, H, rhoH, q1, q2, q3
                const int f = fi == 1 ? F_P : (fi == 2 ? F_H : (fi == 3 ? F_RH : F_Q + fi - 4));
                const T *sf = s1 + f * NT;
                const T own = sf[a];
                T d[3], div = 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    T v = lt.ds[k] * own, w = Wo[k] * own;
#pragma unroll
                    for (int o = 0; o < P; ++o) {
                        const T x = sf[lt.off[k][o]];
                        v += lt.d[k][o] * x;
                        w += Wx[k][o] * x;
                    }
                    d[k] = v;
                    div += w;
                }
                if (fi == 1) {
#pragma unroll
                    for (int i = 0; i < 3; ++i) Rq[i] += He[i][0] * d[0] + He[i][1] * d[1] + He[i][2] * d[2];
                } else if (fi == 2) { // H: q.grad H + rho div(H u)
                    T qt = 0;
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        qt += (He[0][k] * s1[F_Q * NT + a] + He[1][k] * s1[(F_Q + 1) * NT + a] +
                               He[2][k] * s1[(F_Q + 2) * NT + a]) * d[k];
                    Re += T(0.25) * (qt + rho * div);
                } else { // rho H, q_i: u.grad s + div(s u)
                    T ut = 0;
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        ut += (He[0][k] * u[0] + He[1][k] * u[1] + He[2][k] * u[2]) * d[k];
                    const T c = T(0.25) * (ut + div);
                    if (fi == 3) {
                        Re += c;
                    } else {
                        const int i = fi - 4;
                        const T dq = He[i][0] * d[0] + He[i][1] * d[1] + He[i][2] * d[2];
                        divQ += dq;
                        if (i == 0) Rq[0] += c;
                        else if (i == 1) Rq[1] += c;
                        else Rq[2] += c;
                    }
                }
            }
        }
        if constexpr (CONV) {
            const T H = s1[F_H * NT + a], rhoH = s1[F_RH * NT + a];
            Rm += T(0.5) * (divQ + rho * divUw);
#pragma unroll
            for (int i = 0; i < 3; ++i)
                Rq[i] += T(0.25) * (u[i] * divQ + s1[(F_Q + i) * NT + a] * divUw);
            Re += T(0.25) * (H * divQ + rhoH * divUw);
            Rm *= vol;
            Re *= vol;
#pragma unroll
            for (int i = 0; i < 3; ++i) Rq[i] *= vol;
        }

        // ---------------- viscous operator, pass 2
        if constexpr (DIFF) {
            esync(); // every lane's fluxes are in s2
            // pass 2: weak divergence with the transposed coefficients
            // dl(b,k,I_k(a)) of the lane's lines
            T W6[6];
#pragma unroll 1
            for (int fl = 0; fl < 6; ++fl) {
                T acc = 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const T *sf = s2 + (3 * fl + k) * NT;
                    T v = lt.ds[k] * sf[a];
#pragma unroll
                    for (int o = 0; o < P; ++o) v += lt.t[k][o] * sf[lt.off[k][o]];
                    acc += v;
                }
                if (fl == 0) W6[0] = acc;
                else if (fl == 1) W6[1] = acc;
                else if (fl == 2) W6[2] = acc;
                else if (fl == 3) W6[3] = acc;
                else if (fl == 4) W6[4] = acc;
                else W6[5] = acc;
            }
            const T nu_e = A.c_rho * mue / rhon;
            const T kappa = muf * A.Cp / A.Pr + A.c_ener * mue * A.Cp / A.Pr +
                            A.Cp * rhon * musgs / T(0.9);
            Rm += nu_e * W6[0];
            Re += W6[2] + kappa * W6[1];
#pragma unroll
            for (int i = 0; i < 3; ++i) Rq[i] += W6[3 + i];
        }

        if (act) {
            T *o = A.out + 5 * ((size_t)e * NN + pa);
            o[0] = Rm;
            o[1] = Rq[0];
            o[2] = Rq[1];
            o[3] = Rq[2];
            o[4] = Re;
        }
        cur = nxt;
        n_nx = n_nx2;
        esync();
    }
    // no copy may still be writing LDS when the workgroup ends
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);
}

// generic_scalar_convec_ijk (elem_convec.f90:472-594):
// Re = gpvol [ 1/2 (div f + eta div u) + 1/2 u.grad(eta) ]
// FUSEF: the flux is f = u * eta, formed in registers instead of read (the
// time loop's f_eta = u(:,:,1) * eta(:,1) of the same two arrays)
template <int P, typename T, bool FUSEF>
__global__ __launch_bounds__(EG<P>::NT) void k_scalar_convec(RhsArgs<T> A, const T *__restrict__ f,
                                                             const T *__restrict__ eta)
{
    constexpr int NN = EG<P>::NN, NT = EG<P>::NT;
    __shared__ T s[7 * NT]; // eta, u(3), f(3)
    const int lane = threadIdx.x;
    const bool act = lane < NN;
    const int a = act ? lane : 0;
    LineTab<P, T> lt;
    build_linetab<P, T>(a, A.dl, A.inv, A.gI, A.gJ, A.gK, lt);
    const int pa = min(max(A.perm[a], 0), NN - 1);
    struct In {
        T et, u[3], fv[3], He[9], vol;
    };
    auto load = [&](In &x, int e) {
        const size_t ea = (size_t)e * NN + a;
        const int n = act ? A.connecT[ea] : 0;
        x.et = eta[n];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            x.u[d] = A.u[n + (size_t)A.npoin * d];
            x.fv[d] = FUSEF ? x.u[d] * x.et : f[n + (size_t)A.npoin * d];
        }
        const T *hp = A.He + 9 * ea;
#pragma unroll
        for (int i = 0; i < 9; ++i) x.He[i] = hp[i];
        x.vol = A.gpvol[ea];
    };
    const ElemRange er = elem_range(A.nelem);
    if (er.begin >= er.end) return;
    const int elast = er.begin + (er.end - 1 - er.begin) / er.stride * er.stride;
    In cur, nxt;
    load(cur, er.begin);
    for (int e = er.begin; e < er.end; e += er.stride) {
        load(nxt, min(e + er.stride, elast));
        if (act) {
            s[a] = cur.et;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                s[(1 + d) * NT + a] = cur.u[d];
                s[(4 + d) * NT + a] = cur.fv[d];
            }
        }
        lds_barrier();
        T gE[3] = {0, 0, 0}, divU = 0, divF = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const T c0 = lt.ds[k];
            T de = c0 * cur.et, du[3], df[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                du[d] = c0 * cur.u[d];
                df[d] = c0 * cur.fv[d];
            }
#pragma unroll
            for (int o = 0; o < P; ++o) {
                const int b = lt.off[k][o];
                const T c = lt.d[k][o];
                de += c * s[b];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    du[d] += c * s[(1 + d) * NT + b];
                    df[d] += c * s[(4 + d) * NT + b];
                }
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const T h = cur.He[i + 3 * k]; // He(i,k)
                gE[i] += h * de;
                divU += h * du[i];
                divF += h * df[i];
            }
        }
        T r = T(0.5) * (divF + cur.et * divU);
#pragma unroll
        for (int i = 0; i < 3; ++i) r += T(0.5) * (cur.u[i] * gE[i]);
        if (act) A.out[(size_t)e * NN + pa] = cur.vol * r;
        cur = nxt;
        lds_barrier();
    }
}

// eval_divergence (mod_operators.f90:485-570) and compute_vorticity
// (mod_operators.f90:684-775) at each GLL point.  With g_ik = du_i/dxi_k the
// line derivatives and grad u(i,j) = sum_k He(j,k) g_ik: the divergence record
// is gpvol * div u, the vorticity record eps_ijk du_j/dx_k (the reference's
// leviCivi sum, minus the curl).  Records of DIV + 3 CURL values per
// element-node at e*NN + perm[a].
template <int P, typename T, bool DIV, bool CURL>
__global__ __launch_bounds__(EG<P>::NT) void k_div_curl(RhsArgs<T> A)
{
    constexpr int NN = EG<P>::NN, NT = EG<P>::NT, NF = (DIV ? 1 : 0) + (CURL ? 3 : 0);
    __shared__ T s[3 * NT]; // u(3)
    const int lane = threadIdx.x;
    const bool act = lane < NN;
    const int a = act ? lane : 0;
    LineTab<P, T> lt;
    build_linetab<P, T>(a, A.dl, A.inv, A.gI, A.gJ, A.gK, lt);
    const int pa = min(max(A.perm[a], 0), NN - 1);
    struct In {
        T u[3], He[9], vol;
    };
    auto load = [&](In &x, int e) {
        const size_t ea = (size_t)e * NN + a;
        const int n = act ? A.connecT[ea] : 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) x.u[d] = A.u[n + (size_t)A.npoin * d];
        const T *hp = A.He + 9 * ea;
#pragma unroll
        for (int i = 0; i < 9; ++i) x.He[i] = hp[i];
        x.vol = DIV ? A.gpvol[ea] : T(0);
    };
    const ElemRange er = elem_range(A.nelem);
    if (er.begin >= er.end) return;
    const int elast = er.begin + (er.end - 1 - er.begin) / er.stride * er.stride;
    In cur, nxt;
    load(cur, er.begin);
    for (int e = er.begin; e < er.end; e += er.stride) {
        load(nxt, min(e + er.stride, elast));
        if (act) {
#pragma unroll
            for (int d = 0; d < 3; ++d) s[d * NT + a] = cur.u[d];
        }
        lds_barrier();
        T g[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const T c0 = lt.ds[k];
            T du[3] = {c0 * cur.u[0], c0 * cur.u[1], c0 * cur.u[2]};
#pragma unroll
            for (int o = 0; o < P; ++o) {
                const int b = lt.off[k][o];
                const T c = lt.d[k][o];
#pragma unroll
                for (int d = 0; d < 3; ++d) du[d] += c * s[d * NT + b];
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) g[d][k] = du[d];
        }
        T gu[3][3]; // gu[i][j] = du_i/dx_j
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                gu[i][j] = cur.He[j] * g[i][0] + cur.He[j + 3] * g[i][1] + cur.He[j + 6] * g[i][2];
        if (act) {
            T *o = A.out + NF * ((size_t)e * NN + pa);
            int f = 0;
            if constexpr (DIV) o[f++] = cur.vol * (gu[0][0] + gu[1][1] + gu[2][2]);
            if constexpr (CURL) {
                o[f++] = gu[1][2] - gu[2][1];
                o[f++] = gu[2][0] - gu[0][2];
                o[f++] = gu[0][1] - gu[1][0];
            }
        }
        cur = nxt;
        lds_barrier();
    }
}

// curlU(n,:) /= lelpn(n) over the working nodes (compute_vorticity's makeConti)
template <typename T>
__global__ __launch_bounds__(256) void k_per_elem_count(int npoin, int npoin_w,
                                                        const int *__restrict__ lpoin_w,
                                                        const int *__restrict__ lelpn,
                                                        T *__restrict__ curl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const T c = T(lelpn[n]);
#pragma unroll
    for (int d = 0; d < 3; ++d) curl[n + (size_t)npoin * d] /= c;
}

// Node-ordered assembly: R_f(n) = (init ? 0 : R_f(n)) + sum_j fact * scr[src[j]*NF + f], j
// over the node's slots row[n]..row[n+1]-1 in ascending (element, node) order
// (records of NF fields per element-node).  Slots are taken four at a time
// with every record load issued before the additions, which stay in slot
// order; blocks are mapped XCD-aware (the nodes of one XCD are one contiguous
// range, whose records share that XCD's L2).
template <typename TS, typename TR, int NF>
__global__ __launch_bounds__(256) void k_gather(int npoin, const int *__restrict__ row,
                                                const int *__restrict__ src,
                                                const TS *__restrict__ scr, TR *__restrict__ R0,
                                                TR *__restrict__ R1, TR *__restrict__ R2,
                                                TR *__restrict__ R3, TR *__restrict__ R4, int init,
                                                TR fact)
{
    const int per = gridDim.x >> 3; // grid is a multiple of 8
    const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int n = blk * blockDim.x + threadIdx.x;
    if (n >= npoin) return;
    const int j0 = row[n], j1 = row[n + 1];
    TR *R[5] = {R0, R1, R2, R3, R4};
    TR acc[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = init ? TR(0) : R[f][n];
    for (int j = j0; j < j1; j += 4) {
        int sj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) sj[u] = j + u < j1 ? src[j + u] : -1;
        TR v[4][NF];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if constexpr (NF == 5 && sizeof(TS) == 8) {
                // an absent slot reads record 0 (always present) and is not added
                double w[5];
                load_rec5((const double *)scr, max(sj[u], 0), w);
#pragma unroll
                for (int f = 0; f < NF; ++f) v[u][f] = TR(w[f]);
            } else {
                // an absent slot reads record 0 unconditionally (cheaper than
                // masking each load) and is not added
#pragma unroll
                for (int f = 0; f < NF; ++f) v[u][f] = TR(scr[(size_t)max(sj[u], 0) * NF + f]);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (sj[u] >= 0) {
#pragma unroll
                for (int f = 0; f < NF; ++f) acc[f] += fact * v[u][f];
            }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) R[f][n] = acc[f];
}

// Tail of the entropy residual on the scalar gather (working nodes): the
// slots of node n summed as k_gather sums them, times invMl (lumped_solver_scal),
// then the residual of k_entropy_residual; Reta and eta are (npoin,2)
struct EntropyTail {
    int npoin_w;
    const int *lpoin_w;
    const double *invMl, *eta;
    double factor_comp, dt;
    double *Reta, *auxReta;
};

__global__ __launch_bounds__(256) void k_gather_entropy(int npoin, const int *__restrict__ wrow,
                                                        const int *__restrict__ wsrc,
                                                        const double *__restrict__ scr,
                                                        EntropyTail t, double fact)
{
    const int per = gridDim.x >> 3; // grid is a multiple of 8
    const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= t.npoin_w) return;
    const int n = t.lpoin_w[i] - 1;
    const int j0 = wrow[i], j1 = wrow[i + 1];
    double acc = 0.0;
    for (int j = j0; j < j1; j += 4) {
        int sj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) sj[u] = j + u < j1 ? wsrc[j + u] : -1;
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = scr[max(sj[u], 0)]; // absent: record 0, not added
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (sj[u] >= 0) acc += fact * v[u];
    }
    const double r2 = acc * t.invMl[n];
    const double r1 = t.Reta[n];
    t.Reta[n + (size_t)npoin] = r2;
    t.auxReta[n] = (1.5 * r2 - 0.5 * r1) +
                   t.factor_comp * (t.eta[n + (size_t)npoin] - t.eta[n]) / t.dt;
    t.Reta[n] = r2;
}

// ---------------------------------------------------------------- launch helpers
// ---------------------------------------------------------------------------
// Fused-stage element kernel with the work of each element split over two
// waves that share its LDS: wave 0 evaluates the split-form convection, wave 1
// the viscous operator and hands its residuals to wave 0 through LDS; wave 0
// puts their sum back into the same LDS slots and wave 1 stores it during the
// next element.  Each wave holds only its own operator's live state, which
// keeps both free of register spills in FP64 at p=3 (a single wave doing both
// spills).  Requires (p+1)^3 == 64: one wave per operator, one lane per node.
//
// The inputs stream through an LDS ring AH elements deep.  Every input of
// element e (He and gpvol, the five conserved fields at its nodes, mu_e) is
// copied global -> LDS by direct-to-LDS loads, which hold no registers while
// in flight, AH elements before e is evaluated; its connectivity row arrives
// AH elements earlier still, since the field addresses are read from it.
// The copies are issued through inline assembly, outside the compiler's wait
// tracking (which cannot tell ring slots apart and would wait for every copy
// in flight before any read of the ring), and each wave waits for its own
// copies with an explicit vector-memory count: a wave issues the same number
// NOPS of copies per element, always in the same order, and loads complete in
// issue order, so element e's copies are complete once at most (AH-1)*NOPS of
// the wave's operations are outstanding.  The wave's stores count as well and
// can only make that wait longer.
//
// Per element, both waves pass three LDS-only barriers:
//   B1  inputs of e in ring slot e%AH (each wave waited for its own copies),
//       viscous residuals of the previous element handed over in LDS; wave 0
//       then completes that element's residual records
//   B2  nodal fields of e staged, slot e%AH read into registers (the copies
//       of element e+AH may start), previous element's records complete
//   B3  line derivatives of u and rho handed over: wave 0 forms them for the
//       convection terms, and wave 1 takes them for grad u and grad rho
//       instead of forming the same sums again
// values per residual record written by the two-wave kernel.  With FP32
// operators the five residuals are padded to six, so that a record starts on
// an 8-byte boundary and the gather reads it with three 8-byte loads (the
// gather then takes 0.25 instead of 0.43 ms at 64^3 elements); in FP64 the
// 20 % larger records cost the element kernel more than the gather gains.
template <typename T> constexpr int split_rw() { return sizeof(T) == 4 ? 6 : 5; }

template <int P, typename T, bool G16, int RUNV = SODHIP_SPLIT_RUN, bool LPSV = false>
struct SplitLDS {
    static constexpr int NN = (P + 1) * (P + 1) * (P + 1);
    static constexpr int AH = sizeof(T) == 8 ? SODHIP_SPLIT_AHEAD64 : SODHIP_SPLIT_AHEAD32;
    // geometry, He then gpvol, copied in 64-lane rows of GW bytes per lane
    static constexpr int HE_BYTES = 9 * NN * (int)sizeof(T), GV_BYTES = NN * (int)sizeof(T);
    static constexpr int GW = G16 ? 16 : 4, GROW = 64 * GW;
    static constexpr int GROWS = (HE_BYTES + GV_BYTES + GROW - 1) / GROW, GEO = GROWS * GROW;
    // conserved fields rho, q1, q2, q3, E (FP64, node-major), then mu_e
    // mu_e(nelem,ngaus) is element-fastest: the 64 values of one element lie
    // nelem apart.  In FP64 it is copied per run of RUN consecutive elements,
    // RUN values per node, into one of three blocks, two runs ahead (each
    // 64-lane row then reads 8 runs of 32 contiguous bytes instead of 32
    // scattered doubles); with FP32 operators per element into its slot.
    static constexpr int RUN = RUNV;
    static constexpr bool MUBLK = sizeof(T) == 8;
    static constexpr int ST = 5 * NN * 8, MU = MUBLK ? 0 : NN * (int)sizeof(T);
    // LPS: the projection records of the element's nodes (kProjRec doubles
    // per node, lexicographic order: 6 rows of 64 16-byte pieces) and its
    // tau_stab (the first 8 bytes of one 4-byte-per-lane row)
    static constexpr int PJ_AT = GEO + ST + MU, PJ = LPSV ? NN * kProjRec * 8 : 0;
    static constexpr int PJ_ROWS = PJ / 1024, TAU_AT = PJ_AT + PJ, TAU = LPSV ? 256 : 0;
    static_assert(!LPSV || (NN == 64 && PJ == PJ_ROWS * 1024), "LPS records: whole rows");
    static constexpr int SLOT = GEO + ST + MU + PJ + TAU;
    // connectivity rows: AH + 1 in use at a time
    static constexpr int CROWS = AH + 1, CROW = NN * 4;
    static constexpr int MUB = RUN * NN * (int)sizeof(T), MUBS = MUBLK ? 3 : 0;
    static constexpr int MUB_AT = AH * SLOT + CROWS * CROW;
    static constexpr int BYTES = MUB_AT + MUBS * MUB;
    static constexpr int MU_OPS = (int)sizeof(T) / 4;
    static_assert(!MUBLK || ((RUN == 4 || RUN == 2) && MUB == 256 * MU_OPS * RUN),
                  "one run's mu_e block in MU_OPS rows per element");
    // copies per element: wave 0 takes the connectivity row, the even
    // geometry rows and rho, q1, q2; wave 1 the odd rows, q3, E and mu_e
    // (with LPS the viscous wave also copies the projection rows and tau)
    static constexpr int NOPS0 = 1 + (GROWS + 1) / 2 + 6,
                         NOPS1 = GROWS / 2 + 4 + MU_OPS + (LPSV ? PJ_ROWS + 1 : 0);
    static_assert(SLOT % 16 == 0, "ring slots must stay 16-byte aligned");
    static_assert(AH >= 2 && AH * NOPS0 <= 63 && AH * NOPS1 <= 63, "vmcnt range");
};

template <int P, typename T> struct SplitState {
    double cr, cq[3], cE;
    T mue;
};

// direct-to-LDS copy of W (4 or 16) bytes per lane from g to lds + W*lane
// (lds wave-uniform), issued outside the compiler's wait tracking
template <int W> __device__ __forceinline__ void dma_lds(const void *g, unsigned lds)
{
    const unsigned m = __builtin_amdgcn_readfirstlane(lds);
    if constexpr (W == 16)
        asm volatile("s_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "{m0}"(m) : "memory");
    else
        asm volatile("s_nop 0\n\tglobal_load_lds_dword %0, off" ::"v"(g), "{m0}"(m) : "memory");
}

// wait until at most N of this wave's vector-memory operations are outstanding
template <int N> __device__ __forceinline__ void wait_vm()
{
    static_assert(N >= 0 && N <= 63, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | kWaitVm0 | ((N >> 4) << 14));
}

// Copies of element e into the ring slot at LDS address slot and, by wave 0,
// the connectivity row of element ec to LDS address crow.  n0 and n1 are the
// global nodes of e at lexicographic (I fastest) positions lane/2 and
// 32 + lane/2: a 64-lane row of a field moves both 4-byte halves of 32 nodes,
// so the field lands as doubles in lexicographic node order, and on a mesh
// numbered along I each row reads 8 runs of p+1 consecutive nodes.
template <int P, typename T, bool G16, int ROLE>
__device__ __forceinline__ void split_issue(const RhsArgs<T> &A, unsigned slot, unsigned crow,
                                            int e, int ec, int n0, int n1, int lane)
{
    using L = SplitLDS<P, T, G16>;
    constexpr int NN = L::NN;
    const int hw = 4 * (lane & 1), nd = lane >> 1;
    if constexpr (ROLE == 0) dma_lds<4>(A.connecT + (size_t)ec * NN + lane, crow);
    const unsigned char *he = (const unsigned char *)(A.He + (size_t)e * 9 * NN);
    const unsigned char *gv = (const unsigned char *)(A.gpvol + (size_t)e * NN);
#pragma unroll
    for (int r = ROLE; r < L::GROWS; r += 2) {
        // lanes past the end of gpvol re-read its last GW bytes
        const int byte = r * L::GROW + lane * L::GW;
        const unsigned char *src =
            byte < L::HE_BYTES ? he + byte : gv + min(byte - L::HE_BYTES, L::GV_BYTES - L::GW);
        dma_lds<L::GW>(src, slot + r * L::GROW);
    }
    const size_t np = (size_t)A.npoin;
    const double *fld[5] = {A.cr, A.cq, A.cq + np, A.cq + 2 * np, A.cE};
    const unsigned st = slot + L::GEO;
#pragma unroll
    for (int f = ROLE == 0 ? 0 : 3; f < (ROLE == 0 ? 3 : 5); ++f) {
        dma_lds<4>((const char *)(fld[f] + n0) + hw, st + f * NN * 8);
        dma_lds<4>((const char *)(fld[f] + n1) + hw, st + f * NN * 8 + 256);
    }
    if constexpr (ROLE == 1 && !L::MUBLK) {
        const unsigned mu = st + L::ST;
        if constexpr (sizeof(T) == 8) {
            dma_lds<4>((const char *)(A.mu_e + e + (size_t)A.nelem * nd) + hw, mu);
            dma_lds<4>((const char *)(A.mu_e + e + (size_t)A.nelem * (32 + nd)) + hw, mu + 256);
        } else {
            dma_lds<4>(A.mu_e + e + (size_t)A.nelem * lane, mu);
        }
    }
}

// FP64 mu_e copies of one iteration: rows 2j and 2j+1 of the block of the run
// starting at element rfirst (j = iteration within its run).  The block is
// node-major and element-fastest: lane L moves half L&1 of
// mu_e(e0 + (L/2)&3, 8i + L/8) of row i.
template <typename T>
__device__ __forceinline__ void mu_issue(const RhsArgs<T> &A, unsigned blk, int rfirst, int j,
                                         int lane)
{
    const int hw = 4 * (lane & 1), el = (lane >> 1) & 3, nd8 = lane >> 3;
    const int e = min(rfirst + el, A.nelem - 1);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i = 2 * j + h;
        dma_lds<4>((const char *)(A.mu_e + e + (size_t)A.nelem * (8 * i + nd8)) + hw, blk + i * 256);
    }
}

// FP64 mu_e copies of one iteration for runs of two elements (the cluster
// order, where elements pair up as e, e+1 on a box mesh): rows 2j and 2j+1 of
// the block of the run of elements e0, e1 (j = iteration within its run).
// The block is node-major and element-fastest: lane L moves half L&1 of
// mu_e(e_{(L/2)&1}, 16i + L/4) of row i.
template <typename T>
__device__ __forceinline__ void mu_issue2(const RhsArgs<T> &A, unsigned blk, int e0, int e1, int j,
                                          int lane)
{
    const int hw = 4 * (lane & 1), nd16 = lane >> 2;
    const int e = ((lane >> 1) & 1) ? e1 : e0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i = 2 * j + h;
        dma_lds<4>((const char *)(A.mu_e + e + (size_t)A.nelem * (16 * i + nd16)) + hw, blk + i * 256);
    }
}

// LPS copies of element e into the ring slot at LDS address slot (viscous
// wave): row r, lane L moves 16-byte piece (64r+L)%6 of the projection record
// of the node at lexicographic position (64r+L)/6, whose element-local index
// ik[r] it holds (c: the element's connectivity row); then tau_stab(e) as two
// 4-byte halves (lanes 0 and 1; the other lanes repeat them)
template <typename T, int PJ_ROWS, int PJ_AT, int TAU_AT>
__device__ __forceinline__ void lps_issue(const RhsArgs<T> &A, unsigned slot, const int *c,
                                          const int (&ik)[PJ_ROWS], int e, int lane)
{
#pragma unroll
    for (int r = 0; r < PJ_ROWS; ++r) {
        const int piece = (64 * r + lane) % 6;
        const int n = c[ik[r]];
        dma_lds<16>(A.prj + kProjRec * (size_t)n + 2 * piece, slot + PJ_AT + r * 1024);
    }
    dma_lds<4>((const char *)(A.tau + e) + 4 * (lane & 1), slot + TAU_AT);
}

// reads ring slot sl into registers: He, gpvol, the conserved fields and mu_e
// of node a (lexicographic position pa)
template <int P, typename T, bool G16>
__device__ __forceinline__ void split_read(const unsigned char *sl, int a, int pa, bool visc,
                                           T He[3][3], T &vol, SplitState<P, T> &x)
{
    using L = SplitLDS<P, T, G16>;
    constexpr int NN = L::NN;
    const T *hp = (const T *)sl + 9 * a;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) He[i][j] = hp[i + 3 * j];
    vol = ((const T *)sl)[9 * NN + a];
    const double *st = (const double *)(sl + L::GEO);
    x.cr = st[pa];
    x.cq[0] = st[NN + pa];
    x.cq[1] = st[2 * NN + pa];
    x.cq[2] = st[3 * NN + pa];
    x.cE = st[4 * NN + pa];
    if constexpr (!L::MUBLK)
        if (visc) x.mue = ((const T *)(sl + L::GEO + L::ST))[a];
}

// residual records of element e (split_rw values per node: the five
// residuals, and a zero when padded; lexicographic node order) from their
// dense copy in LDS, 16 contiguous bytes per lane
template <int P, typename T>
__device__ __forceinline__ void split_store(const RhsArgs<T> &A, const void *dense, int e, int lane)
{
    constexpr int RW = split_rw<T>();
    constexpr int NN = (P + 1) * (P + 1) * (P + 1), CH = RW * NN * (int)sizeof(T) / 16;
    using U4 = unsigned __attribute__((ext_vector_type(4)));
    const U4 *src = (const U4 *)dense;
    U4 *o = (U4 *)(A.out + (size_t)e * RW * NN);
#pragma unroll
    for (int q = lane; q < CH; q += 64) o[q] = src[q];
}

// The two-wave kernel.  The staged nodal fields and the viscous fluxes are
// node-major records, so one 16-byte LDS read returns two fields of a
// neighbour and the field offsets are immediates of the read instruction:
//   node record (12 values): u1 u2 | u3 rho | p H | rhoH q1 | q2 q3 | T -
//   flux record (18 values): for k = 1..3, G_k of (grad rho, grad T, tau u,
//                            tau_1., tau_2., tau_3.)
//
// SPLIT_PAIR: line derivatives D0, D1 and W-weighted line sums DIV0, DIV1 of
// the field pair at record offset OFF (convection wave).
#define SPLIT_PAIR(OFF, OWN0, OWN1, D0, D1, DIV0, DIV1)                                    \
    do {                                                                                   \
        const T o0_ = (OWN0), o1_ = (OWN1);                                                \
        DIV0 = 0;                                                                          \
        DIV1 = 0;                                                                          \
        _Pragma("unroll") for (int k = 0; k < 3; ++k)                                      \
        {                                                                                  \
            T v0 = lt.ds[k] * o0_, v1 = lt.ds[k] * o1_;                                    \
            T w0 = Wo[k] * o0_, w1 = Wo[k] * o1_;                                          \
            _Pragma("unroll") for (int o = 0; o < P; ++o)                                  \
            {                                                                              \
                const V2 x = LD2(nb[k][o] + (OFF));                                        \
                v0 += lt.d[k][o] * x.x;                                                    \
                v1 += lt.d[k][o] * x.y;                                                    \
                w0 += Wx[k][o] * x.x;                                                      \
                w1 += Wx[k][o] * x.y;                                                      \
            }                                                                              \
            D0[k] = v0;                                                                    \
            D1[k] = v1;                                                                    \
            DIV0 += w0;                                                                    \
            DIV1 += w1;                                                                    \
        }                                                                                  \
    } while (0)
// CL: brick-local assembly (FP64).  The block walks element clusters, one
// after the other; the viscous wave adds each element's residual records into
// an LDS accumulator per cluster-local node instead of storing them, and after
// a cluster's last element finishes its interior nodes (mass inverse and
// stage combine, as k_gather_update) and stores one partial record per
// surface node.  mu_e runs are then pairs of elements.
// LPS: the viscous wave evaluates full_diff_stab_ijk (elem_stab.f90:207-438)
// instead of full_diffusion_ijk: the stabilising terms -c_lps tau_e (Proj -
// grad) enter the fluxes, nu_e and kappa_e multiply the gradients at each
// point before the weak divergence, and the projections (node records) and
// tau_e come through the ring with the element's other inputs.
template <int P, typename T, bool G16, bool CL = false, bool LPS = false>
__global__ __launch_bounds__(128, sizeof(T) == 8 ? SODHIP_SPLIT_WAVES : SODHIP_SPLIT_WAVES_F32) void k_rhs_split(RhsArgs<T> A)
{
    using L = SplitLDS<P, T, G16, CL ? 2 : SODHIP_SPLIT_RUN, LPS>;
    static_assert(!CL || (sizeof(T) == 8 && P == 3), "cluster assembly: FP64, p=3");
    static_assert(!LPS || (sizeof(T) == 8 && P == 3 && !CL), "LPS: FP64, p=3, element order");
    using V2 = T __attribute__((ext_vector_type(2)));
    // node records of REC values: 112 bytes in FP64, an odd multiple of 16,
    // so that 16 lanes reading 16 consecutive records with 16-byte reads hit
    // distinct banks (96-byte records would pair them up); 144-byte flux records
    constexpr int NN = L::NN, REC = sizeof(T) == 8 ? 14 : 12, GREC = 18, AH = L::AH;
    // s2 holds, in turn: the viscous flux records (GREC values per node) and
    // the handed-over viscous residuals (3 pairs per node), then the dense
    // residual records (RW values per node) followed by the line derivatives
    // of u and rho (6 pairs per node) from DREC on
    constexpr int RW = split_rw<T>();
    constexpr int DREC = RW * NN / 2;
    static_assert(DREC + 6 * NN <= 64 * GREC / 2, "derivative records fit in s2");
    static_assert(NN == 64, "one lane per node");
    __shared__ V2 s1[64 * REC / 2]; // records of REC values as pairs
    __shared__ V2 s2[64 * GREC / 2];
    __shared__ __attribute__((aligned(16))) unsigned char ring[L::BYTES];
    const unsigned ring_lds = (unsigned)(uintptr_t)(__attribute__((address_space(3))) unsigned char *)ring;
    // cluster mode: the block's element order and local-node ranges, the
    // local-index rows of the last AH + 1 elements, the accumulator
    constexpr int G = kClusterG, CMAX = kClusterBlockMax, CR = AH + 1;
    __shared__ int cel_s[CL ? CMAX * G : 1];
    __shared__ int cb_s[CL ? 2 * CMAX : 1];
    __shared__ __attribute__((aligned(16))) int cloc_s[CL ? CR * 64 : 1];
    __shared__ double acc_s[CL ? 5 * kClusterLMax : 1];
    const unsigned cloc_lds =
        (unsigned)(uintptr_t)(__attribute__((address_space(3))) int *)cloc_s;

    const int role = threadIdx.x >> 6; // 0: convection, 1: viscous (wave-uniform)
    const int lane = threadIdx.x & 63;
    const bool act = true;
    const bool visc = role == 1;

    constexpr int RUN = L::RUN, RSH = RUN == 4 ? 2 : 1;
    // over runs of RUN elements, or over clusters
    const ElemRange er = CL ? elem_range(A.ncl) : elem_range((A.nelem + RUN - 1) / RUN);
    if (er.begin >= er.end) return;
    const int rlast = er.begin + (er.end - 1 - er.begin) / er.stride * er.stride;
    const int nunits = (er.end - er.begin + er.stride - 1) / er.stride;
    const int nit = (CL ? G : RUN) * nunits;
    if constexpr (CL) {
        // a padding slot (-1) of a cluster evaluates its first element again
        // and adds nothing
        for (int i = threadIdx.x; i < nunits * G; i += 128) {
            const long c = er.begin + (long)(i / G) * er.stride;
            cel_s[i] = A.cel[c * G + i % G];
        }
        for (int i = threadIdx.x; i < nunits; i += 128) {
            const long c = er.begin + (long)i * er.stride;
            cb_s[2 * i] = A.cnb[c];
            cb_s[2 * i + 1] = A.cnb[c + 1];
        }
        for (int i = threadIdx.x; i < 5 * kClusterLMax; i += 128) acc_s[i] = 0.0;
        __syncthreads();
    }
    // element of iteration t.  Past the block's last run, and past nelem in
    // the last run, elements repeat: the copies stay unconditional and a
    // repeated element stores the same records again.  In cluster mode past
    // the block's last element the copies repeat that element.
    auto elem_of = [&](int t) {
        if constexpr (CL) {
            const int i = min(t, nit - 1);
            const int e = cel_s[i];
            return e >= 0 ? e : cel_s[i - i % G];
        } else {
            const long r = min((long)er.begin + (long)(t / RUN) * er.stride, (long)rlast);
            return (int)min(r * RUN + t % RUN, (long)A.nelem - 1);
        }
    };
    const int elast = elem_of(nit - 1);
    // FP64 mu_e blocks: iteration t (t >= -2 RUN) copies its share of the
    // block of run t/RUN + 2 (floor); the block of run r is read from
    // iteration RUN r on
    auto mu_at = [&](int t) {
        if constexpr (L::MUBLK) {
            const int r = (t >> RSH) + 2;
            const unsigned blk = ring_lds + L::MUB_AT + (r % 3) * L::MUB;
            if constexpr (CL) {
                mu_issue2<T>(A, blk, elem_of(r * RUN), elem_of(r * RUN + 1), t & (RUN - 1), lane);
            } else {
                const long g = min((long)er.begin + (long)r * er.stride, (long)rlast);
                mu_issue<T>(A, blk, (int)(g * RUN), t & 3, lane);
            }
        }
    };
    auto mu_read = [&](int t, int node) {
        return ((const T *)(ring + L::MUB_AT + ((t >> RSH) % 3) * L::MUB))[node * RUN + (t & (RUN - 1))];
    };
    // cluster mode: the local-index row of iteration t's element, copied by
    // the viscous wave with its share of that element's inputs
    auto cloc_issue = [&](int t) {
        if constexpr (CL) {
            const int i = min(t, nit - 1);
            const long c = er.begin + (long)(i / G) * er.stride;
            dma_lds<4>(A.cloc + ((size_t)c * G + i % G) * NN + lane, cloc_lds + (t % CR) * 256);
        }
    };
    auto slot_of = [&](int t) { return (t % AH) * L::SLOT; };
    auto crow_of = [&](int t) { return AH * L::SLOT + (t % L::CROWS) * L::CROW; };
    const int nd = lane >> 1;

    // Lane L evaluates the node a at lexicographic position L (I fastest);
    // the LDS records are kept in that order.  Along every line direction
    // the 16 lanes of an LDS access phase then read 16 distinct records
    // (a permutation of 16 consecutive ones).  ia0, ia1: the nodes at
    // lexicographic positions nd and 32 + nd.
    auto lex = [&](int r) { return min(max(A.perm[r], 0), NN - 1); };
    int a, ia0, ia1;
    // LPS: element-local index of the node whose projection piece lane
    // copies in row r (lexicographic position (64r + lane)/6)
    int ik[LPS ? L::PJ_ROWS : 1];
    {
        int *inv = (int *)s1;
        if (role == 0) inv[lex(lane)] = lane;
        lds_barrier();
        a = min(max(inv[lane], 0), NN - 1);
        ia0 = min(max(inv[nd], 0), NN - 1);
        ia1 = min(max(inv[32 + nd], 0), NN - 1);
        if constexpr (LPS)
#pragma unroll
            for (int r = 0; r < L::PJ_ROWS; ++r)
                ik[r] = min(max(inv[(64 * r + lane) / 6], 0), NN - 1);
        lds_barrier();
    }
    auto lps_at = [&](int t, const int *c) {
        if constexpr (LPS)
            lps_issue<T, L::PJ_ROWS, L::PJ_AT, L::TAU_AT>(A, ring_lds + slot_of(t), c, ik,
                                                           elem_of(t), lane);
    };
    const int pa = lane;
    LineTab<P, T> lt;
    build_linetab<P, T>(a, A.dl, A.inv, A.gI, A.gJ, A.gK, lt);
    // neighbour records (element offsets into s1)
    int nb[3][P];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int o = 0; o < P; ++o) nb[k][o] = lex(lt.off[k][o]) * REC;
    const int own = pa * REC;

    // the connectivity of the first AH elements, then the first AH elements
    // in the ring in the order of the main loop
    if (role == 0)
        for (int t = 0; t < AH; ++t)
            dma_lds<4>(A.connecT + (size_t)elem_of(t) * NN + lane, ring_lds + crow_of(t));
    else
        for (int t = -2 * RUN; t < -AH; ++t) mu_at(t);
    wait_vm<0>();
    lds_barrier();
    for (int t = 0; t < AH; ++t) {
        const int *c = (const int *)(ring + crow_of(t));
        const int n0 = c[ia0], n1 = c[ia1];
        if (visc) {
            split_issue<P, T, G16, 1>(A, ring_lds + slot_of(t), ring_lds + crow_of(t + AH), elem_of(t),
                                      elem_of(t + AH), n0, n1, lane);
            mu_at(t - AH);
            cloc_issue(t);
            lps_at(t, c);
        } else
            split_issue<P, T, G16, 0>(A, ring_lds + slot_of(t), ring_lds + crow_of(t + AH), elem_of(t),
                                      elem_of(t + AH), n0, n1, lane);
        lds_barrier(); // the row read here is rewritten in a later pass
    }
    SplitState<P, T> cur;

    // reciprocals of the gas constant and of Pr (one rounding apart from the
    // divisions of the reference's expressions)
    const double rRgas = 1.0 / A.Rgas;
    const T cp_pr = A.Cp / A.Pr;
    double rinv = 0.0;
    auto prim = [&](const SplitState<P, T> &c, double &r, double ud[3], double &pd, double &hd,
                    double &Td, double &umag) {
        r = c.cr;
        const double ri = 1.0 / r;
        umag = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ud[d] = c.cq[d] * ri;
            umag += ud[d] * ud[d];
        }
        const double eint = c.cE * ri - 0.5 * umag;
        pd = r * (A.gamma - 1.0) * eint;
        hd = (A.gamma / (A.gamma - 1.0)) * pd * ri;
        Td = pd * ri * rRgas;
        rinv = ri;
    };
        // indices are even value offsets into the records
#define LD1(i) (s1[(i) >> 1].x)
#define LD2(i) (s1[(i) >> 1])
#define GL2(i) (s2[(i) >> 1])

    if (!visc) {
        // ======================= convection wave
        // convection residuals (times gpvol) of the previous element
        T pend[5] = {0, 0, 0, 0, 0};
        // its records: pend plus the viscous residuals handed over in s2, as
        // dense records in lexicographic node order over the same LDS (all
        // lanes read before any writes; a padded record is exactly the slot
        // the lane read its handover from); the viscous wave stores them
        auto finish = [&]() {
            const V2 h01 = s2[3 * pa], h23 = s2[3 * pa + 1], h4 = s2[3 * pa + 2];
            if constexpr (RW == 6) {
                s2[3 * pa] = V2{pend[0] + h01.x, pend[1] + h01.y};
                s2[3 * pa + 1] = V2{pend[2] + h23.x, pend[3] + h23.y};
                s2[3 * pa + 2] = V2{pend[4] + h4.x, T(0)};
            } else {
                T *d = (T *)s2 + 5 * pa;
                d[0] = pend[0] + h01.x;
                d[1] = pend[1] + h01.y;
                d[2] = pend[2] + h23.x;
                d[3] = pend[3] + h23.y;
                d[4] = pend[4] + h4.x;
            }
        };
        for (int t = 0; t < nit; ++t) {
            wait_vm<(AH - 1) * L::NOPS0>();
            lds_barrier(); // B1
            if (t != 0) finish();
            T He[3][3], vol;
            split_read<P, T, G16>(ring + slot_of(t), a, pa, false, He, vol, cur);
            T rho, u[3], q[3], H, rhoH;
            {
                double r, ud[3], pd, hd, Td, umag;
                prim(cur, r, ud, pd, hd, Td, umag);
                rho = T(r);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    u[d] = T(ud[d]);
                    q[d] = T(cur.cq[d]);
                }
                H = T(hd + 0.5 * umag);
                rhoH = T(r * (hd + 0.5 * umag));
                if (act) {
                    s1[(own + 0) >> 1] = V2{u[0], u[1]};
                    s1[(own + 2) >> 1] = V2{u[2], rho};
                    s1[(own + 4) >> 1] = V2{T(pd - A.p_ref), H};
                    s1[(own + 6) >> 1] = V2{rhoH, q[0]};
                    s1[(own + 8) >> 1] = V2{q[1], q[2]};
                }
            }
            lds_barrier(); // B2
            {
                const int *c = (const int *)(ring + crow_of(t + AH));
                split_issue<P, T, G16, 0>(A, ring_lds + slot_of(t), ring_lds + crow_of(t + 2 * AH),
                                          elem_of(t + AH), elem_of(t + 2 * AH), c[ia0], c[ia1], lane);
            }

            T Rm = 0, Rq[3] = {0, 0, 0}, Re = 0;
            T Ut[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) Ut[k] = He[0][k] * u[0] + He[1][k] * u[1] + He[2][k] * u[2];
            // velocity and density lines: contravariant line weights W_k(m),
            // div(u_i u), div u, q_k d/dxi_k u_i and d/dxi_k rho
            T Wo[3], Wx[3][P];
            T divUw = 0, drho[3];
            {
                T Qt[3], divFuu[3] = {0, 0, 0};
#pragma unroll
                for (int k = 0; k < 3; ++k) Qt[k] = He[0][k] * q[0] + He[1][k] * q[1] + He[2][k] * q[2];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const T c0 = lt.ds[k];
                    T du[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i) du[i] = c0 * u[i];
                    T dr = c0 * rho;
                    Wo[k] = c0 * Ut[k];
                    divUw += Wo[k];
#pragma unroll
                    for (int i = 0; i < 3; ++i) divFuu[i] += Wo[k] * u[i];
#pragma unroll
                    for (int o = 0; o < P; ++o) {
                        const T c = lt.d[k][o];
                        const V2 x01 = LD2(nb[k][o] + 0), x2r = LD2(nb[k][o] + 2);
                        const T ub[3] = {x01.x, x01.y, x2r.x};
#pragma unroll
                        for (int i = 0; i < 3; ++i) du[i] += c * ub[i];
                        dr += c * x2r.y;
                        const T w = c * (He[0][k] * ub[0] + He[1][k] * ub[1] + He[2][k] * ub[2]);
                        Wx[k][o] = w;
                        divUw += w;
#pragma unroll
                        for (int i = 0; i < 3; ++i) divFuu[i] += w * ub[i];
                    }
                    drho[k] = dr;
                    s2[DREC + 6 * pa + 2 * k] = V2{du[0], du[1]};
                    s2[DREC + 6 * pa + 2 * k + 1] = V2{du[2], dr};
#pragma unroll
                    for (int i = 0; i < 3; ++i) Rq[i] += Qt[k] * du[i];
                }
#pragma unroll
                for (int i = 0; i < 3; ++i) Rq[i] = T(0.25) * (Rq[i] + rho * divFuu[i]);
            }
            lds_barrier(); // B3
            {
                const T urho = Ut[0] * drho[0] + Ut[1] * drho[1] + Ut[2] * drho[2];
                Rm += T(0.5) * urho;
#pragma unroll
                for (int i = 0; i < 3; ++i) Rq[i] += T(0.25) * u[i] * urho;
                Re += T(0.25) * H * urho;
            }
            // field pairs (p, H), (rho H, q1), (q2, q3): line derivative d and
            // W-weighted line sum div = div(s u) of both fields
            T divQ = 0;
            auto grad_i = [&](int i, const T d[3]) { return He[i][0] * d[0] + He[i][1] * d[1] + He[i][2] * d[2]; };
            auto utd = [&](const T d[3]) { return Ut[0] * d[0] + Ut[1] * d[1] + Ut[2] * d[2]; };
            {
                T dp[3], dH[3], dv0, dv1;
                SPLIT_PAIR(4, LD1(own + 4), H, dp, dH, dv0, dv1);
#pragma unroll
                for (int i = 0; i < 3; ++i) Rq[i] += grad_i(i, dp);
                T qt = 0;
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    qt += (He[0][k] * q[0] + He[1][k] * q[1] + He[2][k] * q[2]) * dH[k];
                Re += T(0.25) * (qt + rho * dv1);
            }
            __builtin_amdgcn_sched_barrier(0);
            {
                T dRH[3], dq0[3], dv0, dv1;
                SPLIT_PAIR(6, rhoH, q[0], dRH, dq0, dv0, dv1);
                Re += T(0.25) * (utd(dRH) + dv0);
                divQ += grad_i(0, dq0);
                Rq[0] += T(0.25) * (utd(dq0) + dv1);
            }
            __builtin_amdgcn_sched_barrier(0);
            {
                T dq1[3], dq2[3], dv0, dv1;
                SPLIT_PAIR(8, q[1], q[2], dq1, dq2, dv0, dv1);
                divQ += grad_i(1, dq1);
                Rq[1] += T(0.25) * (utd(dq1) + dv0);
                divQ += grad_i(2, dq2);
                Rq[2] += T(0.25) * (utd(dq2) + dv1);
            }
            Rm += T(0.5) * (divQ + rho * divUw);
#pragma unroll
            for (int i = 0; i < 3; ++i) Rq[i] += T(0.25) * (u[i] * divQ + q[i] * divUw);
            Re += T(0.25) * (H * divQ + rhoH * divUw);
            pend[0] = Rm * vol;
            pend[1] = Rq[0] * vol;
            pend[2] = Rq[1] * vol;
            pend[3] = Rq[2] * vol;
            pend[4] = Re * vol;
        }
        // the last element's records
        lds_barrier();
        finish();
        lds_barrier();
    } else {
        // ======================= viscous wave
        T tt[3][P];
        {
            const int own3[3] = {min(max(A.gI[a] - 1, 0), P), min(max(A.gJ[a] - 1, 0), P),
                                 min(max(A.gK[a] - 1, 0), P)};
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int o = 0; o < P; ++o) tt[k][o] = A.dl[lt.off[k][o] + NN * (k + 3 * own3[k])];
        }
        int gnb[3][P];
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int o = 0; o < P; ++o) gnb[k][o] = lex(lt.off[k][o]) * GREC + 6 * k;
        // cluster mode: adds iteration tt's dense records (s2, lexicographic
        // node order) into the accumulator at the element's local nodes (64
        // distinct local nodes per element: no two lanes collide)
        auto cl_accumulate = [&](int tt) {
            if constexpr (CL) {
                const int i = min(tt, nit - 1);
                if (cel_s[i] < 0) return; // padding slot
                const int sl = min(max(cloc_s[(tt % CR) * 64 + lane], 0), kClusterLMax - 1);
                const T *d = (const T *)s2 + RW * lane;
                double *ac = acc_s + 5 * sl;
#pragma unroll
                for (int f = 0; f < 5; ++f) ac[f] += (double)d[f];
            }
        };
        // cluster mode, after the last element of the cluster of iteration tt:
        // interior working nodes get R = fact * sum * invMl and the stage
        // combine (k_gather_update's arithmetic; the sum runs over the same
        // elements in the same ascending order), surface working nodes one
        // partial record; the accumulator is cleared for the next cluster
        auto cl_flush = [&](int tt) {
            if constexpr (CL) {
                const int k = min(tt, nit - 1) / G;
                const int base = cb_s[2 * k], lim = cb_s[2 * k + 1];
                const size_t np = (size_t)A.npoin;
                for (int l = base + lane; l < lim; l += 64) {
                    double v[5];
                    double *ac = acc_s + 5 * (l - base);
#pragma unroll
                    for (int f = 0; f < 5; ++f) {
                        v[f] = ac[f];
                        ac[f] = 0.0;
                    }
                    const int kind = A.ckind[l];
                    if (kind == 1) {
                        double *r = A.rec + 5 * (size_t)l;
#pragma unroll
                        for (int f = 0; f < 5; ++f) r[f] = v[f];
                    } else if (kind == 0) {
                        const size_t n = (size_t)A.cnode[l];
                        const size_t k3[5] = {n, n, n + np, n + 2 * np, n};
                        const double *y1v[5] = {A.y1r, A.y1q, A.y1q, A.y1q, A.y1E};
                        double *yv[5] = {A.yr, A.yq, A.yq, A.yq, A.yE};
                        double *rv[5] = {A.Rm, A.Rq, A.Rq, A.Rq, A.Re};
                        double *xv[5] = {A.nxr, A.nxq, A.nxq, A.nxq, A.nxE};
                        const double m = A.invMl[n];
                        double b[5], y[5];
#pragma unroll
                        for (int f = 0; f < 5; ++f) {
                            b[f] = A.rd1 ? y1v[f][k3[f]] : 0.0;
                            y[f] = A.rdy ? yv[f][k3[f]] : 0.0;
                        }
#pragma unroll
                        for (int f = 0; f < 5; ++f) {
                            const double r = (A.fact * v[f]) * m;
                            const double z = A.lam1 * b[f] + A.lams * y[f] + A.gdt * r;
                            yv[f][k3[f]] = z;
                            if (A.nxr) {
                                const double zero = 0.0, nls = 0.0;
                                xv[f][k3[f]] = A.nx_l1 * z + nls * zero + A.nx_gdt * r;
                            }
                            if (A.Rm) rv[f][k3[f]] = r;
                        }
                    }
                }
            }
        };
        for (int t = 0; t < nit; ++t) {
            wait_vm<(AH - 1) * (L::NOPS1 + (CL ? 1 : 0))>();
            lds_barrier(); // B1
            T He[3][3], vol;
            split_read<P, T, G16>(ring + slot_of(t), a, pa, true, He, vol, cur);
            if constexpr (L::MUBLK) cur.mue = mu_read(t, a);
            // LPS: the node's projection record and tau_e, read before B2 (the
            // copies of element t + AH then start into this slot)
            double pj[LPS ? kProjRec : 1], taus = 0.0;
            if constexpr (LPS) {
                const double *pr = (const double *)(ring + slot_of(t) + L::PJ_AT) + kProjRec * pa;
#pragma unroll
                for (int j = 0; j < kProjRec; ++j) pj[j] = pr[j];
                taus = *(const double *)(ring + slot_of(t) + L::TAU_AT);
            }
            T rho, u[3], Tk, muf;
            const T mue = cur.mue;
            {
                double r, ud[3], pd, hd, Td, umag;
                prim(cur, r, ud, pd, hd, Td, umag);
                rho = T(r);
#pragma unroll
                for (int d = 0; d < 3; ++d) u[d] = T(ud[d]);
                Tk = T(Td);
                const double mfe = A.mu_factor_e ? A.mu_factor_e[elem_of(t)] : A.mu_factor;
                muf = T(sutherland<double>(mfe, Td));
                if (act) s1[(own + 10) >> 1] = V2{Tk, T(0)};
            }
            lds_barrier(); // B2
            // cluster mode: the previous element's records go into the
            // accumulator before its local-index row is overwritten below
            if constexpr (CL)
                if (t != 0) cl_accumulate(t - 1);
            {
                const int *c = (const int *)(ring + crow_of(t + AH));
                split_issue<P, T, G16, 1>(A, ring_lds + slot_of(t), ring_lds + crow_of(t + 2 * AH),
                                          elem_of(t + AH), elem_of(t + 2 * AH), c[ia0], c[ia1], lane);
                mu_at(t);
                cloc_issue(t + AH);
                lps_at(t + AH, c);
            }
            // residual records of the previous element, completed in s2 by the
            // convection wave: their stores complete while this element runs
            // (read here, before this wave's flux records overwrite s2)
            if constexpr (CL) {
                if (t != 0 && t % G == 0) cl_flush(t - 1);
            } else if (t != 0) {
                split_store<P, T>(A, s2, elem_of(t - 1), lane);
            }
            // grad T from the T lines; grad u and grad rho from the line
            // derivatives handed over by the convection wave
            T dT[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                T vt = lt.ds[k] * Tk;
#pragma unroll
                for (int o = 0; o < P; ++o) vt += lt.d[k][o] * LD1(nb[k][o] + 10);
                dT[k] = vt;
            }
            lds_barrier(); // B3
            T gU[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            T dr[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const V2 d01 = s2[DREC + 6 * pa + 2 * k], d2r = s2[DREC + 6 * pa + 2 * k + 1];
                const T du[3] = {d01.x, d01.y, d2r.x};
                dr[k] = d2r.y;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) gU[i][j] += He[j][k] * du[i];
            }
            T gR[3], gt[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                gR[i] = He[i][0] * dr[0] + He[i][1] * dr[1] + He[i][2] * dr[2];
                gt[i] = He[i][0] * dT[0] + He[i][1] * dT[1] + He[i][2] * dT[2];
            }
            // contravariant fluxes G_k(F) = gpvol sum_d He(d,k) F_d into the flux record
            if constexpr (LPS) {
                // full_diff_stab_ijk's fluxes at the point (mu_sgs = 0), in
                // the arithmetic of k_diff_stab; record per k: mass, energy,
                // (unused), momentum 1..3
                const double rhon = rho, mufl = muf, musgs = 0.0;
                const double twoThirds = 2.0 / 3.0;
                const double divU = gU[0][0] + gU[1][1] + gU[2][2];
                const double nu_e = A.c_rho * mue / rhon;
                const double mu_fgp = mufl + rhon * musgs;
                const double kappa_e = mufl * A.Cp / A.Pr + A.c_ener * mue * A.Cp / A.Pr +
                                       A.Cp * rhon * musgs / 0.9;
                const double cl = -A.c_lps * taus;
                // projected tensor rows from its upper triangle t11 t12 t13 t22 t23 t33
                const int at[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
                double F[5][3];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        double tt2 = gU[i][d] + gU[d][i];
                        if (i == d) tt2 = tt2 - twoThirds * divU;
                        F[1 + i][d] = cl * (pj[at[3 * i + d]] - tt2) * rhon;
                    }
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    F[4][d] = cl * (pj[9 + d] - gt[d]) * rhon * A.Cp / A.Pr;
                    F[0][d] = cl * (pj[6 + d] - gR[d]);
                }
                double tauU[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int i = 0; i < 3; ++i) {
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        tauU[i] = tauU[i] + (mu_fgp + mue) * (gU[i][j] + gU[j][i]) * u[j];
                        double tv = (mu_fgp + mue) * (gU[i][j] + gU[j][i]);
                        if (i == j) tv = tv - mu_fgp * twoThirds * divU;
                        F[1 + i][j] = F[1 + i][j] + tv;
                    }
                    tauU[i] = tauU[i] - mu_fgp * twoThirds * divU * u[i];
                }
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    F[4][d] = F[4][d] + kappa_e * gt[d] + tauU[d];
                    F[0][d] = F[0][d] + nu_e * gR[d];
                }
                const int g = pa * GREC;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double Gf[5];
#pragma unroll
                    for (int f = 0; f < 5; ++f)
                        Gf[f] = vol * (He[0][k] * F[f][0] + He[1][k] * F[f][1] + He[2][k] * F[f][2]);
                    if (act) {
                        s2[(g + 6 * k + 0) >> 1] = V2{Gf[0], Gf[4]};
                        s2[(g + 6 * k + 2) >> 1] = V2{T(0), Gf[1]};
                        s2[(g + 6 * k + 4) >> 1] = V2{Gf[2], Gf[3]};
                    }
                }
            } else {
                const T mut = muf + mue;
                const T divU = gU[0][0] + gU[1][1] + gU[2][2];
                const T tdiv = muf * T(2.0 / 3.0) * divU;
                T tau[3][3];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) tau[i][j] = mut * (gU[i][j] + gU[j][i]);
#pragma unroll
                for (int i = 0; i < 3; ++i) tau[i][i] -= tdiv;
                const int g = pa * GREC;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const T h0 = vol * He[0][k], h1 = vol * He[1][k], h2 = vol * He[2][k];
                    T gr[3], tuk = 0;
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        gr[r] = h0 * tau[r][0] + h1 * tau[r][1] + h2 * tau[r][2];
                        tuk += gr[r] * u[r]; // tau is symmetric: contravariant (tau u)_k
                    }
                    if (act) {
                        s2[(g + 6 * k + 0) >> 1] = V2{h0 * gR[0] + h1 * gR[1] + h2 * gR[2],
                                                    h0 * gt[0] + h1 * gt[1] + h2 * gt[2]};
                        s2[(g + 6 * k + 2) >> 1] = V2{tuk, gr[0]};
                        s2[(g + 6 * k + 4) >> 1] = V2{gr[1], gr[2]};
                    }
                }
            }
            // this wave alone reads s2 next: its LDS instructions execute in order
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
            // pass 2: weak divergence sum_k sum_b dl(b,k,I_k(a)) G_k(F)(b)
            T W6[6] = {0, 0, 0, 0, 0, 0};
            {
                const int g = pa * GREC;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const T c0 = lt.ds[k];
                    const V2 o01 = GL2(g + 6 * k), o23 = GL2(g + 6 * k + 2), o45 = GL2(g + 6 * k + 4);
                    T v[6] = {c0 * o01.x, c0 * o01.y, c0 * o23.x, c0 * o23.y, c0 * o45.x, c0 * o45.y};
#pragma unroll
                    for (int o = 0; o < P; ++o) {
                        const T c = tt[k][o];
                        const V2 x01 = GL2(gnb[k][o]), x23 = GL2(gnb[k][o] + 2), x45 = GL2(gnb[k][o] + 4);
                        v[0] += c * x01.x;
                        v[1] += c * x01.y;
                        v[2] += c * x23.x;
                        v[3] += c * x23.y;
                        v[4] += c * x45.x;
                        v[5] += c * x45.y;
                    }
#pragma unroll
                    for (int f = 0; f < 6; ++f) W6[f] += v[f];
                }
            }
            const T nu_e = A.c_rho * mue * T(rinv);
            const T kappa = (muf + A.c_ener * mue) * cp_pr;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
            if (act) {
                if constexpr (LPS) {
                    // nu_e and kappa_e are inside the fluxes
                    s2[3 * pa] = V2{W6[0], W6[3]};
                    s2[3 * pa + 1] = V2{W6[4], W6[5]};
                    s2[3 * pa + 2] = V2{W6[1], T(0)};
                } else {
                    s2[3 * pa] = V2{nu_e * W6[0], W6[3]};
                    s2[3 * pa + 1] = V2{W6[4], W6[5]};
                    s2[3 * pa + 2] = V2{W6[2] + kappa * W6[1], T(0)};
                }
            }
        }
        // the last element's records, completed by the convection wave
        lds_barrier();
        lds_barrier();
        if constexpr (CL) {
            cl_accumulate(nit - 1);
            cl_flush(nit - 1);
        } else {
            split_store<P, T>(A, s2, elast, lane);
        }
    }
    // no copy may still be writing this block's LDS when it retires
    wait_vm<0>();
#undef LD1
#undef SPLIT_PAIR
#undef LD2
#undef GL2
}

template <typename T>
static int launch_rhs(sodhip_ctx ctx, int porder, bool conv, bool diff, bool cons,
                      const RhsArgs<T> &A0, int grid)
{
    // 16-byte geometry copies need every element's He and gpvol block on a
    // 16-byte boundary
    RhsArgs<T> A = A0;
    const size_t nn = (size_t)nnode_of(porder);
    A.geo16 = ((uintptr_t)A.He % 16 == 0) && ((uintptr_t)A.gpvol % 16 == 0) &&
              (9 * nn * sizeof(T)) % 16 == 0 && (nn * sizeof(T)) % 16 == 0;
#define SODHIP_RHS_CASE(PP)                                                                \
    case PP:                                                                               \
        if (conv && diff && cons)                                                          \
            k_rhs<PP, T, true, true, true><<<grid, EG<PP>::NT, 0, ctx->stream>>>(A);       \
        else if (conv && !diff && !cons)                                                   \
            k_rhs<PP, T, true, false, false><<<grid, EG<PP>::NT, 0, ctx->stream>>>(A);     \
        else if (!conv && diff && !cons)                                                   \
            k_rhs<PP, T, false, true, false><<<grid, EG<PP>::NT, 0, ctx->stream>>>(A);     \
        else                                                                               \
            return set_error(SODHIP_ERR_UNSUPPORTED, "unsupported operator combination");  \
        break;
    switch (porder) {
        SODHIP_RHS_CASE(2)
        SODHIP_RHS_CASE(3)
        SODHIP_RHS_CASE(4)
    default:
        return set_error(SODHIP_ERR_UNSUPPORTED, "porder %d not supported (2..4)", porder);
    }
#undef SODHIP_RHS_CASE
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

template <typename TS, typename TR>
static int launch_gather(sodhip_ctx ctx, const MeshCache *m, const TS *scr, size_t EN, int nf,
                         TR *R0, TR *R1, TR *R2, TR *R3, TR *R4, int init, TR fact)
{
    (void)EN;
    const int bs = 256;
    const int grid = ((m->npoin + bs - 1) / bs + 7) / 8 * 8;
#define SODHIP_GATHER(NFV)                                                                     \
    k_gather<TS, TR, NFV><<<grid, bs, 0, ctx->stream>>>(m->npoin, m->row, m->src, scr, R0, R1,  \
                                                        R2, R3, R4, init, fact)
    switch (nf) {
    case 1: SODHIP_GATHER(1); break;
    case 3: SODHIP_GATHER(3); break;
    case 4: SODHIP_GATHER(4); break;
    case 5: SODHIP_GATHER(5); break;
    default: return set_error(SODHIP_ERR_UNSUPPORTED, "gather of %d fields", nf);
    }
#undef SODHIP_GATHER
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

template <typename T>
static int check_common(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec)
{
    SODHIP_REQUIRE(ctx, "null context");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_REQUIRE(nelem > 0 && npoin > 0 && connec, "invalid mesh arguments");
    return SODHIP_OK;
}

// full_convec_ijk's element records only (5 values per element-node, unscaled,
// at e*nnode + perm[a] of out), for stage drivers that add more terms to the
// same records before one gather
template <typename T>
int conv_elements(sodhip_ctx ctx, int porder, const MeshCache *m, const T *He, const T *gpvol,
                  const T *dl, const int *inv, const int *gI, const int *gJ, const int *gK,
                  const T *u, const T *q, const T *rho, const T *pr, const T *E, T *out)
{
    RhsArgs<T> A{};
    A.nelem = m->nelem;
    A.npoin = m->npoin;
    A.EN = (size_t)m->nelem * nnode_of(porder);
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dl;
    A.inv = inv;
    A.gI = gI;
    A.gJ = gJ;
    A.gK = gK;
    A.u = u;
    A.q = q;
    A.rho = rho;
    A.pr = pr;
    A.E = E;
    A.out = out;
    int tok = prof_begin(ctx, 0);
    const int rc = launch_rhs<T>(ctx, porder, true, false, false, A, elem_grid(ctx, m->nelem, 128));
    prof_end(ctx, tok);
    return rc;
}
template int conv_elements<double>(sodhip_ctx, int, const MeshCache *, const double *,
                                   const double *, const double *, const int *, const int *,
                                   const int *, const int *, const double *, const double *,
                                   const double *, const double *, const double *, double *);

template <typename T>
static int full_convec(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
                       const T *He, const T *gpvol, const T *dl, const int *inv, const int *gI,
                       const int *gJ, const int *gK, const T *u, const T *q, const T *rho,
                       const T *pr, const T *E, T *Rmass, T *Rmom, T *Rener, int initialize,
                       T fact)
{
    int rc = check_common<T>(ctx, porder, nelem, npoin, connec);
    if (rc) return rc;
    SODHIP_REQUIRE(He && gpvol && dl && inv && gI && gJ && gK && u && q && rho && pr && E &&
                       Rmass && Rmom && Rener,
                   "null array argument");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    if ((rc = get_mesh(ctx, porder, nelem, npoin, connec, &m))) return rc;
    const size_t EN = (size_t)nelem * nnode_of(porder);
    if ((rc = ensure_buf(ctx->scratch, 5 * EN * sizeof(T)))) return rc;
    RhsArgs<T> A{};
    A.nelem = nelem;
    A.npoin = npoin;
    A.EN = EN;
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dl;
    A.inv = inv;
    A.gI = gI;
    A.gJ = gJ;
    A.gK = gK;
    A.u = u;
    A.q = q;
    A.rho = rho;
    A.pr = pr;
    A.E = E;
    A.out = (T *)ctx->scratch.ptr;
    int tok = prof_begin(ctx, 0);
    rc = launch_rhs<T>(ctx, porder, true, false, false, A, elem_grid(ctx, nelem, 128));
    prof_end(ctx, tok);
    if (rc) return rc;
    tok = prof_begin(ctx, 1);
    rc = launch_gather<T, T>(ctx, m, A.out, EN, 5, Rmass, Rmom, Rmom + npoin,
                             Rmom + 2 * (size_t)npoin, Rener, initialize, fact);
    prof_end(ctx, tok);
    return rc;
}

template <typename T>
static int full_diffusion(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
                          const T *He, const T *gpvol, const T *dl, const int *inv, const int *gI,
                          const int *gJ, const int *gK, T Cp, T Pr, const T *rho_n, const T *rho,
                          const T *u, const T *Tem, const T *mu_fluid, const T *mu_e,
                          const T *mu_sgs, T c_rho, T c_ener, T *Rmass, T *Rmom, T *Rener,
                          int initialize, T fact)
{
    int rc = check_common<T>(ctx, porder, nelem, npoin, connec);
    if (rc) return rc;
    SODHIP_REQUIRE(He && gpvol && dl && inv && gI && gJ && gK && rho_n && rho && u && Tem &&
                       mu_fluid && mu_e && Rmass && Rmom && Rener,
                   "null array argument");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    if ((rc = get_mesh(ctx, porder, nelem, npoin, connec, &m))) return rc;
    const size_t EN = (size_t)nelem * nnode_of(porder);
    if ((rc = ensure_buf(ctx->scratch, 5 * EN * sizeof(T)))) return rc;
    RhsArgs<T> A{};
    A.nelem = nelem;
    A.npoin = npoin;
    A.EN = EN;
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dl;
    A.inv = inv;
    A.gI = gI;
    A.gJ = gJ;
    A.gK = gK;
    A.rho = rho;
    A.u = u;
    A.rho_n = rho_n;
    A.Tem = Tem;
    A.mu_fluid = mu_fluid;
    A.mu_e = mu_e;
    A.mu_sgs = mu_sgs;
    A.Cp = Cp;
    A.Pr = Pr;
    A.c_rho = c_rho;
    A.c_ener = c_ener;
    A.out = (T *)ctx->scratch.ptr;
    int tok = prof_begin(ctx, 0);
    rc = launch_rhs<T>(ctx, porder, false, true, false, A, elem_grid(ctx, nelem, 128));
    prof_end(ctx, tok);
    if (rc) return rc;
    tok = prof_begin(ctx, 1);
    rc = launch_gather<T, T>(ctx, m, A.out, EN, 5, Rmass, Rmom, Rmom + npoin,
                             Rmom + 2 * (size_t)npoin, Rener, initialize, fact);
    prof_end(ctx, tok);
    return rc;
}

// Fused RHS used by the stage driver (ops_stage.hip): convection + viscous
// operator from the conserved FP64 state; writes element partials only.
// values per record of rhs_fused_elements' output (operator precision 8 or 4)
int fused_record_width(int porder, int op_bytes)
{
#if SODHIP_SPLIT_RHS
    if (porder == 3) return op_bytes == 4 ? split_rw<float>() : split_rw<double>();
#endif
    (void)porder;
    (void)op_bytes;
    return 5;
}

template <typename T>
int rhs_fused_elements(sodhip_ctx ctx, int porder, const MeshCache *m, const T *He,
                       const T *gpvol, const T *dl, const int *inv, const int *gI, const int *gJ,
                       const int *gK, const double *rho, const double *q, const double *E,
                       const T *mu_e, double gamma, double Rgas, double Cp, double Pr,
                       double mu_factor, const double *mu_factor_e, double c_rho, double c_ener,
                       double p_ref, T *out)
{
    RhsArgs<T> A{};
    A.nelem = m->nelem;
    A.npoin = m->npoin;
    A.EN = (size_t)m->nelem * nnode_of(porder);
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dl;
    A.inv = inv;
    A.gI = gI;
    A.gJ = gJ;
    A.gK = gK;
    A.cr = rho;
    A.cq = q;
    A.cE = E;
    A.mu_e = mu_e;
    A.gamma = gamma;
    A.Rgas = Rgas;
    A.mu_factor = mu_factor;
    A.mu_factor_e = mu_factor_e;
    A.p_ref = p_ref;
    A.Cp = T(Cp);
    A.Pr = T(Pr);
    A.c_rho = T(c_rho);
    A.c_ener = T(c_ener);
    A.out = out;
    int tok = prof_begin(ctx, 0);
    int rc = SODHIP_OK;
#if SODHIP_SPLIT_RHS
    // p=3: one wave per operator (in FP64 a single wave holding both spills;
    // in FP32 the split is also faster); at p=2 a single wave holds both
    // operators with fewer spills, p=4 needs two waves per operator
    bool done = false;
    {
        if (porder == 3) {
            const size_t nn = (size_t)nnode_of(porder);
            A.geo16 = ctx->stage_geo16 && ((uintptr_t)A.He % 16 == 0) &&
                      ((uintptr_t)A.gpvol % 16 == 0) && (9 * nn * sizeof(T)) % 16 == 0 &&
                      (nn * sizeof(T)) % 16 == 0;
            const int grid = elem_grid(ctx, m->nelem, ctx->stage_bpc);
            if (A.geo16)
                k_rhs_split<3, T, true><<<grid, 128, 0, ctx->stream>>>(A);
            else
                k_rhs_split<3, T, false><<<grid, 128, 0, ctx->stream>>>(A);
            if (hipGetLastError() != hipSuccess)
                rc = set_error(SODHIP_ERR_HIP, "k_rhs_split launch failed");
            done = true;
        }
    }
    if (!done)
#endif
        rc = launch_rhs<T>(ctx, porder, true, true, true, A, elem_grid(ctx, m->nelem, 128));
    prof_end(ctx, tok);
    return rc;
}
int rhs_cluster_stage(sodhip_ctx ctx, const MeshCache *m, const ClusterTables *cl,
                      const double *He, const double *gpvol, const double *dl, const int *inv,
                      const int *gI, const int *gJ, const int *gK, const double *mu_e,
                      double gamma, double Rgas, double Cp, double Pr, double mu_factor,
                      const double *mu_factor_e, double c_rho, double c_ener, double p_ref,
                      const double *invMl, double fact, double dt, double lam1, double lams,
                      double gam, const double *rho1, const double *q1, const double *E1,
                      double *rho, double *q, double *E, double *Rmass, double *Rmom,
                      double *Rener, const NextStage *next, double *rec)
{
    using T = double;
    RhsArgs<T> A{};
    A.nelem = m->nelem;
    A.npoin = m->npoin;
    A.EN = (size_t)m->nelem * 64;
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dl;
    A.inv = inv;
    A.gI = gI;
    A.gJ = gJ;
    A.gK = gK;
    A.cr = rho;
    A.cq = q;
    A.cE = E;
    A.mu_e = mu_e;
    A.gamma = gamma;
    A.Rgas = Rgas;
    A.mu_factor = mu_factor;
    A.mu_factor_e = mu_factor_e;
    A.p_ref = p_ref;
    A.Cp = Cp;
    A.Pr = Pr;
    A.c_rho = c_rho;
    A.c_ener = c_ener;
    A.out = nullptr;
    A.cel = cl->cel;
    A.cloc = cl->cloc;
    A.cnb = cl->cnb;
    A.cnode = cl->cnode;
    A.ckind = cl->ckind;
    A.ncl = cl->ncl;
    A.rec = rec;
    A.invMl = invMl;
    A.fact = fact;
    A.lam1 = lam1;
    A.lams = lams;
    A.gdt = gam * dt;
    A.rd1 = lam1 != 0.0;
    A.rdy = lams != 0.0;
    A.y1r = rho1;
    A.y1q = q1;
    A.y1E = E1;
    A.yr = rho;
    A.yq = q;
    A.yE = E;
    A.Rm = Rmass;
    A.Rq = Rmom;
    A.Re = Rener;
    if (next) {
        A.nx_l1 = next->nl1;
        A.nx_gdt = next->ngdt;
        A.nxr = next->rho;
        A.nxq = next->q;
        A.nxE = next->E;
    }
    A.geo16 = ((uintptr_t)He % 16 == 0) && ((uintptr_t)gpvol % 16 == 0);
    // XCD-aware: every XCD's eighth of the clusters split over its blocks,
    // at most kClusterBlockMax clusters per block
    const long chunk = ((long)cl->ncl + 7) / 8;
    const long per = std::max<long>((chunk + kClusterBlockMax - 1) / kClusterBlockMax, 1);
    const int grid = (int)(8 * per);
    int tok = prof_begin(ctx, 0);
    if (A.geo16)
        k_rhs_split<3, T, true, true><<<grid, 128, 0, ctx->stream>>>(A);
    else
        k_rhs_split<3, T, false, true><<<grid, 128, 0, ctx->stream>>>(A);
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

// blocks of 128 threads of kernel k resident on one CU (at least 1)
template <typename K> static int lps_resident(K k)
{
    int r = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, k, 128, 0) != hipSuccess) r = 1;
    return std::max(r, 1);
}

int rhs_lps_elements(sodhip_ctx ctx, const MeshCache *m, const double *He, const double *gpvol,
                     const double *dl, const int *inv, const int *gI, const int *gJ, const int *gK,
                     const double *rho, const double *q, const double *E, const double *mu_e,
                     const double *prj, const double *tau, double gamma, double Rgas, double Cp,
                     double Pr, double mu_factor, double c_rho, double c_ener, double c_lps,
                     double *out)
{
    using T = double;
    RhsArgs<T> A{};
    A.nelem = m->nelem;
    A.npoin = m->npoin;
    A.EN = (size_t)m->nelem * 64;
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dl;
    A.inv = inv;
    A.gI = gI;
    A.gJ = gJ;
    A.gK = gK;
    A.cr = rho;
    A.cq = q;
    A.cE = E;
    A.mu_e = mu_e;
    A.gamma = gamma;
    A.Rgas = Rgas;
    A.mu_factor = mu_factor;
    A.p_ref = 0.0;
    A.Cp = Cp;
    A.Pr = Pr;
    A.c_rho = c_rho;
    A.c_ener = c_ener;
    A.prj = prj;
    A.tau = tau;
    A.c_lps = c_lps;
    A.out = out;
    A.geo16 = ctx->stage_geo16 && ((uintptr_t)He % 16 == 0) && ((uintptr_t)gpvol % 16 == 0);
    // the grid is a whole number of rounds of the blocks a CU holds at once:
    // the LPS mode's LDS leaves room for 3, where the plain kernel has 4 and
    // the stage's blocks-per-CU setting is tuned for it (16 blocks per CU
    // would leave a third of a round: 1.35 ms against 1.30 ms at 64^3)
    static const int res16 = lps_resident(k_rhs_split<3, T, true, false, true>);
    static const int res4 = lps_resident(k_rhs_split<3, T, false, false, true>);
    const int res = A.geo16 ? res16 : res4;
    const int grid = elem_grid(ctx, m->nelem, std::max(res, ctx->stage_bpc / res * res));
    int tok = prof_begin(ctx, 0);
    if (A.geo16)
        k_rhs_split<3, T, true, false, true><<<grid, 128, 0, ctx->stream>>>(A);
    else
        k_rhs_split<3, T, false, false, true><<<grid, 128, 0, ctx->stream>>>(A);
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

template int rhs_fused_elements<double>(sodhip_ctx, int, const MeshCache *, const double *,
                                        const double *, const double *, const int *, const int *,
                                        const int *, const int *, const double *, const double *,
                                        const double *, const double *, double, double, double,
                                        double, double, const double *, double, double, double,
                                        double *);
template int rhs_fused_elements<float>(sodhip_ctx, int, const MeshCache *, const float *,
                                       const float *, const float *, const int *, const int *,
                                       const int *, const int *, const double *, const double *,
                                       const double *, const float *, double, double, double,
                                       double, double, const double *, double, double, double,
                                       float *);

template <typename T>
static int scalar_convec(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
                         const T *He, const T *gpvol, const T *dl, const int *inv, const int *gI,
                         const int *gJ, const int *gK, const T *q, const T *eta, const T *u,
                         T *Rconvec, int initialize, T fact, const EntropyTail *tail = nullptr)
{
    int rc = check_common<T>(ctx, porder, nelem, npoin, connec);
    if (rc) return rc;
    SODHIP_REQUIRE(He && gpvol && dl && inv && gI && gJ && gK && eta && u && (Rconvec || tail),
                   "null array argument");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    if ((rc = get_mesh(ctx, porder, nelem, npoin, connec, &m))) return rc;
    const size_t EN = (size_t)nelem * nnode_of(porder);
    if ((rc = ensure_buf(ctx->scratch, EN * sizeof(T)))) return rc;
    RhsArgs<T> A{};
    A.nelem = nelem;
    A.npoin = npoin;
    A.EN = EN;
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dl;
    A.inv = inv;
    A.gI = gI;
    A.gJ = gJ;
    A.gK = gK;
    A.u = u;
    A.out = (T *)ctx->scratch.ptr;
    const int grid = elem_grid(ctx, nelem, 128);
    int tok = prof_begin(ctx, 3);
#define SODHIP_SC(PP)                                                                       \
    if (q)                                                                                  \
        k_scalar_convec<PP, T, false><<<grid, EG<PP>::NT, 0, ctx->stream>>>(A, q, eta);     \
    else                                                                                    \
        k_scalar_convec<PP, T, true><<<grid, EG<PP>::NT, 0, ctx->stream>>>(A, q, eta);
    switch (porder) {
    case 2: SODHIP_SC(2) break;
    case 3: SODHIP_SC(3) break;
    case 4: SODHIP_SC(4) break;
    }
#undef SODHIP_SC
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    if (tail) {
        if constexpr (std::is_same<T, double>::value) {
            if ((rc = get_working(ctx, m, tail->npoin_w, tail->lpoin_w))) return rc;
            const int bs = 256;
            const int grid = ((tail->npoin_w + bs - 1) / bs + 7) / 8 * 8;
            tok = prof_begin(ctx, 3);
            if (tail->npoin_w > 0)
                k_gather_entropy<<<grid, bs, 0, ctx->stream>>>(npoin, m->wrow, m->wsrc, A.out,
                                                               *tail, fact);
            prof_end(ctx, tok);
            SODHIP_TRY_HIP(hipGetLastError());
            return SODHIP_OK;
        }
        return set_error(SODHIP_ERR_UNSUPPORTED, "entropy residual in FP64 only");
    }
    return launch_gather<T, T>(ctx, m, A.out, EN, 1, Rconvec, nullptr, nullptr, nullptr, nullptr,
                               initialize, fact);
}

// eval_divergence / compute_vorticity / both in one element pass
template <typename T>
static int div_curl(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                    const int *connec, const int *lpoin_w, const int *lelpn, const T *He,
                    const T *gpvol, const T *dl, const int *inv, const int *gI, const int *gJ,
                    const int *gK, const T *u, T *divU, T *curlU)
{
    int rc = check_common<T>(ctx, porder, nelem, npoin, connec);
    if (rc) return rc;
    SODHIP_REQUIRE(He && dl && inv && gI && gJ && gK && u && (divU || curlU),
                   "null array argument");
    SODHIP_REQUIRE(!divU || gpvol, "null gpvol");
    SODHIP_REQUIRE(!curlU || !lelpn || (lpoin_w && npoin_w > 0 && npoin_w <= npoin),
                   "null lpoin_w or invalid npoin_w");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    if ((rc = get_mesh(ctx, porder, nelem, npoin, connec, &m))) return rc;
    const int nf = (divU ? 1 : 0) + (curlU ? 3 : 0);
    const size_t EN = (size_t)nelem * nnode_of(porder);
    if ((rc = ensure_buf(ctx->scratch, nf * EN * sizeof(T)))) return rc;
    RhsArgs<T> A{};
    A.nelem = nelem;
    A.npoin = npoin;
    A.EN = EN;
    A.connecT = m->connecT;
    A.perm = m->perm;
    A.He = He;
    A.gpvol = gpvol;
    A.dl = dl;
    A.inv = inv;
    A.gI = gI;
    A.gJ = gJ;
    A.gK = gK;
    A.u = u;
    A.out = (T *)ctx->scratch.ptr;
    const int grid = elem_grid(ctx, nelem, 128);
    int tok = prof_begin(ctx, 3);
#define SODHIP_DC(PP, DV, CU) \
    k_div_curl<PP, T, DV, CU><<<grid, EG<PP>::NT, 0, ctx->stream>>>(A)
#define SODHIP_DC_P(PP)                                                        \
    if (divU && curlU) SODHIP_DC(PP, true, true);                              \
    else if (divU) SODHIP_DC(PP, true, false);                                 \
    else SODHIP_DC(PP, false, true);
    switch (porder) {
    case 2: SODHIP_DC_P(2) break;
    case 3: SODHIP_DC_P(3) break;
    case 4: SODHIP_DC_P(4) break;
    }
#undef SODHIP_DC_P
#undef SODHIP_DC
    SODHIP_TRY_HIP(hipGetLastError());
    T *c0 = curlU, *c1 = curlU ? curlU + npoin : nullptr, *c2 = curlU ? curlU + 2 * (size_t)npoin : nullptr;
    if (divU && curlU)
        rc = launch_gather<T, T>(ctx, m, A.out, EN, 4, divU, c0, c1, c2, nullptr, 1, T(1));
    else if (divU)
        rc = launch_gather<T, T>(ctx, m, A.out, EN, 1, divU, nullptr, nullptr, nullptr, nullptr, 1,
                                 T(1));
    else
        rc = launch_gather<T, T>(ctx, m, A.out, EN, 3, c0, c1, c2, nullptr, nullptr, 1, T(1));
    if (rc == SODHIP_OK && curlU && lelpn) {
        k_per_elem_count<T><<<(npoin_w + 255) / 256, 256, 0, ctx->stream>>>(npoin, npoin_w, lpoin_w,
                                                                           lelpn, curlU);
        SODHIP_TRY_HIP(hipGetLastError());
    }
    prof_end(ctx, tok);
    return rc;
}

} // namespace sodhip

using namespace sodhip;

extern "C" int sodhip_full_convec_ijk_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                          const int *connec, const double *He,
                                          const double *gpvol, const double *dlxigp_ip,
                                          const int *invAtoIJK, const int *gmshAtoI,
                                          const int *gmshAtoJ, const int *gmshAtoK,
                                          const double *u, const double *q, const double *rho,
                                          const double *pr, const double *E, double *Rmass,
                                          double *Rmom, double *Rener, int initialize,
                                          double fact)
{
    return full_convec<double>(ctx, porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                               gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener,
                               initialize, fact);
}

extern "C" int sodhip_full_convec_ijk_f32(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                          const int *connec, const float *He, const float *gpvol,
                                          const float *dlxigp_ip, const int *invAtoIJK,
                                          const int *gmshAtoI, const int *gmshAtoJ,
                                          const int *gmshAtoK, const float *u, const float *q,
                                          const float *rho, const float *pr, const float *E,
                                          float *Rmass, float *Rmom, float *Rener, int initialize,
                                          float fact)
{
    return full_convec<float>(ctx, porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                              gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener,
                              initialize, fact);
}

extern "C" int sodhip_full_diffusion_ijk_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec, const double *He,
    const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK, const int *gmshAtoI,
    const int *gmshAtoJ, const int *gmshAtoK, double Cp, double Pr, const double *rho_n,
    const double *rho, const double *u, const double *Tem, const double *mu_fluid,
    const double *mu_e, const double *mu_sgs, double c_rho, double c_ener, double *Rmass,
    double *Rmom, double *Rener, int initialize, double fact)
{
    return full_diffusion<double>(ctx, porder, nelem, npoin, connec, He, gpvol, dlxigp_ip,
                                  invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho_n, rho, u,
                                  Tem, mu_fluid, mu_e, mu_sgs, c_rho, c_ener, Rmass, Rmom, Rener,
                                  initialize, fact);
}

extern "C" int sodhip_full_diffusion_ijk_f32(
    sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec, const float *He,
    const float *gpvol, const float *dlxigp_ip, const int *invAtoIJK, const int *gmshAtoI,
    const int *gmshAtoJ, const int *gmshAtoK, float Cp, float Pr, const float *rho_n,
    const float *rho, const float *u, const float *Tem, const float *mu_fluid, const float *mu_e,
    const float *mu_sgs, float c_rho, float c_ener, float *Rmass, float *Rmom, float *Rener,
    int initialize, float fact)
{
    return full_diffusion<float>(ctx, porder, nelem, npoin, connec, He, gpvol, dlxigp_ip,
                                 invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho_n, rho, u,
                                 Tem, mu_fluid, mu_e, mu_sgs, c_rho, c_ener, Rmass, Rmom, Rener,
                                 initialize, fact);
}

extern "C" int sodhip_generic_scalar_convec_ijk_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec, const double *He,
    const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK, const int *gmshAtoI,
    const int *gmshAtoJ, const int *gmshAtoK, const double *q, const double *eta,
    const double *u, double *Rconvec, int initialize, double fact)
{
    return scalar_convec<double>(ctx, porder, nelem, npoin, connec, He, gpvol, dlxigp_ip,
                                 invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, q, eta, u, Rconvec,
                                 initialize, fact);
}

extern "C" int sodhip_generic_scalar_convec_ijk_f32(
    sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec, const float *He,
    const float *gpvol, const float *dlxigp_ip, const int *invAtoIJK, const int *gmshAtoI,
    const int *gmshAtoJ, const int *gmshAtoK, const float *q, const float *eta, const float *u,
    float *Rconvec, int initialize, float fact)
{
    return scalar_convec<float>(ctx, porder, nelem, npoin, connec, He, gpvol, dlxigp_ip,
                                invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, q, eta, u, Rconvec,
                                initialize, fact);
}

extern "C" int sodhip_entropy_convec_residual_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w, const int *connec,
    const int *lpoin_w, const double *He, const double *gpvol, const double *dlxigp_ip,
    const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
    const double *u, const double *eta, const double *invMl, double factor_comp, double dt,
    double *Reta, double *auxReta)
{
    SODHIP_REQUIRE(lpoin_w && eta && invMl && Reta && auxReta, "null array argument");
    SODHIP_REQUIRE(npoin_w >= 0 && npoin_w <= npoin, "npoin_w outside 0..npoin");
    EntropyTail t{npoin_w, lpoin_w, invMl, eta, factor_comp, dt, Reta, auxReta};
    return scalar_convec<double>(ctx, porder, nelem, npoin, connec, He, gpvol, dlxigp_ip,
                                 invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, nullptr, eta, u,
                                 nullptr, 1, 1.0, &t);
}

extern "C" int sodhip_eval_divergence_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                          const int *connec, const double *He,
                                          const double *gpvol, const double *dlxigp_ip,
                                          const int *invAtoIJK, const int *gmshAtoI,
                                          const int *gmshAtoJ, const int *gmshAtoK,
                                          const double *u, double *divU)
{
    return div_curl<double>(ctx, porder, nelem, npoin, 0, connec, nullptr, nullptr, He, gpvol,
                            dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, u, divU, nullptr);
}

extern "C" int sodhip_compute_vorticity_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                            int npoin_w, const int *lpoin_w, const int *connec,
                                            const int *lelpn, const double *He,
                                            const double *dlxigp_ip, const int *invAtoIJK,
                                            const int *gmshAtoI, const int *gmshAtoJ,
                                            const int *gmshAtoK, const double *u, double *curlU)
{
    return div_curl<double>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w, lelpn, He,
                            nullptr, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, u,
                            nullptr, curlU);
}

extern "C" int sodhip_div_curl_f64(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                                   const int *lpoin_w, const int *connec, const int *lelpn,
                                   const double *He, const double *gpvol,
                                   const double *dlxigp_ip, const int *invAtoIJK,
                                   const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                                   const double *u, double *divU, double *curlU)
{
    SODHIP_REQUIRE(divU && curlU, "null output");
    return div_curl<double>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w, lelpn, He, gpvol,
                            dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, u, divU, curlU);
}

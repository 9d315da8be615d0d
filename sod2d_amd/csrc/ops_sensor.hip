// ops_sensor.hip -- entropy-viscosity sensor smart_visc_spectral
// (mod_entropy_viscosity.f90:14-111) with the intra-element 3x3x3 filter of
// init_filters (mod_filters.f90:15-55).
//
// The entropy normalisation (mean of eta over working nodes in FP64, then the
// maximum deviation) is a deterministic two-level reduction: every block of
// the first kernel sums a fixed slice in a fixed tree order; the second kernel
// re-reduces the partial sums in the same fixed order in every block (all
// blocks obtain the identical mean) and forms partial maxima; the element
// kernel folds the partial maxima (max is order independent).  No host
// synchronisation is needed between the three launches.
#include <algorithm>
#include <cfloat>

#include "common.hpp"
#include "device.hpp"
#include "sensor.hpp"

namespace sodhip {

constexpr int kRedThreads = 256;

template <typename T> __device__ __forceinline__ T block_sum(T v, T *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    T r = sh[0];
    __syncthreads();
    return r;
}

template <typename T> __device__ __forceinline__ T block_max(T v, T *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
        __syncthreads();
    }
    T r = sh[0];
    __syncthreads();
    return r;
}

// Block-wide sum of eta over the lpoin_w positions [i0, i1).  Four strided
// values are loaded ahead and added in the order of the plain strided loop,
// so the partial sums do not depend on the unrolling.
template <typename T>
__device__ __forceinline__ double eta_slice_sum(const int *__restrict__ lpoin_w,
                                                const T *__restrict__ eta, int i0, int i1,
                                                double *sh)
{
    const int bs = blockDim.x;
    double s = 0.0;
    int i = i0 + threadIdx.x;
    for (; i + 3 * bs < i1; i += 4 * bs) {
        const int n0 = lpoin_w[i], n1 = lpoin_w[i + bs], n2 = lpoin_w[i + 2 * bs],
                  n3 = lpoin_w[i + 3 * bs];
        const double v0 = (double)eta[n0 - 1], v1 = (double)eta[n1 - 1],
                     v2 = (double)eta[n2 - 1], v3 = (double)eta[n3 - 1];
        s += v0;
        s += v1;
        s += v2;
        s += v3;
    }
    for (; i < i1; i += bs) s += (double)eta[lpoin_w[i] - 1];
    return block_sum(s, sh);
}

// Block-wide max |eta - mean| over the lpoin_w positions [i0, i1), with mean
// the sum of the nparts partial sums part[] (in a fixed order) over nw.
template <typename T>
__device__ __forceinline__ double eta_slice_maxdev(const int *__restrict__ lpoin_w,
                                                   const T *__restrict__ eta,
                                                   const double *__restrict__ part, int nparts,
                                                   int nw, int i0, int i1, double *sh)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
    const double mean = block_sum(s, sh) / (double)nw;
    const int bs = blockDim.x;
    double m = 0.0;
    int i = i0 + threadIdx.x;
    for (; i + 3 * bs < i1; i += 4 * bs) {
        const int n0 = lpoin_w[i], n1 = lpoin_w[i + bs], n2 = lpoin_w[i + 2 * bs],
                  n3 = lpoin_w[i + 3 * bs];
        const double v0 = (double)eta[n0 - 1], v1 = (double)eta[n1 - 1],
                     v2 = (double)eta[n2 - 1], v3 = (double)eta[n3 - 1];
        m = fmax(fmax(m, fabs(v0 - mean)), fabs(v1 - mean));
        m = fmax(fmax(m, fabs(v2 - mean)), fabs(v3 - mean));
    }
    for (; i < i1; i += bs) m = fmax(m, fabs((double)eta[lpoin_w[i] - 1] - mean));
    return block_max(m, sh);
}

// partial sums of eta over working nodes: block b owns a contiguous slice
// (FP64 sums for either storage type)
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_eta_sum(int npoin_w, const int *__restrict__ lpoin_w,
                                                         const T *__restrict__ eta,
                                                         double *__restrict__ part)
{
    __shared__ double sh[kRedThreads];
    const int per = (npoin_w + gridDim.x - 1) / gridDim.x;
    const int i0 = blockIdx.x * per, i1 = min(npoin_w, i0 + per);
    const double s = eta_slice_sum(lpoin_w, eta, i0, i1, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// every block: mean = (sum of partial sums)/npoin_w in a fixed order, then its
// slice's max |eta - mean|
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_eta_maxdev(int npoin_w,
                                                            const int *__restrict__ lpoin_w,
                                                            const T *__restrict__ eta,
                                                            const double *__restrict__ part,
                                                            int nparts, double *__restrict__ pmax)
{
    __shared__ double sh[kRedThreads];
    const int per = (npoin_w + gridDim.x - 1) / gridDim.x;
    const int i0 = blockIdx.x * per, i1 = min(npoin_w, i0 + per);
    const double m = eta_slice_maxdev(lpoin_w, eta, part, nparts, npoin_w, i0, i1, sh);
    if (threadIdx.x == 0) pmax[blockIdx.x] = m;
}

// The same two passes per instance of a joined mesh: instance i owns the
// lpoin_w positions wseg[i]..wseg[i+1]-1 and the blocks i*nb..i*nb+nb-1,
// which split its segment as nb blocks split a single mesh's list (so the
// partials, mean and maximum have the bits of the instance run alone).
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_eta_sum_seg(const int *__restrict__ wseg, int nb,
                                                             const int *__restrict__ lpoin_w,
                                                             const T *__restrict__ eta,
                                                             double *__restrict__ part)
{
    __shared__ double sh[kRedThreads];
    const int inst = blockIdx.x / nb, b = blockIdx.x % nb;
    const int w0 = wseg[inst], nw = wseg[inst + 1] - w0;
    const int per = (nw + nb - 1) / nb;
    const int i0 = w0 + b * per, i1 = min(w0 + nw, i0 + per);
    const double s = eta_slice_sum(lpoin_w, eta, i0, i1, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_eta_maxdev_seg(const int *__restrict__ wseg,
                                                                int nb,
                                                                const int *__restrict__ lpoin_w,
                                                                const T *__restrict__ eta,
                                                                const double *__restrict__ part,
                                                                double *__restrict__ pmax)
{
    __shared__ double sh[kRedThreads];
    const int inst = blockIdx.x / nb, b = blockIdx.x % nb;
    const int w0 = wseg[inst], nw = wseg[inst + 1] - w0;
    const int per = (nw + nb - 1) / nb;
    const int i0 = w0 + b * per, i1 = min(w0 + nw, i0 + per);
    const double m = eta_slice_maxdev(lpoin_w, eta, part + (size_t)inst * nb, nb, nw, i0, i1, sh);
    if (threadIdx.x == 0) pmax[blockIdx.x] = m;
}

// per instance: the maximum of its nb partial maxima (any order gives the
// same maximum)
__global__ __launch_bounds__(64) void k_norm_seg(int nb, const double *__restrict__ pmax,
                                                 double *__restrict__ norm_i)
{
    double m = 0.0;
    for (int i = threadIdx.x; i < nb; i += 64) m = fmax(m, pmax[(size_t)blockIdx.x * nb + i]);
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) m = fmax(m, __shfl_xor(m, s, 64));
    if (threadIdx.x == 0) norm_i[blockIdx.x] = m;
}

template <typename T> __device__ __forceinline__ T block_min(T v, T *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = fmin(sh[t], sh[t + s]);
        __syncthreads();
    }
    T r = sh[0];
    __syncthreads();
    return r;
}

// partial minima of rho over working nodes (contiguous slices), starting
// from the reference's 100000
__global__ __launch_bounds__(kRedThreads) void k_rho_min(int npoin_w, const int *__restrict__ lpoin_w,
                                                         const double *__restrict__ rho,
                                                         double *__restrict__ part)
{
    __shared__ double sh[kRedThreads];
    const int per = (npoin_w + gridDim.x - 1) / gridDim.x;
    const int i0 = blockIdx.x * per, i1 = min(npoin_w, i0 + per);
    double m = 100000.0;
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) m = fmin(m, rho[lpoin_w[i] - 1]);
    m = block_min(m, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// smart_visc_spectral_imex (mod_entropy_viscosity.f90:113-227), one thread per
// element over the reference's element-fastest connec, mu_e and mue_l (for a
// node position the lanes of a wave read and write consecutive entries):
// mue_l <- incoming mu_e; beta_e = max over the nodes of rho helem_k (cmax/p)
// (|u| + fact c); mu_e = cglob min(ce rho |Reta|/norm (helem_k/p)^2, beta_e).
// norm: 1, or the max of the partials (max |eta - mean|), or the min of the
// partials (min rho), folded per block.  The nodes are visited an i-line at a
// time (build_alex) with all loads of a line issued together.
struct ImexArgs {
    int nelem, nnode, porder;
    const int *connec, *perm;
    const double *Reta, *rho, *u, *csound, *helem_k;
    int npoin;
    double ce, cmax, cglob, fact_low_mach;
    const double *part; // partials of the norm, or nullptr (norm = 1)
    int nparts, norm_min;
    double *mu_e, *mue_l;
};

__global__ __launch_bounds__(256) void k_visc_imex(ImexArgs A)
{
    __shared__ double sh[256];
    __shared__ int alex[125];
    build_alex(A.perm, A.nnode, alex);
    double norm = 1.0;
    if (A.part) {
        double v = A.norm_min ? 100000.0 : 0.0;
        for (int i = threadIdx.x; i < A.nparts; i += blockDim.x)
            v = A.norm_min ? fmin(v, A.part[i]) : fmax(v, A.part[i]);
        norm = A.norm_min ? block_min(v, sh) : block_max(v, sh);
    }
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= A.nelem) return;
    const size_t NE = (size_t)A.nelem;
    const double *u1 = A.u + A.npoin, *u2 = A.u + 2 * (size_t)A.npoin;
    const double hk = A.helem_k[e], cp = A.cmax / (double)A.porder;
    const int KA = A.porder + 1;
    double betae = 0.0;
    for (int l0 = 0; l0 < A.nnode; l0 += KA) {
        size_t ea[5];
        int n[5];
        double mu[5], v0[5], v1[5], v2[5], cs[5], rh[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            ea[j] = e + NE * alex[min(l0 + min(j, KA - 1), A.nnode - 1)];
            n[j] = A.connec[ea[j]] - 1;
            mu[j] = A.mu_e[ea[j]];
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            v0[j] = A.u[n[j]];
            v1[j] = u1[n[j]];
            v2[j] = u2[n[j]];
            cs[j] = A.csound[n[j]];
            rh[j] = A.rho[n[j]];
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            if (j < KA) A.mue_l[ea[j]] = mu[j];
            const double speed =
                sqrt(v0[j] * v0[j] + v1[j] * v1[j] + v2[j] * v2[j]) + A.fact_low_mach * cs[j];
            betae = fmax(betae, (rh[j] * hk) * cp * speed);
        }
    }
    const double hp = hk / (double)A.porder;
    for (int l0 = 0; l0 < A.nnode; l0 += KA) {
        size_t ea[5];
        int n[5];
        double rh[5], re[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            ea[j] = e + NE * alex[min(l0 + min(j, KA - 1), A.nnode - 1)];
            n[j] = A.connec[ea[j]] - 1;
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            rh[j] = A.rho[n[j]];
            re[j] = A.Reta[n[j]];
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            if (j >= KA) break;
            const double R1 = rh[j] * fabs(re[j]) / norm;
            const double Ve = A.ce * R1 * (hp * hp);
            A.mu_e[ea[j]] = A.cglob * fmin(Ve, betae);
        }
    }
}

// ducros_visc_limit (mod_entropy_viscosity.f90:277-314) in two passes.  Of
// the limiter's inputs only helem and mu_e are per element-node; the rest
// depends on the node alone.  A node pass forms, per node, d2 = (divU/Ml)^2,
// phi = min(4/3 d2 / (d2 + |curl u|^2 + eps), 1) and |u| (three arrays, read
// and written in streams); the element pass gathers those three instead of
// the eight node fields (eight gather streams per node overran the caches:
// about 3.7 times the algorithmic bytes were fetched).  Same expressions,
// same bits as one pass.
__global__ __launch_bounds__(256) void k_ducros_node(int npoin, const double *__restrict__ u,
                                                     const double *__restrict__ divU,
                                                     const double *__restrict__ vorti,
                                                     const double *__restrict__ Ml,
                                                     double *__restrict__ nq)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= npoin) return;
    const size_t P1 = (size_t)npoin, P2 = 2 * (size_t)npoin;
    const double dv = divU[n] / Ml[n];
    const double w0 = vorti[n], w1 = vorti[n + P1], w2 = vorti[n + P2];
    const double v0 = u[n], v1 = u[n + P1], v2 = u[n + P2];
    const double d2 = dv * dv;
    nq[n] = fmin((4.0 / 3.0) * d2 / (d2 + (w0 * w0 + w1 * w1 + w2 * w2) + DBL_EPSILON), 1.0);
    nq[n + P1] = d2;
    nq[n + P2] = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
}

// the element pass: one thread per element over the reference's
// element-fastest connec, helem and mu_e, the nodes visited an i-line at a
// time (build_alex) with all loads of a line issued together
__global__ __launch_bounds__(256) void k_ducros(int nelem, int nnode, int porder, int npoin,
                                                const int *__restrict__ connec,
                                                const int *__restrict__ perm,
                                                const double *__restrict__ nq,
                                                const double *__restrict__ helem, double dmin,
                                                double *__restrict__ mu_e)
{
    __shared__ int alex[125];
    build_alex(perm, nnode, alex);
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    const size_t P1 = (size_t)npoin, P2 = 2 * (size_t)npoin;
    const int KA = porder + 1;
    for (int l0 = 0; l0 < nnode; l0 += KA) {
        size_t t[5];
        int n[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            t[j] = (size_t)alex[min(l0 + min(j, KA - 1), nnode - 1)] * nelem + e;
            n[j] = connec[t[j]] - 1;
        }
        double phi[5], d2[5], um[5], he[5], mu[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            phi[j] = nq[n[j]];
            d2[j] = nq[n[j] + P1];
            um[j] = nq[n[j] + P2];
            he[j] = helem[t[j]];
            mu[j] = mu_e[t[j]];
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            if (j >= KA) break;
            const double rr = 1.0 * um[j] / he[j];
            const double phi2 = d2[j] / (d2[j] + rr * rr + DBL_EPSILON);
            mu_e[t[j]] = mu[j] * (fmax(phi[j] - dmin, 0.0) * phi2 + dmin);
        }
    }
}

// Chunked variant: the workgroup (NT = roundup((p+1)^3, 64) lanes, one or two
// waves) walks CH consecutive elements at a time so that the element-fastest
// arrays helem(nelem,nnode), mue_l and mu_e move as whole 128-byte lines: the
// CH x NN block is read and written through LDS with thread t holding (node
// t/CH, element t%CH), while the per-element arithmetic keeps lane = node.
template <int P, int CH, typename S = double>
__global__ __launch_bounds__(SG<P>::NT) void k_smart_visc_ch(SensorArgsT<S> A)
{
    constexpr int N1 = SG<P>::N1, NN = SG<P>::NN, NT = SG<P>::NT, NW = NT / 64;
    constexpr int BLK = CH * NN, ROWS = (BLK + NT - 1) / NT;
    __shared__ double lb[BLK]; // helem, then mue_l, of the chunk: [c][a]
    __shared__ double ob[BLK]; // mu_e of the chunk
    __shared__ double wmax[NW];
    const int lane = threadIdx.x, wl = lane & 63, wv = lane >> 6;
    const bool act = lane < NN;
    const int a = act ? lane : 0;
    double norm = 1.0;
    if (A.pmax) { // every wave folds all partial maxima itself
        norm = 0.0;
        for (int i = wl; i < A.nparts; i += 64) norm = fmax(norm, A.pmax[i]);
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) norm = fmax(norm, __shfl_xor(norm, s, 64));
    }
    int src[27];
    filter_sources<P>(a, A.inv, src);
    const double w[3] = {0.25, 0.5, 0.25};
    // XCD-aware: the blocks of one XCD walk chunks of one contiguous eighth
    const int per = gridDim.x >> 3, xcd = blockIdx.x & 7, rk = blockIdx.x >> 3;
    const int span = (A.nelem + 7) >> 3;
    const int xb = xcd * span, xe = min(A.nelem, xb + span);
    const int nch = xe > xb ? (xe - xb + CH - 1) / CH : 0;
    if (rk >= nch) return;
    struct In {
        double u0, u1, u2, T, rho, Reta, hk;
    };
    auto load = [&](In &x, int e) {
        const int n = act ? A.connecT[(size_t)e * NN + a] : 0;
        x.u0 = A.u[n];
        x.u1 = A.u[n + A.npoin];
        x.u2 = A.u[n + 2 * (size_t)A.npoin];
        x.T = A.Tem[n];
        x.rho = A.rho[n];
        x.Reta = A.Reta[n];
        x.hk = A.helem_k[e];
    };
    In cur, nxt;
    load(cur, xb + rk * CH);
    for (int ci = rk; ci < nch; ci += per) {
        const int c0 = xb + ci * CH, cn = min(CH, xe - c0);
        const int cnext = ci + per < nch ? xb + (ci + per) * CH : c0 + cn - 1;
        // helem of the chunk, whole lines (elements past the range re-read the last)
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int t = r * NT + lane;
            if (t < BLK) {
                const int an = t / CH, c = t % CH;
                lb[c * NN + an] = A.helem[(c0 + min(c, cn - 1)) + (size_t)A.nelem * an];
            }
        }
        lds_barrier();
        for (int c = 0; c < cn; ++c) {
            const int e = c0 + c;
            load(nxt, c + 1 < cn ? e + 1 : cnext);
            const double aux2 = sqrt(cur.u0 * cur.u0 + cur.u1 * cur.u1 + cur.u2 * cur.u2);
            const double aux3 = A.fact_low_mach * sqrt(A.Rgas * A.gamma * cur.T);
            double be = act ? (cur.rho * cur.hk) * (A.cmax / (double)P) * (aux2 + aux3) : 0.0;
#pragma unroll
            for (int s = 32; s > 0; s >>= 1) be = fmax(be, __shfl_xor(be, s, 64));
            if constexpr (NW > 1) {
                if (wl == 0) wmax[wv] = be;
                lds_barrier();
                be = wmax[0];
#pragma unroll
                for (int w2 = 1; w2 < NW; ++w2) be = fmax(be, wmax[w2]);
            }
            const double nrm = A.norm_i ? A.norm_i[A.elem_inst[e]] : norm;
            const double R1 = cur.rho * fabs(cur.Reta) / nrm;
            const double h = lb[c * NN + a];
            const double Ve = A.ce * R1 * (h * h);
            const double ml = A.cglob * fmin(Ve, be);
            if (act) lb[c * NN + a] = ml; // the lane's own h slot
            lds_barrier();
            double acc = 0.0;
#pragma unroll
            for (int l = 0; l < 3; ++l)
#pragma unroll
                for (int m = 0; m < 3; ++m)
#pragma unroll
                    for (int nn = 0; nn < 3; ++nn)
                        acc += w[l] * w[m] * w[nn] * lb[c * NN + src[9 * l + 3 * m + nn]];
            if (act) ob[c * NN + a] = acc;
            cur = nxt;
        }
        lds_barrier();
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int t = r * NT + lane;
            if (t < BLK) {
                const int an = t / CH, c = t % CH;
                if (c < cn) {
                    const size_t g = (c0 + c) + (size_t)A.nelem * an;
                    A.mue_l[g] = lb[c * NN + an];
                    A.mu_e[g] = ob[c * NN + an];
                }
            }
        }
        lds_barrier(); // the next chunk overwrites lb and ob
    }
}

// The entropy normalisation's partial maxima of |eta - mean| over the working
// nodes (kRedBlocks values in ctx->scratch2), or nullptr when the sensor does
// not normalise; for kernels that fold them as the sensor does.
int entropy_norm_partials(sodhip_ctx ctx, int npoin_w, const int *lpoin_w, const double *eta,
                          bool normalise, const double **pmax, int *nparts)
{
    *pmax = nullptr;
    *nparts = kRedBlocks;
    if (!normalise) return SODHIP_OK;
    int rc = ensure_buf(ctx->scratch2, 2 * kRedBlocks * sizeof(double));
    if (rc) return rc;
    double *part = (double *)ctx->scratch2.ptr;
    k_eta_sum<double><<<kRedBlocks, kRedThreads, 0, ctx->stream>>>(npoin_w, lpoin_w, eta, part);
    k_eta_maxdev<double><<<kRedBlocks, kRedThreads, 0, ctx->stream>>>(npoin_w, lpoin_w, eta, part,
                                                              kRedBlocks, part + kRedBlocks);
    SODHIP_TRY_HIP(hipGetLastError());
    *pmax = part + kRedBlocks;
    return SODHIP_OK;
}

int entropy_norm_instances(sodhip_ctx ctx, const sodhip_instances *ens, const int *lpoin_w,
                           const double *eta, const double **norm_i)
{
    SODHIP_REQUIRE(ens && ens->ninst > 0 && ens->wseg && ens->elem_inst, "invalid instances");
    const size_t np = (size_t)ens->ninst * kRedBlocks;
    int rc = ensure_buf(ctx->scratch2, (2 * np + ens->ninst) * sizeof(double));
    if (rc) return rc;
    double *part = (double *)ctx->scratch2.ptr, *pmax = part + np, *nrm = pmax + np;
    const int grid = ens->ninst * kRedBlocks;
    k_eta_sum_seg<double><<<grid, kRedThreads, 0, ctx->stream>>>(ens->wseg, kRedBlocks, lpoin_w,
                                                                 eta, part);
    k_eta_maxdev_seg<double><<<grid, kRedThreads, 0, ctx->stream>>>(ens->wseg, kRedBlocks,
                                                                    lpoin_w, eta, part, pmax);
    k_norm_seg<<<ens->ninst, 64, 0, ctx->stream>>>(kRedBlocks, pmax, nrm);
    SODHIP_TRY_HIP(hipGetLastError());
    *norm_i = nrm;
    return SODHIP_OK;
}

} // namespace sodhip

using namespace sodhip;

template <typename T>
static int smart_visc(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                      const int *connec, const int *lpoin_w, const T *Reta, double gamma_gas,
                      const T *rho, const T *u, const T *Tem, const T *eta, const T *helem,
                      const T *helem_k, const int *invAtoIJK, const sodhip_visc_params *prm,
                      T *mu_e, T *mue_l, const sodhip_instances *ens = nullptr)
{
    SODHIP_REQUIRE(ctx && prm, "null argument");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_REQUIRE(nelem > 0 && npoin > 0 && npoin_w > 0 && npoin_w <= npoin,
                   "invalid mesh sizes");
    SODHIP_REQUIRE(connec && lpoin_w && Reta && rho && u && Tem && eta && helem && helem_k &&
                       invAtoIJK && mu_e && mue_l,
                   "null array argument");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    int rc = get_mesh(ctx, porder, nelem, npoin, connec, &m);
    if (rc) return rc;
    if ((rc = ensure_buf(ctx->scratch2, 2 * kRedBlocks * sizeof(double)))) return rc;
    double *part = (double *)ctx->scratch2.ptr;
    double *pmax = part + kRedBlocks;
    const bool normalise = prm->flag_normalise_entropy == 1;
    int tok = prof_begin(ctx, 2);
    const double *norm_i = nullptr;
    if (normalise && ens) {
        if constexpr (sizeof(T) == 8) {
            if ((rc = entropy_norm_instances(ctx, ens, lpoin_w, eta, &norm_i))) return rc;
        }
    } else if (normalise) {
        k_eta_sum<T><<<kRedBlocks, kRedThreads, 0, ctx->stream>>>(npoin_w, lpoin_w, eta, part);
        k_eta_maxdev<T><<<kRedBlocks, kRedThreads, 0, ctx->stream>>>(npoin_w, lpoin_w, eta, part,
                                                                     kRedBlocks, pmax);
    }
    SensorArgsT<T> A{};
    A.nelem = nelem;
    A.npoin = npoin;
    A.connecT = m->connecT;
    A.Reta = Reta;
    A.rho = rho;
    A.u = u;
    A.Tem = Tem;
    A.helem = helem;
    A.helem_k = helem_k;
    A.inv = invAtoIJK;
    A.gamma = gamma_gas;
    A.Rgas = prm->Rgas;
    A.ce = prm->ce;
    A.cmax = prm->cmax;
    A.cglob = prm->cglob;
    A.fact_low_mach = prm->flag_drop_c_in_envit ? 0.0 : 1.0;
    A.pmax = normalise && !ens ? pmax : nullptr;
    A.nparts = kRedBlocks;
    A.norm_i = norm_i;
    A.elem_inst = norm_i ? ens->elem_inst : nullptr;
    A.mu_e = mu_e;
    A.mue_l = mue_l;
    switch (porder) {
    case 2: k_smart_visc_ch<2, 16, T><<<elem_grid(ctx, (nelem + 15) / 16, 8), 64, 0, ctx->stream>>>(A); break;
    case 3: k_smart_visc_ch<3, 16, T><<<elem_grid(ctx, (nelem + 15) / 16, 8), 64, 0, ctx->stream>>>(A); break;
    case 4: k_smart_visc_ch<4, 16, T><<<elem_grid(ctx, (nelem + 15) / 16, 8), SG<4>::NT, 0, ctx->stream>>>(A); break;
    }
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

extern "C" int sodhip_smart_visc_spectral_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                              int npoin_w, const int *connec, const int *lpoin_w,
                                              const double *Reta, double gamma_gas,
                                              const double *rho, const double *u,
                                              const double *Tem, const double *eta,
                                              const double *helem, const double *helem_k,
                                              const int *invAtoIJK, const sodhip_visc_params *prm,
                                              double *mu_e, double *mue_l)
{
    return smart_visc<double>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w, Reta, gamma_gas,
                              rho, u, Tem, eta, helem, helem_k, invAtoIJK, prm, mu_e, mue_l);
}

extern "C" int sodhip_smart_visc_spectral_ens_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w, const int *connec,
    const int *lpoin_w, const sodhip_instances *ens, const double *Reta, double gamma_gas,
    const double *rho, const double *u, const double *Tem, const double *eta,
    const double *helem, const double *helem_k, const int *invAtoIJK,
    const sodhip_visc_params *prm, double *mu_e, double *mue_l)
{
    SODHIP_REQUIRE(ens && ens->ninst > 0 && ens->wseg && ens->elem_inst, "invalid instances");
    return smart_visc<double>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w, Reta, gamma_gas,
                              rho, u, Tem, eta, helem, helem_k, invAtoIJK, prm, mu_e, mue_l, ens);
}

extern "C" int sodhip_smart_visc_spectral_f32(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                              int npoin_w, const int *connec, const int *lpoin_w,
                                              const float *Reta, float gamma_gas,
                                              const float *rho, const float *u, const float *Tem,
                                              const float *eta, const float *helem,
                                              const float *helem_k, const int *invAtoIJK,
                                              const sodhip_visc_params *prm, float *mu_e,
                                              float *mue_l)
{
    return smart_visc<float>(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w, Reta,
                             (double)gamma_gas, rho, u, Tem, eta, helem, helem_k, invAtoIJK, prm,
                             mu_e, mue_l);
}

extern "C" int sodhip_smart_visc_spectral_imex_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w, const int *connec,
    const int *lpoin_w, const double *Reta, const double *rho, const double *u,
    const double *csound, const double *eta, const double *helem_k, sodhip_visc_params *prm,
    double ce_comp, int flag_high_mach, double *mu_e, double *mue_l)
{
    SODHIP_REQUIRE(ctx && prm, "null argument");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_REQUIRE(nelem > 0 && npoin > 0 && npoin_w > 0 && npoin_w <= npoin,
                   "invalid mesh sizes");
    SODHIP_REQUIRE(connec && lpoin_w && Reta && rho && u && csound && eta && helem_k && mu_e &&
                       mue_l,
                   "null array argument");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    int rc = get_mesh(ctx, porder, nelem, npoin, connec, &m);
    if (rc) return rc;
    if ((rc = ensure_buf(ctx->scratch2, 2 * kRedBlocks * sizeof(double)))) return rc;
    double *part = (double *)ctx->scratch2.ptr;
    double *pmax = part + kRedBlocks;
    const bool normalise = prm->flag_normalise_entropy == 1;
    const bool high_mach = normalise && flag_high_mach;
    int tok = prof_begin(ctx, 2);
    if (high_mach) {
        k_rho_min<<<kRedBlocks, kRedThreads, 0, ctx->stream>>>(npoin_w, lpoin_w, rho, pmax);
        // the reference assigns its module global ce (kept for later calls)
        prm->ce = ce_comp;
    } else if (normalise) {
        k_eta_sum<double><<<kRedBlocks, kRedThreads, 0, ctx->stream>>>(npoin_w, lpoin_w, eta, part);
        k_eta_maxdev<double><<<kRedBlocks, kRedThreads, 0, ctx->stream>>>(npoin_w, lpoin_w, eta, part,
                                                                  kRedBlocks, pmax);
    }
    ImexArgs A{};
    A.nelem = nelem;
    A.nnode = (porder + 1) * (porder + 1) * (porder + 1);
    A.porder = porder;
    A.connec = connec;
    A.perm = m->perm;
    A.Reta = Reta;
    A.rho = rho;
    A.u = u;
    A.csound = csound;
    A.helem_k = helem_k;
    A.npoin = npoin;
    A.ce = prm->ce;
    A.cmax = prm->cmax;
    A.cglob = prm->cglob;
    A.fact_low_mach = prm->flag_drop_c_in_envit ? 0.0 : 1.0;
    A.part = normalise ? pmax : nullptr;
    A.nparts = kRedBlocks;
    A.norm_min = high_mach ? 1 : 0;
    A.mu_e = mu_e;
    A.mue_l = mue_l;
    k_visc_imex<<<tpe_grid(nelem), 256, 0, ctx->stream>>>(A);
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

extern "C" int sodhip_ducros_visc_limit_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                            const int *connec, const double *u,
                                            const double *divU, const double *vorti,
                                            const double *Ml, const double *helem,
                                            double ducros_min_val, double *mu_e)
{
    SODHIP_REQUIRE(ctx, "null context");
    SODHIP_REQUIRE(porder >= 2 && porder <= kMaxP, "porder %d not supported (2..4)", porder);
    SODHIP_REQUIRE(nelem > 0 && npoin > 0, "invalid mesh sizes");
    SODHIP_REQUIRE(connec && u && divU && vorti && Ml && helem && mu_e, "null array argument");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    int rc = get_mesh(ctx, porder, nelem, npoin, connec, &m);
    if (rc) return rc;
    // the node pass's three arrays in the element-node scratch (3 npoin doubles,
    // far below the stage's records; the limiter runs between stages)
    if ((rc = ensure_buf(ctx->scratch, 3 * (size_t)npoin * sizeof(double)))) return rc;
    double *nq = (double *)ctx->scratch.ptr;
    int tok = prof_begin(ctx, 2);
    k_ducros_node<<<(npoin + 255) / 256, 256, 0, ctx->stream>>>(npoin, u, divU, vorti, Ml, nq);
    k_ducros<<<tpe_grid(nelem), 256, 0, ctx->stream>>>(nelem, nnode_of(porder), porder, npoin,
                                                       connec, m->perm, nq, helem, ducros_min_val,
                                                       mu_e);
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

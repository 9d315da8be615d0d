// sensor.hpp -- shared pieces of the entropy-viscosity sensor kernels
// (smart_visc_spectral, mod_entropy_viscosity.f90:14-111): argument block and
// the intra-element 3x3x3 filter of init_filters (mod_filters.f90:15-55).
#pragma once

#include "common.hpp"

namespace sodhip {

template <int P> struct SG {
    static constexpr int N1 = P + 1;
    static constexpr int NN = N1 * N1 * N1;
    static constexpr int NT = (NN + 63) / 64 * 64;
};

// the sensor's arrays in the storage type T (double, or float for the FP32
// interface); the arithmetic is FP64 either way
template <typename T> struct SensorArgsT {
    int nelem, npoin;
    const int *connecT;
    const T *Reta, *rho, *u, *Tem, *helem, *helem_k;
    const int *inv;
    double gamma, Rgas, ce, cmax, cglob, fact_low_mach;
    const double *pmax; // partial maxima (nparts) or nullptr when not normalising
    int nparts;
    // joined instances: the normalisation of element e is norm_i[elem_inst[e]]
    // (norm_i nullptr: the one normalisation above)
    const double *norm_i;
    const int *elem_inst;
    T *mu_e, *mue_l;
};
using SensorArgs = SensorArgsT<double>;

// sod2d 1-D index s (1-based) -> geometric position along the line (1-based)
__device__ __forceinline__ int geo_of(int s, int N1) { return s == 1 ? 1 : (s == 2 ? N1 : s - 1); }

// convertIJK(0:p+2): geometric position (with one ghost each side) -> sod2d
// index, reflecting 0 -> 3 and p+2 -> p (mod_filters.f90:44-50)
__device__ __forceinline__ int convert_ijk(int g, int P)
{
    if (g == 0) return 3;
    if (g == 1) return 1;
    if (g == P + 2) return P;
    if (g == P + 1) return 2;
    return g + 1; // 2..P -> 3..P+1
}

// The 27 element-local sources of node a's filter value: the node at
// geometric offset (l-1, m-1, n-1) of a, reflected at the element boundary by
// convertIJK, in the reference's (l, m, n) loop order.
template <int P>
__device__ __forceinline__ void filter_sources(int a, const int *inv, int src[27])
{
    constexpr int N1 = SG<P>::N1, NN = SG<P>::NN;
    int I = 1, J = 1, K = 1;
    for (int k = 0; k < N1; ++k)
        for (int j = 0; j < N1; ++j)
            for (int i = 0; i < N1; ++i)
                if (inv[i + N1 * (j + N1 * k)] - 1 == a) {
                    I = i + 1;
                    J = j + 1;
                    K = k + 1;
                }
    const int ii = geo_of(I, N1), jj = geo_of(J, N1), kk = geo_of(K, N1);
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const int ci = convert_ijk(ii + l - 1, P) - 1, cj = convert_ijk(jj + m - 1, P) - 1,
                          ck = convert_ijk(kk + n - 1, P) - 1;
                src[9 * l + 3 * m + n] = min(max(inv[ci + N1 * (cj + N1 * ck)] - 1, 0), NN - 1);
            }
}

// blocks of the entropy normalisation's reductions: its partial maxima take
// kRedBlocks doubles after kRedBlocks partial sums at the start of ctx->scratch2
constexpr int kRedBlocks = 512;

// launches the entropy normalisation's reductions (ops_sensor.hip)
int entropy_norm_partials(sodhip_ctx ctx, int npoin_w, const int *lpoin_w, const double *eta,
                          bool normalise, const double **pmax, int *nparts);
// the normalisation of every instance of a joined mesh into *norm_i (ninst
// doubles in ctx->scratch2 after the partials); ops_sensor.hip
int entropy_norm_instances(sodhip_ctx ctx, const sodhip_instances *ens, const int *lpoin_w,
                           const double *eta, const double **norm_i);

} // namespace sodhip

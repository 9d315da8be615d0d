// element.hpp -- element-level helpers shared by the element kernels: element
// sizes, the per-lane GLL line tables and the XCD-aware element ranges.
#pragma once

#include "common.hpp"

namespace sodhip {

template <int P> struct EG {
    static constexpr int N1 = P + 1;
    static constexpr int NN = N1 * N1 * N1;
    static constexpr int NT = (NN + 63) / 64 * 64;
};

// Per-lane line tables: the P other nodes of the lane's three GLL lines.
template <int P, typename T> struct LineTab {
    int off[3][P];
    T ds[3], d[3][P]; // d/dxi_k at the lane's point: own and other coefficients
    T t[3][P];        // transposed: d/dxi_k of the lane's basis function at the
                      // other nodes (weak divergence)
};

template <int P, typename T>
__device__ __forceinline__ void build_linetab(int a, const T *__restrict__ dl,
                                              const int *__restrict__ inv,
                                              const int *__restrict__ gI,
                                              const int *__restrict__ gJ,
                                              const int *__restrict__ gK, LineTab<P, T> &t)
{
    constexpr int N1 = EG<P>::N1, NN = EG<P>::NN;
    // indices are clamped so that malformed tables cannot address outside the
    // element's arrays
    const int own[3] = {min(max(gI[a] - 1, 0), P), min(max(gJ[a] - 1, 0), P),
                        min(max(gK[a] - 1, 0), P)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        t.ds[k] = dl[a + NN * (k + 3 * own[k])];
#pragma unroll
        for (int o = 0; o < P; ++o) {
            const int m = o + (o >= own[k] ? 1 : 0);
            int idx[3] = {own[0], own[1], own[2]};
            idx[k] = m;
            const int b = min(max(inv[idx[0] + N1 * (idx[1] + N1 * idx[2])] - 1, 0), NN - 1);
            t.off[k][o] = b;
            t.d[k][o] = dl[a + NN * (k + 3 * m)];
            t.t[k][o] = dl[b + NN * (k + 3 * own[k])];
        }
    }
}

// XCD-aware grid-stride element range of this block.
struct ElemRange {
    int begin, end, stride;
};
__device__ __forceinline__ ElemRange elem_range(int nelem)
{
    const int nblk = gridDim.x, b = blockIdx.x;
    const int per = nblk >> 3; // grid is a multiple of 8
    const int xcd = b & 7, r = b >> 3;
    const int chunk = (nelem + 7) >> 3;
    int e0 = xcd * chunk;
    int e1 = min(nelem, e0 + chunk);
    return {e0 + r, e1, per};
}

// Loads the 5-value FP64 record sj (40 bytes at 40*sj): it starts on a 16-byte
// boundary (even sj) or 8 bytes past one (odd sj), so two aligned 16-byte
// loads and one 8-byte load at per-lane offsets cover it either way.
__device__ __forceinline__ void load_rec5(const double *__restrict__ scr, int sj, double v[5])
{
    const double *r = scr + 5 * (size_t)sj;
    const int odd = sj & 1;
    const double2 x = *(const double2 *)(r + odd);
    const double2 y = *(const double2 *)(r + 2 + odd);
    const double z = r[odd ? 0 : 4];
    v[0] = odd ? z : x.x;
    v[1] = odd ? x.x : x.y;
    v[2] = odd ? x.y : y.x;
    v[3] = odd ? y.x : y.y;
    v[4] = odd ? y.y : z;
}

} // namespace sodhip

// ops_mlp.hip -- a per-element MLP that stands in for the entropy-viscosity
// sensor smart_visc_spectral (mod_entropy_viscosity.f90:14-111), with the
// plain sensor as its per-element fallback.
//
//  * features: twelve dimensionless statistics per element of the quantities
//    the plain sensor reads (the local viscosity ce rho |Reta| / norm h^2
//    relative to the element's cap beta_e, Mach number, density and
//    temperature spread), taken on 16 of the element's 64 nodes (its vertices
//    and its interior nodes) by a lane pair;
//  * the MLP 12 -> 64 -> 64 -> 1 with SiLU activations: BF16 weights and
//    activations on the matrix cores (v_mfma_f32_32x32x16_bf16), FP32
//    accumulation, biases and master weights in FP32;
//  * its output y_e stands for the element mean of mu_e / (cglob beta_e) on
//    the log scale 1 + log10(ratio + 1e-10)/10, so
//    mu_e(e, :) = cglob beta_e (10^(10 (y_e - 1)) - 1e-10);
//  * training: forward with the pre-activations kept, backward with the
//    weight gradients on the matrix cores (each wave accumulates its tiles,
//    the partials are summed in a fixed order), Adam on the FP32 weights;
//  * deployment: one kernel computes features, MLP and mu_e for tiles of 32
//    elements per wave; an element whose features leave the training range,
//    or whose output leaves [0, 1], is listed and takes the plain sensor's
//    values instead (one wave per listed element).
#include <cfloat>
#include <climits>
#include <cmath>

#include "common.hpp"
#include "device.hpp"
#include "sensor.hpp"

namespace sodhip {

constexpr int kNF = 12, kNFP = 16, kNH = 64;
// FP32 parameter layout: W1 (12 x 64, row = feature), b1, W2 (64 x 64, row =
// input), b2, w3 (64), b3
constexpr int kW1 = 0, kB1 = kW1 + kNF * kNH, kW2 = kB1 + kNH, kB2 = kW2 + kNH * kNH,
              kW3 = kB2 + kNH, kB3 = kW3 + kNH, kNP = kB3 + 1;
// BF16 operand copy: W1 padded to 16 rows (rows 12..15 zero), W2, w3
constexpr int kQW1 = 0, kQW2 = kNFP * kNH, kQW3 = kQW2 + kNH * kNH, kNQ = kQW3 + kNH;
static_assert(kNP == 5057 && kNQ == 5184, "parameter layout");

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// C[32x32] += A[32x16] B[16x32] on the matrix cores: lane l supplies row l%32
// of A and column l%32 of B at k = 8(l/32) .. 8(l/32)+7; element r of the
// result in lane l is C[crow(r, l)][l%32].
__device__ __forceinline__ f32x16 mma(const bf16x8 &a, const bf16x8 &b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ int crow(int r, int l) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }

// operand fragments from row-major LDS tiles (element types __bf16 or float)
template <typename S> // A[i][k] = M[i][k0 + k]
__device__ __forceinline__ bf16x8 frag_a(const S *M, int ld, int k0, int l)
{
    bf16x8 a;
    const S *p = M + (l & 31) * ld + k0 + 8 * (l >> 5);
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = (__bf16)(float)p[q];
    return a;
}
template <typename S> // A[i][k] = M[k0 + k][i0 + i] (the transpose of M)
__device__ __forceinline__ bf16x8 frag_at(const S *M, int ld, int i0, int k0, int l)
{
    bf16x8 a;
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = (__bf16)(float)M[(k0 + 8 * (l >> 5) + q) * ld + i0 + (l & 31)];
    return a;
}
template <typename S> // B[k][j] = M[k0 + k][j0 + j]
__device__ __forceinline__ bf16x8 frag_b(const S *M, int ld, int k0, int j0, int l)
{
    bf16x8 b;
#pragma unroll
    for (int q = 0; q < 8; ++q) b[q] = (__bf16)(float)M[(k0 + 8 * (l >> 5) + q) * ld + j0 + (l & 31)];
    return b;
}
template <typename S> // B[k][j] = M[j0 + j][k0 + k] (the transpose of M)
__device__ __forceinline__ bf16x8 frag_bt(const S *M, int ld, int k0, int j0, int l)
{
    bf16x8 b;
    const S *p = M + (j0 + (l & 31)) * ld + k0 + 8 * (l >> 5);
#pragma unroll
    for (int q = 0; q < 8; ++q) b[q] = (__bf16)(float)p[q];
    return b;
}

// ---------------------------------------------------------------- features
struct FeatArgs {
    int nelem, npoin;
    const int *connecT;
    const double *Reta, *rho, *u, *Tem, *helem, *helem_k;
    const int *inv;
    double gamma, Rgas, ce, cmax, fact_low_mach;
    const double *pmax;
    int nparts;
    // joined instances: the normalisation of element e is norm_i[elem_inst[e]]
    const double *norm_i;
    const int *elem_inst;
};

// the scale of the viscosity ratios and of the MLP's output: a ratio x in
// [0, 1] maps to 1 + log10(x + 1e-10)/10 in [0, 1], so the BF16 operands carry
// ratios over ten decades with the same relative resolution
__device__ __forceinline__ double logscale(double x) { return 1.0 + log10(x + 1e-10) / 10.0; }
__device__ __forceinline__ double unlogscale(double y) { return pow(10.0, 10.0 * (y - 1.0)) - 1e-10; }

__device__ __forceinline__ double wmax(double v)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v = fmax(v, __shfl_xor(v, s, 64));
    return v;
}

// the entropy normalisation folded from its partial maxima by every wave
__device__ __forceinline__ double fold_norm(const FeatArgs &A, int l)
{
    if (!A.pmax) return 1.0;
    double norm = 0.0;
    for (int i = l; i < A.nparts; i += 64) norm = fmax(norm, A.pmax[i]);
    return wmax(norm);
}

// The features read 16 of the element's 64 nodes: its 8 vertices (sod2d
// indices 1..2 in each direction) for h = 0 and its 8 interior nodes (3..4)
// for h = 1.  Returns the element-local index of the q-th of them.
__device__ __forceinline__ int subset_node(const int *inv, int h, int q)
{
    const int i = 2 * h + (q & 1), j = 2 * h + ((q >> 1) & 1), k = 2 * h + (q >> 2);
    return inv[i + 4 * (j + 4 * k)] - 1;
}

// Features of element e, computed by a lane pair: lane h of the pair reads
// the element's vertices (h = 0) or interior nodes (h = 1), and the pair
// combines its partial statistics, so that both lanes return the same values.
// With b_n = rho hk cmax/p (|u| + c), the cap estimate beta = max b_n over the
// 16 nodes, r_n = ce rho |Reta|/norm h_n^2 / beta and m_n = min(r_n, 1):
// max, mean, min and standard deviation of m and its mean over the vertices
// and over the interior nodes (these six on the log scale); the share of
// nodes with r >= 1; log10(1 + max r)/4; the largest Mach number; the mean of
// |u|/(|u| + c); ten times the relative spread (max - min)/mean of density
// and of temperature.
// a double at a 32-bit byte offset from a wave-uniform base (one offset
// register serves every field of a node)
__device__ __forceinline__ double ld_at(const double *base, unsigned off)
{
    return *(const double *)((const char *)base + off);
}

__device__ __forceinline__ void pair_features(const FeatArgs &A, int e, int h, const int sn[8],
                                              double norm, float f[kNF], double &beta)
{
    const double hk = A.helem_k[e];
    const double ccoef = A.cmax / 3.0;
    // byte offsets below 2^32: the host checks nelem * 64 * 8 and npoin * 8
    const double *u1p = A.u + A.npoin, *u2p = A.u + 2 * (size_t)A.npoin;
    int nd[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
        nd[q] = *(const int *)((const char *)A.connecT + ((unsigned)e * 256u + (unsigned)sn[q] * 4u));
    double Ve[8];
    double bmax = 0.0, mach = 0.0, uc = 0.0;
    double rx = -DBL_MAX, rn = DBL_MAX, rs = 0.0, tx = -DBL_MAX, tn = DBL_MAX, ts = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const unsigned o = (unsigned)nd[q] * 8u;
        const double u0 = ld_at(A.u, o), u1 = ld_at(u1p, o), u2 = ld_at(u2p, o);
        const double T = ld_at(A.Tem, o), rho = ld_at(A.rho, o), Reta = ld_at(A.Reta, o);
        const double hn = ld_at(A.helem, ((unsigned)e + (unsigned)A.nelem * (unsigned)sn[q]) * 8u);
        const double um = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
        const double c = sqrt(A.Rgas * A.gamma * T);
        bmax = fmax(bmax, (rho * hk) * ccoef * (um + A.fact_low_mach * c));
        Ve[q] = A.ce * (rho * fabs(Reta) / norm) * (hn * hn);
        mach = fmax(mach, c > 0.0 ? um / c : 0.0);
        uc += um / (um + c);
        rx = fmax(rx, rho);
        rn = fmin(rn, rho);
        rs += rho;
        tx = fmax(tx, T);
        tn = fmin(tn, T);
        ts += T;
    }
    auto pmax2 = [](double v) { return fmax(v, __shfl_xor(v, 1, 64)); };
    auto pmin2 = [](double v) { return fmin(v, __shfl_xor(v, 1, 64)); };
    auto psum2 = [](double v) { return v + __shfl_xor(v, 1, 64); }; // a + b == b + a
    const double be = pmax2(bmax);
    double ms = 0.0, m2 = 0.0, mx = 0.0, mn = DBL_MAX, cap = 0.0, rmax = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const double r = be > 0.0 ? Ve[q] / be : 0.0;
        const double m = fmin(r, 1.0);
        ms += m;
        m2 += m * m;
        mx = fmax(mx, m);
        mn = fmin(mn, m);
        cap += r >= 1.0 ? 1.0 : 0.0;
        rmax = fmax(rmax, r);
    }
    const double other = __shfl_xor(ms, 1, 64);
    const double mvert = h == 0 ? ms : other, minner = h == 0 ? other : ms;
    const double msum = ms + other, mean = msum / 16.0;
    m2 = psum2(m2);
    f[0] = (float)logscale(pmax2(mx));
    f[1] = (float)logscale(mean);
    f[2] = (float)logscale(pmin2(mn));
    f[3] = (float)logscale(sqrt(fmax(m2 / 16.0 - mean * mean, 0.0)));
    f[4] = (float)logscale(mvert / 8.0);
    f[5] = (float)logscale(minner / 8.0);
    f[6] = (float)(psum2(cap) / 16.0);
    f[7] = (float)(log10(1.0 + pmax2(rmax)) / 4.0);
    f[8] = (float)pmax2(mach);
    f[9] = (float)(psum2(uc) / 16.0);
    f[10] = (float)(10.0 * (pmax2(rx) - pmin2(rn)) / (psum2(rs) / 16.0));
    f[11] = (float)(10.0 * (pmax2(tx) - pmin2(tn)) / (psum2(ts) / 16.0));
    beta = be;
}

// One wave per tile of 32 consecutive elements (lane pair 2c, 2c+1 takes
// element c); writes the features (nelem x 16, zero padded) and beta.
__global__ __launch_bounds__(256) void k_sensor_features(FeatArgs A, float *__restrict__ X,
                                                         double *__restrict__ beta)
{
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, h = l & 1;
    const double norm = fold_norm(A, l);
    int sn[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) sn[q] = subset_node(A.inv, h, q);
    const int ntile = (A.nelem + 31) / 32;
    for (int tile = blockIdx.x * 4 + wv; tile < ntile; tile += gridDim.x * 4) {
        const int e = tile * 32 + (l >> 1), ec = min(e, A.nelem - 1);
        float f[kNF];
        double be;
        pair_features(A, ec, h, sn, A.norm_i ? A.norm_i[A.elem_inst[ec]] : norm, f, be);
        if (e < A.nelem) { // lane h writes features 8h .. 8h+7
            float4 *o = (float4 *)(X + (size_t)e * kNFP + 8 * h);
            const float g[8] = {h ? f[8] : f[0], h ? f[9] : f[1], h ? f[10] : f[2],
                                h ? f[11] : f[3], h ? 0.0f : f[4], h ? 0.0f : f[5],
                                h ? 0.0f : f[6], h ? 0.0f : f[7]};
            o[0] = make_float4(g[0], g[1], g[2], g[3]);
            o[1] = make_float4(g[4], g[5], g[6], g[7]);
            if (h == 0) beta[e] = be;
        }
    }
}

// ---------------------------------------------------------------- forward
// The MLP on one tile of 32 rows held by a wave.  xa is the input operand
// (row l%32, features 8(l/32) .. 8(l/32)+7, BF16); the weights are in LDS (sW
// the BF16 copy, sb the FP32 biases b1, b2, b3) and hs is the wave's 32 x 64
// activation tile.  TRAIN keeps the pre-activations of rows r0 + i < n in Z1,
// Z2 (n x 64).  Returns the output layer's accumulator without b3: element r
// of lanes 0 and 32 is row crow(r, l).
template <bool TRAIN>
__device__ __forceinline__ f32x16 mlp_tile(const bf16x8 &xa, const __bf16 *sW, const float *sb,
                                           __bf16 *hs, int l, int r0, int n, float *Z1, float *Z2)
{
    const f32x16 zero = {};
    f32x16 c1[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) c1[nt] = mma(xa, frag_b(sW + kQW1, kNH, 0, 32 * nt, l), zero);
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = crow(r, l), j = 32 * nt + (l & 31);
            const float z = c1[nt][r] + sb[j];
            if (TRAIN && r0 + i < n) Z1[(size_t)(r0 + i) * kNH + j] = z;
            hs[i * kNH + j] = (__bf16)silu(z);
        }
    wave_sync();
    f32x16 c2[2] = {zero, zero};
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int kc = 0; kc < 4; ++kc)
            c2[nt] = mma(frag_a(hs, kNH, 16 * kc, l), frag_b(sW + kQW2, kNH, 16 * kc, 32 * nt, l), c2[nt]);
    wave_sync(); // every lane has read H1 before H2 replaces it
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = crow(r, l), j = 32 * nt + (l & 31);
            const float z = c2[nt][r] + sb[kNH + j];
            if (TRAIN && r0 + i < n) Z2[(size_t)(r0 + i) * kNH + j] = z;
            hs[i * kNH + j] = (__bf16)silu(z);
        }
    wave_sync();
    // output layer: B has w3 in column 0 and zeros elsewhere
    f32x16 c3 = zero;
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
        bf16x8 b;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            b[q] = (l & 31) == 0 ? sW[kQW3 + 16 * kc + 8 * (l >> 5) + q] : (__bf16)0.0f;
        c3 = mma(frag_a(hs, kNH, 16 * kc, l), b, c3);
    }
    return c3;
}

// the BF16 weights and FP32 biases into LDS (block-wide)
__device__ __forceinline__ void load_weights(const __bf16 *wq, const float *w, __bf16 *sW, float *sb)
{
    // 16-byte pieces (kNQ = 5184 BF16 = 648 pieces; wq and sW 16-byte aligned)
    static_assert(kNQ % 8 == 0, "BF16 copy in 16-byte pieces");
    for (int i = threadIdx.x; i < kNQ / 8; i += blockDim.x)
        ((uint4 *)sW)[i] = ((const uint4 *)wq)[i];
    for (int i = threadIdx.x; i < 2 * kNH + 1; i += blockDim.x)
        sb[i] = i < kNH ? w[kB1 + i] : (i < 2 * kNH ? w[kB2 + i - kNH] : w[kB3]);
    __syncthreads();
}

// One wave per tile of 32 rows (elements); the block's four waves share the
// BF16 weights in LDS.  TRAIN keeps the pre-activations Z1, Z2 (n x 64).
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp_forward(int n, const float *__restrict__ X,
                                                     const __bf16 *__restrict__ wq,
                                                     const float *__restrict__ w,
                                                     float *__restrict__ y, float *__restrict__ Z1,
                                                     float *__restrict__ Z2)
{
    __shared__ __align__(16) __bf16 sW[kNQ];
    __shared__ float sb[2 * kNH + 1];
    __shared__ __bf16 sX[4][32 * kNFP];
    __shared__ __bf16 sH[4][32 * kNH];
    load_weights(wq, w, sW, sb);
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    __bf16 *xs = sX[wv], *hs = sH[wv];
    const int ntile = (n + 31) / 32;
    for (int tile = blockIdx.x * 4 + wv; tile < ntile; tile += gridDim.x * 4) {
        const int r0 = tile * 32;
        for (int i = l; i < 32 * kNFP; i += 64) {
            const int rr = r0 + i / kNFP;
            xs[i] = (__bf16)(rr < n ? X[(size_t)rr * kNFP + i % kNFP] : 0.0f);
        }
        wave_sync();
        const f32x16 c3 = mlp_tile<TRAIN>(frag_a(xs, kNFP, 0, l), sW, sb, hs, l, r0, n, Z1, Z2);
        if ((l & 31) == 0)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = crow(r, l);
                if (r0 + i < n) y[r0 + i] = c3[r] + sb[2 * kNH];
            }
        wave_sync(); // the next tile rewrites sX and sH
    }
}

// ---------------------------------------------------------------- backward
// Gradients of L = inv_batch * sum over rows (y - t)^2.  Every wave takes the
// tiles gw, gw + nw, ... (gw its global index, nw the number of waves),
// accumulates their gradients in registers and writes its partial (kNP
// values) to gpart + gw * kNP; the partials are summed in wave order.
__global__ __launch_bounds__(256) void k_mlp_backward(int n, float inv_batch,
                                                      const float *__restrict__ X,
                                                      const float *__restrict__ t,
                                                      const __bf16 *__restrict__ wq,
                                                      const float *__restrict__ y,
                                                      const float *__restrict__ Z1,
                                                      const float *__restrict__ Z2,
                                                      float *__restrict__ gpart)
{
    __shared__ __bf16 sW[kNQ];
    __shared__ float sX[4][32 * kNFP];
    __shared__ float sA[4][32 * kNH]; // H1 of the tile
    __shared__ float sD[4][32 * kNH]; // dZ2, then dZ1
    __shared__ float sdy[4][32];
    for (int i = threadIdx.x; i < kNQ; i += 256) sW[i] = wq[i];
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    float *xs = sX[wv], *as = sA[wv], *ds = sD[wv], *dys = sdy[wv];
    const int gw = blockIdx.x * 4 + wv, nw = gridDim.x * 4;
    const f32x16 zero = {};
    f32x16 gW1[2] = {zero, zero};                   // rows: features (0..15), columns j
    f32x16 gW2[2][2] = {{zero, zero}, {zero, zero}}; // [input tile][output tile]
    float gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f; // lane l: column l
    const float w3l = (float)sW[kQW3 + l];
    const int ntile = (n + 31) / 32;
    for (int tile = gw; tile < ntile; tile += nw) {
        const int r0 = tile * 32;
        for (int i = l; i < 32 * kNFP; i += 64) {
            const int rr = r0 + i / kNFP;
            xs[i] = rr < n ? (float)(__bf16)X[(size_t)rr * kNFP + i % kNFP] : 0.0f;
        }
        if (l < 32) dys[l] = r0 + l < n ? 2.0f * (y[r0 + l] - t[r0 + l]) * inv_batch : 0.0f;
        wave_sync();
        // lane l = column j; rows in order
        for (int i = 0; i < 32; ++i) {
            const bool ok = r0 + i < n;
            const float z1 = ok ? Z1[(size_t)(r0 + i) * kNH + l] : 0.0f;
            const float z2 = ok ? Z2[(size_t)(r0 + i) * kNH + l] : 0.0f;
            const float dyi = dys[i];
            as[i * kNH + l] = ok ? silu(z1) : 0.0f;
            const float dz2 = ok ? dyi * w3l * dsilu(z2) : 0.0f;
            ds[i * kNH + l] = dz2;
            gw3 += (ok ? silu(z2) : 0.0f) * dyi;
            gb2 += dz2;
            if (l == 0) gb3 += dyi;
        }
        wave_sync();
        // dH1 = dZ2 W2^T
        f32x16 dh[2] = {zero, zero};
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int kc = 0; kc < 4; ++kc)
                dh[nt] = mma(frag_a(ds, kNH, 16 * kc, l), frag_bt(sW + kQW2, kNH, 16 * kc, 32 * nt, l), dh[nt]);
        // dW2 += H1^T dZ2
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int kc = 0; kc < 2; ++kc)
                    gW2[mt][nt] = mma(frag_at(as, kNH, 32 * mt, 16 * kc, l), frag_b(ds, kNH, 16 * kc, 32 * nt, l), gW2[mt][nt]);
        wave_sync(); // dZ2 read by every lane before dZ1 replaces it
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = crow(r, l), j = 32 * nt + (l & 31);
                const bool ok = r0 + i < n;
                const float z1 = ok ? Z1[(size_t)(r0 + i) * kNH + j] : 0.0f;
                ds[i * kNH + j] = ok ? dh[nt][r] * dsilu(z1) : 0.0f;
            }
        wave_sync();
        for (int i = 0; i < 32; ++i) gb1 += ds[i * kNH + l];
        // dW1 += X^T dZ1 (rows 16..31 of the A tile are zero)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) {
                bf16x8 a;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int k = 16 * kc + 8 * (l >> 5) + q, m = l & 31;
                    a[q] = (__bf16)(m < kNFP ? xs[k * kNFP + m] : 0.0f);
                }
                gW1[nt] = mma(a, frag_b(ds, kNH, 16 * kc, 32 * nt, l), gW1[nt]);
            }
        wave_sync(); // the next tile rewrites this wave's LDS
    }
    float *g = gpart + (size_t)gw * kNP;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int f = crow(r, l), j = 32 * nt + (l & 31);
            if (f < kNF) g[kW1 + f * kNH + j] = gW1[nt][r];
        }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int k = 32 * mt + crow(r, l), j = 32 * nt + (l & 31);
                g[kW2 + k * kNH + j] = gW2[mt][nt][r];
            }
    g[kB1 + l] = gb1;
    g[kB2 + l] = gb2;
    g[kW3 + l] = gw3;
    if (l == 0) g[kB3] = gb3;
}

// the BF16 operand copy of the FP32 parameters (W1 padded with zero rows)
__device__ __forceinline__ void quantize_one(int p, const float *w, __bf16 *wq)
{
    if (p >= kNQ) return;
    if (p < kQW2) {
        const int f = p / kNH, j = p % kNH;
        wq[p] = (__bf16)(f < kNF ? w[kW1 + f * kNH + j] : 0.0f);
    } else if (p < kQW3) {
        wq[p] = (__bf16)w[kW2 + (p - kQW2)];
    } else {
        wq[p] = (__bf16)w[kW3 + (p - kQW3)];
    }
}
__global__ __launch_bounds__(256) void k_mlp_quantize(const float *w, __bf16 *wq)
{
    quantize_one(blockIdx.x * blockDim.x + threadIdx.x, w, wq);
}

// Adam (one thread per parameter); the gradient is the sum of the wave
// partials in wave order; then the BF16 copy is refreshed
__global__ __launch_bounds__(256) void k_mlp_adam(int nparts, const float *__restrict__ gpart,
                                                  float *__restrict__ w, float *__restrict__ m,
                                                  float *__restrict__ v, float lr, float b1,
                                                  float b2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kNP) {
        float g = 0.0f;
        for (int k = 0; k < nparts; ++k) g += gpart[(size_t)k * kNP + p];
        const float mm = b1 * m[p] + (1.0f - b1) * g;
        const float vv = b2 * v[p] + (1.0f - b2) * g * g;
        m[p] = mm;
        v[p] = vv;
        w[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    }
}

// ---------------------------------------------------------------- deployment
struct SurRange {
    float lo[kNF], hi[kNF];
};

struct SurArgs {
    FeatArgs F;
    const __bf16 *wq;
    const float *w;
    SurRange R;
    double cglob;
    float *X;     // optional: features (nelem x 16)
    double *beta; // optional: the cap estimate per element
    float *y;     // optional: the MLP's output per element
    int *flag;    // per element: 1 where the plain sensor supplies mu_e
    int *list;    // list[0] = number of fallback elements, list[1..] = them
    double *mu_e;
};

// The surrogate sensor, one wave per tile of 32 consecutive elements:
// features (lane pairs), the MLP (the wave's tile on the matrix cores), the
// range checks and mu_e(e, :) = cglob beta_e (10^(10 (y_e - 1)) - 1e-10) for
// the elements that pass them, written in whole lines of the element-fastest
// mu_e.  An element whose features leave [lo, hi] or whose output leaves
// [0, 1] is flagged and appended to the fallback list (one atomic per wave;
// the list's order varies between runs, its set does not).
__global__ __launch_bounds__(256, 2) void k_surrogate(SurArgs A)
{
    __shared__ __align__(16) __bf16 sW[kNQ];
    __shared__ float sb[2 * kNH + 1];
    __shared__ __bf16 sH[4][32 * kNH];
    __shared__ float sy[4][32];
    load_weights(A.wq, A.w, sW, sb);
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, h = l & 1, c = l >> 1;
    const double norm = fold_norm(A.F, l);
    int sn[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) sn[q] = subset_node(A.F.inv, h, q);
    const int nelem = A.F.nelem, ntile = (nelem + 31) / 32;
    for (int tile = blockIdx.x * 4 + wv; tile < ntile; tile += gridDim.x * 4) {
        const int e0 = tile * 32, e = e0 + c;
        const bool valid = e < nelem;
        float f[kNF];
        double be;
        pair_features(A.F, min(e, nelem - 1), h, sn, norm, f, be);
        // input operand: row l%32 is element l%32, held by lane 2 (l%32)
        const int src = 2 * (l & 31);
        bf16x8 xa;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float lo = __shfl(f[q], src, 64);
            const float hi = q < 4 ? __shfl(f[8 + q], src, 64) : 0.0f;
            xa[q] = (__bf16)(l < 32 ? lo : hi);
        }
        const f32x16 c3 = mlp_tile<false>(xa, sW, sb, sH[wv], l, 0, 0, nullptr, nullptr);
        if ((l & 31) == 0)
#pragma unroll
            for (int r = 0; r < 16; ++r) sy[wv][crow(r, l)] = c3[r] + sb[2 * kNH];
        wave_sync();
        const float ye = sy[wv][c];
        bool ok = ye >= 0.0f && ye <= 1.0f; // false for NaN
#pragma unroll
        for (int k = 0; k < kNF; ++k) ok = ok && f[k] >= A.R.lo[k] && f[k] <= A.R.hi[k];
        const bool bad = valid && !ok;
        const double v = A.cglob * be * fmax(unlogscale((double)ye), 0.0);
        if (valid && h == 0) {
            A.flag[e] = bad ? 1 : 0;
            if (A.beta) A.beta[e] = be;
            if (A.y) A.y[e] = ye;
        }
        if (A.X && valid) {
            float4 *o = (float4 *)(A.X + (size_t)e * kNFP + 8 * h);
            const float g[8] = {h ? f[8] : f[0], h ? f[9] : f[1], h ? f[10] : f[2],
                                h ? f[11] : f[3], h ? 0.0f : f[4], h ? 0.0f : f[5],
                                h ? 0.0f : f[6], h ? 0.0f : f[7]};
            o[0] = make_float4(g[0], g[1], g[2], g[3]);
            o[1] = make_float4(g[4], g[5], g[6], g[7]);
        }
        const unsigned long long bm = __ballot(bad && h == 0);
        if (bm) {
            int base = 0;
            if (l == 0) base = atomicAdd(A.list, __popcll(bm));
            base = __shfl(base, 0, 64);
            if (bad && h == 0) A.list[1 + base + __popcll(bm & ((1ull << l) - 1))] = e;
        }
        // mu_e: lane l writes element l%32 at nodes 32(l/32) .. 32(l/32)+31
        const double vc = __shfl(v, src, 64);
        const int keep = __shfl((valid && !bad) ? 1 : 0, src, 64);
        if (keep) {
            double *o = A.mu_e + (e0 + (l & 31)) + (size_t)nelem * 32 * (l >> 5);
#pragma unroll 8
            for (int a = 0; a < 32; ++a) o[(size_t)nelem * a] = vc;
        }
        wave_sync(); // the next tile rewrites sy and sH
    }
}

// smart_visc_spectral (the plain sensor, p = 3) on the listed elements, one
// wave per element; the same arithmetic as k_smart_visc_ch.  With nfall set,
// block 0 also copies the count there.
__global__ __launch_bounds__(64) void k_smart_visc_list(SensorArgs A, const int *__restrict__ list,
                                                        int *__restrict__ nfall)
{
    constexpr int P = 3, NN = 64;
    __shared__ double ml[NN];
    const int a = threadIdx.x;
    const int cnt = list[0];
    if (nfall && blockIdx.x == 0 && a == 0) *nfall = cnt;
    if ((int)blockIdx.x >= cnt) return;
    int src[27];
    filter_sources<P>(a, A.inv, src);
    double norm = 1.0;
    if (A.pmax) {
        norm = 0.0;
        for (int i = a; i < A.nparts; i += 64) norm = fmax(norm, A.pmax[i]);
        norm = wmax(norm);
    }
    const double w[3] = {0.25, 0.5, 0.25};
    for (int i = blockIdx.x; i < cnt; i += gridDim.x) {
        const int e = list[1 + i];
        const int n = A.connecT[(size_t)e * NN + a];
        const double u0 = A.u[n], u1 = A.u[n + A.npoin], u2 = A.u[n + 2 * (size_t)A.npoin];
        const double aux2 = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
        const double aux3 = A.fact_low_mach * sqrt(A.Rgas * A.gamma * A.Tem[n]);
        const double rho = A.rho[n];
        const double be = wmax((rho * A.helem_k[e]) * (A.cmax / (double)P) * (aux2 + aux3));
        const double R1 = rho * fabs(A.Reta[n]) / norm;
        const double h = A.helem[e + (size_t)A.nelem * a];
        const double v = A.cglob * fmin(A.ce * R1 * (h * h), be);
        ml[a] = v;
        __syncthreads();
        double acc = 0.0;
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int nn = 0; nn < 3; ++nn) acc += w[l] * w[m] * w[nn] * ml[src[9 * l + 3 * m + nn]];
        const size_t g = e + (size_t)A.nelem * a;
        A.mue_l[g] = v;
        A.mu_e[g] = acc;
        __syncthreads();
    }
}

__global__ void k_zero_int(int *p) { *p = 0; }

static int mlp_check(const sodhip_mlp *M)
{
    SODHIP_REQUIRE(M && M->w && M->wq, "null model");
    SODHIP_REQUIRE(((uintptr_t)M->wq & 15) == 0, "the BF16 copy wq must be 16-byte aligned");
    return SODHIP_OK;
}

} // namespace sodhip

using namespace sodhip;

extern "C" int sodhip_mlp_sizes(int *nparams, int *nq, int *nfeatures)
{
    if (nparams) *nparams = kNP;
    if (nq) *nq = kNQ;
    if (nfeatures) *nfeatures = kNF;
    return SODHIP_OK;
}

extern "C" int sodhip_mlp_quantize(sodhip_ctx ctx, const sodhip_mlp *M)
{
    SODHIP_REQUIRE(ctx, "null context");
    int rc = mlp_check(M);
    if (rc) return rc;
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    k_mlp_quantize<<<(kNQ + 255) / 256, 256, 0, ctx->stream>>>(M->w, (__bf16 *)M->wq);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

// the feature arguments after the entropy normalisation's partials (pmax)
static FeatArgs feat_args(const MeshCache *m, int nelem, int npoin, const double *Reta,
                          double gamma_gas, const double *rho, const double *u, const double *Tem,
                          const double *helem, const double *helem_k, const int *invAtoIJK,
                          const sodhip_visc_params *prm, const double *pmax, int nparts)
{
    FeatArgs A{};
    A.pmax = pmax;
    A.nparts = nparts;
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
    A.fact_low_mach = prm->flag_drop_c_in_envit ? 0.0 : 1.0;
    return A;
}

// blocks of four waves, one tile of 32 elements per wave and pass
static int tile_grid(sodhip_ctx ctx, int nelem)
{
    const int ntile = (nelem + 31) / 32;
    return std::max(1, std::min((ntile + 3) / 4, ctx->num_cu * 8));
}

static int check_sensor_args(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                             const int *connec, const int *lpoin_w, const double *Reta,
                             const double *rho, const double *u, const double *Tem,
                             const double *eta, const double *helem, const double *helem_k,
                             const int *invAtoIJK, const sodhip_visc_params *prm)
{
    SODHIP_REQUIRE(ctx && prm, "null argument");
    SODHIP_REQUIRE(porder == 3, "the sensor features and the surrogate are defined for porder 3");
    SODHIP_REQUIRE(nelem > 0 && npoin > 0 && npoin_w > 0 && npoin_w <= npoin,
                   "invalid mesh sizes");
    SODHIP_REQUIRE(connec && lpoin_w && Reta && rho && u && Tem && eta && helem && helem_k &&
                       invAtoIJK,
                   "null array argument");
    // the feature kernels address node and element-node arrays with 32-bit byte offsets
    SODHIP_REQUIRE((long long)nelem * 64 * 8 < (1LL << 32) && (long long)npoin * 8 < (1LL << 32),
                   "mesh too large for the feature kernels (nelem < 2^23, npoin < 2^29)");
    return SODHIP_OK;
}

static int sensor_features(sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w,
                           const int *connec, const int *lpoin_w, const double *Reta,
                           double gamma_gas, const double *rho, const double *u, const double *Tem,
                           const double *eta, const double *helem, const double *helem_k,
                           const int *invAtoIJK, const sodhip_visc_params *prm, float *X,
                           double *beta, const sodhip_instances *ens)
{
    int rc = check_sensor_args(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w, Reta, rho, u,
                               Tem, eta, helem, helem_k, invAtoIJK, prm);
    if (rc) return rc;
    SODHIP_REQUIRE(X && beta, "null output array");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    if ((rc = get_mesh(ctx, porder, nelem, npoin, connec, &m))) return rc;
    const double *pmax = nullptr, *norm_i = nullptr;
    int nparts = 0;
    const bool normalise = prm->flag_normalise_entropy == 1;
    int tok = prof_begin(ctx, 2);
    if (ens && normalise) {
        if ((rc = entropy_norm_instances(ctx, ens, lpoin_w, eta, &norm_i))) return rc;
    } else if ((rc = entropy_norm_partials(ctx, npoin_w, lpoin_w, eta, normalise, &pmax,
                                           &nparts))) {
        return rc;
    }
    FeatArgs A = feat_args(m, nelem, npoin, Reta, gamma_gas, rho, u, Tem, helem, helem_k,
                           invAtoIJK, prm, pmax, nparts);
    A.norm_i = norm_i;
    A.elem_inst = norm_i ? ens->elem_inst : nullptr;
    k_sensor_features<<<tile_grid(ctx, nelem), 256, 0, ctx->stream>>>(A, X, beta);
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

extern "C" int sodhip_sensor_features_f64(sodhip_ctx ctx, int porder, int nelem, int npoin,
                                          int npoin_w, const int *connec, const int *lpoin_w,
                                          const double *Reta, double gamma_gas, const double *rho,
                                          const double *u, const double *Tem, const double *eta,
                                          const double *helem, const double *helem_k,
                                          const int *invAtoIJK, const sodhip_visc_params *prm,
                                          float *X, double *beta)
{
    return sensor_features(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w, Reta, gamma_gas,
                           rho, u, Tem, eta, helem, helem_k, invAtoIJK, prm, X, beta, nullptr);
}

extern "C" int sodhip_sensor_features_ens_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w, const int *connec,
    const int *lpoin_w, const sodhip_instances *ens, const double *Reta, double gamma_gas,
    const double *rho, const double *u, const double *Tem, const double *eta,
    const double *helem, const double *helem_k, const int *invAtoIJK,
    const sodhip_visc_params *prm, float *X, double *beta)
{
    SODHIP_REQUIRE(ens && ens->ninst > 0 && ens->wseg && ens->elem_inst, "invalid instances");
    return sensor_features(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w, Reta, gamma_gas,
                           rho, u, Tem, eta, helem, helem_k, invAtoIJK, prm, X, beta, ens);
}

extern "C" int sodhip_mlp_forward(sodhip_ctx ctx, const sodhip_mlp *M, int n, const float *X,
                                  float *y)
{
    SODHIP_REQUIRE(ctx, "null context");
    int rc = mlp_check(M);
    if (rc) return rc;
    SODHIP_REQUIRE(n > 0 && X && y, "invalid arguments");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    const int ntile = (n + 31) / 32;
    const int grid = std::max(1, std::min((ntile + 3) / 4, ctx->num_cu * 4));
    k_mlp_forward<false><<<grid, 256, 0, ctx->stream>>>(n, X, (const __bf16 *)M->wq, M->w, y,
                                                        nullptr, nullptr);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

extern "C" int sodhip_mlp_train_step(sodhip_ctx ctx, sodhip_mlp *M, int n, const float *X,
                                     const float *t, int step, float lr, float *y, float *Z,
                                     float *gpart, int nwaves)
{
    SODHIP_REQUIRE(ctx, "null context");
    int rc = mlp_check(M);
    if (rc) return rc;
    SODHIP_REQUIRE(M->m && M->v, "null Adam moments");
    SODHIP_REQUIRE(n > 0 && X && t && y && Z && gpart && step >= 1 && nwaves >= 4 &&
                       nwaves % 4 == 0,
                   "invalid arguments");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    const int ntile = (n + 31) / 32;
    const int fgrid = std::max(1, std::min((ntile + 3) / 4, ctx->num_cu * 4));
    float *Z1 = Z, *Z2 = Z + (size_t)n * kNH;
    k_mlp_forward<true><<<fgrid, 256, 0, ctx->stream>>>(n, X, (const __bf16 *)M->wq, M->w, y, Z1, Z2);
    k_mlp_backward<<<nwaves / 4, 256, 0, ctx->stream>>>(n, 1.0f / (float)n, X, t,
                                                        (const __bf16 *)M->wq, y, Z1, Z2, gpart);
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f / (1.0f - powf(b1, (float)step)), c2 = 1.0f / (1.0f - powf(b2, (float)step));
    k_mlp_adam<<<(kNP + 255) / 256, 256, 0, ctx->stream>>>(nwaves, gpart, M->w, M->m, M->v, lr, b1,
                                                           b2, 1e-8f, c1, c2);
    k_mlp_quantize<<<(kNQ + 255) / 256, 256, 0, ctx->stream>>>(M->w, (__bf16 *)M->wq);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

extern "C" int sodhip_mlp_gradient(sodhip_ctx ctx, const sodhip_mlp *M, int n, const float *X,
                                   const float *t, float *y, float *Z, float *gpart, int nwaves)
{
    SODHIP_REQUIRE(ctx, "null context");
    int rc = mlp_check(M);
    if (rc) return rc;
    SODHIP_REQUIRE(n > 0 && X && t && y && Z && gpart && nwaves >= 4 && nwaves % 4 == 0,
                   "invalid arguments");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    const int ntile = (n + 31) / 32;
    const int fgrid = std::max(1, std::min((ntile + 3) / 4, ctx->num_cu * 4));
    float *Z1 = Z, *Z2 = Z + (size_t)n * kNH;
    k_mlp_forward<true><<<fgrid, 256, 0, ctx->stream>>>(n, X, (const __bf16 *)M->wq, M->w, y, Z1, Z2);
    k_mlp_backward<<<nwaves / 4, 256, 0, ctx->stream>>>(n, 1.0f / (float)n, X, t,
                                                        (const __bf16 *)M->wq, y, Z1, Z2, gpart);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

extern "C" int sodhip_surrogate_visc_spectral_f64(
    sodhip_ctx ctx, int porder, int nelem, int npoin, int npoin_w, const int *connec,
    const int *lpoin_w, const double *Reta, double gamma_gas, const double *rho, const double *u,
    const double *Tem, const double *eta, const double *helem, const double *helem_k,
    const int *invAtoIJK, const sodhip_visc_params *prm, const sodhip_mlp *M, double *mu_e,
    double *mue_l, float *X, double *beta, float *y, int *flag, int *nfallback)
{
    int rc = check_sensor_args(ctx, porder, nelem, npoin, npoin_w, connec, lpoin_w, Reta, rho, u,
                               Tem, eta, helem, helem_k, invAtoIJK, prm);
    if (rc) return rc;
    if ((rc = mlp_check(M))) return rc;
    SODHIP_REQUIRE(mu_e && mue_l && flag, "null array argument");
    SODHIP_REQUIRE(nelem <= INT_MAX - 32, "nelem too large for 32-element tiles");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    MeshCache *m = nullptr;
    if ((rc = get_mesh(ctx, porder, nelem, npoin, connec, &m))) return rc;
    // scratch2: the normalisation's partials (2 kRedBlocks doubles), then the
    // fallback list (count, elements)
    const size_t list_off = 2 * kRedBlocks * sizeof(double);
    if ((rc = ensure_buf(ctx->scratch2, list_off + sizeof(int) * ((size_t)nelem + 1)))) return rc;
    int *list = (int *)((char *)ctx->scratch2.ptr + list_off);
    const double *pmax = nullptr;
    int nparts = 0;
    int tok = prof_begin(ctx, 2);
    if ((rc = entropy_norm_partials(ctx, npoin_w, lpoin_w, eta, prm->flag_normalise_entropy == 1,
                                    &pmax, &nparts)))
        return rc;
    SurArgs A{};
    A.F = feat_args(m, nelem, npoin, Reta, gamma_gas, rho, u, Tem, helem, helem_k, invAtoIJK, prm,
                    pmax, nparts);
    A.wq = (const __bf16 *)M->wq;
    A.w = M->w;
    for (int f = 0; f < kNF; ++f) {
        A.R.lo[f] = M->lo[f];
        A.R.hi[f] = M->hi[f];
    }
    A.cglob = prm->cglob;
    A.X = X;
    A.beta = beta;
    A.y = y;
    A.flag = flag;
    A.list = list;
    A.mu_e = mu_e;
    k_zero_int<<<1, 1, 0, ctx->stream>>>(list);
    k_surrogate<<<tile_grid(ctx, nelem), 256, 0, ctx->stream>>>(A);
    // the fallback elements take the plain sensor
    SensorArgs S{};
    S.nelem = nelem;
    S.npoin = npoin;
    S.connecT = m->connecT;
    S.Reta = Reta;
    S.rho = rho;
    S.u = u;
    S.Tem = Tem;
    S.helem = helem;
    S.helem_k = helem_k;
    S.inv = invAtoIJK;
    S.gamma = gamma_gas;
    S.Rgas = prm->Rgas;
    S.ce = prm->ce;
    S.cmax = prm->cmax;
    S.cglob = prm->cglob;
    S.fact_low_mach = prm->flag_drop_c_in_envit ? 0.0 : 1.0;
    S.pmax = pmax;
    S.nparts = nparts;
    S.mu_e = mu_e;
    S.mue_l = mue_l;
    k_smart_visc_list<<<std::min(nelem, ctx->num_cu * 16), 64, 0, ctx->stream>>>(S, list, nfallback);
    prof_end(ctx, tok);
    SODHIP_TRY_HIP(hipGetLastError());
    return SODHIP_OK;
}

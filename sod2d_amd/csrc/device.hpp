// device.hpp -- device-side helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace sodhip {

// Workgroup barrier for data exchanged through LDS by ordinary LDS stores:
// the wave's own LDS operations complete (lgkmcnt), then all waves meet.  The
// asm is a compiler barrier for memory accesses.  Neither __syncthreads nor a
// workgroup fence on LDS is used: the former also waits for every
// outstanding global load, the latter for every direct-to-LDS copy in flight,
// which would drain the prefetch of the next elements.  Data written by
// direct-to-LDS copies is read only after an explicit vmcnt wait.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// s_waitcnt immediates (gfx9 encoding: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] |
// vmcnt[5:4] << 14)
constexpr int kWaitVm0 = 0x0F70;   // vmcnt(0)
constexpr int kWaitLgkm0 = 0xC07F; // lgkmcnt(0)

// Element order of the lane-per-node element loops: the blocks of one XCD
// (blockIdx % 8; the grid is a multiple of 8) walk one contiguous eighth of
// the elements, so elements that share cache lines (neighbouring nodes, the
// 16 elements of an element-fastest (nelem, nnode) line) run on one XCD and
// its L2.
struct XcdElems {
    int first, end, stride;
};
__device__ __forceinline__ XcdElems xcd_elems(int nelem)
{
    const int per = gridDim.x >> 3, xcd = blockIdx.x & 7, rk = blockIdx.x >> 3;
    const int span = (nelem + 7) >> 3;
    const int b = xcd * span;
    return {b + rk, min(nelem, b + span), per};
}

// max over the block of NW waves (one value per wave in sh[NW]); every lane
// returns the maximum
template <int NW> __device__ __forceinline__ double block_max_w(double v, double *sh)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v = fmax(v, __shfl_xor(v, s, 64));
    if constexpr (NW == 1) {
        return v;
    } else {
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
        __syncthreads();
        double m = sh[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) m = fmax(m, sh[w]);
        __syncthreads();
        return m;
    }
}

// Node visiting order of the thread-per-element kernels: alex[l] is the node
// at lexicographic position l of the element's (i, j, k) line indices (the
// inverse of the mesh cache's perm), so that a group of P+1 consecutive
// positions is one i-line: nodes that lie next to each other in memory on a
// structured numbering, read while their lines are still cached.  Built per
// block in LDS (nnode <= 125).
__device__ __forceinline__ void build_alex(const int *__restrict__ perm, int nnode, int *alex)
{
    for (int a = threadIdx.x; a < nnode; a += blockDim.x) alex[min(max(perm[a], 0), nnode - 1)] = a;
    __syncthreads();
}

// Block order of the thread-per-element kernels: logical block (blockIdx % 8)
// * (gridDim / 8) + blockIdx / 8, so that the blocks of one XCD take one
// contiguous range of elements, whose shared nodes then sit in that XCD's L2
// (the grid is a multiple of 8: tpe_grid).
__device__ __forceinline__ int xcd_block()
{
    return (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
}
inline int tpe_grid(int nelem) { return ((nelem + 255) / 256 + 7) / 8 * 8; }

} // namespace sodhip

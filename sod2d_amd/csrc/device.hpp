// device.hpp -- device-side helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace sodhip {

// Workgroup barrier that orders LDS accesses only.  __syncthreads also waits
// for every outstanding global load of the wave (its workgroup fence covers
// global memory), which would drain the prefetch of the next element; the
// kernels exchange data between lanes through LDS alone.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// s_waitcnt immediates (gfx9 encoding: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] |
// vmcnt[5:4] << 14)
constexpr int kWaitVm0 = 0x0F70;   // vmcnt(0)
constexpr int kWaitLgkm0 = 0xC07F; // lgkmcnt(0)

} // namespace sodhip

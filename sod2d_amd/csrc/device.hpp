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

} // namespace sodhip

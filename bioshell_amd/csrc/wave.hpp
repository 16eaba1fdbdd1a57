// wave.hpp -- wave-level helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace bs {

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Orders the LDS accesses of one wave before and after this point (the
// lanes of a wave then see each other's LDS writes); no work-group barrier.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

}  // namespace bs

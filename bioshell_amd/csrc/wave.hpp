// wave.hpp -- wave-level helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace bs {

// One DPP move of a double (two 32-bit halves): lanes the row or bank mask
// leaves out, and sources outside the row, read 0.
template <int CTRL, int ROW_MASK, bool BOUND>
__device__ __forceinline__ double dpp_f64(double x)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xF, BOUND);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xF, BOUND);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the 64 lanes (all active), returned in every lane.  Row shifts by
// 1, 2, 4 and 8 lanes turn each row of 16 lanes into its prefix sums (lane 15
// of a row holds the row's total); broadcasting lane 15 of row 0 (2) into row
// 1 (3), then lane 31 into rows 2 and 3, leaves the total in lane 63, which a
// readlane hands to every lane.  DPP moves take a few cycles where a
// cross-lane permute through the LDS unit takes about a hundred.
__device__ inline double wave_sum(double v)
{
    v += dpp_f64<0x111, 0xF, true>(v);   // row_shr:1
    v += dpp_f64<0x112, 0xF, true>(v);   // row_shr:2
    v += dpp_f64<0x114, 0xF, true>(v);   // row_shr:4
    v += dpp_f64<0x118, 0xF, true>(v);   // row_shr:8
    v += dpp_f64<0x142, 0xA, false>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xC, false>(v);  // row_bcast:31 into rows 2 and 3
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Orders the LDS accesses of one wave before and after this point (the
// lanes of a wave then see each other's LDS writes); no work-group barrier.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// One DPP step on an unsigned value: the row-shift and row-broadcast
// pattern of wave_sum, with lanes outside a row or a mask reading the
// identity.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned dpp_u32(unsigned x, unsigned identity)
{
    return (unsigned)__builtin_amdgcn_update_dpp((int)identity, (int)x, CTRL, ROW_MASK, 0xF, false);
}

// Minimum of a 64-bit unsigned value over the 64 lanes (all active), in
// every lane: the same pattern on both halves, compared as one value.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp_min_u64(uint64_t v)
{
    const unsigned lo = dpp_u32<CTRL, ROW_MASK>((unsigned)v, 0xFFFFFFFFu);
    const unsigned hi = dpp_u32<CTRL, ROW_MASK>((unsigned)(v >> 32), 0xFFFFFFFFu);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    return o < v ? o : v;
}

__device__ inline uint64_t wave_min_u64(uint64_t v)
{
    v = dpp_min_u64<0x111, 0xF>(v);
    v = dpp_min_u64<0x112, 0xF>(v);
    v = dpp_min_u64<0x114, 0xF>(v);
    v = dpp_min_u64<0x118, 0xF>(v);
    v = dpp_min_u64<0x142, 0xA>(v);
    v = dpp_min_u64<0x143, 0xC>(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

}  // namespace bs

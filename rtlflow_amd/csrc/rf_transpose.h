// rf_transpose.h -- wavefront-level 64 x 64 bit-matrix transpose for the
// bit-plane kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

// 64 x 64 bit-matrix transpose across wavefront 0: thread j holds row j on
// entry and column j on exit (bit t of the result = bit j of thread t's
// row). Six block-swap stages (32, 16, ..., 1), each one cross-lane exchange
// with lane t ^ M; all 64 threads of the wavefront must execute it together.
// Exchanges within 16-lane rows are DPP moves (no trip through the LDS
// unit): quad permutations for M = 1, 2; for M = 4 the half-row mirror
// (t ^ 7) then the quad reversal (t ^ 3); for M = 8 the row mirror (t ^ 15)
// then the half-row mirror. M = 16 and M = 32 use the lane-swap
// instructions: with both operands v, permlane16_swap returns v with its odd
// 16-lane rows replaced by the even rows (first result) and v with its even
// rows replaced by the odd rows (second result); permlane32_swap does the
// same for the two 32-lane halves. Lane t's partner value is the first
// result where bit M of t is set, the second elsewhere.
template <int CTRL>
__device__ __forceinline__ uint32_t bp_dpp(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, false);
}

template <uint32_t M>
__device__ __forceinline__ uint32_t bp_xor_lane(uint32_t v) {
    if constexpr (M == 1) return bp_dpp<0xB1>(v);                    // quad_perm [1,0,3,2]
    else if constexpr (M == 2) return bp_dpp<0x4E>(v);               // quad_perm [2,3,0,1]
    else if constexpr (M == 4) return bp_dpp<0x1B>(bp_dpp<0x141>(v));  // [3,2,1,0] after half mirror
    else if constexpr (M == 8) return bp_dpp<0x141>(bp_dpp<0x140>(v)); // half mirror after mirror
    else if constexpr (M == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (__lane_id() & 16u) ? r[0] : r[1];
    } else {
        static_assert(M == 32, "exchange distance");
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (__lane_id() & 32u) ? r[0] : r[1];
    }
}

template <uint32_t M>
__device__ __forceinline__ uint64_t bp_stage(uint64_t x, uint32_t t, uint64_t lo) {
    const uint64_t p = ((uint64_t)bp_xor_lane<M>((uint32_t)(x >> 32)) << 32) | bp_xor_lane<M>((uint32_t)x);
    return (t & M) ? ((x & ~lo) | ((p & ~lo) >> M)) : ((x & lo) | ((p & lo) << M));
}

__device__ __forceinline__ uint64_t bp_transpose64(uint64_t x, uint32_t t) {
    x = bp_stage<32>(x, t, 0x00000000FFFFFFFFull);
    x = bp_stage<16>(x, t, 0x0000FFFF0000FFFFull);
    x = bp_stage<8>(x, t, 0x00FF00FF00FF00FFull);
    x = bp_stage<4>(x, t, 0x0F0F0F0F0F0F0F0Full);
    x = bp_stage<2>(x, t, 0x3333333333333333ull);
    x = bp_stage<1>(x, t, 0x5555555555555555ull);
    return x;
}

}  // namespace

// common.hpp -- shared device helpers and the device-side batch layout.
//
// Literal code l = 2v + s, s = 1 for the positive literal (reference
// src/SATSolver/SolverTypes.cu:6-11).  All device work is integer and wave64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace gpsat {

constexpr int kWave = 64;

// Per-instance metadata of a batch in HBM (32 bytes).
struct InstMeta {
    int64_t lit_base;    // first literal of the instance in the arena
    int64_t coff_base;   // first entry of its (m+1) clause offsets (relative to lit_base)
    int64_t word_base;   // sum of W_j over earlier instances; rows start at A * word_base
    int32_t m;           // clauses
    int32_t n;           // variables
};

__host__ __device__ inline int32_t words_per_row(int32_t n) { return (2 * n + 31) >> 5; }

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// xor-swizzle inside each 32-lane group (ds_swizzle bitmask mode: and 0x1f, xor S).
template <int S>
__device__ __forceinline__ uint32_t swz_xor(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (S << 10));
}

// 32x32 bit-matrix transpose held one row per lane in each 32-lane half: stage S swaps
// the off-diagonal SxS blocks of every 2Sx2S block.  Per stage a lane takes its partner's
// row (lane ^ S), rotates it by S one way or the other (a per-lane rotate amount) and
// keeps its own bits outside the half it receives.
struct Transposer {
    uint32_t rot[5];
    uint32_t keep[5];
    __device__ explicit Transposer(int lane)
    {
        const uint32_t H[5] = {0xFFFF0000u, 0xFF00FF00u, 0xF0F0F0F0u, 0xCCCCCCCCu, 0xAAAAAAAAu};
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const uint32_t S = 16u >> k;
            const bool upper = (lane & S) != 0;
            rot[k] = upper ? S : 32u - S;       // rotate right by S, or left by S
            keep[k] = upper ? H[k] : ~H[k];
        }
    }
    template <int K>
    __device__ __forceinline__ uint32_t stage(uint32_t x) const
    {
        const uint32_t y = swz_xor<(16 >> K)>(x);
        const uint32_t t = __builtin_amdgcn_alignbit(y, y, rot[K]);
        return (x & keep[K]) | (t & ~keep[K]);
    }
    // afterwards lane r holds column r (bit c = old row c, bit r)
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const
    {
        x = stage<0>(x);
        x = stage<1>(x);
        x = stage<2>(x);
        x = stage<3>(x);
        return stage<4>(x);
    }
};

__device__ __forceinline__ uint32_t shfl_xor32(uint32_t x)
{
    int l = lane_id();
    return (uint32_t)__builtin_amdgcn_ds_bpermute(((l ^ 32) << 2), (int)x);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int readlane(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ int readfirst(int x) { return __builtin_amdgcn_readfirstlane(x); }

}  // namespace gpsat

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

// One stage of the 32x32 bit-matrix transpose held one row per lane: swap the
// off-diagonal SxS blocks of every 2Sx2S block.
template <int S>
__device__ __forceinline__ uint32_t tstage(uint32_t x, bool upper)
{
    constexpr uint32_t H = S == 16 ? 0xFFFF0000u : S == 8 ? 0xFF00FF00u : S == 4 ? 0xF0F0F0F0u
                         : S == 2 ? 0xCCCCCCCCu : 0xAAAAAAAAu;
    uint32_t y = swz_xor<S>(x);
    uint32_t t = upper ? (y >> S) : (y << S);
    uint32_t keep = upper ? H : ~H;
    return (x & keep) | (t & ~keep);
}

// Transpose the 32x32 bit matrix whose row r is word x of lane (r mod 32) within
// each 32-lane half: afterwards lane r holds column r (bit c = old row c, bit r).
__device__ __forceinline__ uint32_t transpose32(uint32_t x, int lane)
{
    x = tstage<16>(x, lane & 16);
    x = tstage<8>(x, lane & 8);
    x = tstage<4>(x, lane & 4);
    x = tstage<2>(x, lane & 2);
    x = tstage<1>(x, lane & 1);
    return x;
}

__device__ __forceinline__ uint32_t shfl_xor32(uint32_t x)
{
    int l = lane_id();
    return (uint32_t)__builtin_amdgcn_ds_bpermute(((l ^ 32) << 2), (int)x);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int readlane(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ int readfirst(int x) { return __builtin_amdgcn_readfirstlane(x); }

}  // namespace gpsat

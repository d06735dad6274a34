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

// Bit-field insert: (mask & a) | (~mask & b) in one v_bfi_b32 (the compiler otherwise
// splits the select across neighbouring transpose stages).
__device__ __forceinline__ uint32_t bfi(uint32_t mask, uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(mask), "v"(a), "v"(b));
    return r;
}

// xor-swizzle inside each 32-lane group (ds_swizzle bitmask mode: and 0x1f, xor S).
template <int S>
__device__ __forceinline__ uint32_t swz_xor(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (S << 10));
}

// 32x32 bit-matrix transpose held one row per lane in each 32-lane half: stage S swaps
// the off-diagonal SxS blocks of every 2Sx2S block, between lanes i and i ^ S.  The
// partner's row arrives through DPP where one move suffices -- S = 8 (row_ror:8: lane i
// reads lane (i - 8) mod 16 = i ^ 8), S = 2 and S = 1 (quad_perm) -- and through the LDS
// crossbar (ds_swizzle) for S = 16 and S = 4, where DPP would need extra moves.  For
// S = 16 and 8 the blocks are whole bytes and one v_perm_b32 with a per-lane byte
// selector merges own and partner bytes; for S = 4, 2, 1 the partner's row is rotated by
// S one way or the other (a per-lane amount) and merged with a bit-field insert.
struct Transposer {
    uint32_t sel16, sel8;
    uint32_t rot[3];
    uint32_t keep[3];
    __device__ explicit Transposer(int lane)
    {
        // v_perm_b32 byte selectors over {partner (bytes 4-7), own (bytes 0-3)}
        sel16 = (lane & 16) ? 0x03020706u : 0x05040100u;   // [y2 y3 x2 x3] | [x0 x1 y0 y1]
        sel8 = (lane & 8) ? 0x03070105u : 0x06020400u;     // [y1 x1 y3 x3] | [x0 y0 x2 y2]
        const uint32_t H[3] = {0xF0F0F0F0u, 0xCCCCCCCCu, 0xAAAAAAAAu};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const uint32_t S = 4u >> k;
            const bool upper = (lane & S) != 0;
            rot[k] = upper ? S : 32u - S;       // rotate right by S, or left by S
            keep[k] = upper ? H[k] : ~H[k];
        }
    }
    template <int K>
    __device__ __forceinline__ uint32_t bit_stage(uint32_t x, uint32_t y) const
    {
        const uint32_t t = __builtin_amdgcn_alignbit(y, y, rot[K]);
        return bfi(keep[K], x, t);
    }
    // afterwards lane r holds column r (bit c = old row c, bit r)
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const
    {
        uint32_t y = swz_xor<16>(x);
        x = __builtin_amdgcn_perm(y, x, sel16);
        y = __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);      // row_ror:8
        x = __builtin_amdgcn_perm(y, x, sel8);
        y = swz_xor<4>(x);
        x = bit_stage<0>(x, y);
        y = __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);       // quad_perm [2,3,0,1]
        x = bit_stage<1>(x, y);
        y = __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);       // quad_perm [1,0,3,2]
        return bit_stage<2>(x, y);
    }
};

// Orders this wavefront's LDS accesses: LDS requests of one wavefront are served in
// issue order, so it only has to keep the compiler from moving LDS accesses across it.
// Kernels whose workgroups hold several independent wavefronts (each with a private LDS
// region) use it instead of a workgroup barrier.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// OR of x over the 64 lanes, returned in every lane, through DPP and the permlane swaps
// (no LDS): quads, then rows of 16, then the two rows of each half, then the halves.
__device__ __forceinline__ uint32_t wave_or32(uint32_t x)
{
    // bound_ctrl with a zero old value (the OR identity) lets the compiler fold each move
    // into its OR as one v_or_b32 with a DPP source
    x |= __builtin_amdgcn_update_dpp(0u, x, 0xB1, 0xF, 0xF, true);    // quad_perm [1,0,3,2]
    x |= __builtin_amdgcn_update_dpp(0u, x, 0x4E, 0xF, 0xF, true);    // quad_perm [2,3,0,1]
    x |= __builtin_amdgcn_update_dpp(0u, x, 0x124, 0xF, 0xF, true);   // row_ror:4
    x |= __builtin_amdgcn_update_dpp(0u, x, 0x128, 0xF, 0xF, true);   // row_ror:8
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    x = r[0] | r[1];
    const auto q = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return q[0] | q[1];
}

__device__ __forceinline__ uint64_t wave_or64(uint64_t x)
{
    return (uint64_t)wave_or32((uint32_t)x) | ((uint64_t)wave_or32((uint32_t)(x >> 32)) << 32);
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

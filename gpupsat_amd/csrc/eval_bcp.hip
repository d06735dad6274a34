// eval_bcp.hip -- batched clause evaluation and BCP fixpoints over 64 assignments at once.
//
// One wavefront owns one (instance, group of <=64 assignments).  The assignment rows
// (literal-indexed bitmaps, bit l = literal l true) are transposed in LDS into
// literal masks LT[l]: a 64-bit word whose bit a says "literal l is true in
// assignment a".  A clause is then evaluated for all 64 assignments with a handful
// of 32-bit logic ops per literal (lane = clause), and the per-assignment totals are
// recovered by transposing the status masks back and popcounting (lane = assignment).
//
// Clause status follows VariablesStateHandler::clause_status
// (reference src/SATSolver/VariablesStateHandler.cu:180-206): SAT if a literal is
// true; UNSAT if every literal occurrence is false; UNIT if not SAT and exactly
// n-1 occurrences are false.  The BCP fixpoint is the unit-propagation closure that
// the reference's watched-literal BCP reaches (src/BCPStrategy/WatchedClausesList.cu:46-309).
#include "common.hpp"
#include "gpsat_internal.hpp"

namespace gpsat {

// Stage the instance's A rows (A*W words) in LDS and build LT (2 words per literal slot).
__device__ __forceinline__ void build_literal_masks(const uint32_t* __restrict__ rows, int A, int W,
                                                    uint32_t* stage, uint32_t* lt32, int lane)
{
    const int total = A * W;
    for (int w = lane; w < total; w += kWave) stage[w] = __builtin_nontemporal_load(rows + w);
    __syncthreads();
    const int half = lane >> 5, i = lane & 31;
    for (int j = 0; j < W; j++) {
        uint32_t x = (lane < A) ? stage[lane * W + j] : 0u;
        x = transpose32(x, lane);
        lt32[2 * (32 * j + i) + half] = x;
    }
    __syncthreads();
}

struct ClauseMasks {
    uint64_t T;     // some literal true
    uint64_t nf0;   // every occurrence so far false
    uint64_t nf1;   // exactly one occurrence so far non-false
};

__device__ __forceinline__ void lit_masks(const uint4* lt4, int32_t l, uint64_t& Tl, uint64_t& Fl)
{
    uint4 q = lt4[l >> 1];
    uint64_t N = (uint64_t)q.x | ((uint64_t)q.y << 32);   // literal 2v   true (var false)
    uint64_t P = (uint64_t)q.z | ((uint64_t)q.w << 32);   // literal 2v+1 true (var true)
    bool pos = l & 1;
    Tl = pos ? P : N;
    Fl = pos ? N : P;
}

__device__ __forceinline__ ClauseMasks clause_masks(const int32_t* __restrict__ cl, int len, const uint4* lt4)
{
    ClauseMasks r{0ull, ~0ull, 0ull};
    for (int k = 0; k < len; k++) {
        uint64_t Tl, Fl;
        lit_masks(lt4, cl[k], Tl, Fl);
        r.T |= Tl;
        r.nf1 = (r.nf1 & Fl) | (r.nf0 & ~Fl);
        r.nf0 &= Fl;
    }
    return r;
}

// Accumulate per-assignment totals of one 64-clause group.  b0 = SAT|UNIT, b1 = UNSAT|UNIT
// per clause (lane); after the per-half transposes lane r holds, for assignment r
// (low word) and 32+r (high word), the statuses of the 32 clauses of its half.
__device__ __forceinline__ void count_group(uint64_t b0, uint64_t b1, int lane, uint32_t (&acc)[2][3])
{
    uint32_t w0[2] = {transpose32((uint32_t)b0, lane), transpose32((uint32_t)(b0 >> 32), lane)};
    uint32_t w1[2] = {transpose32((uint32_t)b1, lane), transpose32((uint32_t)(b1 >> 32), lane)};
#pragma unroll
    for (int h = 0; h < 2; h++) {
        acc[h][0] += __builtin_popcount(w0[h] & ~w1[h]);   // SAT
        acc[h][1] += __builtin_popcount(w1[h] & ~w0[h]);   // UNSAT
        acc[h][2] += __builtin_popcount(w0[h] & w1[h]);    // UNIT
    }
}

// ---------------------------------------------------------------------------------------
// DO_EVAL: write #SAT/#UNSAT/#UNIT per (instance, assignment) for the input rows.
// DO_BCP:  write the closure rows, conflict flags and implied-literal counts.
template <bool DO_EVAL, bool DO_BCP>
__global__ __launch_bounds__(64) void k_eval_bcp(const InstMeta* __restrict__ meta, const int32_t* __restrict__ lits,
                                                  const int32_t* __restrict__ coff, const uint32_t* __restrict__ rows,
                                                  int A, int32_t* __restrict__ counts, uint32_t* __restrict__ out_rows,
                                                  int32_t* __restrict__ conflict, int32_t* __restrict__ props,
                                                  int32_t* __restrict__ passes_out)
{
    extern __shared__ __align__(16) uint32_t smem[];
    const int inst = blockIdx.x;
    const int lane = threadIdx.x;
    const InstMeta mi = meta[inst];
    const int W = words_per_row(mi.n);
    const int m = mi.m;
    uint32_t* lt32 = smem;                       // 64*W words (2 per literal slot)
    uint32_t* stage = smem + 64 * W;             // A*W words
    uint32_t* conf = stage + 64 * W;             // 2 words
    const uint4* lt4 = reinterpret_cast<const uint4*>(lt32);
    unsigned long long* lt64 = reinterpret_cast<unsigned long long*>(lt32);
    const int64_t row0 = (int64_t)A * mi.word_base;
    const int32_t* ilits = lits + mi.lit_base;
    const int32_t* icoff = coff + mi.coff_base;

    build_literal_masks(rows + row0, A, W, stage, lt32, lane);
    if (lane < 2) conf[lane] = 0u;
    __syncthreads();

    const int ngroups = (m + kWave - 1) / kWave;
    if (DO_EVAL) {
        uint32_t acc[2][3] = {{0, 0, 0}, {0, 0, 0}};
        for (int g = 0; g < ngroups; g++) {
            const int c = g * kWave + lane;
            uint64_t b0 = 0, b1 = 0;
            if (c < m) {
                const int32_t b = icoff[c], e = icoff[c + 1];
                ClauseMasks r = clause_masks(ilits + b, e - b, lt4);
                uint64_t unit = r.nf1 & ~r.T;
                uint64_t unsat = r.nf0 & ~r.T;
                b0 = r.T | unit;
                b1 = unsat | unit;
            }
            count_group(b0, b1, lane, acc);
        }
        // assignment a: low word of lanes a and a^32 (a < 32) or high words (a >= 32)
        uint32_t tot[3];
#pragma unroll
        for (int s = 0; s < 3; s++) {
            uint32_t lo = acc[0][s] + shfl_xor32(acc[0][s]);
            uint32_t hi = acc[1][s] + shfl_xor32(acc[1][s]);
            tot[s] = (lane < 32) ? lo : hi;
        }
        if (lane < A) {
            int32_t* o = counts + ((int64_t)inst * A + lane) * 3;
            o[0] = (int32_t)tot[0];
            o[1] = (int32_t)tot[1];
            o[2] = (int32_t)tot[2];
        }
    }

    if (DO_BCP) {
        // contradictory inputs (both polarities true) are conflicts from the start
        {
            uint64_t cm = 0;
            for (int v = lane; v < 16 * W; v += kWave) {
                uint4 q = lt4[v];
                cm |= ((uint64_t)q.x | ((uint64_t)q.y << 32)) & ((uint64_t)q.z | ((uint64_t)q.w << 32));
            }
            if (cm) { atomicOr(&conf[0], (uint32_t)cm); atomicOr(&conf[1], (uint32_t)(cm >> 32)); }
            __syncthreads();
        }
        int pass = 0;
        bool changed = true;
        while (changed) {
            pass++;
            const uint64_t frozen = (uint64_t)conf[0] | ((uint64_t)conf[1] << 32);
            uint64_t my_conf = 0;
            bool my_change = false;
            for (int g = 0; g < ngroups; g++) {
                const int c = g * kWave + lane;
                if (c < m) {
                    const int32_t b = icoff[c], e = icoff[c + 1];
                    ClauseMasks r = clause_masks(ilits + b, e - b, lt4);
                    my_conf |= r.nf0 & ~r.T;
                    uint64_t U = r.nf1 & ~r.T & ~frozen;
                    if (U) {
                        for (int k = b; k < e; k++) {
                            const int32_t l = ilits[k];
                            uint64_t Tl, Fl;
                            lit_masks(lt4, l, Tl, Fl);
                            uint64_t Uj = U & ~Fl;
                            if (Uj) {
                                atomicOr(&lt64[l], (unsigned long long)Uj);
                                my_change = true;
                            }
                        }
                    }
                }
            }
            my_conf &= ~frozen;
            if (my_conf) { atomicOr(&conf[0], (uint32_t)my_conf); atomicOr(&conf[1], (uint32_t)(my_conf >> 32)); }
            changed = ballot(my_change) != 0ull;
            __syncthreads();
        }
        // complementary implications are conflicts too
        {
            uint64_t cm = 0;
            for (int v = lane; v < 16 * W; v += kWave) {
                uint4 q = lt4[v];
                cm |= ((uint64_t)q.x | ((uint64_t)q.y << 32)) & ((uint64_t)q.z | ((uint64_t)q.w << 32));
            }
            if (cm) { atomicOr(&conf[0], (uint32_t)cm); atomicOr(&conf[1], (uint32_t)(cm >> 32)); }
            __syncthreads();
        }
        const uint64_t cmask = (uint64_t)conf[0] | ((uint64_t)conf[1] << 32);
        const bool my_conflict = (cmask >> lane) & 1ull;
        const int half = lane >> 5, i = lane & 31;
        uint32_t np = 0;
        uint32_t* orow = out_rows + row0;
        for (int j = 0; j < W; j++) {
            uint32_t x = lt32[2 * (32 * j + i) + half];
            x = transpose32(x, lane);            // lane a now holds word j of assignment a
            if (lane < A) {
                const uint32_t in = stage[lane * W + j];
                const uint32_t o = my_conflict ? in : x;
                np += __builtin_popcount(o & ~in);
                orow[lane * W + j] = o;
            }
        }
        if (lane < A) {
            conflict[(int64_t)inst * A + lane] = my_conflict ? 1 : 0;
            props[(int64_t)inst * A + lane] = (int32_t)np;
        }
        if (passes_out && lane == 0) passes_out[inst] = pass;
    }
}

size_t eval_bcp_lds_bytes(int max_n) { return (size_t)(128 * words_per_row(max_n) + 2) * 4 + 16; }

hipError_t launch_eval_bcp(const DeviceBatch& b, const uint32_t* rows, int A, int32_t* counts, uint32_t* out_rows,
                           int32_t* conflict, int32_t* props, int32_t* passes, bool do_eval, bool do_bcp,
                           hipStream_t st)
{
    if (b.n_inst == 0) return hipSuccess;
    const size_t lds = eval_bcp_lds_bytes(b.max_n);
    dim3 grid(b.n_inst), block(kWave);
    if (do_eval && do_bcp)
        hipLaunchKernelGGL((k_eval_bcp<true, true>), grid, block, lds, st, b.meta, b.lits, b.coff, rows, A, counts,
                           out_rows, conflict, props, passes);
    else if (do_eval)
        hipLaunchKernelGGL((k_eval_bcp<true, false>), grid, block, lds, st, b.meta, b.lits, b.coff, rows, A, counts,
                           out_rows, conflict, props, passes);
    else
        hipLaunchKernelGGL((k_eval_bcp<false, true>), grid, block, lds, st, b.meta, b.lits, b.coff, rows, A, counts,
                           out_rows, conflict, props, passes);
    return hipGetLastError();
}

}  // namespace gpsat

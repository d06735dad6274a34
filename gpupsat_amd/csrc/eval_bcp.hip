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

// Copy count words global -> LDS with 8 loads in flight per lane.
__device__ __forceinline__ void bulk_to_lds(uint32_t* dst, const uint32_t* __restrict__ src, int count, int lane)
{
    int w = lane;
    for (; w + 7 * kWave < count; w += 8 * kWave) {
        uint32_t r[8];
#pragma unroll
        for (int u = 0; u < 8; u++) r[u] = __builtin_nontemporal_load(src + w + u * kWave);
#pragma unroll
        for (int u = 0; u < 8; u++) dst[w + u * kWave] = r[u];
    }
    for (; w < count; w += kWave) dst[w] = __builtin_nontemporal_load(src + w);
}

// Stage the instance's A rows (A*W words) in LDS and build LT (2 words per literal slot).
__device__ __forceinline__ void build_literal_masks(const uint32_t* __restrict__ rows, int A, int W,
                                                    uint32_t* stage, uint32_t* lt32, int lane, const Transposer& tp)
{
    bulk_to_lds(stage, rows, A * W, lane);
    __syncthreads();
    const int half = lane >> 5, i = lane & 31;
    for (int j = 0; j < W; j++) {
        uint32_t x = (lane < A) ? stage[lane * W + j] : 0u;
        x = tp(x);
        lt32[2 * (32 * j + i) + half] = x;
    }
    __syncthreads();
}

struct ClauseMasks {
    uint64_t T;     // some literal true
    uint64_t nf0;   // every occurrence so far false
    uint64_t nf1;   // exactly one occurrence so far non-false
};

// True mask of literal l (bit a: literal l true in assignment a); its false mask is the
// true mask of l ^ 1.
__device__ __forceinline__ uint64_t tmask(const uint64_t* lt64, int32_t l) { return lt64[l]; }

__device__ __forceinline__ void lit_masks(const uint64_t* lt64, int32_t l, uint64_t& Tl, uint64_t& Fl)
{
    Tl = lt64[l];
    Fl = lt64[l ^ 1];
}

__device__ __forceinline__ ClauseMasks clause_masks(const int32_t* __restrict__ cl, int len, const uint64_t* lt64)
{
    ClauseMasks r;
    if (len == 3) {
        const int32_t a = cl[0], b = cl[1], c = cl[2];
        const uint64_t Ta = lt64[a], Tb = lt64[b], Tc = lt64[c];
        const uint64_t Fa = lt64[a ^ 1], Fb = lt64[b ^ 1], Fc = lt64[c ^ 1];
        r.T = Ta | Tb | Tc;
        r.nf0 = Fa & Fb & Fc;
        const uint64_t maj = (Fa & Fb) | (Fc & (Fa ^ Fb));   // at least two false
        r.nf1 = maj & ~r.nf0;                                   // exactly two false
        return r;
    }
    r.T = 0ull;
    r.nf0 = ~0ull;
    r.nf1 = 0ull;
    for (int k = 0; k < len; k++) {
        uint64_t Tl, Fl;
        lit_masks(lt64, cl[k], Tl, Fl);
        r.T |= Tl;
        r.nf1 = (r.nf1 & Fl) | (r.nf0 & ~Fl);
        r.nf0 &= Fl;
    }
    return r;
}

// Accumulate per-assignment totals of one 64-clause group.  b0 = SAT|UNIT, b1 = UNSAT|UNIT
// per clause (lane); after the per-half transposes lane r holds, for assignment r
// (low word) and 32+r (high word), the statuses of the 32 clauses of its half.
__device__ __forceinline__ void count_group(uint64_t b0, uint64_t b1, const Transposer& tp, uint32_t (&acc)[2][3])
{
    const uint32_t w0[2] = {tp((uint32_t)b0), tp((uint32_t)(b0 >> 32))};
    const uint32_t w1[2] = {tp((uint32_t)b1), tp((uint32_t)(b1 >> 32))};
#pragma unroll
    for (int h = 0; h < 2; h++) {
        acc[h][0] += __builtin_popcount(w0[h] & ~w1[h]);   // SAT
        acc[h][1] += __builtin_popcount(w1[h] & ~w0[h]);   // UNSAT
        acc[h][2] += __builtin_popcount(w0[h] & w1[h]);    // UNIT
    }
}

__device__ __forceinline__ uint64_t lds_mask(const uint32_t* w) { return (uint64_t)w[0] | ((uint64_t)w[1] << 32); }

__device__ __forceinline__ void publish(uint32_t* conf, uint64_t x)
{
    if (x) {
        atomicOr(&conf[0], (uint32_t)x);
        atomicOr(&conf[1], (uint32_t)(x >> 32));
    }
}

// ---------------------------------------------------------------------------------------
// DO_EVAL: write #SAT/#UNSAT/#UNIT per (instance, assignment) for the input rows.
// DO_BCP:  write the closure rows, conflict flags and implied-literal counts.
// STAGE: the instance's clause arena (literals and offsets) is copied to LDS first, so
// every clause group reads LDS only.
template <bool DO_EVAL, bool DO_BCP, bool STAGE>
__global__ __launch_bounds__(64) void k_eval_bcp(const InstMeta* __restrict__ meta, const int32_t* __restrict__ lits,
                                                  const int32_t* __restrict__ coff, const uint32_t* __restrict__ rows,
                                                  int A, int32_t* __restrict__ counts, uint32_t* __restrict__ out_rows,
                                                  int32_t* __restrict__ conflict, int32_t* __restrict__ props,
                                                  int32_t* __restrict__ passes_out)
{
    extern __shared__ __align__(16) uint32_t smem[];
    const int inst = blockIdx.x;
    const int lane = threadIdx.x;
    const Transposer tp(lane);
    const InstMeta mi = meta[inst];
    const int W = words_per_row(mi.n);
    const int m = mi.m;
    uint32_t* lt32 = smem;                       // 64*W words (2 per literal slot)
    uint32_t* stage = smem + 64 * W;             // A*W words
    uint32_t* conf = stage + 64 * W;             // 2 words
    const uint4* lt4 = reinterpret_cast<const uint4*>(lt32);
    uint64_t* lt64 = reinterpret_cast<uint64_t*>(lt32);
    const int64_t row0 = (int64_t)A * mi.word_base;
    const int32_t* ilits = lits + mi.lit_base;
    const int32_t* icoff = coff + mi.coff_base;
    if (STAGE) {
        const int L = __builtin_amdgcn_readfirstlane(icoff[m]);
        int32_t* s_coff = (int32_t*)(conf + 4);
        int32_t* s_lits = s_coff + ((m + 4) & ~3);
        bulk_to_lds((uint32_t*)s_lits, (const uint32_t*)ilits, L, lane);
        bulk_to_lds((uint32_t*)s_coff, (const uint32_t*)icoff, m + 1, lane);
        ilits = s_lits;
        icoff = s_coff;
    }

    build_literal_masks(rows + row0, A, W, stage, lt32, lane, tp);
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
                ClauseMasks r = clause_masks(ilits + b, e - b, lt64);
                const uint64_t unit = r.nf1 & ~r.T;
                const uint64_t unsat = r.nf0 & ~r.T;
                b0 = r.T | unit;
                b1 = unsat | unit;
            }
            count_group(b0, b1, tp, acc);
        }
        // assignment a: low word of lanes a and a^32 (a < 32) or high words (a >= 32)
        uint32_t tot[3];
#pragma unroll
        for (int s = 0; s < 3; s++) {
            const uint32_t lo = acc[0][s] + shfl_xor32(acc[0][s]);
            const uint32_t hi = acc[1][s] + shfl_xor32(acc[1][s]);
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
        const uint64_t amask = A >= 64 ? ~0ull : ((1ull << A) - 1ull);
        // contradictory inputs (both polarities true) are conflicts from the start
        {
            uint64_t cm = 0;
            for (int v = lane; v < 16 * W; v += kWave) {
                const uint4 q = lt4[v];
                cm |= ((uint64_t)q.x | ((uint64_t)q.y << 32)) & ((uint64_t)q.z | ((uint64_t)q.w << 32));
            }
            publish(conf, cm);
            __syncthreads();
        }
        uint64_t frozen = lds_mask(conf);
        int pass = 0;
        bool changed = (frozen & amask) != amask;
        while (changed) {
            pass++;
            bool my_change = false;
            for (int g = 0; g < ngroups; g++) {
                const int c = g * kWave + lane;
                uint64_t newconf = 0;
                if (c < m) {
                    const int32_t b = icoff[c], e = icoff[c + 1];
                    if (e - b == 3) {
                        // three literals: the unit literal is the one that is not false
                        const int32_t la = ilits[b], lb = ilits[b + 1], lc = ilits[b + 2];
                        const uint64_t Fa = lt64[la ^ 1], Fb = lt64[lb ^ 1], Fc = lt64[lc ^ 1];
                        const uint64_t T = lt64[la] | lt64[lb] | lt64[lc];
                        const uint64_t all = Fa & Fb & Fc;
                        newconf = all & ~T & ~frozen;
                        const uint64_t U = ((Fa & Fb) | (Fc & (Fa ^ Fb))) & ~all & ~T & ~frozen;
                        if (U) {
                            const uint64_t Ua = U & ~Fa, Ub = U & ~Fb, Uc = U & ~Fc;
                            if (Ua) atomicOr((unsigned long long*)&lt64[la], (unsigned long long)Ua);
                            if (Ub) atomicOr((unsigned long long*)&lt64[lb], (unsigned long long)Ub);
                            if (Uc) atomicOr((unsigned long long*)&lt64[lc], (unsigned long long)Uc);
                            my_change = true;
                        }
                    } else {
                    const ClauseMasks r = clause_masks(ilits + b, e - b, lt64);
                    newconf = r.nf0 & ~r.T & ~frozen;
                    const uint64_t U = r.nf1 & ~r.T & ~frozen;
                    if (U) {
                        for (int k = b; k < e; k++) {
                            const int32_t l = ilits[k];
                            const uint64_t Uj = U & ~lt64[l ^ 1];
                            if (Uj) {
                                atomicOr((unsigned long long*)&lt64[l], (unsigned long long)Uj);
                                my_change = true;
                            }
                        }
                    }
                    }
                }
                // a conflicting assignment stops propagating as soon as it is found
                if (ballot(newconf != 0)) {
                    publish(conf, newconf);
                    __syncthreads();
                    frozen = lds_mask(conf);
                    if ((frozen & amask) == amask) break;
                }
            }
            changed = ballot(my_change) != 0ull && (frozen & amask) != amask;
            __syncthreads();
        }
        // complementary implications are conflicts too
        {
            uint64_t cm = 0;
            for (int v = lane; v < 16 * W; v += kWave) {
                const uint4 q = lt4[v];
                cm |= ((uint64_t)q.x | ((uint64_t)q.y << 32)) & ((uint64_t)q.z | ((uint64_t)q.w << 32));
            }
            publish(conf, cm);
            __syncthreads();
        }
        const uint64_t cmask = lds_mask(conf);
        const bool my_conflict = (cmask >> lane) & 1ull;
        const int half = lane >> 5, i = lane & 31;
        uint32_t np = 0;
        uint32_t* orow = out_rows + row0;
        const bool all_conflict = (cmask & amask) == amask;
        for (int j = 0; j < W; j++) {
            uint32_t x = 0;
            if (!all_conflict) x = tp(lt32[2 * (32 * j + i) + half]);   // lane a: word j of row a
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

size_t eval_bcp_lds_bytes(int max_n) { return (size_t)(128 * words_per_row(max_n) + 4) * 4 + 16; }

static size_t stage_lds_bytes(const DeviceBatch& b)
{
    size_t mx = 0;
    for (const InstMeta& mi : b.h_meta) {
        const size_t L = (size_t)b.h_coff[mi.coff_base + mi.m];
        mx = std::max(mx, L + (((size_t)mi.m + 4) & ~(size_t)3));
    }
    return eval_bcp_lds_bytes(b.max_n) + 4 * mx;
}

template <bool E, bool B>
static void launch_variant(const DeviceBatch& b, bool stage, size_t lds, const uint32_t* rows, int A, int32_t* counts,
                           uint32_t* out_rows, int32_t* conflict, int32_t* props, int32_t* passes, hipStream_t st)
{
    dim3 grid(b.n_inst), block(kWave);
    if (stage)
        hipLaunchKernelGGL((k_eval_bcp<E, B, true>), grid, block, lds, st, b.meta, b.lits, b.coff, rows, A, counts,
                           out_rows, conflict, props, passes);
    else
        hipLaunchKernelGGL((k_eval_bcp<E, B, false>), grid, block, lds, st, b.meta, b.lits, b.coff, rows, A, counts,
                           out_rows, conflict, props, passes);
}

hipError_t launch_eval_bcp(const DeviceBatch& b, const uint32_t* rows, int A, int32_t* counts, uint32_t* out_rows,
                           int32_t* conflict, int32_t* props, int32_t* passes, bool do_eval, bool do_bcp,
                           hipStream_t st)
{
    if (b.n_inst == 0) return hipSuccess;
    // stage the arena in LDS when the largest instance keeps the wavefront under 40 KB
    const size_t staged = stage_lds_bytes(b);
    const bool stage = tuning().eval_stage_arena && staged <= 40 * 1024;
    const size_t lds = stage ? staged : eval_bcp_lds_bytes(b.max_n);
    if (do_eval && do_bcp)
        launch_variant<true, true>(b, stage, lds, rows, A, counts, out_rows, conflict, props, passes, st);
    else if (do_eval)
        launch_variant<true, false>(b, stage, lds, rows, A, counts, out_rows, conflict, props, passes, st);
    else
        launch_variant<false, true>(b, stage, lds, rows, A, counts, out_rows, conflict, props, passes, st);
    return hipGetLastError();
}

}  // namespace gpsat

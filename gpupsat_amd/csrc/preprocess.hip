// preprocess.hip -- the reference's preprocessing for a whole batch on the device.
//
// FormulaData::copy_host_clauses_to_dev (reference src/FileManager/FormulaData.cu:82-105)
// runs two removers on the host before solving:
//   RepeatedLiteralsRemover (src/Preprocessing/RepeatedLiteralsRemover.cu:25-62): a clause
//     with a literal and its negation is dropped; repeated literals keep their first
//     occurrence;
//   UnaryClausesRemover (src/Preprocessing/UnaryClausesRemover.cu:13-173): the unit
//     clauses' literals are solved, then passes over the formula remove satisfied clauses,
//     turn clauses with one non-false literal into new solved literals, delete false
//     literals from the rest, until a pass solves nothing new; a complementary pair of
//     solved literals or an all-false clause is UNSAT, an empty formula SAT.  Without any
//     unit clause no pass runs at all (UnaryClausesRemover.cu:39-58), so an empty clause
//     then survives as UNDEF.
// The passes reach the unit-propagation closure of the unit clauses, so the result is
// order-independent: the solved set is that closure, the remaining clauses are the
// non-satisfied ones with their false literals deleted, in their original order.  Here
// one wavefront takes one instance: lanes own clauses, the closure is a literal bitmap in
// LDS updated with atomic ORs (Gauss-Seidel passes), and the compaction is a wave prefix
// sum in clause order.
#include "../../include/gpsat.h"
#include "common.hpp"
#include "gpsat_internal.hpp"

namespace gpsat {

struct PreArgs {
    const InstMeta* meta;
    const int32_t* lits;        // batch arena
    const int32_t* coff;        // per-instance relative clause offsets
    int32_t* dlits;             // [n_lits] deduplicated clauses at their original positions
    int32_t* dlen;              // [n_clauses + n_inst] length after dedup, -1 for a tautology
    int32_t* out_lits;          // [n_lits] remaining literals, instance i from its lit_base
    int32_t* out_coff;          // [n_clauses + n_inst] relative offsets of the remaining clauses
    int32_t* out_m;             // [n_inst] remaining clauses
    int32_t* status;            // [n_inst] 0 SAT, 1 UNSAT, 2 UNDEF
    uint32_t* solved;           // [total_words] solved literals (closure), bitmap rows
    int32_t n_inst;
    int32_t lds_words;          // per wavefront
};

__device__ __forceinline__ bool lit_true(const uint32_t* T, int32_t l) { return (T[l >> 5] >> (l & 31)) & 1u; }

// Complementary pair in the bitmap (bits 2v and 2v+1 both set) -> true in every lane.
__device__ __forceinline__ bool complementary(const uint32_t* T, int32_t W, int lane)
{
    bool bad = false;
    for (int32_t w = lane; w < W; w += kWave) {
        const uint32_t x = T[w];
        bad |= ((x & (x >> 1)) & 0x55555555u) != 0;
    }
    return ballot(bad) != 0ull;
}

__global__ __launch_bounds__(256) void k_preprocess(PreArgs a)
{
    extern __shared__ __align__(16) uint32_t lds_all[];
    const int inst = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    const int lane = threadIdx.x % kWave;
    if (inst >= a.n_inst) return;
    uint32_t* T = lds_all + (size_t)(threadIdx.x / kWave) * a.lds_words;
    const InstMeta mi = a.meta[inst];
    const int32_t m = mi.m, n = mi.n, W = words_per_row(n);
    const int32_t* lits = a.lits + mi.lit_base;
    const int32_t* coff = a.coff + mi.coff_base;
    int32_t* dl = a.dlits + mi.lit_base;
    int32_t* dlen = a.dlen + mi.coff_base;
    for (int32_t w = lane; w < W; w += kWave) T[w] = 0u;
    wave_sync();

    // RepeatedLiteralsRemover: first occurrences kept, tautologies marked -1; unit
    // clauses seed the solved set
    bool any_unit = false;
    for (int32_t c = lane; c < m; c += kWave) {
        const int32_t b = coff[c], e = coff[c + 1];
        int32_t k = 0;
        bool taut = false;
        for (int32_t i = b; i < e && !taut; i++) {
            const int32_t l = lits[i];
            bool dup = false;
            for (int32_t j = b; j < i; j++) {
                const int32_t q = lits[j];
                if (q == l) dup = true;
                else if (q == (l ^ 1)) taut = true;
            }
            if (!dup && !taut) dl[b + k++] = l;
        }
        dlen[c] = taut ? -1 : k;
        if (!taut && k == 1) {
            const int32_t u = dl[b];
            atomicOr(&T[u >> 5], 1u << (u & 31));
            any_unit = true;
        }
    }
    wave_sync();
    const bool units = ballot(any_unit) != 0ull;

    // UnaryClausesRemover passes, to the closure
    bool conflict = false;
    if (units) {
        conflict = complementary(T, W, lane);
        bool changed = !conflict;
        while (changed) {
            bool my_change = false, my_conf = false;
            for (int32_t c = lane; c < m; c += kWave) {
                const int32_t len = dlen[c];
                if (len < 0) continue;
                const int32_t b = coff[c];
                int32_t nfalse = 0, unit = -1;
                bool sat = false;
                for (int32_t i = 0; i < len && !sat; i++) {
                    const int32_t l = dl[b + i];
                    if (lit_true(T, l)) sat = true;
                    else if (lit_true(T, l ^ 1)) nfalse++;
                    else unit = l;
                }
                if (sat) continue;
                if (nfalse == len) {
                    my_conf = true;
                } else if (nfalse == len - 1) {
                    atomicOr(&T[unit >> 5], 1u << (unit & 31));
                    my_change = true;
                }
            }
            wave_sync();
            conflict = ballot(my_conf) != 0ull || complementary(T, W, lane);
            changed = !conflict && ballot(my_change) != 0ull;
        }
    }

    // compaction in clause order: remaining clauses without their false literals
    int32_t* ol = a.out_lits + mi.lit_base;
    int32_t* oc = a.out_coff + mi.coff_base;
    int32_t run_c = 0, run_l = 0;
    if (!conflict) {
        for (int32_t c0 = 0; c0 < m; c0 += kWave) {
            const int32_t c = c0 + lane;
            int32_t keep_len = -1;
            if (c < m) {
                const int32_t len = dlen[c];
                if (len >= 0) {
                    const int32_t b = coff[c];
                    bool sat = false;
                    int32_t k = 0;
                    for (int32_t i = 0; i < len && !sat; i++) {
                        const int32_t l = dl[b + i];
                        if (units && lit_true(T, l)) sat = true;
                        else if (!(units && lit_true(T, l ^ 1))) k++;
                    }
                    if (!sat) keep_len = k;
                }
            }
            const bool keep = keep_len >= 0;
            const uint64_t km = ballot(keep);
            const int32_t before_c = __builtin_popcountll(km & ((1ull << lane) - 1ull));
            // exclusive prefix of kept literal counts across the wave
            int32_t x = keep ? keep_len : 0;
            int32_t inc = x;
            for (int d = 1; d < kWave; d <<= 1) {
                const int32_t y = __shfl_up(inc, d);
                if (lane >= d) inc += y;
            }
            const int32_t total = __shfl(inc, kWave - 1);
            const int32_t my_off = run_l + inc - x;
            if (keep) {
                oc[run_c + before_c] = my_off;
                const int32_t b = coff[c];
                int32_t k = 0;
                for (int32_t i = 0; i < dlen[c]; i++) {
                    const int32_t l = dl[b + i];
                    if (!(units && lit_true(T, l ^ 1))) ol[my_off + k++] = l;
                }
            }
            run_c += __builtin_popcountll(km);
            run_l += total;
        }
    }
    if (lane == 0) {
        oc[run_c] = run_l;
        a.out_m[inst] = conflict ? 0 : run_c;
        a.status[inst] = conflict ? GPSAT_UNSAT : (run_c == 0 ? GPSAT_SAT : GPSAT_UNDEF);
    }
    uint32_t* srow = a.solved + mi.word_base;
    for (int32_t w = lane; w < W; w += kWave) srow[w] = units ? T[w] : 0u;
}

hipError_t launch_preprocess(const DeviceBatch& b, int32_t* dlits, int32_t* dlen, int32_t* out_lits, int32_t* out_coff,
                             int32_t* out_m, int32_t* status, uint32_t* solved, hipStream_t st)
{
    if (b.n_inst == 0) return hipSuccess;
    PreArgs a;
    a.meta = b.meta;
    a.lits = b.lits;
    a.coff = b.coff;
    a.dlits = dlits;
    a.dlen = dlen;
    a.out_lits = out_lits;
    a.out_coff = out_coff;
    a.out_m = out_m;
    a.status = status;
    a.solved = solved;
    a.n_inst = b.n_inst;
    a.lds_words = (words_per_row(b.max_n) + 3) & ~3;
    int wpb = 4;
    while (wpb > 1 && (size_t)wpb * a.lds_words * 4 > 64 * 1024) wpb >>= 1;
    if ((size_t)a.lds_words * 4 > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_preprocess, dim3((b.n_inst + wpb - 1) / wpb), dim3(kWave * wpb), (size_t)wpb * a.lds_words * 4,
                       st, a);
    return hipGetLastError();
}

}  // namespace gpsat
